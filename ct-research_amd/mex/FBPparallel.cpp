// img = FBPparallel(R, ThetaScanRange, Tdomain, Size, t_length, s_length [, filter])
// Drop-in for Analytic_Algorithm/FBP_type_analytic_algorithm_GPU_cpp_sourcecode/FBP_parallel_GPU/FBPparallel.cpp:3-64:
// half Shepp-Logan filter (filter = 1: half Ram-Lak) scaled by the t pitch,
// no pre-weight, linear detector interpolation x dtheta.  theta and t are
// taken in single precision, as the reference reads them.
#include "ctr_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const char *op = "FBPparallel";
    ctr_mex::nargs(nrhs, 6, 7, nlhs, 1, op);
    const std::vector<float> thf = ctr_mex::as_float(prhs[1], "ThetaScanRange", op);
    const std::vector<float> tf = ctr_mex::as_float(prhs[2], "Tdomain", op);
    const std::vector<float> R = ctr_mex::as_float(prhs[0], "R", op, thf.size() * tf.size());
    const std::vector<double> size = ctr_mex::as_double(prhs[3], "Size", op);
    if (size.size() < 2) ctr_mex::raise(op, "size", "Size must have 2 elements");
    const int nt = ctr_mex::as_count(prhs[4], "t_length", op), ns = ctr_mex::as_count(prhs[5], "s_length", op);
    const int filter = nrhs > 6 ? (int)ctr_mex::as_scalar(prhs[6], "filter", op) : CTR_FILTER_SL_HALF;
    const std::vector<double> theta(thf.begin(), thf.end()), t(tf.begin(), tf.end());
    ctr_par_geom g = {nt, ns, {size[0], size[1]}, (int)theta.size(), (int)t.size(), theta.data(), t.data()};
    ctr_mex::check(ctr_par_geom_check(&g, 1), op);
    std::vector<float> img((size_t)nt * ns);
    ctr_mex::check(ctr_par_fbp(ctr_mex::context(op), &g, filter, R.data(), img.data(), 0), op);
    plhs[0] = ctr_mex::double_column(img);
}
