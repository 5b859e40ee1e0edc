// img = FBPfan(R, beta, Pdomain, Size, Nt, Ns, Rscan [, filter])
// Drop-in for Analytic_Algorithm/FBP_type_analytic_algorithm_GPU_cpp_sourcecode/FBP_fan_GPU/FBPfan.cpp:3-69:
// flat-detector fan FBP, pre-weight Rscan/sqrt(Rscan^2+p^2), half
// Shepp-Logan filter (filter = 1: half Ram-Lak), backprojection / U^2 x dbeta.
#include "ctr_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const char *op = "FBPfan";
    ctr_mex::nargs(nrhs, 7, 8, nlhs, 1, op);
    const std::vector<float> bf = ctr_mex::as_float(prhs[1], "BetaScanRange", op);
    const std::vector<float> pf = ctr_mex::as_float(prhs[2], "Pdomain", op);
    const std::vector<float> R = ctr_mex::as_float(prhs[0], "R", op, bf.size() * pf.size());
    const std::vector<double> size = ctr_mex::as_double(prhs[3], "Size", op);
    if (size.size() < 2) ctr_mex::raise(op, "size", "Size must have 2 elements");
    const int nt = ctr_mex::as_count(prhs[4], "Nt", op), ns = ctr_mex::as_count(prhs[5], "Ns", op);
    const int filter = nrhs > 7 ? (int)ctr_mex::as_scalar(prhs[7], "filter", op) : CTR_FILTER_SL_HALF;
    const std::vector<double> beta(bf.begin(), bf.end()), p(pf.begin(), pf.end());
    ctr_fan_geom g = {nt, ns, {size[0], size[1]}, ctr_mex::as_scalar(prhs[6], "Rscan", op), (int)beta.size(),
                      (int)p.size(), beta.data(), p.data()};
    ctr_mex::check(ctr_fan_geom_check(&g, 1), op);
    std::vector<float> img((size_t)nt * ns);
    ctr_mex::check(ctr_fan_fbp(ctr_mex::context(op), &g, CTR_FAN_FLAT, filter, R.data(), img.data(), 0), op);
    plhs[0] = ctr_mex::double_column(img);
}
