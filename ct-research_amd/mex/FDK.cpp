// V = FDK(R, Xigamadomain, Pdomain, beta, Distance, Size, Nt, Ns, Nz [, filter [, refQuirks]])
// Drop-in for Analytic_Algorithm/FBP_type_analytic_algorithm_GPU_cpp_sourcecode/FDK_GPU/FDK.cpp:3-61:
// cosine pre-weight, half Shepp-Logan ramp filter (filter = 1 selects the
// half Ram-Lak), bilinear backprojection x mag^2 x dbeta.  refQuirks = 1
// reproduces the reference's stale sample above the detector.
#include "ctr_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const char *op = "FDK";
    ctr_mex::nargs(nrhs, 9, 11, nlhs, 1, op);
    const std::vector<float> xi = ctr_mex::as_float(prhs[1], "Xigamadomain", op);
    const std::vector<float> p = ctr_mex::as_float(prhs[2], "Pdomain", op);
    const std::vector<float> beta = ctr_mex::as_float(prhs[3], "BetaScanRange", op);
    const std::vector<float> R = ctr_mex::as_float(prhs[0], "R", op, beta.size() * p.size() * xi.size());
    const std::vector<double> size = ctr_mex::as_double(prhs[5], "Size", op);
    if (size.size() < 3) ctr_mex::raise(op, "size", "Size must have 3 elements");
    const int nt = ctr_mex::as_count(prhs[6], "Nt", op), ns = ctr_mex::as_count(prhs[7], "Ns", op),
              nz = ctr_mex::as_count(prhs[8], "Nz", op);
    const int filter = nrhs > 9 ? (int)ctr_mex::as_scalar(prhs[9], "filter", op) : CTR_FILTER_SL_HALF;
    const unsigned flags = (nrhs > 10 && ctr_mex::as_scalar(prhs[10], "refQuirks", op) != 0) ? CTR_REF_QUIRKS : 0u;
    ctr_cone_geom g = {nt, ns, nz, {size[0], size[1], size[2]}, ctr_mex::as_scalar(prhs[4], "Distance", op),
                       (int)beta.size(), (int)p.size(), (int)xi.size(), beta.data(), p.data(), xi.data()};
    ctr_mex::check(ctr_cone_geom_check(&g, 1), op);
    std::vector<float> vol((size_t)nt * ns * nz);
    ctr_mex::check(ctr_fdk(ctr_mex::context(op), &g, filter, R.data(), vol.data(), flags), op);
    plhs[0] = ctr_mex::double_column(vol);
}
