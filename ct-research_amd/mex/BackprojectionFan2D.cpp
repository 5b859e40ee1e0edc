// img = BackprojectionFan2D(R, beta, Pdomain, Size, Nt, Ns, Rscan [, refQuirks])
// Drop-in for GPU_accelerated_Operator/BackprojectionFan_2D/BackprojectionFan2D.cpp:3-66
// (beta and Pdomain are single in the reference and are rounded to single
// here).  refQuirks = 1 reproduces the reference's entry-face test and its
// chord measured to the ray's far end point.
#include "ctr_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const char *op = "BackprojectionFan2D";
    ctr_mex::nargs(nrhs, 7, 8, nlhs, 1, op);
    const std::vector<float> bf = ctr_mex::as_float(prhs[1], "BetaScanRange", op);
    const std::vector<float> pf = ctr_mex::as_float(prhs[2], "Pdomain", op);
    const std::vector<float> R = ctr_mex::as_float(prhs[0], "R", op, bf.size() * pf.size());
    const std::vector<double> size = ctr_mex::as_double(prhs[3], "Size", op);
    if (size.size() < 2) ctr_mex::raise(op, "size", "Size must have 2 elements");
    const int nt = ctr_mex::as_count(prhs[4], "Nt", op), ns = ctr_mex::as_count(prhs[5], "Ns", op);
    const unsigned flags = (nrhs > 7 && ctr_mex::as_scalar(prhs[7], "refQuirks", op) != 0) ? CTR_REF_QUIRKS : 0u;
    const std::vector<double> beta(bf.begin(), bf.end()), p(pf.begin(), pf.end());
    ctr_fan_geom g = {nt, ns, {size[0], size[1]}, ctr_mex::as_scalar(prhs[6], "Rscan", op), (int)beta.size(),
                      (int)p.size(), beta.data(), p.data()};
    ctr_mex::check(ctr_fan_geom_check(&g, 1), op);
    std::vector<float> img((size_t)nt * ns);
    ctr_mex::check(ctr_fan_bp_ref(ctr_mex::context(op), &g, R.data(), img.data(), flags), op);
    plhs[0] = ctr_mex::double_column(img);
}
