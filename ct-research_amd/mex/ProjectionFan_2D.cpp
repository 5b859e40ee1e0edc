// R = ProjectionFan_2D(img, Nt, Ns, Size, beta, Pdomain, Distance)
// Drop-in for GPU_accelerated_Operator/ProjectionFan_2D/ProjectionFan_2D.cpp:3-54
// (img single in the Img2vec_Mat2Cpp2D layout; beta and Pdomain double
// columns, single accepted too).
#include "ctr_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const char *op = "ProjectionFan_2D";
    ctr_mex::nargs(nrhs, 7, 7, nlhs, 1, op);
    const int nt = ctr_mex::as_count(prhs[1], "Nt", op), ns = ctr_mex::as_count(prhs[2], "Ns", op);
    const std::vector<float> img = ctr_mex::as_float(prhs[0], "picvector", op, (size_t)nt * ns);
    const std::vector<double> size = ctr_mex::as_double(prhs[3], "Size", op);
    if (size.size() < 2) ctr_mex::raise(op, "size", "Size must have 2 elements");
    const std::vector<double> beta = ctr_mex::as_double(prhs[4], "BetaScanRange", op);
    const std::vector<double> p = ctr_mex::as_double(prhs[5], "Pdomain", op);
    ctr_fan_geom g = {nt, ns, {size[0], size[1]}, ctr_mex::as_scalar(prhs[6], "Distance", op), (int)beta.size(),
                      (int)p.size(), beta.data(), p.data()};
    ctr_mex::check(ctr_fan_geom_check(&g, 0), op);
    std::vector<float> proj(beta.size() * p.size());
    ctr_mex::check(ctr_fan_fp(ctr_mex::context(op), &g, img.data(), proj.data(), 0), op);
    plhs[0] = ctr_mex::double_column(proj);
}
