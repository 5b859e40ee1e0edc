// SysMatrix = GenSysMatFan(height, width, Size, BetaScanRange, Pdomain, RScan, Center_x, Center_y)
// Drop-in for Projection_computation_matlab/GenSysMatFan.m:6-85: the fan-beam
// Siddon system matrix as a sparse double matrix [LBeta*LP x height*width],
// rows (beta, p) with p fastest, columns (y, x) with x fastest, values the
// segment lengths.  Square pixels of pitch max(Size)/height (:14); the
// rotation centre must be the image centre.
#include "ctr_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const char *op = "GenSysMatFan";
    ctr_mex::nargs(nrhs, 8, 8, nlhs, 1, op);
    const int height = ctr_mex::as_count(prhs[0], "height", op), width = ctr_mex::as_count(prhs[1], "width", op);
    const std::vector<double> size = ctr_mex::as_double(prhs[2], "Size", op);
    if (size.empty()) ctr_mex::raise(op, "size", "Size must not be empty");
    const std::vector<double> beta = ctr_mex::as_double(prhs[3], "BetaScanRange", op);
    const std::vector<double> p = ctr_mex::as_double(prhs[4], "Pdomain", op);
    const double rscan = ctr_mex::as_scalar(prhs[5], "RScan", op);
    const double cx = ctr_mex::as_scalar(prhs[6], "Center_x", op), cy = ctr_mex::as_scalar(prhs[7], "Center_y", op);
    double smax = 0.0;
    for (double v : size) smax = v > smax ? v : smax;
    const double res = smax / height, wx = width * res, hy = height * res;
    if (std::fabs(cx - wx / 2) > 1e-9 * wx || std::fabs(cy - hy / 2) > 1e-9 * hy)
        ctr_mex::raise(op, "centre", "Center_x, Center_y must be the image centre (width, height) * max(Size)/height / 2");
    ctr_fan_geom g = {width, height, {wx, hy}, rscan, (int)beta.size(), (int)p.size(), beta.data(), p.data()};
    ctr_mex::check(ctr_fan_geom_check(&g, 0), op);
    ctr_ctx *ctx = ctr_mex::context(op);
    plhs[0] = ctr_mex::sparse_system_matrix(
        [&](int64_t *rp, int32_t *c, float *v, int64_t *nnz, unsigned f) { return ctr_fan_sysmat(ctx, &g, rp, c, v, nnz, f); },
        beta.size() * p.size(), (size_t)height * width, op);
}
