// SysMatrix = GenSysMatParal(height, width, Size, Center_x, Center_y, thetaRange, t_range)
// Drop-in for Projection_computation_matlab/GenSysMatParal.m: the parallel-beam
// Siddon system matrix as a sparse double matrix [Ltheta*Lt x height*width],
// rows (theta, t) with t fastest, columns (y, x) with x fastest.  Square
// pixels of pitch max(Size)/height; the rotation centre must be the image
// centre.  Rays are whole lines (the builder traced +-t_range(end) about the
// foot point).
#include "ctr_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const char *op = "GenSysMatParal";
    ctr_mex::nargs(nrhs, 7, 7, nlhs, 1, op);
    const int height = ctr_mex::as_count(prhs[0], "height", op), width = ctr_mex::as_count(prhs[1], "width", op);
    const std::vector<double> size = ctr_mex::as_double(prhs[2], "Size", op);
    if (size.empty()) ctr_mex::raise(op, "size", "Size must not be empty");
    const double cx = ctr_mex::as_scalar(prhs[3], "Center_x", op), cy = ctr_mex::as_scalar(prhs[4], "Center_y", op);
    const std::vector<double> theta = ctr_mex::as_double(prhs[5], "thetaRange", op);
    const std::vector<double> t = ctr_mex::as_double(prhs[6], "t_range", op);
    double smax = 0.0;
    for (double v : size) smax = v > smax ? v : smax;
    const double res = smax / height, wx = width * res, hy = height * res;
    if (std::fabs(cx - wx / 2) > 1e-9 * wx || std::fabs(cy - hy / 2) > 1e-9 * hy)
        ctr_mex::raise(op, "centre", "Center_x, Center_y must be the image centre (width, height) * max(Size)/height / 2");
    ctr_par_geom g = {width, height, {wx, hy}, (int)theta.size(), (int)t.size(), theta.data(), t.data()};
    ctr_mex::check(ctr_par_geom_check(&g, 0), op);
    ctr_ctx *ctx = ctr_mex::context(op);
    plhs[0] = ctr_mex::sparse_system_matrix(
        [&](int64_t *rp, int32_t *c, float *v, int64_t *nnz, unsigned f) { return ctr_par_sysmat(ctx, &g, rp, c, v, nnz, f); },
        theta.size() * t.size(), (size_t)height * width, op);
}
