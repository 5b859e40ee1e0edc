// R = ProjectionParallel_2D(img, height, width, Size, thetaRange, t_range)
// Drop-in for GPU_accelerated_Operator/ProjectionParallel_2D/ProjectionParallel_2D.cpp:3-50
// (img single, [height][width] with width fastest; Size(1) is the vertical
// and Size(2) the horizontal extent; theta and t double columns, single
// accepted too).
#include "ctr_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const char *op = "ProjectionParallel_2D";
    ctr_mex::nargs(nrhs, 6, 6, nlhs, 1, op);
    const int height = ctr_mex::as_count(prhs[1], "height", op), width = ctr_mex::as_count(prhs[2], "width", op);
    const std::vector<float> img = ctr_mex::as_float(prhs[0], "picvector", op, (size_t)height * width);
    const std::vector<double> size = ctr_mex::as_double(prhs[3], "Size", op);
    if (size.size() < 2) ctr_mex::raise(op, "size", "Size must have 2 elements");
    const std::vector<double> theta = ctr_mex::as_double(prhs[4], "thetaRange", op);
    const std::vector<double> t = ctr_mex::as_double(prhs[5], "t_range", op);
    // the t pixel axis is the image width (horizontal, Size(2))
    ctr_par_geom g = {width, height, {size[1], size[0]}, (int)theta.size(), (int)t.size(), theta.data(), t.data()};
    ctr_mex::check(ctr_par_geom_check(&g, 0), op);
    std::vector<float> proj(theta.size() * t.size());
    ctr_mex::check(ctr_par_fp(ctr_mex::context(op), &g, img.data(), proj.data(), 0), op);
    plhs[0] = ctr_mex::double_column(proj);
}
