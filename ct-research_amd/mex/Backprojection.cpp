// V = Backprojection(R, Xigamadomain, Pdomain, beta, Distance, Size, Nt, Ns, Nz)
// Drop-in for GPU_accelerated_Operator/BackprojectionCone_3D/Backprojection3D.cpp:3-60
// (also built as BackprojectionCone_3D and Backprojection3D): bilinear
// detector sample x chord of the source-to-voxel line through the voxel.
#include "ctr_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const char *op = "Backprojection";
    ctr_mex::nargs(nrhs, 9, 9, nlhs, 1, op);
    const std::vector<float> xi = ctr_mex::as_float(prhs[1], "Xigamadomain", op);
    const std::vector<float> p = ctr_mex::as_float(prhs[2], "Pdomain", op);
    const std::vector<float> beta = ctr_mex::as_float(prhs[3], "BetaScanRange", op);
    const std::vector<float> R = ctr_mex::as_float(prhs[0], "R", op, beta.size() * p.size() * xi.size());
    const std::vector<double> size = ctr_mex::as_double(prhs[5], "Size", op);
    if (size.size() < 3) ctr_mex::raise(op, "size", "Size must have 3 elements");
    const int nt = ctr_mex::as_count(prhs[6], "Nt", op), ns = ctr_mex::as_count(prhs[7], "Ns", op),
              nz = ctr_mex::as_count(prhs[8], "Nz", op);
    ctr_cone_geom g = {nt, ns, nz, {size[0], size[1], size[2]}, ctr_mex::as_scalar(prhs[4], "Distance", op),
                       (int)beta.size(), (int)p.size(), (int)xi.size(), beta.data(), p.data(), xi.data()};
    ctr_mex::check(ctr_cone_geom_check(&g, 1), op);
    std::vector<float> vol((size_t)nt * ns * nz);
    ctr_mex::check(ctr_cone_bp_ref(ctr_mex::context(op), &g, R.data(), vol.data(), 0), op);
    plhs[0] = ctr_mex::double_column(vol);
}
