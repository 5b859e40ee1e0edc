// Display = ImplementUI(R, Xigamadomain, Gamadomain, BetaScanRange, betaStartAngle, MaxGama,
//                       Distance, Distance_s2d, Size, z_length [, refQuirks])
// Drop-in for My_toolbox/Implement_UnitedImaging_Data/ImplementUI.cpp:3-57: curved
// (cylindrical, equiangular) detector, short scan with Parker weights,
// per-row equiangular Shepp-Logan filter, voxel-driven backprojection.
// Size(1:2) are the in-plane pixel counts at 1 mm and Size(3) the z extent
// (mm), each truncated to an integer as the reference does (:20-22).
// refQuirks = 1 places voxel centres at (i - 0.5) as ImplementUI.cu:110 does.
#include "ctr_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const char *op = "ImplementUI";
    ctr_mex::nargs(nrhs, 10, 11, nlhs, 1, op);
    const std::vector<double> xi = ctr_mex::as_double(prhs[1], "Xigamadomain", op);
    const std::vector<double> gamma = ctr_mex::as_double(prhs[2], "Gamadomain", op);
    const std::vector<double> beta = ctr_mex::as_double(prhs[3], "BetaScanRange", op);
    const std::vector<float> R = ctr_mex::as_float(prhs[0], "R", op, beta.size() * gamma.size() * xi.size());
    const std::vector<double> size = ctr_mex::as_double(prhs[8], "Size", op);
    if (size.size() < 3) ctr_mex::raise(op, "size", "Size must have 3 elements");
    const int sz[3] = {(int)size[0], (int)size[1], (int)size[2]};
    const int nz = ctr_mex::as_count(prhs[9], "z_length", op);
    const unsigned flags = (nrhs > 10 && ctr_mex::as_scalar(prhs[10], "refQuirks", op) != 0) ? CTR_REF_QUIRKS : 0u;
    ctr_curved_geom g = {sz[0], sz[1], nz, {(double)sz[0], (double)sz[1], (double)sz[2]},
                         ctr_mex::as_scalar(prhs[6], "Distance", op), ctr_mex::as_scalar(prhs[7], "Distance_s2d", op),
                         (int)beta.size(), (int)gamma.size(), (int)xi.size(), beta.data(), gamma.data(), xi.data(),
                         ctr_mex::as_scalar(prhs[4], "betaStartAngle", op), ctr_mex::as_scalar(prhs[5], "MaxGama", op)};
    ctr_mex::check(ctr_curved_geom_check(&g), op);
    std::vector<float> vol((size_t)sz[0] * sz[1] * nz);
    ctr_mex::check(ctr_curved_fdk(ctr_mex::context(op), &g, R.data(), vol.data(), flags), op);
    plhs[0] = ctr_mex::double_column(vol);
}
