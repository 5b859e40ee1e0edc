// SysMatrix = GenSysMatCircular3D(Size, Center_t, Center_s, Center_z, Pdomain, Xigamadomain,
//                                 BetaScanRange, Distance, Resolution)
// Drop-in for Projection_computation_matlab/GenSysMatCircular3D.m:8-100: the
// cone-beam Siddon system matrix as a sparse double matrix
// [LBeta*LXigama*LP x Lt*Ls*Lz], rows (beta, xi, p) with p fastest, columns
// (z, s, t) with t fastest.  Size = [Lt Ls Lz] voxel counts, Resolution the
// three voxel pitches; BetaScanRange in degrees (:20); the rotation centre
// must be the volume centre.
#include "ctr_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const char *op = "GenSysMatCircular3D";
    ctr_mex::nargs(nrhs, 9, 9, nlhs, 1, op);
    const std::vector<double> size = ctr_mex::as_double(prhs[0], "Size", op);
    const std::vector<double> res = ctr_mex::as_double(prhs[8], "Resolution", op);
    if (size.size() < 3 || res.size() < 3) ctr_mex::raise(op, "size", "Size and Resolution must have 3 elements");
    int n[3];
    double ext[3];
    const double centre[3] = {ctr_mex::as_scalar(prhs[1], "Center_t", op), ctr_mex::as_scalar(prhs[2], "Center_s", op),
                              ctr_mex::as_scalar(prhs[3], "Center_z", op)};
    for (int a = 0; a < 3; ++a) {
        if (!(size[a] >= 1) || size[a] != std::floor(size[a])) ctr_mex::raise(op, "value", "Size must hold voxel counts");
        n[a] = (int)size[a];
        ext[a] = n[a] * res[a];
        if (std::fabs(centre[a] - ext[a] / 2) > 1e-9 * std::fabs(ext[a]))
            ctr_mex::raise(op, "centre", "Center_t, Center_s, Center_z must be the volume centre Size .* Resolution / 2");
    }
    const std::vector<float> p = ctr_mex::as_float(prhs[4], "Pdomain", op);
    const std::vector<float> xi = ctr_mex::as_float(prhs[5], "Xigamadomain", op);
    const std::vector<double> beta_deg = ctr_mex::as_double(prhs[6], "BetaScanRange", op);
    std::vector<float> beta(beta_deg.size());
    for (size_t i = 0; i < beta.size(); ++i) beta[i] = (float)(beta_deg[i] * 3.141592653589793 / 180.0);
    ctr_cone_geom g = {n[0], n[1], n[2], {ext[0], ext[1], ext[2]}, ctr_mex::as_scalar(prhs[7], "Distance", op),
                       (int)beta.size(), (int)p.size(), (int)xi.size(), beta.data(), p.data(), xi.data()};
    ctr_mex::check(ctr_cone_geom_check(&g, 0), op);
    ctr_ctx *ctx = ctr_mex::context(op);
    plhs[0] = ctr_mex::sparse_system_matrix(
        [&](int64_t *rp, int32_t *c, float *v, int64_t *nnz, unsigned f) { return ctr_cone_sysmat(ctx, &g, rp, c, v, nnz, f); },
        beta.size() * xi.size() * p.size(), (size_t)n[0] * n[1] * n[2], op);
}
