// R = ProjectionCone_3D(vol, Nt, Ns, Nz, Size, beta, Pdomain, Xigamadomain, Distance)
// Drop-in for GPU_accelerated_Operator/ProjectionCone_3D/ProjectionCone_3D.cpp:3-59:
// same arguments (vol single [Nt*Ns*Nz,1]; Size [3] mm; beta, Pdomain,
// Xigamadomain single columns; Distance = SOD), double [LP*LXigama*LBeta,1] out.
#include "ctr_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const char *op = "ProjectionCone_3D";
    ctr_mex::nargs(nrhs, 9, 9, nlhs, 1, op);
    const int nt = ctr_mex::as_count(prhs[1], "Nt", op), ns = ctr_mex::as_count(prhs[2], "Ns", op),
              nz = ctr_mex::as_count(prhs[3], "Nz", op);
    const std::vector<float> vol = ctr_mex::as_float(prhs[0], "vol", op, (size_t)nt * ns * nz);
    const std::vector<double> size = ctr_mex::as_double(prhs[4], "Size", op);
    if (size.size() < 3) ctr_mex::raise(op, "size", "Size must have 3 elements");
    const std::vector<float> beta = ctr_mex::as_float(prhs[5], "BetaScanRange", op);
    const std::vector<float> p = ctr_mex::as_float(prhs[6], "Pdomain", op);
    const std::vector<float> xi = ctr_mex::as_float(prhs[7], "Xigamadomain", op);
    ctr_cone_geom g = {nt, ns, nz, {size[0], size[1], size[2]}, ctr_mex::as_scalar(prhs[8], "Distance", op),
                       (int)beta.size(), (int)p.size(), (int)xi.size(), beta.data(), p.data(), xi.data()};
    ctr_mex::check(ctr_cone_geom_check(&g, 0), op);
    std::vector<float> proj(beta.size() * p.size() * xi.size());
    ctr_mex::check(ctr_cone_fp(ctr_mex::context(op), &g, vol.data(), proj.data(), 0), op);
    plhs[0] = ctr_mex::double_column(proj);
}
