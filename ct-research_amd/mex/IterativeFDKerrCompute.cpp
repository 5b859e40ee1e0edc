// ErrorSlice = IterativeFDKerrCompute(PriorImage, t_length, s_length, z_length, Size,
//                                     BetaScanRange, Udomain, Pdomain, Distance, G, t)
// Drop-in for My_novel_algorithm/IterativeFDK/IterativeFDKerrCompute.cpp:3-56: the
// error slice of the iterative FDK correction for voxel slice t (1-based, as
// MATLAB passes it) of PriorImage, as a double column [LBeta*z_length*s_length]
// with s fastest (IterativeFDK.m:332 reshapes it to s x z x beta).  Image,
// BetaScanRange and Udomain single, Pdomain and G double (either class is
// accepted for each).
#include "ctr_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const char *op = "IterativeFDKerrCompute";
    ctr_mex::nargs(nrhs, 11, 11, nlhs, 1, op);
    const int nt = ctr_mex::as_count(prhs[1], "t_length", op), ns = ctr_mex::as_count(prhs[2], "s_length", op),
              nz = ctr_mex::as_count(prhs[3], "z_length", op);
    const std::vector<float> pic = ctr_mex::as_float(prhs[0], "PriorImage", op, (size_t)nt * ns * nz);
    const std::vector<double> size = ctr_mex::as_double(prhs[4], "Size", op);
    if (size.size() < 3) ctr_mex::raise(op, "size", "Size must have 3 elements");
    const std::vector<float> beta = ctr_mex::as_float(prhs[5], "BetaScanRange", op);
    const std::vector<float> u = ctr_mex::as_float(prhs[6], "Udomain", op);
    const std::vector<double> p = ctr_mex::as_double(prhs[7], "Pdomain", op);
    const std::vector<double> g = ctr_mex::as_double(prhs[9], "G", op);
    const int t = ctr_mex::as_count(prhs[10], "t", op);
    ctr_fdkerr_geom a = {nt, ns, nz, {size[0], size[1], size[2]}, ctr_mex::as_scalar(prhs[8], "Distance", op),
                         (int)beta.size(), (int)u.size(), (int)p.size(), (int)g.size(), beta.data(), u.data(),
                         p.data(), g.data(), t - 1};
    std::vector<float> err(beta.size() * (size_t)nz * ns);
    ctr_mex::check(ctr_fdk_error_slice(ctr_mex::context(op), &a, pic.data(), err.data(), 0), op);
    plhs[0] = ctr_mex::double_column(err);
}
