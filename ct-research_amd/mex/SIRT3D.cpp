// [x, res] = SIRT3D(R, Xigamadomain, Pdomain, beta, Distance, Size, Nt, Ns, Nz, nIter, lambda
//                   [, x0 [, minLim, maxLim]])
// No reference MEX exists: Iterative_algorithm/SIRT3D.m:67-92 is a MATLAB
// script over an explicit sparse system matrix.  This gateway runs the same
// update x <- max(min(x + lambda A'((R - Ax)./(A1+1e-9))./(A'1+1e-9), maxLim), minLim)
// on the GPU with the matrix-free Siddon pair.  beta is in radians (the
// script uses degrees, SIRT3D.m:37-40).  res = ||R - A x_k|| / ||R||, k = 0..nIter.
#include <limits>

#include "ctr_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const char *op = "SIRT3D";
    ctr_mex::nargs(nrhs, 11, 14, nlhs, 2, op);
    if (nrhs == 13) ctr_mex::raise(op, "nrhs", "minLim and maxLim go together");
    const std::vector<float> xi = ctr_mex::as_float(prhs[1], "Xigamadomain", op);
    const std::vector<float> p = ctr_mex::as_float(prhs[2], "Pdomain", op);
    const std::vector<float> beta = ctr_mex::as_float(prhs[3], "BetaScanRange", op);
    const std::vector<float> R = ctr_mex::as_float(prhs[0], "R", op, beta.size() * p.size() * xi.size());
    const std::vector<double> size = ctr_mex::as_double(prhs[5], "Size", op);
    if (size.size() < 3) ctr_mex::raise(op, "size", "Size must have 3 elements");
    const int nt = ctr_mex::as_count(prhs[6], "Nt", op), ns = ctr_mex::as_count(prhs[7], "Ns", op),
              nz = ctr_mex::as_count(prhs[8], "Nz", op);
    const double niter = ctr_mex::as_scalar(prhs[9], "nIter", op);
    if (!(niter >= 0 && niter <= 1e7)) ctr_mex::raise(op, "value", "nIter must be a non-negative integer");
    const size_t nvox = (size_t)nt * ns * nz;
    std::vector<float> x(nvox, 0.0f);
    int use_x0 = 0;
    if (nrhs > 11 && mxGetNumberOfElements(prhs[11]) > 0) {
        x = ctr_mex::as_float(prhs[11], "x0", op, nvox);
        use_x0 = 1;
    }
    float lo = 0.0f, hi = std::numeric_limits<float>::infinity();
    if (nrhs == 14) {
        lo = (float)ctr_mex::as_scalar(prhs[12], "minLim", op);
        hi = (float)ctr_mex::as_scalar(prhs[13], "maxLim", op);
    }
    const int n = (int)niter;
    std::vector<double> res(n + 1, 0.0);
    ctr_sirt_opts o = {n, (float)ctr_mex::as_scalar(prhs[10], "lambda", op), lo, hi, use_x0, res.data(), nullptr};
    ctr_cone_geom g = {nt, ns, nz, {size[0], size[1], size[2]}, ctr_mex::as_scalar(prhs[4], "Distance", op),
                       (int)beta.size(), (int)p.size(), (int)xi.size(), beta.data(), p.data(), xi.data()};
    ctr_mex::check(ctr_cone_geom_check(&g, 1), op);
    ctr_mex::check(ctr_sirt3d(ctr_mex::context(op), &g, R.data(), x.data(), &o, 0), op);
    plhs[0] = ctr_mex::double_column(x);
    if (nlhs > 1) plhs[1] = ctr_mex::double_column(res);
}
