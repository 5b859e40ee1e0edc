// ctr_mex.h - shared helpers of the MEX gateways (header only: every MEX
// file is its own shared library).
//
// Differences from the reference gateways, all deliberate:
//   - argument count, class (single or double) and shape are checked; a bad
//     argument raises a MATLAB error "ctr:<op>:<what>" (the reference read
//     raw pointers unchecked, e.g. ProjectionCone_3D.cpp:5-20);
//   - a failed computation raises "ctr:<op>:<status>" instead of returning a
//     zero-filled output (ProjectionCone_3D.cpp:38-42);
//   - one library context per MEX file persists across calls and is released
//     at mexAtExit, instead of cudaDeviceReset() on every call (:53).
#pragma once

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "ctr/ctr.h"
#include "mex.h"

namespace ctr_mex {

inline ctr_ctx *&ctx_slot()
{
    static ctr_ctx *ctx = nullptr;
    return ctx;
}

inline void release_ctx()
{
    ctr_ctx_destroy(ctx_slot());
    ctx_slot() = nullptr;
}

inline const char *status_name(int st)
{
    switch (st) {
    case CTR_EINVAL: return "EINVAL";
    case CTR_ENOMEM: return "ENOMEM";
    case CTR_EHIP: return "EHIP";
    case CTR_ETOOBIG: return "ETOOBIG";
    case CTR_EFAULT_INJECTED: return "EFAULT_INJECTED";
    case CTR_ENODEV: return "ENODEV";
    default: return "EUNKNOWN";
    }
}

[[noreturn]] inline void raise(const char *op, const char *what, const char *msg)
{
    char id[128], text[768];
    snprintf(id, sizeof id, "ctr:%s:%s", op, what);
    snprintf(text, sizeof text, "%s: %s", op, msg);
    mexErrMsgIdAndTxt(id, "%s", text);
    throw 0;   // not reached: mexErrMsgIdAndTxt does not return
}

inline void check(int st, const char *op)
{
    if (st != CTR_OK) raise(op, status_name(st), ctr_last_error(ctx_slot()));
}

inline ctr_ctx *context(const char *op)
{
    if (!ctx_slot()) {
        ctr_ctx *c = nullptr;
        const int st = ctr_ctx_create(0, &c);
        if (st != CTR_OK) raise(op, status_name(st), ctr_last_error(nullptr));
        ctx_slot() = c;
        mexAtExit(release_ctx);
    }
    return ctx_slot();
}

inline void nargs(int nrhs, int lo, int hi, int nlhs, int max_out, const char *op)
{
    char msg[128];
    if (nrhs < lo || nrhs > hi) {
        snprintf(msg, sizeof msg, "expected %d to %d inputs, got %d", lo, hi, nrhs);
        raise(op, "nrhs", msg);
    }
    if (nlhs > max_out) {
        snprintf(msg, sizeof msg, "at most %d outputs", max_out);
        raise(op, "nlhs", msg);
    }
}

// A real numeric vector (column or row) of class single or double.
inline size_t check_vector(const mxArray *a, const char *name, const char *op)
{
    char msg[160];
    if (!a || mxIsComplex(a) || !(mxIsSingle(a) || mxIsDouble(a))) {
        snprintf(msg, sizeof msg, "%s must be a real single or double array", name);
        raise(op, "class", msg);
    }
    if (mxGetM(a) != 1 && mxGetN(a) != 1) {
        snprintf(msg, sizeof msg, "%s must be a vector", name);
        raise(op, "shape", msg);
    }
    return mxGetNumberOfElements(a);
}

inline std::vector<float> as_float(const mxArray *a, const char *name, const char *op, size_t expect = 0)
{
    const size_t n = check_vector(a, name, op);
    if (expect && n != expect) {
        char msg[160];
        snprintf(msg, sizeof msg, "%s has %zu elements, expected %zu", name, n, expect);
        raise(op, "size", msg);
    }
    std::vector<float> v(n);
    if (mxIsSingle(a)) {
        std::memcpy(v.data(), mxGetData(a), n * sizeof(float));
    } else {
        const double *d = static_cast<const double *>(mxGetData(a));
        for (size_t i = 0; i < n; ++i) v[i] = (float)d[i];
    }
    return v;
}

inline std::vector<double> as_double(const mxArray *a, const char *name, const char *op)
{
    const size_t n = check_vector(a, name, op);
    std::vector<double> v(n);
    if (mxIsDouble(a)) {
        std::memcpy(v.data(), mxGetData(a), n * sizeof(double));
    } else {
        const float *f = static_cast<const float *>(mxGetData(a));
        for (size_t i = 0; i < n; ++i) v[i] = f[i];
    }
    return v;
}

inline double as_scalar(const mxArray *a, const char *name, const char *op)
{
    if (!a || !(mxIsSingle(a) || mxIsDouble(a)) || mxGetNumberOfElements(a) != 1) {
        char msg[128];
        snprintf(msg, sizeof msg, "%s must be a numeric scalar", name);
        raise(op, "class", msg);
    }
    return mxGetScalar(a);
}

inline int as_count(const mxArray *a, const char *name, const char *op)
{
    const double v = as_scalar(a, name, op);
    if (!(v >= 1 && v <= 1e9) || v != std::floor(v)) {
        char msg[128];
        snprintf(msg, sizeof msg, "%s must be a positive integer", name);
        raise(op, "value", msg);
    }
    return (int)v;
}

// Double column vector output, converted from the float result
// (mxCreateDoubleMatrix(L,1) + float->double loop of ProjectionCone_3D.cpp:30,46-47).
inline mxArray *double_column(const std::vector<float> &v)
{
    mxArray *out = mxCreateDoubleMatrix(v.size(), 1, mxREAL);
    double *d = mxGetPr(out);
    for (size_t i = 0; i < v.size(); ++i) d[i] = (double)v[i];
    return out;
}

inline mxArray *double_column(const std::vector<double> &v)
{
    mxArray *out = mxCreateDoubleMatrix(v.size(), 1, mxREAL);
    std::memcpy(mxGetPr(out), v.data(), v.size() * sizeof(double));
    return out;
}

// The system matrix of one geometry as a MATLAB sparse double matrix
// (rays x voxels, as sparse(L, J, W, ...) of GenSysMatFan.m:85 builds it).
// sysmat(row_ptr, col, val, &nnz, flags) is one of ctr_*_sysmat bound to its
// context and geometry.  MATLAB stores compressed columns with ascending row
// indices, so the CSR rows are transposed by a counting sort over columns.
template <class Sysmat>
inline mxArray *sparse_system_matrix(Sysmat sysmat, size_t n_rays, size_t n_vox, const char *op)
{
    std::vector<int64_t> row_ptr(n_rays + 1);
    int64_t nnz = 0;
    check(sysmat(row_ptr.data(), nullptr, nullptr, &nnz, 0u), op);
    std::vector<int32_t> col((size_t)(nnz > 0 ? nnz : 1));
    std::vector<float> val(col.size());
    check(sysmat(row_ptr.data(), col.data(), val.data(), &nnz, 0u), op);
    mxArray *out = mxCreateSparse(n_rays, n_vox, (size_t)nnz, mxREAL);
    mwIndex *jc = mxGetJc(out), *ir = mxGetIr(out);
    double *pr = mxGetPr(out);
    for (size_t j = 0; j <= n_vox; ++j) jc[j] = 0;
    for (int64_t k = 0; k < nnz; ++k) ++jc[(size_t)col[k] + 1];
    for (size_t j = 0; j < n_vox; ++j) jc[j + 1] += jc[j];
    std::vector<mwIndex> next(jc, jc + n_vox);
    for (size_t r = 0; r < n_rays; ++r)
        for (int64_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
            const mwIndex dst = next[(size_t)col[k]]++;
            ir[dst] = (mwIndex)r;
            pr[dst] = (double)val[k];
        }
    return out;
}

}  // namespace ctr_mex
