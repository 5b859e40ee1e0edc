// fakemex.cpp - mxArray storage and the MEX API subset of mex.h, plus a C
// entry point that calls a gateway's mexFunction and reports MATLAB errors.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mex.h"

struct mxArray_tag {
    mxClassID cls;
    size_t m, n;
    std::vector<unsigned char> data;   // dense: m*n elements; sparse: nzmax doubles
    bool sparse = false;
    std::vector<mwIndex> ir, jc;       // sparse: row indices [nzmax], column starts [n+1]
};

namespace {
struct MexError {
    std::string id, msg;
};
std::vector<void (*)(void)> g_atexit;
size_t elsize(mxClassID c) { return c == mxDOUBLE_CLASS ? 8 : 4; }
}  // namespace

extern "C" {

double *mxGetPr(const mxArray *a) { return (double *)a->data.data(); }
void *mxGetData(const mxArray *a) { return (void *)a->data.data(); }
size_t mxGetM(const mxArray *a) { return a->m; }
size_t mxGetN(const mxArray *a) { return a->n; }
size_t mxGetNumberOfElements(const mxArray *a) { return a->m * a->n; }
mwIndex *mxGetIr(const mxArray *a) { return const_cast<mwIndex *>(a->ir.data()); }
mwIndex *mxGetJc(const mxArray *a) { return const_cast<mwIndex *>(a->jc.data()); }
int mxIsSparse(const mxArray *a) { return a->sparse ? 1 : 0; }
int mxIsSingle(const mxArray *a) { return a->cls == mxSINGLE_CLASS; }
int mxIsDouble(const mxArray *a) { return a->cls == mxDOUBLE_CLASS; }
int mxIsComplex(const mxArray *) { return 0; }
mxClassID mxGetClassID(const mxArray *a) { return a->cls; }

double mxGetScalar(const mxArray *a)
{
    if (a->m * a->n == 0) return 0.0;
    switch (a->cls) {
    case mxDOUBLE_CLASS: return ((const double *)a->data.data())[0];
    case mxSINGLE_CLASS: return ((const float *)a->data.data())[0];
    case mxINT32_CLASS: return ((const int *)a->data.data())[0];
    default: return 0.0;
    }
}

mxArray *mxCreateDoubleMatrix(size_t m, size_t n, mxComplexity)
{
    mxArray *a = new mxArray_tag{mxDOUBLE_CLASS, m, n, {}};
    a->data.assign(m * n * 8, 0);
    return a;
}

mxArray *mxCreateSparse(size_t m, size_t n, size_t nzmax, mxComplexity)
{
    if (nzmax == 0) nzmax = 1;   // as MATLAB does
    mxArray *a = new mxArray_tag{mxDOUBLE_CLASS, m, n, {}};
    a->sparse = true;
    a->data.assign(nzmax * 8, 0);
    a->ir.assign(nzmax, 0);
    a->jc.assign(n + 1, 0);
    return a;
}

void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw MexError{id, buf};
}

int mexPrintf(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    int r = vprintf(fmt, ap);
    va_end(ap);
    return r;
}

int mexAtExit(void (*fn)(void))
{
    g_atexit.push_back(fn);
    return 0;
}

// ---- harness entry points (ctypes) ----
mxArray *fm_create(int cls, size_t m, size_t n, const void *src)
{
    mxArray *a = new mxArray_tag{(mxClassID)cls, m, n, {}};
    a->data.resize(m * n * elsize((mxClassID)cls));
    if (src) std::memcpy(a->data.data(), src, a->data.size());
    return a;
}

void fm_destroy(mxArray *a) { delete a; }
const void *fm_data(const mxArray *a) { return a->data.data(); }
size_t fm_m(const mxArray *a) { return a->m; }
size_t fm_n(const mxArray *a) { return a->n; }
int fm_class(const mxArray *a) { return (int)a->cls; }
int fm_is_sparse(const mxArray *a) { return a->sparse ? 1 : 0; }
const mwIndex *fm_ir(const mxArray *a) { return a->ir.data(); }
const mwIndex *fm_jc(const mxArray *a) { return a->jc.data(); }

typedef void (*mexfn_t)(int, mxArray **, int, const mxArray **);

int fm_call(void *fn, int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs, char *errid, char *errmsg)
{
    try {
        ((mexfn_t)fn)(nlhs, plhs, nrhs, prhs);
        return 0;
    } catch (const MexError &e) {
        snprintf(errid, 128, "%s", e.id.c_str());
        snprintf(errmsg, 1024, "%s", e.msg.c_str());
        return 1;
    }
}

void fm_unload(void)
{
    for (auto it = g_atexit.rbegin(); it != g_atexit.rend(); ++it) (*it)();
    g_atexit.clear();
}
}
