/* mex.h - the subset of MATLAB's MEX API used by the ctr gateways, for
 * exercising them without MATLAB (test harness only).  mexErrMsgIdAndTxt
 * throws a C++ exception that fakemex_call() turns into an error id and
 * message, as MATLAB would unwind to the caller. */
#ifndef FAKEMEX_MEX_H_
#define FAKEMEX_MEX_H_
#include <stddef.h>

typedef enum { mxUNKNOWN_CLASS = 0, mxDOUBLE_CLASS = 6, mxSINGLE_CLASS = 7, mxINT32_CLASS = 12 } mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;
typedef struct mxArray_tag mxArray;
typedef size_t mwIndex;   /* 64-bit MATLAB: sparse row indices and column starts */

#ifdef __cplusplus
extern "C" {
#endif
double *mxGetPr(const mxArray *a);
void *mxGetData(const mxArray *a);
size_t mxGetM(const mxArray *a);
size_t mxGetN(const mxArray *a);
size_t mxGetNumberOfElements(const mxArray *a);
double mxGetScalar(const mxArray *a);
int mxIsSingle(const mxArray *a);
int mxIsDouble(const mxArray *a);
int mxIsComplex(const mxArray *a);
mxClassID mxGetClassID(const mxArray *a);
mxArray *mxCreateDoubleMatrix(size_t m, size_t n, mxComplexity c);
/* sparse (compressed column): pr/ir hold nzmax entries, jc holds n + 1 column starts */
mxArray *mxCreateSparse(size_t m, size_t n, size_t nzmax, mxComplexity c);
mwIndex *mxGetIr(const mxArray *a);
mwIndex *mxGetJc(const mxArray *a);
int mxIsSparse(const mxArray *a);
void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...);
int mexPrintf(const char *fmt, ...);
int mexAtExit(void (*fn)(void));
/* the gateway's entry point has C linkage, as in MATLAB's own mex.h */
void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);
#ifdef __cplusplus
}
#endif
#endif
