/*
 * mex.h -- minimal MATLAB mx/mex API, written for this project so the mex-ABI
 * host files (asets*_mex.c) can be built and driven without MATLAB: from C
 * tests and from Python through ctypes.  Under MATLAB, build the same host
 * files with MATLAB's own mex.h instead (the `mex` command puts it first on
 * the include path); no source change is needed.
 *
 * Only the subset the solvers use is provided.  Class IDs carry MATLAB's
 * numeric values.  mexErrMsgTxt does not return: it long-jumps to the
 * innermost mexshim_call() on the calling thread, as MATLAB unwinds to its
 * prompt.
 */
#ifndef ASETS_MEXSHIM_MEX_H
#define ASETS_MEXSHIM_MEX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef size_t mwSize;
typedef size_t mwIndex;

typedef enum {
    mxUNKNOWN_CLASS = 0,
    mxDOUBLE_CLASS = 6,
    mxSINGLE_CLASS = 7,
    mxINT8_CLASS = 8,
    mxUINT8_CLASS = 9,
    mxINT16_CLASS = 10,
    mxUINT16_CLASS = 11,
    mxINT32_CLASS = 12,
    mxUINT32_CLASS = 13
} mxClassID;

typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

typedef struct mxArray_tag mxArray;

mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls, mxComplexity cx);
mxArray *mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity cx);
void mxDestroyArray(mxArray *a);
void *mxGetData(const mxArray *a);
mwSize mxGetNumberOfElements(const mxArray *a);
mwSize mxGetNumberOfDimensions(const mxArray *a);
const mwSize *mxGetDimensions(const mxArray *a);
mxClassID mxGetClassID(const mxArray *a);
int mxIsSingle(const mxArray *a);
int mxIsComplex(const mxArray *a);

void mexErrMsgTxt(const char *msg);
void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...);
int mexPrintf(const char *fmt, ...);

/* ---- harness (not part of MATLAB's API) ---- */
typedef void (*mex_function_t)(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);
/* Calls fn; returns 0 on normal return, 1 if it raised mexErrMsgTxt (message
 * copied into err).  plhs has max(nlhs, 1) slots.  As in MATLAB, arrays fn
 * created and did not return in plhs are freed when it returns, and every array
 * it created is freed when it raises an error. */
int mexshim_call(mex_function_t fn, int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[],
                 char *err, size_t errlen);
/* 0: mexPrintf writes to stdout (default); 1: mexPrintf output is dropped. */
void mexshim_set_quiet(int quiet);
/* Arrays created and not yet freed, all threads (a leak check for tests). */
long mexshim_live_arrays(void);

#ifdef __cplusplus
}
#endif
#endif
