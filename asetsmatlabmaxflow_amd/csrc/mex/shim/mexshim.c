/*
 * mexshim.c -- implementation of the minimal mx/mex API in mex.h.
 * mxCreateNumericArray zero-fills, as MATLAB does.
 */
#include "mex.h"

#include <setjmp.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

struct mxArray_tag {
    mxClassID cls;
    mwSize ndim;
    mwSize dims[8];
    mwSize numel;
    void *data;
};

/* Arrays created inside mexshim_call are recorded, so that -- as MATLAB does --
 * the ones the function does not return are freed when it ends, normally or
 * through mexErrMsgTxt. */
#define MAX_LIVE 64

static __thread jmp_buf *t_jmp = NULL;
static __thread char t_msg[1024];
static __thread mxArray **t_outs = NULL;
static __thread int t_nouts = 0;
static __thread mxArray *t_live[MAX_LIVE];
static __thread int t_nlive = 0;
static int g_quiet = 0;
static long g_arrays = 0;   /* arrays allocated and not yet freed (all threads) */

static void live_add(mxArray *a) {
    if (t_jmp && t_nlive < MAX_LIVE) t_live[t_nlive++] = a;
}

static void live_remove(mxArray *a) {
    int i;
    for (i = 0; i < t_nlive; ++i)
        if (t_live[i] == a) {
            t_live[i] = t_live[--t_nlive];
            return;
        }
}

static size_t elem_size(mxClassID c) {
    switch (c) {
    case mxDOUBLE_CLASS: return 8;
    case mxSINGLE_CLASS: return 4;
    case mxINT8_CLASS:
    case mxUINT8_CLASS: return 1;
    case mxINT16_CLASS:
    case mxUINT16_CLASS: return 2;
    case mxINT32_CLASS:
    case mxUINT32_CLASS: return 4;
    default: return 0;
    }
}

mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls, mxComplexity cx) {
    mxArray *a;
    mwSize i, n = 1;
    size_t es = elem_size(cls);
    if (cx != mxREAL || es == 0 || ndim > 8) return NULL;
    a = (mxArray *)calloc(1, sizeof(mxArray));
    if (!a) return NULL;
    a->cls = cls;
    a->ndim = ndim < 2 ? 2 : ndim;
    a->dims[0] = a->dims[1] = 1;
    for (i = 0; i < ndim; ++i) {
        a->dims[i] = dims[i];
        n *= dims[i];
    }
    a->numel = n;
    /* at least 8 bytes so a 1x1 output can absorb a wider store */
    a->data = calloc(n ? n : 1, es < 8 ? 8 : es);
    if (!a->data) {
        free(a);
        return NULL;
    }
    live_add(a);
    __sync_fetch_and_add(&g_arrays, 1);
    return a;
}

mxArray *mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity cx) {
    mwSize d[2];
    d[0] = m;
    d[1] = n;
    return mxCreateNumericArray(2, d, cls, cx);
}

void mxDestroyArray(mxArray *a) {
    if (!a) return;
    live_remove(a);
    free(a->data);
    free(a);
    __sync_fetch_and_sub(&g_arrays, 1);
}

long mexshim_live_arrays(void) { return __sync_fetch_and_add(&g_arrays, 0); }

void *mxGetData(const mxArray *a) { return a ? a->data : NULL; }
mwSize mxGetNumberOfElements(const mxArray *a) { return a ? a->numel : 0; }
mwSize mxGetNumberOfDimensions(const mxArray *a) { return a ? a->ndim : 0; }
const mwSize *mxGetDimensions(const mxArray *a) { return a ? a->dims : NULL; }
mxClassID mxGetClassID(const mxArray *a) { return a ? a->cls : mxUNKNOWN_CLASS; }
int mxIsSingle(const mxArray *a) { return a && a->cls == mxSINGLE_CLASS; }
int mxIsComplex(const mxArray *a) { (void)a; return 0; }

void mexErrMsgTxt(const char *msg) {
    snprintf(t_msg, sizeof t_msg, "%s", msg ? msg : "");
    if (t_jmp) longjmp(*t_jmp, 1);
    fprintf(stderr, "mexErrMsgTxt outside mexshim_call: %s\n", t_msg);
    abort();
}

void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    (void)id;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    mexErrMsgTxt(buf);
}

int mexPrintf(const char *fmt, ...) {
    int r;
    va_list ap;
    if (g_quiet) return 0;
    va_start(ap, fmt);
    r = vprintf(fmt, ap);
    va_end(ap);
    fflush(stdout);
    return r;
}

void mexshim_set_quiet(int quiet) { g_quiet = quiet; }

int mexshim_call(mex_function_t fn, int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[],
                 char *err, size_t errlen) {
    jmp_buf jb;
    jmp_buf *prev = t_jmp;
    mxArray **prev_outs = t_outs;
    int prev_n = t_nouts;
    /* plhs has max(nlhs, 1) slots (MATLAB's ans) */
    const int nslots = nlhs < 1 ? 1 : nlhs;
    int i, j;
    for (i = 0; i < nslots; ++i) plhs[i] = NULL;
    t_nlive = 0;
    t_jmp = &jb;
    t_outs = plhs;
    t_nouts = nlhs;
    if (setjmp(jb) == 0) {
        fn(nlhs, plhs, nrhs, prhs);
        t_jmp = prev;
        t_outs = prev_outs;
        t_nouts = prev_n;
        /* free the temporaries: every recorded array not returned in plhs */
        for (i = 0; i < t_nlive; ++i) {
            int returned = 0;
            for (j = 0; j < nslots; ++j) returned |= (plhs[j] == t_live[i]);
            if (!returned) {
                free(t_live[i]->data);
                free(t_live[i]);
                __sync_fetch_and_sub(&g_arrays, 1);
            }
        }
        t_nlive = 0;
        if (err && errlen) err[0] = '\0';
        return 0;
    }
    /* error path: free everything the function had created */
    t_jmp = prev;
    for (i = 0; i < t_nlive; ++i) {
        free(t_live[i]->data);
        free(t_live[i]);
        __sync_fetch_and_sub(&g_arrays, 1);
    }
    t_nlive = 0;
    for (i = 0; i < nslots; ++i) plhs[i] = NULL;
    t_outs = prev_outs;
    t_nouts = prev_n;
    if (err && errlen) snprintf(err, errlen, "%s", t_msg);
    return 1;
}
