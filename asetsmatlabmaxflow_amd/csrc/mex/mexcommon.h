/*
 * mexcommon.h -- argument checking and output marshalling shared by the four
 * mex-ABI host files.  Each host file exports mexFunction exactly as the
 * reference solver of the same name (maxflow/asets*_mex.c) and forwards to
 * libasets_hip.so through include/asets.h.
 *
 * Differences from the reference, all deliberate:
 *  - nrhs, classes and element counts are validated (the reference reads
 *    whatever it is given);
 *  - alpha <= 0 is rejected before iterating, with the reference's message,
 *    instead of from inside the loop (which leaked the scratch buffers);
 *  - numIter is stored as a real uint16 and time as a real float of wall
 *    seconds (the reference stores an int through the uint16 slot and a double
 *    through the single slot, and times process CPU time).
 */
#ifndef ASETS_MEXCOMMON_H
#define ASETS_MEXCOMMON_H

#include <mex.h>
#include <stdint.h>
#include <stdio.h>

#include "asets.h"

static void amx_require_single(const mxArray *a, const char *what, size_t numel) {
    char buf[256];
    if (!a || !mxIsSingle(a) || mxIsComplex(a)) {
        snprintf(buf, sizeof buf, "%s must be a real single array.", what);
        mexErrMsgTxt(buf);
    }
    if (numel != (size_t)-1 && mxGetNumberOfElements(a) != numel) {
        snprintf(buf, sizeof buf, "%s has %zu elements; pars implies %zu.", what,
                 (size_t)mxGetNumberOfElements(a), numel);
        mexErrMsgTxt(buf);
    }
}

/* pars as float32, at least n entries (the reference reads pars as float*). */
static const float *amx_pars(const mxArray *a, size_t n) {
    char buf[128];
    amx_require_single(a, "pars", (size_t)-1);
    if (mxGetNumberOfElements(a) < n) {
        snprintf(buf, sizeof buf, "pars needs %zu entries.", n);
        mexErrMsgTxt(buf);
    }
    return (const float *)mxGetData(a);
}

static int64_t amx_dim(float v, const char *what) {
    char buf[128];
    int64_t d = (int64_t)(int)v; /* (int) pars[k], as the reference */
    if (d <= 0) {
        snprintf(buf, sizeof buf, "pars: %s must be positive.", what);
        mexErrMsgTxt(buf);
    }
    return d;
}

/* Outputs u and erriter (asetsBinaryMF3D_mex.c:91-104); numIter and time are
 * created by amx_finish once the solve has returned.  plhs has max(nlhs, 1)
 * slots: with nlhs < 2 erriter is not returned, so the solver writes it into a
 * temporary array (*tmp, destroyed by amx_finish; MATLAB also frees it if the
 * call ends in an error). */
static float *amx_outputs(int nlhs, mxArray *plhs[], mwSize ndim, const mwSize *udims,
                          int32_t max_iter, float **cvg, mxArray **tmp) {
    mwSize d2[2];
    mxArray *c;
    plhs[0] = mxCreateNumericArray(ndim, udims, mxSINGLE_CLASS, mxREAL);
    d2[0] = (mwSize)max_iter;
    d2[1] = 1;
    c = mxCreateNumericArray(2, d2, mxSINGLE_CLASS, mxREAL);
    if (!plhs[0] || !c) mexErrMsgTxt("out of memory");
    if (nlhs >= 2) {
        plhs[1] = c;
        *tmp = NULL;
    } else {
        *tmp = c;
    }
    *cvg = (float *)mxGetData(c);
    return (float *)mxGetData(plhs[0]);
}

/* numIter (uint16) and time (single) outputs (asetsBinaryMF3D_mex.c:106-118)
 * and the reference's one-line report (:132); tmp: the erriter array
 * amx_outputs made when it is not returned. */
static void amx_finish(int nlhs, mxArray *plhs[], int32_t it_num, double secs,
                       const char *label, mxArray *tmp) {
    mwSize d2[2] = {1, 1};
    mxDestroyArray(tmp);
    if (nlhs > 2) {
        plhs[2] = mxCreateNumericArray(2, d2, mxUINT16_CLASS, mxREAL);
        if (!plhs[2]) mexErrMsgTxt("out of memory");
        *(uint16_t *)mxGetData(plhs[2]) = (uint16_t)it_num; /* low 16 bits, as MATLAB sees them */
    }
    if (nlhs > 3) {
        plhs[3] = mxCreateNumericArray(2, d2, mxSINGLE_CLASS, mxREAL);
        if (!plhs[3]) mexErrMsgTxt("out of memory");
        *(float *)mxGetData(plhs[3]) = (float)secs;
    }
    mexPrintf("%s: number of iterations = %i; time = %.4f sec\n", label, (int)it_num, secs);
}

static void amx_check(asets_status st) {
    if (st != ASETS_OK) mexErrMsgTxt(asets_last_error());
}

#endif
