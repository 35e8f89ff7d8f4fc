/*
 * asetsBinaryMF2D_mex -- mex-ABI drop-in for maxflow/asetsBinaryMF2D_mex.c.
 *
 *   [u, erriter, numIter, time] = asetsBinaryMF2D_mex(Cs, Ct, alpha, pars)
 *   [u, erriter, numIter, time] = asetsBinaryMF2D_mex(Cs, Ct, alpha, pars, u0, pt0)
 *
 * pars = single([Ny; Nx; maxIt; errBound; cc; steps])
 * (asetsBinaryMF2D_mex.c:74-79).  The six-argument form is the warm start
 * (no reference counterpart).
 */
#include "mexcommon.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    const float *pars;
    int64_t ny, nx;
    size_t n;
    asets_params p;
    float *u, *cvg;
    mxArray *cvg_tmp = NULL;
    int32_t it_num = 0;
    double secs = 0.0;
    mwSize dims[2];
    const float *u0 = NULL, *pt0 = NULL;

    if (nrhs != 4 && nrhs != 6)
        mexErrMsgTxt("asetsBinaryMF2D_mex: expected (Cs, Ct, alpha, pars[, u0, pt0]).");
    pars = amx_pars(prhs[3], 6);
    ny = amx_dim(pars[0], "Ny");
    nx = amx_dim(pars[1], "Nx");
    p.max_iter = (int32_t)pars[2];
    p.err_bound = pars[3];
    p.cc = pars[4];
    p.steps = pars[5];
    if (p.max_iter < 0) mexErrMsgTxt("pars: maxIt must be >= 0.");
    n = (size_t)(ny * nx);
    amx_require_single(prhs[0], "Cs", n);
    amx_require_single(prhs[1], "Ct", n);
    amx_require_single(prhs[2], "alpha", n);
    if (nrhs == 6) {
        amx_require_single(prhs[4], "u0", n);
        amx_require_single(prhs[5], "pt0", n);
        u0 = (const float *)mxGetData(prhs[4]);
        pt0 = (const float *)mxGetData(prhs[5]);
    }
    dims[0] = (mwSize)ny;
    dims[1] = (mwSize)nx;
    u = amx_outputs(nlhs, plhs, 2, dims, p.max_iter, &cvg, &cvg_tmp);
    amx_check(asets_binary_warm(2, (const float *)mxGetData(prhs[0]),
                                (const float *)mxGetData(prhs[1]),
                                (const float *)mxGetData(prhs[2]), ny, nx, 1, &p, u0, pt0, u,
                                cvg, &it_num, &secs, NULL));
    amx_finish(nlhs, plhs, it_num, secs, "binary max flow 2D", cvg_tmp);
}
