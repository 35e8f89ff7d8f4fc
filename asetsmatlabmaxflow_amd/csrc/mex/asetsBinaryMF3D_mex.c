/*
 * asetsBinaryMF3D_mex -- mex-ABI drop-in for maxflow/asetsBinaryMF3D_mex.c.
 *
 *   [u, erriter, numIter, time] = asetsBinaryMF3D_mex(Cs, Ct, alpha, pars)
 *   [u, erriter, numIter, time] = asetsBinaryMF3D_mex(Cs, Ct, alpha, pars, u0, pt0)
 *
 * pars = single([Ny; Nx; Nz; maxIt; errBound; cc; steps])
 * (asetsBinaryMF3D_mex.c:79-85).  The six-argument form (warm start: u = u0,
 * ps = pt = pt0, flows 0) has no reference counterpart; the four-argument form
 * behaves as the reference.
 */
#include "mexcommon.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    const float *pars;
    int64_t ny, nx, nz;
    size_t n;
    asets_params p;
    float *u, *cvg;
    mxArray *cvg_tmp = NULL;
    int32_t it_num = 0;
    double secs = 0.0;
    mwSize dims[3];
    const float *u0 = NULL, *pt0 = NULL;

    if (nrhs != 4 && nrhs != 6)
        mexErrMsgTxt("asetsBinaryMF3D_mex: expected (Cs, Ct, alpha, pars[, u0, pt0]).");
    pars = amx_pars(prhs[3], 7);
    ny = amx_dim(pars[0], "Ny");
    nx = amx_dim(pars[1], "Nx");
    nz = amx_dim(pars[2], "Nz");
    p.max_iter = (int32_t)pars[3];
    p.err_bound = pars[4];
    p.cc = pars[5];
    p.steps = pars[6];
    if (p.max_iter < 0) mexErrMsgTxt("pars: maxIt must be >= 0.");
    n = (size_t)(ny * nx * nz);
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
    dims[2] = (mwSize)nz;
    u = amx_outputs(nlhs, plhs, 3, dims, p.max_iter, &cvg, &cvg_tmp);
    amx_check(asets_binary_warm(3, (const float *)mxGetData(prhs[0]),
                                (const float *)mxGetData(prhs[1]),
                                (const float *)mxGetData(prhs[2]), ny, nx, nz, &p, u0, pt0, u,
                                cvg, &it_num, &secs, NULL));
    amx_finish(nlhs, plhs, it_num, secs, "binary max flow 3D", cvg_tmp);
}
