/*
 * asetsPotts2D_mex -- mex-ABI drop-in for maxflow/asetsPotts2D_mex.c.
 *
 *   [u, erriter, numIter, time] = asetsPotts2D_mex(Ct, alpha, pars)
 *
 * pars = single([Ny; Nx; L; maxIt; errBound; cc; steps])
 * (asetsPotts2D_mex.c:80-86); Ct and alpha are [Ny Nx L].
 */
#include "mexcommon.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    const float *pars;
    int64_t ny, nx, nl;
    size_t n;
    asets_params p;
    float *u, *cvg;
    mxArray *cvg_tmp = NULL;
    int32_t it_num = 0;
    double secs = 0.0;
    mwSize dims[3];

    if (nrhs != 3) mexErrMsgTxt("asetsPotts2D_mex: expected (Ct, alpha, pars).");
    pars = amx_pars(prhs[2], 7);
    ny = amx_dim(pars[0], "Ny");
    nx = amx_dim(pars[1], "Nx");
    nl = amx_dim(pars[2], "L");
    p.max_iter = (int32_t)pars[3];
    p.err_bound = pars[4];
    p.cc = pars[5];
    p.steps = pars[6];
    if (p.max_iter < 0) mexErrMsgTxt("pars: maxIt must be >= 0.");
    n = (size_t)(ny * nx * nl);
    amx_require_single(prhs[0], "Ct", n);
    amx_require_single(prhs[1], "alpha", n);
    dims[0] = (mwSize)ny;
    dims[1] = (mwSize)nx;
    dims[2] = (mwSize)nl;
    u = amx_outputs(nlhs, plhs, 3, dims, p.max_iter, &cvg, &cvg_tmp);
    amx_check(asets_potts2d((const float *)mxGetData(prhs[0]), (const float *)mxGetData(prhs[1]),
                            ny, nx, nl, &p, u, cvg, &it_num, &secs, NULL));
    amx_finish(nlhs, plhs, it_num, secs, "potts model max flow 2D", cvg_tmp);
}
