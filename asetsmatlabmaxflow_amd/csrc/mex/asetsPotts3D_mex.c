/*
 * asetsPotts3D_mex -- mex-ABI drop-in for maxflow/asetsPotts3D_mex.c.
 *
 *   [u, erriter, numIter, time] = asetsPotts3D_mex(Ct, alpha, pars)
 *
 * pars = single([Ny; Nx; Nz; L; maxIt; errBound; cc; steps])
 * (asetsPotts3D_mex.c:84-91); Ct and alpha are [Ny Nx Nz L].
 */
#include "mexcommon.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    const float *pars;
    int64_t ny, nx, nz, nl;
    size_t n;
    asets_params p;
    float *u, *cvg;
    mxArray *cvg_tmp = NULL;
    int32_t it_num = 0;
    double secs = 0.0;
    mwSize dims[4];

    if (nrhs != 3) mexErrMsgTxt("asetsPotts3D_mex: expected (Ct, alpha, pars).");
    pars = amx_pars(prhs[2], 8);
    ny = amx_dim(pars[0], "Ny");
    nx = amx_dim(pars[1], "Nx");
    nz = amx_dim(pars[2], "Nz");
    nl = amx_dim(pars[3], "L");
    p.max_iter = (int32_t)pars[4];
    p.err_bound = pars[5];
    p.cc = pars[6];
    p.steps = pars[7];
    if (p.max_iter < 0) mexErrMsgTxt("pars: maxIt must be >= 0.");
    n = (size_t)(ny * nx * nz * nl);
    amx_require_single(prhs[0], "Ct", n);
    amx_require_single(prhs[1], "alpha", n);
    dims[0] = (mwSize)ny;
    dims[1] = (mwSize)nx;
    dims[2] = (mwSize)nz;
    dims[3] = (mwSize)nl;
    u = amx_outputs(nlhs, plhs, 4, dims, p.max_iter, &cvg, &cvg_tmp);
    amx_check(asets_potts3d((const float *)mxGetData(prhs[0]), (const float *)mxGetData(prhs[1]),
                            ny, nx, nz, nl, &p, u, cvg, &it_num, &secs, NULL));
    amx_finish(nlhs, plhs, it_num, secs, "potts model max flow 3D", cvg_tmp);
}
