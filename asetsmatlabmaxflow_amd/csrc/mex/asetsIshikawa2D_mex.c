/*
 * asetsIshikawa2D_mex -- mex-ABI drop-in for maxflow/asetsIshikawa2D_mex.c.
 *
 *   [u, erriter, numIter, time] = asetsIshikawa2D_mex(Ct, alpha, pars)
 *
 * pars = single([Ny; Nx; L; maxIt; errBound; cc; steps])
 * (asetsIshikawa2D_mex.c:77-83); Ct is [Ny Nx L]; alpha holds L-1 or L label fields of
 * the same size (only the first L-1 are used, as in the reference); u is [Ny Nx L-1].
 * Stops on cvg < errBound (asetsIshikawa2D_mex.c:205).
 */
#include "mexcommon.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    const float *pars;
    int64_t ny, nx, nz = 1, nl;
    size_t n, na;
    asets_params p;
    float *u, *cvg;
    mxArray *cvg_tmp = NULL;
    int32_t it_num = 0;
    double secs = 0.0;
    mwSize dims[4];
    int k = 0;

    if (nrhs != 3) mexErrMsgTxt("asetsIshikawa2D_mex: expected (Ct, alpha, pars).");
    pars = amx_pars(prhs[2], 7);
    ny = amx_dim(pars[k++], "Ny");
    nx = amx_dim(pars[k++], "Nx");
    nl = amx_dim(pars[k++], "L");
    p.max_iter = (int32_t)pars[k++];
    p.err_bound = pars[k++];
    p.cc = pars[k++];
    p.steps = pars[k++];
    if (p.max_iter < 0) mexErrMsgTxt("pars: maxIt must be >= 0.");
    if (nl < 2) mexErrMsgTxt("pars: the Ishikawa model needs L >= 2.");
    n = (size_t)(ny * nx * nz);
    amx_require_single(prhs[0], "Ct", n * (size_t)nl);
    amx_require_single(prhs[1], "alpha", (size_t)-1);
    na = (size_t)mxGetNumberOfElements(prhs[1]);
    if (na != n * (size_t)(nl - 1) && na != n * (size_t)nl)
        mexErrMsgTxt("alpha must hold L-1 or L label fields of the volume size.");
    dims[0] = (mwSize)ny;
    dims[1] = (mwSize)nx;
    dims[2] = (mwSize)(nl - 1);
    u = amx_outputs(nlhs, plhs, 3, dims, p.max_iter, &cvg, &cvg_tmp);
    amx_check(asets_ishikawa2d((const float *)mxGetData(prhs[0]),
                               (const float *)mxGetData(prhs[1]), ny, nx, nl, &p, u, cvg,
                               &it_num, &secs, NULL));
    amx_finish(nlhs, plhs, it_num, secs, "ishikawa model max flow 2D", cvg_tmp);
}
