/*
 * oracle_impl.h -- body of the CPU restatement of the ASETS continuous max-flow
 * solvers, instantiated twice by asets_oracle.c:
 *   REAL=float  : reference operation order, bit-for-bit the arithmetic of
 *                 maxflow/asetsBinaryMF{2D,3D}_mex.c and asetsPotts{2D,3D}_mex.c
 *   REAL=double : the same algorithm carried out in FP64 (error yardstick)
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls this
 * code; it is the checker the HIP kernels are compared against.
 *
 * Requires: REAL, SFX(name) defined by the includer.
 *
 * Index convention (column-major, as the reference): g = z*Nx*Ny + x*Ny + y,
 * Ny (pars[0], rows) is the fastest axis; label l lives at offset l*N.
 *
 * Face storage is this oracle's own: each voxel owns the face on its low side
 * along every axis (bx[g] sits between x-1 and x).  A face with axis index 0 is
 * never updated and stays 0; the high face of the last voxel on an axis reads 0.
 * This reproduces the reference's zero-flux boundary, which it gets from never
 * writing the index-0 faces and from a calloc'd padding row/plane
 * (asetsBinaryMF3D_mex.c:148-150, 252, 273, 290, 321-323).
 */

/* ---- binary model (Yuan-Bae-Tai), 2D and 3D --------------------------------
 * Follows asetsBinaryMF3D_mex.c:137-203 (runMaxFlow) / :205-438 (passes) and
 * asetsBinaryMF2D_mex.c:132-358.  is3d selects the 3D expression forms (6-term
 * norm and divergence); with is3d == 0, nz must be 1 and the 2D forms (4 terms)
 * of asetsBinaryMF2D_mex.c:277-278, 335-336 are used.
 */
static int SFX(binary)(const REAL *Cs, const REAL *Ct, const REAL *alpha,
                       int64_t ny, int64_t nx, int64_t nz, int is3d,
                       int max_it, REAL err_bound, REAL cc, REAL steps,
                       const REAL *u0, const REAL *pt0, int stop_rule,
                       REAL *u, float *cvg_ref, double *cvg64, int32_t *it_num,
                       REAL *out_ps, REAL *out_pt, REAL *out_bx, REAL *out_by,
                       REAL *out_bz)
{
    const int64_t N = nx * ny * nz;
    const int64_t sx = ny, sz = nx * ny;
    int64_t g, x, y, z;
    int it;
    REAL *bx, *by, *bz, *dv, *gk, *ps, *pt;

    if (N <= 0) return ORC_EBADARG;
    if (!is3d && nz != 1) return ORC_EBADARG;
    /* asetsBinaryMF3D_mex.c:317-319 raises inside the first projection pass:
       the error is observable iff at least one iteration runs. */
    if (max_it > 0)
        for (g = 0; g < N; ++g)
            if (alpha[g] <= 0) return ORC_EALPHA;

    bx = (REAL *)calloc((size_t)N, sizeof(REAL));
    by = (REAL *)calloc((size_t)N, sizeof(REAL));
    bz = (REAL *)calloc((size_t)N, sizeof(REAL));
    dv = (REAL *)calloc((size_t)N, sizeof(REAL));
    gk = (REAL *)calloc((size_t)N, sizeof(REAL));
    ps = (REAL *)calloc((size_t)N, sizeof(REAL));
    pt = (REAL *)calloc((size_t)N, sizeof(REAL));
    if (!bx || !by || !bz || !dv || !gk || !ps || !pt) {
        free(bx); free(by); free(bz); free(dv); free(gk); free(ps); free(pt);
        return ORC_ENOMEM;
    }

    /* init: asetsBinaryMF3D_mex.c:205-224 (u = [Cs-Ct >= 0], ps = pt = min).
       Warm start (no reference counterpart): u = u0, ps = pt = pt0, flows 0. */
    for (g = 0; g < N; ++g) {
        if (u0) {
            u[g] = u0[g];
            ps[g] = pt[g] = pt0[g];
        } else {
            u[g] = (Cs[g] - Ct[g] >= 0) ? (REAL)1 : (REAL)0;
            ps[g] = ORC_MIN(Cs[g], Ct[g]);
            pt[g] = ps[g];
        }
    }

    for (it = 0; it < max_it; ++it) {
        REAL err = 0;          /* sequential sum, as asetsBinaryMF3D_mex.c:401,430 */
        double err_d = 0.0;    /* FP64 sum of the same terms */

        /* gk = dv - ((ps - pt) + u/cc)      (:238-239) */
        for (g = 0; g < N; ++g)
            gk[g] = dv[g] - ((ps[g] - pt[g]) + u[g] / cc);

        /* gradient step on the interior faces (:258, :278, :297) */
        for (z = 0; z < nz; ++z)
            for (x = 0; x < nx; ++x)
                for (y = 0; y < ny; ++y) {
                    g = z * sz + x * sx + y;
                    if (x > 0) bx[g] = steps * (gk[g] - gk[g - sx]) + bx[g];
                    if (y > 0) by[g] = steps * (gk[g] - gk[g - 1]) + by[g];
                    if (is3d && z > 0) bz[g] = steps * (gk[g] - gk[g - sz]) + bz[g];
                }

        /* projection |p| <= alpha (:317-331): squares, sum and sqrt in double,
           high face first, then narrowed; gk becomes the rescale factor. */
        for (z = 0; z < nz; ++z)
            for (x = 0; x < nx; ++x)
                for (y = 0; y < ny; ++y) {
                    double t, xr, yr, zr;
                    REAL n;
                    g = z * sz + x * sx + y;
                    xr = (x + 1 < nx) ? (double)bx[g + sx] : 0.0;
                    yr = (y + 1 < ny) ? (double)by[g + 1] : 0.0;
                    t = xr * xr + (double)bx[g] * (double)bx[g];
                    t = t + yr * yr;
                    t = t + (double)by[g] * (double)by[g];
                    if (is3d) {
                        zr = (z + 1 < nz) ? (double)bz[g + sz] : 0.0;
                        t = t + zr * zr;
                        t = t + (double)bz[g] * (double)bz[g];
                    }
                    n = (REAL)sqrt(t * 0.5);
                    if (n > alpha[g]) n = n / alpha[g];
                    else n = 1;
                    gk[g] = (REAL)1 / n;
                }

        /* rescale the interior faces by the mean factor of their two voxels
           (:351-352, :372-373, :392-393); the reference multiplies by the
           double literal 0.5, i.e. one rounding of the exact product. */
        for (z = 0; z < nz; ++z)
            for (x = 0; x < nx; ++x)
                for (y = 0; y < ny; ++y) {
                    g = z * sz + x * sx + y;
                    if (x > 0)
                        bx[g] = (REAL)((double)(gk[g] + gk[g - sx]) * 0.5 * (double)bx[g]);
                    if (y > 0)
                        by[g] = (REAL)(0.5 * (double)(gk[g - 1] + gk[g]) * (double)by[g]);
                    if (is3d && z > 0)
                        bz[g] = (REAL)(0.5 * (double)(gk[g - sz] + gk[g]) * (double)bz[g]);
                }

        /* divergence, source/sink flows, multiplier (:414-432), memory order */
        for (z = 0; z < nz; ++z)
            for (x = 0; x < nx; ++x)
                for (y = 0; y < ny; ++y) {
                    REAL d, f, uc, xr, yr;
                    g = z * sz + x * sx + y;
                    xr = (x + 1 < nx) ? bx[g + sx] : (REAL)0;
                    yr = (y + 1 < ny) ? by[g + 1] : (REAL)0;
                    d = ((yr - by[g]) + xr) - bx[g];
                    if (is3d) {
                        REAL zr = (z + 1 < nz) ? bz[g + sz] : (REAL)0;
                        d = (d + zr) - bz[g];
                    }
                    dv[g] = d;
                    uc = u[g] / cc;
                    f = ((pt[g] - uc) + d) + (REAL)1 / cc;
                    ps[g] = ORC_MIN(f, Cs[g]);
                    f = (ps[g] + uc) - d;
                    pt[g] = ORC_MIN(f, Ct[g]);
                    f = cc * ((pt[g] + d) - ps[g]);
                    err += ORC_ABS(f);
                    err_d += fabs((double)f);
                    u[g] -= f;
                }

        cvg_ref[it] = (float)(err / (REAL)N);
        if (cvg64) cvg64[it] = err_d / (double)N;
        if (stop_rule == 0) {
            if ((float)(err / (REAL)N) <= (float)err_bound) break;
        } else {
            if ((float)(err_d / (double)N) <= (float)err_bound) break;
        }
    }
    *it_num = it;

    if (out_ps) memcpy(out_ps, ps, (size_t)N * sizeof(REAL));
    if (out_pt) memcpy(out_pt, pt, (size_t)N * sizeof(REAL));
    if (out_bx) memcpy(out_bx, bx, (size_t)N * sizeof(REAL));
    if (out_by) memcpy(out_by, by, (size_t)N * sizeof(REAL));
    if (out_bz) memcpy(out_bz, bz, (size_t)N * sizeof(REAL));
    free(bx); free(by); free(bz); free(dv); free(gk); free(ps); free(pt);
    return ORC_OK;
}

/* ---- Potts model (horizontal max-flow, per-label alpha), 2D and 3D ----------
 * Follows asetsPotts3D_mex.c:144-217 (runMaxFlow), :219-258 (init), :260-471
 * (per-label passes) and :474-521 (updatePSU); 2D: asetsPotts2D_mex.c:138-434.
 * Labels are processed in order inside an iteration; each label's passes read
 * only its own fields and the previous iteration's ps, so they are independent.
 */
static int SFX(potts)(const REAL *Ct, const REAL *alpha,
                      int64_t ny, int64_t nx, int64_t nz, int64_t L, int is3d,
                      int max_it, REAL err_bound, REAL cc, REAL steps, int stop_rule,
                      REAL *u, float *cvg_ref, double *cvg64, int32_t *it_num,
                      REAL *out_ps, REAL *out_pt, REAL *out_bx, REAL *out_by,
                      REAL *out_bz)
{
    const int64_t N = nx * ny * nz;
    const int64_t NL = N * L;
    const int64_t sx = ny, sz = nx * ny;
    int64_t g, x, y, z, l;
    int it;
    REAL *bx, *by, *bz, *dv, *gk, *ps, *pt, *ft;

    if (N <= 0 || L <= 0) return ORC_EBADARG;
    if (!is3d && nz != 1) return ORC_EBADARG;
    if (max_it > 0)
        for (g = 0; g < NL; ++g)
            if (alpha[g] <= 0) return ORC_EALPHA;   /* asetsPotts3D_mex.c:356-358 */

    bx = (REAL *)calloc((size_t)NL, sizeof(REAL));
    by = (REAL *)calloc((size_t)NL, sizeof(REAL));
    bz = (REAL *)calloc((size_t)NL, sizeof(REAL));
    dv = (REAL *)calloc((size_t)NL, sizeof(REAL));
    pt = (REAL *)calloc((size_t)NL, sizeof(REAL));
    gk = (REAL *)calloc((size_t)N, sizeof(REAL));
    ps = (REAL *)calloc((size_t)N, sizeof(REAL));
    ft = (REAL *)calloc((size_t)L, sizeof(REAL));
    if (!bx || !by || !bz || !dv || !pt || !gk || !ps || !ft) {
        free(bx); free(by); free(bz); free(dv); free(pt); free(gk); free(ps); free(ft);
        return ORC_ENOMEM;
    }

    /* init (asetsPotts3D_mex.c:219-258): first strict arg-min of Ct over labels
       starting from 1e30; ps = min, every pt_l = ps, u one-hot at the arg-min. */
    for (g = 0; g < N; ++g) {
        REAL mv = (REAL)1e30;
        int64_t mi = -1;
        for (l = 0; l < L; ++l)
            if (mv > Ct[g + l * N]) { mv = Ct[g + l * N]; mi = l; }
        ps[g] = mv;
        for (l = 0; l < L; ++l) {
            pt[g + l * N] = mv;
            u[g + l * N] = (l == mi) ? (REAL)1 : (REAL)0;
        }
    }

    for (it = 0; it < max_it; ++it) {
        REAL err = 0;
        double err_d = 0.0;
        int64_t k;
        for (k = 0; k < L; ++k) {
            REAL *bxk = bx + k * N, *byk = by + k * N, *bzk = bz + k * N;
            REAL *dvk = dv + k * N, *ptk = pt + k * N, *uk = u + k * N;
            const REAL *ak = alpha + k * N, *ctk = Ct + k * N;

            for (g = 0; g < N; ++g)                                   /* :273-274 */
                gk[g] = dvk[g] - ((ps[g] - ptk[g]) + uk[g] / cc);

            for (z = 0; z < nz; ++z)                                  /* :294,315,335 */
                for (x = 0; x < nx; ++x)
                    for (y = 0; y < ny; ++y) {
                        g = z * sz + x * sx + y;
                        if (x > 0) bxk[g] = steps * (gk[g] - gk[g - sx]) + bxk[g];
                        if (y > 0) byk[g] = steps * (gk[g] - gk[g - 1]) + byk[g];
                        if (is3d && z > 0) bzk[g] = steps * (gk[g] - gk[g - sz]) + bzk[g];
                    }

            for (z = 0; z < nz; ++z)                                  /* :356-370 */
                for (x = 0; x < nx; ++x)
                    for (y = 0; y < ny; ++y) {
                        double t, xr, yr, zr;
                        REAL n;
                        g = z * sz + x * sx + y;
                        xr = (x + 1 < nx) ? (double)bxk[g + sx] : 0.0;
                        yr = (y + 1 < ny) ? (double)byk[g + 1] : 0.0;
                        t = xr * xr + (double)bxk[g] * (double)bxk[g];
                        t = t + yr * yr;
                        t = t + (double)byk[g] * (double)byk[g];
                        if (is3d) {
                            zr = (z + 1 < nz) ? (double)bzk[g + sz] : 0.0;
                            t = t + zr * zr;
                            t = t + (double)bzk[g] * (double)bzk[g];
                        }
                        n = (REAL)sqrt(t * 0.5);
                        if (n > ak[g]) n = n / ak[g];
                        else n = 1;
                        gk[g] = (REAL)1 / n;
                    }

            for (z = 0; z < nz; ++z)                                  /* :391-434 */
                for (x = 0; x < nx; ++x)
                    for (y = 0; y < ny; ++y) {
                        g = z * sz + x * sx + y;
                        if (x > 0)
                            bxk[g] = (REAL)((double)(gk[g] + gk[g - sx]) * 0.5 * (double)bxk[g]);
                        if (y > 0)
                            byk[g] = (REAL)(0.5 * (double)(gk[g - 1] + gk[g]) * (double)byk[g]);
                        if (is3d && z > 0)
                            bzk[g] = (REAL)(0.5 * (double)(gk[g - sz] + gk[g]) * (double)bzk[g]);
                    }

            for (z = 0; z < nz; ++z)                                  /* :457-466 */
                for (x = 0; x < nx; ++x)
                    for (y = 0; y < ny; ++y) {
                        REAL d, t, xr, yr;
                        g = z * sz + x * sx + y;
                        xr = (x + 1 < nx) ? bxk[g + sx] : (REAL)0;
                        yr = (y + 1 < ny) ? byk[g + 1] : (REAL)0;
                        d = ((yr - byk[g]) + xr) - bxk[g];
                        if (is3d) {
                            REAL zr = (z + 1 < nz) ? bzk[g + sz] : (REAL)0;
                            d = (d + zr) - bzk[g];
                        }
                        dvk[g] = d;
                        t = (ps[g] + uk[g] / cc) - d;
                        ptk[g] = (t < ctk[g]) ? t : ctk[g];
                    }
        }

        /* source flow and multipliers, labels coupled per voxel (:495-514) */
        for (g = 0; g < N; ++g) {
            REAL s = 0;
            for (l = 0; l < L; ++l) {
                ft[l] = dv[g + l * N] + pt[g + l * N];
                s += (ft[l] - u[g + l * N] / cc);
            }
            ps[g] = s / (REAL)L + (REAL)1 / (cc * (REAL)L);
            for (l = 0; l < L; ++l) {
                REAL e = cc * (ft[l] - ps[g]);
                u[g + l * N] -= e;
                err += ORC_ABS(e);
                err_d += fabs((double)e);
            }
        }

        cvg_ref[it] = (float)(err / (REAL)NL);
        if (cvg64) cvg64[it] = err_d / (double)NL;
        if (stop_rule == 0) {
            if ((float)(err / (REAL)NL) <= (float)err_bound) break;
        } else {
            if ((float)(err_d / (double)NL) <= (float)err_bound) break;
        }
    }
    *it_num = it;

    if (out_ps) memcpy(out_ps, ps, (size_t)N * sizeof(REAL));
    if (out_pt) memcpy(out_pt, pt, (size_t)NL * sizeof(REAL));
    if (out_bx) memcpy(out_bx, bx, (size_t)NL * sizeof(REAL));
    if (out_by) memcpy(out_by, by, (size_t)NL * sizeof(REAL));
    if (out_bz) memcpy(out_bz, bz, (size_t)NL * sizeof(REAL));
    free(bx); free(by); free(bz); free(dv); free(pt); free(gk); free(ps); free(ft);
    return ORC_OK;
}

/* ---- Ishikawa model (linearly ordered labels), 2D and 3D ------------------
 * Follows asetsIshikawa3D_mex.c:141-228 (runMaxFlow), :230-270 (init) and the
 * passes :272-610; asetsIshikawa2D_mex.c:135-481 for 2D (4-face forms, stop rule
 * cvg < errBound instead of <=, :205).  L labels of Ct; L-1 cut functions u_l,
 * flow fields and alpha_l (alpha holds at least L-1 fields).  pt has L fields;
 * the top one (l = L-1) starts at 0 (init never writes it, :253-265).
 */
static int SFX(ishikawa)(const REAL *Ct, const REAL *alpha,
                         int64_t ny, int64_t nx, int64_t nz, int64_t L, int is3d,
                         int max_it, REAL err_bound, REAL cc, REAL steps, int stop_rule,
                         REAL *u, float *cvg_ref, double *cvg64, int32_t *it_num,
                         REAL *out_pt, REAL *out_bx, REAL *out_by, REAL *out_bz)
{
    const int64_t N = nx * ny * nz;
    const int64_t M = L - 1;            /* cut functions / flow fields */
    const int64_t sx = ny, sz = nx * ny;
    int64_t g, x, y, z, l;
    int it;
    REAL *bx, *by, *bz, *dv, *gk, *pt;

    if (N <= 0 || L < 2) return ORC_EBADARG;
    if (!is3d && nz != 1) return ORC_EBADARG;
    if (max_it > 0)
        for (g = 0; g < N * M; ++g)
            if (alpha[g] <= 0) return ORC_EALPHA;   /* asetsIshikawa3D_mex.c:377-379 */

    bx = (REAL *)calloc((size_t)(N * M), sizeof(REAL));
    by = (REAL *)calloc((size_t)(N * M), sizeof(REAL));
    bz = (REAL *)calloc((size_t)(N * M), sizeof(REAL));
    dv = (REAL *)calloc((size_t)(N * M), sizeof(REAL));
    gk = (REAL *)calloc((size_t)N, sizeof(REAL));
    pt = (REAL *)calloc((size_t)(N * L), sizeof(REAL));
    if (!bx || !by || !bz || !dv || !gk || !pt) {
        free(bx); free(by); free(bz); free(dv); free(gk); free(pt);
        return ORC_ENOMEM;
    }

    /* init (:230-270): first strict arg-min of Ct from 1e30; pt_l = min for
       l < L-1; u_l = 1 below the arg-min label, 0 from it on */
    for (g = 0; g < N; ++g) {
        REAL mv = (REAL)1e30;
        int64_t mi = 1000000000;
        for (l = 0; l < L; ++l)
            if (mv > Ct[g + l * N]) { mv = Ct[g + l * N]; mi = l; }
        for (l = 0; l < M; ++l) {
            pt[g + l * N] = mv;
            u[g + l * N] = (l >= mi) ? (REAL)0 : (REAL)1;
        }
    }

    for (it = 0; it < max_it; ++it) {
        REAL err = 0;
        double err_d = 0.0;
        int64_t k;
        for (k = 0; k < M; ++k) {
            REAL *bxk = bx + k * N, *byk = by + k * N, *bzk = bz + k * N;
            REAL *dvk = dv + k * N, *uk = u + k * N;
            const REAL *ptk = pt + k * N, *ptk1 = pt + (k + 1) * N;
            const REAL *ak = alpha + k * N;

            for (g = 0; g < N; ++g)                                   /* :287-288 */
                gk[g] = dvk[g] - ((ptk[g] - ptk1[g]) + uk[g] / cc);

            for (z = 0; z < nz; ++z)                                  /* :296-359 */
                for (x = 0; x < nx; ++x)
                    for (y = 0; y < ny; ++y) {
                        g = z * sz + x * sx + y;
                        if (x > 0) bxk[g] = steps * (gk[g] - gk[g - sx]) + bxk[g];
                        if (y > 0) byk[g] = steps * (gk[g] - gk[g - 1]) + byk[g];
                        if (is3d && z > 0) bzk[g] = steps * (gk[g] - gk[g - sz]) + bzk[g];
                    }

            for (z = 0; z < nz; ++z)                                  /* :361-397 */
                for (x = 0; x < nx; ++x)
                    for (y = 0; y < ny; ++y) {
                        double t, xr, yr;
                        REAL n;
                        g = z * sz + x * sx + y;
                        xr = (x + 1 < nx) ? (double)bxk[g + sx] : 0.0;
                        yr = (y + 1 < ny) ? (double)byk[g + 1] : 0.0;
                        t = xr * xr + (double)bxk[g] * (double)bxk[g];
                        t = t + yr * yr;
                        t = t + (double)byk[g] * (double)byk[g];
                        if (is3d) {
                            double zr = (z + 1 < nz) ? (double)bzk[g + sz] : 0.0;
                            t = t + zr * zr;
                            t = t + (double)bzk[g] * (double)bzk[g];
                        }
                        n = (REAL)sqrt(t * 0.5);
                        if (n > ak[g]) n = n / ak[g];
                        else n = 1;
                        gk[g] = (REAL)1 / n;
                    }

            for (z = 0; z < nz; ++z)                                  /* :399-468 */
                for (x = 0; x < nx; ++x)
                    for (y = 0; y < ny; ++y) {
                        g = z * sz + x * sx + y;
                        if (x > 0)
                            bxk[g] = (REAL)((double)(gk[g] + gk[g - sx]) * 0.5 * (double)bxk[g]);
                        if (y > 0)
                            byk[g] = (REAL)(0.5 * (double)(gk[g - 1] + gk[g]) * (double)byk[g]);
                        if (is3d && z > 0)
                            bzk[g] = (REAL)(0.5 * (double)(gk[g - sz] + gk[g]) * (double)bzk[g]);
                    }

            for (z = 0; z < nz; ++z)                                  /* :470-496 */
                for (x = 0; x < nx; ++x)
                    for (y = 0; y < ny; ++y) {
                        REAL d, xr, yr;
                        g = z * sz + x * sx + y;
                        xr = (x + 1 < nx) ? bxk[g + sx] : (REAL)0;
                        yr = (y + 1 < ny) ? byk[g + 1] : (REAL)0;
                        d = ((yr - byk[g]) + xr) - bxk[g];
                        if (is3d) {
                            REAL zr = (z + 1 < nz) ? bzk[g + sz] : (REAL)0;
                            d = (d + zr) - bzk[g];
                        }
                        dvk[g] = d;
                    }
        }

        /* sink flows, labels in order; layer l reads the already updated
           pt_{l-1} and the previous pt_{l+1} (:189-201, :498-576) */
        for (l = 0; l < L; ++l)
            for (g = 0; g < N; ++g) {
                const int64_t i = g + l * N;
                REAL f;
                if (l == 0) {
                    f = ((dv[i] + pt[i + N]) - u[i] / cc) + (REAL)1 / cc;
                } else if (l < L - 1) {
                    f = (-dv[i - N] + pt[i - N]) + u[i - N] / cc;
                    f += (dv[i] + pt[i + N]) - u[i] / cc;
                    f /= (REAL)2;
                } else {
                    f = (-dv[i - N] + pt[i - N]) + u[i - N] / cc;
                }
                pt[i] = ORC_MIN(f, Ct[i]);
            }

        /* multipliers (updateU, :578-612): voxel-major, labels inner */
        for (g = 0; g < N; ++g)
            for (l = 0; l < M; ++l) {
                const int64_t i = g + l * N;
                const REAL e = cc * ((dv[i] + pt[i + N]) - pt[i]);
                u[i] -= e;
                err += ORC_ABS(e);
                err_d += fabs((double)e);
            }

        cvg_ref[it] = (float)(err / (REAL)(N * M));
        if (cvg64) cvg64[it] = err_d / (double)(N * M);
        {
            const float c = (stop_rule == 0) ? (float)(err / (REAL)(N * M))
                                             : (float)(err_d / (double)(N * M));
            /* 3D stops on <= (asetsIshikawa3D_mex.c:213), 2D on < (2D :205) */
            if (is3d ? (c <= (float)err_bound) : (c < (float)err_bound)) break;
        }
    }
    *it_num = it;

    if (out_pt) memcpy(out_pt, pt, (size_t)(N * L) * sizeof(REAL));
    if (out_bx) memcpy(out_bx, bx, (size_t)(N * M) * sizeof(REAL));
    if (out_by) memcpy(out_by, by, (size_t)(N * M) * sizeof(REAL));
    if (out_bz) memcpy(out_bz, bz, (size_t)(N * M) * sizeof(REAL));
    free(bx); free(by); free(bz); free(dv); free(gk); free(pt);
    return ORC_OK;
}
