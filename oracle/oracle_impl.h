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

/* ---- hierarchical max-flow (HMF), full-flow variant -------------------------
 * Follows maxflow/asetsHMF2D.m and asetsHMF3D.m: MaxFullFlow (2D :78-88, 3D
 * :77-87), InitializeFullFlow (2D :107-127, 3D :105-130), UpdateSinkFlows (2D
 * :309-338, 3D :307-336), UpdateLabels (2D :206-217, 3D :191-202),
 * UpdateSpatialFlows (2D :220-306, 3D :205-304), CalculateCostsFullFlow (2D
 * :184-203).  Every statement of the reference is a whole-array expression in the
 * class of Ct; here each element is computed in REAL in the same operation order.
 *
 * Tree: parent[n] is the index of node n's parent, -1 for the single root; the
 * children of a node are visited in increasing index order (the order of the
 * reference's cell array h.C).  Node rows of N elements: Ct (read for leaves),
 * alpha (per voxel; a scalar alpha is a constant row), l[n][3] (the shape-
 * constraint directions lx, ly, lz along dims 1, 2, 3, used when has_l[n]).
 *
 * Faces (own-face, as above): f[n][0] along dim 1 (the reference's px, stride 1),
 * f[n][1] along dim 2 (py, stride Ny), f[n][2] along dim 3 (pz, stride Nx*Ny).
 * The reference's px(i,j,k), the face between rows i and i+1 (1-based), is
 * f[n][0] at row i (0-based), i.e. px == f[n][0][1:, :, :].
 *
 * Choices the reference leaves to MATLAB's runtime, fixed here:
 *   .^2 is x*x, .^0.5 is the correctly rounded square root;
 *   min(Ct, x) is (Ct < x) ? Ct : x;  max(x, 0) is (x > 0) ? x : +0;
 *   the masks (px>0 & l>0).*l.*a select one product or +0;
 *   cc is MATLAB's double: every operation with REAL arrays uses (REAL)cc, but the
 *   root's 1/cc .* ones(h.D) (asetsHMF3D.m:320, asetsHMF2D.m likewise) is a double
 *   array, rounded to REAL only where it meets the children's pt: (REAL)(1.0/cc);
 *   the 0.001 of the normalisation is (REAL)0.001;
 *   the costs are sums, in double, of the REAL products u*Ct and u*div.
 */
#define HMF_MAXN 4096

static void SFX(hmf_sink)(int n, const int32_t *parent, const int32_t *first_child,
                          const int32_t *next_sib, int64_t N, REAL cc, REAL root_cc,
                          const REAL *Ct, REAL *u, REAL *pt, REAL *div)
{
    /* children first (bottom-up), then this node (asetsHMF3D.m:307-336) */
    int c, nc = 0;
    int64_t g;
    REAL *ptn = pt + (int64_t)n * N;
    const REAL *dvn = div + (int64_t)n * N, *un = u + (int64_t)n * N;
    for (c = first_child[n]; c >= 0; c = next_sib[c]) {
        SFX(hmf_sink)(c, parent, first_child, next_sib, N, cc, root_cc, Ct, u, pt, div);
        ++nc;
    }
    if (nc == 0) {
        /* leaf: pt = min(Ct, P.pt - div + u/cc) */
        const REAL *ppt = pt + (int64_t)parent[n] * N, *ct = Ct + (int64_t)n * N;
        for (g = 0; g < N; ++g) {
            const REAL f = (ppt[g] - dvn[g]) + un[g] / cc;
            ptn[g] = (ct[g] < f) ? ct[g] : f;
        }
        return;
    }
    for (g = 0; g < N; ++g) {
        REAL acc = (parent[n] < 0) ? root_cc
                                   : (pt[(int64_t)parent[n] * N + g] - dvn[g]) + un[g] / cc;
        for (c = first_child[n]; c >= 0; c = next_sib[c]) {
            const int64_t i = (int64_t)c * N + g;
            acc = ((acc + pt[i]) + div[i]) - u[i] / cc;
        }
        ptn[g] = acc / (REAL)((parent[n] < 0) ? nc : nc + 1);
    }
}

static int SFX(hmf)(int64_t ny, int64_t nx, int64_t nz, int is3d, int nnodes,
                    const int32_t *parent, const REAL *Ct, const REAL *alpha,
                    const REAL *l_in, const int32_t *has_l, int num_it, REAL steps, double cc_d,
                    REAL *u, REAL *pt, REAL *div, REAL *f, REAL *l_out,
                    double *data_cost, double *regu_cost)
{
    const int64_t N = nx * ny * nz;
    const int64_t st[3] = {1, ny, nx * ny};
    const int nd = is3d ? 3 : 2;
    const REAL cc = (REAL)cc_d, root_cc = (REAL)(1.0 / cc_d);
    int32_t first_child[HMF_MAXN], next_sib[HMF_MAXN], last[HMF_MAXN], order[HMF_MAXN];
    int n, k, it, root = -1, norder = 0;
    int64_t g;
    REAL *G, *A, *EX[3], *L;

    if (N <= 0 || nnodes < 2 || nnodes > HMF_MAXN) return ORC_EBADARG;
    if (!is3d && nz != 1) return ORC_EBADARG;
    for (n = 0; n < nnodes; ++n) first_child[n] = next_sib[n] = last[n] = -1;
    for (n = 0; n < nnodes; ++n) {
        if (parent[n] < 0) {
            if (root >= 0) return ORC_EBADARG;
            root = n;
        } else {
            if (parent[n] >= nnodes || parent[n] == n) return ORC_EBADARG;
            if (last[parent[n]] < 0) first_child[parent[n]] = n;
            else next_sib[last[parent[n]]] = n;
            last[parent[n]] = n;
        }
    }
    if (root < 0 || first_child[root] < 0) return ORC_EBADARG;
    /* every node reachable from the root (no cycles): breadth-first order */
    order[norder++] = root;
    for (k = 0; k < norder; ++k)
        for (n = first_child[order[k]]; n >= 0; n = next_sib[n]) {
            if (norder >= nnodes) return ORC_EBADARG;
            order[norder++] = n;
        }
    if (norder != nnodes) return ORC_EBADARG;

    G = (REAL *)calloc((size_t)N, sizeof(REAL));
    A = (REAL *)calloc((size_t)N, sizeof(REAL));
    L = (REAL *)calloc((size_t)(nnodes * 3) * (size_t)N, sizeof(REAL));
    EX[0] = (REAL *)calloc((size_t)N, sizeof(REAL));
    EX[1] = (REAL *)calloc((size_t)N, sizeof(REAL));
    EX[2] = (REAL *)calloc((size_t)N, sizeof(REAL));
    if (!G || !A || !L || !EX[0] || !EX[1] || !EX[2]) {
        free(G); free(A); free(L); free(EX[0]); free(EX[1]); free(EX[2]);
        return ORC_ENOMEM;
    }
    /* InitializeFullFlow: zero state; shape directions normalised where
       |l| > 0.001, else 0 */
    memset(u, 0, (size_t)(nnodes * N) * sizeof(REAL));
    memset(pt, 0, (size_t)(nnodes * N) * sizeof(REAL));
    memset(div, 0, (size_t)(nnodes * N) * sizeof(REAL));
    memset(f, 0, (size_t)(nnodes * 3 * N) * sizeof(REAL));
    for (n = 0; n < nnodes; ++n) {
        if (!has_l || !has_l[n]) continue;
        for (g = 0; g < N; ++g) {
            REAL q[3] = {0, 0, 0}, s2, den;
            for (k = 0; k < nd; ++k) q[k] = l_in[((int64_t)n * 3 + k) * N + g];
            s2 = q[0] * q[0] + q[1] * q[1];
            if (is3d) s2 = s2 + q[2] * q[2];
            den = (REAL)sqrt((double)s2);
            for (k = 0; k < nd; ++k)
                L[((int64_t)n * 3 + k) * N + g] = (den > (REAL)0.001) ? q[k] / den : (REAL)0;
        }
    }

    for (it = 0; it < num_it; ++it) {
        SFX(hmf_sink)(root, parent, first_child, next_sib, N, cc, root_cc, Ct, u, pt, div);

        /* labels (top-down; order-free): u = u + cc*(P.pt - div - pt) */
        for (n = 0; n < nnodes; ++n) {
            if (n == root) continue;
            for (g = 0; g < N; ++g) {
                const int64_t i = (int64_t)n * N + g;
                u[i] = u[i] + cc * ((pt[(int64_t)parent[n] * N + g] - div[i]) - pt[i]);
            }
        }

        /* spatial flows of every non-root node */
        for (n = 0; n < nnodes; ++n) {
            REAL *fn[3], *un, *dvn;
            const REAL *ptn, *ppt, *al, *ln[3];
            int shape;
            int64_t x, y, z;
            if (n == root) continue;
            un = u + (int64_t)n * N;
            dvn = div + (int64_t)n * N;
            ptn = pt + (int64_t)n * N;
            ppt = pt + (int64_t)parent[n] * N;
            al = alpha + (int64_t)n * N;
            shape = has_l && has_l[n];
            for (k = 0; k < 3; ++k) {
                fn[k] = f + ((int64_t)n * 3 + k) * N;
                ln[k] = L + ((int64_t)n * 3 + k) * N;
            }
            /* g = steps*(div + pt - P.pt - u/cc) */
            for (g = 0; g < N; ++g) G[g] = steps * (((dvn[g] + ptn[g]) - ppt[g]) - un[g] / cc);
            /* faces with axis index >= 1 step by the difference of g across them:
               2D p = p + (g_hi - g_lo) (asetsHMF2D.m:225-226), 3D p = p + g_hi - g_lo
               (asetsHMF3D.m:210-212) */
            for (z = 0; z < nz; ++z)
                for (x = 0; x < nx; ++x)
                    for (y = 0; y < ny; ++y) {
                        const int64_t c3[3] = {y, x, z};
                        g = z * st[2] + x * st[1] + y;
                        for (k = 0; k < nd; ++k) {
                            if (c3[k] == 0) continue;
                            if (is3d) fn[k][g] = (fn[k][g] + G[g]) - G[g - st[k]];
                            else fn[k][g] = fn[k][g] + (G[g] - G[g - st[k]]);
                        }
                    }
            if (shape) {
                /* exemption amount per voxel (asetsHMF3D.m:217-223): dim-1 low face
                   if negative, dim-1 high face if positive, then dims 2 and 3 high
                   face if positive, low face if negative, each times l, clipped at 0 */
                for (z = 0; z < nz; ++z)
                    for (x = 0; x < nx; ++x)
                        for (y = 0; y < ny; ++y) {
                            const int64_t c3[3] = {y, x, z}, n3[3] = {ny, nx, nz};
                            REAL a = 0, t;
                            g = z * st[2] + x * st[1] + y;
                            for (k = 0; k < nd; ++k) {
                                const REAL lo = (c3[k] > 0) ? fn[k][g] : (REAL)0;
                                const REAL hi = (c3[k] + 1 < n3[k]) ? fn[k][g + st[k]] : (REAL)0;
                                const REAL lk = ln[k][g];
                                if (k == 0) {
                                    t = (lo < 0) ? lo * lk : (REAL)0;
                                    a = (t > 0) ? t : (REAL)0;
                                    t = (hi > 0) ? hi * lk : (REAL)0;
                                    a = a + ((t > 0) ? t : (REAL)0);
                                } else {
                                    t = (hi > 0) ? hi * lk : (REAL)0;
                                    a = a + ((t > 0) ? t : (REAL)0);
                                    t = (lo < 0) ? lo * lk : (REAL)0;
                                    a = a + ((t > 0) ? t : (REAL)0);
                                }
                            }
                            A[g] = a;
                        }
                /* exemption per face (:226-233) and its removal (:236-238) */
                for (k = 0; k < nd; ++k) memset(EX[k], 0, (size_t)N * sizeof(REAL));
                for (z = 0; z < nz; ++z)
                    for (x = 0; x < nx; ++x)
                        for (y = 0; y < ny; ++y) {
                            const int64_t c3[3] = {y, x, z};
                            g = z * st[2] + x * st[1] + y;
                            for (k = 0; k < nd; ++k) {
                                const int64_t w = g - st[k];   /* the low voxel of the face */
                                REAL p, e = 0;
                                if (c3[k] == 0) continue;
                                p = fn[k][g];
                                if (p > 0 && ln[k][w] > 0) e = ln[k][w] * A[w];
                                else if (p < 0 && ln[k][g] < 0) e = ln[k][g] * A[g];
                                EX[k][g] = e;
                                fn[k][g] = p - e;
                            }
                        }
            }
            /* flow magnitude and capacity factor (:241-258 / :266-282) */
            for (z = 0; z < nz; ++z)
                for (x = 0; x < nx; ++x)
                    for (y = 0; y < ny; ++y) {
                        const int64_t c3[3] = {y, x, z}, n3[3] = {ny, nx, nz};
                        REAL t = 0, m;
                        g = z * st[2] + x * st[1] + y;
                        for (k = 0; k < nd; ++k) {
                            const REAL lo = (c3[k] > 0) ? fn[k][g] : (REAL)0;
                            const REAL hi = (c3[k] + 1 < n3[k]) ? fn[k][g + st[k]] : (REAL)0;
                            if (k == 0) t = lo * lo + hi * hi;
                            else t = (t + hi * hi) + lo * lo;
                        }
                        m = (REAL)sqrt((double)t);
                        G[g] = (m <= al[g]) ? (REAL)1 : al[g] / m;
                    }
            /* rescale: p = [ex +] 0.5*p.*(g_hi + g_lo) (:259-261 / :283-285) */
            for (z = 0; z < nz; ++z)
                for (x = 0; x < nx; ++x)
                    for (y = 0; y < ny; ++y) {
                        const int64_t c3[3] = {y, x, z};
                        g = z * st[2] + x * st[1] + y;
                        for (k = 0; k < nd; ++k) {
                            REAL r;
                            if (c3[k] == 0) continue;
                            r = ((REAL)0.5 * fn[k][g]) * (G[g] + G[g - st[k]]);
                            fn[k][g] = shape ? EX[k][g] + r : r;
                        }
                    }
            /* divergence (:288-295) */
            for (z = 0; z < nz; ++z)
                for (x = 0; x < nx; ++x)
                    for (y = 0; y < ny; ++y) {
                        const int64_t c3[3] = {y, x, z}, n3[3] = {ny, nx, nz};
                        REAL d = 0;
                        g = z * st[2] + x * st[1] + y;
                        for (k = 0; k < nd; ++k) {
                            const REAL lo = (c3[k] > 0) ? fn[k][g] : (REAL)0;
                            const REAL hi = (c3[k] + 1 < n3[k]) ? fn[k][g + st[k]] : (REAL)0;
                            d = (k == 0) ? hi - lo : (d + hi) - lo;
                        }
                        dvn[g] = d;
                    }
        }
    }

    /* CalculateCostsFullFlow: own terms, then subtree totals in reverse
       breadth-first order (children before parents; children added in order) */
    for (n = 0; n < nnodes; ++n) {
        double dc = 0.0, rc = 0.0;
        const int64_t o = (int64_t)n * N;
        if (first_child[n] < 0)
            for (g = 0; g < N; ++g) dc += (double)(u[o + g] * Ct[o + g]);
        if (n != root)
            for (g = 0; g < N; ++g) rc += (double)(u[o + g] * div[o + g]);
        data_cost[n] = dc;
        regu_cost[n] = rc;
    }
    for (k = norder - 1; k >= 0; --k) {
        int c;
        n = order[k];
        for (c = first_child[n]; c >= 0; c = next_sib[c]) {
            data_cost[n] += data_cost[c];
            regu_cost[n] += regu_cost[c];
        }
    }
    if (l_out) memcpy(l_out, L, (size_t)(nnodes * 3 * N) * sizeof(REAL));
    free(G); free(A); free(L); free(EX[0]); free(EX[1]); free(EX[2]);
    return ORC_OK;
}
