/*
 * asets_oracle.c -- CPU oracle for the ASETS continuous max-flow solvers.
 *
 * TEST INFRASTRUCTURE ONLY: this library is the checker.  It may be loaded by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py;
 * the product (libasets_hip.so and the mex-ABI host files) never links it.
 *
 * Two instantiations of oracle_impl.h:
 *   orc_*_f32 : single precision in the reference's operation order
 *               (maxflow/asetsBinaryMF2D_mex.c, asetsBinaryMF3D_mex.c,
 *               asetsPotts2D_mex.c, asetsPotts3D_mex.c, asetsIshikawa2D_mex.c,
 *               asetsIshikawa3D_mex.c).  Built with
 *               -ffp-contract=off so no multiply-add is fused, as in the
 *               reference's x86-64 SSE build.
 *   orc_*_f64 : the same algorithm in double precision.
 *
 * Both report two convergence traces per iteration: cvg_ref, the reference's
 * sequential running sum of |e| in the working precision divided by N
 * (asetsBinaryMF3D_mex.c:183, 401, 430), and cvg64, an FP64 sum divided by N.
 * stop_rule 0 stops on cvg_ref <= errBound (the reference's rule, :187-188);
 * stop_rule 1 stops on (float)cvg64 <= errBound (the GPU library's rule).
 *
 * Parity pins (tests/test_oracle_*.py): closed-form init, the one-voxel scalar
 * fixed point, spatially uniform costs, an independent FP64 sparse-matrix
 * restatement of one iteration, the Nz=1 == 2D identity, the Potts
 * sum-to-one invariant, label-permutation and cost-shift equivariance,
 * convergence of FP32 to the FP64 run.  The reference C itself cannot be built
 * in this image (it needs MATLAB's mex.h), so no bitwise golden trajectory of
 * the reference exists to compare against.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_OK 0
#define ORC_EBADARG 1
#define ORC_EALPHA 2
#define ORC_ENOMEM 3

#define ORC_MIN(a, b) (((a) < (b)) ? (a) : (b))        /* asetsBinaryMF3D_mex.c:17 */
#define ORC_ABS(x) (((x) > 0.0) ? (x) : -(x))           /* asetsBinaryMF3D_mex.c:19 */

#define REAL float
#define SFX(name) name##_f32
#include "oracle_impl.h"
#undef REAL
#undef SFX

#define REAL double
#define SFX(name) name##_f64
#include "oracle_impl.h"
#undef REAL
#undef SFX

#define EXPORT __attribute__((visibility("default")))

EXPORT int orc_binary_f32(const float *Cs, const float *Ct, const float *alpha,
                          int64_t ny, int64_t nx, int64_t nz, int is3d,
                          int max_it, float err_bound, float cc, float steps,
                          const float *u0, const float *pt0, int stop_rule,
                          float *u, float *cvg_ref, double *cvg64, int32_t *it_num,
                          float *ps, float *pt, float *bx, float *by, float *bz)
{
    return binary_f32(Cs, Ct, alpha, ny, nx, nz, is3d, max_it, err_bound, cc, steps,
                      u0, pt0, stop_rule, u, cvg_ref, cvg64, it_num, ps, pt, bx, by, bz);
}

EXPORT int orc_binary_f64(const double *Cs, const double *Ct, const double *alpha,
                          int64_t ny, int64_t nx, int64_t nz, int is3d,
                          int max_it, double err_bound, double cc, double steps,
                          const double *u0, const double *pt0, int stop_rule,
                          double *u, float *cvg_ref, double *cvg64, int32_t *it_num,
                          double *ps, double *pt, double *bx, double *by, double *bz)
{
    return binary_f64(Cs, Ct, alpha, ny, nx, nz, is3d, max_it, err_bound, cc, steps,
                      u0, pt0, stop_rule, u, cvg_ref, cvg64, it_num, ps, pt, bx, by, bz);
}

EXPORT int orc_potts_f32(const float *Ct, const float *alpha,
                         int64_t ny, int64_t nx, int64_t nz, int64_t L, int is3d,
                         int max_it, float err_bound, float cc, float steps, int stop_rule,
                         float *u, float *cvg_ref, double *cvg64, int32_t *it_num,
                         float *ps, float *pt, float *bx, float *by, float *bz)
{
    return potts_f32(Ct, alpha, ny, nx, nz, L, is3d, max_it, err_bound, cc, steps,
                     stop_rule, u, cvg_ref, cvg64, it_num, ps, pt, bx, by, bz);
}

EXPORT int orc_potts_f64(const double *Ct, const double *alpha,
                         int64_t ny, int64_t nx, int64_t nz, int64_t L, int is3d,
                         int max_it, double err_bound, double cc, double steps, int stop_rule,
                         double *u, float *cvg_ref, double *cvg64, int32_t *it_num,
                         double *ps, double *pt, double *bx, double *by, double *bz)
{
    return potts_f64(Ct, alpha, ny, nx, nz, L, is3d, max_it, err_bound, cc, steps,
                     stop_rule, u, cvg_ref, cvg64, it_num, ps, pt, bx, by, bz);
}

EXPORT int orc_ishikawa_f32(const float *Ct, const float *alpha,
                            int64_t ny, int64_t nx, int64_t nz, int64_t L, int is3d,
                            int max_it, float err_bound, float cc, float steps, int stop_rule,
                            float *u, float *cvg_ref, double *cvg64, int32_t *it_num,
                            float *pt, float *bx, float *by, float *bz)
{
    return ishikawa_f32(Ct, alpha, ny, nx, nz, L, is3d, max_it, err_bound, cc, steps,
                        stop_rule, u, cvg_ref, cvg64, it_num, pt, bx, by, bz);
}

EXPORT int orc_ishikawa_f64(const double *Ct, const double *alpha,
                            int64_t ny, int64_t nx, int64_t nz, int64_t L, int is3d,
                            int max_it, double err_bound, double cc, double steps, int stop_rule,
                            double *u, float *cvg_ref, double *cvg64, int32_t *it_num,
                            double *pt, double *bx, double *by, double *bz)
{
    return ishikawa_f64(Ct, alpha, ny, nx, nz, L, is3d, max_it, err_bound, cc, steps,
                        stop_rule, u, cvg_ref, cvg64, it_num, pt, bx, by, bz);
}

/* ---- energy evaluator (FP64) ------------------------------------------------
 * No reference counterpart: the C-MEX solvers report only the mean |du| trace.
 * The primal energy of the binary model,
 *     E(u) = sum_g (1 - u_g) Cs_g + u_g Ct_g  +  sum_g alpha_g |grad u|_g,
 * with the voxel-centred norm built from the same face differences as the
 * projection of asetsBinaryMF3D_mex.c:321-324 (the high and the low face of each
 * axis, squared, summed, times 0.5, square root; a face outside the volume
 * contributes 0).  An axis-aligned unit edge therefore costs sqrt(2)*alpha, which
 * is the capacity the projection gives a single face (|q| <= sqrt(2) alpha).
 * Potts: E = sum_l sum_g u_lg Ct_lg + alpha_lg |grad u_l|_g.
 * out[0] = E, out[1] = data term, out[2] = regularisation term.  is3d == 0 uses
 * the 4-face 2D norm (nz must be 1).
 */
static void energy_field(const double *Cs, const double *Ct, const double *alpha,
                         const double *u, int64_t ny, int64_t nx, int64_t nz, int is3d,
                         double *data, double *reg)
{
    const int64_t sx = ny, sz = nx * ny;
    double d = 0.0, r = 0.0;
    for (int64_t z = 0; z < nz; ++z)
        for (int64_t x = 0; x < nx; ++x)
            for (int64_t y = 0; y < ny; ++y) {
                const int64_t g = z * sz + x * sx + y;
                const double ug = u[g];
                d += (Cs ? (1.0 - ug) * Cs[g] : 0.0) + ug * Ct[g];
                double t = 0.0, q;
                if (x > 0)      { q = ug - u[g - sx]; t += q * q; }
                if (x < nx - 1) { q = u[g + sx] - ug; t += q * q; }
                if (y > 0)      { q = ug - u[g - 1];  t += q * q; }
                if (y < ny - 1) { q = u[g + 1] - ug;  t += q * q; }
                if (is3d) {
                    if (z > 0)      { q = ug - u[g - sz]; t += q * q; }
                    if (z < nz - 1) { q = u[g + sz] - ug; t += q * q; }
                }
                r += alpha[g] * sqrt(0.5 * t);
            }
    *data += d;
    *reg += r;
}

EXPORT int orc_energy_binary(const double *Cs, const double *Ct, const double *alpha,
                             int64_t ny, int64_t nx, int64_t nz, int is3d, const double *u,
                             double *out)
{
    if (ny <= 0 || nx <= 0 || nz <= 0 || (!is3d && nz != 1)) return ORC_EBADARG;
    double d = 0.0, r = 0.0;
    energy_field(Cs, Ct, alpha, u, ny, nx, nz, is3d, &d, &r);
    out[0] = d + r;
    out[1] = d;
    out[2] = r;
    return ORC_OK;
}

EXPORT int orc_energy_potts(const double *Ct, const double *alpha, int64_t ny, int64_t nx,
                            int64_t nz, int64_t L, int is3d, const double *u, double *out)
{
    if (ny <= 0 || nx <= 0 || nz <= 0 || L <= 0 || (!is3d && nz != 1)) return ORC_EBADARG;
    const int64_t N = ny * nx * nz;
    double d = 0.0, r = 0.0;
    for (int64_t l = 0; l < L; ++l)
        energy_field(NULL, Ct + l * N, alpha + l * N, u + l * N, ny, nx, nz, is3d, &d, &r);
    out[0] = d + r;
    out[1] = d;
    out[2] = r;
    return ORC_OK;
}

/* ---- hierarchical max-flow, full flow (oracle_impl.h: hmf) ---------------- */
EXPORT int orc_hmf_f32(int64_t ny, int64_t nx, int64_t nz, int is3d, int nnodes,
                       const int32_t *parent, const float *Ct, const float *alpha,
                       const float *l_in, const int32_t *has_l, int num_it, float steps,
                       double cc, float *u, float *pt, float *div, float *f, float *l_out,
                       double *data_cost, double *regu_cost)
{
    return hmf_f32(ny, nx, nz, is3d, nnodes, parent, Ct, alpha, l_in, has_l, num_it, steps,
                   cc, u, pt, div, f, l_out, data_cost, regu_cost);
}

EXPORT int orc_hmf_f64(int64_t ny, int64_t nx, int64_t nz, int is3d, int nnodes,
                       const int32_t *parent, const double *Ct, const double *alpha,
                       const double *l_in, const int32_t *has_l, int num_it, double steps,
                       double cc, double *u, double *pt, double *div, double *f, double *l_out,
                       double *data_cost, double *regu_cost)
{
    return hmf_f64(ny, nx, nz, is3d, nnodes, parent, Ct, alpha, l_in, has_l, num_it, steps,
                   cc, u, pt, div, f, l_out, data_cost, regu_cost);
}
