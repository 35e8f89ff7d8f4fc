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
 *               asetsPotts2D_mex.c, asetsPotts3D_mex.c).  Built with
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
