// kernels_bounds.h -- FP64 primal-dual bounds of the binary continuous max-flow
// model for a state (u, bx, by, bz).  No reference counterpart: the reference
// stops on the mean |du| trace (asetsBinaryMF3D_mex.c:183-188) and its only energy
// terms are HMF's (asetsHMF2D.m:184-203).
//
// Model: max sum ps s.t. ps <= Cs, pt <= Ct, div p - ps + pt = 0 and the per-voxel
// capacity 0.5 * (sum of the voxel's six faces^2) <= alpha^2 (the norm of projStep,
// asetsBinaryMF3D_mex.c:321-331).
//   dual   D(p~) = sum min(Cs, Ct + div p~), p~ = each face scaled by the smaller
//          feasibility factor s = min(1, alpha / n) of its two voxels;
//   primal P(u) = sum (1-u) Cs + u Ct + F(lambda), u clipped to [0, 1], w the face
//          differences of u, and F(lambda) = sum_v lambda_v alpha_v^2 +
//          sum_f w_f^2 / (2 (lambda_v1 + lambda_v2)) -- an upper bound of the
//          regulariser for any lambda >= 0 (Lagrange duality over the capacity
//          constraints), lowered by red-black coordinate descent (a per-voxel Newton
//          solve with the neighbours fixed).  The starting lambda, the equal split of
//          every face, gives the voxel-centred TV bound.
// D(p~) <= min P <= P(u).  Faces use the solver's own-face convention (face k of
// voxel v lies between v - e_k and v; 0 on the low boundary); sums over the six
// faces run k = x, y, z, own face then the +axis neighbour's, with the same
// guards as the CPU restatement, so that each voxel's value matches it to
// rounding.  Volume sums use a fixed grid and a fixed-order final sum
// (deterministic).
#pragma once
#include "common.h"

namespace asets {

constexpr int kBoundsBlocks = 1024;
constexpr int kBoundsThreads = 256;
constexpr int kBoundsNewton = 40;   // Newton steps per voxel and sweep

struct BGeo {
    int64_t ny, nx, nz, sx, sz, n;   // extents, strides of x and z, voxels
};

// the voxel's (x, y, z) and its neighbour offsets along k = x, y, z
struct BVox {
    int64_t c[3], ext[3], st[3];
};
__device__ __forceinline__ BVox bvox(const BGeo &G, int64_t g) {
    BVox v;
    v.c[1] = g % G.ny;
    v.c[0] = (g / G.ny) % G.nx;
    v.c[2] = g / G.sz;
    v.ext[0] = G.nx;
    v.ext[1] = G.ny;
    v.ext[2] = G.nz;
    v.st[0] = G.sx;
    v.st[1] = 1;
    v.st[2] = G.sz;
    return v;
}

__device__ __forceinline__ double face(const float *f, int64_t g) {
    return f ? (double)f[g] : 0.0;
}

// s_v = n_v > alpha ? alpha / n_v : 1, n_v = sqrt(0.5 * sum of the six faces^2)
__global__ void k_bnd_scale(BGeo G, const float *alpha, const float *bx, const float *by,
                            const float *bz, double *s) {
    const float *f[3] = {bx, by, bz};
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < G.n;
         g += (int64_t)gridDim.x * blockDim.x) {
        const BVox v = bvox(G, g);
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double a = face(f[k], g);
            const double b = v.c[k] + 1 < v.ext[k] ? face(f[k], g + v.st[k]) : 0.0;
            t += a * a + b * b;
        }
        const double n = __builtin_sqrt(0.5 * t);
        const double al = alpha[g];
        s[g] = n > al ? al / n : 1.0;
    }
}

// dual terms min(Cs, Ct + div p~), summed per block
__global__ __launch_bounds__(kBoundsThreads) void k_bnd_dual(BGeo G, const float *Cs,
                                                             const float *Ct, const float *bx,
                                                             const float *by, const float *bz,
                                                             const double *s, double *part) {
    __shared__ double red[16];
    const float *f[3] = {bx, by, bz};
    double acc = 0.0;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < G.n;
         g += (int64_t)gridDim.x * blockDim.x) {
        const BVox v = bvox(G, g);
        const double sv = s[g];
        double dv = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double own = face(f[k], g) * fmin(sv, v.c[k] > 0 ? s[g - v.st[k]] : 0.0);
            const double up = v.c[k] + 1 < v.ext[k]
                                  ? face(f[k], g + v.st[k]) * fmin(s[g + v.st[k]], sv)
                                  : 0.0;
            dv += up - own;
        }
        acc += fmin((double)Cs[g], (double)Ct[g] + dv);
    }
    const double b = block_sum(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = b;
}

// u clipped to [0, 1]
__device__ __forceinline__ double uclip(const float *u, int64_t g) {
    const double x = u[g];
    return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x);
}
// w_k(v) = u(v) - u(v - e_k), 0 on the low boundary; lo/hi outside the volume: 0
__device__ __forceinline__ double wface(const BGeo &G, const BVox &v, const float *u, int64_t g,
                                        int k) {
    return v.c[k] > 0 ? uclip(u, g) - uclip(u, g - v.st[k]) : 0.0;
}

// lambda at the voxel-centred split, lam = |w|_v / (2 sqrt(2) alpha), and the data
// term (1 - u) Cs + u Ct, summed per block
__global__ __launch_bounds__(kBoundsThreads) void k_bnd_init(BGeo G, const float *Cs,
                                                             const float *Ct, const float *alpha,
                                                             const float *u, double *lam,
                                                             double *part) {
    __shared__ double red[16];
    double acc = 0.0;
    const double c22 = 2.0 * __builtin_sqrt(2.0);
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < G.n;
         g += (int64_t)gridDim.x * blockDim.x) {
        const BVox v = bvox(G, g);
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double a = wface(G, v, u, g, k);
            double b = 0.0;
            if (v.c[k] + 1 < v.ext[k]) {
                BVox vn = v;
                vn.c[k] += 1;
                b = wface(G, vn, u, g + v.st[k], k);
            }
            t += a * a + b * b;
        }
        lam[g] = __builtin_sqrt(t) / (c22 * (double)alpha[g]);
        const double ub = uclip(u, g);
        acc += (1.0 - ub) * (double)Cs[g] + ub * (double)Ct[g];
    }
    const double b = block_sum(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = b;
}

// One colour of a red-black sweep: on the voxels with (x + y + z) % 2 == color,
// lambda_v = argmin F over lambda_v >= 0, the neighbours (all of the other colour)
// fixed.  F's derivative in lambda_v is alpha^2 - sum_m C_m / (lambda_v + M_m)^2
// over the six faces m (C_m = w_m^2 / 2, M_m the neighbour's lambda across face m);
// where it is >= 0 at 0 the minimum is lambda_v = 0, else Newton from a point below
// the root (safeguarded: a step that leaves (0, inf) quarters lambda instead).
__global__ void k_bnd_sweep(BGeo G, const float *alpha, const float *u, double *lam,
                            int color) {
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < G.n;
         g += (int64_t)gridDim.x * blockDim.x) {
        const BVox v = bvox(G, g);
        if ((int)((v.c[0] + v.c[1] + v.c[2]) & 1) != color) continue;
        double C[6], M[6];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double a = wface(G, v, u, g, k);
            C[2 * k] = 0.5 * a * a;
            M[2 * k] = v.c[k] > 0 ? lam[g - v.st[k]] : 0.0;
            double b = 0.0, mb = 0.0;
            if (v.c[k] + 1 < v.ext[k]) {
                BVox vn = v;
                vn.c[k] += 1;
                b = wface(G, vn, u, g + v.st[k], k);
                mb = lam[g + v.st[k]];
            }
            C[2 * k + 1] = 0.5 * b * b;
            M[2 * k + 1] = mb;
        }
        const double al = alpha[g], a2 = al * al;
        double csum = C[0];
        for (int m = 1; m < 6; ++m) csum += C[m];
        double q = 0.0;
        double mmin = __builtin_inf();
        for (int m = 0; m < 6; ++m) {
            if (C[m] > 0.0) {
                const double mm = M[m] > 0.0 ? M[m] : 0.0;
                q += C[m] / (mm * mm);
                mmin = fmin(mmin, M[m]);
            }
        }
        double d0 = a2 - q;
        if (d0 != d0) d0 = -__builtin_inf();
        const bool active = csum > 0.0 && d0 < 0.0;
        double x = active ? __builtin_sqrt(csum) / al - (mmin < __builtin_inf() ? mmin : 0.0)
                          : 1.0;
        if (!(x > 0.0)) x = 1e-3;
        for (int it = 0; it < kBoundsNewton; ++it) {
            double gs = 0.0, hs = 0.0;
            for (int m = 0; m < 6; ++m) {
                double den = x + M[m];
                den = den > 0.0 ? den : 1.0;
                const double t = C[m] > 0.0 ? C[m] / (den * den) : 0.0;
                gs += t;
                hs += C[m] > 0.0 ? t / den : 0.0;
            }
            const double gr = a2 - gs, h = 2.0 * hs;
            const double xn = x - gr / (h > 0.0 ? h : 1.0);
            x = xn > 0.0 ? xn : x / 4.0;
        }
        lam[g] = active ? x : 0.0;
    }
}

// F(lambda) terms: lambda_v alpha_v^2 + sum over v's own faces of
// (w^2 / 2) / (lambda_v + lambda_{v - e_k}), summed per block
__global__ __launch_bounds__(kBoundsThreads) void k_bnd_F(BGeo G, const float *alpha,
                                                          const float *u, const double *lam,
                                                          double *part) {
    __shared__ double red[16];
    double acc = 0.0;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < G.n;
         g += (int64_t)gridDim.x * blockDim.x) {
        const BVox v = bvox(G, g);
        const double al = alpha[g];
        const double lv = lam[g];
        double t = al * al * lv;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double a = wface(G, v, u, g, k);
            const double c = 0.5 * a * a;
            const double L = lv + (v.c[k] > 0 ? lam[g - v.st[k]] : 0.0);
            t += c > 0.0 ? c / (L > 0.0 ? L : 1.0) : 0.0;
        }
        acc += t;
    }
    const double b = block_sum(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = b;
}

// out[i] = fixed-order sum of part[i * nb .. i * nb + nb - 1], i < nsum
__global__ void k_bnd_sum(const double *part, int nb, int nsum, double *out) {
    __shared__ double red[16];
    for (int i = 0; i < nsum; ++i) {
        double a = 0.0;
        for (int b = threadIdx.x; b < nb; b += blockDim.x) a += part[(size_t)i * nb + b];
        const double s = block_sum(a, red);
        if (threadIdx.x == 0) out[i] = s;
        __syncthreads();
    }
}

}  // namespace asets
