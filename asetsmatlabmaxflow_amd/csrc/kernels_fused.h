// kernels_fused.h -- one kernel per iteration of the binary continuous max-flow
// (asetsBinaryMF3D_mex.c:163-190: updateP1, updatePX/PY/PZ, projStep,
// updateBX/BY/BZ, updateDIVPSPTU), fused by streaming along z.
//
// Dependency radius of one iteration at voxel v (L1 distance):
//   gk      = dv - ((ps - pt) + u/cc)           needed within radius 2 of v
//   b1      = steps*grad(gk) + b                faces within radius 1
//   g       = 1/max(|b1|/alpha, 1)              within radius 1
//   b2      = 0.5*(g_lo + g_hi)*b1              v's six faces
//   dv, ps, pt, u                               at v
// dv at the start of an iteration equals div(b) of the stored faces (it is
// computed from them at the end of the previous iteration with the same
// expression), so it is recomputed instead of stored: per voxel-iteration the
// kernel reads Cs, Ct, alpha and u, ps, pt, bx, by, bz and writes the latter
// six -- 60 bytes of compulsory HBM traffic in fp32.
//
// Work decomposition: a workgroup owns an in-plane tile of TY (=64, the
// contiguous y axis, one wavefront) by TX voxels and a range of z planes.  It
// streams through the planes keeping
//   GK[4][TX+4][TY+4]  gk of planes p..p+3 over the tile + 2-voxel halo
//   G [2][TX+2][TY+2]  g  of planes p, p+1 over the tile + 1-voxel halo
// in LDS.  Step p: (A) g of plane p+1; barrier; (B) output plane p and gk of
// plane p+3; barrier.  The z face below plane p is carried in registers from
// the step before.  Halo values are recomputed by the neighbouring workgroups
// (same inputs, same operation order, same bits).
//
// The state is double-buffered (in -> out) because neighbouring workgroups
// read the halo of the previous iterate.  Every face with axis index 0 is
// written as 0 and high faces past the last voxel read as 0: zero-flux
// boundary as in the reference.
#pragma once
#include "common.h"

namespace asets {

template <typename T>
struct FusedState {
    T *u, *ps, *pt, *bx, *by, *bz;
};

template <typename T>
struct FusedArgs {
    const float *Cs, *Ct, *alpha;
    FusedState<T> in, out;
    int ny, nx, nz;
    int64_t plane;          // ny*nx
    float cc, steps;
    int zchunk;             // planes per workgroup
    int ntile_y, ntile_x, nzc;
    int nblocks;            // ntile_y*ntile_x*nzc (logical)
    double *partials;       // [gridDim.x]
    const Flags *flags;
};

template <typename T> __device__ __forceinline__ float ld(const T *p);
template <> __device__ __forceinline__ float ld<float>(const float *p) { return *p; }
template <> __device__ __forceinline__ float ld<__half>(const __half *p) { return __half2float(*p); }
template <typename T> __device__ __forceinline__ void st(T *p, float v);
template <> __device__ __forceinline__ void st<float>(float *p, float v) { *p = v; }
template <> __device__ __forceinline__ void st<__half>(__half *p, float v) { *p = __float2half_rn(v); }

template <bool IS3D, int TX, typename T>
__global__ __launch_bounds__(256) void kf_binary(FusedArgs<T> a) {
    constexpr int TY = 64;
    constexpr int NT = 256;
    constexpr int W2 = TY + 4, H2 = TX + 4;   // gk region (radius 2)
    constexpr int W1 = TY + 2, H1 = TX + 2;   // g region (radius 1)
    constexpr int KOUT = (TY * TX) / NT;      // output voxels per thread per plane
    static_assert((TY * TX) % NT == 0, "tile must be a multiple of the block");

    __shared__ float GK[4][H2][W2];
    __shared__ float G[2][H1][W1];
    __shared__ double red[16];

    if (a.flags->done) return;

    // XCD-aware block order: consecutive logical tiles land on the same XCD
    // (hardware dispatches block b to XCD b % 8) so in-plane halos are shared
    // through one L2.
    const int nb = gridDim.x;
    int lb = blockIdx.x;
    if ((nb & 7) == 0) lb = (blockIdx.x & 7) * (nb >> 3) + (blockIdx.x >> 3);
    const int tid = threadIdx.x;
    double e_acc = 0.0;

    if (lb < a.nblocks) {
        const int tyi = lb % a.ntile_y;
        const int txi = (lb / a.ntile_y) % a.ntile_x;
        const int zci = lb / (a.ntile_y * a.ntile_x);
        const int y0 = tyi * TY, x0 = txi * TX;
        const int z0 = zci * a.zchunk;
        const int z1 = min(z0 + a.zchunk, a.nz);
        const int ny = a.ny, nx = a.nx, nz = a.nz;
        const int64_t plane = a.plane;
        const float cc = a.cc, steps = a.steps;

        // ---- stage S1: gk of plane q over the radius-2 region ------------
        auto s1 = [&](int q) {
            float(*gk)[W2] = GK[(q + 4) & 3];
            const bool qin = (q >= 0) && (q < nz);
            const int64_t base = (int64_t)q * plane;
            for (int idx = tid; idx < W2 * H2; idx += NT) {
                const int i = idx % W2, j = idx / W2;
                const int y = y0 - 2 + i, x = x0 - 2 + j;
                float v = 0.0f;
                if (qin && y >= 0 && y < ny && x >= 0 && x < nx) {
                    const int64_t o = base + (int64_t)x * ny + y;
                    const float bxl = ld(a.in.bx + o);
                    const float bxr = (x + 1 < nx) ? ld(a.in.bx + o + ny) : 0.0f;
                    const float byl = ld(a.in.by + o);
                    const float byr = (y + 1 < ny) ? ld(a.in.by + o + 1) : 0.0f;
                    float d = ((byr - byl) + bxr) - bxl;
                    if (IS3D) {
                        const float bzl = ld(a.in.bz + o);
                        const float bzr = (q + 1 < nz) ? ld(a.in.bz + o + plane) : 0.0f;
                        d = (d + bzr) - bzl;
                    }
                    v = d - ((ld(a.in.ps + o) - ld(a.in.pt + o)) + ld(a.in.u + o) / cc);
                }
                gk[j][i] = v;
            }
        };

        // ---- stage S2: projection factor g of plane q over radius 1 -------
        auto s2 = [&](int q) {
            float(*gq)[W2] = GK[(q + 4) & 3];
            float(*gm)[W2] = GK[(q + 3) & 3];
            float(*gp)[W2] = GK[(q + 5) & 3];
            float(*go)[W1] = G[q & 1];
            const bool qin = (q >= 0) && (q < nz);
            const int64_t base = (int64_t)q * plane;
            for (int idx = tid; idx < W1 * H1; idx += NT) {
                const int i1 = idx % W1, j1 = idx / W1;
                const int y = y0 - 1 + i1, x = x0 - 1 + j1;
                float v = 1.0f;
                if (qin && y >= 0 && y < ny && x >= 0 && x < nx) {
                    const int i = i1 + 1, j = j1 + 1;
                    const int64_t o = base + (int64_t)x * ny + y;
                    const float c = gq[j][i];
                    const float bxl = (x > 0) ? steps * (c - gq[j - 1][i]) + ld(a.in.bx + o) : 0.0f;
                    const float bxr = (x + 1 < nx) ? steps * (gq[j + 1][i] - c) + ld(a.in.bx + o + ny) : 0.0f;
                    const float byl = (y > 0) ? steps * (c - gq[j][i - 1]) + ld(a.in.by + o) : 0.0f;
                    const float byr = (y + 1 < ny) ? steps * (gq[j][i + 1] - c) + ld(a.in.by + o + 1) : 0.0f;
                    float bzl = 0.0f, bzr = 0.0f;
                    if (IS3D) {
                        bzl = (q > 0) ? steps * (c - gm[j][i]) + ld(a.in.bz + o) : 0.0f;
                        bzr = (q + 1 < nz) ? steps * (gp[j][i] - c) + ld(a.in.bz + o + plane) : 0.0f;
                    }
                    v = proj_factor<IS3D>(bxr, bxl, byr, byl, bzr, bzl, a.alpha[o]);
                }
                go[j1][i1] = v;
            }
        };

        float carry[KOUT];   // rescaled z face below plane p, per output voxel

        // ---- prologue: gk of z0-2..z0+1, g of z0-1 and z0, carry ----------
        s1(z0 - 2); s1(z0 - 1); s1(z0); s1(z0 + 1);
        __syncthreads();
        s2(z0 - 1);
        __syncthreads();
        s2(z0);
        __syncthreads();
        {
            float(*gq)[W2] = GK[(z0 + 4) & 3];
            float(*gm)[W2] = GK[(z0 + 3) & 3];
            float(*g0)[W1] = G[z0 & 1];
            float(*gmm)[W1] = G[(z0 + 1) & 1];   // plane z0-1
            const int64_t base = (int64_t)z0 * plane;
#pragma unroll
            for (int k = 0; k < KOUT; ++k) {
                const int idx = tid + k * NT;
                const int i = idx % TY, j = idx / TY;
                const int y = y0 + i, x = x0 + j;
                float v = 0.0f;
                if (IS3D && z0 > 0 && y < ny && x < nx) {
                    const int64_t o = base + (int64_t)x * ny + y;
                    const float b1 = steps * (gq[j + 2][i + 2] - gm[j + 2][i + 2]) + ld(a.in.bz + o);
                    v = rescale_face(g0[j + 1][i + 1], gmm[j + 1][i + 1], b1);
                }
                carry[k] = v;
            }
        }
        s1(z0 + 2);
        __syncthreads();

        // ---- main loop ----------------------------------------------------
        for (int p = z0; p < z1; ++p) {
            s2(p + 1);
            __syncthreads();
            {
                float(*gq)[W2] = GK[(p + 4) & 3];
                float(*gn)[W2] = GK[(p + 5) & 3];
                float(*g0)[W1] = G[p & 1];
                float(*g1)[W1] = G[(p + 1) & 1];
                const int64_t base = (int64_t)p * plane;
#pragma unroll
                for (int k = 0; k < KOUT; ++k) {
                    const int idx = tid + k * NT;
                    const int i = idx % TY, j = idx / TY;
                    const int y = y0 + i, x = x0 + j;
                    if (y < ny && x < nx) {
                        const int64_t o = base + (int64_t)x * ny + y;
                        const int I = i + 2, J = j + 2;      // gk coordinates
                        const int I1 = i + 1, J1 = j + 1;    // g coordinates
                        const float c = gq[J][I];
                        const float gc = g0[J1][I1];
                        // x faces
                        float bxl = 0.0f, bxr = 0.0f;
                        if (x > 0) {
                            const float b1 = steps * (c - gq[J - 1][I]) + ld(a.in.bx + o);
                            bxl = rescale_face(gc, g0[J1 - 1][I1], b1);
                        }
                        if (x + 1 < nx) {
                            const float b1 = steps * (gq[J + 1][I] - c) + ld(a.in.bx + o + ny);
                            bxr = rescale_face(g0[J1 + 1][I1], gc, b1);
                        }
                        // y faces
                        float byl = 0.0f, byr = 0.0f;
                        if (y > 0) {
                            const float b1 = steps * (c - gq[J][I - 1]) + ld(a.in.by + o);
                            byl = rescale_face(gc, g0[J1][I1 - 1], b1);
                        }
                        if (y + 1 < ny) {
                            const float b1 = steps * (gq[J][I + 1] - c) + ld(a.in.by + o + 1);
                            byr = rescale_face(g0[J1][I1 + 1], gc, b1);
                        }
                        float d = ((byr - byl) + bxr) - bxl;
                        float bzl = 0.0f, bzr = 0.0f;
                        if (IS3D) {
                            bzl = carry[k];
                            if (p + 1 < nz) {
                                const float b1 = steps * (gn[J][I] - c) + ld(a.in.bz + o + plane);
                                bzr = rescale_face(g1[J1][I1], gc, b1);
                            }
                            d = (d + bzr) - bzl;
                            carry[k] = bzr;
                        }
                        // source/sink flows and multiplier (asetsBinaryMF3D_mex.c:419-432)
                        const float uo = ld(a.in.u + o);
                        const float uc = uo / cc;
                        float f = ((ld(a.in.pt + o) - uc) + d) + 1.0f / cc;
                        const float nps = fmin_ref(f, a.Cs[o]);
                        f = (nps + uc) - d;
                        const float npt = fmin_ref(f, a.Ct[o]);
                        const float e = cc * ((npt + d) - nps);
                        e_acc += fabs((double)e);
                        st(a.out.u + o, uo - e);
                        st(a.out.ps + o, nps);
                        st(a.out.pt + o, npt);
                        st(a.out.bx + o, bxl);
                        st(a.out.by + o, byl);
                        if (IS3D) st(a.out.bz + o, bzl);
                    }
                }
            }
            s1(p + 3);
            __syncthreads();
        }
    }
    const double s = block_sum(e_acc, red);
    if (tid == 0) a.partials[blockIdx.x] = s;
}

}  // namespace asets
