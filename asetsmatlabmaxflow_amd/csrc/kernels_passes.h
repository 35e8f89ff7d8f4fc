// kernels_passes.h -- one kernel per pass of the reference iteration.
//
// The reference (maxflow/asetsBinaryMF3D_mex.c:163-190, asetsPotts3D_mex.c:
// 170-204) runs 9 full-volume passes per iteration (9L+1 for Potts).  Here the
// face-step passes are one kernel, the rescale passes one kernel, and all
// labels of a Potts model run in the same launch (labels are independent
// inside an iteration; each has its own gk plane).  One thread per element.
// This family serves every model and shape; the z-streaming fused kernel in
// kernels_fused.h is the fast path for the binary model.
//
// Face storage: each voxel owns its low face on every axis (bx[g] between x-1
// and x).  Index-0 faces stay 0 and the high face of the last voxel reads 0:
// the reference's zero-flux boundary.
#pragma once
#include "common.h"

namespace asets {

struct Geo {
    int64_t ny, nx, nz;  // volume extent (Ny fastest)
    int64_t N;           // ny*nx*nz
    int64_t L;           // labels (1 for the binary model)
};

struct Coord {
    int64_t g;  // voxel index within the label volume
    int64_t y, x, z;
};

__device__ __forceinline__ Coord decode(int64_t gl, const Geo &G) {
    Coord c;
    c.g = gl % G.N;
    c.y = c.g % G.ny;
    int64_t t = c.g / G.ny;
    c.x = t % G.nx;
    c.z = t / G.nx;
    return c;
}

__device__ __forceinline__ int64_t gid() {
    return (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
}

__global__ void k_fill_f32(float *p, float v, int64_t n) {
    int64_t i = gid();
    if (i < n) p[i] = v;
}

// Start of a solve in one launch: nflag flag words to 0, *it_num = max_iter and
// cvg[0, n) = 0.
__global__ void k_solve_setup(int32_t *flags, int nflag, int32_t *it_num, int32_t max_iter,
                              float *cvg, int64_t n) {
    const int64_t i = gid();
    if (i < nflag) flags[i] = 0;
    if (i == 0) *it_num = max_iter;
    if (i < n) cvg[i] = 0.0f;
}

// asetsBinaryMF3D_mex.c:205-224; warm start (u0, pt0 and optionally ps0) has no
// reference counterpart.
__global__ void k_init_binary(const float *Cs, const float *Ct, const float *u0,
                              const float *ps0, const float *pt0, float *u, float *ps,
                              float *pt, int64_t N) {
    int64_t g = gid();
    if (g >= N) return;
    if (u0) {
        u[g] = u0[g];
        ps[g] = ps0 ? ps0[g] : pt0[g];
        pt[g] = pt0[g];
    } else {
        float cs = Cs[g], ct = Ct[g];
        u[g] = (cs - ct >= 0.0f) ? 1.0f : 0.0f;
        float m = fmin_ref(cs, ct);
        ps[g] = m;
        pt[g] = m;
    }
}

// asetsPotts3D_mex.c:219-258: first strict arg-min of Ct from 1e30.
__global__ void k_init_potts(const float *Ct, float *u, float *ps, float *pt, int64_t N,
                             int64_t L) {
    int64_t g = gid();
    if (g >= N) return;
    float mv = 1e30f;
    int64_t mi = -1;
    for (int64_t l = 0; l < L; ++l) {
        float c = Ct[g + l * N];
        if (mv > c) { mv = c; mi = l; }
    }
    ps[g] = mv;
    for (int64_t l = 0; l < L; ++l) {
        if (pt) pt[g + l * N] = mv;   // null: the caller sets pt itself
        u[g + l * N] = (l == mi) ? 1.0f : 0.0f;
    }
}

// gk = dv - ((ps - pt) + u/cc)   (asetsBinaryMF3D_mex.c:238-239, Potts :273-274)
__global__ void kp_gk(const float *dv, const float *ps, const float *pt, const float *u,
                      float *gk, float cc, Geo G, const Flags *f) {
    if (f->done) return;
    int64_t gl = gid();
    if (gl >= G.N * G.L) return;
    int64_t g = gl % G.N;
    gk[gl] = dv[gl] - ((ps[g] - pt[gl]) + u[gl] / cc);
}

// b = steps*(gk[g] - gk[g-s]) + b on interior faces (:258, :278, :297)
template <bool IS3D>
__global__ void kp_step(const float *gk, float *bx, float *by, float *bz, float steps, Geo G,
                        const Flags *f) {
    if (f->done) return;
    int64_t gl = gid();
    if (gl >= G.N * G.L) return;
    Coord c = decode(gl, G);
    const float gc = gk[gl];
    if (c.x > 0) bx[gl] = steps * (gc - gk[gl - G.ny]) + bx[gl];
    if (c.y > 0) by[gl] = steps * (gc - gk[gl - 1]) + by[gl];
    if (IS3D && c.z > 0) bz[gl] = steps * (gc - gk[gl - G.ny * G.nx]) + bz[gl];
}

// projection factor into gk (:317-331)
template <bool IS3D>
__global__ void kp_proj(const float *bx, const float *by, const float *bz, const float *alpha,
                        float *gk, Geo G, const Flags *f) {
    if (f->done) return;
    int64_t gl = gid();
    if (gl >= G.N * G.L) return;
    Coord c = decode(gl, G);
    const float xr = (c.x + 1 < G.nx) ? bx[gl + G.ny] : 0.0f;
    const float yr = (c.y + 1 < G.ny) ? by[gl + 1] : 0.0f;
    float zr = 0.0f, zl = 0.0f;
    if (IS3D) {
        zr = (c.z + 1 < G.nz) ? bz[gl + G.ny * G.nx] : 0.0f;
        zl = bz[gl];
    }
    gk[gl] = proj_factor<IS3D>(xr, bx[gl], yr, by[gl], zr, zl, alpha[gl]);
}

// rescale interior faces (:351-352, :372-373, :392-393)
template <bool IS3D>
__global__ void kp_rescale(const float *gk, float *bx, float *by, float *bz, Geo G,
                           const Flags *f) {
    if (f->done) return;
    int64_t gl = gid();
    if (gl >= G.N * G.L) return;
    Coord c = decode(gl, G);
    const float gc = gk[gl];
    if (c.x > 0) bx[gl] = rescale_face(gc, gk[gl - G.ny], bx[gl]);
    if (c.y > 0) by[gl] = rescale_face(gc, gk[gl - 1], by[gl]);
    if (IS3D && c.z > 0) bz[gl] = rescale_face(gc, gk[gl - G.ny * G.nx], bz[gl]);
}

template <bool IS3D>
__device__ __forceinline__ float divergence(const float *bx, const float *by, const float *bz,
                                            int64_t gl, const Coord &c, const Geo &G) {
    const float xr = (c.x + 1 < G.nx) ? bx[gl + G.ny] : 0.0f;
    const float yr = (c.y + 1 < G.ny) ? by[gl + 1] : 0.0f;
    float d = ((yr - by[gl]) + xr) - bx[gl];
    if (IS3D) {
        const float zr = (c.z + 1 < G.nz) ? bz[gl + G.ny * G.nx] : 0.0f;
        d = (d + zr) - bz[gl];
    }
    return d;
}

// dv = div b of a given flow field (the dv a resumed solve starts from; the same
// expression as the update below, asetsBinaryMF3D_mex.c:414-415)
__global__ void kp_div(const float *bx, const float *by, const float *bz, float *dv, Geo G,
                       bool is3d) {
    int64_t g = gid();
    if (g >= G.N) return;
    Coord c = decode(g, G);
    dv[g] = is3d ? divergence<true>(bx, by, bz, g, c, G) : divergence<false>(bx, by, bz, g, c, G);
}

// divergence, ps, pt, u and |e| (asetsBinaryMF3D_mex.c:414-432)
template <bool IS3D>
__global__ void kp_update_binary(const float *bx, const float *by, const float *bz, float *dv,
                                 float *ps, float *pt, float *u, const float *Cs,
                                 const float *Ct, float cc, Geo G, double *partials,
                                 const Flags *f) {
    __shared__ double red[16];
    if (f->done) return;
    int64_t g = gid();
    double e_abs = 0.0;
    if (g < G.N) {
        Coord c = decode(g, G);
        const float d = divergence<IS3D>(bx, by, bz, g, c, G);
        dv[g] = d;
        const float uc = u[g] / cc;
        float v = ((pt[g] - uc) + d) + 1.0f / cc;
        const float nps = fmin_ref(v, Cs[g]);
        ps[g] = nps;
        v = (nps + uc) - d;
        const float npt = fmin_ref(v, Ct[g]);
        pt[g] = npt;
        const float e = cc * ((npt + d) - nps);
        e_abs = fabs((double)e);
        u[g] = u[g] - e;
    }
    double s = block_sum(e_abs, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// Potts per-label divergence and sink flow (asetsPotts3D_mex.c:457-466)
template <bool IS3D>
__global__ void kp_ptdiv(const float *bx, const float *by, const float *bz, float *dv,
                         const float *ps, float *pt, const float *u, const float *Ct, float cc,
                         Geo G, const Flags *f) {
    if (f->done) return;
    int64_t gl = gid();
    if (gl >= G.N * G.L) return;
    Coord c = decode(gl, G);
    const float d = divergence<IS3D>(bx, by, bz, gl, c, G);
    dv[gl] = d;
    const float t = (ps[c.g] + u[gl] / cc) - d;
    const float ct = Ct[gl];
    pt[gl] = (t < ct) ? t : ct;
}

// Potts source flow and multipliers, labels in order (asetsPotts3D_mex.c:495-514).
// ft_l = dv_l + pt_l is recomputed in the second loop (same bits) instead of
// being kept in an array of L entries.
__global__ void kp_psu(const float *dv, const float *pt, float *u, float *ps, float cc, Geo G,
                       double *partials, const Flags *f) {
    __shared__ double red[16];
    if (f->done) return;
    int64_t g = gid();
    double e_abs = 0.0;
    if (g < G.N) {
        const float fl = (float)G.L;
        float s = 0.0f;
        for (int64_t l = 0; l < G.L; ++l) {
            const int64_t i = g + l * G.N;
            const float ft = dv[i] + pt[i];
            s += (ft - u[i] / cc);
        }
        const float nps = s / fl + 1.0f / (cc * fl);
        ps[g] = nps;
        for (int64_t l = 0; l < G.L; ++l) {
            const int64_t i = g + l * G.N;
            const float ft = dv[i] + pt[i];
            const float e = cc * (ft - nps);
            u[i] = u[i] - e;
            e_abs += fabs((double)e);
        }
    }
    double s = block_sum(e_abs, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// Sum the per-block partials in a fixed order, store cvg[it] and raise the stop
// flag when cvg <= err_bound (asetsBinaryMF3D_mex.c:183-188).
// STRICT: stop on cvg < err_bound (asetsIshikawa2D_mex.c:205) instead of <=.
template <bool STRICT = false>
__global__ void k_finalize(const double *partials, int64_t nparts, double denom, int32_t it,
                           float err_bound, float *cvg, int32_t *it_num, Flags *f) {
    __shared__ double red[16];
    if (f->done || f->stop_at) return;
    const double v = strided_sum(partials, nparts, threadIdx.x, blockDim.x);
    double s = block_sum(v, red);
    if (threadIdx.x == 0) {
        const float c = (float)(s / denom);
        cvg[it] = c;
        if (STRICT ? (c < err_bound) : (c <= err_bound)) {
            f->done = 1;
            f->it_num = it;
            *it_num = it;
        }
    }
}

}  // namespace asets
