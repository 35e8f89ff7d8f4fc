// kernels_batched.h -- many binary 3D volumes of uneven size per launch.
//
// The work list holds one item per (volume, tile, z chunk).  A persistent grid
// (sized to the resident capacity of the device) pulls items from a device
// counter, so volumes of any size and volumes that have already converged
// (their items are skipped) balance across the CUs.  Each item writes the
// partial |e| sum of its own slot, and a per-volume finalize kernel reduces a
// volume's slots in a fixed order: results do not depend on which workgroup ran
// which item.  Iteration counters, stop flags and the work counter live in
// device memory, so a sequence of iterations can be captured once in a HIP
// graph and replayed until every volume has stopped.
#pragma once
#include "kernels_fused.h"

namespace asets {

// One volume's two state buffers: planar fields, or (BLK kernels) one
// block-interleaved array each (BlkLayout), u[k] its base.
template <typename T>
struct BatchVol {
    const float *Cs, *Ct, *alpha;
    float *u[2];
    T *ps[2], *pt[2], *bx[2], *by[2], *bz[2];
    int ny, nx, nz;
    int ntile_y, ntile_x, nzc, zchunk, nblocks;
    int work0;   // first partial slot of this volume
    int nyb, pb; // blocks per column, state bytes per plane (block-interleaved state)
};

struct BatchMeta {
    int work0, nblocks;
    double n;          // voxels
    float *cvg;        // [max_iter]
    int32_t *it_num;
    float *u_out;      // the caller's u
};

struct BatchCtl {
    int32_t remaining;   // volumes still iterating
    int32_t bad_alpha;   // some volume has alpha <= 0 (first iteration)
    int32_t work_next;   // work-list cursor of the running iteration
    int32_t pad[61];
};

// Values read through a wave-uniform index are uniform, but the compiler cannot
// see it: readfirstlane moves them to scalar registers.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
template <typename P>
__device__ __forceinline__ P *uni(P *p) {
    const uint64_t x = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return (P *)(((uint64_t)hi << 32) | lo);
}

// FIRST: the first iteration of every volume (cold start, fused_tile), which also
// checks alpha (vbad[v] = 1 and ctl->bad_alpha = 1 when alpha <= 0 somewhere in
// volume v).
// BLK: block-interleaved state buffers (BlkLayout).
template <int TX, int NT, int WAVES, typename T, bool FIRST = false, bool BLK = true>
__global__ __launch_bounds__(NT, WAVES) void kf_binary_batched(
    const BatchVol<T> *__restrict__ vols, const int2 *__restrict__ work, int nwork,
    const int32_t *vdone, const int32_t *viter, double *partials, BatchCtl *ctl, float cc,
    float steps, int fast_cc, int32_t *vbad) {
    __shared__ double red[16];
    __shared__ int item;
    for (;;) {
        if (threadIdx.x == 0) item = atomicAdd(&ctl->work_next, 1);
        __syncthreads();
        const int w = uni(item);
        __syncthreads();
        if (w >= nwork) break;
        const int v = uni(work[w].x), lb = uni(work[w].y);
        if (uni(vdone[v])) continue;
        const BatchVol<T> &V = vols[v];
        const int i = uni(viter[v]);
        const int ib = i & 1, ob = (i + 1) & 1;
        FusedArgs<T> a;
        a.Cs = uni(V.Cs);
        a.Ct = uni(V.Ct);
        a.alpha = uni(V.alpha);
        a.in = FusedState<T>{uni(V.u[ib]), uni(V.ps[ib]), uni(V.pt[ib]), uni(V.bx[ib]),
                             uni(V.by[ib]), uni(V.bz[ib])};
        a.out = FusedState<T>{uni(V.u[ob]), uni(V.ps[ob]), uni(V.pt[ob]), uni(V.bx[ob]),
                              uni(V.by[ob]), uni(V.bz[ob])};
        a.ny = uni(V.ny);
        a.nx = uni(V.nx);
        a.nz = uni(V.nz);
        a.plane = a.ny * a.nx;
        a.cc = cc;
        a.steps = steps;
        a.zchunk = uni(V.zchunk);
        a.ntile_y = uni(V.ntile_y);
        a.ntile_x = uni(V.ntile_x);
        a.nzc = uni(V.nzc);
        a.nblocks = uni(V.nblocks);
        a.fast_cc = fast_cc;
        a.partials = nullptr;
        a.flags = nullptr;
        a.alpha_bad = FIRST ? vbad + v : nullptr;
        a.alpha_stop = FIRST ? &ctl->bad_alpha : nullptr;
        a.nyb = uni(V.nyb);
        a.pb = uni(V.pb);
        // normal cache policy: streaming stores measured slower for these volumes
        const double e = fused_tile<true, TX, NT, T, 0, FIRST, false, BLK>(a, lb);
        const double s = block_sum(e, red);
        if (threadIdx.x == 0) partials[uni(V.work0) + lb] = s;
        __syncthreads();
    }
}

// One workgroup per volume: fixed-order sum of the volume's partial slots, cvg,
// stop test (asetsBinaryMF3D_mex.c:183-188), iteration counter; resets the work
// cursor for the next iteration.
__global__ void k_finalize_batched(const BatchMeta *meta, int32_t *vdone, int32_t *viter,
                                   const double *partials, float err_bound, int32_t max_iter,
                                   BatchCtl *ctl) {
    __shared__ double red[16];
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->work_next = 0;
    const int v = blockIdx.x;
    if (vdone[v]) return;
    const BatchMeta m = meta[v];
    double acc = 0.0;
    for (int i = threadIdx.x; i < m.nblocks; i += blockDim.x) acc += partials[m.work0 + i];
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) {
        const int it = viter[v];
        const float c = (float)(s / m.n);
        m.cvg[it] = c;
        if (c <= err_bound) {
            vdone[v] = 1;
            *m.it_num = it;
            atomicSub(&ctl->remaining, 1);
        } else if (it + 1 >= max_iter) {
            vdone[v] = 1;
            *m.it_num = max_iter;
            atomicSub(&ctl->remaining, 1);
        } else {
            viter[v] = it + 1;
        }
    }
}

}  // namespace asets
