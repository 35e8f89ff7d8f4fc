// asets_hip.hip -- libasets_hip.so: host orchestration and the extern "C"
// entry points declared in include/asets.h.
//
// A solve is: init kernel(s) -> max_iter iterations, each a few kernels plus a
// finalize kernel that reduces the per-block FP64 |e| partials in a fixed
// order, stores cvg[it] and raises a device stop flag when cvg <= err_bound
// (asetsBinaryMF3D_mex.c:183-188).  Every kernel of a later iteration reads the
// flag first and returns at once when it is set, so the host can enqueue many
// iterations ahead and poll the flag only every poll_every iterations; u is
// the state after iteration it_num, exactly as in the reference loop.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "asets.h"
#include "common.h"
#include "kernels_batched.h"
#include "kernels_bounds.h"
#include "kernels_energy.h"
#include "kernels_prepost.h"
#include "kernels_fused.h"
#include "kernels_hmf.h"
#include "kernels_membench.h"
#include "kernels_passes.h"

using namespace asets;

namespace {

thread_local std::string g_last_error;

asets_status fail(asets_status s, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return s;
}

#define HIPCHK(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(ASETS_EDEVICE, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

constexpr int kBlock = 256;
constexpr int kMaxDevices = 64;
// Most partial sums for which every workgroup of a fused binary launch runs the
// previous iteration's stop test itself (each reads all of them: 16 KiB at most).
constexpr int64_t kStopAllParts = 2048;

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Per-device cached workspace, grown on demand and reused across calls.  A
// per-device mutex serialises solves on one device (they share the buffers).
struct DeviceCache {
    std::mutex mu;     // guards dev and host_flag for the duration of a solve
    std::mutex io_mu;  // guards io for the duration of a host-pointer call
    void *dev = nullptr;
    size_t dev_bytes = 0;
    void *io = nullptr;  // staging for the host-pointer entry points
    size_t io_bytes = 0;
    int32_t *host_flag = nullptr;  // pinned
};
DeviceCache g_cache[kMaxDevices];

asets_status ensure(void **p, size_t *cap, size_t need) {
    if (*cap >= need) return ASETS_OK;
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
        *cap = 0;
    }
    if (hipMalloc(p, need) != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(ASETS_ENOMEM, "hipMalloc of %zu bytes failed", need);
    }
    *cap = need;
    return ASETS_OK;
}

// Bump allocator over a workspace; 256-byte aligned slices.
struct Carver {
    char *base;
    size_t off = 0;
    explicit Carver(void *b) : base((char *)b) {}
    template <typename T>
    T *take(int64_t n) {
        T *p = (T *)(base + off);
        off += ((size_t)n * sizeof(T) + 255) & ~(size_t)255;
        return p;
    }
    static size_t sz(int64_t n, size_t es) { return ((size_t)n * es + 255) & ~(size_t)255; }
};

// Kernel-timing events of one call, destroyed on every return path.
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

asets_status check_params(const asets_params *p) {
    if (!p) return fail(ASETS_EBADARG, "params is NULL");
    if (p->max_iter < 0) return fail(ASETS_EBADARG, "max_iter must be >= 0 (got %d)", p->max_iter);
    if (!(p->cc > 0.0f) && !(p->cc < 0.0f))
        return fail(ASETS_EBADARG, "cc must be non-zero");
    return ASETS_OK;
}

asets_status check_dims(int64_t ny, int64_t nx, int64_t nz, int64_t L) {
    if (ny <= 0 || nx <= 0 || nz <= 0 || L <= 0)
        return fail(ASETS_EBADARG, "sizes must be positive (ny=%lld nx=%lld nz=%lld L=%lld)",
                    (long long)ny, (long long)nx, (long long)nz, (long long)L);
    if (ny * nx > ((int64_t)1 << 28))
        return fail(ASETS_ETOOBIG, "a plane of %lld voxels exceeds the 2^28 limit",
                    (long long)(ny * nx));
    if (nz > (int64_t)INT32_MAX / 2 || ny * nx * nz * L > ((int64_t)1 << 40))
        return fail(ASETS_ETOOBIG, "volume too large");
    return ASETS_OK;
}

// The fused kernels address one field with 32-bit byte offsets, planes -2 to
// nz+1 included (out-of-range planes must fall past the field and read 0): the
// field plus two planes must stay below 2^31 bytes (about 800^3 voxels).
inline bool fused_fits(int64_t ny, int64_t nx, int64_t nz) {
    return ny * nx * (nz + 2) * 4 < ((int64_t)1 << 31);
}

// Bytes per plane of the block-interleaved binary state (BlkLayout) for flow
// element size es.
inline int64_t blk_plane_bytes(int64_t ny, int64_t nx, size_t es) {
    return nx * cdiv(ny, 64) * (256 + 320 * (int64_t)es);
}
// The block-interleaved binary kernel addresses one state buffer with 32-bit byte
// offsets, planes -3 to nz+2 (out-of-range planes must fall past the buffer): the
// buffer plus four planes must stay below 2^31 bytes (about 430^3 voxels with fp32
// flows).  Larger volumes run the planar fused kernel.
inline bool fused_fits_blk(int64_t ny, int64_t nx, int64_t nz, size_t es) {
    return (nz + 4) * blk_plane_bytes(ny, nx, es) < ((int64_t)1 << 31);
}

// alpha <= 0 anywhere is an error when at least one iteration runs
// (asetsBinaryMF3D_mex.c:317-319 raises it inside the first projection pass).
__global__ void k_alpha_check(const float *alpha, int64_t n, int32_t *bad, int32_t *done) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int found = 0;
    for (; i < n; i += stride)
        if (alpha[i] <= 0.0f) found = 1;
    if (found) {
        atomicOr(bad, 1);
        if (done) atomicOr(done, 1);
    }
}

// Enqueue the alpha > 0 check into the solve's flags (cleared before): a
// violation stops every following iteration at once, and the host reports it
// at its first poll of the flags (run_loop), before any output is written.
void alpha_check_async(const float *alpha, int64_t n, Flags *f, hipStream_t s) {
    const int64_t blocks = std::min<int64_t>(cdiv(n, kBlock), 4096);
    k_alpha_check<<<(unsigned)blocks, kBlock, 0, s>>>(alpha, n, &f->bad_alpha, &f->done);
}

asets_status alpha_error() { return fail(ASETS_EALPHA, "alpha(x) must be positive. Exiting..."); }

inline void fill(float *p, float v, int64_t n, hipStream_t s) {
    if (n > 0) k_fill_f32<<<(unsigned)cdiv(n, kBlock), kBlock, 0, s>>>(p, v, n);
}

// A solve's flags cleared, *it_num = max_iter and cvg[0, max_iter) = 0, in one
// launch (before any alpha check writes into the flags).
void solve_setup(Flags *f, int32_t *it_num, int32_t max_iter, float *cvg, hipStream_t s) {
    constexpr int kWords = (int)(sizeof(Flags) / sizeof(int32_t));
    const int64_t n = std::max<int64_t>(std::max<int64_t>(max_iter, kWords), 1);
    k_solve_setup<<<(unsigned)cdiv(n, kBlock), kBlock, 0, s>>>((int32_t *)f, kWords, it_num,
                                                              max_iter, cvg, max_iter);
}

// Enqueue iterations [0, max_iter) via enqueue_iter(it), polling the stop flag
// on the host every poll iterations.  Returns the number of iterations whose
// kernels were enqueued.
// hflag: host memory for the four flag words read at each poll.
template <typename F>
asets_status run_loop(int32_t max_iter, float err_bound, int32_t poll_opt, hipStream_t s,
                      int32_t *hflag, Flags *flags, int32_t *itn, F &&enqueue_iter) {
    int32_t poll = poll_opt > 0 ? poll_opt : (err_bound > 0.0f ? 16 : 256);
    int32_t it = 0;
    *itn = max_iter;   // no stop before max_iter
    while (it < max_iter) {
        int32_t k = std::min(poll, max_iter - it);
        for (int32_t j = 0; j < k; ++j) {
            asets_status st = enqueue_iter(it + j);
            if (st != ASETS_OK) return st;
        }
        it += k;
        HIPCHK(hipGetLastError());
        // done, bad_alpha, it_num and stop_at in one read
        HIPCHK(hipMemcpyAsync(hflag, &flags->done, 4 * sizeof(int32_t), hipMemcpyDeviceToHost,
                              s));
        HIPCHK(hipStreamSynchronize(s));
        if (hflag[1]) return alpha_error();
        if (hflag[0] || hflag[3]) {
            *itn = hflag[2];
            break;
        }
        if (poll < 128 && err_bound > 0.0f && poll_opt <= 0) poll *= 2;
    }
    return ASETS_OK;
}

struct TileCfg {
    int tx;
    int zchunk;
};

// z-chunk length of the single-volume fused kernels for `tiles` tiles per z
// chunk layer (all labels or flow fields of a launch).  A workgroup streams its
// chunk plus the 4-plane prologue; the 64 x 16 tile's LDS allows two resident
// workgroups per CU, and two give about 1.25x the throughput of one.  A CU given
// k = ceil(W / CUs) of the W workgroups runs them in pairs, the odd one alone:
// about (k / 2) * 2 / 1.25 + (k % 2) workgroup-times, and the chunk (at least 8
// planes) with the least modelled time wins.  A grid just past one resident round
// is what it avoids: Ishikawa at 192^3 with 5 flow fields (180 tiles) took 0.731
// ms per iteration with 64-plane chunks (540 workgroups); a sweep of 20 to 64
// planes ranked the chunks as this model does (28 best, 0.591 ms; 48, 0.600 ms;
// 32, 0.626 ms), and so for Potts (28, 0.630 ms; 48, 0.661 ms).  For the binary
// kernel it picks what a resident-round rule picked (256^3: 32).
int pick_zchunk(int64_t tiles, int64_t nz) {
    if (nz <= 8) return (int)std::max<int64_t>(nz, 1);
    int cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0) {
        (void)hipGetLastError();
        cus = 256;
    }
    double best = 0.0;
    int64_t best_zc = nz;
    for (int64_t nzc = 1; nzc <= nz / 8; ++nzc) {
        const int64_t zc = cdiv(nz, nzc);
        const int64_t k = cdiv(tiles * cdiv(nz, zc), cus);
        const double cost = ((double)(k / 2) * (2.0 / 1.25) + (double)(k % 2)) *
                            (double)(zc + 4);
        if (nzc == 1 || cost < best) {
            best = cost;
            best_zc = zc;
        }
    }
    return (int)best_zc;
}

// Tile shape for the fused kernel.  Default: 64 x 16 tiles (512 threads, two
// workgroups per CU by LDS) and z chunks from pick_zchunk.  A single plane with
// fewer 64 x 16 tiles than CUs takes 64 x 8 tiles: twice the workgroups on its
// short, latency-bound iteration (256^2: 9.15 vs 9.43 us).  Overridden by
// asets_opts.
TileCfg pick_tiles(int64_t ny, int64_t nx, int64_t nz, const asets_opts *o) {
    TileCfg t;
    if (o && o->tile_x > 0) t.tx = o->tile_x;
    else t.tx = (nz == 1 && cdiv(ny, 64) * cdiv(nx, 16) < 256) ? 8 : 16;
    if (t.tx != 8 && t.tx != 16 && t.tx != 32) t.tx = 16;
    const int64_t tiles = cdiv(ny, 64) * cdiv(nx, t.tx);
    const int64_t zc = (o && o->zchunk > 0) ? o->zchunk : pick_zchunk(tiles, nz);
    t.zchunk = (int)std::min<int64_t>(zc, nz);
    return t;
}

// Fused kernel instances: (tile width TX, threads per workgroup).  Wider tiles
// use more threads so each thread keeps few positions in registers.
template <bool IS3D, int TX, int NT, typename T>
void launch_fused_t(const FusedArgs<T> &a, int grid, hipStream_t s, bool first, bool blk) {
    // occupancy target: 4 waves per SIMD (<= 128 VGPRs) for the 512- and
    // 1024-thread tiles, 3 for the 256-thread tile (LDS allows 3 per CU)
    constexpr int W = (NT >= 512) ? 4 : 3;
    if (blk) {
        if (first) kf_binary<IS3D, TX, NT, W, T, true, true><<<grid, NT, 0, s>>>(a);
        else kf_binary<IS3D, TX, NT, W, T, false, true><<<grid, NT, 0, s>>>(a);
    } else {
        if (first) kf_binary<IS3D, TX, NT, W, T, true, false><<<grid, NT, 0, s>>>(a);
        else kf_binary<IS3D, TX, NT, W, T, false, false><<<grid, NT, 0, s>>>(a);
    }
}

// first: the cold-start first iteration (the initial cut formed from Cs and Ct,
// alpha checked into a.alpha_bad / a.alpha_stop); blk: block-interleaved state
template <typename T>
void launch_fused(bool is3d, int tx, const FusedArgs<T> &a, int grid, hipStream_t s,
                  bool first, bool blk) {
    if (is3d) {
        if (tx == 8) launch_fused_t<true, 8, 256, T>(a, grid, s, first, blk);
        else if (tx == 32) launch_fused_t<true, 32, 1024, T>(a, grid, s, first, blk);
        else launch_fused_t<true, 16, 512, T>(a, grid, s, first, blk);
    } else {
        if (tx == 8) launch_fused_t<false, 8, 256, T>(a, grid, s, first, blk);
        else if (tx == 32) launch_fused_t<false, 32, 1024, T>(a, grid, s, first, blk);
        else launch_fused_t<false, 16, 512, T>(a, grid, s, first, blk);
    }
}

__global__ void k_selftest_div(int mode, float d, unsigned long long *count) {
    unsigned long long bad = 0;
    const float rd = rcp_refined(d);
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < (1ull << 32);
         b += (uint64_t)gridDim.x * blockDim.x) {
        const float x = __uint_as_float((uint32_t)b);
        float got, want;
        if (mode == 0) {
            got = div_fast(x, d, rd);
            want = x / d;
        } else {
            if (!div_fast_ok(x)) continue;
            got = div_fast(1.0f, x, rcp_refined(x));
            want = 1.0f / x;
        }
        const bool same = (__float_as_uint(got) == __float_as_uint(want)) ||
                          (got != got && want != want);
        bad += same ? 0 : 1;
    }
    if (bad) atomicAdd(count, bad);
}

__global__ void k_copy_f32(float *dst, const float *src, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

template <typename T>
__global__ void k_fill_t(T *p, float v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) Io<T>::put(p + i, v);
}

template <typename T>
__global__ void k_init_binary_t(const float *Cs, const float *Ct, const float *u0,
                                const float *ps0, const float *pt0, float *u, T *ps, T *pt,
                                int64_t N) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N) return;
    float uu, ms, mt;
    if (u0) {
        uu = u0[g];
        mt = pt0[g];
        ms = ps0 ? ps0[g] : mt;
    } else {
        const float cs = Cs[g], ct = Ct[g];
        uu = (cs - ct >= 0.0f) ? 1.0f : 0.0f;   // asetsBinaryMF3D_mex.c:215-219
        ms = mt = fmin_ref(cs, ct);
    }
    u[g] = uu;
    Io<T>::put(ps + kPair * g, ms);   // paired binary state: pt == ps + 1
    Io<T>::put(pt + kPair * g, mt);
}

// A resumed state's source and sink flows when the caller gives only u and the faces:
// the fixed point of the source/sink update (asetsBinaryMF3D_mex.c:411-429 at du = 0)
// pairs them as dv = div b (the faces with axis index 0 taken as 0), pt = Ct where
// u >= 1/2 and min(Cs - dv, Ct) elsewhere, ps = pt + dv, with u clamped to [0, 1].
// The same operations, in the same order, as asets_nn_warm_state (2D: no z terms).
__device__ __forceinline__ void pair_voxel(const float *u0, const float *bx0, const float *by0,
                                           const float *bz0, const float *Cs, const float *Ct,
                                           int64_t g, int x, int y, int z, int ny, int nx,
                                           int nz, int64_t plane, float &u, float &ps,
                                           float &pt) {
    const float byl = y > 0 ? by0[g] : 0.0f;
    const float bxl = x > 0 ? bx0[g] : 0.0f;
    const float byh = y + 1 < ny ? by0[g + 1] : 0.0f;
    const float bxh = x + 1 < nx ? bx0[g + ny] : 0.0f;
    float dv = byh - byl + bxh - bxl;
    if (bz0) {
        const float bzl = z > 0 ? bz0[g] : 0.0f;
        const float bzh = z + 1 < nz ? bz0[g + plane] : 0.0f;
        dv = dv + bzh - bzl;
    }
    float uu = u0[g];
    uu = uu < 0.0f ? 0.0f : (uu > 1.0f ? 1.0f : uu);
    const float cs = Cs[g], ct = Ct[g];
    float t = cs - dv;
    t = t < ct ? t : ct;
    u = uu;
    pt = uu >= 0.5f ? ct : t;
    ps = pt + dv;
}

// u, ps and pt of a resumed solve paired as pair_voxel, into planar state buffers
// (ps and pt every `stride` elements: 1, or kPair for the fused kernels' pairs).
template <typename T>
__global__ void k_pair_init_t(const float *Cs, const float *Ct, const float *u0,
                              const float *bx0, const float *by0, const float *bz0, int ny,
                              int nx, int nz, float *u, T *ps, T *pt, int stride) {
    const int64_t plane = (int64_t)ny * nx, N = plane * nz;
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N) return;
    const int y = (int)(g % ny), x = (int)((g / ny) % nx), z = (int)(g / plane);
    float uu, ms, mt;
    pair_voxel(u0, bx0, by0, bz0, Cs, Ct, g, x, y, z, ny, nx, nz, plane, uu, ms, mt);
    u[g] = uu;
    Io<T>::put(ps + (int64_t)stride * g, ms);
    Io<T>::put(pt + (int64_t)stride * g, mt);
}

// The whole initial state of a cold start in one pass: u and the (ps, pt) pair
// as k_init_binary_t, the (bx, by) pair and bz zero.
template <typename T>
__global__ void k_init_binary_cold_t(const float *Cs, const float *Ct, float *u, T *pspt, T *bxby,
                                     T *bz, int64_t N) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N) return;
    const float cs = Cs[g], ct = Ct[g];
    const float m = fmin_ref(cs, ct);
    u[g] = (cs - ct >= 0.0f) ? 1.0f : 0.0f;   // asetsBinaryMF3D_mex.c:215-219
    Io<T>::put(pspt + kPair * g, m);
    Io<T>::put(pspt + kPair * g + 1, m);
    Io<T>::put(bxby + kPair * g, 0.0f);
    Io<T>::put(bxby + kPair * g + 1, 0.0f);
    Io<T>::put(bz + g, 0.0f);
}

// Initial flow faces of a resumed solve: dst = src, except the faces with axis
// index 0, which are the zero-flux boundary (never updated; asetsBinaryMF3D_mex.c
// leaves index 0 of bx, by, bz at 0) and are forced to 0.  axis: 0 = y (by),
// 1 = x (bx), 2 = z (bz).
template <typename T>
__global__ void k_init_faces(T *dst, const float *src, int64_t ny, int64_t nx, int64_t N,
                             int axis, int stride) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N) return;
    int64_t c = axis == 0 ? g % ny : (axis == 1 ? (g / ny) % nx : g / (ny * nx));
    Io<T>::put(dst + stride * g, c == 0 ? 0.0f : src[g]);
}

// Zero the three face fields (pair == 2: bx and by interleaved, by == bx + 1).
template <typename T>
void zero_faces(T *bx, T *by, T *bz, int64_t N, int pair, hipStream_t s) {
    const unsigned nb = (unsigned)cdiv(N, kBlock);
    if (pair == 2) {
        k_fill_t<T><<<(unsigned)cdiv(2 * N, kBlock), kBlock, 0, s>>>(bx, 0.0f, 2 * N);
    } else {
        k_fill_t<T><<<nb, kBlock, 0, s>>>(bx, 0.0f, N);
        k_fill_t<T><<<nb, kBlock, 0, s>>>(by, 0.0f, N);
    }
    k_fill_t<T><<<nb, kBlock, 0, s>>>(bz, 0.0f, N);
}

template <typename T>
void init_faces(T *const dst[3], const float *const src[3], int64_t ny, int64_t nx, int64_t nz,
                int pair, hipStream_t s) {
    const int64_t N = ny * nx * nz;
    const unsigned nb = (unsigned)cdiv(N, kBlock);
    const int axis[3] = {1, 0, 2};   // bx, by, bz
    const int stride[3] = {pair, pair, 1};
    zero_faces<T>(dst[0], dst[1], dst[2], N, pair, s);
    for (int k = 0; k < 3; ++k)
        if (src[k])
            k_init_faces<T><<<nb, kBlock, 0, s>>>(dst[k], src[k], ny, nx, N, axis[k], stride[k]);
}

template <typename T>
__global__ void k_to_f32(float *dst, const T *src, int64_t n, int stride) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = Io<T>::get(src + stride * i);
}

// Initial block-interleaved binary state (BlkLayout).  Cold start (u0 null): u =
// (Cs - Ct >= 0), ps = pt = min(Cs, Ct), flows 0 (asetsBinaryMF3D_mex.c:215-219).
// Warm start or resume: u0, ps0 (null: pt0), pt0 and the flows b0 (null: 0), whose faces
// with axis index 0 -- the zero-flux boundary, never updated -- are forced to 0.  Padding
// lanes y >= ny: all 0.  Flows are rounded as the storage type stores them.  One
// workgroup per column (x, z) of the volume (blk_col_grid), its threads over the lanes
// of the column's nyb blocks: the column's position is divided out once per workgroup,
// not per lane.
template <typename T>
// alpha (may be NULL): checked on the way (alpha <= 0 somewhere sets *bad and *done, as
// k_alpha_check does), for a resumed solve whose first iteration does not check it.
__global__ void k_blk_init(const float *Cs, const float *Ct, const float *u0, const float *ps0,
                           const float *pt0, const float *bx0, const float *by0,
                           const float *bz0, char *st, int ny, int nx, int nz, int nyb,
                           int pair, const float *alpha, int32_t *bad, int32_t *done) {
    using BL = BlkLayout<T>;
    // 32-bit indices: the layout's buffers stay below 2^31 bytes (fused_fits_blk)
    const int c = (int)blockIdx.x;            // column z * nx + x
    const int x = c % nx, z = c / nx;
    const int64_t g0 = (int64_t)c * ny;       // voxel (x, y = 0, z)
    int alpha_bad = 0;
    for (int i = threadIdx.x; i < nyb * 64; i += blockDim.x) {
        const int yb = i >> 6, ln = i & 63;
        const int y = i;
        float u = 0.0f, ps = 0.0f, pt = 0.0f, bx = 0.0f, by = 0.0f, bz = 0.0f;
        if (y < ny) {
            const int64_t g = g0 + y;
            if (alpha && alpha[g] <= 0.0f) alpha_bad = 1;
            if (pair) {   // ps, pt paired to u and the faces (pair_voxel)
                pair_voxel(u0, bx0, by0, bz0, Cs, Ct, g, x, y, z, ny, nx, nz,
                           (int64_t)ny * nx, u, ps, pt);
            } else if (u0) {
                u = u0[g];
                pt = pt0[g];
                ps = ps0 ? ps0[g] : pt;
            } else {
                const float cs = Cs[g], ct = Ct[g];
                u = (cs - ct >= 0.0f) ? 1.0f : 0.0f;
                ps = pt = fmin_ref(cs, ct);
            }
            if (bx0 && x > 0) bx = bx0[g];
            if (by0 && y > 0) by = by0[g];
            if (bz0 && z > 0) bz = bz0[g];
        }
        char *blk = st + ((size_t)c * nyb + yb) * BL::kBytes;
        ((float *)(blk + BL::kU))[ln] = u;
        T *pp = (T *)(blk + BL::kPS) + 2 * ln;
        T *bb = (T *)(blk + BL::kBB) + 2 * ln;
        Io<T>::put(pp, ps);
        Io<T>::put(pp + 1, pt);
        Io<T>::put(bb, bx);
        Io<T>::put(bb + 1, by);
        Io<T>::put((T *)(blk + BL::kBZ) + ln, bz);
    }
    if (alpha_bad) {
        atomicOr(bad, 1);
        atomicOr(done, 1);
    }
}

// Launch shape of the per-column block-interleaved kernels: one workgroup per column,
// up to 256 threads over its 64 * nyb lanes.
struct BlkColGrid {
    unsigned grid, threads;
};
inline BlkColGrid blk_col_grid(int64_t nx, int64_t nz, int nyb) {
    const int lanes = 64 * nyb;
    return BlkColGrid{(unsigned)(nx * nz), (unsigned)(lanes < 256 ? lanes : 256)};
}

// Initial block-interleaved Potts flows (PottsBlkLayout) of every label: pt = ps, the
// initial source flow (asetsPotts3D_mex.c:219-258), and faces 0; padding lanes 0.  One
// thread per lane of a block of a label; st holds L labels of nz planes of pb bytes.
__global__ void k_init_potts_blk(const float *ps, char *st, int ny, int nx, int nz, int nyb,
                                 int pb, int L) {
    const int64_t per = (int64_t)nz * nx * nyb * 64;   // lanes per label
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= per * L) return;
    const int l = (int)(t / per);
    const int rem = (int)(t % per);
    const int b = rem >> 6, ln = rem & 63;
    const int yb = b % nyb, xz = b / nyb;
    const int y = yb * 64 + ln;
    float *blk = (float *)(st + (size_t)l * nz * pb + (size_t)b * PottsBlkLayout::kBytes);
    blk[ln] = (y < ny) ? ps[(int64_t)xz * ny + y] : 0.0f;   // pt
    blk[64 + ln] = 0.0f;                                   // bx
    blk[128 + ln] = 0.0f;                                  // by
    blk[192 + ln] = 0.0f;                                  // bz
}

// u and, where requested, the flows of a block-interleaved state to planar fp32
// arrays (own-face convention).  One workgroup per column, as k_blk_init.
template <typename T>
__global__ void k_blk_take(const char *st, float *u, float *ps, float *pt, float *bx, float *by,
                           float *bz, int ny, int nx, int nz, int nyb) {
    using BL = BlkLayout<T>;
    // 32-bit indices: the layout's buffers stay below 2^31 bytes (fused_fits_blk)
    const int c = (int)blockIdx.x;            // column z * nx + x
    const int64_t g0 = (int64_t)c * ny;
    for (int y = threadIdx.x; y < ny; y += blockDim.x) {
        const int yb = y >> 6, ln = y & 63;
        const int64_t g = g0 + y;
        const char *blk = st + ((size_t)c * nyb + yb) * BL::kBytes;
        const T *pp = (const T *)(blk + BL::kPS) + 2 * ln;
        const T *bb = (const T *)(blk + BL::kBB) + 2 * ln;
        if (u) u[g] = ((const float *)(blk + BL::kU))[ln];
        if (ps) ps[g] = Io<T>::get(pp);
        if (pt) pt[g] = Io<T>::get(pp + 1);
        if (bx) bx[g] = Io<T>::get(bb);
        if (by) by[g] = Io<T>::get(bb + 1);
        if (bz) bz[g] = Io<T>::get((const T *)(blk + BL::kBZ) + ln);
    }
}

// Everything the fused solve loop needs (one volume).
struct FusedRun {
    bool is3d;
    const float *Cs, *Ct, *alpha;
    int64_t ny, nx, nz;
    const asets_params *p;
    const float *u0, *ps0, *pt0;
    const float *b0[3];   // initial bx, by, bz (nullptr: 0)
    float *u, *cvg;
    int32_t *it_num;
    float *o_ps, *o_pt, *o_bx, *o_by, *o_bz;
    TileCfg tc;
    int ntile_y, ntile_x, nzc, nblocks, grid;
    int64_t nparts;
    double *partials;
    Flags *flags;
    hipStream_t s;
    hipEvent_t ev0, ev1;
    int32_t poll;
    int32_t *itn;   // host: iteration at which the stop test fired, or max_iter
    bool blk;       // block-interleaved state (BlkLayout) instead of planar fields
    int nyb;        // blocks per column (blk)
    int pb;         // state bytes per plane (blk)
    bool pair;      // resume: ps and pt paired to u0 and the faces (pair_voxel)
};

// Workspace bytes of a fused solve with flow element size es.
size_t fused_bytes(int64_t N, size_t es) {
    return 2 * (Carver::sz(N, sizeof(float)) + 5 * Carver::sz(N, es));
}

template <typename T>
asets_status run_fused(const FusedRun &r, Carver &cv, int32_t *hflag) {
    const int64_t N = r.ny * r.nx * r.nz;
    const hipStream_t s = r.s;
    const unsigned nb = (unsigned)cdiv(N, kBlock);
    // two state buffers; iteration i reads buf[i&1] and writes buf[(i+1)&1]
    FusedState<T> buf[2];
    for (int b = 0; b < 2; ++b) {
        if (r.blk) {   // one block-interleaved array; every pointer is its base
            char *base = cv.take<char>(r.nz * (int64_t)r.pb);
            buf[b] = FusedState<T>{(float *)base, (T *)base, (T *)base, (T *)base, (T *)base,
                                   (T *)base};
            continue;
        }
        buf[b].u = cv.take<float>(N);
        buf[b].ps = cv.take<T>(kPair * N);
        buf[b].pt = buf[b].ps + 1;
        buf[b].bx = cv.take<T>(kPair * N);
        buf[b].by = buf[b].bx + 1;
        buf[b].bz = cv.take<T>(N);
    }
    const int32_t max_iter = r.p->max_iter;
    // cold start: iteration 0 forms the initial state itself (and checks alpha);
    // with no iteration buf[0] must hold it for the outputs
    const bool first = !r.u0;
    if (r.blk) {
        if (!first || max_iter == 0) {
            const BlkColGrid cg = blk_col_grid(r.nx, r.nz, r.nyb);
            k_blk_init<T><<<cg.grid, cg.threads, 0, s>>>(r.Cs, r.Ct, r.u0, r.ps0, r.pt0,
                                                         r.b0[0], r.b0[1], r.b0[2],
                                                         (char *)buf[0].u, (int)r.ny,
                                                         (int)r.nx, (int)r.nz, r.nyb,
                                                         r.pair ? 1 : 0,
                                                         max_iter > 0 ? r.alpha : nullptr,
                                                         &r.flags->bad_alpha, &r.flags->done);
        }
    } else if (first && max_iter == 0) {
        k_init_binary_cold_t<T><<<nb, kBlock, 0, s>>>(r.Cs, r.Ct, buf[0].u, buf[0].ps,
                                                     buf[0].bx, buf[0].bz, N);
    } else if (!first) {
        if (r.pair)
            k_pair_init_t<T><<<nb, kBlock, 0, s>>>(r.Cs, r.Ct, r.u0, r.b0[0], r.b0[1], r.b0[2],
                                                  (int)r.ny, (int)r.nx, (int)r.nz, buf[0].u,
                                                  buf[0].ps, buf[0].pt, kPair);
        else
            k_init_binary_t<T><<<nb, kBlock, 0, s>>>(r.Cs, r.Ct, r.u0, r.ps0, r.pt0, buf[0].u,
                                                    buf[0].ps, buf[0].pt, N);
        T *const f0[3] = {buf[0].bx, buf[0].by, buf[0].bz};
        init_faces<T>(f0, r.b0, r.ny, r.nx, r.nz, kPair, s);
    }
    // buf[1] needs no initialisation: iteration 0 writes every voxel of it
    HIPCHK(hipGetLastError());

    FusedArgs<T> fa{};
    fa.Cs = r.Cs;
    fa.Ct = r.Ct;
    fa.alpha = r.alpha;
    fa.ny = (int)r.ny;
    fa.nx = (int)r.nx;
    fa.nz = (int)r.nz;
    fa.plane = (int)(r.ny * r.nx);
    fa.fast_cc = (std::fabs(r.p->cc) >= 0x1p-16f && std::fabs(r.p->cc) <= 0x1p16f) ? 1 : 0;
    fa.cc = r.p->cc;
    fa.steps = r.p->steps;
    fa.zchunk = r.tc.zchunk;
    fa.ntile_y = r.ntile_y;
    fa.ntile_x = r.ntile_x;
    fa.nzc = r.nzc;
    fa.nblocks = r.nblocks;
    fa.partials = r.partials;
    fa.flags = r.flags;
    fa.alpha_bad = &r.flags->bad_alpha;
    fa.alpha_stop = &r.flags->done;
    fa.nyb = r.nyb;
    fa.pb = r.pb;
    // err_bound 0 (a fixed iteration count; the stop test fires only on an exact
    // fixed point): no launch has to wait for an iteration's stop test, so it runs
    // inside the next iteration's kernel (DeferredStop, same bits as k_finalize),
    // on partials double-buffered by iteration parity; k_finalize only after the
    // last.  A stop found one launch late is still exact: that launch wrote the
    // other state buffer, and every later launch returns at once.
    // err_bound > 0: every workgroup of the next launch runs the stop test at its
    // start (stop_all) and returns when it fires, which costs each workgroup one
    // read of the partials instead of a k_finalize launch per iteration; the launch
    // after the last iteration then costs no more than an empty one.  Only where
    // the partials are few (each workgroup reads all of them).
    const bool stop_all = (r.p->err_bound != 0.0f) && r.nparts <= kStopAllParts;
    const bool deferred = (r.p->err_bound == 0.0f) || stop_all;
    // and when only u is asked for, the launch that finds the stop also writes it
    // (k_blk_take's work, without its launch after the host's poll)
    const bool take_u = stop_all && r.blk && !r.o_ps && !r.o_pt && !r.o_bx && !r.o_by &&
                        !r.o_bz && r.ny % 4 == 0 && ((uintptr_t)r.u & 15) == 0;
    double *const P[2] = {r.partials, r.partials + r.nparts};
    if (r.ev0) HIPCHK(hipEventRecord(r.ev0, s));
    auto iter = [&](int32_t it) -> asets_status {
        fa.in = buf[it & 1];
        fa.out = buf[(it + 1) & 1];
        fa.partials = deferred ? P[it & 1] : r.partials;
        fa.stop = DeferredStop{};
        fa.stop_all = stop_all ? 1 : 0;
        fa.take_u = take_u ? r.u : nullptr;
        if (deferred && it > 0)
            fa.stop = DeferredStop{P[(it - 1) & 1], r.nparts, (double)N, it - 1,
                                   r.p->err_bound, r.cvg, r.it_num, r.flags};
        launch_fused<T>(r.is3d, r.tc.tx, fa, r.grid, s, first && it == 0, r.blk);
        if (!deferred || it == max_iter - 1)
            k_finalize<<<1, 1024, 0, s>>>(fa.partials, r.nparts, (double)N, it, r.p->err_bound,
                                          r.cvg, r.it_num, r.flags);
        return ASETS_OK;
    };
    int32_t itn = 0;
    asets_status st = run_loop(max_iter, r.p->err_bound, r.poll, s, hflag, r.flags, &itn, iter);
    if (st) return st;
    if (r.ev1) HIPCHK(hipEventRecord(r.ev1, s));
    *r.itn = itn;
    // final state lives in buf[n_exec & 1]
    const int32_t n_exec = (itn < max_iter) ? itn + 1 : max_iter;
    const FusedState<T> &fin = buf[n_exec & 1];
    if (r.blk) {
        if (take_u && hflag[3] != 0) return ASETS_OK;   // the stop's launch wrote u
        const BlkColGrid cg = blk_col_grid(r.nx, r.nz, r.nyb);
        k_blk_take<T><<<cg.grid, cg.threads, 0, s>>>((const char *)fin.u, r.u, r.o_ps, r.o_pt,
                                                     r.o_bx, r.o_by, r.o_bz, (int)r.ny,
                                                     (int)r.nx, (int)r.nz, r.nyb);
        HIPCHK(hipGetLastError());
        return ASETS_OK;
    }
    HIPCHK(hipMemcpyAsync(r.u, fin.u, N * 4, hipMemcpyDeviceToDevice, s));
    const T *srcs[5] = {fin.ps, fin.pt, fin.bx, fin.by, fin.bz};
    float *dsts[5] = {r.o_ps, r.o_pt, r.o_bx, r.o_by, r.o_bz};
    for (int k = 0; k < 5; ++k)
        if (dsts[k]) k_to_f32<T><<<nb, kBlock, 0, s>>>(dsts[k], srcs[k], N, k < 4 ? kPair : 1);
    HIPCHK(hipGetLastError());
    return ASETS_OK;
}

template <bool IS3D, int TX, int NT>
void launch_potts_flow_t(const PottsArgs &p, int grid, hipStream_t s, bool blk) {
    constexpr int W = (NT >= 512) ? 4 : 3;
    if (blk) kf_potts_flow<IS3D, TX, NT, W, true><<<grid, NT, 0, s>>>(p);
    else kf_potts_flow<IS3D, TX, NT, W, false><<<grid, NT, 0, s>>>(p);
}

void launch_potts_flow(bool is3d, int tx, const PottsArgs &p, int grid, hipStream_t s,
                       bool blk) {
    if (is3d) {
        if (tx == 8) launch_potts_flow_t<true, 8, 256>(p, grid, s, blk);
        else if (tx == 32) launch_potts_flow_t<true, 32, 1024>(p, grid, s, blk);
        else launch_potts_flow_t<true, 16, 512>(p, grid, s, blk);
    } else {
        if (tx == 8) launch_potts_flow_t<false, 8, 256>(p, grid, s, blk);
        else if (tx == 32) launch_potts_flow_t<false, 32, 1024>(p, grid, s, blk);
        else launch_potts_flow_t<false, 16, 512>(p, grid, s, blk);
    }
}

void launch_potts_couple(int L, const float *ft, float *u, float *ps, int64_t N, float cc,
                         int fast_cc, double *partials, const Flags *flags, hipStream_t s) {
    const unsigned nb = (unsigned)cdiv(N, kBlock);
#define ASETS_COUPLE(LT) \
    kp_potts_couple<LT><<<nb, kBlock, 0, s>>>(ft, u, ps, N, L, cc, fast_cc, partials, flags)
    switch (L) {
    case 2: ASETS_COUPLE(2); break;
    case 3: ASETS_COUPLE(3); break;
    case 4: ASETS_COUPLE(4); break;
    case 5: ASETS_COUPLE(5); break;
    case 6: ASETS_COUPLE(6); break;
    case 8: ASETS_COUPLE(8); break;
    case 10: ASETS_COUPLE(10); break;
    default: ASETS_COUPLE(0); break;
    }
#undef ASETS_COUPLE
}

template <bool IS3D, int TX, int NT>
void launch_ishikawa_flow_t(const IshiArgs &p, int grid, hipStream_t s) {
    constexpr int W = (NT >= 512) ? 4 : 3;
    kf_ishikawa_flow<IS3D, TX, NT, W><<<grid, NT, 0, s>>>(p);
}
void launch_ishikawa_flow(bool is3d, int tx, const IshiArgs &p, int grid, hipStream_t s) {
    if (is3d) {
        if (tx == 8) launch_ishikawa_flow_t<true, 8, 256>(p, grid, s);
        else if (tx == 32) launch_ishikawa_flow_t<true, 32, 1024>(p, grid, s);
        else launch_ishikawa_flow_t<true, 16, 512>(p, grid, s);
    } else {
        if (tx == 8) launch_ishikawa_flow_t<false, 8, 256>(p, grid, s);
        else if (tx == 32) launch_ishikawa_flow_t<false, 32, 1024>(p, grid, s);
        else launch_ishikawa_flow_t<false, 16, 512>(p, grid, s);
    }
}
void launch_ishikawa_couple(int L, const float *dv, float *pt, float *u, const float *Ct,
                            int64_t N, float cc, int fast_cc, double *partials,
                            const Flags *flags, hipStream_t s) {
    const unsigned nb = (unsigned)cdiv(N, kBlock);
#define ASETS_ICOUPLE(LT) \
    k_ishikawa_couple<LT><<<nb, kBlock, 0, s>>>(dv, pt, u, Ct, N, L, cc, fast_cc, partials, flags)
    switch (L) {
    case 2: ASETS_ICOUPLE(2); break;
    case 3: ASETS_ICOUPLE(3); break;
    case 4: ASETS_ICOUPLE(4); break;
    case 5: ASETS_ICOUPLE(5); break;
    case 6: ASETS_ICOUPLE(6); break;
    case 8: ASETS_ICOUPLE(8); break;
    default: ASETS_ICOUPLE(0); break;
    }
#undef ASETS_ICOUPLE
}

// ---------------------------------------------------------------------------
// binary model on device memory
// ---------------------------------------------------------------------------
// Initial state of a solve: all null = the reference's initial cut; u0 and pt0 =
// warm start (ps0 = pt0, flows 0 unless given); every field = resume.
struct WarmIn {
    const float *u0 = nullptr, *ps0 = nullptr, *pt0 = nullptr;
    const float *b0[3] = {nullptr, nullptr, nullptr};
    bool pair = false;   // ps0, pt0 NULL: paired to u0 and the faces (pair_voxel)
};

// A workspace the caller fixes instead of the library's cached one (the
// allocation-free entry): a device region, its size, the mutex that serialises
// its users, and host memory for the stop-flag polls.
struct FixedWs {
    void *dev;
    size_t bytes;
    std::mutex *mu;
    int32_t *hflag;
};

// fw == nullptr: the cached per-device workspace (grown with hipMalloc as needed).
asets_status binary_dev(int ndim, const float *Cs, const float *Ct, const float *alpha,
                        int64_t ny, int64_t nx, int64_t nz, const asets_params *p,
                        const WarmIn &w, float *u, float *cvg,
                        int32_t *it_num, double *seconds, float *o_ps, float *o_pt,
                        float *o_bx, float *o_by, float *o_bz, const asets_opts *o,
                        const FixedWs *fw = nullptr) {
    auto t_start = std::chrono::steady_clock::now();
    if (ndim != 2 && ndim != 3) return fail(ASETS_EBADARG, "ndim must be 2 or 3");
    if (ndim == 2 && nz != 1) return fail(ASETS_EBADARG, "2D solve needs nz == 1");
    asets_status st = check_params(p);
    if (st) return st;
    st = check_dims(ny, nx, nz, 1);
    if (st) return st;
    if (!Cs || !Ct || !alpha || !u || !it_num || (p->max_iter > 0 && !cvg))
        return fail(ASETS_EBADARG, "null array argument");
    const float *u0 = w.u0, *pt0 = w.pt0;
    if (w.pair ? !u0 : (u0 == nullptr) != (pt0 == nullptr))
        return fail(ASETS_EBADARG, "warm start needs both u0 and pt0");
    if (!u0 && (w.ps0 || w.b0[0] || w.b0[1] || w.b0[2]))
        return fail(ASETS_EBADARG, "a resumed state needs u0 and pt0");
    if (ndim == 2 && w.b0[2]) return fail(ASETS_EBADARG, "a 2D state has no bz");
    const int flow16 = o ? o->flow_dtype : ASETS_FLOW_F32;
    if (flow16 != ASETS_FLOW_F32 && flow16 != ASETS_FLOW_F16)
        return fail(ASETS_EBADARG, "unknown flow_dtype %d", flow16);
    const int kern = o ? o->kernel : ASETS_KERNEL_AUTO;
    bool fused = (kern != ASETS_KERNEL_PASSES);
    if (flow16 == ASETS_FLOW_F16 && !fused)
        return fail(ASETS_EBADARG, "fp16 flow storage needs the fused kernel");
    if (fused && !fused_fits(ny, nx, nz)) {
        if (kern == ASETS_KERNEL_FUSED || flow16 == ASETS_FLOW_F16)
            return fail(ASETS_ETOOBIG, "volume too large for the fused kernel (fields of "
                                       "2^31 bytes or more); use the per-pass kernels");
        fused = false;   // auto: the per-pass kernels have no such limit
    }

    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(ASETS_EDEVICE, "device id %d", dev);
    if (fw && o && o->kernel_ms)
        return fail(ASETS_EBADARG, "kernel timing needs HIP events: not available without "
                                   "allocation");
    DeviceCache &dc = g_cache[dev];
    std::lock_guard<std::mutex> lock(fw ? *fw->mu : dc.mu);
    hipStream_t s = o ? (hipStream_t)o->stream : nullptr;
    if (!fw && !dc.host_flag) HIPCHK(hipHostMalloc((void **)&dc.host_flag, 64));
    int32_t *const hflag = fw ? fw->hflag : dc.host_flag;

    const bool is3d = (ndim == 3);
    const int64_t N = ny * nx * nz;
    const Geo G{ny, nx, nz, N, 1};
    const int32_t max_iter = p->max_iter;
    TileCfg tc = pick_tiles(ny, nx, nz, o);
    const int ntile_y = (int)cdiv(ny, 64), ntile_x = (int)cdiv(nx, tc.tx);
    const int nzc = (int)cdiv(nz, tc.zchunk);
    const int nblocks = ntile_y * ntile_x * nzc;
    const int grid_fused = (int)cdiv(nblocks, 8) * 8;
    const int64_t nparts = fused ? grid_fused : cdiv(N, kBlock);

    // workspace
    const size_t es = (flow16 == ASETS_FLOW_F16) ? sizeof(__half) : sizeof(float);
    // fused: two partial buffers (DeferredStop)
    size_t need = Carver::sz(1, sizeof(Flags)) + Carver::sz(2 * nparts, sizeof(double));
    // the block-interleaved state unless a buffer would pass its 2^31-byte limit
    const bool blk = fused && fused_fits_blk(ny, nx, nz, es);
    const int64_t pb = blk_plane_bytes(ny, nx, es);
    if (blk)
        need += 2 * Carver::sz(nz * pb, 1);
    else if (fused)
        need += fused_bytes(N, es);
    else
        need += 7 * Carver::sz(N, sizeof(float));
    if (fw) {
        if (need > fw->bytes)
            return fail(ASETS_ETOOBIG, "the solve needs %zu bytes of workspace; the "
                                       "allocation-free entry has %zu", need, fw->bytes);
    } else {
        st = ensure(&dc.dev, &dc.dev_bytes, need);
        if (st) return st;
    }
    Carver cv(fw ? fw->dev : dc.dev);
    Flags *flags = cv.take<Flags>(1);
    double *partials = cv.take<double>(2 * nparts);

    solve_setup(flags, it_num, max_iter, cvg, s);
    // a fused cold start checks alpha in its first iteration, a fused block-interleaved
    // resume while it lays the state out (k_blk_init)
    if (max_iter > 0 && !(fused && (!u0 || blk))) alpha_check_async(alpha, N, flags, s);
    HIPCHK(hipGetLastError());
    int32_t itn = max_iter;

    EventPair evp;
    hipEvent_t &ev0 = evp.a, &ev1 = evp.b;
    if (o && o->kernel_ms) {
        HIPCHK(hipEventCreate(&ev0));
        HIPCHK(hipEventCreate(&ev1));
    }

    if (!fused) {
        float *ps = cv.take<float>(N), *pt = cv.take<float>(N), *bx = cv.take<float>(N);
        float *by = cv.take<float>(N), *bz = cv.take<float>(N), *dv = cv.take<float>(N);
        float *gk = cv.take<float>(N);
        float *const f0[3] = {bx, by, bz};
        init_faces<float>(f0, w.b0, ny, nx, nz, 1, s);
        fill(dv, 0.0f, N, s);
        if (w.pair)
            k_pair_init_t<float><<<(unsigned)cdiv(N, kBlock), kBlock, 0, s>>>(
                Cs, Ct, u0, w.b0[0], w.b0[1], w.b0[2], (int)ny, (int)nx, (int)nz, u, ps, pt, 1);
        else
            k_init_binary<<<(unsigned)cdiv(N, kBlock), kBlock, 0, s>>>(Cs, Ct, u0, w.ps0, pt0, u,
                                                                       ps, pt, N);
        if (w.b0[0] || w.b0[1] || w.b0[2]) {
            // dv of the resumed flows: the divergence the previous iteration left
            kp_div<<<(unsigned)cdiv(N, kBlock), kBlock, 0, s>>>(bx, by, bz, dv, G, is3d);
        }
        HIPCHK(hipGetLastError());
        const unsigned nb = (unsigned)cdiv(N, kBlock);
        if (ev0) HIPCHK(hipEventRecord(ev0, s));
        auto iter = [&](int32_t it) -> asets_status {
            kp_gk<<<nb, kBlock, 0, s>>>(dv, ps, pt, u, gk, p->cc, G, flags);
            if (is3d) {
                kp_step<true><<<nb, kBlock, 0, s>>>(gk, bx, by, bz, p->steps, G, flags);
                kp_proj<true><<<nb, kBlock, 0, s>>>(bx, by, bz, alpha, gk, G, flags);
                kp_rescale<true><<<nb, kBlock, 0, s>>>(gk, bx, by, bz, G, flags);
                kp_update_binary<true><<<nb, kBlock, 0, s>>>(bx, by, bz, dv, ps, pt, u, Cs, Ct,
                                                            p->cc, G, partials, flags);
            } else {
                kp_step<false><<<nb, kBlock, 0, s>>>(gk, bx, by, bz, p->steps, G, flags);
                kp_proj<false><<<nb, kBlock, 0, s>>>(bx, by, bz, alpha, gk, G, flags);
                kp_rescale<false><<<nb, kBlock, 0, s>>>(gk, bx, by, bz, G, flags);
                kp_update_binary<false><<<nb, kBlock, 0, s>>>(bx, by, bz, dv, ps, pt, u, Cs, Ct,
                                                             p->cc, G, partials, flags);
            }
            k_finalize<<<1, 1024, 0, s>>>(partials, nparts, (double)N, it, p->err_bound, cvg,
                                          it_num, flags);
            return ASETS_OK;
        };
        st = run_loop(max_iter, p->err_bound, o ? o->poll_every : 0, s, hflag, flags, &itn, iter);
        if (st) return st;
        if (ev1) HIPCHK(hipEventRecord(ev1, s));
        if (o_ps) HIPCHK(hipMemcpyAsync(o_ps, ps, N * 4, hipMemcpyDeviceToDevice, s));
        if (o_pt) HIPCHK(hipMemcpyAsync(o_pt, pt, N * 4, hipMemcpyDeviceToDevice, s));
        if (o_bx) HIPCHK(hipMemcpyAsync(o_bx, bx, N * 4, hipMemcpyDeviceToDevice, s));
        if (o_by) HIPCHK(hipMemcpyAsync(o_by, by, N * 4, hipMemcpyDeviceToDevice, s));
        if (o_bz) HIPCHK(hipMemcpyAsync(o_bz, bz, N * 4, hipMemcpyDeviceToDevice, s));
    } else {
        FusedRun fr{is3d, Cs, Ct, alpha, ny, nx, nz, p, u0, w.ps0, pt0,
                    {w.b0[0], w.b0[1], w.b0[2]}, u, cvg, it_num,
                    o_ps, o_pt, o_bx, o_by, o_bz, tc, ntile_y, ntile_x, nzc, nblocks,
                    grid_fused, nparts, partials, flags, s, ev0, ev1,
                    o ? o->poll_every : 0, &itn, blk, (int)cdiv(ny, 64), (int)pb, w.pair};
        st = (flow16 == ASETS_FLOW_F16) ? run_fused<__half>(fr, cv, hflag)
                                          : run_fused<float>(fr, cv, hflag);
        if (st) return st;
    }
    HIPCHK(hipStreamSynchronize(s));
    if (ev0) {
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
        const int32_t n_exec = (itn < max_iter) ? itn + 1 : max_iter;
        *o->kernel_ms = n_exec > 0 ? (double)ms / n_exec : 0.0;
    }
    if (seconds)
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    return ASETS_OK;
}

// ---------------------------------------------------------------------------
// Potts model on device memory
// ---------------------------------------------------------------------------
asets_status potts_dev(int ndim, const float *Ct, const float *alpha, int64_t ny, int64_t nx,
                       int64_t nz, int64_t L, const asets_params *p, float *u, float *cvg,
                       int32_t *it_num, double *seconds, const asets_opts *o) {
    auto t_start = std::chrono::steady_clock::now();
    if (ndim != 2 && ndim != 3) return fail(ASETS_EBADARG, "ndim must be 2 or 3");
    if (ndim == 2 && nz != 1) return fail(ASETS_EBADARG, "2D solve needs nz == 1");
    asets_status st = check_params(p);
    if (st) return st;
    st = check_dims(ny, nx, nz, L);
    if (st) return st;
    if (!Ct || !alpha || !u || !it_num || (p->max_iter > 0 && !cvg))
        return fail(ASETS_EBADARG, "null array argument");
    if (o && o->flow_dtype != ASETS_FLOW_F32)
        return fail(ASETS_EBADARG, "fp16 flow storage is available for the binary model only");
    const int kern = o ? o->kernel : ASETS_KERNEL_AUTO;
    bool fused = (kern != ASETS_KERNEL_PASSES);
    if (fused && !fused_fits(ny, nx, nz)) {
        if (kern == ASETS_KERNEL_FUSED)
            return fail(ASETS_ETOOBIG, "volume too large for the fused kernel (fields of "
                                       "2^31 bytes or more); use the per-pass kernels");
        fused = false;
    }

    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(ASETS_EDEVICE, "device id %d", dev);
    DeviceCache &dc = g_cache[dev];
    std::lock_guard<std::mutex> lock(dc.mu);
    hipStream_t s = o ? (hipStream_t)o->stream : nullptr;
    if (!dc.host_flag) HIPCHK(hipHostMalloc((void **)&dc.host_flag, 64));

    const bool is3d = (ndim == 3);
    const int64_t N = ny * nx * nz, NL = N * L;
    const Geo G{ny, nx, nz, N, L};
    const int32_t max_iter = p->max_iter;
    const int64_t nparts = cdiv(N, kBlock);
    // fused: each label's flows block-interleaved (PottsBlkLayout) unless a label's
    // buffer plus four planes would reach 2^31 bytes (then planar fields)
    const int nyb = (int)cdiv(ny, 64);
    const int64_t ppb = nx * nyb * (int64_t)PottsBlkLayout::kBytes;
    const bool blk = fused && (nz + 4) * ppb < ((int64_t)1 << 31);
    size_t need = Carver::sz(1, sizeof(Flags)) + Carver::sz(nparts, sizeof(double)) +
                  Carver::sz(1, sizeof(int32_t)) + Carver::sz(N, 4);
    if (blk) need += 2 * Carver::sz(L * nz * ppb, 1) + Carver::sz(NL, 4);
    else need += (fused ? 9 : 6) * Carver::sz(NL, 4);
    st = ensure(&dc.dev, &dc.dev_bytes, need);
    if (st) return st;
    Carver cv(dc.dev);
    Flags *flags = cv.take<Flags>(1);
    double *partials = cv.take<double>(nparts);
    float *ps = cv.take<float>(N);

    solve_setup(flags, it_num, max_iter, cvg, s);
    if (max_iter > 0) alpha_check_async(alpha, NL, flags, s);

    EventPair evp;
    hipEvent_t &ev0 = evp.a, &ev1 = evp.b;
    if (o && o->kernel_ms) {
        HIPCHK(hipEventCreate(&ev0));
        HIPCHK(hipEventCreate(&ev1));
    }
    const unsigned nbl = (unsigned)cdiv(NL, kBlock), nbv = (unsigned)cdiv(N, kBlock);
    std::function<asets_status(int32_t)> iter;
    // per-pass family state
    float *bx = nullptr, *by = nullptr, *bz = nullptr, *dv = nullptr, *pt = nullptr, *gk = nullptr;
    // fused family state
    PottsArgs pa{};
    float *ptb[2] = {nullptr, nullptr}, *bb[2][3] = {};
    TileCfg tc{};
    int grid = 0;
    if (!fused) {
        bx = cv.take<float>(NL);
        by = cv.take<float>(NL);
        bz = cv.take<float>(NL);
        dv = cv.take<float>(NL);
        pt = cv.take<float>(NL);
        gk = cv.take<float>(NL);
        fill(bx, 0.0f, NL, s);
        fill(by, 0.0f, NL, s);
        fill(bz, 0.0f, NL, s);
        fill(dv, 0.0f, NL, s);
        k_init_potts<<<(unsigned)cdiv(N, kBlock), kBlock, 0, s>>>(Ct, u, ps, pt, N, L);
        iter = [&](int32_t it) -> asets_status {
            kp_gk<<<nbl, kBlock, 0, s>>>(dv, ps, pt, u, gk, p->cc, G, flags);
            if (is3d) {
                kp_step<true><<<nbl, kBlock, 0, s>>>(gk, bx, by, bz, p->steps, G, flags);
                kp_proj<true><<<nbl, kBlock, 0, s>>>(bx, by, bz, alpha, gk, G, flags);
                kp_rescale<true><<<nbl, kBlock, 0, s>>>(gk, bx, by, bz, G, flags);
                kp_ptdiv<true><<<nbl, kBlock, 0, s>>>(bx, by, bz, dv, ps, pt, u, Ct, p->cc, G, flags);
            } else {
                kp_step<false><<<nbl, kBlock, 0, s>>>(gk, bx, by, bz, p->steps, G, flags);
                kp_proj<false><<<nbl, kBlock, 0, s>>>(bx, by, bz, alpha, gk, G, flags);
                kp_rescale<false><<<nbl, kBlock, 0, s>>>(gk, bx, by, bz, G, flags);
                kp_ptdiv<false><<<nbl, kBlock, 0, s>>>(bx, by, bz, dv, ps, pt, u, Ct, p->cc, G, flags);
            }
            kp_psu<<<nbv, kBlock, 0, s>>>(dv, pt, u, ps, p->cc, G, partials, flags);
            k_finalize<<<1, 1024, 0, s>>>(partials, nparts, (double)NL, it, p->err_bound, cvg,
                                          it_num, flags);
            return ASETS_OK;
        };
    } else {
        // label-parallel fused flow kernel + per-voxel coupling kernel; u is
        // updated in place in the caller's buffer, pt and faces ping-pong
        if (blk) {   // one block-interleaved array per buffer, every label's flows
            for (int b = 0; b < 2; ++b) {
                ptb[b] = (float *)cv.take<char>(L * nz * ppb);
                for (int k = 0; k < 3; ++k) bb[b][k] = ptb[b];
            }
        } else {
            for (int b = 0; b < 2; ++b) {
                ptb[b] = cv.take<float>(NL);
                for (int k = 0; k < 3; ++k) bb[b][k] = cv.take<float>(NL);
            }
        }
        float *ft = cv.take<float>(NL);
        if (blk) {
            k_init_potts<<<(unsigned)cdiv(N, kBlock), kBlock, 0, s>>>(Ct, u, ps, nullptr, N, L);
            // buffer 1 needs none: the first flow launch writes every lane of it
            k_init_potts_blk<<<(unsigned)cdiv(L * nz * nx * nyb * 64, kBlock), kBlock, 0, s>>>(
                ps, (char *)ptb[0], (int)ny, (int)nx, (int)nz, nyb, (int)ppb, (int)L);
        } else {
            for (int b = 0; b < 2; ++b)
                for (int k = 0; k < 3; ++k) fill(bb[b][k], 0.0f, NL, s);
            k_init_potts<<<(unsigned)cdiv(N, kBlock), kBlock, 0, s>>>(Ct, u, ps, ptb[0], N, L);
        }
        tc.tx = (o && o->tile_x > 0) ? o->tile_x : 16;
        if (tc.tx != 8 && tc.tx != 32) tc.tx = 16;
        const int64_t tiles = cdiv(ny, 64) * cdiv(nx, tc.tx) * L;
        tc.zchunk = (int)std::min<int64_t>(
            nz, (o && o->zchunk > 0) ? (int64_t)o->zchunk : pick_zchunk(tiles, nz));
        pa.Ct = Ct;
        pa.alpha = alpha;
        pa.u = u;
        pa.ps = ps;
        pa.ft = ft;
        pa.ny = (int)ny;
        pa.nx = (int)nx;
        pa.nz = (int)nz;
        pa.nlab = (int)L;
        pa.cc = p->cc;
        pa.steps = p->steps;
        pa.zchunk = tc.zchunk;
        pa.ntile_y = (int)cdiv(ny, 64);
        pa.ntile_x = (int)cdiv(nx, tc.tx);
        pa.nzc = (int)cdiv(nz, tc.zchunk);
        pa.nblocks = pa.ntile_y * pa.ntile_x * pa.nzc;
        pa.fast_cc = (std::fabs(p->cc) >= 0x1p-16f && std::fabs(p->cc) <= 0x1p16f) ? 1 : 0;
        pa.flags = flags;
        pa.nyb = nyb;
        pa.pb = (int)ppb;
        grid = (int)cdiv((int64_t)pa.nblocks * L, 8) * 8;
        // errBound 0 (a fixed iteration count): iteration it's stop test runs in
        // the first workgroup of the next flow launch, and k_finalize only after
        // the last iteration
        const bool deferred = (p->err_bound == 0.0f);
        iter = [&, deferred](int32_t it) -> asets_status {
            pa.pt_in = ptb[it & 1];
            pa.pt_out = ptb[(it + 1) & 1];
            for (int k = 0; k < 3; ++k) {
                pa.b_in[k] = bb[it & 1][k];
                pa.b_out[k] = bb[(it + 1) & 1][k];
            }
            pa.stop = DeferredStop{};
            if (deferred && it > 0)
                pa.stop = DeferredStop{partials, nparts, (double)NL, it - 1, p->err_bound, cvg,
                                       it_num, flags};
            launch_potts_flow(is3d, tc.tx, pa, grid, s, blk);
            launch_potts_couple((int)L, pa.ft, u, ps, N, p->cc, pa.fast_cc, partials, flags, s);
            if (!deferred || it == max_iter - 1)
                k_finalize<<<1, 1024, 0, s>>>(partials, nparts, (double)NL, it, p->err_bound,
                                              cvg, it_num, flags);
            return ASETS_OK;
        };
    }
    HIPCHK(hipGetLastError());
    if (ev0) HIPCHK(hipEventRecord(ev0, s));
    int32_t itn = max_iter;
    st = run_loop(max_iter, p->err_bound, o ? o->poll_every : 0, s, dc.host_flag, flags, &itn, iter);
    if (st) return st;
    if (ev1) HIPCHK(hipEventRecord(ev1, s));
    HIPCHK(hipStreamSynchronize(s));
    if (ev0) {
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
        const int32_t n_exec = (itn < max_iter) ? itn + 1 : max_iter;
        *o->kernel_ms = n_exec > 0 ? (double)ms / n_exec : 0.0;
    }
    if (seconds)
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    return ASETS_OK;
}

// ---------------------------------------------------------------------------
// autotuner: tile shape of the fused kernel per shape bucket
// ---------------------------------------------------------------------------
// Candidates are timed on the caller's own data (a warm-up iteration and two
// timed iterations each, state re-initialised afterwards) and the fastest is
// cached in memory under a key made of the volume count, the total voxel count
// rounded to a quarter octave and the flow storage type.  With ASETS_TUNE_CACHE
// set to a file path the cache is also read from and appended to that file.
struct TuneCache {
    std::mutex mu;
    std::map<std::string, TileCfg> m;
    bool loaded = false;
};
TuneCache g_tune;

std::string tune_key(const char *kind, int32_t nvol, int64_t vox, int es) {
    int bucket = (int)std::lround(4.0 * std::log2((double)std::max<int64_t>(vox, 1)));
    char buf[128];
    snprintf(buf, sizeof buf, "%s:nvol=%d:vox4log2=%d:es=%d", kind, nvol, bucket, es);
    return buf;
}

// The tilings the batched solver runs: the three kernel tile widths, and z chunks of
// at least kBatchMinZchunk planes (its work list and partials are sized for that).
constexpr int kBatchMinZchunk = 8;
bool batched_cfg_ok(TileCfg c) {
    return (c.tx == 8 || c.tx == 16 || c.tx == 32) && c.zchunk >= kBatchMinZchunk &&
           c.zchunk <= (1 << 20);
}

bool tune_lookup(const std::string &key, TileCfg *out) {
    std::lock_guard<std::mutex> lk(g_tune.mu);
    const char *path = std::getenv("ASETS_TUNE_CACHE");
    if (!g_tune.loaded && path) {
        g_tune.loaded = true;
        if (FILE *f = std::fopen(path, "r")) {
            char k[256];
            int tx, zc;
            // a line a later build or a hand edit left with a tiling the batched kernel
            // does not have (tile widths 8, 16, 32; chunks of 8 planes or more) is dropped
            while (std::fscanf(f, "%255s %d %d", k, &tx, &zc) == 3)
                if (batched_cfg_ok(TileCfg{tx, zc})) g_tune.m[k] = TileCfg{tx, zc};
            std::fclose(f);
        }
    }
    auto it = g_tune.m.find(key);
    if (it == g_tune.m.end() || !batched_cfg_ok(it->second)) return false;
    *out = it->second;
    return true;
}

void tune_store(const std::string &key, TileCfg c) {
    std::lock_guard<std::mutex> lk(g_tune.mu);
    g_tune.m[key] = c;
    if (const char *path = std::getenv("ASETS_TUNE_CACHE")) {
        if (FILE *f = std::fopen(path, "a")) {
            std::fprintf(f, "%s %d %d\n", key.c_str(), c.tx, c.zchunk);
            std::fclose(f);
        }
    }
}

// ---------------------------------------------------------------------------
// batched binary 3D
// ---------------------------------------------------------------------------
template <int TX, int NT, typename T, bool BLK>
int batched_grid() {
    constexpr int W = (NT >= 512) ? 4 : 3;
    int dev = 0, cus = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per, kf_binary_batched<TX, NT, W, T, false, BLK>, NT, 0);
    return std::max(1, cus) * std::max(1, per);
}

template <int TX, int NT, int W, typename T, bool BLK>
void launch_batched_t(const BatchVol<T> *vols, const int2 *work, int nwork, const int32_t *vdone,
                      const int32_t *viter, double *partials, BatchCtl *ctl, float cc,
                      float steps, int fast_cc, bool first, int32_t *vbad, hipStream_t s) {
    const int g = std::min(nwork, batched_grid<TX, NT, T, BLK>());
    if (first)
        kf_binary_batched<TX, NT, W, T, true, BLK><<<g, NT, 0, s>>>(
            vols, work, nwork, vdone, viter, partials, ctl, cc, steps, fast_cc, vbad);
    else
        kf_binary_batched<TX, NT, W, T, false, BLK><<<g, NT, 0, s>>>(
            vols, work, nwork, vdone, viter, partials, ctl, cc, steps, fast_cc, vbad);
}

// first: the cold-start first iteration of every volume (checks alpha into vbad);
// blk: block-interleaved state buffers
template <typename T>
void launch_batched(int tx, const BatchVol<T> *vols, const int2 *work, int nwork,
                    const int32_t *vdone, const int32_t *viter, double *partials, BatchCtl *ctl,
                    float cc, float steps, int fast_cc, bool first, int32_t *vbad,
                    hipStream_t s, bool blk) {
    if (blk) {
        if (tx == 8)
            launch_batched_t<8, 256, 3, T, true>(vols, work, nwork, vdone, viter, partials, ctl,
                                                 cc, steps, fast_cc, first, vbad, s);
        else if (tx == 32)
            launch_batched_t<32, 1024, 4, T, true>(vols, work, nwork, vdone, viter, partials,
                                                   ctl, cc, steps, fast_cc, first, vbad, s);
        else
            launch_batched_t<16, 512, 4, T, true>(vols, work, nwork, vdone, viter, partials,
                                                  ctl, cc, steps, fast_cc, first, vbad, s);
        return;
    }
    if (tx == 8)
        launch_batched_t<8, 256, 3, T, false>(vols, work, nwork, vdone, viter, partials, ctl, cc,
                                              steps, fast_cc, first, vbad, s);
    else if (tx == 32)
        launch_batched_t<32, 1024, 4, T, false>(vols, work, nwork, vdone, viter, partials, ctl,
                                                cc, steps, fast_cc, first, vbad, s);
    else
        launch_batched_t<16, 512, 4, T, false>(vols, work, nwork, vdone, viter, partials, ctl,
                                               cc, steps, fast_cc, first, vbad, s);
}

// One workgroup per volume: its stop flag, iteration counter, cvg (zeros) and
// it_num (max_iter, the value when the stop test never fires); the control words.
__global__ void k_batch_reset(const BatchMeta *meta, int32_t *vdone, int32_t *viter,
                              int32_t *vbad, int nvol, int32_t max_iter, BatchCtl *ctl) {
    const int v = blockIdx.x;
    const BatchMeta m = meta[v];
    for (int32_t i = threadIdx.x; i < max_iter; i += blockDim.x) m.cvg[i] = 0.0f;
    if (threadIdx.x == 0) {
        *m.it_num = max_iter;
        vdone[v] = (max_iter <= 0) ? 1 : 0;
        viter[v] = 0;
        vbad[v] = 0;
        if (v == 0) {
            ctl->remaining = (max_iter <= 0) ? 0 : nvol;
            ctl->bad_alpha = 0;
            ctl->work_next = 0;
        }
    }
}

// u of every volume from the state buffer its last iteration wrote ((iterations
// executed) & 1) to the caller's array; blockIdx.y = volume.  BLK: the buffers are
// block-interleaved (one thread per lane of a block; 32-bit indices, as the
// layout's buffers stay below 2^31 bytes).
template <typename T, bool BLK>
__global__ void k_batch_take_u(const BatchVol<T> *vols, const BatchMeta *meta, int32_t max_iter) {
    const int v = blockIdx.y;
    const BatchMeta m = meta[v];
    const int32_t itn = *m.it_num;
    const int32_t n_exec = (itn < max_iter) ? itn + 1 : max_iter;
    const float *src = vols[v].u[n_exec & 1];
    if constexpr (BLK) {
        const BatchVol<T> &V = vols[v];
        const int nyb = V.nyb, ny = V.ny;
        const int lanes = V.nz * V.nx * nyb * 64;
        for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < lanes;
             t += gridDim.x * blockDim.x) {
            const int b = t >> 6, ln = t & 63;
            const int yb = b % nyb, xz = b / nyb;
            const int y = yb * 64 + ln;
            if (y >= ny) continue;
            const int64_t g = (int64_t)xz * ny + y;
            m.u_out[g] = ((const float *)((const char *)src + (size_t)b *
                                          BlkLayout<T>::kBytes + BlkLayout<T>::kU))[ln];
        }
        return;
    }
    const int64_t N = (int64_t)m.n;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * blockDim.x)
        m.u_out[i] = src[i];
}

template <typename T>
asets_status batched_dev(int32_t nvol, const asets_volume *vols, const asets_params *p,
                         double *seconds, const asets_opts *o) {
    auto t_start = std::chrono::steady_clock::now();
    if (nvol <= 0 || !vols) return fail(ASETS_EBADARG, "need at least one volume");
    if (nvol > 65535) return fail(ASETS_EBADARG, "at most 65535 volumes per call (got %d)", nvol);
    asets_status st = check_params(p);
    if (st) return st;
    int64_t total = 0;
    for (int v = 0; v < nvol; ++v) {
        const asets_volume &V = vols[v];
        st = check_dims(V.ny, V.nx, V.nz, 1);
        if (st) return fail(st, "volume %d: %s", v, asets_last_error());
        if (!V.Cs || !V.Ct || !V.alpha || !V.u || !V.it_num || (p->max_iter > 0 && !V.cvg))
            return fail(ASETS_EBADARG, "volume %d: null array", v);
        if (!fused_fits(V.ny, V.nx, V.nz))
            return fail(ASETS_ETOOBIG, "volume %d too large for the fused kernel", v);
        total += V.ny * V.nx * V.nz;
    }
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    DeviceCache &dc = g_cache[dev];
    std::lock_guard<std::mutex> lock(dc.mu);
    if (!dc.host_flag) HIPCHK(hipHostMalloc((void **)&dc.host_flag, 64));
    hipStream_t us = o ? (hipStream_t)o->stream : nullptr;
    hipStream_t s = us;
    hipStream_t own = nullptr;
    if (!s) {   // graphs cannot be captured on the null stream
        HIPCHK(hipStreamCreateWithFlags(&own, hipStreamNonBlocking));
        s = own;
    }
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() {
            if (s) (void)hipStreamDestroy(s);
        }
    } sg{own};
    const int32_t max_iter = p->max_iter;
    const int fast_cc = (std::fabs(p->cc) >= 0x1p-16f && std::fabs(p->cc) <= 0x1p16f) ? 1 : 0;

    // work-list capacity: the finest tiling any source may pick (tx 8, zchunk 8); an
    // explicit tiling outside what batched_cfg_ok allows is rejected here, and
    // set_tiling checks the count it builds against the capacity
    if (o && o->tile_x > 0 && o->tile_x != 8 && o->tile_x != 16 && o->tile_x != 32)
        return fail(ASETS_EBADARG, "tile_x must be 8, 16 or 32 (got %d)", o->tile_x);
    if (o && o->zchunk > 0 && o->zchunk < kBatchMinZchunk)
        return fail(ASETS_EBADARG, "the batched solver needs zchunk >= %d (got %d)",
                    kBatchMinZchunk, o->zchunk);
    int64_t max_items = 0;
    for (int v = 0; v < nvol; ++v)
        max_items += cdiv(vols[v].ny, 64) * cdiv(vols[v].nx, 8) *
                     cdiv(vols[v].nz, kBatchMinZchunk);
    size_t need = Carver::sz(1, sizeof(BatchCtl)) + Carver::sz(nvol, sizeof(BatchVol<T>)) +
                  Carver::sz(nvol, sizeof(BatchMeta)) + 2 * Carver::sz(nvol, 4) +
                  Carver::sz(max_items, sizeof(int2)) + Carver::sz(max_items, sizeof(double)) +
                  Carver::sz(nvol, 4);
    // block-interleaved state buffers when every volume's fit (fused_fits_blk), else planar
    bool blk = true;
    for (int v = 0; v < nvol; ++v)
        blk = blk && fused_fits_blk(vols[v].ny, vols[v].nx, vols[v].nz, sizeof(T));
    for (int v = 0; v < nvol; ++v) {
        const asets_volume &V = vols[v];
        need += blk ? 2 * Carver::sz(V.nz * blk_plane_bytes(V.ny, V.nx, sizeof(T)), 1)
                    : fused_bytes(V.ny * V.nx * V.nz, sizeof(T));
    }
    st = ensure(&dc.dev, &dc.dev_bytes, need);
    if (st) return st;
    Carver cv(dc.dev);
    BatchCtl *ctl = cv.take<BatchCtl>(1);
    BatchVol<T> *d_vols = cv.take<BatchVol<T>>(nvol);
    BatchMeta *d_meta = cv.take<BatchMeta>(nvol);
    int32_t *vdone = cv.take<int32_t>(nvol);
    int32_t *viter = cv.take<int32_t>(nvol);
    int2 *d_work = cv.take<int2>(max_items);
    double *partials = cv.take<double>(max_items);
    int32_t *vbad = cv.take<int32_t>(nvol);   // alpha <= 0 somewhere in volume v

    std::vector<BatchVol<T>> hv(nvol);
    std::vector<BatchMeta> hm(nvol);
    for (int v = 0; v < nvol; ++v) {
        const asets_volume &V = vols[v];
        const int64_t N = V.ny * V.nx * V.nz;
        BatchVol<T> &b = hv[v];
        b.Cs = V.Cs;
        b.Ct = V.Ct;
        b.alpha = V.alpha;
        b.nyb = (int)cdiv(V.ny, 64);
        b.pb = (int)blk_plane_bytes(V.ny, V.nx, sizeof(T));
        for (int k = 0; k < 2; ++k) {
            if (blk) {   // one block-interleaved array; every pointer is its base
                char *base = cv.take<char>(V.nz * (int64_t)b.pb);
                b.u[k] = (float *)base;
                b.ps[k] = b.pt[k] = b.bx[k] = b.by[k] = b.bz[k] = (T *)base;
                continue;
            }
            b.u[k] = cv.take<float>(N);
            b.ps[k] = cv.take<T>(kPair * N);
            b.pt[k] = b.ps[k] + 1;
            b.bx[k] = cv.take<T>(kPair * N);
            b.by[k] = b.bx[k] + 1;
            b.bz[k] = cv.take<T>(N);
        }
        b.ny = (int)V.ny;
        b.nx = (int)V.nx;
        b.nz = (int)V.nz;
        hm[v].n = (double)N;
        hm[v].cvg = V.cvg;
        hm[v].it_num = V.it_num;
        hm[v].u_out = V.u;
    }

    // reset the control words and (with_state) every volume's initial state; a solve
    // that iterates starts with the cold-start iteration instead
    // (needs d_meta: after set_tiling)
    auto init_all = [&](bool with_state) -> asets_status {
        for (int v = 0; with_state && v < nvol; ++v) {
            const BatchVol<T> &b = hv[v];
            const int64_t N = (int64_t)b.ny * b.nx * b.nz;
            const unsigned nb = (unsigned)cdiv(N, kBlock);
            // buffer 1 needs none: a volume's first iteration writes every voxel of it
            if (blk) {
                const BlkColGrid cg = blk_col_grid(b.nx, b.nz, b.nyb);
                k_blk_init<T><<<cg.grid, cg.threads, 0, s>>>(b.Cs, b.Ct, nullptr, nullptr,
                                                             nullptr, nullptr, nullptr, nullptr,
                                                             (char *)b.u[0], b.ny, b.nx, b.nz,
                                                             b.nyb, 0, nullptr, nullptr,
                                                             nullptr);
            } else
                k_init_binary_cold_t<T><<<nb, kBlock, 0, s>>>(b.Cs, b.Ct, b.u[0], b.ps[0],
                                                             b.bx[0], b.bz[0], N);
        }
        k_batch_reset<<<(unsigned)nvol, 256, 0, s>>>(d_meta, vdone, viter, vbad, nvol, max_iter,
                                                     ctl);
        HIPCHK(hipGetLastError());
        return ASETS_OK;
    };

    // tiling -> work list
    int nwork = 0;
    auto set_tiling = [&](TileCfg c) -> asets_status {
        std::vector<int2> w;
        for (int v = 0; v < nvol; ++v) {
            BatchVol<T> &b = hv[v];
            b.ntile_y = (int)cdiv(b.ny, 64);
            b.ntile_x = (int)cdiv(b.nx, c.tx);
            b.zchunk = std::min(c.zchunk, b.nz);
            b.nzc = (int)cdiv(b.nz, b.zchunk);
            b.nblocks = b.ntile_y * b.ntile_x * b.nzc;
            b.work0 = (int)w.size();
            hm[v].work0 = b.work0;
            hm[v].nblocks = b.nblocks;
            for (int lb = 0; lb < b.nblocks; ++lb) w.push_back(make_int2(v, lb));
        }
        if ((int64_t)w.size() > max_items)
            return fail(ASETS_EBADARG, "tiling (%d, %d) exceeds the work-list capacity", c.tx,
                        c.zchunk);
        nwork = (int)w.size();
        HIPCHK(hipMemcpyAsync(d_work, w.data(), w.size() * sizeof(int2), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d_vols, hv.data(), nvol * sizeof(BatchVol<T>), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d_meta, hm.data(), nvol * sizeof(BatchMeta), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        return ASETS_OK;
    };
    auto one_iter = [&](int tx, bool first = false) {
        launch_batched<T>(tx, d_vols, d_work, nwork, vdone, viter, partials, ctl, p->cc, p->steps,
                          fast_cc, first, vbad, s, blk);
        k_finalize_batched<<<nvol, 256, 0, s>>>(d_meta, vdone, viter, partials, p->err_bound,
                                                max_iter, ctl);
    };

    // tile shape: options, tuning cache, or tune now
    TileCfg cfg{16, 32};
    const bool fixed = o && o->tile_x > 0;
    if (fixed) {
        cfg.tx = o->tile_x;
        cfg.zchunk = o->zchunk > 0 ? o->zchunk : 32;
    } else if (max_iter > 0) {
        const std::string key = tune_key("batched3d", nvol, total, (int)sizeof(T));
        if (!tune_lookup(key, &cfg)) {
            float best = 1e30f;
            EventPair tp;
            hipEvent_t &e0 = tp.a, &e1 = tp.b;
            HIPCHK(hipEventCreate(&e0));
            HIPCHK(hipEventCreate(&e1));
            for (int tx : {8, 16, 32})
                for (int zc : {8, 16, 32, 64}) {
                    st = set_tiling(TileCfg{tx, zc});
                    if (st) return st;
                    st = init_all(true);
                    if (st) return st;
                    one_iter(tx);   // warm-up
                    HIPCHK(hipEventRecord(e0, s));
                    one_iter(tx);
                    one_iter(tx);
                    HIPCHK(hipEventRecord(e1, s));
                    HIPCHK(hipEventSynchronize(e1));
                    float ms = 0.0f;
                    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                    if (ms < best) {
                        best = ms;
                        cfg = TileCfg{tx, zc};
                    }
                }
            tune_store(key, cfg);
        }
    }
    st = set_tiling(cfg);
    if (st) return st;
    st = init_all(max_iter == 0);
    if (st) return st;

    EventPair evp;
    hipEvent_t &ev0 = evp.a, &ev1 = evp.b;
    if (o && o->kernel_ms) {
        HIPCHK(hipEventCreate(&ev0));
        HIPCHK(hipEventCreate(&ev1));
        HIPCHK(hipEventRecord(ev0, s));
    }
    // the cold-start iteration, then K iterations captured once as a graph, replayed
    // until every volume stopped; each poll reads (remaining, bad_alpha)
    if (max_iter > 0) {
        one_iter(cfg.tx, true);
        int32_t done_iters = 1;
        const int K = std::max(1, std::min(8, (int)max_iter - 1));
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        if (max_iter > 1) {
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < K; ++k) one_iter(cfg.tx);
            HIPCHK(hipStreamEndCapture(s, &graph));
            HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        }
        asets_status lst = ASETS_OK;
        for (;;) {
            if (done_iters < max_iter) {
                if (hipGraphLaunch(exec, s) != hipSuccess) {
                    lst = fail(ASETS_EDEVICE, "hipGraphLaunch failed");
                    break;
                }
                done_iters += K;
            }
            if (hipMemcpyAsync(dc.host_flag, &ctl->remaining, 2 * sizeof(int32_t),
                               hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                lst = fail(ASETS_EDEVICE, "stream error in the batched loop");
                break;
            }
            if (dc.host_flag[1]) {   // alpha <= 0: name the first such volume
                std::vector<int32_t> bad(nvol);
                if (hipMemcpy(bad.data(), vbad, nvol * sizeof(int32_t), hipMemcpyDeviceToHost) !=
                    hipSuccess) {
                    lst = fail(ASETS_EDEVICE, "stream error in the batched loop");
                    break;
                }
                int v = 0;
                while (v < nvol - 1 && !bad[v]) ++v;
                lst = fail(ASETS_EALPHA, "volume %d: alpha(x) must be positive. Exiting...", v);
                break;
            }
            if (dc.host_flag[0] <= 0 || done_iters >= max_iter) break;
        }
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (lst) return lst;
    }
    if (ev1) HIPCHK(hipEventRecord(ev1, s));
    // u of each volume lives in buffer (iterations executed) & 1: one copy kernel
    // reads each volume's it_num on the device
    if (blk) k_batch_take_u<T, true><<<dim3(256, (unsigned)nvol), 256, 0, s>>>(d_vols, d_meta,
                                                                               max_iter);
    else k_batch_take_u<T, false><<<dim3(256, (unsigned)nvol), 256, 0, s>>>(d_vols, d_meta,
                                                                             max_iter);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    if (ev0) {
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
        *o->kernel_ms = ms;   // whole batched loop (ms)
    }
    if (seconds)
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    return ASETS_OK;
}

// ---------------------------------------------------------------------------
// host-pointer wrappers (the C-MEX call surface)
// ---------------------------------------------------------------------------
struct HostIo {
    float *d_in[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    float *d_u = nullptr, *d_cvg = nullptr;
    int32_t *d_it = nullptr;
};

// Device staging for host arrays; a separate cache from the solver workspace.
// n_in[i]: element count of input i.
asets_status stage(int nin, const float *const *h_in, const int64_t *n_in, int64_t n_u,
                   int32_t max_iter, HostIo &io, hipStream_t s) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    DeviceCache &dc = g_cache[dev];
    size_t need = Carver::sz(n_u, 4) + Carver::sz(std::max(max_iter, 1), 4) + Carver::sz(1, 4);
    for (int i = 0; i < nin; ++i) need += Carver::sz(n_in[i], 4);
    asets_status st = ensure(&dc.io, &dc.io_bytes, need);
    if (st) return st;
    Carver cv(dc.io);
    for (int i = 0; i < nin; ++i) {
        io.d_in[i] = cv.take<float>(n_in[i]);
        HIPCHK(hipMemcpyAsync(io.d_in[i], h_in[i], n_in[i] * 4, hipMemcpyHostToDevice, s));
    }
    io.d_u = cv.take<float>(n_u);
    io.d_cvg = cv.take<float>(std::max(max_iter, 1));
    io.d_it = cv.take<int32_t>(1);
    return ASETS_OK;
}

asets_status unstage(const HostIo &io, int64_t n_u, int32_t max_iter, float *u, float *cvg,
                     int32_t *it_num, hipStream_t s) {
    HIPCHK(hipMemcpyAsync(u, io.d_u, n_u * 4, hipMemcpyDeviceToHost, s));
    if (max_iter > 0 && cvg)
        HIPCHK(hipMemcpyAsync(cvg, io.d_cvg, (size_t)max_iter * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(it_num, io.d_it, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return ASETS_OK;
}

asets_status binary_host(int ndim, const float *Cs, const float *Ct, const float *alpha,
                         int64_t ny, int64_t nx, int64_t nz, const asets_params *p,
                         const float *u0, const float *pt0, float *u, float *cvg,
                         int32_t *it_num, double *seconds, const asets_opts *o) {
    auto t0 = std::chrono::steady_clock::now();
    asets_status st = check_params(p);
    if (st) return st;
    st = check_dims(ny, nx, nz, 1);
    if (st) return st;
    if (!Cs || !Ct || !alpha || !u || !it_num || (p->max_iter > 0 && !cvg))
        return fail(ASETS_EBADARG, "null array argument");
    if ((u0 == nullptr) != (pt0 == nullptr))
        return fail(ASETS_EBADARG, "warm start needs both u0 and pt0");
    const int64_t N = ny * nx * nz;
    hipStream_t s = o ? (hipStream_t)o->stream : nullptr;
    const float *in[5] = {Cs, Ct, alpha, u0, pt0};
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> io_lock(g_cache[dev].io_mu);
    HostIo io;
    const int64_t sizes[5] = {N, N, N, N, N};
    st = stage(u0 ? 5 : 3, in, sizes, N, p->max_iter, io, s);
    if (st) return st;
    WarmIn w;
    w.u0 = io.d_in[3];
    w.pt0 = io.d_in[4];
    st = binary_dev(ndim, io.d_in[0], io.d_in[1], io.d_in[2], ny, nx, nz, p, w, io.d_u,
                    io.d_cvg, io.d_it, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, o);
    if (st) return st;
    st = unstage(io, N, p->max_iter, u, cvg, it_num, s);
    if (st) return st;
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return ASETS_OK;
}

asets_status potts_host(int ndim, const float *Ct, const float *alpha, int64_t ny, int64_t nx,
                        int64_t nz, int64_t L, const asets_params *p, float *u, float *cvg,
                        int32_t *it_num, double *seconds, const asets_opts *o) {
    auto t0 = std::chrono::steady_clock::now();
    asets_status st = check_params(p);
    if (st) return st;
    st = check_dims(ny, nx, nz, L);
    if (st) return st;
    if (!Ct || !alpha || !u || !it_num || (p->max_iter > 0 && !cvg))
        return fail(ASETS_EBADARG, "null array argument");
    const int64_t NL = ny * nx * nz * L;
    hipStream_t s = o ? (hipStream_t)o->stream : nullptr;
    const float *in[2] = {Ct, alpha};
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> io_lock(g_cache[dev].io_mu);
    HostIo io;
    const int64_t sizes[2] = {NL, NL};
    st = stage(2, in, sizes, NL, p->max_iter, io, s);
    if (st) return st;
    st = potts_dev(ndim, io.d_in[0], io.d_in[1], ny, nx, nz, L, p, io.d_u, io.d_cvg, io.d_it,
                   nullptr, o);
    if (st) return st;
    st = unstage(io, NL, p->max_iter, u, cvg, it_num, s);
    if (st) return st;
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return ASETS_OK;
}

asets_status ishikawa_dev(int ndim, const float *Ct, const float *alpha, int64_t ny,
                          int64_t nx, int64_t nz, int64_t L, const asets_params *p, float *u,
                          float *cvg, int32_t *it_num, double *seconds, const asets_opts *o);

asets_status ishikawa_host(int ndim, const float *Ct, const float *alpha, int64_t ny,
                           int64_t nx, int64_t nz, int64_t L, const asets_params *p, float *u,
                           float *cvg, int32_t *it_num, double *seconds, const asets_opts *o) {
    auto t0 = std::chrono::steady_clock::now();
    asets_status st = check_params(p);
    if (st) return st;
    st = check_dims(ny, nx, nz, L);
    if (st) return st;
    if (L < 2) return fail(ASETS_EBADARG, "the Ishikawa model needs at least 2 labels");
    if (!Ct || !alpha || !u || !it_num || (p->max_iter > 0 && !cvg))
        return fail(ASETS_EBADARG, "null array argument");
    const int64_t N = ny * nx * nz;
    hipStream_t s = o ? (hipStream_t)o->stream : nullptr;
    const float *in[2] = {Ct, alpha};
    const int64_t sizes[2] = {N * L, N * (L - 1)};
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> io_lock(g_cache[dev].io_mu);
    HostIo io;
    st = stage(2, in, sizes, N * (L - 1), p->max_iter, io, s);
    if (st) return st;
    st = ishikawa_dev(ndim, io.d_in[0], io.d_in[1], ny, nx, nz, L, p, io.d_u, io.d_cvg, io.d_it,
                      nullptr, o);
    if (st) return st;
    st = unstage(io, N * (L - 1), p->max_iter, u, cvg, it_num, s);
    if (st) return st;
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return ASETS_OK;
}

// ---------------------------------------------------------------------------
// Ishikawa model (linearly ordered labels) on device memory
// ---------------------------------------------------------------------------
asets_status ishikawa_dev(int ndim, const float *Ct, const float *alpha, int64_t ny,
                          int64_t nx, int64_t nz, int64_t L, const asets_params *p, float *u,
                          float *cvg, int32_t *it_num, double *seconds, const asets_opts *o) {
    auto t_start = std::chrono::steady_clock::now();
    if (ndim != 2 && ndim != 3) return fail(ASETS_EBADARG, "ndim must be 2 or 3");
    if (ndim == 2 && nz != 1) return fail(ASETS_EBADARG, "2D solve needs nz == 1");
    asets_status st = check_params(p);
    if (st) return st;
    st = check_dims(ny, nx, nz, L);
    if (st) return st;
    if (L < 2) return fail(ASETS_EBADARG, "the Ishikawa model needs at least 2 labels");
    if (!Ct || !alpha || !u || !it_num || (p->max_iter > 0 && !cvg))
        return fail(ASETS_EBADARG, "null array argument");
    if (o && o->flow_dtype != ASETS_FLOW_F32)
        return fail(ASETS_EBADARG, "fp16 flow storage is available for the binary model only");
    if (o && o->kernel == ASETS_KERNEL_PASSES)
        return fail(ASETS_EBADARG, "the Ishikawa model runs on the fused kernel only");
    if (!fused_fits(ny, nx, nz))
        return fail(ASETS_ETOOBIG, "volume too large for the fused kernel (fields of 2^31 "
                                   "bytes or more)");

    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(ASETS_EDEVICE, "device id %d", dev);
    DeviceCache &dc = g_cache[dev];
    std::lock_guard<std::mutex> lock(dc.mu);
    hipStream_t s = o ? (hipStream_t)o->stream : nullptr;
    if (!dc.host_flag) HIPCHK(hipHostMalloc((void **)&dc.host_flag, 64));

    const bool is3d = (ndim == 3);
    const int64_t N = ny * nx * nz, M = L - 1, NM = N * M;
    const int32_t max_iter = p->max_iter;
    const int64_t nparts = cdiv(N, kBlock);
    const size_t need = Carver::sz(1, sizeof(Flags)) + Carver::sz(nparts, sizeof(double)) +
                        Carver::sz(1, sizeof(int32_t)) + Carver::sz(N * L, 4) +
                        7 * Carver::sz(NM, 4);
    st = ensure(&dc.dev, &dc.dev_bytes, need);
    if (st) return st;
    Carver cv(dc.dev);
    Flags *flags = cv.take<Flags>(1);
    double *partials = cv.take<double>(nparts);
    float *pt = cv.take<float>(N * L);
    float *dv = cv.take<float>(NM);
    float *bb[2][3];
    for (int b = 0; b < 2; ++b)
        for (int k = 0; k < 3; ++k) bb[b][k] = cv.take<float>(NM);

    solve_setup(flags, it_num, max_iter, cvg, s);
    // only the first L-1 alpha fields are used (asetsIshikawa3D_mex.c:377-379)
    if (max_iter > 0) alpha_check_async(alpha, NM, flags, s);
    for (int b = 0; b < 2; ++b)
        for (int k = 0; k < 3; ++k) fill(bb[b][k], 0.0f, NM, s);
    k_init_ishikawa<<<(unsigned)cdiv(N, kBlock), kBlock, 0, s>>>(Ct, u, pt, N, (int)L);
    HIPCHK(hipGetLastError());

    EventPair evp;
    hipEvent_t &ev0 = evp.a, &ev1 = evp.b;
    if (o && o->kernel_ms) {
        HIPCHK(hipEventCreate(&ev0));
        HIPCHK(hipEventCreate(&ev1));
    }
    TileCfg tc{};
    tc.tx = (o && o->tile_x > 0) ? o->tile_x : 16;
    if (tc.tx != 8 && tc.tx != 32) tc.tx = 16;
    const int64_t tiles = cdiv(ny, 64) * cdiv(nx, tc.tx) * M;
    tc.zchunk = (int)std::min<int64_t>(
        nz, (o && o->zchunk > 0) ? (int64_t)o->zchunk : pick_zchunk(tiles, nz));
    IshiArgs ia{};
    ia.alpha = alpha;
    ia.u = u;
    ia.pt = pt;
    ia.dv = dv;
    ia.ny = (int)ny;
    ia.nx = (int)nx;
    ia.nz = (int)nz;
    ia.nflow = (int)M;
    ia.cc = p->cc;
    ia.steps = p->steps;
    ia.zchunk = tc.zchunk;
    ia.ntile_y = (int)cdiv(ny, 64);
    ia.ntile_x = (int)cdiv(nx, tc.tx);
    ia.nzc = (int)cdiv(nz, tc.zchunk);
    ia.nblocks = ia.ntile_y * ia.ntile_x * ia.nzc;
    ia.fast_cc = (std::fabs(p->cc) >= 0x1p-16f && std::fabs(p->cc) <= 0x1p16f) ? 1 : 0;
    ia.flags = flags;
    const int grid = (int)cdiv((int64_t)ia.nblocks * M, 8) * 8;
    // errBound 0 (a fixed iteration count): iteration it's stop test runs in the
    // first workgroup of the next flow launch, and k_finalize only after the last
    // iteration.  The deferred test stops on cvg <= bound; 2D's strict cvg < 0
    // is cvg <= the largest float below 0 (never, as cvg >= 0).
    const bool deferred = (p->err_bound == 0.0f);
    const float deferred_bound = is3d ? p->err_bound : std::nextafter(0.0f, -1.0f);
    auto iter = [&](int32_t it) -> asets_status {
        for (int k = 0; k < 3; ++k) {
            ia.b_in[k] = bb[it & 1][k];
            ia.b_out[k] = bb[(it + 1) & 1][k];
        }
        ia.stop = DeferredStop{};
        if (deferred && it > 0)
            ia.stop = DeferredStop{partials, nparts, (double)NM, it - 1, deferred_bound, cvg,
                                   it_num, flags};
        launch_ishikawa_flow(is3d, tc.tx, ia, grid, s);
        launch_ishikawa_couple((int)L, dv, pt, u, Ct, N, p->cc, ia.fast_cc, partials, flags, s);
        if (deferred && it != max_iter - 1) return ASETS_OK;
        // 3D stops on cvg <= errBound (asetsIshikawa3D_mex.c:213), 2D on < (2D :205)
        if (is3d)
            k_finalize<false><<<1, 1024, 0, s>>>(partials, nparts, (double)NM, it,
                                                 p->err_bound, cvg, it_num, flags);
        else
            k_finalize<true><<<1, 1024, 0, s>>>(partials, nparts, (double)NM, it,
                                                p->err_bound, cvg, it_num, flags);
        return ASETS_OK;
    };
    if (ev0) HIPCHK(hipEventRecord(ev0, s));
    int32_t itn = max_iter;
    st = run_loop(max_iter, p->err_bound, o ? o->poll_every : 0, s, dc.host_flag, flags, &itn, iter);
    if (st) return st;
    if (ev1) HIPCHK(hipEventRecord(ev1, s));
    HIPCHK(hipStreamSynchronize(s));
    if (ev0) {
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
        const int32_t n_exec = (itn < max_iter) ? itn + 1 : max_iter;
        *o->kernel_ms = n_exec > 0 ? (double)ms / n_exec : 0.0;
    }
    if (seconds)
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    return ASETS_OK;
}

// ---------------------------------------------------------------------------
// FP64 energy of a labelling (binary: Cs != nullptr, nlab == 1; Potts: Cs null)
// ---------------------------------------------------------------------------
asets_status energy_dev(int ndim, const float *Cs, const float *Ct, const float *alpha,
                        int64_t ny, int64_t nx, int64_t nz, int64_t nlab, const float *u,
                        double *out, const asets_opts *o) {
    if (ndim != 2 && ndim != 3) return fail(ASETS_EBADARG, "ndim must be 2 or 3");
    if (ndim == 2 && nz != 1) return fail(ASETS_EBADARG, "2D energy needs nz == 1");
    asets_status st = check_dims(ny, nx, nz, nlab);
    if (st) return st;
    if (!Ct || !alpha || !u || !out) return fail(ASETS_EBADARG, "null array argument");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(ASETS_EDEVICE, "device id %d", dev);
    DeviceCache &dc = g_cache[dev];
    std::lock_guard<std::mutex> lock(dc.mu);
    hipStream_t s = o ? (hipStream_t)o->stream : nullptr;
    if (!dc.host_flag) HIPCHK(hipHostMalloc((void **)&dc.host_flag, 64));
    const size_t need = Carver::sz(2 * kEnergyBlocks, 8) + Carver::sz(3, 8);
    st = ensure(&dc.dev, &dc.dev_bytes, need);
    if (st) return st;
    Carver cv(dc.dev);
    double *partials = cv.take<double>(2 * kEnergyBlocks);
    double *res = cv.take<double>(3);
    if (ndim == 3)
        k_energy<true><<<kEnergyBlocks, kEnergyThreads, 0, s>>>(Cs, Ct, alpha, u, (int)ny, (int)nx,
                                                                (int)nz, (int)nlab, partials);
    else
        k_energy<false><<<kEnergyBlocks, kEnergyThreads, 0, s>>>(Cs, Ct, alpha, u, (int)ny,
                                                                 (int)nx, 1, (int)nlab, partials);
    k_energy_finalize<<<1, 256, 0, s>>>(partials, kEnergyBlocks, res);
    HIPCHK(hipGetLastError());
    double *h = (double *)((char *)dc.host_flag + 32);   // 3 doubles of the pinned slot
    HIPCHK(hipMemcpyAsync(h, res, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    out[0] = h[0];
    out[1] = h[1];
    out[2] = h[2];
    return ASETS_OK;
}

// FP64 primal-dual bounds of a binary state (kernels_bounds.h): out = {primal
// P_lambda(u) after `sweeps` red-black sweeps, the voxel-centred primal bound, dual
// D(p~)}.  Scratch: one double per voxel (the feasibility factors, then lambda).
asets_status bounds_dev(int ndim, const float *Cs, const float *Ct, const float *alpha,
                        const float *u, const float *bx, const float *by, const float *bz,
                        int64_t ny, int64_t nx, int64_t nz, int32_t sweeps, double *out,
                        const asets_opts *o) {
    if (ndim != 2 && ndim != 3) return fail(ASETS_EBADARG, "ndim must be 2 or 3");
    if (ndim == 2 && nz != 1) return fail(ASETS_EBADARG, "2D bounds need nz == 1");
    asets_status st = check_dims(ny, nx, nz, 1);
    if (st) return st;
    if (!Cs || !Ct || !alpha || !u || !bx || !by || (ndim == 3 && !bz) || !out)
        return fail(ASETS_EBADARG, "null array argument");
    if (sweeps < 0 || sweeps > 100000) return fail(ASETS_EBADARG, "sweeps must be in [0, 1e5]");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(ASETS_EDEVICE, "device id %d", dev);
    DeviceCache &dc = g_cache[dev];
    std::lock_guard<std::mutex> lock(dc.mu);
    hipStream_t s = o ? (hipStream_t)o->stream : nullptr;
    if (!dc.host_flag) HIPCHK(hipHostMalloc((void **)&dc.host_flag, 64));
    const BGeo G{ny, nx, nz, ny, ny * nx, ny * nx * nz};
    const int nb = kBoundsBlocks;
    const size_t need = Carver::sz(G.n, 8) + Carver::sz(4 * nb, 8) + Carver::sz(4, 8);
    st = ensure(&dc.dev, &dc.dev_bytes, need);
    if (st) return st;
    Carver cv(dc.dev);
    double *sl = cv.take<double>(G.n);   // feasibility factors, then lambda
    double *part = cv.take<double>(4 * nb);
    double *res = cv.take<double>(4);
    const float *fz = ndim == 3 ? bz : nullptr;
    const unsigned gv = (unsigned)std::min<int64_t>(cdiv(G.n, 256), 65536);
    k_bnd_scale<<<gv, 256, 0, s>>>(G, alpha, bx, by, fz, sl);
    k_bnd_dual<<<nb, kBoundsThreads, 0, s>>>(G, Cs, Ct, bx, by, fz, sl, part);
    k_bnd_init<<<nb, kBoundsThreads, 0, s>>>(G, Cs, Ct, alpha, u, sl, part + nb);
    k_bnd_F<<<nb, kBoundsThreads, 0, s>>>(G, alpha, u, sl, part + 2 * nb);
    for (int32_t i = 0; i < sweeps; ++i) k_bnd_sweep<<<gv, 256, 0, s>>>(G, alpha, u, sl, i & 1);
    k_bnd_F<<<nb, kBoundsThreads, 0, s>>>(G, alpha, u, sl, part + 3 * nb);
    k_bnd_sum<<<1, 256, 0, s>>>(part, nb, 4, res);
    HIPCHK(hipGetLastError());
    double *h = (double *)dc.host_flag + 4;   // 4 doubles of the pinned slot
    HIPCHK(hipMemcpyAsync(h, res, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    out[0] = h[1] + h[3];   // data + F(lambda after the sweeps)
    out[1] = h[1] + h[2];   // data + F(voxel-centred lambda)
    out[2] = h[0];          // dual
    return ASETS_OK;
}

// ---------------------------------------------------------------------------
// costs, weights and labels on device memory
// ---------------------------------------------------------------------------
struct PrepCtx {
    DeviceCache *dc = nullptr;
    hipStream_t s = nullptr;
};

asets_status prep_begin(const asets_opts *o, size_t need, PrepCtx &c,
                        std::unique_lock<std::mutex> &lock) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(ASETS_EDEVICE, "device id %d", dev);
    c.dc = &g_cache[dev];
    lock = std::unique_lock<std::mutex>(c.dc->mu);
    c.s = o ? (hipStream_t)o->stream : nullptr;
    if (!c.dc->host_flag) HIPCHK(hipHostMalloc((void **)&c.dc->host_flag, 64));
    return ensure(&c.dc->dev, &c.dc->dev_bytes, need);
}

// min and max of x into mm[2] (device); the values are copied to host[2]
asets_status minmax(const float *x, int64_t n, float *part, float *mm, float *host, hipStream_t s,
                    DeviceCache &dc) {
    const int nb = (int)std::min<int64_t>(kPrepBlocks, cdiv(n, 256));
    k_minmax_part<<<nb, 256, 0, s>>>(x, n, part);
    k_minmax_final<<<1, 64, 0, s>>>(part, nb, mm);
    HIPCHK(hipGetLastError());
    float *h = (float *)((char *)dc.host_flag + 32);
    HIPCHK(hipMemcpyAsync(h, mm, 2 * sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    host[0] = h[0];
    host[1] = h[1];
    return ASETS_OK;
}

asets_status cost_loglik(const float *img, const int32_t *scrib, int64_t n, int32_t nlab,
                         double eps, float *Ct, const asets_opts *o) {
    if (!img || !scrib || !Ct) return fail(ASETS_EBADARG, "null array argument");
    if (n <= 0 || nlab <= 0 || nlab > 64)
        return fail(ASETS_EBADARG, "need n > 0 and 1 <= nlab <= 64");
    if (n > ((int64_t)1 << 40)) return fail(ASETS_ETOOBIG, "too many voxels");
    PrepCtx c;
    std::unique_lock<std::mutex> lock;
    const size_t need = Carver::sz(2 * kPrepBlocks, 4) + Carver::sz(2, 4) +
                        Carver::sz(256 * nlab, 4) + Carver::sz(nlab, 8);
    asets_status st = prep_begin(o, need, c, lock);
    if (st) return st;
    Carver cv(c.dc->dev);
    float *part = cv.take<float>(2 * kPrepBlocks);
    float *mm = cv.take<float>(2);
    uint32_t *hist = cv.take<uint32_t>(256 * nlab);
    uint64_t *tot = cv.take<uint64_t>(nlab);
    float h[2];
    st = minmax(img, n, part, mm, h, c.s, *c.dc);
    if (st) return st;
    if (!(h[1] > h[0]))
        return fail(ASETS_EBADARG, "the image is constant (or holds NaN): no normalisation");
    HIPCHK(hipMemsetAsync(hist, 0, 256 * nlab * sizeof(uint32_t), c.s));
    const int nb = (int)std::min<int64_t>(kPrepBlocks, cdiv(n, 256));
    k_loglik_hist<<<nb, 256, 256 * nlab * sizeof(uint32_t), c.s>>>(img, scrib, n, nlab, mm, hist);
    k_loglik_totals<<<1, 64, 0, c.s>>>(hist, nlab, tot);
    HIPCHK(hipGetLastError());
    uint64_t *ht = (uint64_t *)c.dc->host_flag;   // 64 pinned bytes: 4 totals at a time
    for (int l0 = 0; l0 < nlab; l0 += 4) {
        const int k = std::min(4, nlab - l0);
        HIPCHK(hipMemcpyAsync(ht, tot + l0, k * sizeof(uint64_t), hipMemcpyDeviceToHost, c.s));
        HIPCHK(hipStreamSynchronize(c.s));
        for (int j = 0; j < k; ++j)
            if (ht[j] == 0)
                return fail(ASETS_EBADARG, "label %d has no scribbled voxel", l0 + j + 1);
    }
    k_loglik_cost<<<(unsigned)cdiv(n, 256), 256, 0, c.s>>>(img, n, nlab, mm, hist, tot, eps, Ct);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c.s));
    return ASETS_OK;
}

asets_status alpha_contrast(int ndim, const float *img, int64_t ny, int64_t nx, int64_t nz,
                            double a0, double a1, double beta, int power, int normalize,
                            float *alpha, const asets_opts *o) {
    if (ndim != 2 && ndim != 3) return fail(ASETS_EBADARG, "ndim must be 2 or 3");
    if (ndim == 2 && nz != 1) return fail(ASETS_EBADARG, "2D needs nz == 1");
    asets_status st = check_dims(ny, nx, nz, 1);
    if (st) return st;
    if (!img || !alpha) return fail(ASETS_EBADARG, "null array argument");
    if (power != 1 && power != 2) return fail(ASETS_EBADARG, "power must be 1 or 2");
    PrepCtx c;
    std::unique_lock<std::mutex> lock;
    st = prep_begin(o, Carver::sz(2 * kPrepBlocks, 4) + Carver::sz(2, 4), c, lock);
    if (st) return st;
    Carver cv(c.dc->dev);
    float *part = cv.take<float>(2 * kPrepBlocks);
    float *mm = cv.take<float>(2);
    const int64_t N = ny * nx * nz;
    if (normalize) {
        float h[2];
        st = minmax(img, N, part, mm, h, c.s, *c.dc);
        if (st) return st;
        if (!(h[1] > h[0])) return fail(ASETS_EBADARG, "the image is constant: no normalisation");
    }
    k_alpha_contrast<<<(unsigned)cdiv(N, 256), 256, 0, c.s>>>(img, ny, nx, nz, ndim == 3, mm,
                                                             normalize, a0, a1, beta, power,
                                                             alpha);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c.s));
    return ASETS_OK;
}

asets_status cost_l1(const float *img, int64_t n, const float *models, int32_t nlab,
                     float *Ct, const asets_opts *o) {
    if (!img || !models || !Ct) return fail(ASETS_EBADARG, "null array argument");
    if (n <= 0 || nlab <= 0 || nlab > kMaxCostLabels)
        return fail(ASETS_EBADARG, "need n > 0 and 1 <= nlab <= %d", kMaxCostLabels);
    if (n > ((int64_t)1 << 40)) return fail(ASETS_ETOOBIG, "too many voxels");
    PrepCtx c;
    std::unique_lock<std::mutex> lock;
    asets_status st = prep_begin(o, Carver::sz(2 * kPrepBlocks, 4) + Carver::sz(2, 4), c, lock);
    if (st) return st;
    Carver cv(c.dc->dev);
    float *part = cv.take<float>(2 * kPrepBlocks);
    float *mm = cv.take<float>(2);
    float h[2];
    st = minmax(img, n, part, mm, h, c.s, *c.dc);
    if (st) return st;
    if (!(h[1] > h[0])) return fail(ASETS_EBADARG, "the image is constant: no normalisation");
    CostModels ms{};
    for (int l = 0; l < nlab; ++l) ms.m[l] = models[l];
    k_l1_costs<<<(unsigned)cdiv(n, 256), 256, 0, c.s>>>(img, n, mm, ms, nlab, Ct);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c.s));
    return ASETS_OK;
}

asets_status cost_smooth_labels(const int32_t *lab, int64_t ny, int64_t nx, const int32_t *ids,
                                int32_t nlab, const double *h, int32_t K, float *Ct,
                                const asets_opts *o) {
    if (!lab || !ids || !h || !Ct) return fail(ASETS_EBADARG, "null array argument");
    asets_status st = check_dims(ny, nx, 1, 1);
    if (st) return st;
    if (nlab <= 0 || nlab > kMaxCostLabels)
        return fail(ASETS_EBADARG, "need 1 <= nlab <= %d", kMaxCostLabels);
    if (K < 1 || K > 15 || K % 2 == 0)
        return fail(ASETS_EBADARG, "the kernel size must be odd, 1 to 15 (got %d)", K);
    PrepCtx c;
    std::unique_lock<std::mutex> lock;
    st = prep_begin(o, Carver::sz(K * K, 8), c, lock);
    if (st) return st;
    Carver cv(c.dc->dev);
    double *dh = cv.take<double>(K * K);
    HIPCHK(hipMemcpyAsync(dh, h, K * K * sizeof(double), hipMemcpyHostToDevice, c.s));
    CostIds cid{};
    for (int l = 0; l < nlab; ++l) cid.id[l] = ids[l];
    k_smooth_labels<<<(unsigned)cdiv(ny * nx, 256), 256, 0, c.s>>>(lab, ny, nx, cid, nlab, dh, K,
                                                                  Ct);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c.s));
    return ASETS_OK;
}

#include "hmf_host.h"

}  // namespace

// ---------------------------------------------------------------------------
// the allocation-free binary solve (asets_binary_dev_noalloc)
// ---------------------------------------------------------------------------
// Its whole device workspace (two state buffers, partial sums, flags) is carved from
// this array: each device the library runs on holds its own copy until the process
// ends.  In host code `scratch` names the host-side stand-in, not device memory: the
// device address is taken once per device with hipGetSymbolAddress (that device
// current) and only that pointer is used.
static __device__ __align__(256) unsigned char scratch[ASETS_NOALLOC_SCRATCH_BYTES];

namespace {
struct ScratchDev {
    std::mutex init_mu;   // guards ptr
    std::mutex mu;        // serialises the solves that use the array
    void *ptr = nullptr;
};
ScratchDev g_scratch[kMaxDevices];

// Device address of this device's scratch (current device = dev).
asets_status scratch_address(int dev, void **out) {
    ScratchDev &sd = g_scratch[dev];
    std::lock_guard<std::mutex> lk(sd.init_mu);
    if (!sd.ptr) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return fail(ASETS_EDEVICE, "hipGetSymbolAddress(scratch) failed: %s",
                        hipGetErrorString(e));
        }
        sd.ptr = p;
    }
    *out = sd.ptr;
    return ASETS_OK;
}
}  // namespace

// ---------------------------------------------------------------------------
// extern "C" entry points
// ---------------------------------------------------------------------------
extern "C" {

asets_status asets_binary2d(const float *Cs, const float *Ct, const float *alpha, int64_t ny,
                            int64_t nx, const asets_params *p, float *u, float *cvg,
                            int32_t *it_num, double *seconds, const asets_opts *o) {
    return binary_host(2, Cs, Ct, alpha, ny, nx, 1, p, nullptr, nullptr, u, cvg, it_num, seconds,
                       o);
}

asets_status asets_binary3d(const float *Cs, const float *Ct, const float *alpha, int64_t ny,
                            int64_t nx, int64_t nz, const asets_params *p, float *u, float *cvg,
                            int32_t *it_num, double *seconds, const asets_opts *o) {
    return binary_host(3, Cs, Ct, alpha, ny, nx, nz, p, nullptr, nullptr, u, cvg, it_num,
                       seconds, o);
}

asets_status asets_binary_warm(int ndim, const float *Cs, const float *Ct, const float *alpha,
                               int64_t ny, int64_t nx, int64_t nz, const asets_params *p,
                               const float *u0, const float *pt0, float *u, float *cvg,
                               int32_t *it_num, double *seconds, const asets_opts *o) {
    if (ndim != 2 && ndim != 3) return fail(ASETS_EBADARG, "ndim must be 2 or 3");
    if (ndim == 2 && nz != 1) return fail(ASETS_EBADARG, "2D solve needs nz == 1");
    return binary_host(ndim, Cs, Ct, alpha, ny, nx, nz, p, u0, pt0, u, cvg, it_num, seconds, o);
}

asets_status asets_potts2d(const float *Ct, const float *alpha, int64_t ny, int64_t nx,
                           int64_t nlab, const asets_params *p, float *u, float *cvg,
                           int32_t *it_num, double *seconds, const asets_opts *o) {
    return potts_host(2, Ct, alpha, ny, nx, 1, nlab, p, u, cvg, it_num, seconds, o);
}

asets_status asets_potts3d(const float *Ct, const float *alpha, int64_t ny, int64_t nx,
                           int64_t nz, int64_t nlab, const asets_params *p, float *u,
                           float *cvg, int32_t *it_num, double *seconds, const asets_opts *o) {
    return potts_host(3, Ct, alpha, ny, nx, nz, nlab, p, u, cvg, it_num, seconds, o);
}

asets_status asets_ishikawa2d(const float *Ct, const float *alpha, int64_t ny, int64_t nx,
                              int64_t nlab, const asets_params *p, float *u, float *cvg,
                              int32_t *it_num, double *seconds, const asets_opts *o) {
    return ishikawa_host(2, Ct, alpha, ny, nx, 1, nlab, p, u, cvg, it_num, seconds, o);
}

asets_status asets_ishikawa3d(const float *Ct, const float *alpha, int64_t ny, int64_t nx,
                              int64_t nz, int64_t nlab, const asets_params *p, float *u,
                              float *cvg, int32_t *it_num, double *seconds, const asets_opts *o) {
    return ishikawa_host(3, Ct, alpha, ny, nx, nz, nlab, p, u, cvg, it_num, seconds, o);
}

asets_status asets_ishikawa_dev(int ndim, const float *Ct, const float *alpha, int64_t ny,
                                int64_t nx, int64_t nz, int64_t nlab, const asets_params *p,
                                float *u, float *cvg, int32_t *it_num, double *seconds,
                                const asets_opts *o) {
    return ishikawa_dev(ndim, Ct, alpha, ny, nx, nz, nlab, p, u, cvg, it_num, seconds, o);
}

asets_status asets_binary_dev(int ndim, const float *Cs, const float *Ct, const float *alpha,
                              int64_t ny, int64_t nx, int64_t nz, const asets_params *p,
                              const float *u0, const float *pt0, float *u, float *cvg,
                              int32_t *it_num, double *seconds, const asets_opts *o) {
    WarmIn w;
    w.u0 = u0;
    w.pt0 = pt0;
    return binary_dev(ndim, Cs, Ct, alpha, ny, nx, nz, p, w, u, cvg, it_num, seconds, nullptr,
                      nullptr, nullptr, nullptr, nullptr, o);
}

asets_status asets_binary_dev_noalloc(int ndim, const float *Cs, const float *Ct,
                                      const float *alpha, int64_t ny, int64_t nx, int64_t nz,
                                      const asets_params *p, const float *u0, const float *pt0,
                                      float *u, float *cvg, int32_t *it_num, double *seconds,
                                      const asets_opts *o) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(ASETS_EDEVICE, "device id %d", dev);
    void *ws = nullptr;
    asets_status st = scratch_address(dev, &ws);
    if (st) return st;
    int32_t hflag[4] = {0, 0, 0, 0};   // stop-flag polls read into pageable host memory
    const FixedWs fw{ws, (size_t)ASETS_NOALLOC_SCRATCH_BYTES, &g_scratch[dev].mu, hflag};
    WarmIn w;
    w.u0 = u0;
    w.pt0 = pt0;
    return binary_dev(ndim, Cs, Ct, alpha, ny, nx, nz, p, w, u, cvg, it_num, seconds, nullptr,
                      nullptr, nullptr, nullptr, nullptr, o, &fw);
}

asets_status asets_binary_dev_state(int ndim, const float *Cs, const float *Ct,
                                    const float *alpha, int64_t ny, int64_t nx, int64_t nz,
                                    const asets_params *p, const float *u0, const float *pt0,
                                    float *u, float *cvg, int32_t *it_num, double *seconds,
                                    float *ps, float *pt, float *bx, float *by, float *bz,
                                    const asets_opts *o) {
    WarmIn w;
    w.u0 = u0;
    w.pt0 = pt0;
    return binary_dev(ndim, Cs, Ct, alpha, ny, nx, nz, p, w, u, cvg, it_num, seconds, ps, pt, bx,
                      by, bz, o);
}

asets_status asets_binary_dev_resume(int ndim, const float *Cs, const float *Ct,
                                     const float *alpha, int64_t ny, int64_t nx, int64_t nz,
                                     const asets_params *p, const float *u0, const float *ps0,
                                     const float *pt0, const float *bx0, const float *by0,
                                     const float *bz0, float *u, float *cvg, int32_t *it_num,
                                     double *seconds, float *ps, float *pt, float *bx, float *by,
                                     float *bz, const asets_opts *o) {
    if (!u0 || !bx0 || !by0 || (ndim == 3 && !bz0))
        return fail(ASETS_EBADARG, "resume needs u0 and the flow faces");
    if (!ps0 != !pt0)
        return fail(ASETS_EBADARG, "resume takes ps0 and pt0 both, or neither (paired)");
    WarmIn w;
    w.u0 = u0;
    w.ps0 = ps0;
    w.pt0 = pt0;
    w.pair = !ps0;
    w.b0[0] = bx0;
    w.b0[1] = by0;
    w.b0[2] = ndim == 3 ? bz0 : nullptr;
    return binary_dev(ndim, Cs, Ct, alpha, ny, nx, nz, p, w, u, cvg, it_num, seconds, ps, pt, bx,
                      by, bz, o);
}

asets_status asets_potts_dev(int ndim, const float *Ct, const float *alpha, int64_t ny,
                             int64_t nx, int64_t nz, int64_t nlab, const asets_params *p,
                             float *u, float *cvg, int32_t *it_num, double *seconds,
                             const asets_opts *o) {
    return potts_dev(ndim, Ct, alpha, ny, nx, nz, nlab, p, u, cvg, it_num, seconds, o);
}

asets_status asets_binary3d_batched_dev(int32_t nvol, const asets_volume *vols,
                                        const asets_params *p, double *seconds,
                                        const asets_opts *o) {
    const int flow16 = o ? o->flow_dtype : ASETS_FLOW_F32;
    if (flow16 == ASETS_FLOW_F16) return batched_dev<__half>(nvol, vols, p, seconds, o);
    if (flow16 != ASETS_FLOW_F32) return fail(ASETS_EBADARG, "unknown flow_dtype %d", flow16);
    return batched_dev<float>(nvol, vols, p, seconds, o);
}

asets_status asets_binary_energy_dev(int ndim, const float *Cs, const float *Ct,
                                     const float *alpha, int64_t ny, int64_t nx, int64_t nz,
                                     const float *u, double *energy, const asets_opts *o) {
    if (!Cs) return fail(ASETS_EBADARG, "null array argument");
    return energy_dev(ndim, Cs, Ct, alpha, ny, nx, nz, 1, u, energy, o);
}

asets_status asets_potts_energy_dev(int ndim, const float *Ct, const float *alpha, int64_t ny,
                                    int64_t nx, int64_t nz, int64_t nlab, const float *u,
                                    double *energy, const asets_opts *o) {
    return energy_dev(ndim, nullptr, Ct, alpha, ny, nx, nz, nlab, u, energy, o);
}

asets_status asets_binary_bounds_dev(int ndim, const float *Cs, const float *Ct,
                                     const float *alpha, const float *u, const float *bx,
                                     const float *by, const float *bz, int64_t ny, int64_t nx,
                                     int64_t nz, int32_t sweeps, double *bounds,
                                     const asets_opts *o) {
    return bounds_dev(ndim, Cs, Ct, alpha, u, bx, by, bz, ny, nx, nz, sweeps, bounds, o);
}

asets_status asets_cost_loglik_dev(const float *img, const int32_t *scribbles, int64_t n,
                                   int32_t nlab, double epsilon, float *Ct,
                                   const asets_opts *o) {
    return cost_loglik(img, scribbles, n, nlab, epsilon, Ct, o);
}

asets_status asets_cost_neglog_dev(const float *p, int64_t n, double epsilon, float *Ct,
                                   const asets_opts *o) {
    if (!p || !Ct) return fail(ASETS_EBADARG, "null array argument");
    if (n <= 0 || n > ((int64_t)1 << 40)) return fail(ASETS_EBADARG, "bad element count");
    hipStream_t s = o ? (hipStream_t)o->stream : nullptr;
    k_neglog<<<(unsigned)cdiv(n, 256), 256, 0, s>>>(p, n, epsilon, Ct);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return ASETS_OK;
}

asets_status asets_alpha_contrast_dev(int ndim, const float *img, int64_t ny, int64_t nx,
                                      int64_t nz, double a0, double a1, double beta, int power,
                                      int normalize, float *alpha, const asets_opts *o) {
    return alpha_contrast(ndim, img, ny, nx, nz, a0, a1, beta, power, normalize, alpha, o);
}

asets_status asets_labels_dev(int model, const float *u, int64_t n, int32_t nlab,
                              uint8_t *labels, const asets_opts *o) {
    if (!u || !labels) return fail(ASETS_EBADARG, "null array argument");
    if (n <= 0 || n > ((int64_t)1 << 40)) return fail(ASETS_EBADARG, "bad element count");
    if (model < 0 || model > 2) return fail(ASETS_EBADARG, "model must be 0, 1 or 2");
    if (model == 0) nlab = 1;
    if (nlab < 1 || nlab > 255 || (model == 2 && nlab < 2))
        return fail(ASETS_EBADARG, "bad label count %d", nlab);
    hipStream_t s = o ? (hipStream_t)o->stream : nullptr;
    k_labels<<<(unsigned)cdiv(n, 256), 256, 0, s>>>(model, u, n, nlab, labels);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return ASETS_OK;
}

asets_status asets_selftest_division(int mode, float d, uint64_t *mismatches) {
    if (!mismatches) return fail(ASETS_EBADARG, "mismatches is NULL");
    if (mode != 0 && mode != 1) return fail(ASETS_EBADARG, "mode must be 0 or 1");
    if (mode == 0 && !(std::fabs(d) >= 0x1p-16f && std::fabs(d) <= 0x1p16f))
        return fail(ASETS_EBADARG, "divisor outside [2^-16, 2^16]");
    unsigned long long *dcount = nullptr;
    HIPCHK(hipMalloc(&dcount, sizeof(unsigned long long)));
    HIPCHK(hipMemset(dcount, 0, sizeof(unsigned long long)));
    k_selftest_div<<<4096, 256>>>(mode, d, dcount);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    unsigned long long h = 0;
    if (e == hipSuccess) e = hipMemcpy(&h, dcount, sizeof h, hipMemcpyDeviceToHost);
    (void)hipFree(dcount);
    if (e != hipSuccess) return fail(ASETS_EDEVICE, "selftest: %s", hipGetErrorString(e));
    *mismatches = h;
    return ASETS_OK;
}

asets_status asets_hbm_stream_dev(int mode, void *buf, int64_t bytes, int32_t flags,
                                  int32_t grid, int32_t iters, double *ms, int64_t *traffic,
                                  const asets_opts *o) {
    if (!buf || !ms || !traffic) return fail(ASETS_EBADARG, "null argument");
    if (mode < kStreamRead || mode > kStreamMixBlk)
        return fail(ASETS_EBADARG, "mode 0 to 5");
    if (((uintptr_t)buf & 255) != 0) return fail(ASETS_EBADARG, "buf must be 256-byte aligned");
    if (iters < 1 || grid < 0) return fail(ASETS_EBADARG, "iters >= 1 and grid >= 0");
    const bool nt = (flags & 1) != 0, unroll1 = (flags & 2) != 0;
    int dev = 0, cus = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int g = grid > 0 ? grid : 8 * cus;
    hipStream_t s = o ? (hipStream_t)o->stream : nullptr;
    char *b = (char *)buf;
    std::function<void()> launch;
    if (mode == kStreamRead) {
        const int64_t sink = (int64_t)g * 256 * 4;
        const int64_t n16 = (bytes - sink) / 16;
        if (n16 < 1) return fail(ASETS_EBADARG, "buffer too small");
        *traffic = n16 * 16;
        const f32x4 *src = (const f32x4 *)b;
        float *snk = (float *)(b + n16 * 16);
        launch = [=]() {
            if (nt && unroll1) k_stream_read<1, true><<<g, 256, 0, s>>>(src, n16, snk);
            else if (nt) k_stream_read<4, true><<<g, 256, 0, s>>>(src, n16, snk);
            else if (unroll1) k_stream_read<1, false><<<g, 256, 0, s>>>(src, n16, snk);
            else k_stream_read<4, false><<<g, 256, 0, s>>>(src, n16, snk);
        };
    } else if (mode == kStreamCopy) {
        const int64_t n16 = (bytes / 2) / 256 * 16;   // halves 256-byte aligned
        if (n16 < 1) return fail(ASETS_EBADARG, "buffer too small");
        *traffic = 2 * n16 * 16;
        const f32x4 *src = (const f32x4 *)b;
        f32x4 *dst = (f32x4 *)(b + n16 * 16);
        launch = [=]() {
            if (nt && unroll1) k_stream_copy<1, true><<<g, 256, 0, s>>>(src, dst, n16);
            else if (nt) k_stream_copy<4, true><<<g, 256, 0, s>>>(src, dst, n16);
            else if (unroll1) k_stream_copy<1, false><<<g, 256, 0, s>>>(src, dst, n16);
            else k_stream_copy<4, false><<<g, 256, 0, s>>>(src, dst, n16);
        };
    } else if (mode == kStreamMixState) {
        const int64_t N = bytes / 60 / 1024 * 1024;
        if (N < 1024) return fail(ASETS_EBADARG, "buffer too small");
        *traffic = 60 * N;
        const f32x4 *cs = (const f32x4 *)b, *ct = (const f32x4 *)(b + 4 * N);
        const f32x4 *al = (const f32x4 *)(b + 8 * N);
        const f32x4 *sin = (const f32x4 *)(b + 12 * N);
        f32x4 *sout = (f32x4 *)(b + 36 * N);
        const int64_t n16 = N / 4;
        launch = [=]() {
            if (nt) k_stream_mix_state<true><<<g, 256, 0, s>>>(cs, ct, al, sin, sout, n16);
            else k_stream_mix_state<false><<<g, 256, 0, s>>>(cs, ct, al, sin, sout, n16);
        };
    } else if (mode == kStreamMixBlk) {
        const int64_t N = bytes / 60 / 1024 * 1024;
        if (N < 1024) return fail(ASETS_EBADARG, "buffer too small");
        *traffic = 60 * N;
        const float *cs = (const float *)b, *ct = (const float *)(b + 4 * N);
        const float *al = (const float *)(b + 8 * N);
        const char *sin = b + 12 * N;
        char *sout = b + 36 * N;
        launch = [=]() {
            if (nt) k_stream_mix_blk<true><<<g, 256, 0, s>>>(cs, ct, al, sin, sout, N);
            else k_stream_mix_blk<false><<<g, 256, 0, s>>>(cs, ct, al, sin, sout, N);
        };
    } else if (mode == kStreamMixPacked) {
        const int64_t N = bytes / 60 / 1024 * 1024;
        if (N < 1024) return fail(ASETS_EBADARG, "buffer too small");
        *traffic = 60 * N;
        const f32x4 *in = (const f32x4 *)b;
        const f32x4 *sin = (const f32x4 *)(b + 12 * N);
        f32x4 *sout = (f32x4 *)(b + 36 * N);
        const int64_t n16 = N / 4;
        launch = [=]() {
            if (nt) k_stream_mix_packed<true><<<g, 256, 0, s>>>(in, sin, sout, n16);
            else k_stream_mix_packed<false><<<g, 256, 0, s>>>(in, sin, sout, n16);
        };
    } else {
        const int64_t N = bytes / 60 / 1024 * 1024;
        if (N < 1024) return fail(ASETS_EBADARG, "buffer too small");
        *traffic = 60 * N;
        MixFields f;
        char *q = b;
        for (int k = 0; k < 5; ++k, q += 4 * N) f.in4[k] = (const f32x4 *)q;
        for (int k = 0; k < 2; ++k, q += 8 * N) f.in8[k] = (const f32x4 *)q;
        for (int k = 0; k < 2; ++k, q += 4 * N) f.out4[k] = (f32x4 *)q;
        for (int k = 0; k < 2; ++k, q += 8 * N) f.out8[k] = (f32x4 *)q;
        const int64_t n16 = N / 4;
        launch = [=]() {
            if (nt) k_stream_mix<true><<<g, 256, 0, s>>>(f, n16);
            else k_stream_mix<false><<<g, 256, 0, s>>>(f, n16);
        };
    }
    EventPair ev;
    HIPCHK(hipEventCreate(&ev.a));
    HIPCHK(hipEventCreate(&ev.b));
    launch();   // untimed
    HIPCHK(hipEventRecord(ev.a, s));
    for (int32_t i = 0; i < iters; ++i) launch();
    HIPCHK(hipEventRecord(ev.b, s));
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventSynchronize(ev.b));
    float t = 0.0f;
    HIPCHK(hipEventElapsedTime(&t, ev.a, ev.b));
    *ms = (double)t / iters;
    return ASETS_OK;
}

asets_status asets_cost_l1_dev(const float *img, int64_t n, const float *models, int32_t nlab,
                               float *Ct, const asets_opts *o) {
    return cost_l1(img, n, models, nlab, Ct, o);
}

asets_status asets_cost_smooth_labels_dev(const int32_t *labels, int64_t ny, int64_t nx,
                                          const int32_t *ids, int32_t nlab, const double *h,
                                          int32_t K, float *Ct, const asets_opts *o) {
    return cost_smooth_labels(labels, ny, nx, ids, nlab, h, K, Ct, o);
}

asets_status asets_hmf_dev(int ndim, int64_t ny, int64_t nx, int64_t nz, asets_hmf_node *nodes,
                           int32_t nnodes, int32_t num_iter, float steps, double cc,
                           double *seconds, const asets_opts *o) {
    return hmf_dev(ndim, ny, nx, nz, nodes, nnodes, num_iter, steps, cc, seconds, o);
}

const char *asets_last_error(void) { return g_last_error.c_str(); }

const char *asets_version(void) { return "asets-mi355x 0.1 (gfx950)"; }

void asets_release(void) {
    for (int d = 0; d < kMaxDevices; ++d) {
        DeviceCache &dc = g_cache[d];
        std::lock_guard<std::mutex> lock(dc.mu);
        if (!dc.dev && !dc.io && !dc.host_flag) continue;
        int prev = 0;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(d);
        if (dc.dev) (void)hipFree(dc.dev);
        if (dc.io) (void)hipFree(dc.io);
        if (dc.host_flag) (void)hipHostFree(dc.host_flag);
        dc.dev = dc.io = nullptr;
        dc.host_flag = nullptr;
        dc.dev_bytes = dc.io_bytes = 0;
        (void)hipSetDevice(prev);
    }
}

}  // extern "C"
