// nn_kernels.hip -- libasets_nn.so: kernels of the warm-start network, a small
// 3D U-Net that maps the cost volumes (Cs, Ct, alpha) of a binary max-flow
// problem to a starting point (u0, pt0) for the iteration.  No reference
// counterpart: the reference always starts from the initial cut of
// asetsBinaryMF3D_mex.c:205-224.
//
// Tensors are NDHWC (channels fastest) in bf16, accumulation in fp32.  A batch
// of column-major volumes maps onto NDHWC with D = Nz, H = Nx, W = Ny: voxel
// p = n*N + g has the same flat index as in the solver's layout, so no
// transposition is needed.  3x3x3 convolutions (stride 1, zero padding 1) run
// as implicit GEMMs on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16):
// M = voxels, N = output channels, K = 27 taps x input channels.  The input
// gradient is the same convolution with the taps mirrored and the weight
// matrix transposed; the weight gradient is an MFMA GEMM per tap over fixed voxel
// chunks, whose partials one kernel adds in order (deterministic).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>

#include "asets_nn.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

thread_local std::string g_nn_err;

asets_status nn_fail(asets_status s, const char *fmt, ...) {
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_nn_err = buf;
    return s;
}

#define NNCHK()                                                                  \
    do {                                                                         \
        hipError_t e_ = hipGetLastError();                                       \
        if (e_ != hipSuccess)                                                    \
            return nn_fail(ASETS_EDEVICE, "launch: %s", hipGetErrorString(e_));  \
    } while (0)

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

constexpr int64_t kMaxElems = (int64_t)1 << 31;   // 32-bit element indexing

__device__ __forceinline__ float bf2f(u16 b) { return __uint_as_float((uint32_t)b << 16); }
// fp32 -> bf16, round to nearest even; NaN stays NaN
__device__ __forceinline__ u16 f2bf(float f) {
    uint32_t x = __float_as_uint(f);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (u16)((x >> 16) | 0x40);
    x += 0x7fffu + ((x >> 16) & 1u);
    return (u16)(x >> 16);
}

struct Geo5 {
    int n, d, h, w;   // batch, depth (z), height (x), width (y)
};

// ---------------------------------------------------------------------------
// 3x3x3 convolution: out[p, co] = act(b[co] + sum_{k, ci} in[p + o(k), ci] *
// w[k, co, ci]), weights [27][COUT][cin] (ci contiguous), tap k = (dz+1)*9 +
// (dy+1)*3 + (dx+1) with (dz, dy, dx) the offsets along (d, h, w).  A 256-thread
// workgroup covers 256 consecutive voxels; each wave 64 voxels (4 MFMA row
// tiles) x all COUT (COUT/16 column tiles).  MFMA 16x16x32 operand layout:
// lane l supplies row l%16 and the 8 consecutive K indices 8*(l/16)..+7 of A
// (and column l%16 of B); it receives rows 4*(l/16)+i, i < 4, of column l%16.
// cin is a multiple of 8, so the 8 K indices of a lane share one tap and are
// one 16-byte load.
//
// UPCAT: the input is the channel concatenation [up2(in), skip] of a decoder
// stage, gathered without materialising it: channels ci < ca read the coarse
// tensor `in` (ca channels, geometry g halved) at the nearest coarse voxel,
// channels ci >= ca read `skip` (cin - ca channels) at the voxel itself.  ca is a
// multiple of 8, so a lane's 8 channels come from one source, and the values are
// those asets_nn_up2 + asets_nn_chan_copy would have stored: same bits.
// ---------------------------------------------------------------------------
template <int COUT, bool UPCAT = false, int MT = 4>
__global__ __launch_bounds__(256) void k_conv3_fwd(const u16 *__restrict__ in, Geo5 g, int cin,
                                                   const u16 *__restrict__ w,
                                                   const float *__restrict__ bias, int relu,
                                                   u16 *__restrict__ out,
                                                   const u16 *__restrict__ skip = nullptr,
                                                   int ca = 0) {
    constexpr int NT = COUT / 16;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int P = g.n * g.d * g.h * g.w;
    const int p0 = ((int)blockIdx.x * 4 + wv) * (16 * MT);
    const int r = lane & 15, kg = lane >> 4;
    const int ch = g.h >> 1, cw = g.w >> 1, cs = cin - ca;   // coarse extents (UPCAT)
    int pz[MT], py[MT], px[MT], pb[MT], pc[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        int p = p0 + m * 16 + r;
        const bool v = p < P;
        if (!v) p = 0;
        px[m] = p % g.w;
        int t = p / g.w;
        py[m] = t % g.h;
        t /= g.h;
        const int z = t % g.d;
        pb[m] = p - (z * g.h + py[m]) * g.w - px[m];   // voxel index of (n, 0, 0, 0)
        pc[m] = (t / g.d) * (g.d >> 1) * ch * cw;       // coarse voxel index of (n, 0, 0, 0)
        pz[m] = v ? z : -1000;   // an invalid row never passes the bounds test
    }
    f32x4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int ktot = 27 * cin;
    for (int k0 = 0; k0 < ktot; k0 += 32) {
        const int kk = k0 + 8 * kg;
        const int tap = kk / cin, ci = kk - tap * cin;
        const bool tv = tap < 27;
        const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
        bf16x8 a[MT], b[NT];
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const int z = pz[m] + dz, y = py[m] + dy, x = px[m] + dx;
            const bool ok = tv && (unsigned)z < (unsigned)g.d && (unsigned)y < (unsigned)g.h &&
                            (unsigned)x < (unsigned)g.w;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (ok) {
                if constexpr (UPCAT) {
                    if (ci < ca) {
                        const int q = pc[m] + ((z >> 1) * ch + (y >> 1)) * cw + (x >> 1);
                        v = *(const uint4 *)(in + (size_t)q * ca + ci);
                    } else {
                        const int q = pb[m] + (z * g.h + y) * g.w + x;
                        v = *(const uint4 *)(skip + (size_t)q * cs + (ci - ca));
                    }
                } else {
                    const int q = pb[m] + (z * g.h + y) * g.w + x;
                    v = *(const uint4 *)(in + (size_t)q * cin + ci);
                }
            }
            a[m] = __builtin_bit_cast(bf16x8, v);
        }
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            uint4 v = make_uint4(0, 0, 0, 0);
            if (tv) v = *(const uint4 *)(w + ((size_t)tap * COUT + n * 16 + r) * cin + ci);
            b[n] = __builtin_bit_cast(bf16x8, v);
        }
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int n = 0; n < NT; ++n)
                acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b[n], acc[m][n], 0, 0, 0);
    }
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int p = p0 + m * 16 + 4 * kg + i;
            if (p < P)
#pragma unroll
                for (int n = 0; n < NT; ++n) {
                    const int co = n * 16 + r;
                    float v = acc[m][n][i] + (bias ? bias[co] : 0.0f);
                    if (relu) v = v > 0.0f ? v : 0.0f;
                    out[(size_t)p * COUT + co] = f2bf(v);
                }
        }
}

// ---------------------------------------------------------------------------
// The same convolution with its operands staged in LDS, for rows of at least 64
// voxels.  A box is 64 (w) x 2 (h) x 2 (d) voxels, one 64-voxel row per wave (4
// MFMA row tiles, all COUT as in k_conv3_fwd).  The box with its one-voxel halo,
// 66 x 4 x 4 voxels (zero outside the volume), is copied to LDS once per chunk of
// CC input channels; every tap then reads its A operand from there instead of
// gathering it from global memory.  The chunk sits in LDS channel-group major
// ([group][voxel], 16 bytes = 8 channels per entry), so the 16 rows of an A
// fragment are 16 consecutive entries.  K runs over taps and channels within a
// chunk, chunk after chunk: for cin <= 16 that is k_conv3_fwd's order, the same
// bits.  Outputs go through LDS and leave as 16 contiguous bytes per lane.
//
// Such a kernel is bound by instruction issue unless its index arithmetic stays
// out of the per-box path, so what does not depend on the box is computed once
// per thread: the staging slots' offsets from the box origin and their boundary
// classes (a box whose faces all lie inside the volume needs no bounds test; at
// a face, one mask test per slot) and each K step's LDS offset (the K loop is
// unrolled at compile time).  Volumes whose extents are not multiples of the box
// take a per-slot bounds test instead.
//
// k_conv3_box (cin = CC = 8 or 16, one chunk): the weights of the whole layer are
// staged in LDS, in MFMA fragment order, once per workgroup; the grid is
// persistent, and a workgroup loads its next box into registers while it
// computes the current one (its compute loop issues no global load, which the
// in-order load counter would make wait for that prefetch).
// k_conv3_boxm (cin a multiple of 16, several chunks, and the decoder's
// concatenated input): one box per workgroup, chunk after chunk, with the
// weights read from global memory (cache hits) one K step ahead.  UPCAT stages
// [up2(in), skip] as k_conv3_fwd<COUT, true> gathers it.
// ---------------------------------------------------------------------------
constexpr int kBoxW = 64, kBoxH = 2, kBoxD = 2;
constexpr int kHaloW = kBoxW + 2, kHaloH = kBoxH + 2, kHaloD = kBoxD + 2;
constexpr int kHaloVox = kHaloW * kHaloH * kHaloD;   // 1056

// fp32 -> bf16 with the gfx950 conversion instruction (round to nearest even)
__device__ __forceinline__ u16 f2bf_hw(float f) { return __builtin_bit_cast(u16, (__bf16)f); }

template <int COUT, int CC>
struct BoxCfg {
    static constexpr int NT = COUT / 16, NG = CC / 8;
    static constexpr int KS = (27 * CC + 31) / 32;          // K steps per chunk
    static constexpr int NA = NG * kHaloVox;                // input entries
    static constexpr int SI = (NA + 255) / 256;             // input slots per thread
    static constexpr int NB = KS * NT * 64;                 // weight fragments
    static constexpr int BI = (NB + 255) / 256;             // weight slots per thread
    static constexpr int NOUT = 4 * 64 * COUT * 2 / 16;     // output staging (uint4)
    static constexpr int NAO = NA > NOUT ? NA : NOUT;
};

template <int COUT, int CC, bool UPCAT>
__global__ __launch_bounds__(256) void k_conv3_box(const u16 *__restrict__ in, Geo5 g, int cin,
                                                   const u16 *__restrict__ w,
                                                   const float *__restrict__ bias, int relu,
                                                   u16 *__restrict__ out,
                                                   const u16 *__restrict__ skip, int ca) {
    using B = BoxCfg<COUT, CC>;
    constexpr int MT = 4, NT = B::NT, NG = B::NG, KS = B::KS;
    static_assert(!UPCAT, "one chunk: cin = CC, no concatenation");
    (void)skip;
    (void)ca;
    __shared__ uint4 tile[B::NAO];   // input box; then the outputs of the box
    __shared__ uint4 btile[B::NB];   // weights (one chunk: the whole layer)
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int r = lane & 15, kg = lane >> 4;
    const int nbw = (g.w + kBoxW - 1) / kBoxW, nbh = (g.h + kBoxH - 1) / kBoxH;
    const int nbd = (g.d + kBoxD - 1) / kBoxD;
    const int nbox = g.n * nbd * nbh * nbw;
    const int wz = wv >> 1, wy = wv & 1;   // this wave's row within the box
    const int plane = g.h * g.w, vol = g.d * plane;
    const bool full = g.w % kBoxW == 0 && g.h % kBoxH == 0 && g.d % kBoxD == 0;
    if ((int)blockIdx.x >= nbox) return;

    // ---- weights: staged once per workgroup ----------------------------------
#pragma unroll
    for (int k = 0; k < B::BI; ++k) {
        const int j = tid + k * 256, ln = j & 63, t = j >> 6;
        const int q = t % NT, ks = t / NT;
        const int kk = ks * 32 + 8 * (ln >> 4);
        const int tap = kk / CC, grp = (kk % CC) / 8;
        if (j < B::NB) {
            uint4 v = make_uint4(0, 0, 0, 0);
            if (tap < 27) v = *(const uint4 *)(w + (tap * COUT + q * 16 + (ln & 15)) * cin + 8 * grp);
            btile[j] = v;
        }
    }
    // ---- per-thread tables (independent of the box) --------------------------
    // input slot k: LDS entry (12 bits), halo position, boundary class (bit 0/1:
    // first/last halo column, 2/3: row, 4/5: plane), valid; voxel offset from the
    // box origin
    unsigned spk[B::SI];   // lds | hx << 12 | hy << 19 | hz << 21 | cls << 23 | valid << 29
    int soff[B::SI];
#pragma unroll
    for (int k = 0; k < B::SI; ++k) {
        const int i = tid + k * 256;
        const int v = i / NG, grp = i % NG;   // group fastest: coalesced loads
        const int hx = v % kHaloW, hy = (v / kHaloW) % kHaloH, hz = v / (kHaloW * kHaloH);
        const unsigned cls = (hx == 0 ? 1u : 0u) | (hx == kHaloW - 1 ? 2u : 0u) |
                             (hy == 0 ? 4u : 0u) | (hy == kHaloH - 1 ? 8u : 0u) |
                             (hz == 0 ? 16u : 0u) | (hz == kHaloD - 1 ? 32u : 0u);
        spk[k] = (unsigned)(grp * kHaloVox + v) | ((unsigned)hx << 12) | ((unsigned)hy << 19) |
                 ((unsigned)hz << 21) | (cls << 23) | ((i < B::NA ? 1u : 0u) << 29);
        soff[k] = (((hz - 1) * g.h + (hy - 1)) * g.w + (hx - 1)) * cin + 8 * grp;
    }
    // LDS entry of K step ks's A operand for this lane (row tile 0); the last step
    // has taps past 26 in some lane groups
    int aofs[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
        const int kk = ks * 32 + 8 * kg;
        const int tap = min(kk / CC, 26), grp = (kk % CC) / 8;
        const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
        aofs[ks] = grp * kHaloVox + ((wz + 1 + dz) * kHaloH + (wy + 1 + dy)) * kHaloW + 1 + dx + r;
    }
    const bool last_ok = (KS - 1) * 32 + 8 * kg < 27 * CC;

    struct Box {
        int n, x0, y0, z0;
        unsigned cls;   // boundary classes whose halo positions lie outside the volume
    };
    auto box_of = [&](int b) {
        Box o;
        const int bw = b % nbw;
        b /= nbw;
        const int bh = b % nbh;
        b /= nbh;
        const int bd = b % nbd;
        o.n = b / nbd;
        o.x0 = bw * kBoxW;
        o.y0 = bh * kBoxH;
        o.z0 = bd * kBoxD;
        o.cls = (bw == 0 ? 1u : 0u) | (bw == nbw - 1 ? 2u : 0u) | (bh == 0 ? 4u : 0u) |
                (bh == nbh - 1 ? 8u : 0u) | (bd == 0 ? 16u : 0u) | (bd == nbd - 1 ? 32u : 0u);
        return o;
    };
    uint4 sv[B::SI];   // the next box's input
    auto load_box = [&](const Box &X) {
        const u16 *src = in + ((size_t)X.n * vol + X.z0 * plane + X.y0 * g.w + X.x0) * cin;
#pragma unroll
        for (int k = 0; k < B::SI; ++k) {
            const unsigned pk = spk[k];
            bool ok = (pk >> 29) & 1u;
            if (full) {
                ok = ok && ((pk >> 23) & X.cls) == 0u;
            } else {
                const int x = X.x0 - 1 + (int)((pk >> 12) & 127u);
                const int y = X.y0 - 1 + (int)((pk >> 19) & 3u);
                const int z = X.z0 - 1 + (int)((pk >> 21) & 3u);
                ok = ok && (unsigned)x < (unsigned)g.w && (unsigned)y < (unsigned)g.h &&
                     (unsigned)z < (unsigned)g.d;
            }
            sv[k] = make_uint4(0, 0, 0, 0);
            if (ok) sv[k] = *(const uint4 *)(src + soff[k]);
        }
    };

    int b = blockIdx.x;
    Box X = box_of(b);
    load_box(X);
    float bco[NT];
#pragma unroll
    for (int q = 0; q < NT; ++q) bco[q] = bias ? bias[q * 16 + r] : 0.0f;
    for (;;) {
        const int nb = b + gridDim.x;
        const bool more = nb < nbox;
        __syncthreads();   // the previous box's LDS reads are done (and the weights stored)
#pragma unroll
        for (int k = 0; k < B::SI; ++k)
            if ((spk[k] >> 29) & 1u) tile[spk[k] & 4095u] = sv[k];
        __syncthreads();
        const Box Xc = X;
        if (more) {
            X = box_of(nb);
            load_box(X);   // in flight during this box's MFMAs
        }
        f32x4 acc[MT][NT];
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int q = 0; q < NT; ++q) acc[m][q] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            bf16x8 a[MT], bb[NT];
#pragma unroll
            for (int m = 0; m < MT; ++m) {
                uint4 v = tile[aofs[ks] + 16 * m];
                if (ks == KS - 1 && !last_ok) v = make_uint4(0, 0, 0, 0);
                a[m] = __builtin_bit_cast(bf16x8, v);
            }
#pragma unroll
            for (int q = 0; q < NT; ++q)
                bb[q] = __builtin_bit_cast(bf16x8, btile[(ks * NT + q) * 64 + lane]);
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int q = 0; q < NT; ++q)
                    acc[m][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], bb[q], acc[m][q], 0, 0, 0);
        }
        // outputs: into LDS in the MFMA layout, then 16 contiguous bytes per lane
        const int z = Xc.z0 + wz, y = Xc.y0 + wy;
        const bool rowok = z < g.d && y < g.h;
        const size_t prow = (size_t)Xc.n * vol + (size_t)z * plane + (size_t)y * g.w;
        __syncthreads();   // every wave has read its last A operand
        u16 *ot = (u16 *)tile + wv * 64 * COUT;
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int q = 0; q < NT; ++q) {
                    float v = acc[m][q][i] + bco[q];
                    if (relu) v = v > 0.0f ? v : 0.0f;
                    ot[(m * 16 + 4 * kg + i) * COUT + q * 16 + r] = f2bf_hw(v);
                }
        __syncthreads();
        constexpr int G = COUT / 8;   // 16-byte groups per voxel
#pragma unroll
        for (int e = lane; e < 64 * G; e += 64) {
            const int x = Xc.x0 + e / G;
            if (rowok && x < g.w)
                *(uint4 *)(out + (prow + x) * COUT + (e % G) * 8) = *(const uint4 *)(ot + e * 8);
        }
        if (!more) break;
        b = nb;
    }
}

// Layers of more than one chunk (cin > 16) and the decoder's concatenated input:
// one box per workgroup, chunk after chunk.  Each chunk's box is staged with all
// of a thread's loads issued before its first LDS store; the weights come from
// global memory (cache hits), loaded one K step ahead; outputs go through LDS.
template <int COUT, bool UPCAT>
__global__ __launch_bounds__(256) void k_conv3_boxm(const u16 *__restrict__ in, Geo5 g, int cin,
                                                    const u16 *__restrict__ w,
                                                    const float *__restrict__ bias, int relu,
                                                    u16 *__restrict__ out,
                                                    const u16 *__restrict__ skip, int ca) {
    using B = BoxCfg<COUT, 16>;
    constexpr int MT = 4, NT = B::NT, SI = B::SI;
    __shared__ uint4 tile[B::NAO];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int r = lane & 15, kg = lane >> 4;
    const int nbw = (g.w + kBoxW - 1) / kBoxW, nbh = (g.h + kBoxH - 1) / kBoxH;
    const int nbd = (g.d + kBoxD - 1) / kBoxD;
    int b = blockIdx.x;
    const int bw = b % nbw;
    b /= nbw;
    const int bh = b % nbh;
    b /= nbh;
    const int bd = b % nbd;
    const int n = b / nbd;
    const int x0 = bw * kBoxW, y0 = bh * kBoxH, z0 = bd * kBoxD;
    const int wz = wv >> 1, wy = wv & 1;   // this wave's row within the box
    const int ch = g.h >> 1, cw = g.w >> 1, cs = cin - ca;
    const int vb = n * g.d * g.h * g.w;       // voxel index of (n, 0, 0, 0)
    const int vc = n * (g.d >> 1) * ch * cw;  // coarse voxel of (n, 0, 0, 0)
    f32x4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int q = 0; q < NT; ++q) acc[m][q] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int cb = 0; cb < cin; cb += 16) {
        // all of a thread's staging loads are issued before the first LDS store
        uint4 sv[SI];
        int sidx[SI];
#pragma unroll
        for (int k = 0; k < SI; ++k) {
            const int i = tid + k * 256;
            const int v = i >> 1, grp = i & 1;   // group fastest: coalesced
            const int hx = v % kHaloW, t = v / kHaloW;
            const int hy = t % kHaloH, hz = t / kHaloH;
            const int x = x0 - 1 + hx, y = y0 - 1 + hy, z = z0 - 1 + hz;
            sv[k] = make_uint4(0, 0, 0, 0);
            sidx[k] = i < B::NA ? grp * kHaloVox + v : -1;
            if (sidx[k] >= 0 && (unsigned)x < (unsigned)g.w && (unsigned)y < (unsigned)g.h &&
                (unsigned)z < (unsigned)g.d) {
                const int ci = cb + 8 * grp;
                if constexpr (UPCAT) {
                    if (ci < ca) {
                        const int q = vc + ((z >> 1) * ch + (y >> 1)) * cw + (x >> 1);
                        sv[k] = *(const uint4 *)(in + (size_t)q * ca + ci);
                    } else {
                        const int q = vb + (z * g.h + y) * g.w + x;
                        sv[k] = *(const uint4 *)(skip + (size_t)q * cs + (ci - ca));
                    }
                } else {
                    const int q = vb + (z * g.h + y) * g.w + x;
                    sv[k] = *(const uint4 *)(in + (size_t)q * cin + ci);
                }
            }
        }
        // weights of K step ks (zero past the last tap), loaded one step ahead
        auto load_b = [&](int ks, bf16x8 *bb) {
            const int kk = ks * 32 + 8 * kg;
            const int tap = kk >> 4, grp = (kk & 15) >> 3;
#pragma unroll
            for (int q = 0; q < NT; ++q) {
                uint4 v = make_uint4(0, 0, 0, 0);
                if (tap < 27)
                    v = *(const uint4 *)(w + ((size_t)tap * COUT + q * 16 + r) * cin + cb + 8 * grp);
                bb[q] = __builtin_bit_cast(bf16x8, v);
            }
        };
        bf16x8 bcur[NT], bnext[NT];
        load_b(0, bcur);
        __syncthreads();   // the previous chunk's reads are done
#pragma unroll
        for (int k = 0; k < SI; ++k)
            if (sidx[k] >= 0) tile[sidx[k]] = sv[k];
        __syncthreads();
#pragma unroll 1
        for (int ks = 0; ks < B::KS; ++ks) {
            const int kk = ks * 32 + 8 * kg;
            const int tap = kk >> 4, grp = (kk & 15) >> 3;
            const bool tv = tap < 27;
            const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
            const int base = grp * kHaloVox + ((wz + 1 + dz) * kHaloH + (wy + 1 + dy)) * kHaloW +
                             1 + dx + r;
            if (ks + 1 < B::KS) load_b(ks + 1, bnext);
            bf16x8 a[MT];
#pragma unroll
            for (int m = 0; m < MT; ++m) {
                uint4 v = make_uint4(0, 0, 0, 0);
                if (tv) v = tile[base + 16 * m];
                a[m] = __builtin_bit_cast(bf16x8, v);
            }
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int q = 0; q < NT; ++q)
                    acc[m][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], bcur[q], acc[m][q], 0, 0, 0);
#pragma unroll
            for (int q = 0; q < NT; ++q) bcur[q] = bnext[q];
        }
    }
    const int z = z0 + wz, y = y0 + wy;
    const bool rowok = z < g.d && y < g.h;
    const int prow = vb + (z * g.h + y) * g.w;
    // outputs: into LDS in the MFMA layout, then 16 contiguous bytes per lane
    __syncthreads();   // every wave has read its last A operand
    u16 *ot = (u16 *)tile + wv * 64 * COUT;
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int q = 0; q < NT; ++q) {
                const int co = q * 16 + r;
                float v = acc[m][q][i] + (bias ? bias[co] : 0.0f);
                if (relu) v = v > 0.0f ? v : 0.0f;
                ot[(m * 16 + 4 * kg + i) * COUT + co] = f2bf_hw(v);
            }
    __syncthreads();
    constexpr int G = COUT / 8;   // 16-byte groups per voxel
#pragma unroll
    for (int e = lane; e < 64 * G; e += 64) {
        const int x = x0 + e / G;
        if (rowok && x < g.w)
            *(uint4 *)(out + (size_t)(prow + x) * COUT + (e % G) * 8) = *(const uint4 *)(ot + e * 8);
    }
}

// [27][cout][cin] -> [27][cin][cout] with the taps mirrored (k -> 26-k): the
// weights of the input-gradient convolution.
__global__ void k_flip_transpose(const u16 *w, u16 *wt, int cout, int cin) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 27 * cout * cin) return;
    const int ci = i % cin;
    const int t = i / cin;
    const int co = t % cout;
    const int k = t / cout;
    wt[((26 - k) * cin + ci) * cout + co] = w[i];
}

// g *= (y > 0): gradient through a ReLU whose output is y
__global__ void k_relu_mask(u16 *g, const u16 *y, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(bf2f(y[i]) > 0.0f)) g[i] = 0;
}

// ---------------------------------------------------------------------------
// weight gradient on the matrix cores: for each tap k,
//   dw[k, co, ci] = sum_p g[p, co] * in[p + o(k), ci],
// a GEMM with M = cout, N = cin and K = voxels.  A workgroup (256 threads) owns
// one 16-channel chunk of cin (8 for cin = 8) and walks boxes of 64 (w) x 2 (h) x
// 2 (d) output voxels (blockIdx.x, blockIdx.x + gridDim.x, ...), accumulating all
// 27 taps of its chunk in registers; each workgroup writes its own partial and
// k_sum_chunks adds them in order (deterministic).
//
// Per box, two LDS tiles, voxels contiguous so the 8 consecutive K values of an
// MFMA lane are one 16-byte read:
//   Gt[co][4 box rows][64]        the output gradient of the box;
//   Xt[ci][16 rows][80]           the input over the 4 x 4 (z, y) rows of the box and
//                                 its halo, w from x0 - 8 (16-byte aligned windows);
// both filled by 8 x 8 (voxel x channel) blocks transposed in registers and stored
// as 16-byte rows, zero outside the volume.  The waves split the nine (dz, dy) rows
// of taps (wave w: w, w + 4, w + 8).  For each of a box row's two 32-voxel K steps
// and each of its (dz, dy) rows, a lane reads its aligned 8-voxel input window and
// the 4 bytes either side; the dx = -1 and +1 windows are the same bytes shifted by
// one voxel (v_alignbyte), so the three taps of the row share one LDS read.  Tile
// rows are padded by 16 bytes so the 16 channels (co) of a fragment fall in
// different banks.
// ---------------------------------------------------------------------------
constexpr int kWgCoMax = 64, kWgCiMax = 128;
constexpr int kWbLen = 80;                          // input row: w0 - 8 .. w0 + 71
constexpr int kWbXs = 16 * kWbLen + 8;              // Xt elements per channel
constexpr int kWbGs = 4 * 64 + 8;                   // Gt elements per channel

// 8 x 8 transpose of 16-bit values: v[j] holds channels 0..7 of voxel j (dwords of
// channel pairs); o[c] receives voxels 0..7 of channel c
__device__ __forceinline__ void tr8x8(const uint4 *v, uint4 *o) {
    uint32_t d[8][4];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        d[j][0] = v[j].x;
        d[j][1] = v[j].y;
        d[j][2] = v[j].z;
        d[j][3] = v[j].w;
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        uint32_t r[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const uint32_t a = d[2 * m][c >> 1], b = d[2 * m + 1][c >> 1];
            r[m] = (c & 1) ? ((a >> 16) | (b & 0xffff0000u)) : ((a & 0xffffu) | (b << 16));
        }
        o[c] = make_uint4(r[0], r[1], r[2], r[3]);
    }
}

template <int NCT>
__global__ __launch_bounds__(256) void k_conv3_wgrad_box(const u16 *__restrict__ in, Geo5 g,
                                                         int cin, const u16 *__restrict__ go,
                                                         float *__restrict__ part,
                                                         float *__restrict__ bpart) {
    constexpr int COUT = NCT * 16;
    __shared__ __attribute__((aligned(16))) u16 xt[16 * kWbXs];
    __shared__ __attribute__((aligned(16))) u16 gt[COUT * kWbGs];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int r = lane & 15, kg = lane >> 4;
    const int cb = blockIdx.y * 16;                      // first input channel of the chunk
    const int ng = (cin - cb) >= 16 ? 2 : 1;             // 8-channel groups in the chunk
    const int nbw = (g.w + 63) / 64, nbh = (g.h + 1) / 2, nbd = (g.d + 1) / 2;
    const int nbox = g.n * nbd * nbh * nbw;
    f32x4 acc[3][3][NCT];
#pragma unroll
    for (int rr = 0; rr < 3; ++rr)
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int ct = 0; ct < NCT; ++ct) acc[rr][q][ct] = f32x4{0.f, 0.f, 0.f, 0.f};
    // bias gradient (first channel chunk only): thread t < 64*NCT stages one block of
    // 8 voxels x 8 output channels per box and sums it, voxels in order
    const bool bias = bpart && blockIdx.y == 0;
    float bsum[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) bsum[c] = 0.0f;
    for (int b = blockIdx.x; b < nbox; b += gridDim.x) {
        int t = b;
        const int x0 = (t % nbw) * 64;
        t /= nbw;
        const int y0 = (t % nbh) * 2;
        t /= nbh;
        const int z0 = (t % nbd) * 2;
        const int n = t / nbd;
        __syncthreads();   // the previous box's reads are done
        // input rows: 16 rows x 10 blocks of 8 voxels x 2 channel halves
        for (int blk = tid; blk < 16 * 10 * 2; blk += 256) {
            const int half = blk & 1, tt = blk >> 1;
            const int eb = tt % 10, row = tt / 10;
            const int zz = z0 - 1 + (row >> 2), yy = y0 - 1 + (row & 3);
            const bool rok = half < ng && (unsigned)zz < (unsigned)g.d &&
                             (unsigned)yy < (unsigned)g.h;
            const size_t rb = (((size_t)n * g.d + zz) * g.h + yy) * g.w;
            uint4 v[8], o[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int x = x0 - 8 + eb * 8 + j;
                v[j] = make_uint4(0, 0, 0, 0);
                if (rok && (unsigned)x < (unsigned)g.w)
                    v[j] = *(const uint4 *)(in + (rb + x) * cin + cb + 8 * half);
            }
            tr8x8(v, o);
#pragma unroll
            for (int c = 0; c < 8; ++c)
                *(uint4 *)&xt[(half * 8 + c) * kWbXs + row * kWbLen + eb * 8] = o[c];
        }
        // output gradient: 4 box rows x 8 blocks of 8 voxels x NCT*2 channel groups
        for (int blk = tid; blk < 4 * 8 * 2 * NCT; blk += 256) {
            const int cg = blk % (2 * NCT), tt = blk / (2 * NCT);
            const int vb = tt & 7, br = tt >> 3;
            const int zz = z0 + (br >> 1), yy = y0 + (br & 1);
            const bool rok = zz < g.d && yy < g.h;
            const size_t rb = (((size_t)n * g.d + zz) * g.h + yy) * g.w;
            uint4 v[8], o[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int x = x0 + vb * 8 + j;
                v[j] = make_uint4(0, 0, 0, 0);
                if (rok && x < g.w) v[j] = *(const uint4 *)(go + (rb + x) * COUT + cg * 8);
            }
            tr8x8(v, o);
#pragma unroll
            for (int c = 0; c < 8; ++c) *(uint4 *)&gt[(cg * 8 + c) * kWbGs + br * 64 + vb * 8] = o[c];
            if (bias) {
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    const uint32_t q[4] = {o[c].x, o[c].y, o[c].z, o[c].w};
#pragma unroll
                    for (int m = 0; m < 4; ++m) {
                        bsum[c] += bf2f((u16)(q[m] & 0xffffu));
                        bsum[c] += bf2f((u16)(q[m] >> 16));
                    }
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) {
            const int br = ks >> 1, hh = ks & 1;
            const int bz = br >> 1, by = br & 1;
            bf16x8 a[NCT];
#pragma unroll
            for (int ct = 0; ct < NCT; ++ct)
                a[ct] = *(const bf16x8 *)&gt[(ct * 16 + r) * kWbGs + br * 64 + hh * 32 + kg * 8];
            const int e0 = 8 + hh * 32 + kg * 8;   // element of voxel x0 + hh*32 + kg*8
#pragma unroll
            for (int rr = 0; rr < 3; ++rr) {
                const int row9 = wv + 4 * rr;   // (dz + 1) * 3 + (dy + 1)
                if (row9 < 9) {
                    const int dz = row9 / 3 - 1, dy = row9 % 3 - 1;
                    const u16 *rp = &xt[r * kWbXs + ((bz + 1 + dz) * 4 + (by + 1 + dy)) * kWbLen];
                    const uint4 cur = *(const uint4 *)(rp + e0);
                    const uint32_t pv = *(const uint32_t *)(rp + e0 - 2);
                    const uint32_t nx = *(const uint32_t *)(rp + e0 + 8);
                    const uint4 wm = make_uint4(__builtin_amdgcn_alignbyte(cur.x, pv, 2),
                                                __builtin_amdgcn_alignbyte(cur.y, cur.x, 2),
                                                __builtin_amdgcn_alignbyte(cur.z, cur.y, 2),
                                                __builtin_amdgcn_alignbyte(cur.w, cur.z, 2));
                    const uint4 wp = make_uint4(__builtin_amdgcn_alignbyte(cur.y, cur.x, 2),
                                                __builtin_amdgcn_alignbyte(cur.z, cur.y, 2),
                                                __builtin_amdgcn_alignbyte(cur.w, cur.z, 2),
                                                __builtin_amdgcn_alignbyte(nx, cur.w, 2));
                    const bf16x8 bm = __builtin_bit_cast(bf16x8, wm);
                    const bf16x8 b0 = __builtin_bit_cast(bf16x8, cur);
                    const bf16x8 bp = __builtin_bit_cast(bf16x8, wp);
#pragma unroll
                    for (int ct = 0; ct < NCT; ++ct) {
                        acc[rr][0][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ct], bm, acc[rr][0][ct], 0, 0, 0);
                        acc[rr][1][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ct], b0, acc[rr][1][ct], 0, 0, 0);
                        acc[rr][2][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ct], bp, acc[rr][2][ct], 0, 0, 0);
                    }
                }
            }
        }
    }
    if (bias) {   // the workgroup's bias partial: the 32 threads of each channel group in order
        __syncthreads();   // every wave is done with xt
        float *bred = (float *)xt;
        if (tid < 64 * NCT)
#pragma unroll
            for (int c = 0; c < 8; ++c) bred[tid * 8 + c] = bsum[c];
        __syncthreads();
        if (tid < COUT) {
            const int cg = tid >> 3, c = tid & 7;
            float t = 0.0f;
            for (int q = cg; q < 64 * NCT; q += 2 * NCT) t += bred[q * 8 + c];
            bpart[(size_t)blockIdx.x * COUT + tid] = t;
        }
    }
    // lane holds rows (co) 4*kg + i of column (ci) r of each (tap, co tile)
    const int ci = cb + r;
    if (r >= 8 * ng) return;
#pragma unroll
    for (int rr = 0; rr < 3; ++rr) {
        const int row9 = wv + 4 * rr;
        if (row9 >= 9) continue;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const int tap = row9 * 3 + q;
#pragma unroll
            for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int co = ct * 16 + 4 * kg + i;
                    part[(((size_t)blockIdx.x * 27 + tap) * COUT + co) * cin + ci] = acc[rr][q][ct][i];
                }
        }
    }
}

// workgroups per input chunk of the weight gradient: about 512 in all, at most one
// per box
inline int wg_groups(int64_t nbox, int cin) {
    const int64_t chunks = cin < 16 ? 1 : cin / 16;
    int64_t gx = (512 + chunks - 1) / chunks;
    if (gx > nbox) gx = nbox;
    return (int)(gx < 1 ? 1 : gx);
}

// out[i] (+)= sum over the chunks c of part[c*n + i].  A block of 256 threads owns
// EPB elements; each of its 256/EPB groups adds a contiguous run of chunks in order
// (loads issued eight ahead of their adds), and the group sums are added in group
// order: a fixed order for every element (deterministic).
template <int EPB>
__global__ __launch_bounds__(256) void k_sum_chunks(const float *part, int nchunk, int n,
                                                    float *out, int accumulate) {
    constexpr int NG = 256 / EPB;
    __shared__ float red[256];
    const int e = threadIdx.x % EPB, gi = threadIdx.x / EPB;
    const int i = blockIdx.x * EPB + e;
    const int per = (nchunk + NG - 1) / NG;
    const int c0 = gi * per, c1 = min(nchunk, c0 + per);
    float s = 0.0f;
    if (i < n) {
        int c = c0;
        for (; c + 8 <= c1; c += 8) {
            float t[8];
#pragma unroll
            for (int b = 0; b < 8; ++b) t[b] = part[(size_t)(c + b) * n + i];
#pragma unroll
            for (int b = 0; b < 8; ++b) s += t[b];
        }
        for (; c < c1; ++c) s += part[(size_t)c * n + i];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    if (gi == 0 && i < n) {
        float t = 0.0f;
        for (int q = 0; q < NG; ++q) t += red[q * EPB + e];
        out[i] = accumulate ? out[i] + t : t;
    }
}

void sum_chunks(const float *part, int nchunk, int n, float *out, int accumulate,
                hipStream_t s) {
    if (n >= 2048)
        k_sum_chunks<32><<<(unsigned)cdiv(n, 32), 256, 0, s>>>(part, nchunk, n, out, accumulate);
    else
        k_sum_chunks<8><<<(unsigned)cdiv(n, 8), 256, 0, s>>>(part, nchunk, n, out, accumulate);
}

// ---------------------------------------------------------------------------
// resampling, channel copies, features, head, optimiser
// ---------------------------------------------------------------------------
// 2x2x2 average pool, g the fine geometry (even extents)
__global__ void k_pool2_fwd(const u16 *in, Geo5 g, int c, u16 *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int D = g.d / 2, H = g.h / 2, W = g.w / 2;
    if (i >= g.n * D * H * W * c) return;
    const int ch = i % c;
    int t = i / c;
    const int x = t % W;
    t /= W;
    const int y = t % H;
    t /= H;
    const int z = t % D;
    const int n = t / D;
    float s = 0.0f;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int e = 0; e < 2; ++e)
                s += bf2f(in[((((size_t)n * g.d + 2 * z + a) * g.h + 2 * y + b) * g.w + 2 * x + e) * c + ch]);
    out[i] = f2bf(s * 0.125f);
}

// pool gradient: each fine voxel gets 1/8 of its coarse gradient (added to gin
// when accumulate != 0)
__global__ void k_pool2_bwd(const u16 *gout, Geo5 g, int c, u16 *gin, int accumulate) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.n * g.d * g.h * g.w * c) return;
    const int ch = i % c;
    int t = i / c;
    const int x = t % g.w;
    t /= g.w;
    const int y = t % g.h;
    t /= g.h;
    const int z = t % g.d;
    const int n = t / g.d;
    const int D = g.d / 2, H = g.h / 2, W = g.w / 2;
    float v = bf2f(gout[((((size_t)n * D + z / 2) * H + y / 2) * W + x / 2) * c + ch]) * 0.125f;
    if (accumulate) v += bf2f(gin[i]);
    gin[i] = f2bf(v);
}

// nearest-neighbour 2x upsample into channels [off, off+c) of out (cd channels);
// g the fine geometry
__global__ void k_up2_fwd(const u16 *in, Geo5 g, int c, u16 *out, int cd, int off) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.n * g.d * g.h * g.w * c) return;
    const int ch = i % c;
    const int p = i / c;
    int t = p;
    const int x = t % g.w;
    t /= g.w;
    const int y = t % g.h;
    t /= g.h;
    const int z = t % g.d;
    const int n = t / g.d;
    const int D = g.d / 2, H = g.h / 2, W = g.w / 2;
    out[(size_t)p * cd + off + ch] = in[((((size_t)n * D + z / 2) * H + y / 2) * W + x / 2) * c + ch];
}

// upsample gradient: sum of the 8 fine gradients read from channels [off,
// off+c) of gout (cd channels); g the fine geometry
__global__ void k_up2_bwd(const u16 *gout, int cd, int off, Geo5 g, int c, u16 *gin) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int D = g.d / 2, H = g.h / 2, W = g.w / 2;
    if (i >= g.n * D * H * W * c) return;
    const int ch = i % c;
    int t = i / c;
    const int x = t % W;
    t /= W;
    const int y = t % H;
    t /= H;
    const int z = t % D;
    const int n = t / D;
    float s = 0.0f;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int e = 0; e < 2; ++e)
                s += bf2f(gout[((((size_t)n * g.d + 2 * z + a) * g.h + 2 * y + b) * g.w + 2 * x + e) * cd + off + ch]);
    gin[i] = f2bf(s);
}

// channel block copy: dst[p, doff + c] = src[p, soff + c] for c < nc (or +=)
__global__ void k_chan_copy(u16 *dst, int cd, int doff, const u16 *src, int cs, int soff, int nc,
                            int nvox, int accumulate) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nvox * nc) return;
    const int p = i / nc, c = i % nc;
    const size_t di = (size_t)p * cd + doff + c;
    float v = bf2f(src[(size_t)p * cs + soff + c]);
    if (accumulate) v += bf2f(dst[di]);
    dst[di] = f2bf(v);
}

// input features, 8 channels: Cs/s, Ct/s, alpha, (Ct-Cs)/s, min(Cs,Ct)/s, 0, 0, 0
// with s = kCostScale
constexpr float kCostScale = 10.0f;

__global__ void k_features(const float *Cs, const float *Ct, const float *alpha, int P, u16 *out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const float s = Cs[p], t = Ct[p], a = alpha[p];
    const float k = 1.0f / kCostScale;
    uint4 v;
    v.x = (uint32_t)f2bf(s * k) | ((uint32_t)f2bf(t * k) << 16);
    v.y = (uint32_t)f2bf(a) | ((uint32_t)f2bf((t - s) * k) << 16);
    v.z = (uint32_t)f2bf((s < t ? s : t) * k);
    v.w = 0;
    *(uint4 *)(out + (size_t)p * 8) = v;
}

// head: 1x1x1 convolution of the 16-channel x to (a, b);
//   u0 = a,  pt0 = min(Cs, Ct) + kCostScale * b
// w = [w_u(16), w_pt(16)], bias = [b_u, b_pt]
__global__ void k_head_fwd(const u16 *x, int P, const float *w, const float *bias,
                           const float *Cs, const float *Ct, float *u0, float *pt0) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    float a = bias[0], b = bias[1];
    for (int c = 0; c < 16; ++c) {
        const float v = bf2f(x[(size_t)p * 16 + c]);
        a += w[c] * v;
        b += w[16 + c] * v;
    }
    const float cs = Cs[p], ct = Ct[p];
    u0[p] = a;
    pt0[p] = (cs < ct ? cs : ct) + kCostScale * b;
}

// loss = mean_p (u0 - tu)^2 + mean_p ((pt0 - tpt) / kCostScale)^2.
// Writes dx (bf16) and per-block partials [35] = dw(32), db(2), loss sum.
__global__ __launch_bounds__(256) void k_head_bwd(const u16 *x, int P, const float *w,
                                                  const float *u0, const float *pt0,
                                                  const float *tu, const float *tpt, u16 *dx,
                                                  float *part) {
    __shared__ float red[35][257];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    float gu = 0.0f, gp = 0.0f, l = 0.0f;
    float xv[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) xv[c] = 0.0f;
    if (p < P) {
        const float eu = u0[p] - tu[p];
        const float ep = (pt0[p] - tpt[p]) * (1.0f / kCostScale);
        l = eu * eu + ep * ep;
        const float inv = 2.0f / (float)P;
        gu = eu * inv;   // d loss / d a
        gp = ep * inv;   // d loss / d b  (pt0 = base + s*b, residual divided by s)
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            xv[c] = bf2f(x[(size_t)p * 16 + c]);
            dx[(size_t)p * 16 + c] = f2bf(gu * w[c] + gp * w[16 + c]);
        }
    }
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        red[c][threadIdx.x] = gu * xv[c];
        red[16 + c][threadIdx.x] = gp * xv[c];
    }
    red[32][threadIdx.x] = gu;
    red[33][threadIdx.x] = gp;
    red[34][threadIdx.x] = l;
    __syncthreads();
    if (threadIdx.x < 35) {
        float s = 0.0f;
        for (int t = 0; t < 256; ++t) s += red[threadIdx.x][t];
        part[(size_t)blockIdx.x * 35 + threadIdx.x] = s;
    }
}

// ---------------------------------------------------------------------------
// Space-to-depth variant: the network runs at half resolution on 2x2x2 blocks of
// voxels and predicts the full solver state of every voxel of the block.
// Coarse voxel (v, z2, x2, y2) of geometry gc (the fine extents halved) holds the
// sub-voxels s = (dz*2 + dx)*2 + dy at fine (2*z2+dz, 2*x2+dx, 2*y2+dy).
// ---------------------------------------------------------------------------
// fine voxel index of sub-voxel s of coarse voxel (v, z2, x2, y2)
__device__ __forceinline__ int s2d_fine(const Geo5 &gc, int v, int z2, int x2, int y2, int s) {
    const int dz = s >> 2, dx = (s >> 1) & 1, dy = s & 1;
    const int H = 2 * gc.h, W = 2 * gc.w;
    return ((v * 2 * gc.d + 2 * z2 + dz) * H + 2 * x2 + dx) * W + 2 * y2 + dy;
}

// input features [P2, 32]: channel s*4 + f of sub-voxel s, f = Cs/s, Ct/s, alpha,
// (Ct-Cs)/s with s = kCostScale
__global__ void k_features_s2d(const float *Cs, const float *Ct, const float *alpha, Geo5 gc,
                               u16 *out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int P2 = gc.n * gc.d * gc.h * gc.w;
    if (p >= P2) return;
    const int y2 = p % gc.w, x2 = (p / gc.w) % gc.h, z2 = (p / (gc.w * gc.h)) % gc.d;
    const int v = p / (gc.w * gc.h * gc.d);
    const float k = 1.0f / kCostScale;
    uint32_t wd[16];
#pragma unroll
    for (int zx = 0; zx < 4; ++zx) {              // the dy = 0, 1 pair is one 8-byte load
        const int g = s2d_fine(gc, v, z2, x2, y2, 2 * zx);
        const float2 cs2 = *(const float2 *)(Cs + g), ct2 = *(const float2 *)(Ct + g);
        const float2 a2 = *(const float2 *)(alpha + g);
        const float cs[2] = {cs2.x, cs2.y}, ct[2] = {ct2.x, ct2.y}, a[2] = {a2.x, a2.y};
#pragma unroll
        for (int dy = 0; dy < 2; ++dy) {
            const int s = 2 * zx + dy;
            wd[2 * s] = (uint32_t)f2bf(cs[dy] * k) | ((uint32_t)f2bf(ct[dy] * k) << 16);
            wd[2 * s + 1] = (uint32_t)f2bf(a[dy]) | ((uint32_t)f2bf((ct[dy] - cs[dy]) * k) << 16);
        }
    }
    uint4 *o = (uint4 *)(out + (size_t)p * 32);
#pragma unroll
    for (int q = 0; q < 4; ++q) o[q] = make_uint4(wd[4 * q], wd[4 * q + 1], wd[4 * q + 2], wd[4 * q + 3]);
}

// head: 1x1x1 convolution of x [P2, 16] to 32 outputs o = s*4 + q, written to
// sub-voxel s of the fine field q (u0, bx0, by0, bz0).  w [32][16], bias [32], staged in
// LDS once per workgroup (every lane then reads the same 16 bytes: a broadcast).  The
// two sub-voxels dy = 0, 1 of a (dz, dx) pair are adjacent in a fine field: one 8-byte
// store, contiguous across the lanes of a row.
__global__ __launch_bounds__(256) void k_head_s2d_fwd(const u16 *x, Geo5 gc, const float *w,
                                                      const float *bias, float *u0, float *bx0,
                                                      float *by0, float *bz0) {
    __shared__ float4 Wl[32 * 4];
    __shared__ float Bl[32];
    for (int i = threadIdx.x; i < 32 * 4; i += blockDim.x) Wl[i] = ((const float4 *)w)[i];
    if (threadIdx.x < 32) Bl[threadIdx.x] = bias[threadIdx.x];
    __syncthreads();
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int P2 = gc.n * gc.d * gc.h * gc.w;
    if (p >= P2) return;
    const int y2 = p % gc.w, x2 = (p / gc.w) % gc.h, z2 = (p / (gc.w * gc.h)) % gc.d;
    const int v = p / (gc.w * gc.h * gc.d);
    float xv[16];
    const uint4 *xp = (const uint4 *)(x + (size_t)p * 16);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const uint4 t = xp[q];
        const uint32_t wd[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            xv[8 * q + 2 * j] = bf2f((u16)(wd[j] & 0xffffu));
            xv[8 * q + 2 * j + 1] = bf2f((u16)(wd[j] >> 16));
        }
    }
    float *dst[4] = {u0, bx0, by0, bz0};
#pragma unroll
    for (int zx = 0; zx < 4; ++zx) {              // (dz, dx) = (zx >> 1, zx & 1)
        const int g = s2d_fine(gc, v, z2, x2, y2, 2 * zx);   // dy = 0; dy = 1 is g + 1
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float a2[2];
#pragma unroll
            for (int dy = 0; dy < 2; ++dy) {
                const int o = (2 * zx + dy) * 4 + q;
                float a = Bl[o];
#pragma unroll
                for (int c4 = 0; c4 < 4; ++c4) {
                    const float4 wv = Wl[o * 4 + c4];
                    a += wv.x * xv[4 * c4];
                    a += wv.y * xv[4 * c4 + 1];
                    a += wv.z * xv[4 * c4 + 2];
                    a += wv.w * xv[4 * c4 + 3];
                }
                a2[dy] = a;
            }
            *(float2 *)(dst[q] + g) = make_float2(a2[0], a2[1]);
        }
    }
}

// loss = (1/P) sum over the P fine voxels of (u0 - tu)^2 + lam * sum_q (b_q - tb_q)^2
// (q over bx, by, bz).  Per coarse voxel: the 32 output gradients into gbuf [P2][32],
// the input gradient dx [P2, 16] (bf16) and the voxel's loss sum into lossv [P2].
__global__ void k_head_s2d_bwd(const u16 *x, Geo5 gc, const float *w, const float *u0,
                               const float *bx0, const float *by0, const float *bz0,
                               const float *tu, const float *tbx, const float *tby,
                               const float *tbz, float lam, u16 *dx, float *gbuf,
                               float *lossv) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int P2 = gc.n * gc.d * gc.h * gc.w;
    if (p >= P2) return;
    const int y2 = p % gc.w, x2 = (p / gc.w) % gc.h, z2 = (p / (gc.w * gc.h)) % gc.d;
    const int v = p / (gc.w * gc.h * gc.d);
    const float inv = 2.0f / (8.0f * (float)P2);
    const float *out[4] = {u0, bx0, by0, bz0};
    const float *tgt[4] = {tu, tbx, tby, tbz};
    float gx[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) gx[c] = 0.0f;
    float l = 0.0f;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int g = s2d_fine(gc, v, z2, x2, y2, s);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int o = s * 4 + q;
            const float wq = q == 0 ? 1.0f : lam;
            const float e = out[q][g] - tgt[q][g];
            l += wq * e * e;
            const float go = wq * e * inv;
            gbuf[(size_t)p * 32 + o] = go;
#pragma unroll
            for (int c = 0; c < 16; ++c) gx[c] += go * w[o * 16 + c];
        }
    }
    uint32_t wd[8];
#pragma unroll
    for (int j = 0; j < 8; ++j)
        wd[j] = (uint32_t)f2bf(gx[2 * j]) | ((uint32_t)f2bf(gx[2 * j + 1]) << 16);
    uint4 *d = (uint4 *)(dx + (size_t)p * 16);
    d[0] = make_uint4(wd[0], wd[1], wd[2], wd[3]);
    d[1] = make_uint4(wd[4], wd[5], wd[6], wd[7]);
    lossv[p] = l;
}

// Per chunk of kHeadChunk coarse voxels: partial sums part[chunk][545] of
// dL/dw[o][c] = sum_p g[p][o] x[p][c] (512), dL/db[o] = sum_p g[p][o] (32) and the
// loss sum (1), each thread one value, voxels in order (k_sum_chunks adds the chunks
// in order: deterministic).
constexpr int kHeadChunk = 128;
constexpr int kHeadGrads = 545;
__global__ __launch_bounds__(576) void k_head_s2d_wgrad(const u16 *x, const float *gbuf,
                                                        const float *lossv, int P2,
                                                        float *part) {
    __shared__ float G[kHeadChunk][33];
    __shared__ float X[kHeadChunk][17];
    __shared__ float Lv[kHeadChunk];
    const int p0 = blockIdx.x * kHeadChunk;
    const int np = min(kHeadChunk, P2 - p0);
    for (int i = threadIdx.x; i < kHeadChunk * 32; i += blockDim.x) {
        const int pp = i / 32, o = i % 32;
        G[pp][o] = pp < np ? gbuf[(size_t)(p0 + pp) * 32 + o] : 0.0f;
    }
    for (int i = threadIdx.x; i < kHeadChunk * 16; i += blockDim.x) {
        const int pp = i / 16, c = i % 16;
        X[pp][c] = pp < np ? bf2f(x[(size_t)(p0 + pp) * 16 + c]) : 0.0f;
    }
    for (int i = threadIdx.x; i < kHeadChunk; i += blockDim.x) Lv[i] = i < np ? lossv[p0 + i] : 0.0f;
    __syncthreads();
    const int t = threadIdx.x;
    if (t >= kHeadGrads) return;
    float s = 0.0f;
    if (t < 512) {
        const int o = t / 16, c = t % 16;
        for (int pp = 0; pp < kHeadChunk; ++pp) s += G[pp][o] * X[pp][c];
    } else if (t < 544) {
        const int o = t - 512;
        for (int pp = 0; pp < kHeadChunk; ++pp) s += G[pp][o];
    } else {
        for (int pp = 0; pp < kHeadChunk; ++pp) s += Lv[pp];
    }
    part[(size_t)blockIdx.x * kHeadGrads + t] = s;
}

// Solver state from a predicted (u0, bx0, by0, bz0), per voxel of n volumes of
// (ny, nx, nz) (column-major, back to back), in place: u0 clamped to [0, 1], the
// zero-flux faces (axis index 0) set to 0; ps, pt the source and sink flows that
// the iteration's fixed point pairs with them (asetsBinaryMF3D_mex.c:411-429 at
// du = 0): dv = div b; u >= 1/2: pt = Ct (the sink saturated), else
// pt = min(Cs - dv, Ct); ps = pt + dv.
__global__ void k_warm_state(float *u0, float *bx0, float *by0, float *bz0, const float *Cs,
                             const float *Ct, int ny, int nx, int nz, int64_t total, float *ps,
                             float *pt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int64_t plane = (int64_t)nx * ny, N = plane * nz;
    const int64_t gl = i % N;
    const int y = (int)(gl % ny), x = (int)((gl / ny) % nx), z = (int)(gl / plane);
    const float byl = y > 0 ? by0[i] : 0.0f;
    const float bxl = x > 0 ? bx0[i] : 0.0f;
    const float bzl = z > 0 ? bz0[i] : 0.0f;
    const float byh = y + 1 < ny ? by0[i + 1] : 0.0f;
    const float bxh = x + 1 < nx ? bx0[i + ny] : 0.0f;
    const float bzh = z + 1 < nz ? bz0[i + plane] : 0.0f;
    if (y == 0) by0[i] = 0.0f;
    if (x == 0) bx0[i] = 0.0f;
    if (z == 0) bz0[i] = 0.0f;
    const float dv = byh - byl + bxh - bxl + bzh - bzl;
    float u = u0[i];
    u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
    u0[i] = u;
    const float cs = Cs[i], ct = Ct[i];
    float t = cs - dv;
    t = t < ct ? t : ct;
    const float p_t = u >= 0.5f ? ct : t;
    pt[i] = p_t;
    ps[i] = p_t + dv;
}

// AdamW on fp32 master weights; refreshes the bf16 copy when wb != null
__global__ void k_adamw(float *w, u16 *wb, const float *g, float *m, float *v, int n, float lr,
                        float b1, float b2, float eps, float wd, float bc1, float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / bc1, vh = vi / bc2;
    const float wi = w[i] * (1.0f - lr * wd) - lr * mh / (sqrtf(vh) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

__global__ void k_f32_to_bf16(const float *a, u16 *b, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = f2bf(a[i]);
}

constexpr int kB = 256;

inline unsigned nblk(int64_t n) { return (unsigned)cdiv(n, kB); }

// Persistent grid of the LDS-staged kernel: the workgroups resident at once on
// the current device (occupancy x CUs, queried once per kernel and device), or
// one per box when there are fewer boxes.
template <int COUT, int CC>
unsigned box_grid(const Geo5 &g) {
    static std::mutex mu;
    static std::map<int, int> resident;   // device -> workgroups resident at once
    int dev = 0;
    (void)hipGetDevice(&dev);
    int res;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = resident.find(dev);
        if (it == resident.end()) {
            int per_cu = 0, cus = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &per_cu, k_conv3_box<COUT, CC, false>, 256, 0) != hipSuccess ||
                per_cu < 1)
                per_cu = 1;
            if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                    hipSuccess ||
                cus < 1)
                cus = 1;
            it = resident.emplace(dev, per_cu * cus).first;
        }
        res = it->second;
    }
    const int64_t nbox = (int64_t)g.n * cdiv(g.d, kBoxD) * cdiv(g.h, kBoxH) * cdiv(g.w, kBoxW);
    return (unsigned)(nbox < res ? nbox : res);
}

// Which convolution kernel a layer runs: rows of at least one box width take the
// LDS-staged kernels -- the persistent one for cin = 8 or 16 and 16 or 32 output
// channels, the per-box one for cin a multiple of 16 up to 64 output channels
// (for UPCAT with a coarse channel count that is a multiple of 16).  Other layers
// (the coarse levels of small inputs among them) keep the gathering kernel.
// voxel count below which the gathering convolution runs 16 voxels per wave
constexpr int64_t kSmallConv = (int64_t)1 << 18;

template <int COUT, bool UPCAT>
void launch_conv(const u16 *x, const Geo5 &g, int cin, const u16 *wb, const float *bias,
                 int relu, u16 *y, const u16 *skip, int ca, hipStream_t s) {
    if constexpr (COUT <= 64) {
        if (g.w >= kBoxW) {
            if constexpr (COUT <= 32 && !UPCAT) {
                if (cin == 8) {
                    k_conv3_box<COUT, 8, false><<<box_grid<COUT, 8>(g), 256, 0, s>>>(
                        x, g, cin, wb, bias, relu, y, skip, ca);
                    return;
                }
                if (cin == 16) {
                    k_conv3_box<COUT, 16, false><<<box_grid<COUT, 16>(g), 256, 0, s>>>(
                        x, g, cin, wb, bias, relu, y, skip, ca);
                    return;
                }
            }
            if (cin % 16 == 0 && (!UPCAT || ca % 16 == 0)) {
                const unsigned nbox = (unsigned)((int64_t)g.n * cdiv(g.d, kBoxD) * cdiv(g.h, kBoxH) *
                                                 cdiv(g.w, kBoxW));
                k_conv3_boxm<COUT, UPCAT><<<nbox, 256, 0, s>>>(x, g, cin, wb, bias, relu, y, skip, ca);
                return;
            }
        }
    }
    // small tensors (the coarse levels): 16 voxels per wave instead of 64, so the grid
    // fills the chip; each output keeps its K order (same bits).  (The gathered decoder
    // input measured slower that way: 112 -> 124 us at 32^3, 128 -> 64 channels.)
    const int64_t P = (int64_t)g.n * g.d * g.h * g.w;
    if (!UPCAT && P < kSmallConv) {
        k_conv3_fwd<COUT, UPCAT, 1><<<(unsigned)cdiv(P, 64), 256, 0, s>>>(
            x, g, cin, wb, bias, relu, y, skip, ca);
        return;
    }
    k_conv3_fwd<COUT, UPCAT><<<(unsigned)cdiv(P, 256), 256, 0, s>>>(
        x, g, cin, wb, bias, relu, y, skip, ca);
}

asets_status check_geo(int n, int d, int h, int w, int c) {
    if (n <= 0 || d <= 0 || h <= 0 || w <= 0 || c <= 0)
        return nn_fail(ASETS_EBADARG, "sizes must be positive");
    if ((int64_t)n * d * h * w * c >= kMaxElems)
        return nn_fail(ASETS_ETOOBIG, "tensor of more than 2^31 elements");
    return ASETS_OK;
}

}  // namespace

extern "C" {

const char *asets_nn_last_error(void) { return g_nn_err.c_str(); }

asets_status asets_nn_conv3(const void *in, int n, int d, int h, int w, int cin, const void *wt,
                            const float *bias, int cout, int relu, void *out, void *stream) {
    if (cin % 8 || cin > 512) return nn_fail(ASETS_EBADARG, "cin must be a multiple of 8, <= 512");
    asets_status st = check_geo(n, d, h, w, cin > cout ? cin : cout);
    if (st) return st;
    if (!in || !wt || !out) return nn_fail(ASETS_EBADARG, "null pointer");
    const Geo5 g{n, d, h, w};
    hipStream_t s = (hipStream_t)stream;
    const u16 *x = (const u16 *)in, *wb = (const u16 *)wt;
    u16 *y = (u16 *)out;
    switch (cout) {
    case 16: launch_conv<16, false>(x, g, cin, wb, bias, relu, y, nullptr, 0, s); break;
    case 32: launch_conv<32, false>(x, g, cin, wb, bias, relu, y, nullptr, 0, s); break;
    case 48: launch_conv<48, false>(x, g, cin, wb, bias, relu, y, nullptr, 0, s); break;
    case 64: launch_conv<64, false>(x, g, cin, wb, bias, relu, y, nullptr, 0, s); break;
    case 96: launch_conv<96, false>(x, g, cin, wb, bias, relu, y, nullptr, 0, s); break;
    case 128: launch_conv<128, false>(x, g, cin, wb, bias, relu, y, nullptr, 0, s); break;
    default: return nn_fail(ASETS_EBADARG, "cout must be 16, 32, 48, 64, 96 or 128");
    }
    NNCHK();
    return ASETS_OK;
}

asets_status asets_nn_conv3_upcat(const void *coarse, int ca, const void *skip, int n, int d,
                                  int h, int w, int cin, const void *wt, const float *bias,
                                  int cout, int relu, void *out, void *stream) {
    if (cin % 8 || cin > 512) return nn_fail(ASETS_EBADARG, "cin must be a multiple of 8, <= 512");
    if (ca <= 0 || ca % 8 || ca >= cin)
        return nn_fail(ASETS_EBADARG, "ca must be a multiple of 8 in (0, cin)");
    asets_status st = check_geo(n, d, h, w, cin > cout ? cin : cout);
    if (st) return st;
    if (d % 2 || h % 2 || w % 2) return nn_fail(ASETS_EBADARG, "upsample needs even extents");
    if (!coarse || !skip || !wt || !out) return nn_fail(ASETS_EBADARG, "null pointer");
    const Geo5 g{n, d, h, w};
    hipStream_t s = (hipStream_t)stream;
    const u16 *x = (const u16 *)coarse, *k = (const u16 *)skip, *wb = (const u16 *)wt;
    u16 *y = (u16 *)out;
    switch (cout) {
    case 16: launch_conv<16, true>(x, g, cin, wb, bias, relu, y, k, ca, s); break;
    case 32: launch_conv<32, true>(x, g, cin, wb, bias, relu, y, k, ca, s); break;
    case 64: launch_conv<64, true>(x, g, cin, wb, bias, relu, y, k, ca, s); break;
    default: return nn_fail(ASETS_EBADARG, "cout must be 16, 32 or 64");
    }
    NNCHK();
    return ASETS_OK;
}

asets_status asets_nn_flip_transpose(const void *wt, void *wt_bwd, int cout, int cin,
                                     void *stream) {
    if (cout <= 0 || cin <= 0) return nn_fail(ASETS_EBADARG, "sizes must be positive");
    k_flip_transpose<<<nblk(27LL * cout * cin), kB, 0, (hipStream_t)stream>>>(
        (const u16 *)wt, (u16 *)wt_bwd, cout, cin);
    NNCHK();
    return ASETS_OK;
}

asets_status asets_nn_relu_mask(void *g, const void *y, int64_t n, void *stream) {
    if (n < 0 || n >= kMaxElems) return nn_fail(ASETS_EBADARG, "bad element count");
    if (n == 0) return ASETS_OK;
    k_relu_mask<<<nblk(n), kB, 0, (hipStream_t)stream>>>((u16 *)g, (const u16 *)y, (int)n);
    NNCHK();
    return ASETS_OK;
}

size_t asets_nn_wgrad_scratch_bytes(int n, int d, int h, int w, int cin, int cout) {
    const int64_t nbox = (int64_t)n * cdiv(d, 2) * cdiv(h, 2) * cdiv(w, 64);
    const int64_t gx = wg_groups(nbox, cin);
    return (size_t)(gx * 27LL * cout * cin + gx * cout) * sizeof(float);
}

asets_status asets_nn_conv3_wgrad(const void *in, int n, int d, int h, int w, int cin,
                                  const void *gout, int cout, float *dw, float *dbias,
                                  int accumulate, void *scratch, size_t scratch_bytes,
                                  void *stream) {
    if (cin % 8 || cout % 16 || (cin > 8 && cin % 16) || cout > kWgCoMax || cin > kWgCiMax)
        return nn_fail(ASETS_EBADARG, "cout must be a multiple of 16 up to 64, cin 8 or a "
                                      "multiple of 16 up to 128");
    asets_status st = check_geo(n, d, h, w, cin > cout ? cin : cout);
    if (st) return st;
    if (!scratch || scratch_bytes < asets_nn_wgrad_scratch_bytes(n, d, h, w, cin, cout))
        return nn_fail(ASETS_EBADARG, "scratch too small");
    const Geo5 g{n, d, h, w};
    const int64_t nbox = (int64_t)n * cdiv(d, 2) * cdiv(h, 2) * cdiv(w, 64);
    const int gx = wg_groups(nbox, cin);
    const dim3 grid((unsigned)gx, (unsigned)(cin < 16 ? 1 : cin / 16));
    hipStream_t s = (hipStream_t)stream;
    const int nw = 27 * cout * cin;
    float *part = (float *)scratch;
    float *bp = dbias ? part + (size_t)gx * nw : nullptr;
    const u16 *x = (const u16 *)in, *gg = (const u16 *)gout;
    switch (cout) {
    case 16: k_conv3_wgrad_box<1><<<grid, 256, 0, s>>>(x, g, cin, gg, part, bp); break;
    case 32: k_conv3_wgrad_box<2><<<grid, 256, 0, s>>>(x, g, cin, gg, part, bp); break;
    case 48: k_conv3_wgrad_box<3><<<grid, 256, 0, s>>>(x, g, cin, gg, part, bp); break;
    default: k_conv3_wgrad_box<4><<<grid, 256, 0, s>>>(x, g, cin, gg, part, bp); break;
    }
    sum_chunks(part, gx, nw, dw, accumulate, s);
    if (dbias) sum_chunks(bp, gx, cout, dbias, accumulate, s);
    NNCHK();
    return ASETS_OK;
}

asets_status asets_nn_pool2(const void *in, int n, int d, int h, int w, int c, void *out,
                            void *stream) {
    asets_status st = check_geo(n, d, h, w, c);
    if (st) return st;
    if (d % 2 || h % 2 || w % 2) return nn_fail(ASETS_EBADARG, "pool needs even extents");
    k_pool2_fwd<<<nblk((int64_t)n * d * h * w * c / 8), kB, 0, (hipStream_t)stream>>>(
        (const u16 *)in, Geo5{n, d, h, w}, c, (u16 *)out);
    NNCHK();
    return ASETS_OK;
}

asets_status asets_nn_pool2_bwd(const void *gout, int n, int d, int h, int w, int c, void *gin,
                                int accumulate, void *stream) {
    asets_status st = check_geo(n, d, h, w, c);
    if (st) return st;
    if (d % 2 || h % 2 || w % 2) return nn_fail(ASETS_EBADARG, "pool needs even extents");
    k_pool2_bwd<<<nblk((int64_t)n * d * h * w * c), kB, 0, (hipStream_t)stream>>>(
        (const u16 *)gout, Geo5{n, d, h, w}, c, (u16 *)gin, accumulate);
    NNCHK();
    return ASETS_OK;
}

asets_status asets_nn_up2(const void *in, int n, int d, int h, int w, int c, void *out, int cd,
                          int off, void *stream) {
    asets_status st = check_geo(n, d, h, w, cd);
    if (st) return st;
    if (d % 2 || h % 2 || w % 2) return nn_fail(ASETS_EBADARG, "upsample needs even extents");
    if (off < 0 || off + c > cd) return nn_fail(ASETS_EBADARG, "channel range out of bounds");
    k_up2_fwd<<<nblk((int64_t)n * d * h * w * c), kB, 0, (hipStream_t)stream>>>(
        (const u16 *)in, Geo5{n, d, h, w}, c, (u16 *)out, cd, off);
    NNCHK();
    return ASETS_OK;
}

asets_status asets_nn_up2_bwd(const void *gout, int cd, int off, int n, int d, int h, int w,
                              int c, void *gin, void *stream) {
    asets_status st = check_geo(n, d, h, w, cd);
    if (st) return st;
    if (d % 2 || h % 2 || w % 2) return nn_fail(ASETS_EBADARG, "upsample needs even extents");
    if (off < 0 || off + c > cd) return nn_fail(ASETS_EBADARG, "channel range out of bounds");
    k_up2_bwd<<<nblk((int64_t)n * d * h * w * c / 8), kB, 0, (hipStream_t)stream>>>(
        (const u16 *)gout, cd, off, Geo5{n, d, h, w}, c, (u16 *)gin);
    NNCHK();
    return ASETS_OK;
}

asets_status asets_nn_chan_copy(void *dst, int cd, int doff, const void *src, int cs, int soff,
                                int nc, int64_t nvox, int accumulate, void *stream) {
    if (nvox <= 0 || nc <= 0 || doff < 0 || soff < 0 || doff + nc > cd || soff + nc > cs)
        return nn_fail(ASETS_EBADARG, "channel range out of bounds");
    if (nvox * (cd > cs ? cd : cs) >= kMaxElems) return nn_fail(ASETS_ETOOBIG, "too large");
    k_chan_copy<<<nblk(nvox * nc), kB, 0, (hipStream_t)stream>>>(
        (u16 *)dst, cd, doff, (const u16 *)src, cs, soff, nc, (int)nvox, accumulate);
    NNCHK();
    return ASETS_OK;
}

asets_status asets_nn_features(const float *Cs, const float *Ct, const float *alpha, int64_t P,
                               void *out, void *stream) {
    if (P <= 0 || P * 8 >= kMaxElems) return nn_fail(ASETS_EBADARG, "bad voxel count");
    k_features<<<nblk(P), kB, 0, (hipStream_t)stream>>>(Cs, Ct, alpha, (int)P, (u16 *)out);
    NNCHK();
    return ASETS_OK;
}

asets_status asets_nn_head(const void *x, int64_t P, const float *w, const float *b,
                           const float *Cs, const float *Ct, float *u0, float *pt0,
                           void *stream) {
    if (P <= 0 || P * 16 >= kMaxElems) return nn_fail(ASETS_EBADARG, "bad voxel count");
    k_head_fwd<<<nblk(P), kB, 0, (hipStream_t)stream>>>((const u16 *)x, (int)P, w, b, Cs, Ct,
                                                        u0, pt0);
    NNCHK();
    return ASETS_OK;
}

size_t asets_nn_head_scratch_bytes(int64_t P) { return (size_t)cdiv(P, kB) * 35 * sizeof(float); }

asets_status asets_nn_head_bwd(const void *x, int64_t P, const float *w, const float *u0,
                               const float *pt0, const float *tu, const float *tpt, void *dx,
                               float *grads, void *scratch, size_t scratch_bytes,
                               void *stream) {
    if (P <= 0 || P * 16 >= kMaxElems) return nn_fail(ASETS_EBADARG, "bad voxel count");
    if (!scratch || scratch_bytes < asets_nn_head_scratch_bytes(P))
        return nn_fail(ASETS_EBADARG, "scratch too small");
    hipStream_t s = (hipStream_t)stream;
    const int nb = (int)cdiv(P, kB);
    float *part = (float *)scratch;
    k_head_bwd<<<nb, kB, 0, s>>>((const u16 *)x, (int)P, w, u0, pt0, tu, tpt, (u16 *)dx, part);
    sum_chunks(part, nb, 35, grads, 0, s);
    NNCHK();
    return ASETS_OK;
}

static asets_status s2d_geo(int n, int d, int h, int w, Geo5 *gc) {
    if (n <= 0 || d <= 0 || h <= 0 || w <= 0 || (d | h | w) & 1)
        return nn_fail(ASETS_EBADARG, "extents must be positive and even");
    if ((int64_t)n * d * h * w * 4 >= kMaxElems)
        return nn_fail(ASETS_ETOOBIG, "tensor of more than 2^31 elements");
    *gc = Geo5{n, d / 2, h / 2, w / 2};
    return ASETS_OK;
}

asets_status asets_nn_features_s2d(const float *Cs, const float *Ct, const float *alpha, int n,
                                   int d, int h, int w, void *out, void *stream) {
    Geo5 gc;
    asets_status st = s2d_geo(n, d, h, w, &gc);
    if (st) return st;
    if (!Cs || !Ct || !alpha || !out) return nn_fail(ASETS_EBADARG, "null pointer");
    if (((uintptr_t)Cs | (uintptr_t)Ct | (uintptr_t)alpha) & 7)
        return nn_fail(ASETS_EBADARG, "Cs, Ct and alpha must be 8-byte aligned");
    const int64_t P2 = (int64_t)gc.n * gc.d * gc.h * gc.w;
    k_features_s2d<<<nblk(P2), kB, 0, (hipStream_t)stream>>>(Cs, Ct, alpha, gc, (u16 *)out);
    NNCHK();
    return ASETS_OK;
}

asets_status asets_nn_head_s2d(const void *x, int n, int d, int h, int w, const float *wt,
                               const float *b, float *u0, float *bx0, float *by0, float *bz0,
                               void *stream) {
    Geo5 gc;
    asets_status st = s2d_geo(n, d, h, w, &gc);
    if (st) return st;
    if (!x || !wt || !b || !u0 || !bx0 || !by0 || !bz0)
        return nn_fail(ASETS_EBADARG, "null pointer");
    if (((uintptr_t)u0 | (uintptr_t)bx0 | (uintptr_t)by0 | (uintptr_t)bz0) & 7)
        return nn_fail(ASETS_EBADARG, "output fields must be 8-byte aligned");
    if ((uintptr_t)wt & 15) return nn_fail(ASETS_EBADARG, "head weights must be 16-byte aligned");
    const int64_t P2 = (int64_t)gc.n * gc.d * gc.h * gc.w;
    k_head_s2d_fwd<<<nblk(P2), kB, 0, (hipStream_t)stream>>>((const u16 *)x, gc, wt, b, u0, bx0,
                                                             by0, bz0);
    NNCHK();
    return ASETS_OK;
}

size_t asets_nn_head_s2d_scratch_bytes(int n, int d, int h, int w) {
    const int64_t P2 = (int64_t)n * (d / 2) * (h / 2) * (w / 2);
    if (P2 <= 0) return 0;
    return (size_t)(P2 * 33 + cdiv(P2, kHeadChunk) * kHeadGrads) * sizeof(float);
}

asets_status asets_nn_head_s2d_bwd(const void *x, int n, int d, int h, int w, const float *wt,
                                   const float *u0, const float *bx0, const float *by0,
                                   const float *bz0, const float *tu, const float *tbx,
                                   const float *tby, const float *tbz, float lam, void *dx,
                                   float *grads, void *scratch, size_t scratch_bytes,
                                   void *stream) {
    Geo5 gc;
    asets_status st = s2d_geo(n, d, h, w, &gc);
    if (st) return st;
    if (!x || !wt || !u0 || !bx0 || !by0 || !bz0 || !tu || !tbx || !tby || !tbz || !dx ||
        !grads)
        return nn_fail(ASETS_EBADARG, "null pointer");
    if (!scratch || scratch_bytes < asets_nn_head_s2d_scratch_bytes(n, d, h, w))
        return nn_fail(ASETS_EBADARG, "scratch too small");
    hipStream_t s = (hipStream_t)stream;
    const int64_t P2 = (int64_t)gc.n * gc.d * gc.h * gc.w;
    float *gbuf = (float *)scratch;
    float *lossv = gbuf + P2 * 32;
    float *part = lossv + P2;
    k_head_s2d_bwd<<<nblk(P2), kB, 0, s>>>((const u16 *)x, gc, wt, u0, bx0, by0, bz0, tu, tbx,
                                           tby, tbz, lam, (u16 *)dx, gbuf, lossv);
    const int nch = (int)cdiv(P2, kHeadChunk);
    k_head_s2d_wgrad<<<nch, 576, 0, s>>>((const u16 *)x, gbuf, lossv, (int)P2, part);
    sum_chunks(part, nch, kHeadGrads, grads, 0, s);
    NNCHK();
    return ASETS_OK;
}

asets_status asets_nn_warm_state(float *u0, float *bx0, float *by0, float *bz0, const float *Cs,
                                 const float *Ct, int64_t ny, int64_t nx, int64_t nz, int n,
                                 float *ps, float *pt, void *stream) {
    if (ny <= 0 || nx <= 0 || nz <= 0 || n <= 0 || ny * nx >= kMaxElems || nz >= kMaxElems ||
        ny * nx * nz >= ((int64_t)1 << 40) / n)
        return nn_fail(ASETS_EBADARG, "bad extents");
    if (!u0 || !bx0 || !by0 || !bz0 || !Cs || !Ct || !ps || !pt)
        return nn_fail(ASETS_EBADARG, "null pointer");
    const int64_t total = ny * nx * nz * n;
    k_warm_state<<<(unsigned)cdiv(total, kB), kB, 0, (hipStream_t)stream>>>(
        u0, bx0, by0, bz0, Cs, Ct, (int)ny, (int)nx, (int)nz, total, ps, pt);
    NNCHK();
    return ASETS_OK;
}

asets_status asets_nn_adamw(float *w, void *wb, const float *g, float *m, float *v, int64_t n,
                            float lr, float b1, float b2, float eps, float wd, int step,
                            void *stream) {
    if (n <= 0 || n >= kMaxElems || step < 1) return nn_fail(ASETS_EBADARG, "bad arguments");
    const float bc1 = 1.0f - std::pow(b1, (float)step), bc2 = 1.0f - std::pow(b2, (float)step);
    k_adamw<<<nblk(n), kB, 0, (hipStream_t)stream>>>(w, (u16 *)wb, g, m, v, (int)n, lr, b1, b2,
                                                     eps, wd, bc1, bc2);
    NNCHK();
    return ASETS_OK;
}

asets_status asets_nn_to_bf16(const float *a, void *b, int64_t n, void *stream) {
    if (n <= 0 || n >= kMaxElems) return nn_fail(ASETS_EBADARG, "bad element count");
    k_f32_to_bf16<<<nblk(n), kB, 0, (hipStream_t)stream>>>(a, (u16 *)b, (int)n);
    NNCHK();
    return ASETS_OK;
}

}  // extern "C"
