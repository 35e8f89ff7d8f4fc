// kernels_fused.h -- one kernel per iteration of the binary continuous max-flow
// (asetsBinaryMF3D_mex.c:163-190: updateP1, updatePX/PY/PZ, projStep,
// updateBX/BY/BZ, updateDIVPSPTU), fused by streaming along z.
//
// Dependency radius of one iteration at voxel v (L1 distance):
//   gk  = dv - ((ps - pt) + u/cc)        within radius 2 of v
//   b1  = steps*grad(gk) + b             faces within radius 1
//   g   = 1/max(|b1|/alpha, 1)           within radius 1
//   b2  = 0.5*(g_lo + g_hi)*b1           v's six faces
//   dv, ps, pt, u                        at v
// dv at the start of an iteration equals div(b) of the stored faces (computed
// from them with the same expression at the end of the previous iteration), so
// it is recomputed, not stored.  Per voxel-iteration the compulsory traffic is
// reading Cs, Ct, alpha and u, ps, pt, bx, by, bz and writing the latter six:
// 60 bytes in fp32.
//
// Work decomposition.  A workgroup (NT threads) owns an in-plane tile of
// TY (=64, the contiguous y axis) x TX voxels and a range of z planes, and
// streams through the planes.  Every raw plane is loaded from HBM once per
// workgroup; its loads are issued at the end of the previous step and consumed
// after the step's output stage.  The output stage's own operands are issued
// right after the previous output stage, a whole step ahead.  Per plane step:
//   S1 (plane s+2)  gk over the tile + 2-voxel halo -> LDS GK ring (2), and
//                   the stepped z face below plane s+2 -> LDS B1Z ring (3)
//   S2a(plane s+2)  stepped in-plane faces b1 from gk -> LDS B1X/B1Y ring (3)
//   S2b(plane s+1)  projection factor g over tile + 1 halo -> LDS G ring (2)
//   S3 (plane s)    rescale, divergence, source/sink/multiplier update, store
// in two barrier phases: {S3(s-1), S1(s+2)} | {S2a(s+2), S2b(s+1)}.  The
// rescaled z face below the current output plane is carried in registers.
// Halo values are recomputed by neighbouring workgroups from identical inputs in
// the same operation order (same bits).
//
// Global accesses go through one buffer descriptor per field (size = the
// field) with the plane offset in the bounds-checked vector offset: reads of a
// plane outside the volume, or past the end of the field, return 0 -- exactly
// the zero-flux value of a high face past the last voxel.  The state is
// double-buffered (in -> out) because neighbouring workgroups read the halo of
// the previous iterate.  The first iteration of a cold start reads no state: it
// forms the initial cut from Cs and Ct (fused_tile's FIRST).  The single-volume
// 3D kernels store with the streaming cache policy (kStream).
#pragma once
#include <hip/hip_fp16.h>

#include <type_traits>

#include "common.h"

namespace asets {

// u is always fp32; the flows ps, pt, bx, by, bz are stored as T (float, or
// __half for the fp16-storage variant; arithmetic stays fp32 either way).
template <typename T>
struct FusedState {
    float *u;
    T *ps, *pt, *bx, *by, *bz;
};

// Stop test of iteration `it`, folded into the next iteration's launch (solves
// with err_bound 0, where no launch waits on it): the partial |e| sums of `it`,
// its cvg slot and the stop outputs.  partials == nullptr: none.
struct DeferredStop {
    const double *partials;
    int64_t nparts;
    double denom;
    int32_t it;
    float err_bound;
    float *cvg;
    int32_t *it_num;
    Flags *flags;
};

template <typename T>
struct FusedArgs {
    const float *Cs, *Ct, *alpha;
    FusedState<T> in, out;
    int ny, nx, nz;
    int plane;              // ny*nx (< 2^30)
    float cc, steps;
    int zchunk;             // planes per workgroup
    int ntile_y, ntile_x, nzc;
    int nblocks;            // ntile_y*ntile_x*nzc (logical)
    int fast_cc;            // cc is in the range where div_fast is exact
    double *partials;       // [gridDim.x]
    const Flags *flags;
    // cold-start first iteration: set to 1 when alpha <= 0 at some voxel (stop: also
    // set, when not null, so that the following iterations return at once)
    int32_t *alpha_bad, *alpha_stop;
    DeferredStop stop;      // the previous iteration's stop test (kf_binary)
    int stop_all;           // 1: every workgroup runs it at its start (else workgroup 0, last)
    float *take_u;          // stop_all: planar u that the launch finding the stop writes
    // block-interleaved binary state (BlkLayout; in.u / out.u are the buffers' bases)
    int nyb;                // 64-row blocks per column: ceil(ny / 64)
    int pb;                 // bytes per plane of the state: nx * nyb * block bytes
};

// ---- buffer access --------------------------------------------------------

// A plane of a field: a descriptor over the whole field (num_records = the
// field's bytes) and the element offset of plane q.  Reads go through the
// vector offset, which the hardware bounds-checks: a plane outside [0, nz)
// (negative offsets wrap to above 2^31) and a position past the field read 0 --
// the zero-flux value of a face past the volume.  Valid while
// (N + 2*plane) * bytes < 2^31 (fused_fits on the host).
struct PRef {
    __amdgpu_buffer_rsrc_t r;
    int b;   // element offset of the plane
};

__device__ __forceinline__ PRef pref(const void *base, int q, int nz, int bytes_per_elem,
                                     int plane) {
    // num_records in unsigned arithmetic: an 8-byte pair field may pass 2^31 bytes
    const unsigned bytes = (unsigned)nz * (unsigned)plane * (unsigned)bytes_per_elem;
    return PRef{__builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, (int)bytes, 0x00020000),
                q * plane};
}

// Byte offset of element `e` as the 32-bit vector offset of a buffer access,
// computed modulo 2^32 (a negative element wraps above the field and reads 0).
__device__ __forceinline__ int boff(int e, unsigned bytes) { return (int)((unsigned)e * bytes); }

// Element index whose byte offset (0xfffffff0) is past every field the fused
// kernels address: a store there is dropped by the buffer bounds check.  Used to
// keep stores of lanes outside the volume unconditional (no wave-level branch
// around them).
__device__ __forceinline__ int drop_elem(unsigned bytes) { return (int)(0xfffffff0u / bytes); }

// Cache-policy operand of the buffer instructions: 0, or non-temporal
// (streaming).  The single-volume 3D kernels store their output planes, which no
// workgroup reads back within the launch, and load the output stage's Cs and Ct,
// which only their own workgroup reads, as streaming: they then displace less of
// the raw planes' halo that neighbouring tiles re-read from L2.  (Measured: at
// 256^3 the binary kernel reads 10 % fewer bytes and runs 3.7 % faster.  Where the
// working set stays in the caches from one launch to the next -- 2D, and the
// batched kernel's many small volumes -- streaming stores measured slower, so
// those kernels use 0.)
constexpr int kStream = 2;

template <typename T> struct Io;
template <> struct Io<float> {
    static constexpr int kBytes = 4;
    // the b32 buffer builtins move raw 32-bit words: bit-cast, never convert
    static __device__ __forceinline__ float load(__amdgpu_buffer_rsrc_t r, int off) {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, boff(off, 4), 0, 0));
    }
    template <int CP = 0>
    static __device__ __forceinline__ void store(__amdgpu_buffer_rsrc_t r, int off, float v) {
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, boff(off, 4), 0, CP);
    }
    static __device__ __forceinline__ float get(const float *p) { return *p; }
    static __device__ __forceinline__ void put(float *p, float v) { *p = v; }
    // the value a store and a load give back
    static __device__ __forceinline__ float stored(float v) { return v; }
};
template <> struct Io<__half> {
    static constexpr int kBytes = 2;
    static __device__ __forceinline__ float load(__amdgpu_buffer_rsrc_t r, int off) {
        unsigned short b = __builtin_amdgcn_raw_buffer_load_b16(r, boff(off, 2), 0, 0);
        return __half2float(__ushort_as_half(b));
    }
    template <int CP = 0>
    static __device__ __forceinline__ void store(__amdgpu_buffer_rsrc_t r, int off, float v) {
        __builtin_amdgcn_raw_buffer_store_b16(__half_as_ushort(__float2half_rn(v)), r,
                                              boff(off, 2), 0, CP);
    }
    static __device__ __forceinline__ float get(const __half *p) { return __half2float(*p); }
    static __device__ __forceinline__ void put(__half *p, float v) { *p = __float2half_rn(v); }
    static __device__ __forceinline__ float stored(float v) {
        return __half2float(__float2half_rn(v));
    }
};

__device__ __forceinline__ float ldf(__amdgpu_buffer_rsrc_t r, int off) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, boff(off, 4), 0, 0));
}
__device__ __forceinline__ float ldf(const PRef &p, int off) { return ldf(p.r, p.b + off); }
// fp32 load with cache policy CP
template <int CP>
__device__ __forceinline__ float ldf_cp(const PRef &p, int off) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(p.r, boff(p.b + off, 4), 0, CP));
}
template <typename IO>
__device__ __forceinline__ float ld(const PRef &p, int off) { return IO::load(p.r, p.b + off); }
template <typename IO, int CP = 0>
__device__ __forceinline__ void st(const PRef &p, int off, float v) {
    IO::template store<CP>(p.r, p.b + off, v);
}
// One in-plane offset for all stores of an output position to fields of one plane
// (plane offset b): invalid lanes get the element 2^29-1 relative to the field,
// past the end of every field of 2-, 4- or 8-byte elements that fused_fits allows
// (N < 2^29), so all their stores are dropped with one select per position.
__device__ __forceinline__ int masked_off(bool ok, int off, int b) {
    return ok ? off : 0x1fffffff - b;
}

// The binary model's state keeps (ps, pt) and (bx, by) interleaved, one pair per
// voxel (the first of the two at the lower address): a pair is one 64-bit (fp32)
// or 32-bit (fp16) access, so a plane needs 6 loads and 4 stores per position
// instead of 8 and 6.  Element stride of ps, pt, bx, by in binary state: kPair;
// bz, u and the Potts and Ishikawa fields are not paired.
constexpr int kPair = 2;

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

template <typename T>
__device__ __forceinline__ float2 ld_pair(const PRef &p, int off);
template <>
__device__ __forceinline__ float2 ld_pair<float>(const PRef &p, int off) {
    const u32x2 w = __builtin_amdgcn_raw_buffer_load_b64(p.r, boff(p.b + off, 8), 0, 0);
    return make_float2(__uint_as_float(w.x), __uint_as_float(w.y));
}
template <>
__device__ __forceinline__ float2 ld_pair<__half>(const PRef &p, int off) {
    const uint32_t w = __builtin_amdgcn_raw_buffer_load_b32(p.r, boff(p.b + off, 4), 0, 0);
    return make_float2(__half2float(__ushort_as_half((unsigned short)(w & 0xffffu))),
                       __half2float(__ushort_as_half((unsigned short)(w >> 16))));
}
// pair store with cache policy CP; ok == false stores past the field (dropped)
template <typename T, int CP = 0>
__device__ __forceinline__ void st_pair(const PRef &p, int off, float lo, float hi,
                                        bool ok = true) {
    if constexpr (std::is_same<T, float>::value) {
        u32x2 w;
        w.x = __float_as_uint(lo);
        w.y = __float_as_uint(hi);
        __builtin_amdgcn_raw_buffer_store_b64(w, p.r, boff(ok ? p.b + off : drop_elem(8), 8), 0,
                                              CP);
    } else {
        const uint32_t w = (uint32_t)__half_as_ushort(__float2half_rn(lo)) |
                           ((uint32_t)__half_as_ushort(__float2half_rn(hi)) << 16);
        __builtin_amdgcn_raw_buffer_store_b32(w, p.r, boff(ok ? p.b + off : drop_elem(4), 4), 0,
                                              CP);
    }
}

// ---- block-interleaved binary state ---------------------------------------
// The fused binary solver (kf_binary) stores a state buffer as one array of
// blocks: block (x, z, yb) holds the 64 voxels y = 64*yb .. 64*yb + 63 of column
// x of plane z, field after field -- u (fp32), the (ps, pt) pairs, the (bx, by)
// pairs, bz -- so a plane of the whole state is one contiguous run and an
// iteration reads one state array and writes one (a plain stream of the
// iteration's bytes through five arrays instead of eleven runs about 11 % faster
// on MI355X).  Block index (z * nx + x) * nyb + yb, nyb = ceil(ny / 64); the lanes
// y >= ny of a column's last block are padding whose faces are kept at 0.  A high
// face past the volume reads 0 as in the planar layout: the y + 1 face of the last
// row is the next column's y = 0 face (or a padding lane), the x + 1 face of the
// last column the next plane's x = 0 face, a plane past the volume lies past the
// buffer (bounds check).
template <typename T>
struct BlkLayout {
    static constexpr int E = (int)sizeof(T);
    static constexpr int kU = 0;                 // u: 64 x 4 B
    static constexpr int kPS = 256;              // (ps, pt): 64 x 2E
    static constexpr int kBB = 256 + 128 * E;    // (bx, by): 64 x 2E
    static constexpr int kBZ = 256 + 256 * E;    // bz: 64 x E
    static constexpr int kBytes = 256 + 320 * E; // block bytes (fp32 1536, fp16 896)
};

// The fused Potts solver (kf_potts_flow) keeps each label's flows the same way:
// block (x, z, yb) of label l holds pt, bx, by and bz of its 64 voxels, 1 KiB,
// with the label's blocks one contiguous run; u, ps, Ct and alpha stay planar.
struct PottsBlkLayout {
    static constexpr int kPT = 0, kBX = 256, kBY = 512, kBZ = 768;
    static constexpr int kBytes = 1024;
};

// raw buffer access at a byte offset (unsigned: negative planes wrap above the buffer)
__device__ __forceinline__ float ldb_f32(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
template <typename T>
__device__ __forceinline__ float ldb_t(__amdgpu_buffer_rsrc_t r, unsigned off) {
    if constexpr (std::is_same<T, float>::value) {
        return ldb_f32(r, off);
    } else {
        const unsigned short b = __builtin_amdgcn_raw_buffer_load_b16(r, (int)off, 0, 0);
        return __half2float(__ushort_as_half(b));
    }
}
template <typename T>
__device__ __forceinline__ float2 ldb_pair(__amdgpu_buffer_rsrc_t r, unsigned off) {
    if constexpr (std::is_same<T, float>::value) {
        const u32x2 w = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0);
        return make_float2(__uint_as_float(w.x), __uint_as_float(w.y));
    } else {
        const uint32_t w = __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0);
        return make_float2(__half2float(__ushort_as_half((unsigned short)(w & 0xffffu))),
                           __half2float(__ushort_as_half((unsigned short)(w >> 16))));
    }
}
template <int CP>
__device__ __forceinline__ void stb_f32(__amdgpu_buffer_rsrc_t r, unsigned off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)off, 0, CP);
}
template <typename T, int CP>
__device__ __forceinline__ void stb_t(__amdgpu_buffer_rsrc_t r, unsigned off, float v) {
    if constexpr (std::is_same<T, float>::value)
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)off, 0, CP);
    else
        __builtin_amdgcn_raw_buffer_store_b16(__half_as_ushort(__float2half_rn(v)), r, (int)off,
                                              0, CP);
}
template <typename T, int CP>
__device__ __forceinline__ void stb_pair(__amdgpu_buffer_rsrc_t r, unsigned off, float lo,
                                         float hi) {
    if constexpr (std::is_same<T, float>::value) {
        u32x2 w;
        w.x = __float_as_uint(lo);
        w.y = __float_as_uint(hi);
        __builtin_amdgcn_raw_buffer_store_b64(w, r, (int)off, 0, CP);
    } else {
        const uint32_t w = (uint32_t)__half_as_ushort(__float2half_rn(lo)) |
                           ((uint32_t)__half_as_ushort(__float2half_rn(hi)) << 16);
        __builtin_amdgcn_raw_buffer_store_b32(w, r, (int)off, 0, CP);
    }
}
// descriptor over a whole block-interleaved state buffer of nz planes of pb bytes
__device__ __forceinline__ __amdgpu_buffer_rsrc_t blk_rsrc(const void *base, int nz, int pb) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0,
                                             (int)((unsigned)nz * (unsigned)pb), 0x00020000);
}
// Base offset of a store to drop: past every block-interleaved buffer fused_fits_blk
// allows (< 2^31 bytes), and adding a field and lane offset (< 4096) stays below 2^32,
// so the bounds check drops it.
constexpr unsigned kBlkDrop = 0xfffff000u;

// ---- correctly rounded division, short form ------------------------------
// The compiler's IEEE fp32 division is: refine a hardware reciprocal of the
// (scaled) divisor with two FMAs, multiply, then two remainder/correction FMA
// steps, a scale-back and a special-case fixup.  When the exponents of x and d
// are moderate the scaling is the identity and the fixup passes the value
// through, so the same sequence without them returns the same correctly
// rounded quotient.  div_fast applies it when |x| is in [2^-64, 2^64] and d
// (precondition, checked by the caller) is in [2^-16, 2^16], and falls back to
// '/' otherwise (tiny, huge, inf, NaN).  rcp_refined(d) is the refined
// reciprocal of that sequence.
__device__ __forceinline__ float rcp_refined(float d) {
    const float e0 = __builtin_amdgcn_rcpf(d);
    const float t = __builtin_fmaf(-d, e0, 1.0f);
    return __builtin_fmaf(t, e0, e0);
}

__device__ __forceinline__ bool div_fast_ok(float d) {
    const float ad = __builtin_fabsf(d);
    return ad >= 0x1p-16f && ad <= 0x1p16f;
}

// True on every active lane of the wavefront: lets a rare exact fallback be a
// wave-uniform branch instead of exec-masked divergent code.
__device__ __forceinline__ bool wave_all(bool p) { return __all(p ? 1 : 0) != 0; }

// x/d with d's refined reciprocal r; exact fallback to '/' only when some lane
// has x outside the short form's range (or d outside it: d_ok false).
__device__ __forceinline__ float div_fast(float x, float d, float r, bool d_ok = true) {
    const float ax = __builtin_fabsf(x);
    const bool ok = d_ok && ((ax >= 0x1p-64f && ax <= 0x1p64f) || ax == 0.0f);
    const float q0 = x * r;
    const float e0 = __builtin_fmaf(-d, q0, x);
    const float q1 = __builtin_fmaf(e0, r, q0);
    const float e1 = __builtin_fmaf(-d, q1, x);
    // x = +-0: x*r already is the signed zero x/d (the corrections could turn
    // -0 into +0)
    float q = (ax == 0.0f) ? q0 : __builtin_fmaf(e1, r, q1);
    if (!wave_all(ok)) q = ok ? q : x / d;
    return q;
}

// Projection factor (asetsBinaryMF3D_mex.c:321-331).  The sum of squares is
// exact-per-term in double and accumulated high face first, as the reference's
// pow() chain.  If t*0.5 <= alpha^2 then sqrt(t*0.5) <= alpha, the correctly
// rounded double and then float square root cannot exceed alpha, the branch
// n > alpha is not taken and the factor is exactly 1: the square root and both
// divisions are needed only when some lane of the wave is above that bound.
// Lanes that do not use a division feed it benign operands (1) so they never
// trigger the exact fallback.
//
// There is no fp32 screen ahead of the FP64 sum: non-packed FP64 FMAs issue at
// the fp32 rate on gfx950, so a screen could save at most the six FP64 FMAs of a
// wave it clears, while costing as many fp32 operations on every wave; on the
// 256^3 workload almost no wave is cleared (most hold a voxel on the |p| = alpha
// boundary), and dropping the screen measured 2.4 % faster.
template <bool IS3D>
__device__ __forceinline__ float proj_factor_fast(float xr, float xl, float yr, float yl,
                                                  float zr, float zl, float a) {
    double t = (double)xr * (double)xr;
    t = __builtin_fma((double)xl, (double)xl, t);
    t = __builtin_fma((double)yr, (double)yr, t);
    t = __builtin_fma((double)yl, (double)yl, t);
    if (IS3D) {
        t = __builtin_fma((double)zr, (double)zr, t);
        t = __builtin_fma((double)zl, (double)zl, t);
    }
    const double h = t * 0.5;
    const double ad = (double)a;
    const bool maybe = !(h <= ad * ad);
    float g = 1.0f;
    if (!wave_all(!maybe)) {
        const float n = (float)__builtin_sqrt(h);
        const bool big = maybe && (n > a);
        const float nn = big ? n : 1.0f, aa = big ? a : 1.0f;
        const float q = div_fast(nn, aa, rcp_refined(aa), div_fast_ok(aa));   // n/alpha
        const float qq = big ? q : 1.0f;
        const float r = div_fast(1.0f, qq, rcp_refined(qq), div_fast_ok(qq)); // 1/(n/alpha)
        g = big ? r : 1.0f;
    }
    return g;
}

// Face rescale (g_hi + g_lo)*0.5*b with the reference's double literal: one
// rounding of the exact product.  (s*0.5f)*b gives the same float whenever the
// halving is exact, i.e. unless 0 < |s| < 2^-125.  Projection factors are >= 0,
// so s >= max(g_hi, g_lo): a voxel whose own factor is >= 2^-124 can take the
// float form for all six of its faces (rescale_f32); otherwise rescale_f64.
__device__ __forceinline__ float rescale_f32(float g_hi, float g_lo, float b) {
    return ((g_hi + g_lo) * 0.5f) * b;
}
__device__ __forceinline__ float rescale_f64(float g_hi, float g_lo, float b) {
    return (float)((double)(g_hi + g_lo) * 0.5 * (double)b);
}
// rescale with the per-voxel choice made wave-uniform in the common case
__device__ __forceinline__ float rescale(float g_hi, float g_lo, float b, bool tiny,
                                         bool any_tiny) {
    const float f = rescale_f32(g_hi, g_lo, b);
    if (any_tiny) return tiny ? rescale_f64(g_hi, g_lo, b) : f;
    return f;
}

// ---- the kernel -------------------------------------------------------------

// One work item: tile lb (of a.ntile_y x a.ntile_x tiles x a.nzc z chunks) of
// one volume, all its planes.  Returns this thread's share of sum |e|.  Ends
// with a barrier, so a workgroup can run several items back to back.
// POTTS: the same flow update for one label of the Potts model
// (asetsPotts3D_mex.c:260-471, label k): in.ps is the label-shared source flow,
// and the output stage instead updates the sink flow pt_k (:457-466) and writes
// ft_k = dv_k + pt_k (the first term of updatePSU, :498) to out.ps; u is only
// read.  The coupling over labels runs afterwards (kp_potts_couple).
// MODE 2 (Ishikawa, asetsIshikawa3D_mex.c:184-199, label k): in.ps holds pt_k and
// in.pt holds pt_{k+1} (gk = dv - ((pt_k - pt_{k+1}) + u_k/cc), :287-288); the
// output stage stores only the faces and dv_k (:470-496) to out.ps; the sink-flow
// chain and the multipliers run afterwards (k_ishikawa_couple).
// FIRST (binary only): the first iteration of a cold start.  The state it reads
// is the initial cut of asetsBinaryMF3D_mex.c:215-219 -- u = (Cs - Ct >= 0),
// ps = pt = min(Cs, Ct) (rounded as the flow storage type stores it), all flows
// 0 -- formed in registers from Cs and Ct instead of loaded, and the alpha
// values it loads are checked (> 0, asetsBinaryMF3D_mex.c:317-319).
// STREAM: output stores and the output stage's Cs/Ct loads are non-temporal (kStream).
// BLK: the state is block-interleaved -- binary: BlkLayout, a.in.u and a.out.u the
// two buffers; Potts: the label's flows in PottsBlkLayout, a.in.pt and a.out.pt the
// two buffers -- while Cs, Ct, alpha (and for Potts u, ps and ft) stay planar.
template <bool IS3D, int TX, int NT, typename T, int MODE = 0, bool FIRST = false,
          bool STREAM = IS3D, bool BLK = false>
__device__ __forceinline__ double fused_tile(const FusedArgs<T> a, int lb) {
    using IO = Io<T>;
    using BL = BlkLayout<T>;
    static_assert(!BLK || MODE <= 1, "the block-interleaved state is the binary and Potts "
                                     "models'");
    using PL = PottsBlkLayout;
    constexpr int CP = STREAM ? kStream : 0;
    constexpr bool POTTS = (MODE == 1), ISHI = (MODE == 2);
    static_assert(!FIRST || MODE == 0, "the cold-start iteration is the binary model's");
    constexpr bool PAIR = (MODE == 0);   // binary state: (ps, pt), (bx, by) interleaved
    constexpr bool PRIO = (MODE == 0);   // loads issued at raised wave priority
    constexpr bool ONEB = (MODE != 0);   // one barrier per plane step (pipeline below)
    constexpr int PB = 2 * IO::kBytes;   // bytes of a pair
    static_assert(PAIR || IO::kBytes == 4, "fp16 flow storage is for the binary model");
    constexpr int TY = 64;
    constexpr int W2 = TY + 4, H2 = TX + 4;   // gk / b1 region (tile + radius 2)
    constexpr int W1 = TY + 2, H1 = TX + 2;   // g region (tile + radius 1)
    constexpr int N2 = W2 * H2, N1 = W1 * H1;
    constexpr int K2 = (N2 + NT - 1) / NT;    // radius-2 positions per thread
    constexpr int K1 = (N1 + NT - 1) / NT;    // radius-1 positions per thread
    constexpr int KO = (TY * TX) / NT;        // output voxels per thread
    constexpr int NH = N2 - TY * TX;          // halo positions of the radius-2 region
    constexpr int NH1 = N1 - TY * TX;         // halo positions of the radius-1 region
    static_assert((TY * TX) % NT == 0, "tile must be a multiple of the block");
    static_assert(KO + (NH + NT - 1) / NT == K2, "radius-2 slots: own positions, then halo");
    static_assert(KO + (NH1 + NT - 1) / NT == K1, "radius-1 slots: own positions, then halo");

    __shared__ float GK[2][N2];
    __shared__ float B1X[3][N2], B1Y[3][N2], B1Z[IS3D ? 3 : 1][IS3D ? N2 : 1];
    __shared__ float G[2][N1];
    const int tid = threadIdx.x;
    // radius-2 slot k holds a position (only the last halo slot can run past NH)
    auto in2 = [&](int k) { return k < KO || (NH % NT) == 0 || tid + (k - KO) * NT < NH; };
    auto in1 = [&](int k) { return k < KO || (NH1 % NT) == 0 || tid + (k - KO) * NT < NH1; };
    double e_acc = 0.0;
    bool abad = false;   // FIRST: alpha <= 0 seen
    {
        const int tyi = lb % a.ntile_y;
        const int txi = (lb / a.ntile_y) % a.ntile_x;
        const int zci = lb / (a.ntile_y * a.ntile_x);
        const int y0 = tyi * TY, x0 = txi * TX;
        const int z0 = zci * a.zchunk;
        const int z1 = min(z0 + a.zchunk, a.nz);
        const int ny = a.ny, nx = a.nx, nz = a.nz, plane = a.plane;
        const float cc = a.cc, steps = a.steps;
        const float rcc = rcp_refined(cc);
        const float one_cc = 1.0f / cc;
        const bool fast_cc = a.fast_cc != 0;

        // ---- per-thread position tables (fixed for all planes) ------------
        // radius-2 region (W2 x H2, region row i, column j): slots 0 .. KO-1
        // are the tile's own positions in the output stage's order (row
        // tid % TY, column tid / TY + k * (NT / TY)), so the u and pt a thread
        // loads for S1 are the ones its output stage uses three steps later;
        // slots KO .. K2-1 hold the NH halo positions, halo index h = tid +
        // (k - KO) * NT: the two columns either side of the tile (h < 4 * TY,
        // TY rows each), then the two rows either side (4 per column).
        // Per slot: in-plane offset, LDS index, and flags
        //   bit0 inside the volume, bit1 x>0, bit2 y>0, bit3 i>=1 && j>=1
        static_assert(K2 <= 16, "flag word holds 16 positions");
        int off2[K2];
        int l2h[K2 - KO];
        // BLK: byte offsets within a state plane of the slot's u, its pairs, its bz
        // and the pairs of the voxel at y + 1 (fp32: bz at the u offset; fp16: the
        // pairs at the u offset).  Potts: bu2 is the slot's offset for all four of
        // its fields and byp2 that of the voxel at y + 1.
        unsigned bu2[BLK ? K2 : 1], bp2[BLK ? K2 : 1], bz2[BLK ? K2 : 1], byp2[BLK ? K2 : 1];
        const int l2o = (2 + tid / TY) * W2 + 2 + tid % TY;   // LDS index of slot 0
        uint64_t fl2 = 0;
#pragma unroll
        for (int k = 0; k < K2; ++k) {
            int i, j;
            if (k < KO) {
                i = 2 + tid % TY;
                j = 2 + tid / TY + k * (NT / TY);
            } else {
                const int h = tid + (k - KO) * NT;
                if (h < 4 * TY) {
                    const int c = h / TY;
                    i = 2 + h % TY;
                    j = (c < 2) ? c : H2 - 4 + c;
                } else {
                    const int h2 = h - 4 * TY, r4 = h2 & 3;
                    i = (r4 < 2) ? r4 : W2 - 4 + r4;
                    j = in2(k) ? h2 >> 2 : 0;
                }
                l2h[k - KO] = j * W2 + i;
            }
            const int y = y0 - 2 + i, x = x0 - 2 + j;
            off2[k] = x * ny + y;
            if constexpr (BLK && POTTS) {
                const int yy = y + 64, yb = yy / 64 - 1, ln = yy % 64;   // y >= -2
                const unsigned ob = (unsigned)((x * a.nyb + yb) * PL::kBytes);
                bu2[k] = ob + 4u * (unsigned)ln;
                byp2[k] = (ln < 63) ? bu2[k] + 4u : ob + (unsigned)PL::kBytes;
            } else if constexpr (BLK) {
                const int yy = y + 64, yb = yy / 64 - 1, ln = yy % 64;   // y >= -2
                const unsigned ob = (unsigned)((x * a.nyb + yb) * BL::kBytes);
                bu2[k] = ob + 4u * (unsigned)ln;
                bp2[k] = ob + (unsigned)(2 * BL::E * ln);
                bz2[k] = ob + (unsigned)(BL::E * ln);
                byp2[k] = (ln < 63) ? bp2[k] + 2u * BL::E : ob + (unsigned)BL::kBytes;
            }
            const bool ok = in2(k);
            const bool in = ok && y >= 0 && y < ny && x >= 0 && x < nx;
            const uint64_t f = (in ? 1u : 0u) | (x > 0 ? 2u : 0u) | (y > 0 ? 4u : 0u) |
                               ((ok && i >= 1 && j >= 1) ? 8u : 0u);
            fl2 |= f << (4 * k);
        }
        auto l2 = [&](int k) { return k < KO ? l2o + k * (NT / TY) * W2 : l2h[k - KO]; };
        // radius-1 region (W1 x H1, row i1, column j1), slots ordered as the
        // radius-2 ones: own positions first (output order), then the NH1 halo
        // positions (the column either side, then the row either side, 2 per
        // column).  Per slot: offset, index in the G (N1) layout and in the N2
        // layout, inside flag.
        static_assert(K1 <= 32, "flag word holds 32 positions");
        int off1[K1], l1h[K1 - KO], l2i1h[K1 - KO];
        const int l1o = (1 + tid / TY) * W1 + 1 + tid % TY;   // G index of slot 0
        uint32_t fl1 = 0;
#pragma unroll
        for (int k = 0; k < K1; ++k) {
            int i1, j1;
            if (k < KO) {
                i1 = 1 + tid % TY;
                j1 = 1 + tid / TY + k * (NT / TY);
            } else {
                const int h = tid + (k - KO) * NT;
                if (h < 2 * TY) {
                    i1 = 1 + h % TY;
                    j1 = (h < TY) ? 0 : H1 - 1;
                } else {
                    const int h2 = h - 2 * TY;
                    i1 = (h2 & 1) ? W1 - 1 : 0;
                    j1 = in1(k) ? h2 >> 1 : 0;
                }
                l1h[k - KO] = j1 * W1 + i1;
                l2i1h[k - KO] = (j1 + 1) * W2 + (i1 + 1);
            }
            const int y = y0 - 1 + i1, x = x0 - 1 + j1;
            off1[k] = x * ny + y;
            const bool in = in1(k) && y >= 0 && y < ny && x >= 0 && x < nx;
            fl1 |= (in ? 1u : 0u) << k;
        }
        auto l1 = [&](int k) { return k < KO ? l1o + k * (NT / TY) * W1 : l1h[k - KO]; };
        auto l2i1 = [&](int k) { return k < KO ? l2(k) : l2i1h[k - KO]; };

        // ---- raw plane registers (prefetched one step ahead) --------------
        struct Raw {
            float u[K2], ps[K2], pt[K2], bxl[K2], bxr[K2], byl[K2], byr[K2], bzn[K2];
            float al[K1];
            float uc[KO];   // u/cc of the own positions, formed by S1 (kept for S3)
        };
        // bz own face of the plane S1 is working on; the next plane's own bz
        // (loaded with the raw plane) becomes it one step later.
        float bz_cur[K2];
#pragma unroll
        for (int k = 0; k < K2; ++k) bz_cur[k] = 0.0f;

        // load raw plane q (u, ps, pt, bx own/high, by own/high), bz own of
        // plane q+1, and alpha of plane q-1 (for S2b one step later)
        auto load_raw = [&](int q, Raw &r) {
            if constexpr (FIRST) {
                const PRef rcs = pref(a.Cs, q, nz, 4, plane);
                const PRef rct = pref(a.Ct, q, nz, 4, plane);
#pragma unroll
                for (int k = 0; k < K2; ++k) {
                    const float cs = ldf(rcs, off2[k]), ct = ldf(rct, off2[k]);
                    r.u[k] = (cs - ct >= 0.0f) ? 1.0f : 0.0f;
                    r.ps[k] = r.pt[k] = IO::stored(fmin_ref(cs, ct));
                    r.bxl[k] = r.bxr[k] = r.byl[k] = r.byr[k] = r.bzn[k] = 0.0f;
                }
                const PRef ra = pref(a.alpha, q - 1, nz, 4, plane);
#pragma unroll
                for (int k = 0; k < K1; ++k) r.al[k] = ldf(ra, off1[k]);
                return;
            }
            if constexpr (BLK && POTTS) {
                const __amdgpu_buffer_rsrc_t rs = blk_rsrc(a.in.pt, nz, a.pb);
                const unsigned P = (unsigned)q * (unsigned)a.pb;
                const unsigned XR = (unsigned)(a.nyb * PL::kBytes);   // the next column
                const PRef ru = pref(a.in.u, q, nz, 4, plane);
                const PRef rps = pref(a.in.ps, q, nz, 4, plane);
#pragma unroll
                for (int k = 0; k < K2; ++k) {
                    r.u[k] = ldf(ru, off2[k]);
                    r.ps[k] = ldf(rps, off2[k]);
                    r.pt[k] = ldb_f32(rs, P + bu2[k] + PL::kPT);
                    r.bxl[k] = ldb_f32(rs, P + bu2[k] + PL::kBX);
                    r.bxr[k] = ldb_f32(rs, P + XR + bu2[k] + PL::kBX);
                    r.byl[k] = ldb_f32(rs, P + bu2[k] + PL::kBY);
                    r.byr[k] = ldb_f32(rs, P + byp2[k] + PL::kBY);
                }
                if (IS3D) {
#pragma unroll
                    for (int k = 0; k < K2; ++k)
                        r.bzn[k] = ldb_f32(rs, P + (unsigned)a.pb + bu2[k] + PL::kBZ);
                }
                const PRef ra = pref(a.alpha, q - 1, nz, 4, plane);
#pragma unroll
                for (int k = 0; k < K1; ++k) r.al[k] = ldf(ra, off1[k]);
                return;
            }
            if constexpr (BLK) {
                const __amdgpu_buffer_rsrc_t rs = blk_rsrc(a.in.u, nz, a.pb);
                const unsigned P = (unsigned)q * (unsigned)a.pb;
                const unsigned XR = (unsigned)(a.nyb * BL::kBytes);   // the next column
#pragma unroll
                for (int k = 0; k < K2; ++k) {
                    r.u[k] = ldb_f32(rs, P + bu2[k] + BL::kU);
                    const float2 pp = ldb_pair<T>(rs, P + bp2[k] + BL::kPS);
                    r.ps[k] = pp.x;
                    r.pt[k] = pp.y;
                    const float2 bb = ldb_pair<T>(rs, P + bp2[k] + BL::kBB);
                    r.bxl[k] = bb.x;
                    r.byl[k] = bb.y;
                    r.bxr[k] = ldb_pair<T>(rs, P + XR + bp2[k] + BL::kBB).x;
                    r.byr[k] = ldb_pair<T>(rs, P + byp2[k] + BL::kBB).y;
                }
                if (IS3D) {
#pragma unroll
                    for (int k = 0; k < K2; ++k)
                        r.bzn[k] = ldb_t<T>(rs, P + (unsigned)a.pb + bz2[k] + BL::kBZ);
                }
                const PRef ra = pref(a.alpha, q - 1, nz, 4, plane);
#pragma unroll
                for (int k = 0; k < K1; ++k) r.al[k] = ldf(ra, off1[k]);
                return;
            }
            const PRef ru = pref(a.in.u, q, nz, 4, plane);
            // the high x face of the last column is the next plane's (or past the
            // field's) x = 0 face, and the high y face of the last row the next
            // column's y = 0 face: both are 0
            if constexpr (PAIR) {
                const PRef rpp = pref(a.in.ps, q, nz, PB, plane);   // (ps, pt) pairs
                const PRef rbb = pref(a.in.bx, q, nz, PB, plane);   // (bx, by) pairs
#pragma unroll
                for (int k = 0; k < K2; ++k) {
                    const int o = off2[k];
                    r.u[k] = ldf(ru, o);
                    const float2 pp = ld_pair<T>(rpp, o);
                    r.ps[k] = pp.x;
                    r.pt[k] = pp.y;
                    const float2 bb = ld_pair<T>(rbb, o);
                    r.bxl[k] = bb.x;
                    r.byl[k] = bb.y;
                    r.bxr[k] = ld_pair<T>(rbb, o + ny).x;
                    r.byr[k] = ld_pair<T>(rbb, o + 1).y;
                }
            } else {
                const PRef rps = pref(a.in.ps, q, nz, IO::kBytes, plane);
                const PRef rpt = pref(a.in.pt, q, nz, IO::kBytes, plane);
                const PRef rbx = pref(a.in.bx, q, nz, IO::kBytes, plane);
                const PRef rby = pref(a.in.by, q, nz, IO::kBytes, plane);
#pragma unroll
                for (int k = 0; k < K2; ++k) {
                    const int o = off2[k];
                    r.u[k] = ldf(ru, o);
                    r.ps[k] = ld<IO>(rps, o);
                    r.pt[k] = ld<IO>(rpt, o);
                    r.bxl[k] = ld<IO>(rbx, o);
                    r.bxr[k] = ld<IO>(rbx, o + ny);
                    r.byl[k] = ld<IO>(rby, o);
                    r.byr[k] = ld<IO>(rby, o + 1);
                }
            }
            if (IS3D) {
                const PRef rbz = pref(a.in.bz, q + 1, nz, IO::kBytes, plane);
#pragma unroll
                for (int k = 0; k < K2; ++k) r.bzn[k] = ld<IO>(rbz, off2[k]);
            }
            const PRef ra = pref(a.alpha, q - 1, nz, 4, plane);
#pragma unroll
            for (int k = 0; k < K1; ++k) r.al[k] = ldf(ra, off1[k]);
        };

        // S1: gk of plane q over the radius-2 region, and the stepped z face
        // below plane q (needs gk of q-1 at the same position: kept in gkp)
        float gkp[K2];
#pragma unroll
        for (int k = 0; k < K2; ++k) gkp[k] = 0.0f;
        auto s1 = [&](int q, Raw &r) {
            float *gk = GK[q & 1];
            const int sl = (q + 3) % 3;
            const bool qin = q >= 0 && q < nz;
#pragma unroll
            for (int k = 0; k < K2; ++k) {
                const int idx = l2(k);
                const uint32_t f = (uint32_t)(fl2 >> (4 * k));
                const bool live = (f & 1) && qin;
                float d = ((r.byr[k] - r.byl[k]) + r.bxr[k]) - r.bxl[k];
                if (IS3D) d = (d + r.bzn[k]) - bz_cur[k];
                const float uc = div_fast(r.u[k], cc, rcc, fast_cc);
                if (k < KO) r.uc[k] = uc;
                const float v = live ? d - ((r.ps[k] - r.pt[k]) + uc) : 0.0f;
                if (in2(k)) {
                    gk[idx] = v;
                    if (IS3D) {
                        const float bz1 =
                            (live && q > 0) ? steps * (v - gkp[k]) + bz_cur[k] : 0.0f;
                        B1Z[sl][idx] = bz1;
                    }
                }
                gkp[k] = v;
                if (IS3D) bz_cur[k] = r.bzn[k];   // the next plane's own z face is current
            }
        };

        // S2a: stepped in-plane own faces of plane q (needs gk of q only)
        // fxl, fyl: the stored own faces bx, by of plane q (the raw plane S1(q) read)
        auto s2a = [&](int q, const float (&fxl)[K2], const float (&fyl)[K2]) {
            const float *gq = GK[q & 1];
            const int sl = (q + 3) % 3;
            const bool qin = q >= 0 && q < nz;
#pragma unroll
            for (int k = 0; k < K2; ++k) {
                const int idx = l2(k);
                const uint32_t f = (uint32_t)(fl2 >> (4 * k));
                const bool live = (f & 1) && qin;
                // positions on the first row/column of the region are never read
                if (in2(k)) {
                    const float c = gq[idx];
                    const float cx = gq[(f & 8) ? idx - W2 : idx];
                    const float cy = gq[(f & 8) ? idx - 1 : idx];
                    const float bx1 = (live && (f & 2)) ? steps * (c - cx) + fxl[k] : 0.0f;
                    const float by1 = (live && (f & 4)) ? steps * (c - cy) + fyl[k] : 0.0f;
                    B1X[sl][idx] = bx1;
                    B1Y[sl][idx] = by1;
                }
            }
        };

        // S2b: projection factor g of plane q over the radius-1 region
        // (alpha of plane q arrives with the raw plane q+1)
        // al: alpha of plane q; the own positions' g also go to g_hi (read by the
        // one-barrier pipeline's output stage instead of LDS, same thread)
        float g_lo[KO], g_hi[KO];   // g at the own positions: planes q-1 and q
#pragma unroll
        for (int k = 0; k < KO; ++k) g_lo[k] = g_hi[k] = 1.0f;
        auto s2b = [&](int q, const float (&al)[K1]) {
            const int sl = (q + 3) % 3, sn = (q + 4) % 3;
            float *go = G[q & 1];
            const bool qin = q >= 0 && q < nz;
#pragma unroll
            for (int k = 0; k < K1; ++k) {
                const int idx = l1(k);
                const int l = l2i1(k);
                const float xl = B1X[sl][l], xr = B1X[sl][l + W2];
                const float yl = B1Y[sl][l], yr = B1Y[sl][l + 1];
                float zl = 0.0f, zr = 0.0f;
                if (IS3D) {
                    zl = B1Z[sl][l];
                    zr = B1Z[sn][l];
                }
                const float v = proj_factor_fast<IS3D>(xr, xl, yr, yl, zr, zl, al[k]);
                const bool live = ((fl1 >> k) & 1) && qin;
                const float gv = live ? v : 1.0f;
                if (in1(k)) go[idx] = gv;
                if (k < KO) g_hi[k] = gv;
                if constexpr (FIRST) abad |= live && al[k] <= 0.0f;
            }
        };

        // S3 operands of plane s at the output positions, loaded one step
        // ahead (binary: Cs, Ct; Potts: -, Ct)
        struct Out {
            float cs[KO], ct[KO];
        };
        auto load_out = [&](int s, Out &o) {
            if constexpr (ISHI) return;   // the Ishikawa output stage needs none of them
            const int sv = (s >= z0 && s < z1) ? s : -1;   // outside: empty descriptors
            const PRef rcs = pref(a.Cs, POTTS ? -1 : sv, nz, 4, plane);
            const PRef rct = pref(a.Ct, sv, nz, 4, plane);
#pragma unroll
            for (int k = 0; k < KO; ++k) {
                const int off = off2[k];   // slot k < KO: output position k
                if (!POTTS) o.cs[k] = ldf_cp<CP>(rcs, off);
                o.ct[k] = ldf_cp<CP>(rct, off);
            }
        };
        // S3 operands kept from the raw plane S1 read (own-position slots):
        // binary u and pt, Potts u and ps (the label-shared source flow), and the
        // u/cc that S1 formed from the same u
        struct Keep {
            float u[KO], p[KO], uc[KO];
        };
        auto keep = [&](Keep &kp, const Raw &r) {
#pragma unroll
            for (int k = 0; k < KO; ++k) {
                kp.u[k] = r.u[k];
                kp.p[k] = POTTS ? r.ps[k] : r.pt[k];
                kp.uc[k] = r.uc[k];
            }
        };

        float carry[KO];   // rescaled z face below plane s
#pragma unroll
        for (int k = 0; k < KO; ++k) carry[k] = 0.0f;

        // the rescaled z face above plane s, at output position k
        auto z_face_up = [&](int s, float gc, float gup, int l2) {
            const int sn = (s + 4) % 3;
            if (!(IS3D && s + 1 < nz)) return 0.0f;
            const bool tiny = !(gc >= 0x1p-124f);
            return rescale(gup, gc, B1Z[sn][l2], tiny, !wave_all(!tiny));
        };

        // S3 light: only the carried z face for plane s+1 (step z0)
        auto s3_carry = [&](int s) {
            const float *g0 = G[s & 1];
            const float *g1 = G[(s + 1) & 1];
#pragma unroll
            for (int k = 0; k < KO; ++k) {
                const int i = tid % TY, j = tid / TY + k * (NT / TY);
                const int l2 = (j + 2) * W2 + (i + 2);
                const int l1 = (j + 1) * W1 + (i + 1);
                carry[k] = ONEB ? z_face_up(s, g_lo[k], g_hi[k], l2)
                                : z_face_up(s, g0[l1], g1[l1], l2);
            }
        };

        // S3: output plane s (z0 <= s < z1)
        auto s3 = [&](int s, const Out &o, const Keep &kp) {
            const int sl = (s + 3) % 3;
            const float *g0 = G[s & 1];
            const float *g1 = G[(s + 1) & 1];
            const PRef wu = pref((POTTS || ISHI) ? a.in.u : a.out.u, s, nz, 4, plane);
            const PRef wps = pref(a.out.ps, s, nz, PAIR ? PB : IO::kBytes, plane);
            const PRef wpt = pref(a.out.pt, s, nz, IO::kBytes, plane);
            const PRef wbx = pref(a.out.bx, s, nz, PAIR ? PB : IO::kBytes, plane);
            const PRef wby = pref(a.out.by, s, nz, IO::kBytes, plane);
            const PRef wbz = pref(a.out.bz, s, nz, IO::kBytes, plane);
#pragma unroll
            for (int k = 0; k < KO; ++k) {
                const int i = tid % TY, j = tid / TY + k * (NT / TY);
                const int y = y0 + i, x = x0 + j;
                const int l2 = (j + 2) * W2 + (i + 2);   // b1 (N2 layout)
                const int l1 = (j + 1) * W1 + (i + 1);   // g  (N1 layout)
                const float gc = ONEB ? g_lo[k] : g0[l1];
                const float bzr = z_face_up(s, gc, ONEB ? g_hi[k] : g1[l1], l2);
                const float b_xl = B1X[sl][l2], b_xr = B1X[sl][l2 + W2];
                const float b_yl = B1Y[sl][l2], b_yr = B1Y[sl][l2 + 1];
                const float g_xl = g0[l1 - W1], g_xr = g0[l1 + W1];
                const float g_yl = g0[l1 - 1], g_yr = g0[l1 + 1];
                const bool tiny = !(gc >= 0x1p-124f);
                const bool any_tiny = !wave_all(!tiny);
                float bxl = rescale(gc, g_xl, b_xl, tiny, any_tiny);
                float bxr = rescale(g_xr, gc, b_xr, tiny, any_tiny);
                float byl = rescale(gc, g_yl, b_yl, tiny, any_tiny);
                float byr = rescale(g_yr, gc, b_yr, tiny, any_tiny);
                // index-0 faces and high faces past the volume are zero
                bxl = (x > 0) ? bxl : 0.0f;
                bxr = (x + 1 < nx) ? bxr : 0.0f;
                byl = (y > 0) ? byl : 0.0f;
                byr = (y + 1 < ny) ? byr : 0.0f;
                float d = ((byr - byl) + bxr) - bxl;
                float bzl = 0.0f;
                if (IS3D) {
                    bzl = carry[k];
                    d = (d + bzr) - bzl;
                    carry[k] = bzr;
                }
                // lanes outside the volume store past the field (dropped)
                const bool ok = y < ny && x < nx;
                const int off = masked_off(ok, x * ny + y, wu.b);
                if constexpr (ISHI) {
                    st<IO, CP>(wps, off, d);
                    st<IO, CP>(wbx, off, bxl);
                    st<IO, CP>(wby, off, byl);
                    if (IS3D) st<IO, CP>(wbz, off, bzl);
                    continue;
                }
                const float uo = kp.u[k];
                const float uc = kp.uc[k];
                if (POTTS) {
                    // sink flow of this label (asetsPotts3D_mex.c:457-466), ft = dv + pt
                    const float t = (kp.p[k] + uc) - d;   // kp.p holds ps here
                    const float npt = fmin_ref(t, o.ct[k]);
                    st<IO, CP>(wps, off, d + npt);
                    if constexpr (BLK) {
                        // every lane of a column inside the volume stores its flows (the
                        // padding lanes' faces are 0 here); a column past it is dropped
                        const __amdgpu_buffer_rsrc_t ws = blk_rsrc(a.out.pt, nz, a.pb);
                        const unsigned ob = (x < nx) ? (unsigned)s * (unsigned)a.pb + bu2[k]
                                                     : kBlkDrop;
                        stb_f32<CP>(ws, ob + PL::kPT, npt);
                        stb_f32<CP>(ws, ob + PL::kBX, bxl);
                        stb_f32<CP>(ws, ob + PL::kBY, byl);
                        if (IS3D) stb_f32<CP>(ws, ob + PL::kBZ, bzl);
                        continue;
                    }
                    st<IO, CP>(wpt, off, npt);
                    st<IO, CP>(wbx, off, bxl);
                    st<IO, CP>(wby, off, byl);
                    if (IS3D) st<IO, CP>(wbz, off, bzl);
                } else {
                    // source/sink flows and multiplier (asetsBinaryMF3D_mex.c:419-432)
                    float f = ((kp.p[k] - uc) + d) + one_cc;
                    const float nps = fmin_ref(f, o.cs[k]);
                    f = (nps + uc) - d;
                    const float npt = fmin_ref(f, o.ct[k]);
                    const float e = cc * ((npt + d) - nps);
                    e_acc += ok ? fabs((double)e) : 0.0;
                    if constexpr (BLK) {
                        // every lane of a column inside the volume stores (the padding
                        // lanes y >= ny keep their faces at 0: they are 0 here); a
                        // column past the volume is dropped
                        const __amdgpu_buffer_rsrc_t ws = blk_rsrc(a.out.u, nz, a.pb);
                        const unsigned P = (unsigned)s * (unsigned)a.pb;
                        const bool okx = x < nx;
                        const unsigned ou = okx ? P + bu2[k] : kBlkDrop;
                        const unsigned op = okx ? P + bp2[k] : kBlkDrop;
                        const unsigned oz = okx ? P + bz2[k] : kBlkDrop;
                        stb_f32<CP>(ws, ou + BL::kU, uo - e);
                        stb_pair<T, CP>(ws, op + BL::kPS, nps, npt);
                        stb_pair<T, CP>(ws, op + BL::kBB, bxl, byl);
                        if (IS3D) stb_t<T, CP>(ws, oz + BL::kBZ, bzl);
                        continue;
                    }
                    st<Io<float>, CP>(wu, off, uo - e);
                    if constexpr (PAIR) {
                        st_pair<T, CP>(wps, off, nps, npt);
                        st_pair<T, CP>(wbx, off, bxl, byl);
                    } else {
                        st<IO, CP>(wps, off, nps);
                        st<IO, CP>(wpt, off, npt);
                        st<IO, CP>(wbx, off, bxl);
                        st<IO, CP>(wby, off, byl);
                    }
                    if (IS3D) st<IO, CP>(wbz, off, bzl);
                }
            }
        };

        // ---- pipeline ----------------------------------------------------
        // Step s, phase X: S3(s-1), the loads of the S3 operands of plane s,
        // and S1(s+2); barrier; phase Y: S2a(s+2) and S2b(s+1), then the loads
        // of raw plane s+3; barrier.  Steps z0-4 .. z0-1 are the prologue (gk,
        // b1 and g of the planes below the chunk, no output), step z0 sets the
        // carried z face of plane z0, steps z0+1 .. z1 output planes
        // z0 .. z1-1.  In the binary model (PRIO) a wave issues its loads at
        // raised priority (s_setprio 2, back to 0 after): the arbiter then lets it
        // get a plane's requests out ahead of the other waves' arithmetic, so they
        // are in flight sooner (256^3: -7 %).  The Potts and Ishikawa flow kernels
        // measured slower with it (+0.4 %, +10 %) and keep one priority.
        // S3's u and pt (Potts: ps) are the raw plane's own-position
        // slots, kept in registers from S1 three steps earlier (k0, k1, k2), so
        // the output stage loads only Cs and Ct.  Every load is issued right after the last use of its
        // registers, so it stays in flight across the barriers and the stages
        // between: the S3 operands behind S1, both barriers, S2a and S2b; the
        // raw plane behind S3.  The hardware counts loads and stores together
        // in issue order, and the compiler's wait counts are exact only when
        // every path issues the same memory operations after a load: so the
        // phases are separate loops whose steps all issue the same loads in the
        // same order, past the chunk's last raw plane the raw loads read plane
        // -3 (never used), and lanes outside the volume store past the field
        // instead of branching around their stores.
        // With fp32 flows the raw plane's loads go out before S2a and S2b, whose
        // operands from that plane's registers (stored in-plane faces, alpha) are
        // copied first, so the loads stay in flight behind both stages as well (256^3:
        // -2.1 %; with fp16 flows +0.6 %, so those issue after them).
        constexpr bool EARLY_RAW = std::is_same<T, float>::value;
        float tfx[EARLY_RAW ? K2 : 1], tfy[EARLY_RAW ? K2 : 1], tal[EARLY_RAW ? K1 : 1];
        auto tail = [&](int s, Raw &r) {
            if constexpr (EARLY_RAW) {
#pragma unroll
                for (int k = 0; k < K2; ++k) {
                    tfx[k] = r.bxl[k];
                    tfy[k] = r.byl[k];
                }
#pragma unroll
                for (int k = 0; k < K1; ++k) tal[k] = r.al[k];
                __syncthreads();
                __builtin_amdgcn_sched_barrier(0);
                if constexpr (PRIO) __builtin_amdgcn_s_setprio(2);
                load_raw(s + 3 <= z1 + 1 ? s + 3 : -3, r);   // S1 at the next step
                if constexpr (PRIO) __builtin_amdgcn_s_setprio(0);
                __builtin_amdgcn_sched_barrier(0);
                if (s + 2 <= z1 + 1) s2a(s + 2, tfx, tfy);
                if (s + 1 >= z0 - 1 && s + 1 <= z1) s2b(s + 1, tal);
            } else {
                __syncthreads();
                if (s + 2 <= z1 + 1) s2a(s + 2, r.bxl, r.byl);
                if (s + 1 >= z0 - 1 && s + 1 <= z1) s2b(s + 1, r.al);
                // the loads below reuse the registers S2a and S2b just read: keep the
                // scheduler from hoisting them (that would need two raw planes live)
                __builtin_amdgcn_sched_barrier(0);
                if constexpr (PRIO) __builtin_amdgcn_s_setprio(2);
                load_raw(s + 3 <= z1 + 1 ? s + 3 : -3, r);   // S1 at the next step
                if constexpr (PRIO) __builtin_amdgcn_s_setprio(0);
            }
            __syncthreads();
        };
        Raw r;
        Out o;
        // kept S3 operands of planes s-1, s, s+1 at the start of step s
        Keep k0, k1, k2;
        if constexpr (IS3D) {
            // bz own of plane z0-2 (raw plane z0-3 would carry it as its "next" bz)
            if constexpr (BLK && POTTS) {
                const __amdgpu_buffer_rsrc_t rs = blk_rsrc(a.in.pt, nz, a.pb);
                const unsigned P = (unsigned)(z0 - 2) * (unsigned)a.pb;
#pragma unroll
                for (int k = 0; k < K2; ++k) bz_cur[k] = ldb_f32(rs, P + bu2[k] + PL::kBZ);
            } else if constexpr (!FIRST && BLK) {
                const __amdgpu_buffer_rsrc_t rs = blk_rsrc(a.in.u, nz, a.pb);
                const unsigned P = (unsigned)(z0 - 2) * (unsigned)a.pb;
#pragma unroll
                for (int k = 0; k < K2; ++k) bz_cur[k] = ldb_t<T>(rs, P + bz2[k] + BL::kBZ);
            } else if constexpr (!FIRST) {
                const PRef rbz = pref(a.in.bz, z0 - 2, nz, IO::kBytes, plane);
#pragma unroll
                for (int k = 0; k < K2; ++k) bz_cur[k] = ld<IO>(rbz, off2[k]);
            }
            if constexpr (!ONEB) {
                load_raw(z0 - 2, r);
                for (int s = z0 - 4; s < z0; ++s) {   // prologue (s + 2 <= z0 + 1 <= z1 + 1)
                    s1(s + 2, r);
                    k0 = k1;
                    k1 = k2;
                    keep(k2, r);
                    tail(s, r);
                }
                s3_carry(z0 - 1);                     // step z0
                __builtin_amdgcn_sched_barrier(0);
                load_out(z0, o);                      // consumed by S3(z0) at the next step
                __builtin_amdgcn_sched_barrier(0);
                s1(z0 + 2, r);
                k0 = k1;
                k1 = k2;
                keep(k2, r);
                tail(z0, r);
                for (int s = z0 + 1; s <= z1; ++s) {
                    s3(s - 1, o, k0);
                    // S3 has read o: refill it; S1 and S2 are free of memory operations
                    __builtin_amdgcn_sched_barrier(0);
                    if constexpr (PRIO) __builtin_amdgcn_s_setprio(2);
                    load_out(s, o);
                    if constexpr (PRIO) __builtin_amdgcn_s_setprio(0);
                    __builtin_amdgcn_sched_barrier(0);
                    k0 = k1;
                    k1 = k2;
                    if (s + 2 <= z1 + 1) {
                        s1(s + 2, r);
                        keep(k2, r);
                    }
                    tail(s, r);
                }
            } else {
                // One barrier per plane step (the Potts and Ishikawa flow kernels,
                // measured 1.5 % and 3 % faster with it; the binary kernel 2.9 %
                // slower).  Step t: S2b(t), S3(t-1) and the loads of the S3 operands
                // of plane t, S2a(t+1), S1(t+2), then the loads of raw plane t+3; one
                // barrier.  Every stage reads in LDS only what earlier steps wrote:
                // S2a(t+1) the gk of S1(t+1), S2b(t) the faces of S2a(t), S1(t) and
                // S1(t+1), S3(t-1) the g of S2b(t-1) and the faces of S2a(t-1) and
                // S1(t); each writes a ring slot whose last reader ran in an earlier
                // step.  What a stage takes from the same step or from a raw plane the
                // registers no longer hold goes in registers: g of plane t at S3(t-1)'s
                // own positions (g_hi, computed by the same thread in S2b(t)), alpha of
                // plane t for S2b (alc) and the stored bx, by of plane t+1 for S2a
                // (fxl, fyl), taken when S1 read the raw plane carrying them.
                float alc[K1], fxl[K2], fyl[K2];
                auto keep_raw = [&](const Raw &rr) {
#pragma unroll
                    for (int k = 0; k < K1; ++k) alc[k] = rr.al[k];
#pragma unroll
                    for (int k = 0; k < K2; ++k) {
                        fxl[k] = rr.bxl[k];
                        fyl[k] = rr.byl[k];
                    }
                };
                auto s2b_step = [&](int t) {
#pragma unroll
                    for (int k = 0; k < KO; ++k) g_lo[k] = g_hi[k];
                    s2b(t, alc);
                };
                auto raw_next = [&](int t) {
                    __builtin_amdgcn_sched_barrier(0);
                    load_raw(t + 3 <= z1 + 1 ? t + 3 : -3, r);   // S1 at the next step
                    __syncthreads();
                };
                load_raw(z0 - 2, r);
                for (int t = z0 - 4; t < z0; ++t) {   // prologue
                    if (t >= z0 - 1) s2b_step(t);
                    if (t >= z0 - 2) s2a(t + 1, fxl, fyl);
                    s1(t + 2, r);
                    k0 = k1;
                    k1 = k2;
                    keep(k2, r);
                    keep_raw(r);
                    raw_next(t);
                }
                s2b_step(z0);                         // step z0
                s3_carry(z0 - 1);
                __builtin_amdgcn_sched_barrier(0);
                load_out(z0, o);                      // consumed by S3(z0) at the next step
                __builtin_amdgcn_sched_barrier(0);
                s2a(z0 + 1, fxl, fyl);
                s1(z0 + 2, r);
                k0 = k1;
                k1 = k2;
                keep(k2, r);
                keep_raw(r);
                raw_next(z0);
                for (int t = z0 + 1; t <= z1; ++t) {
                    s2b_step(t);
                    s3(t - 1, o, k0);
                    // S3 has read o: refill it; S2a and S1 are free of memory operations
                    __builtin_amdgcn_sched_barrier(0);
                    load_out(t, o);
                    __builtin_amdgcn_sched_barrier(0);
                    k0 = k1;
                    k1 = k2;
                    if (t + 1 <= z1) s2a(t + 1, fxl, fyl);
                    if (t + 2 <= z1 + 1) {
                        s1(t + 2, r);
                        keep(k2, r);
                    }
                    keep_raw(r);
                    raw_next(t);
                }
            }
        } else {
            // one plane: its stages in dependency order, no z pipeline; alpha of
            // plane 0 (load_raw fetches the previous plane's for the pipeline)
            load_raw(0, r);
            const PRef ra = pref(a.alpha, 0, nz, 4, plane);
#pragma unroll
            for (int k = 0; k < K1; ++k) r.al[k] = ldf(ra, off1[k]);
            load_out(0, o);
            s1(0, r);
            keep(k0, r);
            __syncthreads();
            s2a(0, r.bxl, r.byl);
            __syncthreads();
            s2b(0, r.al);
#pragma unroll
            for (int k = 0; k < KO; ++k) g_lo[k] = g_hi[k];   // S3(0) reads g of plane 0
            __syncthreads();
            s3(0, o, k0);
        }
    }
    if constexpr (FIRST) {
        if (abad && a.alpha_bad) {
            atomicOr(a.alpha_bad, 1);
            if (a.alpha_stop) atomicOr(a.alpha_stop, 1);
        }
    }
    return e_acc;
}

// The stop test of a previous iteration, as k_finalize (1024 threads) computes it,
// by a workgroup of blockDim.x threads (a divisor of 1024): each thread plays
// 1024 / blockDim.x of k_finalize's threads -- the same strided sums, the same
// wave sums of the same 64 values, the 16 wave sums added in the same order --
// so cvg has the same bits.  Every thread forms the final sum, so every thread
// returns the decision; thread 0 writes cvg and the stop outputs when `write`:
// flags->done, or, with `at` (every workgroup of the launch runs the test),
// flags->stop_at.  scratch: >= 16 doubles of LDS, free on entry and on return.
__device__ __forceinline__ bool stop_test(const DeferredStop &d, double *scratch, bool write,
                                          bool at = false) {
    constexpr int kFinalizeThreads = 1024;
    for (int vt = threadIdx.x; vt < kFinalizeThreads; vt += blockDim.x) {
        double v = strided_sum(d.partials, d.nparts, vt, kFinalizeThreads);
        v = wave_sum(v);
        if ((vt & (kWave - 1)) == 0) scratch[vt / kWave] = v;
    }
    __syncthreads();
    double r = 0.0;
    for (int w = 0; w < kFinalizeThreads / kWave; ++w) r += scratch[w];
    __syncthreads();
    const float c = (float)(r / d.denom);
    const bool stop = c <= d.err_bound;   // asetsBinaryMF3D_mex.c:183-188
    if (write && threadIdx.x == 0) {
        d.cvg[d.it] = c;
        if (stop) {
            d.flags->it_num = d.it;
            if (at)   // read by other workgroups of this launch
                __hip_atomic_store(&d.flags->stop_at, d.it + 1, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            else
                d.flags->done = 1;
            *d.it_num = d.it;
        }
    }
    return stop;
}

__device__ __forceinline__ void deferred_stop(const DeferredStop &d, double *scratch) {
    (void)stop_test(d, scratch, true);
}

// This workgroup's share of u out of the block-interleaved input state a.in into
// the planar a.take_u, four voxels per 16-byte access (ny a multiple of 4, take_u
// 16-byte aligned; the state's blocks are).
template <typename T>
__device__ void take_u_share(const FusedArgs<T> &a) {
    using BL = BlkLayout<T>;
    const char *st = (const char *)a.in.u;
    const int ny = a.ny;
    const int nq = a.plane / 4 * a.nz;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gridDim.x * blockDim.x) {
        const int g = 4 * q;
        const int c = g / ny, y = g - c * ny;
        const float4 v = *(const float4 *)(st + ((size_t)c * a.nyb + (y >> 6)) * BL::kBytes +
                                           BL::kU + 4 * (y & 63));
        *(float4 *)(a.take_u + g) = v;
    }
}

// XCD-aware block order: hardware dispatches block b to XCD b % 8; remap so
// that consecutive logical tiles share an XCD and its L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    return ((nb & 7) == 0) ? (b & 7) * (nb >> 3) + (b >> 3) : b;
}

// One volume: one workgroup per work item, partial |e| sums per workgroup.
// The first dispatched workgroup also runs the previous iteration's deferred
// stop test, when there is one (it starts first, so it has the most slack).
// BLK: block-interleaved state (BlkLayout), else planar fields.
template <bool IS3D, int TX, int NT, int WAVES, typename T, bool FIRST = false, bool BLK = true>
__global__ __launch_bounds__(NT, WAVES) void kf_binary(FusedArgs<T> a) {
    __shared__ double red[16];
    if (a.flags->done) return;
    // stop_all: every workgroup runs the previous iteration's stop test first (workgroup
    // 0 writes its outputs).  When it fires, the workgroup writes its share of u
    // (take_u) from this launch's input, the final state, and returns, so this launch
    // costs no more than that copy.  stop_at (that iteration + 1) is written once;
    // a workgroup that reads it skips the test: it returns at once in a later launch
    // and takes its share in this one.  A stale 0 only makes it run the test.  One
    // thread reads stop_at for the whole workgroup (its waves must agree).
    if (a.stop_all && a.stop.partials) {
        __shared__ int32_t at_s;
        if (threadIdx.x == 0)
            at_s = __hip_atomic_load(&a.flags->stop_at, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const int32_t at = at_s;
        if (at != 0 && at - 1 != a.stop.it) return;
        if (at != 0 || stop_test(a.stop, red, blockIdx.x == 0, true)) {
            if constexpr (BLK) {
                if (a.take_u) take_u_share<T>(a);
            }
            return;
        }
    }
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const double e =
        (lb < a.nblocks) ? fused_tile<IS3D, TX, NT, T, 0, FIRST, IS3D, BLK>(a, lb) : 0.0;
    const double s = block_sum(e, red);
    if (threadIdx.x == 0) a.partials[blockIdx.x] = s;
    if (blockIdx.x == 0 && a.stop.partials && !a.stop_all) {
        __syncthreads();   // thread 0 has read red
        deferred_stop(a.stop, red);
    }
}

// ---- Potts model -----------------------------------------------------------

struct PottsArgs {
    const float *Ct, *alpha;   // [L*N]
    const float *u;            // [L*N] (read)
    const float *ps;           // [N]   label-shared source flow (read)
    float *pt_in, *pt_out;     // [L*N]
    float *b_in[3], *b_out[3]; // [L*N] faces x, y, z
    float *ft;                 // [L*N] out: dv + pt per label
    int ny, nx, nz, nlab;
    float cc, steps;
    int zchunk, ntile_y, ntile_x, nzc;
    int nblocks;               // per label
    int fast_cc;
    const Flags *flags;
    DeferredStop stop;         // the previous iteration's stop test (partials null: none)
    // block-interleaved flows (PottsBlkLayout; pt_in / pt_out the buffers' bases)
    int nyb;                   // 64-row blocks per column
    int pb;                    // bytes per plane of one label's flows
};

// All labels' flow updates of one iteration in one launch: workgroup b handles
// tile b % nblocks of label b / nblocks (labels are independent inside an
// iteration, asetsPotts3D_mex.c:172-191).  Label-major, so that each XCD holds
// whole slabs of one label and neighbouring tiles share their halo in L2 (label
// fastest, sharing the label-independent ps instead, measured 3 % slower).
// The first dispatched workgroup also runs the previous iteration's deferred stop
// test, when there is one: this launch writes only the other pt and face buffers
// and ft, so a stop found here leaves that iteration's state whole, and the
// coupling kernel after it (the only writer of u and ps) returns at once.
// BLK: block-interleaved label flows (PottsBlkLayout), else planar fields.
template <bool IS3D, int TX, int NT, int WAVES, bool BLK = true>
__global__ __launch_bounds__(NT, WAVES) void kf_potts_flow(PottsArgs p) {
    __shared__ double red[16];
    if (p.flags->done) return;
    const int total = p.nblocks * p.nlab;
    const int gb = xcd_remap(blockIdx.x, gridDim.x);
    if (gb >= total) return;
    const int l = gb / p.nblocks, lb = gb % p.nblocks;
    const int64_t off = (int64_t)l * p.ny * p.nx * p.nz;
    FusedArgs<float> a;
    a.Cs = nullptr;
    a.Ct = p.Ct + off;
    a.alpha = p.alpha + off;
    a.in = FusedState<float>{(float *)p.u + off, (float *)p.ps, p.pt_in + off, p.b_in[0] + off,
                             p.b_in[1] + off, p.b_in[2] + off};
    a.out = FusedState<float>{nullptr, p.ft + off, p.pt_out + off, p.b_out[0] + off,
                              p.b_out[1] + off, p.b_out[2] + off};
    if constexpr (BLK) {   // label l's flows start l * nz * pb bytes into each buffer
        const size_t lb_off = (size_t)l * p.nz * p.pb;
        a.in.pt = (float *)((char *)p.pt_in + lb_off);
        a.out.pt = (float *)((char *)p.pt_out + lb_off);
        a.in.bx = a.in.by = a.in.bz = a.in.pt;
        a.out.bx = a.out.by = a.out.bz = a.out.pt;
    }
    a.nyb = p.nyb;
    a.pb = p.pb;
    a.ny = p.ny;
    a.nx = p.nx;
    a.nz = p.nz;
    a.plane = p.ny * p.nx;
    a.cc = p.cc;
    a.steps = p.steps;
    a.zchunk = p.zchunk;
    a.ntile_y = p.ntile_y;
    a.ntile_x = p.ntile_x;
    a.nzc = p.nzc;
    a.nblocks = p.nblocks;
    a.fast_cc = p.fast_cc;
    a.partials = nullptr;
    a.flags = nullptr;
    a.alpha_bad = a.alpha_stop = nullptr;
    (void)fused_tile<IS3D, TX, NT, float, 1, false, IS3D, BLK>(a, lb);
    if (blockIdx.x == 0 && p.stop.partials) deferred_stop(p.stop, red);
}

// Source flow and multipliers, labels coupled per voxel in label order
// (asetsPotts3D_mex.c:495-514): s += ft_l - u_l/cc; ps = s/L + 1/(cc*L);
// e_l = cc*(ft_l - ps); u_l -= e_l; sum |e_l| in FP64 per workgroup.
// LT > 0: the label count is the compile-time LT and ft stays in registers;
// LT == 0: any label count, ft is read again in the second pass.
template <int LT>
__global__ __launch_bounds__(256) void kp_potts_couple(const float *ft, float *u, float *ps,
                                                       int64_t N, int nlab_rt, float cc,
                                                       int fast_cc, double *partials,
                                                       const Flags *flags) {
    __shared__ double red[16];
    if (flags->done) return;
    const int nlab = LT > 0 ? LT : nlab_rt;
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = g < N;
    const int64_t gg = in ? g : 0;
    const float rcc = rcp_refined(cc);
    const float fl = (float)nlab;
    const float tail = 1.0f / (cc * fl);
    constexpr int R = LT > 0 ? LT : 1;
    float ftr[R], ur[R];
    float s = 0.0f;
#pragma unroll
    for (int l = 0; l < (LT > 0 ? LT : 1); ++l) {
        if (LT > 0) {
            ftr[l] = ft[gg + l * N];
            ur[l] = u[gg + l * N];
            s += (ftr[l] - div_fast(ur[l], cc, rcc, fast_cc != 0));
        }
    }
    if (LT == 0)
        for (int l = 0; l < nlab; ++l)
            s += (ft[gg + l * N] - div_fast(u[gg + l * N], cc, rcc, fast_cc != 0));
    const float nps = s / fl + tail;
    double e_abs = 0.0;
    if (LT > 0) {
#pragma unroll
        for (int l = 0; l < (LT > 0 ? LT : 1); ++l) {
            const float e = cc * (ftr[l] - nps);
            if (in) {
                u[gg + l * N] = ur[l] - e;
                e_abs += fabs((double)e);
            }
        }
    } else {
        for (int l = 0; l < nlab; ++l) {
            const float e = cc * (ft[gg + l * N] - nps);
            if (in) {
                u[gg + l * N] = u[gg + l * N] - e;
                e_abs += fabs((double)e);
            }
        }
    }
    if (in) ps[gg] = nps;
    const double sum = block_sum(e_abs, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = sum;
}

// ---- Ishikawa model (linearly ordered labels) --------------------------------

struct IshiArgs {
    const float *alpha;        // [M*N], M = L-1
    const float *u;            // [M*N] (read)
    const float *pt;           // [L*N] (read)
    float *b_in[3], *b_out[3]; // [M*N] faces x, y, z
    float *dv;                 // [M*N] out
    int ny, nx, nz, nflow;     // nflow = M
    float cc, steps;
    int zchunk, ntile_y, ntile_x, nzc;
    int nblocks;               // per flow field
    int fast_cc;
    const Flags *flags;
    DeferredStop stop;         // the previous iteration's stop test (partials null: none)
};

// All M flow updates of one iteration in one launch (the label loop of
// asetsIshikawa3D_mex.c:162-187 has no dependence between labels).  The first
// dispatched workgroup runs the previous iteration's deferred stop test as in
// kf_potts_flow: this launch writes only dv and the other face buffers, and the
// coupling kernel (the only writer of pt and u) returns at once after a stop.
template <bool IS3D, int TX, int NT, int WAVES>
__global__ __launch_bounds__(NT, WAVES) void kf_ishikawa_flow(IshiArgs p) {
    __shared__ double red[16];
    if (p.flags->done) return;
    const int total = p.nblocks * p.nflow;
    const int gb = xcd_remap(blockIdx.x, gridDim.x);
    if (gb >= total) return;
    const int l = gb / p.nblocks, lb = gb % p.nblocks;
    const int64_t N = (int64_t)p.ny * p.nx * p.nz;
    const int64_t off = (int64_t)l * N;
    FusedArgs<float> a;
    a.Cs = nullptr;
    a.Ct = nullptr;
    a.alpha = p.alpha + off;
    a.in = FusedState<float>{(float *)p.u + off, (float *)p.pt + off, (float *)p.pt + off + N,
                             p.b_in[0] + off, p.b_in[1] + off, p.b_in[2] + off};
    a.out = FusedState<float>{nullptr, p.dv + off, nullptr, p.b_out[0] + off, p.b_out[1] + off,
                              p.b_out[2] + off};
    a.ny = p.ny;
    a.nx = p.nx;
    a.nz = p.nz;
    a.plane = p.ny * p.nx;
    a.cc = p.cc;
    a.steps = p.steps;
    a.zchunk = p.zchunk;
    a.ntile_y = p.ntile_y;
    a.ntile_x = p.ntile_x;
    a.nzc = p.nzc;
    a.nblocks = p.nblocks;
    a.fast_cc = p.fast_cc;
    a.partials = nullptr;
    a.flags = nullptr;
    a.alpha_bad = a.alpha_stop = nullptr;
    (void)fused_tile<IS3D, TX, NT, float, 2>(a, lb);
    if (blockIdx.x == 0 && p.stop.partials) deferred_stop(p.stop, red);
}

// Sink flows (asetsIshikawa3D_mex.c:498-576: bottom, middle and top layer, in
// label order, each middle layer reading the already updated layer below) and
// multipliers (updateU, :578-612) per voxel; sum |e| in FP64 per workgroup.
// LT > 0: compile-time label count (values stay in registers); LT == 0: any.
template <int LT>
__global__ __launch_bounds__(256) void k_ishikawa_couple(const float *dv, float *pt, float *u,
                                                         const float *Ct, int64_t N,
                                                         int nlab_rt, float cc, int fast_cc,
                                                         double *partials,
                                                         const Flags *flags) {
    __shared__ double red[16];
    if (flags->done) return;
    const int L = LT > 0 ? LT : nlab_rt;
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = g < N;
    const int64_t gg = in ? g : 0;
    const float rcc = rcp_refined(cc);
    const bool fcc = fast_cc != 0;
    const float one_cc = 1.0f / cc;
    double e_abs = 0.0;
    if constexpr (LT > 0) {
        float d[LT - 1], uu[LT - 1], uc[LT - 1], q[LT];
#pragma unroll
        for (int l = 0; l < LT - 1; ++l) {
            d[l] = dv[gg + l * N];
            uu[l] = u[gg + l * N];
            uc[l] = div_fast(uu[l], cc, rcc, fcc);
        }
#pragma unroll
        for (int l = 0; l < LT; ++l) q[l] = pt[gg + l * N];
#pragma unroll
        for (int l = 0; l < LT; ++l) {
            float f;
            if (l == 0) {
                f = ((d[0] + q[1]) - uc[0]) + one_cc;
            } else if (l < LT - 1) {
                f = (-d[l - 1] + q[l - 1]) + uc[l - 1];
                f += (d[l] + q[l + 1]) - uc[l];
                f /= 2.0f;
            } else {
                f = (-d[l - 1] + q[l - 1]) + uc[l - 1];
            }
            q[l] = fmin_ref(f, Ct[gg + l * N]);
        }
#pragma unroll
        for (int l = 0; l < LT - 1; ++l) {
            const float e = cc * ((d[l] + q[l + 1]) - q[l]);
            if (in) {
                u[gg + l * N] = uu[l] - e;
                e_abs += fabs((double)e);
            }
        }
        if (in)
#pragma unroll
            for (int l = 0; l < LT; ++l) pt[gg + l * N] = q[l];
    } else {
        // chain with the neighbours in registers: layer l needs the new pt_{l-1}
        // and the old pt_{l+1}; pt is rewritten in place behind the chain
        float qprev = 0.0f, dprev = 0.0f, ucprev = 0.0f;
        float qcur = pt[gg];
        for (int l = 0; l < L; ++l) {
            const float qnext = (l + 1 < L) ? pt[gg + (l + 1) * N] : 0.0f;
            float dl = 0.0f, ucl = 0.0f;
            if (l < L - 1) {
                dl = dv[gg + l * N];
                ucl = div_fast(u[gg + l * N], cc, rcc, fcc);
            }
            float f;
            if (l == 0) {
                f = ((dl + qnext) - ucl) + one_cc;
            } else if (l < L - 1) {
                f = (-dprev + qprev) + ucprev;
                f += (dl + qnext) - ucl;
                f /= 2.0f;
            } else {
                f = (-dprev + qprev) + ucprev;
            }
            const float qn = fmin_ref(f, Ct[gg + l * N]);
            if (in) pt[gg + l * N] = qn;
            qprev = qn;
            dprev = dl;
            ucprev = ucl;
            qcur = qnext;
        }
        (void)qcur;
        for (int l = 0; l < L - 1; ++l) {
            const float e = cc * ((dv[gg + l * N] + pt[gg + (l + 1) * N]) - pt[gg + l * N]);
            if (in) {
                u[gg + l * N] = u[gg + l * N] - e;
                e_abs += fabs((double)e);
            }
        }
    }
    const double sum = block_sum(e_abs, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = sum;
}

// asetsIshikawa3D_mex.c:230-270: first strict arg-min of Ct from 1e30; pt_l =
// min for l < L-1, the top layer 0; u_l = [l < arg-min]
__global__ void k_init_ishikawa(const float *Ct, float *u, float *pt, int64_t N, int L) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N) return;
    float mv = 1e30f;
    int mi = 1000000000;
    for (int l = 0; l < L; ++l) {
        const float c = Ct[g + l * N];
        if (mv > c) { mv = c; mi = l; }
    }
    for (int l = 0; l < L - 1; ++l) {
        pt[g + l * N] = mv;
        u[g + l * N] = (l >= mi) ? 0.0f : 1.0f;
    }
    pt[g + (int64_t)(L - 1) * N] = 0.0f;
}

}  // namespace asets
