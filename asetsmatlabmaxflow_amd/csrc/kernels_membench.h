// kernels_membench.h -- HBM ceiling kernels: what this chip gives a plain stream,
// the yardstick for the solver kernels' roofline fractions.  No reference
// counterpart.
//
// Every kernel is a grid-stride loop over 16-byte (dwordx4) elements with UNROLL
// independent elements in flight per lane, optionally with the non-temporal
// (streaming) cache policy.  Modes:
//   kStreamRead  read n16 elements, fold them into one value per lane (written
//                once, so the loads are not dead)
//   kStreamCopy  dst[i] = src[i] for n16 elements (1:1 read:write)
//   kStreamMix   the binary solver's per-voxel traffic over N voxels (N a multiple
//                of 4): read Cs, Ct, alpha, u, bz (4 B per voxel) and the (ps, pt),
//                (bx, by) pairs (8 B), write u, bz and the two pairs: 36 B read and
//                24 B written per voxel, as the fused iteration moves them, without
//                its stencil.  One element = 4 voxels.
//   kStreamMixPacked  the same 36 B read and 24 B written per voxel through three
//                arrays instead of eleven: the inputs (12 B per voxel), the state read
//                (24 B) and the state written (24 B), each interleaved by blocks --
//                for every 64 elements (one wave), field k is a contiguous 1 KiB
//                block -- so each load and store is one coalesced 1 KiB wave access.
//   kStreamMixState  the same bytes with the inputs in their own three arrays (Cs,
//                Ct, alpha as the caller holds them) and only the state block-
//                interleaved (state read, state written): five arrays.
//   kStreamMixBlk  the same five arrays with the fused binary kernel's own layout and
//                access widths: one voxel per lane, Cs, Ct, alpha as 4-byte loads, the
//                state in 1536-byte blocks of 64 voxels (u at 4 B, the (ps, pt) and
//                (bx, by) pairs at 8 B, bz at 4 B per lane), read from one buffer and
//                written to another.
#pragma once
#include "common.h"

namespace asets {

enum StreamMode {
    kStreamRead = 0,
    kStreamCopy = 1,
    kStreamMix = 2,
    kStreamMixPacked = 3,
    kStreamMixState = 4,
    kStreamMixBlk = 5
};

// 16-byte element (a clang vector: the non-temporal builtins take those)
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct MixFields {
    const f32x4 *in4[5];    // Cs, Ct, alpha, u, bz       [N/4]
    const f32x4 *in8[2];    // (ps, pt), (bx, by) pairs    [N/2]
    f32x4 *out4[2];         // u, bz                       [N/4]
    f32x4 *out8[2];         // (ps, pt), (bx, by)          [N/2]
};

template <bool NT>
__device__ __forceinline__ f32x4 ld16(const f32x4 *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT>
__device__ __forceinline__ void st16(f32x4 *p, f32x4 v) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

template <int UNROLL, bool NT>
__global__ __launch_bounds__(256) void k_stream_read(const f32x4 *src, int64_t n16,
                                                     float *sink) {
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float acc = 0.0f;
    for (; i + (UNROLL - 1) * st < n16; i += UNROLL * st) {
        f32x4 v[UNROLL];
#pragma unroll
        for (int k = 0; k < UNROLL; ++k) v[k] = ld16<NT>(src + i + k * st);
#pragma unroll
        for (int k = 0; k < UNROLL; ++k) acc += (v[k].x + v[k].y) + (v[k].z + v[k].w);
    }
    for (; i < n16; i += st) {
        const f32x4 v = ld16<NT>(src + i);
        acc += (v.x + v.y) + (v.z + v.w);
    }
    sink[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

template <int UNROLL, bool NT>
__global__ __launch_bounds__(256) void k_stream_copy(const f32x4 *src, f32x4 *dst,
                                                     int64_t n16) {
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + (UNROLL - 1) * st < n16; i += UNROLL * st) {
        f32x4 v[UNROLL];
#pragma unroll
        for (int k = 0; k < UNROLL; ++k) v[k] = ld16<NT>(src + i + k * st);
#pragma unroll
        for (int k = 0; k < UNROLL; ++k) st16<NT>(dst + i + k * st, v[k]);
    }
    for (; i < n16; i += st) st16<NT>(dst + i, ld16<NT>(src + i));
}

// n16 = N/4 elements of 4 voxels; the outputs are cheap functions of all nine
// inputs, so every load is live.
template <bool NT>
__global__ __launch_bounds__(256) void k_stream_mix(MixFields f, int64_t n16) {
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += st) {
        f32x4 a[5], p[2][2];
#pragma unroll
        for (int k = 0; k < 5; ++k) a[k] = ld16<NT>(f.in4[k] + i);
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            p[k][0] = ld16<NT>(f.in8[k] + 2 * i);
            p[k][1] = ld16<NT>(f.in8[k] + 2 * i + 1);
        }
        f32x4 s;
        s.x = (a[0].x + a[1].x) + a[2].x;
        s.y = (a[0].y + a[1].y) + a[2].y;
        s.z = (a[0].z + a[1].z) + a[2].z;
        s.w = (a[0].w + a[1].w) + a[2].w;
        f32x4 o0 = a[3], o1 = a[4];
        o0.x += s.x; o0.y += s.y; o0.z += s.z; o0.w += s.w;
        o1.x -= s.x; o1.y -= s.y; o1.z -= s.z; o1.w -= s.w;
        st16<NT>(f.out4[0] + i, o0);
        st16<NT>(f.out4[1] + i, o1);
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            f32x4 q0 = p[k][0], q1 = p[k][1];
            q0.x += s.x; q0.y += s.y; q0.z += s.z; q0.w += s.w;
            q1.x += s.w; q1.y += s.z; q1.z += s.y; q1.w += s.x;
            st16<NT>(f.out8[k] + 2 * i, q0);
            st16<NT>(f.out8[k] + 2 * i + 1, q1);
        }
    }
}

// n16 = N/4 elements of 4 voxels (a multiple of 64): 3 input, 6 state-in and 6
// state-out f32x4 each.
template <bool NT>
__global__ __launch_bounds__(256) void k_stream_mix_packed(const f32x4 *in, const f32x4 *sin,
                                                           f32x4 *sout, int64_t n16) {
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += st) {
        const int64_t blk = i >> 6, lane = i & 63;
        f32x4 a[3], q[6];
#pragma unroll
        for (int k = 0; k < 3; ++k) a[k] = ld16<NT>(in + (blk * 3 + k) * 64 + lane);
#pragma unroll
        for (int k = 0; k < 6; ++k) q[k] = ld16<NT>(sin + (blk * 6 + k) * 64 + lane);
        const f32x4 s = (a[0] + a[1]) + a[2];
#pragma unroll
        for (int k = 0; k < 6; ++k) st16<NT>(sout + (blk * 6 + k) * 64 + lane, q[k] + s);
    }
}

// kStreamMixState: three planar input arrays (n16 elements each) and the state
// block-interleaved as in k_stream_mix_packed.
template <bool NT>
__global__ __launch_bounds__(256) void k_stream_mix_state(const f32x4 *cs, const f32x4 *ct,
                                                          const f32x4 *al, const f32x4 *sin,
                                                          f32x4 *sout, int64_t n16) {
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += st) {
        const int64_t blk = i >> 6, lane = i & 63;
        f32x4 q[6];
        const f32x4 a0 = ld16<NT>(cs + i), a1 = ld16<NT>(ct + i), a2 = ld16<NT>(al + i);
#pragma unroll
        for (int k = 0; k < 6; ++k) q[k] = ld16<NT>(sin + (blk * 6 + k) * 64 + lane);
        const f32x4 s = (a0 + a1) + a2;
#pragma unroll
        for (int k = 0; k < 6; ++k) st16<NT>(sout + (blk * 6 + k) * 64 + lane, q[k] + s);
    }
}

typedef float f32x2 __attribute__((ext_vector_type(2)));

template <bool NT, typename V>
__device__ __forceinline__ V ldn(const V *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT, typename V>
__device__ __forceinline__ void stn(V *p, V v) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// kStreamMixBlk: N voxels (a multiple of 64), one per lane per step.
template <bool NT>
__global__ __launch_bounds__(256) void k_stream_mix_blk(const float *cs, const float *ct,
                                                        const float *al, const char *sin,
                                                        char *sout, int64_t N) {
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += st) {
        const int64_t base = (i >> 6) * 1536;
        const int ln = (int)(i & 63);
        const float a = ldn<NT>(cs + i), b = ldn<NT>(ct + i), c = ldn<NT>(al + i);
        const float u = ldn<NT>((const float *)(sin + base) + ln);
        const f32x2 p1 = ldn<NT>((const f32x2 *)(sin + base + 256) + ln);
        const f32x2 p2 = ldn<NT>((const f32x2 *)(sin + base + 768) + ln);
        const float bz = ldn<NT>((const float *)(sin + base + 1280) + ln);
        const float q = (a + b) + c;
        stn<NT>((float *)(sout + base) + ln, u + q);
        stn<NT>((f32x2 *)(sout + base + 256) + ln, p1 + q);
        stn<NT>((f32x2 *)(sout + base + 768) + ln, p2 + q);
        stn<NT>((float *)(sout + base + 1280) + ln, bz + q);
    }
}

}  // namespace asets
