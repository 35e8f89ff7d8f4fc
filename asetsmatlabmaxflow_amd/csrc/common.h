// common.h -- shared device helpers for the ASETS gfx950 kernels.
//
// Arithmetic contract (every kernel): no contraction of a*b+c into FMA (the
// library is compiled with -ffp-contract=off), correctly rounded fp32 division
// (HIP's default lowering of '/'), fp64 square root.  Together with keeping the
// reference's evaluation order this makes u bit-identical to the reference
// C-MEX solvers at a fixed iteration count.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace asets {

constexpr int kWave = 64;

// Flag words shared by all kernels of one solve (device memory).
//   flags[0]  : 1 once some iteration met cvg <= err_bound; later kernels no-op
struct Flags {
    int32_t done;       // stop: converged, or alpha invalid (every later iteration returns)
    int32_t bad_alpha;  // alpha <= 0 somewhere (set before the first iteration)
    int32_t it_num;     // iteration at which the stop test fired (valid when done or stop_at)
    int32_t stop_at;    // fused binary, stop test in every workgroup: that iteration + 1,
                        // written once by the launch that finds the stop (0: none)
    int32_t pad[60];
};

__device__ __forceinline__ float fmin_ref(float a, float b) {
    // MIN(a,b) of asetsBinaryMF3D_mex.c:17 -- returns b when a is NaN.
    return (a < b) ? a : b;
}

// |p| <= alpha projection factor of asetsBinaryMF3D_mex.c:321-331: squares and
// sum in double with the high face first, sqrt in double, narrowed to float,
// then the two float divisions n/alpha and 1/n.
template <bool IS3D>
__device__ __forceinline__ float proj_factor(float xr, float xl, float yr, float yl,
                                             float zr, float zl, float a) {
    double t = (double)xr * (double)xr + (double)xl * (double)xl;
    t = t + (double)yr * (double)yr;
    t = t + (double)yl * (double)yl;
    if (IS3D) {
        t = t + (double)zr * (double)zr;
        t = t + (double)zl * (double)zl;
    }
    float n = (float)__builtin_sqrt(t * 0.5);
    n = (n > a) ? n / a : 1.0f;
    return 1.0f / n;
}

// Rescale of one interior face by the mean factor of its two voxels
// (asetsBinaryMF3D_mex.c:351-352): the float sum, then *0.5 and *b in double,
// i.e. one rounding of the exact product.
__device__ __forceinline__ float rescale_face(float g_hi, float g_lo, float b) {
    float s = g_hi + g_lo;
    return (float)((double)s * 0.5 * (double)b);
}

// Sum of a double over the 64 lanes of a wave (butterfly; every lane ends with
// the total).  Fixed order, so the result is deterministic.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// ((0 + p[i]) + p[i + st]) + ... over i < n, added in index order; the loads
// are issued eight at a time ahead of their adds, so a long strided sum waits
// on memory once per eight elements instead of once per element.
__device__ __forceinline__ double strided_sum(const double *p, int64_t n, int64_t i,
                                              int64_t st) {
    constexpr int B = 8;
    double v = 0.0;
    for (; i + (B - 1) * st < n; i += B * st) {
        double t[B];
#pragma unroll
        for (int b = 0; b < B; ++b) t[b] = p[i + b * st];
#pragma unroll
        for (int b = 0; b < B; ++b) v += t[b];
    }
    for (; i < n; i += st) v += p[i];
    return v;
}

// Block-wide sum (blockDim.x a multiple of 64, <= 1024); result valid in
// thread 0.  scratch: >= 16 doubles of LDS.
__device__ __forceinline__ double block_sum(double v, double *scratch) {
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    if (lane == 0) scratch[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        const int nw = blockDim.x / kWave;
        for (int i = 0; i < nw; ++i) r += scratch[i];
    }
    return r;
}

}  // namespace asets
