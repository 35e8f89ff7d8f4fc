// kernels_prepost.h -- the steps either side of the solve, on the GPU: data
// costs and regularisation weights before it, discrete labels after it.
//
//   loglik costs     applications/interactiveGraphCutSegmentation3D.m:96-118
//                    (computeLogLikelihoodCost): image normalised to [0, 255] in
//                    fp32, 256-bin histogram of the scribbled voxels with histc's
//                    edges 0..255 (bin floor(x), x == 255 in the last bin), the
//                    lookup at uint16(x) (round half away from zero),
//                    cost = -log10(P + eps) in FP64, stored as fp32
//   -log maps        applications/atlasProbabilityMapsRegularization3D.m:47-53:
//                    Ct = -log(p + eps)
//   contrast alpha   applications/smoothManualSegmentations2D.m:73-82: image
//                    normalised to [0, 1], MATLAB gradient() (central differences
//                    inside, one-sided at the borders), alpha = a0 + a1*exp(-beta*m)
//                    with m = |grad I| (power 1) or |grad I|^2 (power 2)
//   L1 costs         applications/L1IntensitySegmentation2D.m:36-44: image normalised
//                    to [0, 1] in fp32, Ct_l = |x - m_l|
//   smoothed labels  applications/smoothManualSegmentations2D.m:48-55: Ct_l =
//                    imfilter(1 - (man == id_l), h, 'replicate') for a caller's K x K
//                    kernel h (fspecial('gaussian', [5 5], 1.5) there); FP64 sums of the
//                    weights in h's column-major order, stored as fp32
//   labels           tests/runTestBinaryMF.m:55 (u > 0.5), tests/runTestPotts.m:58
//                    (1-based index of the first maximum over labels),
//                    tests/runTestIshikawa.m:57-60 (number of cut functions > 0.5)
#pragma once
#include "common.h"

namespace asets {

constexpr int kPrepBlocks = 1024;

// per-block min and max (fp32 values compare exactly; order-independent)
__global__ void k_minmax_part(const float *x, int64_t n, float *part) {
    __shared__ float smin[256], smax[256];
    float mn = INFINITY, mx = -INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float v = x[i];
        mn = fminf(mn, v);
        mx = fmaxf(mx, v);
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = smin[0];
        part[2 * blockIdx.x + 1] = smax[0];
    }
}

__global__ void k_minmax_final(const float *part, int nb, float *out) {
    if (threadIdx.x != 0) return;
    float mn = INFINITY, mx = -INFINITY;
    for (int b = 0; b < nb; ++b) {
        mn = fminf(mn, part[2 * b]);
        mx = fmaxf(mx, part[2 * b + 1]);
    }
    out[0] = mn;
    out[1] = mx;
}

// x in [0, 255]: img normalised as ((img - mn) / (mx - mn)) * 255 in fp32
__device__ __forceinline__ float norm255(float v, float mn, float range) {
    return ((v - mn) / range) * 255.0f;
}

// histograms [nlab][256] of the voxels whose scribble is l+1 (histc bins)
__global__ void k_loglik_hist(const float *img, const int32_t *scrib, int64_t n, int nlab,
                              const float *mm, uint32_t *hist) {
    extern __shared__ uint32_t sh[];   // nlab * 256
    for (int i = threadIdx.x; i < nlab * 256; i += blockDim.x) sh[i] = 0;
    __syncthreads();
    const float mn = mm[0], range = mm[1] - mm[0];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t s = scrib[i];
        if (s >= 1 && s <= nlab) {
            const float x = norm255(img[i], mn, range);
            const int bin = min((int)x, 255);
            atomicAdd(&sh[(s - 1) * 256 + bin], 1u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nlab * 256; i += blockDim.x)
        if (sh[i]) atomicAdd(&hist[i], sh[i]);
}

// number of scribbled voxels of each label (sum of its bins)
__global__ void k_loglik_totals(const uint32_t *hist, int nlab, uint64_t *tot) {
    const int l = threadIdx.x;
    if (l >= nlab) return;
    uint64_t t = 0;
    for (int k = 0; k < 256; ++k) t += hist[l * 256 + k];
    tot[l] = t;
}

// Ct[l*n + i] = (float)(-log10(count_l[round(x_i)] / total_l + eps))
__global__ void k_loglik_cost(const float *img, int64_t n, int nlab, const float *mm,
                              const uint32_t *hist, const uint64_t *tot, double eps,
                              float *Ct) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float mn = mm[0], range = mm[1] - mm[0];
    const float x = norm255(img[i], mn, range);
    int b = (int)x;                      // x >= 0
    if (x - (float)b >= 0.5f) ++b;       // uint16(): round half away from zero
    for (int l = 0; l < nlab; ++l) {
        const double P = (double)hist[l * 256 + b] / (double)tot[l];
        Ct[l * n + i] = (float)(-log10(P + eps));
    }
}

__global__ void k_neglog(const float *p, int64_t n, double eps, float *Ct) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) Ct[i] = (float)(-log((double)p[i] + eps));
}

// MATLAB gradient() along an axis of extent m at index k, stride s (FP64)
__device__ __forceinline__ double grad1(const float *I, int64_t g, int64_t k, int64_t m,
                                        int64_t s, double mn, double inv) {
    if (m < 2) return 0.0;
    auto at = [&](int64_t q) { return ((double)I[q] - mn) * inv; };
    if (k == 0) return at(g + s) - at(g);
    if (k == m - 1) return at(g) - at(g - s);
    return (at(g + s) - at(g - s)) * 0.5;
}

__global__ void k_alpha_contrast(const float *I, int64_t ny, int64_t nx, int64_t nz, int is3d,
                                 const float *mm, int normalize, double a0, double a1,
                                 double beta, int power, float *alpha) {
    const int64_t N = ny * nx * nz;
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N) return;
    const int64_t y = g % ny, x = (g / ny) % nx, z = g / (ny * nx);
    double mn = 0.0, inv = 1.0;
    if (normalize) {
        mn = (double)mm[0];
        inv = 1.0 / ((double)mm[1] - (double)mm[0]);
    }
    const double gx = grad1(I, g, x, nx, ny, mn, inv);
    const double gy = grad1(I, g, y, ny, 1, mn, inv);
    const double gz = is3d ? grad1(I, g, z, nz, ny * nx, mn, inv) : 0.0;
    const double q = gx * gx + gy * gy + gz * gz;
    const double m = (power == 1) ? sqrt(q) : q;
    alpha[g] = (float)(a0 + a1 * exp(-beta * m));
}

constexpr int kMaxCostLabels = 64;
struct CostModels {
    float m[kMaxCostLabels];
};
struct CostIds {
    int32_t id[kMaxCostLabels];
};

// Ct[l*n + i] = |x_i - m_l|, x = (img - min) / (max - min) in fp32
__global__ void k_l1_costs(const float *img, int64_t n, const float *mm, CostModels ms, int nlab,
                           float *Ct) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = (img[i] - mm[0]) / (mm[1] - mm[0]);
    for (int l = 0; l < nlab; ++l) Ct[l * n + i] = fabsf(x - ms.m[l]);
}

// Ct[l*n + g] = sum over the K x K window around g (borders replicated) of
// h(a, b) * (lab != id_l), h column-major (a along dim 1), summed in that order in FP64
__global__ void k_smooth_labels(const int32_t *lab, int64_t ny, int64_t nx, CostIds ids,
                                int nlab, const double *h, int K, float *Ct) {
    const int64_t n = ny * nx;
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int64_t y = g % ny, x = g / ny;
    const int r = K / 2;
    for (int l = 0; l < nlab; ++l) {
        double sacc = 0.0;
        for (int b = 0; b < K; ++b) {
            const int64_t xx = min(max(x + b - r, (int64_t)0), nx - 1);
            for (int a = 0; a < K; ++a) {
                const int64_t yy = min(max(y + a - r, (int64_t)0), ny - 1);
                if (lab[xx * ny + yy] != ids.id[l]) sacc += h[b * K + a];
            }
        }
        Ct[l * n + g] = (float)sacc;
    }
}

// model 0: u > 0.5; 1: 1-based first arg-max over nlab fields; 2: number of the
// nlab-1 cut functions above 0.5
__global__ void k_labels(int model, const float *u, int64_t n, int nlab, uint8_t *lab) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int r = 0;
    if (model == 0) {
        r = u[i] > 0.5f ? 1 : 0;
    } else if (model == 1) {
        float best = u[i];
        r = 1;
        for (int l = 1; l < nlab; ++l) {
            const float v = u[l * n + i];
            if (v > best) {   // ties keep the first, as MATLAB max
                best = v;
                r = l + 1;
            }
        }
    } else {
        for (int l = 0; l < nlab - 1; ++l) r += u[l * n + i] > 0.5f ? 1 : 0;
    }
    lab[i] = (uint8_t)r;
}

}  // namespace asets
