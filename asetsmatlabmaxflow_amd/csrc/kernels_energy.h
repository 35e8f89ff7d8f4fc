// kernels_energy.h -- FP64 primal energy of a labelling u (no reference
// counterpart; the C-MEX solvers report only the mean |du| trace).
//
//   binary: E(u) = sum_g (1 - u_g) Cs_g + u_g Ct_g + alpha_g |grad u|_g
//   Potts : E(u) = sum_l sum_g u_lg Ct_lg + alpha_lg |grad u_l|_g
//
// |grad u|_g = sqrt(0.5 * sum over the low and the high face of each axis of the
// squared difference across that face), a face outside the volume counting 0:
// the same face set and 0.5 factor as the projection of
// asetsBinaryMF3D_mex.c:321-324, so an axis-aligned unit edge costs
// sqrt(2)*alpha, the capacity the projection gives one face.
//
// A fixed grid strides over the voxels; every block writes its (data, reg)
// partials and one block sums them in order, so the result is deterministic.
#pragma once
#include "common.h"

namespace asets {

constexpr int kEnergyBlocks = 1024;
constexpr int kEnergyThreads = 256;

template <bool IS3D>
__global__ __launch_bounds__(kEnergyThreads) void k_energy(
    const float *__restrict__ Cs, const float *__restrict__ Ct,
    const float *__restrict__ alpha, const float *__restrict__ u, int ny, int nx, int nz,
    int nlab, double *__restrict__ partials) {
    __shared__ double red[16];
    const int64_t sx = ny, sz = (int64_t)nx * ny;
    const int64_t N = sz * nz;
    const int64_t total = N * nlab;
    double d = 0.0, r = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = i % N;
        const int y = (int)(g % ny);
        const int x = (int)((g / ny) % nx);
        const int z = (int)(g / sz);
        const double ug = u[i];
        d += (Cs ? (1.0 - ug) * (double)Cs[i] : 0.0) + ug * (double)Ct[i];
        double t = 0.0, q;
        if (x > 0) { q = ug - (double)u[i - sx]; t += q * q; }
        if (x < nx - 1) { q = (double)u[i + sx] - ug; t += q * q; }
        if (y > 0) { q = ug - (double)u[i - 1]; t += q * q; }
        if (y < ny - 1) { q = (double)u[i + 1] - ug; t += q * q; }
        if (IS3D) {
            if (z > 0) { q = ug - (double)u[i - sz]; t += q * q; }
            if (z < nz - 1) { q = (double)u[i + sz] - ug; t += q * q; }
        }
        r += (double)alpha[i] * __builtin_sqrt(0.5 * t);
    }
    const double sd = block_sum(d, red);
    __syncthreads();
    const double sr = block_sum(r, red);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = sd;
        partials[2 * blockIdx.x + 1] = sr;
    }
}

// out[0] = E, out[1] = data term, out[2] = regularisation term.  One workgroup;
// thread t sums partials t, t + blockDim, ... and the block sum is fixed-order.
__global__ void k_energy_finalize(const double *partials, int nblocks, double *out) {
    __shared__ double red[16];
    double d = 0.0, r = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
        d += partials[2 * b];
        r += partials[2 * b + 1];
    }
    const double sd = block_sum(d, red);
    __syncthreads();
    const double sr = block_sum(r, red);
    if (threadIdx.x == 0) {
        out[0] = sd + sr;
        out[1] = sd;
        out[2] = sr;
    }
}

}  // namespace asets
