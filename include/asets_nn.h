/*
 * asets_nn.h -- C ABI of libasets_nn.so, the kernels of the learned warm start
 * for the binary max-flow solver: a small 3D U-Net (3x3x3 convolutions on the
 * gfx950 matrix cores, bf16 activations and weights, fp32 accumulation) that
 * maps (Cs, Ct, alpha) to a starting point (u0, pt0) for asets_binary_warm /
 * asets_binary_dev, or, in its space-to-depth form, to a full solver state
 * (u, ps, pt and the face flows) for asets_binary_dev_resume.
 *
 * None of these entry points has a counterpart in the reference: its solvers
 * always start from the initial cut of asetsBinaryMF3D_mex.c:205-224
 * (u = [Cs >= Ct], ps = pt = min(Cs, Ct)).  The warm start replaces exactly that
 * initialisation and leaves the iteration (and its stop rule) unchanged.
 *
 * Conventions: every pointer is device memory; tensors are NDHWC (channels
 * fastest) bf16 (16-bit storage passed as void *) unless typed float.  A batch
 * of n column-major volumes (Ny fastest) is NDHWC with D = Nz, H = Nx, W = Ny,
 * so voxel p = v*N + g keeps its solver index.  Tensors hold fewer than 2^31
 * elements.  Conv weights are [27][cout][cin] (tap-major, cin contiguous), tap
 * k = (dz+1)*9 + (dh+1)*3 + (dw+1).  Work is enqueued on `stream` (NULL = the
 * default stream); nothing synchronises.  Errors return a status;
 * asets_nn_last_error() holds the message.
 */
#ifndef ASETS_NN_H
#define ASETS_NN_H

#include <stddef.h>
#include <stdint.h>

#include "asets.h"

#ifdef __cplusplus
extern "C" {
#endif

/* out = act(conv3x3x3(in) + bias); in [n,d,h,w,cin], out [n,d,h,w,cout];
 * cin a multiple of 8; cout in {16,32,48,64,96,128}; bias may be NULL;
 * relu != 0 applies max(0, .).  Zero padding 1, stride 1. */
asets_status asets_nn_conv3(const void *in, int n, int d, int h, int w, int cin,
                            const void *wt, const float *bias, int cout, int relu, void *out,
                            void *stream);

/* A decoder convolution over the concatenation [up2(coarse), skip] without
 * materialising it: the same result, bit for bit, as asets_nn_up2(coarse, ...,
 * ca, cat, cin, 0) + asets_nn_chan_copy(cat, cin, ca, skip, cin - ca, 0, cin - ca,
 * ...) + asets_nn_conv3(cat, ...).  coarse [n,d/2,h/2,w/2,ca], skip
 * [n,d,h,w,cin-ca]; d, h, w even; ca a multiple of 8 in (0, cin); cout in
 * {16,32,64}. */
asets_status asets_nn_conv3_upcat(const void *coarse, int ca, const void *skip, int n, int d,
                                  int h, int w, int cin, const void *wt, const float *bias,
                                  int cout, int relu, void *out, void *stream);

/* Weights of the input-gradient convolution: wt [27][cout][cin] ->
 * wt_bwd [27][cin][cout] with taps mirrored (k -> 26-k).  Then
 * asets_nn_conv3(gout, ..., cout, wt_bwd, NULL, cin, 0, gin) is d loss / d in. */
asets_status asets_nn_flip_transpose(const void *wt, void *wt_bwd, int cout, int cin,
                                     void *stream);

/* g[i] = 0 where y[i] <= 0 (gradient through a ReLU with output y). */
asets_status asets_nn_relu_mask(void *g, const void *y, int64_t n, void *stream);

/* Weight and bias gradient of asets_nn_conv3: dw [27][cout][cin] fp32,
 * dbias [cout] fp32 (may be NULL); accumulate != 0 adds to them.  cout a
 * multiple of 16 up to 64; cin 8 or a multiple of 16 up to 128.  scratch: device memory of
 * asets_nn_wgrad_scratch_bytes().  Computed on the matrix cores (bf16 products, fp32
 * sums): each workgroup accumulates all 27 taps of one 16-channel input chunk over its
 * share of 64x2x2-voxel boxes, and the workgroups' partials are added in a fixed order
 * (deterministic). */
size_t asets_nn_wgrad_scratch_bytes(int n, int d, int h, int w, int cin, int cout);
asets_status asets_nn_conv3_wgrad(const void *in, int n, int d, int h, int w, int cin,
                                  const void *gout, int cout, float *dw, float *dbias,
                                  int accumulate, void *scratch, size_t scratch_bytes,
                                  void *stream);

/* 2x2x2 average pool, (d, h, w) the fine extents (even): out [n,d/2,h/2,w/2,c];
 * backward: gin [n,d,h,w,c] = gout/8 broadcast (added when accumulate != 0). */
asets_status asets_nn_pool2(const void *in, int n, int d, int h, int w, int c, void *out,
                            void *stream);
asets_status asets_nn_pool2_bwd(const void *gout, int n, int d, int h, int w, int c, void *gin,
                                int accumulate, void *stream);

/* Nearest 2x upsample of in [n,d/2,h/2,w/2,c] into channels [off, off+c) of
 * out [n,d,h,w,cd] (a concatenation buffer); backward sums the 8 fine gradients
 * of that channel range of gout [n,d,h,w,cd] into gin [n,d/2,h/2,w/2,c]. */
asets_status asets_nn_up2(const void *in, int n, int d, int h, int w, int c, void *out, int cd,
                          int off, void *stream);
asets_status asets_nn_up2_bwd(const void *gout, int cd, int off, int n, int d, int h, int w,
                              int c, void *gin, void *stream);

/* dst[p, doff+c] (+)= src[p, soff+c], c < nc, over nvox voxels (skip
 * connections into a concatenation buffer and their gradients back). */
asets_status asets_nn_chan_copy(void *dst, int cd, int doff, const void *src, int cs, int soff,
                                int nc, int64_t nvox, int accumulate, void *stream);

/* Input features [P, 8] of P voxels: Cs/10, Ct/10, alpha, (Ct-Cs)/10,
 * min(Cs,Ct)/10, 0, 0, 0. */
asets_status asets_nn_features(const float *Cs, const float *Ct, const float *alpha, int64_t P,
                               void *out, void *stream);

/* Head: 1x1x1 convolution of x [P,16] with w = [w_u(16), w_pt(16)], b = [b_u, b_pt]:
 *   u0 = w_u.x + b_u,   pt0 = min(Cs, Ct) + 10 * (w_pt.x + b_pt)
 * (pt0 is predicted relative to the reference's initial pt). */
asets_status asets_nn_head(const void *x, int64_t P, const float *w, const float *b,
                           const float *Cs, const float *Ct, float *u0, float *pt0,
                           void *stream);

/* Loss mean((u0-tu)^2) + mean(((pt0-tpt)/10)^2) and its gradients: dx [P,16]
 * bf16; grads (fp32, 35 values) = d/dw (32), d/db (2), then the loss * P.
 * scratch: device memory of asets_nn_head_scratch_bytes(P). */
size_t asets_nn_head_scratch_bytes(int64_t P);
asets_status asets_nn_head_bwd(const void *x, int64_t P, const float *w, const float *u0,
                               const float *pt0, const float *tu, const float *tpt, void *dx,
                               float *grads, void *scratch, size_t scratch_bytes,
                               void *stream);

/* ---- space-to-depth network: half resolution in, the full solver state out ----
 * Coarse voxel (v, z2, x2, y2) of the half-resolution tensors [n, d/2, h/2, w/2, c]
 * stands for the 2x2x2 fine voxels s = (dz*2 + dx)*2 + dy at (2*z2+dz, 2*x2+dx,
 * 2*y2+dy) (dz along d = Nz, dx along h = Nx, dy along w = Ny).  (d, h, w) are the
 * fine extents, all even; the fine float fields are 8-byte aligned (a dy pair is one
 * 8-byte access). */

/* Input features [n, d/2, h/2, w/2, 32]: channel s*4 + f of sub-voxel s holds
 * f = Cs/10, Ct/10, alpha, (Ct-Cs)/10. */
asets_status asets_nn_features_s2d(const float *Cs, const float *Ct, const float *alpha, int n,
                                   int d, int h, int w, void *out, void *stream);

/* Head: 1x1x1 convolution of x [n, d/2, h/2, w/2, 16] with w [32][16] and b [32];
 * output o = s*4 + q goes to sub-voxel s of field q: u0, then the face flows bx0,
 * by0, bz0 in the solver's own-face convention (bx0[g] is the face between x-1 and
 * x).  Fine fields of n*d*h*w floats; wt 16-byte aligned. */
asets_status asets_nn_head_s2d(const void *x, int n, int d, int h, int w, const float *wt,
                               const float *b, float *u0, float *bx0, float *by0, float *bz0,
                               void *stream);

/* Loss (1/P) sum over the P = n*d*h*w fine voxels of (u0-tu)^2 + lam *
 * ((bx0-tbx)^2 + (by0-tby)^2 + (bz0-tbz)^2) and its gradients: dx [n,d/2,h/2,w/2,16]
 * bf16; grads (fp32, 545 values) = d/dw (512, [32][16]), d/db (32), then the loss * P.
 * Deterministic (fixed-order sums).  scratch: asets_nn_head_s2d_scratch_bytes(). */
size_t asets_nn_head_s2d_scratch_bytes(int n, int d, int h, int w);
asets_status asets_nn_head_s2d_bwd(const void *x, int n, int d, int h, int w, const float *wt,
                                   const float *u0, const float *bx0, const float *by0,
                                   const float *bz0, const float *tu, const float *tbx,
                                   const float *tby, const float *tbz, float lam, void *dx,
                                   float *grads, void *scratch, size_t scratch_bytes,
                                   void *stream);

/* Solver state for asets_binary_dev_resume from a predicted (u0, bx0, by0, bz0) of n
 * column-major volumes (ny, nx, nz) stored back to back, in place: u0 clamped to
 * [0, 1] and the zero-flux faces (axis index 0) set to 0; ps and pt written as the
 * fixed point of the reference's source/sink update (asetsBinaryMF3D_mex.c:411-429)
 * pairs them with u and the faces: dv = div b, pt = Ct where u >= 1/2 and
 * min(Cs - dv, Ct) elsewhere, ps = pt + dv. */
asets_status asets_nn_warm_state(float *u0, float *bx0, float *by0, float *bz0, const float *Cs,
                                 const float *Ct, int64_t ny, int64_t nx, int64_t nz, int n,
                                 float *ps, float *pt, void *stream);

/* AdamW step `step` (1-based) on fp32 weights w with moments m, v; refreshes the
 * bf16 copy wb when it is not NULL. */
asets_status asets_nn_adamw(float *w, void *wb, const float *g, float *m, float *v, int64_t n,
                            float lr, float b1, float b2, float eps, float wd, int step,
                            void *stream);

/* b = bf16(a), round to nearest even. */
asets_status asets_nn_to_bf16(const float *a, void *b, int64_t n, void *stream);

const char *asets_nn_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* ASETS_NN_H */
