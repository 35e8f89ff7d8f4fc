/*
 * asets.h -- C ABI of libasets_hip.so, the gfx950 (MI355X) implementation of
 * ASETS' continuous max-flow primal-dual iteration.
 *
 * No HIP, torch or C++ types cross this boundary: plain pointers, sizes and
 * POD structs.  Every entry point returns an asets_status and never aborts the
 * process; asets_last_error() gives the message of the calling thread's last
 * failure.
 *
 * Array convention (as the MATLAB C-MEX solvers it replaces): float32,
 * column-major, index g = z*Nx*Ny + x*Ny + y with Ny (rows, pars[0]) fastest;
 * Potts label l at offset l*Nx*Ny*Nz.
 *
 * Outputs follow maxflow/asetsBinaryMF3D_mex.c:91-118:
 *   u       [Ny Nx (Nz) (L)]  labelling function
 *   cvg     [max_iter]        per-iteration mean |delta u|; entries after the
 *                             stop index are 0
 *   it_num                    0-based index of the iteration at which
 *                             cvg <= err_bound held, or max_iter if none did
 *                             (asetsBinaryMF3D_mex.c:163-192)
 *   seconds                   wall time of the solve (the reference reports
 *                             process CPU time, asetsBinaryMF3D_mex.c:122-131)
 * cvg is the FP64 sum of |e| over the volume divided by N and rounded to
 * float; the reference sums in float32 sequentially, so near err_bound the two
 * can stop on different iterations.  At a fixed iteration count u is
 * bit-identical to the reference operation order.
 *
 * alpha <= 0 anywhere (checked when max_iter > 0, as the reference checks it
 * inside its first projection pass) returns ASETS_EALPHA with the reference's
 * message.  The check runs on the device ahead of (or, for a cold start of the
 * fused binary solver, within) the first iteration and is read at the first
 * stop-flag poll.  On that or any other error the contents of the device
 * output arrays (u, cvg, it_num) are unspecified; the host-pointer entry
 * points (and the mex files) write no output on error.
 */
#ifndef ASETS_H
#define ASETS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum asets_status {
    ASETS_OK = 0,
    ASETS_EBADARG = 1,   /* null pointer, non-positive size, bad option       */
    ASETS_EALPHA = 2,    /* some alpha <= 0 (asetsBinaryMF3D_mex.c:317-319)   */
    ASETS_ENOMEM = 3,    /* device or pinned-host allocation failed           */
    ASETS_EDEVICE = 4,   /* HIP runtime error                                 */
    ASETS_ETOOBIG = 5    /* size beyond a documented limit                    */
} asets_status;

/* Solver parameters: the numeric tail of the reference's pars vector
 * (asetsBinaryMF3D_mex.c:82-85, asetsPotts3D_mex.c:88-91). */
typedef struct asets_params {
    int32_t max_iter;    /* pars[maxIt]     */
    float err_bound;     /* pars[errBound]  */
    float cc;            /* pars[cc]        step size of the augmented Lagrangian */
    float steps;         /* pars[steps]     step size of the flow gradient step   */
} asets_params;

/* Flow storage precision. */
#define ASETS_FLOW_F32 0
#define ASETS_FLOW_F16 1   /* ps, pt, bx, by, bz stored in fp16; arithmetic in fp32 */

/* Kernel family. */
#define ASETS_KERNEL_AUTO 0
#define ASETS_KERNEL_FUSED 1    /* one z-streaming kernel per iteration (binary 3D/2D) */
#define ASETS_KERNEL_PASSES 2   /* one kernel per reference pass (all models)          */

/* Optional execution settings; pass NULL for defaults. */
typedef struct asets_opts {
    void *stream;          /* hipStream_t to run on (device-pointer entries); NULL = null stream */
    int32_t flow_dtype;    /* ASETS_FLOW_*                       (binary only)                     */
    int32_t kernel;        /* ASETS_KERNEL_*                                                        */
    int32_t poll_every;    /* iterations enqueued between host polls of the stop flag; 0 = auto     */
    int32_t tile_x;        /* fused kernel tile width along x (8, 16, 32); 0 = library's choice     */
    int32_t zchunk;        /* fused kernel planes per workgroup; 0 = library's choice (one volume:  */
                           /* a per-CU load model; batched: timed on the caller's data, cached)     */
    int32_t reserved[3];
    double *kernel_ms;     /* out (optional): mean device time per iteration of the solve loop      */
} asets_opts;

/* ---- host-pointer entry points: drop-in for the four C-MEX solvers ------- */

/* asetsBinaryMF2D_mex(Cs, Ct, alpha, [Ny Nx maxIt errBound cc steps])
 * replaces maxflow/asetsBinaryMF2D_mex.c:36-129 (mexFunction) and :132-358. */
asets_status asets_binary2d(const float *Cs, const float *Ct, const float *alpha,
                            int64_t ny, int64_t nx, const asets_params *p,
                            float *u, float *cvg, int32_t *it_num, double *seconds,
                            const asets_opts *o);

/* asetsBinaryMF3D_mex(Cs, Ct, alpha, [Ny Nx Nz maxIt errBound cc steps])
 * replaces maxflow/asetsBinaryMF3D_mex.c:39-134 (mexFunction) and :137-438. */
asets_status asets_binary3d(const float *Cs, const float *Ct, const float *alpha,
                            int64_t ny, int64_t nx, int64_t nz, const asets_params *p,
                            float *u, float *cvg, int32_t *it_num, double *seconds,
                            const asets_opts *o);

/* Binary 2D (ndim 2, nz == 1) or 3D solve on host arrays with an optional warm
 * start: u0/pt0 both NULL reproduces asets_binary2d/3d; both given start from
 * u = u0, ps = pt = pt0 and zero spatial flows.  The warm start has no
 * reference counterpart (asetsBinaryMF3D_mex.c:205-224 always cold-starts);
 * it backs the six-argument form of the mex host files. */
asets_status asets_binary_warm(int ndim, const float *Cs, const float *Ct, const float *alpha,
                               int64_t ny, int64_t nx, int64_t nz, const asets_params *p,
                               const float *u0, const float *pt0,
                               float *u, float *cvg, int32_t *it_num, double *seconds,
                               const asets_opts *o);

/* asetsPotts2D_mex(Ct, alpha, [Ny Nx L maxIt errBound cc steps])
 * replaces maxflow/asetsPotts2D_mex.c:42-135 (mexFunction) and :138-434. */
asets_status asets_potts2d(const float *Ct, const float *alpha,
                           int64_t ny, int64_t nx, int64_t nlab, const asets_params *p,
                           float *u, float *cvg, int32_t *it_num, double *seconds,
                           const asets_opts *o);

/* asetsPotts3D_mex(Ct, alpha, [Ny Nx Nz L maxIt errBound cc steps])
 * replaces maxflow/asetsPotts3D_mex.c:44-141 (mexFunction) and :144-521. */
asets_status asets_potts3d(const float *Ct, const float *alpha,
                           int64_t ny, int64_t nx, int64_t nz, int64_t nlab,
                           const asets_params *p,
                           float *u, float *cvg, int32_t *it_num, double *seconds,
                           const asets_opts *o);

/* asetsIshikawa2D_mex(Ct, alpha, [Ny Nx L maxIt errBound cc steps])
 * replaces maxflow/asetsIshikawa2D_mex.c:40-133 (mexFunction) and :136-481.
 * Ct has L label fields, alpha L-1 (asetsIshikawa2D_mex.c reads only those);
 * u receives the L-1 cut functions [Ny Nx L-1].  Stops on cvg < errBound
 * (asetsIshikawa2D_mex.c:205; the 3D solver uses <=). */
asets_status asets_ishikawa2d(const float *Ct, const float *alpha,
                              int64_t ny, int64_t nx, int64_t nlab, const asets_params *p,
                              float *u, float *cvg, int32_t *it_num, double *seconds,
                              const asets_opts *o);

/* asetsIshikawa3D_mex(Ct, alpha, [Ny Nx Nz L maxIt errBound cc steps])
 * replaces maxflow/asetsIshikawa3D_mex.c:42-139 (mexFunction) and :141-612;
 * u receives [Ny Nx Nz L-1]; stops on cvg <= errBound (:213). */
asets_status asets_ishikawa3d(const float *Ct, const float *alpha,
                              int64_t ny, int64_t nx, int64_t nz, int64_t nlab,
                              const asets_params *p,
                              float *u, float *cvg, int32_t *it_num, double *seconds,
                              const asets_opts *o);

/* ---- device-pointer entry points (inputs already resident in HBM) -------- */

/* Binary solve on device memory.  ndim is 2 (nz must be 1; the 2D operation
 * order of asetsBinaryMF2D_mex.c) or 3.  u0/pt0 (device, optional, both or
 * neither): warm start u = u0, ps = pt = pt0, spatial flows 0 -- a superset of
 * the reference, which always starts from asetsBinaryMF3D_mex.c:205-224.
 * cvg and it_num are device pointers; seconds is a host pointer (optional). */
asets_status asets_binary_dev(int ndim, const float *Cs, const float *Ct, const float *alpha,
                              int64_t ny, int64_t nx, int64_t nz, const asets_params *p,
                              const float *u0, const float *pt0,
                              float *u, float *cvg, int32_t *it_num, double *seconds,
                              const asets_opts *o);

/* Workspace of asets_binary_dev_noalloc: a file-scope device array of this many
 * bytes per device, held from the library's first use of a device until the
 * process ends. */
#define ASETS_NOALLOC_SCRATCH_BYTES 104857600

/* asets_binary_dev without allocation.  No counterpart in the reference; it is
 * for callers that cannot allocate: code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc.  Same arguments and results as
 * asets_binary_dev (u bit for bit), but the solve's device workspace (two
 * state buffers, per-workgroup partial sums, flags) is carved from the
 * library's ASETS_NOALLOC_SCRATCH_BYTES array, and the stop flag is polled
 * into host stack memory: no hipMalloc, hipHostMalloc or hipEventCreate.  A
 * solve needing more workspace returns ASETS_ETOOBIG before any device work:
 * the fused fp32 solver needs about 48 bytes per voxel (128^3 fits), fp16
 * flows 28, the per-pass kernels 28.  o->kernel_ms must be NULL.  The array
 * keeps the last call's contents; every call initialises what it reads.  All
 * calls on one device share it: the library serialises them (each returns
 * after its work on the stream is complete), so calls from several host
 * threads wait for each other instead of racing. */
asets_status asets_binary_dev_noalloc(int ndim, const float *Cs, const float *Ct,
                                      const float *alpha, int64_t ny, int64_t nx, int64_t nz,
                                      const asets_params *p, const float *u0, const float *pt0,
                                      float *u, float *cvg, int32_t *it_num, double *seconds,
                                      const asets_opts *o);

/* Same as asets_binary_dev, additionally returning the final flows
 * (device pointers, N floats each, any may be NULL): ps, pt and the three
 * face fields in own-face convention (bx[g] is the face between x-1 and x). */
asets_status asets_binary_dev_state(int ndim, const float *Cs, const float *Ct,
                                    const float *alpha, int64_t ny, int64_t nx, int64_t nz,
                                    const asets_params *p, const float *u0, const float *pt0,
                                    float *u, float *cvg, int32_t *it_num, double *seconds,
                                    float *ps, float *pt, float *bx, float *by, float *bz,
                                    const asets_opts *o);

/* Resume a binary solve from a full iteration state (no reference counterpart: the
 * reference always starts from the initial cut, asetsBinaryMF3D_mex.c:205-224).  The
 * state is (u, ps, pt, bx, by, bz) exactly as asets_binary_dev_state returns it:
 * bx[g] is the flow on the face between voxels x-1 and x (by: y-1|y, bz: z-1|z);
 * the faces with axis index 0 are the zero-flux boundary and are taken as 0 whatever
 * the input holds.  A solve of k iterations followed by a resume of m iterations from
 * its state gives the same u (bit for bit) as one solve of k+m iterations.  bz0 is
 * ignored (may be NULL) when ndim == 2; ps..bz out may be NULL.
 * ps0 and pt0 may both be NULL (a predicted state of u and the faces): they are then
 * paired to them as the fixed point of the source/sink update
 * (asetsBinaryMF3D_mex.c:411-429): dv = div b, pt = Ct where u >= 1/2 and
 * min(Cs - dv, Ct) elsewhere, ps = pt + dv, with u clamped to [0, 1] -- the pairing of
 * asets_nn_warm_state, formed while the state is laid out, without its passes. */
asets_status asets_binary_dev_resume(int ndim, const float *Cs, const float *Ct,
                                     const float *alpha, int64_t ny, int64_t nx, int64_t nz,
                                     const asets_params *p, const float *u0, const float *ps0,
                                     const float *pt0, const float *bx0, const float *by0,
                                     const float *bz0, float *u, float *cvg, int32_t *it_num,
                                     double *seconds, float *ps, float *pt, float *bx, float *by,
                                     float *bz, const asets_opts *o);

/* Potts solve on device memory; ndim 2 (nz == 1) or 3. */
asets_status asets_potts_dev(int ndim, const float *Ct, const float *alpha,
                             int64_t ny, int64_t nx, int64_t nz, int64_t nlab,
                             const asets_params *p,
                             float *u, float *cvg, int32_t *it_num, double *seconds,
                             const asets_opts *o);

/* Ishikawa solve on device memory; ndim 2 (nz == 1, stop on <) or 3 (<=). */
asets_status asets_ishikawa_dev(int ndim, const float *Ct, const float *alpha,
                                int64_t ny, int64_t nx, int64_t nz, int64_t nlab,
                                const asets_params *p,
                                float *u, float *cvg, int32_t *it_num, double *seconds,
                                const asets_opts *o);

/* ---- batched binary 3D: many volumes of uneven size, one launch per iteration
 * No reference counterpart (the reference solves one volume per call).  Each
 * volume has its own sizes, arrays, cvg/it_num outputs and stop decision; a
 * converged volume drops out of the work list.  1 <= nvol <= 65535. */
typedef struct asets_volume {
    const float *Cs, *Ct, *alpha;   /* device */
    int64_t ny, nx, nz;
    float *u;                       /* device, out */
    float *cvg;                     /* device, out [max_iter] */
    int32_t *it_num;                /* device, out */
} asets_volume;

asets_status asets_binary3d_batched_dev(int32_t nvol, const asets_volume *vols,
                                        const asets_params *p, double *seconds,
                                        const asets_opts *o);

/* ---- energy of a labelling (FP64) ----------------------------------------
 * No reference counterpart: asetsBinaryMF3D_mex.c:183 and asetsPotts3D_mex.c:203
 * report only the mean |du| trace.  Binary:
 *   E(u) = sum (1-u) Cs + u Ct + alpha |grad u|;  Potts: sum_l u_l Ct_l + alpha_l |grad u_l|,
 * where |grad u| at a voxel is sqrt(0.5 * sum of the squared differences across its
 * low and high face on each axis) -- the face set and 0.5 factor of the projection
 * at asetsBinaryMF3D_mex.c:321-324; faces outside the volume count 0.  Device
 * arrays, column-major; energy (host) receives {E, data term, regularisation term}.
 * Accumulated in FP64 with a fixed reduction order (deterministic). */
asets_status asets_binary_energy_dev(int ndim, const float *Cs, const float *Ct,
                                     const float *alpha, int64_t ny, int64_t nx, int64_t nz,
                                     const float *u, double *energy, const asets_opts *o);
asets_status asets_potts_energy_dev(int ndim, const float *Ct, const float *alpha, int64_t ny,
                                    int64_t nx, int64_t nz, int64_t nlab, const float *u,
                                    double *energy, const asets_opts *o);

/* ---- primal-dual bounds of a binary state (FP64) ------------------------------
 * No reference counterpart (the reference stops on the mean |du| trace,
 * asetsBinaryMF3D_mex.c:183-188).  For the model max sum ps, ps <= Cs, pt <= Ct,
 * div p - ps + pt = 0, per-voxel capacity 0.5 * (sum of the six faces^2) <= alpha^2
 * (projStep, asetsBinaryMF3D_mex.c:321-331), and a state u, bx, by, bz (device
 * arrays, column-major, own-face convention as asets_binary_dev_state returns them;
 * bz unused and may be NULL in 2D):
 *   bounds[2] = D(p~) = sum min(Cs, Ct + div p~), the flows scaled face by face to
 *               the smaller feasibility factor of their two voxels (a lower bound of
 *               the optimum);
 *   bounds[0] = P(u) with the regulariser bounded by F(lambda) after `sweeps`
 *               red-black Newton sweeps over lambda (an upper bound);
 *   bounds[1] = the same with the voxel-centred lambda (the looser TV bound).
 * u is clipped to [0, 1]; alpha must be > 0.  On the reference iteration the bracket
 * does not close: its rescale (asetsBinaryMF3D_mex.c:337-397) averages the two
 * voxels' factors, so its fixed point is not this problem's saddle point, and the
 * relative width levels off (about 0.7 % on the 256^3 workload).  It certifies how
 * far a state can be from the optimum; it is not a stop rule. */
asets_status asets_binary_bounds_dev(int ndim, const float *Cs, const float *Ct,
                                     const float *alpha, const float *u, const float *bx,
                                     const float *by, const float *bz, int64_t ny, int64_t nx,
                                     int64_t nz, int32_t sweeps, double *bounds,
                                     const asets_opts *o);

/* ---- before and after the solve: data costs, weights, labels (device arrays) --
 * The steps the reference's applications and tests run around the solvers, on the
 * GPU.  No reference C counterpart (they are MATLAB code); semantics as cited. */

/* Log-likelihood data costs, applications/interactiveGraphCutSegmentation3D.m:96-118
 * (computeLogLikelihoodCost) for labels 1..nlab: the image is normalised to
 * x = ((img - min) / (max - min)) * 255 in fp32; label l's 256-bin histogram
 * counts its scribbled voxels (scribbles == l) in bin floor(x) (histc with edges
 * 0..255, x == 255 in the last bin); Ct[(l-1)*n + i] = -log10(P_l(uint16(x_i)) +
 * epsilon) with P the normalised histogram and uint16() rounding half away from
 * zero, in FP64 stored as fp32.  A constant image or a label without scribbles is
 * an error (MATLAB would produce NaN costs).  nlab <= 64. */
asets_status asets_cost_loglik_dev(const float *img, const int32_t *scribbles, int64_t n,
                                   int32_t nlab, double epsilon, float *Ct,
                                   const asets_opts *o);

/* Ct[i] = -log(p[i] + epsilon) (natural logarithm, FP64 stored as fp32), the data term
 * from probability maps of applications/atlasProbabilityMapsRegularization3D.m:47-53. */
asets_status asets_cost_neglog_dev(const float *p, int64_t n, double epsilon, float *Ct,
                                   const asets_opts *o);

/* Contrast-sensitive regularisation weight, applications/smoothManualSegmentations2D.m:
 * 73-82: I normalised to [0, 1] when normalize != 0, the gradient of MATLAB's gradient()
 * (central differences inside, one-sided at the borders; ndim 2 or 3), and
 *   alpha = a0 + a1 * exp(-beta * m),  m = |grad I| (power 1) or |grad I|^2 (power 2),
 * in FP64 stored as fp32.  The reference application uses a0 = 0, a1 = 0.2, beta = 10,
 * power 1. */
asets_status asets_alpha_contrast_dev(int ndim, const float *img, int64_t ny, int64_t nx,
                                      int64_t nz, double a0, double a1, double beta, int power,
                                      int normalize, float *alpha, const asets_opts *o);

/* Discrete labels from a solution u of n voxels: model 0 (binary) u > 0.5
 * (tests/runTestBinaryMF.m:55); model 1 (Potts, nlab fields) the 1-based index of the
 * first maximum over labels (tests/runTestPotts.m:58, MATLAB max); model 2 (Ishikawa,
 * nlab-1 cut functions) the number of cut functions > 0.5 (tests/runTestIshikawa.m:57-60). */
asets_status asets_labels_dev(int model, const float *u, int64_t n, int32_t nlab,
                              uint8_t *labels, const asets_opts *o);

/* L1 intensity costs, applications/L1IntensitySegmentation2D.m:36-44: the image is
 * normalised to [0, 1] in fp32, x = (img - min) / (max - min), and
 * Ct[l*n + i] = |x_i - models[l]| for nlab <= 64 model intensities (host array). */
asets_status asets_cost_l1_dev(const float *img, int64_t n, const float *models, int32_t nlab,
                               float *Ct, const asets_opts *o);

/* Smoothed inverse label indicators, applications/smoothManualSegmentations2D.m:48-55:
 * Ct[l*n + g] = imfilter(1 - (labels == ids[l]), h, 'replicate') on an ny x nx label
 * image (int32, column-major), h a host K x K kernel in double, column-major (K odd,
 * <= 15; the script uses fspecial('gaussian', [5 5], 1.5)).  The window's weights are
 * summed in FP64 in h's element order and the sum stored as fp32.  nlab <= 64. */
asets_status asets_cost_smooth_labels_dev(const int32_t *labels, int64_t ny, int64_t nx,
                                          const int32_t *ids, int32_t nlab, const double *h,
                                          int32_t K, float *Ct, const asets_opts *o);

/* ---- hierarchical max-flow (HMF), full-flow variant -----------------------
 * Replaces the MATLAB classes maxflow/asetsHMF2D.m and asetsHMF3D.m: the tree
 * built by the constructor asetsHMF2D(children, alpha, Ct) (:57-76, one node per
 * object), the public properties lx, ly(, lz) (shape constraint) and Name, and
 * MaxFullFlow(numIts, steps, cc) (2D :78-88, 3D :77-87: InitializeFullFlow, then
 * numIts times UpdateSinkFlows, UpdateLabels, UpdateSpatialFlows, then
 * CalculateCostsFullFlow).  The pseudo-flow variant (MaxPseudoFlow) and the
 * plotting of named nodes are not provided.
 *
 * One entry per object of the tree.  parent is the index of the parent entry (-1
 * for the single root); the children of a node are taken in increasing index
 * order, the order of the reference's cell array.  A node without children is a
 * leaf and needs Ct.  Arithmetic is fp32 (the reference computes in the class of
 * Ct; single-precision Ct gives these results element for element).  Device
 * arrays of N = ny*nx*nz floats, column-major; faces in own-face layout: faces[0]
 * along dim 1 (px), faces[1] along dim 2 (py), faces[2] along dim 3 (pz), each
 * voxel holding the face to its lower neighbour, so px(i,j,k) = faces[0][g] at
 * row i (0-based) and the index-0 faces are 0.  Output pointers may be NULL.
 * ndim == 2 needs nz == 1.  At most 4096 nodes and 16 levels of internal nodes.
 * cc is a double, as MATLAB's: the fp32 updates use (float)cc, while the root's
 * sink term 1/cc .* ones(h.D) (asetsHMF3D.m:320) is evaluated in double and
 * rounded to fp32 where it meets the children's sink flows. */
typedef struct asets_hmf_node {
    int32_t parent;          /* parent entry, -1 for the root                              */
    float alpha;             /* scalar alpha, used when alpha_map is NULL                  */
    const float *alpha_map;  /* per-voxel alpha [N], or NULL                               */
    const float *Ct;         /* data cost [N] of a leaf                                    */
    const float *l[3];       /* shape-constraint directions lx, ly(, lz) [N]: all or none  */
    float *u, *pt, *div;     /* out [N]: label, sink flow, divergence of the spatial flow  */
    float *faces[3];         /* out [N]: spatial flows px, py(, pz) in own-face layout     */
    float *l_out[3];         /* out [N]: the normalised directions (InitializeFullFlow)    */
    double data_cost;        /* out: dataCost of the subtree (CalculateCostsFullFlow)      */
    double regu_cost;        /* out: reguCost of the subtree                               */
} asets_hmf_node;

asets_status asets_hmf_dev(int ndim, int64_t ny, int64_t nx, int64_t nz,
                           asets_hmf_node *nodes, int32_t nnodes, int32_t num_iter,
                           float steps, double cc, double *seconds, const asets_opts *o);

/* ---- misc ---------------------------------------------------------------- */
/* Self-check of the short correctly-rounded division form the kernels use for
 * x/cc, n/alpha and 1/n (no reference counterpart): compares it with IEEE
 * division over all 2^32 float bit patterns x.  mode 0: x/d for the divisor d
 * (|d| in [2^-16, 2^16]); mode 1: 1/x for every x in that range.  Returns the
 * number of bitwise mismatches in *mismatches (0 expected). */
asets_status asets_selftest_division(int mode, float d, uint64_t *mismatches);

/* HBM ceiling (measurement only; no reference counterpart): times a plain stream
 * over the caller's device buffer buf of `bytes` bytes (256-byte aligned) with
 * 16-byte accesses per lane.  mode 0 reads the buffer (less a small sink the lanes
 * write their folded value to); mode 1 copies its first half to its second half;
 * mode 2 moves the binary solver's per-voxel traffic over the N voxels the buffer
 * holds (N = bytes/60 rounded down to a multiple of 1024): read Cs, Ct, alpha, u, bz
 * (4 B) and the (ps, pt), (bx, by) pairs (8 B), write u, bz and the two pairs --
 * 36 B read and 24 B written per voxel, without the stencil; mode 3 moves the
 * same bytes through three arrays (inputs interleaved at 12 B per voxel, state
 * read at 24 B, state written at 24 B); mode 4 keeps Cs, Ct, alpha in three
 * arrays and interleaves only the state (five arrays); mode 5 moves the same five
 * arrays with the fused binary kernel's layout and access widths (one voxel per
 * lane: 4-byte inputs, 1536-byte state blocks of 64 voxels read as 4- and 8-byte
 * fields).  flags bit 0:
 * non-temporal cache policy; bit 1: one element per lane in flight instead of four
 * (modes 0, 1).  grid: workgroups of 256 threads, 0 = 8 per CU.  One untimed launch,
 * then `iters` launches timed with HIP events on the stream of o: *ms = mean time
 * per launch, *traffic = bytes read plus written per launch. */
asets_status asets_hbm_stream_dev(int mode, void *buf, int64_t bytes, int32_t flags,
                                  int32_t grid, int32_t iters, double *ms, int64_t *traffic,
                                  const asets_opts *o);

const char *asets_last_error(void);
const char *asets_version(void);
/* Free the library's cached device workspaces (all devices). */
void asets_release(void);

#ifdef __cplusplus
}
#endif
#endif /* ASETS_H */
