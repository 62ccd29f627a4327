/*
 * kevernals_amd -- C ABI of the MI355X (gfx950) tomosynthesis projection and
 * reconstruction path.
 *
 * Every entry point takes plain pointers and sizes and returns a kv_status;
 * none of them exits the process.  Device-pointer entry points enqueue work on
 * the context's HIP stream and return without synchronising; the *_host
 * variants copy in, run and copy out synchronously (the calling convention of
 * the reference's Projector, whose wrappers were synchronous).
 *
 * Layouts (as KevLoq/Kevernals source/modules/reconstruction/Projector.cu):
 *   volume       float[nz][ny][nx], x fastest         (Projector.cu:361-363,401)
 *   projections  float[np][nv][nu], u fastest         (Projector.cu:264,649,658)
 *   world frame  millimetres, volume centred on the origin (Projector.cu:229-230)
 *
 * Operator modes:
 *   KV_MODE_R  reference discretisation: weighted-mean ray-driven forward
 *              projector (Projector.cu:227-372) and weighted-mean voxel-driven
 *              +-5 pixel backprojector (Projector.cu:376-785).
 *   KV_MODE_A  matched pair: Siddon line integrals in mm and their exact
 *              transpose (no reference counterpart; used by the likelihood
 *              and the learned steps).
 */
#ifndef KEVERNALS_AMD_H
#define KEVERNALS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum kv_status {
    KV_OK = 0,
    KV_ERR_INVALID = 1,      /* bad argument or geometry */
    KV_ERR_DEVICE = 2,       /* HIP runtime failure */
    KV_ERR_NOMEM = 3,        /* device allocation failed */
    KV_ERR_STATE = 4,        /* call order (e.g. no geometry set) */
    KV_ERR_PRECONDITION = 5  /* algorithmic precondition (ShiftAndAdd checks) */
} kv_status;

typedef enum kv_mode { KV_MODE_R = 0, KV_MODE_A = 1 } kv_mode;

/* The kernel argument list of CudaPerformProjection /
 * CudaPerformBackProjection (Projector.cu:227-236, 376-385) with per-view
 * arrays as host pointers: det[3p+0..2] = detector bottom-left (x, y) and
 * plane z of view p (Projector.cu:917-923); src[3p+0..2] = source position of
 * view p (Projector.cu:930-936).  The reference's unused volume origin
 * argument is dropped. */
typedef struct kv_geometry {
    int32_t nx, ny, nz;  /* volume size in voxels */
    float sx, sy, sz;    /* voxel spacing, mm */
    int32_t nu, nv, np;  /* detector pixels per row, rows, views */
    float pu, pv;        /* pixel pitch, mm */
    const float *det;    /* host, np*3 floats */
    const float *src;    /* host, np*3 floats */
} kv_geometry;

typedef struct kv_context kv_context;

/* ---- context ----------------------------------------------------------- */
/* Replaces the per-call cudaMalloc/cudaMemcpy/launch/sync/free sequence of
 * Projector::PerformProjection (Projector.cu:893-986): a context owns a
 * stream, the device copy of the geometry and its precomputed tables. */
kv_status kv_create(int device, kv_context **out);
kv_status kv_destroy(kv_context *ctx);
/* Use an external HIP stream (hipStream_t) for all later work; NULL selects
 * the null (default) stream.  A new context uses a stream of its own; a
 * caller that allocates with a stream-ordered allocator (torch's caching
 * allocator) must enqueue on the allocator's stream. */
kv_status kv_set_stream(kv_context *ctx, void *hip_stream);
void *kv_get_stream(kv_context *ctx);
kv_status kv_synchronize(kv_context *ctx);
/* Message of the last failure on the calling thread ("" if none). */
const char *kv_last_error(void);
/* Library version string. */
const char *kv_version(void);

/* Geometry -> kernel arguments (Projector.cu:876-940).  Copies the per-view
 * arrays to the device and builds the per-view crossing tables. */
kv_status kv_set_geometry(kv_context *ctx, const kv_geometry *geom);

/* ---- device memory helpers (for hosts without their own allocator) ----- */
kv_status kv_malloc(kv_context *ctx, size_t bytes, void **d_ptr);
kv_status kv_free(kv_context *ctx, void *d_ptr);
kv_status kv_memcpy_h2d(kv_context *ctx, void *d_dst, const void *h_src, size_t bytes);
kv_status kv_memcpy_d2h(kv_context *ctx, void *h_dst, const void *d_src, size_t bytes);
kv_status kv_memset(kv_context *ctx, void *d_dst, int value, size_t bytes);
/* Device-to-device copy on the context's stream (asynchronous). */
kv_status kv_memcpy_d2d(kv_context *ctx, void *d_dst, const void *d_src, size_t bytes);

/* ---- operators (device pointers, asynchronous) --------------------------- */
/* Forward projection: CudaPerformProjection (Projector.cu:227-372) for
 * KV_MODE_R; line integrals in mm for KV_MODE_A. */
kv_status kv_forward_project(kv_context *ctx, const float *d_vol, float *d_proj, int mode);
/* Backprojection: CudaPerformBackProjection (Projector.cu:376-785) for
 * KV_MODE_R; the exact transpose of the mode-A forward projector for
 * KV_MODE_A. */
kv_status kv_back_project(kv_context *ctx, const float *d_proj, float *d_vol, int mode);

/* One iteration of recons::MLEM (Reconstructors.h:182-242), fused:
 * the ratio relax*(data/FP) (Reconstructors.h:207-213) is the forward
 * projector's epilogue and the multiplicative update (:238-241) the
 * backprojector's.  d_vol is updated in place.  KV_MODE_A runs the standard
 * x <- x/s * A^T(data/Ax) on line-integral data (relax ignored). */
kv_status kv_mlem_iteration(kv_context *ctx, float *d_vol, const float *d_data, float relax, int mode);
/* The two fused halves of kv_mlem_iteration, for callers that interleave
 * work or time the passes separately: kv_mlem_forward writes the ratio
 * (Reconstructors.h:185,207-213) into the context's work buffer,
 * kv_mlem_backward backprojects it and applies the update
 * (Reconstructors.h:225,238-241).  Other calls that use the work buffer
 * (kv_back_project in mode A, kv_log_likelihood, kv_sart_iteration,
 * kv_count_pairs) between the two overwrite the ratio. */
kv_status kv_mlem_forward(kv_context *ctx, const float *d_vol, const float *d_data, float relax, int mode);
kv_status kv_mlem_backward(kv_context *ctx, float *d_vol, int mode);
/* One iteration of recons::ART (Reconstructors.h:65-122), which is a
 * simultaneous (SART-style) update: relax*(data - FP) fused into the forward
 * projector, x += BP(...) into the backprojector.  Mode R only. */
kv_status kv_sart_iteration(kv_context *ctx, float *d_vol, const float *d_data, float relax);

/* The elementwise steps of recons::ART and recons::MLEM on their own, for
 * callers that need the intermediate images (the verbose TIFF dumps,
 * Reconstructors.h:78-81,98-101,195-198,218-221).  kv_correct_projections
 * turns a mode-R forward projection d_proj (kv_forward_project) into the
 * corrected projection in place: relax*(data - FP) for KV_LOOP_ART
 * (Reconstructors.h:87-93), relax*(data/FP) except where FP is zero for
 * KV_LOOP_MLEM (:204-214).  kv_apply_update applies a mode-R backprojection
 * d_bp (kv_back_project) to d_vol: x += b (ART, :113-121) or x *= b (MLEM,
 * :235-241).  With the two projections these give the fused iterations'
 * values bit for bit. */
#define KV_LOOP_ART 0
#define KV_LOOP_MLEM 1
kv_status kv_correct_projections(kv_context *ctx, int loop, float relax, const float *d_data, float *d_proj);
kv_status kv_apply_update(kv_context *ctx, int loop, float *d_vol, const float *d_bp);

/* recons::ShiftAndAdd (Reconstructors.h:263-379).  src_y (host, np floats)
 * are the source y positions, plane_z (host, nplanes floats) the slice
 * heights, sid the source-to-imager distance, fulcrum_z the fulcrum height,
 * voxel_dy the volume y spacing.  d_out is float[nplanes][nv][nu].
 * Returns KV_ERR_PRECONDITION for the reference's error exits
 * (Reconstructors.h:275-285, 301-305, 319-323). */
kv_status kv_shift_and_add(kv_context *ctx, const float *d_proj, const float *src_y, int nplanes,
                           const float *plane_z, float sid, float fulcrum_z, float voxel_dy, float *d_out);

/* kv_shift_and_add without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * The source y positions, plane heights and the row-shift table are staged in
 * a fixed array that the library holds on each GPU until the process ends
 * ((KV_SAA_NOALLOC_MAX_VIEWS + KV_SAA_NOALLOC_MAX_PLANES +
 * KV_SAA_NOALLOC_MAX_SHIFTS) * 4 bytes = 276 KiB); a call with more than
 * KV_SAA_NOALLOC_MAX_VIEWS views, KV_SAA_NOALLOC_MAX_PLANES planes or
 * KV_SAA_NOALLOC_MAX_SHIFTS plane-view pairs returns KV_ERR_INVALID.  The
 * array is shared by every call on a GPU and keeps the last call's contents:
 * calls that overlap on one GPU (from several threads or streams) race.
 * Output bit-identical to kv_shift_and_add; synchronous, like it. */
#define KV_SAA_NOALLOC_MAX_VIEWS 1024
#define KV_SAA_NOALLOC_MAX_PLANES 4096
#define KV_SAA_NOALLOC_MAX_SHIFTS 65536
kv_status kv_shift_and_add_noalloc(kv_context *ctx, const float *d_proj, const float *src_y, int nplanes,
                                   const float *plane_z, float sid, float fulcrum_z, float voxel_dy, float *d_out);

/* Poisson log-likelihood sum over rays with Ax > 0 of data*log(Ax) - Ax,
 * Ax the mode-A forward projection of d_vol; reduced deterministically in
 * FP64.  Synchronous (writes *out on return). */
kv_status kv_log_likelihood(kv_context *ctx, const float *d_data, const float *d_vol, double *out);
/* kv_log_likelihood of n <= KV_LL_MAX volumes (out[i] for d_vols[i], each
 * bit-identical to a kv_log_likelihood call) with one ray walk for all of
 * them: the likelihood gates compare up to three candidates.  Uses two more
 * projection-sized buffers of the context and, for n >= 2, a buffer of 16
 * bytes per voxel into which the volumes are interleaved (all allocated on
 * first use and kept with the context).  Synchronous. */
#define KV_LL_MAX 3
kv_status kv_log_likelihoods(kv_context *ctx, const float *d_data, const float *const *d_vols, int n, double *out);

/* Mode-A sensitivity s = A^T 1 into d_vol. */
kv_status kv_sensitivity(kv_context *ctx, float *d_vol);

/* ---- inspection (host pointers, synchronous) ----------------------------- */
/* Walk sampled rays with the production kernels' walk and record the
 * accepted (voxel index, weight) pairs: rays[i] = p*nu*nv + v*nu + u.
 * idx/w hold nrays*cap entries, counts[i] the number of accepted pairs
 * (may exceed cap; only cap are stored).  Mode R records the segment
 * alpha-weights of Projector.cu:356-366, mode A the alpha-lengths. */
kv_status kv_trace_rays(kv_context *ctx, int mode, int nrays, const int32_t *rays, int cap, int64_t *idx,
                        float *w, int32_t *counts);

/* Number of accepted (ray, voxel) pairs of the forward walk (segments that
 * pass Projector.cu:358-359) and of the backprojector (window pixels that
 * pass Projector.cu:710-725); for mode A the pairs with positive length.
 * These give the algorithmic bytes of a pass.  Synchronous. */
kv_status kv_count_pairs(kv_context *ctx, int mode, int64_t *fp_pairs, int64_t *bp_pairs);

/* ---- batched reconstructions (no reference counterpart) ------------------
 * Many reconstructions of uneven volume and projection-count sizes share one
 * persistent forward launch and one persistent backprojection launch per
 * iteration (fused MLEM in mode R -- the reference's recons::MLEM -- or mode A,
 * x <- x/s * A^T(data/Ax); the per-reconstruction result is bit-identical to
 * kv_mlem_iteration on its own); the iterations can be captured once into a
 * hipGraph and replayed.  The batch owns its device buffers and stream. */
typedef struct kv_batch kv_batch;
kv_status kv_batch_create(int device, int n, const kv_geometry *geoms, int mode, kv_batch **out);
kv_status kv_batch_destroy(kv_batch *b);
/* Host <-> device copies of reconstruction i's volume (initial estimate in,
 * result out) and data (projections in). */
kv_status kv_batch_set_volume(kv_batch *b, int i, const float *h_vol);
kv_status kv_batch_set_data(kv_batch *b, int i, const float *h_data);
kv_status kv_batch_get_volume(kv_batch *b, int i, float *h_vol);
/* iters[i] MLEM iterations for reconstruction i (0 = untouched), relaxation
 * `relax` as recons::MLEM.  use_graph != 0 captures the iteration sequence in
 * a hipGraph (re-captured only when the counts or tile shapes change) and
 * replays it.  *ms (optional) receives the device time of the run.
 * Synchronous. */
kv_status kv_batch_mlem(kv_batch *b, const int32_t *iters, float relax, int use_graph, float *ms);
/* Tile shapes of reconstruction i: forward layout 1, 2, 4 or 8 (detector
 * rows per wavefront), backprojection brick depth 4 or 8 slices.  Results do
 * not depend on them. */
kv_status kv_batch_set_tiles(kv_batch *b, int i, int fp_layout, int bp_slices);
/* Reconstruction i's context (owned by the batch: do not destroy it; its
 * geometry and tables are those of the batch) and device buffers, for
 * per-reconstruction work between batch calls (likelihood, learned step).
 * Work enqueued through the context must be complete (kv_synchronize) before
 * the next kv_batch_* call. */
kv_status kv_batch_context(kv_batch *b, int i, kv_context **ctx);
kv_status kv_batch_buffers(kv_batch *b, int i, float **d_vol, float **d_data);
kv_status kv_batch_get_tiles(kv_batch *b, int i, int *fp_layout, int *bp_slices);
/* Times the tile shapes per size class (volume voxel count < 16 Mi, < 40 Mi,
 * larger) on this batch and applies the fastest; with a cache file, classes
 * already tuned on this device are read from it instead and new results are
 * written back.  Volumes are restored afterwards. */
kv_status kv_batch_autotune(kv_batch *b, const char *cache_path);

/* ---- learned step: 3D residual CNN (no reference counterpart) ------------
 * y = x + f(s*x)/s, f = four 3x3x3 convolutions 1->16->16->16->1 with ReLU
 * after the first three; bfloat16 weights and activations, FP32
 * accumulation on the matrix cores.  Parameters are one FP32 blob of
 * KV_CNN_NPARAMS values: W1[16][1][27] b1[16] W2[16][16][27] b2[16]
 * W3[16][16][27] b3[16] W4[1][16][27] b4[1], W[o][i][t] with tap
 * t = 9(dz+1) + 3(dy+1) + (dx+1).  Volumes float[nz][ny][nx], at most 2^30
 * voxels and 22369621 voxels per slice (KV_ERR_INVALID beyond).  The caller
 * provides a 256-byte aligned device workspace of kv_cnn_workspace_size
 * bytes.  Asynchronous on the context's stream unless stated. */
#define KV_CNN_NPARAMS 14737
kv_status kv_cnn_workspace_size(int nx, int ny, int nz, int training, size_t *bytes);
/* Byte offsets in a training workspace of the stored intermediates
 * a1, a2, a3 (bf16 [V][16]), dz4 (float [V], dL/dz4), dz3, dz2, dz1
 * (bf16 [V][16], gradients after the ReLU masks): offsets[0..6]. */
kv_status kv_cnn_workspace_offsets(int nx, int ny, int nz, size_t *offsets);
/* Inference: d_y = d_x + f(scale*d_x)/scale (d_y must not alias d_x). */
kv_status kv_cnn_forward(kv_context *ctx, const float *d_params, float scale, const float *d_x, float *d_y, int nx,
                         int ny, int nz, void *d_work, size_t work_bytes);
/* Loss L = mean((scale*(y - target))^2) and its gradient dL/dparams into
 * d_grad (KV_CNN_NPARAMS floats).  Deterministic.  Synchronous when loss is
 * non-NULL (the loss is copied back). */
kv_status kv_cnn_gradient(kv_context *ctx, const float *d_params, float scale, const float *d_x,
                          const float *d_target, int nx, int ny, int nz, void *d_work, size_t work_bytes,
                          float *d_grad, double *loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected, step >= 1)
 * on the parameter blob; d_m and d_v are the moment buffers (start at 0). */
kv_status kv_cnn_adam(kv_context *ctx, float *d_params, const float *d_grad, float *d_m, float *d_v, int step,
                      float lr);

/* ---- data path ------------------------------------------------------------
 * PhantomMaker::GetPhantom (source/modules/dataHandling/PhantomMaker.cpp:9-51)
 * on the context's volume grid: every voxel starts at `background`, then
 * adds, in list order, the density of each box (Pave::Contains, Pave.h:51-59:
 * inclusive bounds up to a relative FP32 epsilon) and then of each sphere
 * (Sphere::Contains, Sphere.h:37-40: distance < radius) that contains the
 * voxel's grid point.  Grid points are the voxels' lower corners, sequential
 * FP32 sums of the spacing from the bottom-left-front corner
 * (TomoVolume.cpp:193-229).  World mm, volume-centred frame.  Cylinders are
 * stored but never applied by the reference (PhantomMaker.h:52 vs
 * PhantomMaker.cpp:15-31) and have no counterpart here. */
typedef struct kv_pave {
    float x, y, z;     /* bottom-left-front corner */
    float wx, wy, wz;  /* size */
    float density;
} kv_pave;
typedef struct kv_sphere {
    float cx, cy, cz, radius, density;
} kv_sphere;
kv_status kv_phantom(kv_context *ctx, float background, const kv_pave *paves, int npaves, const kv_sphere *spheres,
                     int nspheres, float *d_vol);
/* vtkImageLogarithmicScale with constant c, as the reference's caller turns
 * raw projection images into data (mainReconstruct.cpp:120-124, c = 1):
 * out = c*log(1 + x) for x > 0, -c*log(1 - x) otherwise, evaluated in FP64
 * and stored as FP32.  n values; d_out may be d_in. */
kv_status kv_log_scale(kv_context *ctx, const float *d_in, float *d_out, int64_t n, double c);

/* Seeded transmission data (no counterpart in the reference, which reads
 * measured projections): for each of the n rays, counts
 * I ~ Poisson(i0 * exp(-d_ell[i])) and d_out[i] = -ln(clamp(I, 1, i0) / i0)
 * (FP64, stored as FP32); d_ell holds mode-A line integrals in mm.  Draws
 * come from Philox4x64-10 (counter (i, block, 0, 0), key (seed, 0x6b76),
 * four uniforms (w >> 11) * 2^-53 per block); means below 10 are sampled by
 * inversion, larger ones by Hoermann's PTRS.  exp, log and log k! are
 * evaluated with IEEE +, -, *, / only, so results do not depend on a math
 * library.  d_out may be d_ell.  Asynchronous on the context's stream. */
kv_status kv_poisson_data(kv_context *ctx, const float *d_ell, int64_t n, float i0, uint64_t seed, float *d_out);

/* ---- trajectory extrapolation of mode-A MLEM ------------------------------
 * No counterpart in the reference: these serve the accelerated
 * reconstruction built on the mode-A MLEM update (a jump from iteration k to
 * an estimate of a later iterate).  d_iterates is a host array of h+1 device
 * pointers to consecutive MLEM-A iterates x_0..x_h of the context's geometry,
 * oldest first, 1 <= h <= KV_TRAJ_MAX.  The log-ratios are
 * d_i = log(x_{i+1}/x_i), i = 0..h-1 (0 where either iterate is not positive).
 *
 * kv_traj_gram: gram[i*h + j] = sum_v w(v) d_i(v) d_j(v), w = sens*x_h with
 * sens = A^T 1 (the context's mode-A sensitivity; 0 where sens or x_h is not
 * positive).  FP64 accumulation, deterministic; synchronous.
 * kv_traj_combine: d_out = x_h * exp(sum_i coef[i] d_i) (coef: h host
 * doubles; x_h kept where it is not positive).  d_out may be one of the
 * iterates.  Asynchronous on the context's stream. */
#define KV_TRAJ_MAX 32
kv_status kv_traj_gram(kv_context *ctx, const float *const *d_iterates, int h, double *gram);
kv_status kv_traj_combine(kv_context *ctx, const float *const *d_iterates, int h, const double *coef, float *d_out);

/* Rayleigh-Ritz coefficients of a trajectory jump (host arithmetic, FP64):
 * from the h x h Gram matrix of h log-ratios (kv_traj_gram) and npairs pairs
 * (pairs[2k], pairs[2k+1]) = (i, j) with d_j = J d_i (consecutive log-ratios
 * of one stretch of MLEM iterations; several stretches may contribute),
 * coef[0..h-1] such that sum_i coef[i] d_i ~ p(J) J d_target with
 * p(t) = sum_{k=1..horizon} t^k.  Gram directions of the pairs' first
 * members below rel_tol times the largest eigenvalue are dropped; the Ritz
 * values are clamped to [0, ritz_max] and returned in ritz (*nritz of them,
 * at most npairs; either may be NULL).  Symmetric eigenproblems by cyclic
 * Jacobi rotations.  No device work. */
kv_status kv_ritz_coefficients(const double *gram, int h, const int32_t *pairs, int npairs, int target, int horizon,
                               double rel_tol, double ritz_max, double *coef, double *ritz, int *nritz);

/* ---- accelerated mode-A MLEM ----------------------------------------------
 * No counterpart in the reference; the sibling of recons::MLEM
 * (Reconstructors.h:136-253) in its mode-A form on line-integral data.
 * Two-stretch jump schedule: MLEM-A runs `first` iterations, recording the
 * iterates from `start` on (stretch A); a probe jumps to an estimate of
 * iteration `mid` and runs `second` MLEM iterations from there (stretch B);
 * the final jump starts again from iteration `first` with the Ritz basis of
 * both stretches and lands on an estimate of iteration `horizon`.  The jump
 * replaces the estimate only if its Poisson log-likelihood is not lower than
 * iteration `first`'s; with CNN parameters (d_params, KV_CNN_NPARAMS floats
 * on the device, input scale `scale`), the CNN step on the jump's result is
 * kept only if it does not lower the jump's log-likelihood.  *kept (optional)
 * receives the steps kept (0: iteration `first`, 1: the jump, 2: jump and CNN
 * step).  Cost: first + second MLEM iterations, two Gram passes, two
 * combinations, one likelihood walk.  Defaults: {2, 10, 40, 8, 100}.
 * Allocates (first - start + second + 4) volumes of device memory for the
 * call.  Synchronous. */
typedef struct kv_accel_schedule {
    int32_t start, first, mid, second, horizon;
} kv_accel_schedule;
/* d_vol: initial volume in (recons::MLEM's default is uniform 1), result out;
 * d_data: line integrals (float[np][nv][nu]). */
kv_status kv_mlem_accelerated(kv_context *ctx, float *d_vol, const float *d_data, const kv_accel_schedule *schedule,
                              const float *d_params, float scale, int *kept);
/* The same schedule on every reconstruction of a mode-A batch: the MLEM
 * iterations of all reconstructions run as the batch's persistent launches;
 * each reconstruction's jumps and gates run through its batch context.
 * Volumes start from the batch's current contents; kept[i] (optional) per
 * reconstruction.  Each result is bit-identical to kv_mlem_accelerated on
 * that reconstruction alone.  Synchronous. */
kv_status kv_batch_mlem_accelerated(kv_batch *b, const kv_accel_schedule *schedule, const float *d_params,
                                    float scale, int32_t *kept);

/* ---- host-buffer variants (synchronous; used by the C++ Projector) ------- */
kv_status kv_forward_project_host(kv_context *ctx, const float *h_vol, float *h_proj, int mode);
kv_status kv_back_project_host(kv_context *ctx, const float *h_proj, float *h_vol, int mode);

#ifdef __cplusplus
}
#endif
#endif
