// Internal declarations shared by the HIP sources of libkevernals_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "kevernals_amd.h"

// Geometry as the kernels see it: passed by value as a kernel argument.
// Per-view arrays live in device memory (det/src: np*3 floats each).
struct KvGeom {
    int nx, ny, nz;
    int nu, nv, np;
    float sx, sy, sz;
    float pu, pv;
    const float *det;
    const float *src;
    // Ray constants in FP32, evaluated by kv_set_geometry as the forward
    // kernels would: rc[KV_RC*p + 0..6] = source x, y, z over the spacing,
    // dz (detector z / sz - source z / sz), the z-face alphas
    // (-nz/2 - sz_p) / dz and (nz/2 - sz_p) / dz and 1 / dz (the last three
    // 0 when |dz| <= 1e-6); for the backprojectors rc[KV_RC*p + 8, 9] =
    // detector corner x over sx and the far detector edge
    // (det x + nu * pu) / sx (Projector.cu:473-492); then per view and
    // detector column, at rc + KV_RC*np, four floats: the pixel-centre
    // coordinate qx = (det x + (u + 0.5) pu) / sx, the x-face alphas
    // (-nx/2 - sx_p) / dx and (nx/2 - sx_p) / dx (0 unless |dx| > 1e-6,
    // dx = qx - sx_p) and 1 / dx (0 when dx = 0); then the same per view
    // and detector row for y.
    const float *rc;
    int walk_fast;  // forward walk may use end thresholds and 24-bit indices (see kv_fp.hip)
    int bp_posz;    // every view's z planes at alpha >= 1e-6: clamped backprojection weights (see kv_bp.hip)
};

constexpr int KV_RC = 16;

// Unsigned division by a divisor fixed for a launch: n / d =
// (umulhi(n, m) + n) >> l for every n < 2^31, with l = ceil(log2 d) and
// m = floor(2^32 (2^l - d) / d) + 1 (the round-up method of Granlund and
// Montgomery, "Division by invariant integers using multiplication", 1994;
// the 33-bit sum cannot overflow for n < 2^31).
struct KvUDiv {
    unsigned m, l;
};
inline KvUDiv kv_udiv(unsigned d)
{
    if (d < 1)
        d = 1;  // no divisor of a valid launch is 0; keep the host arithmetic defined
    unsigned l = 0;
    while (l < 31 && (1ull << l) < d)
        ++l;
    const unsigned long long m = ((1ull << 32) * ((1ull << l) - d)) / d + 1;
    return KvUDiv{(unsigned)m, l};
}

// Mapping of a one-dimensional forward launch onto tiles (kv_fp.hip
// fp_xcd_tile): tile layout, stripe width in tiles, stripes per tile row and
// tile rows per view with their divisors, first view of the launch.
struct KvFpMap {
    int layout, chunk, nch, gy, p0;
    KvUDiv dchunk, dnch, dgy;
};

// Per-view plane-crossing tables.  For mode R an entry is the reference's
// alpha1/alpha2 of Projector.cu:653-669 for plane k of a pixel's ray:
//   x[p][k][u] = (|rx| > 1e-24) ? ((k - nx/2) - sx) / rx : 0
// For mode A it is the mode-A plane crossing ((k - n/2) - s) * (1/d), with
// +-INF encoding of rays parallel to the planes, plus the squared mm
// components of each ray (lx2[p][u], ly2[p][v], lz2[p]).
struct KvTables {
    float *x = nullptr;   // [np][nx+1][nu]
    float *y = nullptr;   // [np][ny+1][nv], then 16 bytes of padding (read, never used)
    float *z = nullptr;   // [np][nz+1]
    float *lx2 = nullptr; // mode A: [np][nu]
    float *ly2 = nullptr; // mode A: [np][nv]
    float *lz2 = nullptr; // mode A: [np]
    // Per view and z-plane l = 0..nz, KV_ZS floats (both modes):
    //  [0] T  = (qz - sz) / ((l - nz/2) - sz), the lateral scale of plane l's
    //          shadow on the detector (Projector.cu:473-480 form)
    //  [1] AZ = the same ratio at the slice centre l + 0.5 (0 if singular)
    //  [2] Kx1, [3] Kx0: detector column (pixel-centre units) of lateral x
    //          coordinate c on plane l, u ~= Kx0 + c * Kx1 (bounds only)
    //  [4] Ky1, [5] Ky0: the same for detector rows
    float *zs = nullptr;
    // Per view, y row j and z slice l (index (p * ny + j) * nz + l: the
    // slices of one row are adjacent, as a brick reads them), KV_YZ
    // words, the backprojectors' terms shared by the voxels of one (y, z)
    // row: [0] first and [1] last detector row of the voxel box's shadow
    // (intersected with the +-5 window around the voxel centre's row in mode
    // R; empty, first > last, when mode R's centre row lies off the widened
    // detector), [2] 1 if the shadow is defined (both planes in front of the
    // source), as ints; [4] lz = max(0, min z-plane alpha), [5] uz =
    // min(1, max z-plane alpha), [6] lz with a +0 floor in bit order, as
    // floats.
    float *yz = nullptr;
};
constexpr int KV_ZS = 8;
constexpr int KV_YZ = 8;

// Epilogues of the projection kernels: plain result, fused MLEM ratio/update,
// fused SART residual/update, or the count of accepted (ray, voxel) pairs.
enum KvEpilogue { KV_EP_PLAIN = 0, KV_EP_MLEM = 1, KV_EP_SART = 2, KV_EP_COUNT = 3 };

// One reconstruction of a batch, as the persistent kernels read it (device
// memory, wave-uniform loads).  Work items are numbered consecutively over
// the batch.  Forward tiles sit in KV_XCDS lists, one per XCD (stripes of
// fp_chunk tile columns, stripe i in list i % KV_XCDS, as fp_xcd_tile): this
// reconstruction holds items [fp_first, fp_first + views*fp_ty*fp_tx/KV_XCDS)
// of every list, fp_tx being padded to whole stripes for all lists.
// Backprojection bricks: [bp_first, bp_first + bp_bx*bp_by*bp_bz).
struct KvProblem {
    KvGeom g;
    KvTables t;
    float *vol;
    const float *data;
    float *ratio;
    const float *sens;         // mode A: sensitivity A^T 1
    int iters;                 // MLEM iterations for this reconstruction
    int fp_layout;             // forward tile layout: 1, 2, 4 or 8 rows per wavefront
    int fp_tx, fp_ty;          // forward tiles per view (fp_tx padded to a multiple of KV_XCDS * fp_chunk)
    int fp_chunk;              // tile columns per XCD stripe
    int fp_nch;                // stripes per tile row in one list: fp_tx / (KV_XCDS * fp_chunk)
    KvUDiv fp_dchunk, fp_dnch, fp_dty;  // divisors fp_chunk, fp_nch, fp_ty
    long long fp_first;
    int bp_zt;                 // backprojection brick depth: 4 or 8 slices
    int bp_bx, bp_by, bp_bz;   // bricks per axis
    long long bp_first;
};

struct kv_context {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    bool has_geom = false;
    KvGeom g{};
    float *d_det = nullptr;
    float *d_src = nullptr;
    float *d_rc = nullptr;     // KvGeom::rc
    KvTables tab_r{};
    bool tab_r_ok = false;
    KvTables tab_a{};
    bool tab_a_ok = false;
    float *d_ratio = nullptr;  // projection-sized work buffer of fused iterations
    float *d_sens = nullptr;   // mode-A sensitivity A^T 1
    bool sens_ok = false;
    double *d_red = nullptr;   // partial sums of deterministic reductions
    size_t red_cap = 0;        // doubles in d_red
    float *d_ll = nullptr;     // projections of the further volumes of kv_log_likelihoods
    float *d_pack = nullptr;   // kv_log_likelihoods' volumes interleaved per voxel (4 floats per voxel)
};

// Reconstruction count and operator mode of a batch (0 / -1 for null).
int kv_internal_batch_size(const kv_batch *b);
int kv_internal_batch_mode(const kv_batch *b);

// Thread-local error message.
void kv_set_error(const char *fmt, ...);

// Builds the crossing tables of `mode` for the context's geometry (once).
kv_status kv_internal_ensure_tables(kv_context *ctx, int mode);
// Computes the mode-A sensitivity A^T 1 (once).
kv_status kv_internal_ensure_sens(kv_context *ctx);

#define KV_TRY_HIP(expr)                                                                        \
    do {                                                                                        \
        hipError_t kv_e_ = (expr);                                                              \
        if (kv_e_ != hipSuccess) {                                                              \
            kv_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(kv_e_), __FILE__, \
                         __LINE__);                                                             \
            return KV_ERR_DEVICE;                                                               \
        }                                                                                       \
    } while (0)

// a * b with 0 * anything = +-0, also for INF and NaN (v_mul_legacy_f32); the
// IEEE product when neither operand is zero.  A sum s + kv_mul_legacy(0, x)
// is s bit for bit unless s is -0.
__device__ __forceinline__ float kv_mul_legacy(float a, float b)
{
    float r;
    asm("v_mul_legacy_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// AlmostEqualRelative(a, 0.F) of commons/Maths.h:25-34: the projection
// values MLEM leaves unchanged (Reconstructors.h:207-213).
__device__ __forceinline__ bool almost_equal_zero(float a)
{
    const float diff = fabsf(a - 0.f);
    const float A = fabsf(a), B = 0.f;
    const float m = (A < B) ? B : A;  // std::max
    return diff <= __FLT_EPSILON__ * m;
}

inline int64_t kv_nvox(const KvGeom &g) { return (int64_t)g.nx * g.ny * g.nz; }
inline int64_t kv_nray(const KvGeom &g) { return (int64_t)g.np * g.nv * g.nu; }

// ---- launchers (kv_fp.hip, kv_bp.hip, kv_misc.hip) ----------------------
// Forward tiling (fp_tile_pixel: rows per wavefront; fp_xcd_tile and the
// batch kernels' per-XCD lists: detector columns per XCD stripe).  Measured
// at 1024x1024x64 with a 2048^2 detector and 25 views: 8x8 rays per
// wavefront in 32-column stripes, mode R forward 10.9 -> 8.8 ms and mode A
// 10.0 -> 8.3 ms against 64x1 rays per wavefront round-robin.  The batch
// kernels' default layouts: mode R 8, mode A 4 (eight full-size mode-A
// reconstructions, 100 batched iterations: 9.56 s against 10.45 s in 8x8);
// the autotuner may choose others.
// The multi-volume likelihood walk takes its layout by volume count: one
// gather per step, of one volume or of a record holding two or three
// (interleaved first by pack_kernel).  Measured at the size above (walk,
// reductions and packing per call): 6.9 / 7.7 / 10.1 ms for one / two /
// three volumes in 8x8 / 16x4 / 16x4 rays per wavefront; before the current
// mode-A walk, two and three volumes took 9.4 / 11.7 ms in these layouts,
// 10.3 / 13.4 ms in 8x8, three 13.6 ms in 32x2, and one gather per volume
// and step 12.5 / 21.5 ms.
constexpr int KV_XCDS = 8;
constexpr int KV_FP_LAYOUT = 8;
constexpr int KV_FP_LAYOUT_MULTI[KV_LL_MAX] = {8, 4, 4};
constexpr int KV_FP_STRIPE = 32;
constexpr int KV_FP_BATCH_LAYOUT_R = 8;
constexpr int KV_FP_BATCH_LAYOUT_A = 4;

hipError_t kv_launch_fp_r(const KvGeom &g, const float *vol, float *out, int epilogue, const float *data,
                          float relax, hipStream_t s);
// Mode-A forward pass.  KV_EP_MLEM writes the MLEM ratio data/Ax
// premultiplied by the ray's length factor from the mode-A tables t (the
// input the mode-A backprojector expects); the other epilogues ignore t.
// Mode-A forward projections of n <= KV_LL_MAX volumes in one walk per ray.
// For n >= 2 the volumes are first interleaved into pack (4 floats per
// voxel), so that one gather per step reads every volume.
hipError_t kv_launch_fp_a_multi(const KvGeom &g, int n, const float *const *vols, float *pack, float *const *outs,
                                hipStream_t s);
hipError_t kv_launch_fp_a(const KvGeom &g, const KvTables &t, const float *vol, float *out, int epilogue,
                          const float *data, hipStream_t s);
hipError_t kv_launch_trace(const KvGeom &g, int mode, int nrays, const int32_t *rays, int cap, int64_t *idx,
                           float *w, int32_t *counts, hipStream_t s);
hipError_t kv_launch_tables_r(const KvGeom &g, const KvTables &t, hipStream_t s);
hipError_t kv_launch_tables_a(const KvGeom &g, const KvTables &t, hipStream_t s);
hipError_t kv_launch_bp_r(const KvGeom &g, const KvTables &t, const float *proj, float *vol, int epilogue,
                          hipStream_t s);
// Mode-A backprojection of projections already multiplied by each ray's
// length factor (kv_launch_scale_a, or the forward pass's MLEM epilogue);
// unit_input backprojects the length factors themselves (sensitivity).
hipError_t kv_launch_bp_a(const KvGeom &g, const KvTables &t, const float *proj, float *vol, int epilogue,
                          const float *sens, bool unit_input, hipStream_t s);
// out = proj * length factor of each ray (mode-A tables), np*nv*nu values.
hipError_t kv_launch_scale_a(const KvGeom &g, const KvTables &t, const float *proj, float *out, hipStream_t s);
// Batched persistent passes (MLEM epilogues, mode R or A).  `grid` workgroups
// of 256 threads claim work items: the forward pass from counters[0 ..
// KV_XCDS-1] (one per XCD list of `total` items), the backprojection from
// *counter; all zero before the launch.
hipError_t kv_launch_fp_batch(int mode, const KvProblem *probs, int nprob, long long total,
                              unsigned long long *counters, int iteration, float relax, int grid, hipStream_t s);
hipError_t kv_launch_bp_batch(int mode, const KvProblem *probs, int nprob, long long total,
                              unsigned long long *counter, int iteration, int grid, hipStream_t s);
hipError_t kv_launch_zero_u64(unsigned long long *p, int n, hipStream_t s);
int kv_fp_batch_blocks_per_cu(int mode);
int kv_bp_batch_blocks_per_cu(int mode);
// Per-device address of the no-allocation shift-and-add staging array
// (kv_misc.hip): _init resolves it once per device, with that device current;
// kv_internal_scratch returns it (nullptr before _init).  Word layout:
// [src_y: KV_SAA_NOALLOC_MAX_VIEWS][plane_z: KV_SAA_NOALLOC_MAX_PLANES]
// [shifts: KV_SAA_NOALLOC_MAX_SHIFTS].
hipError_t kv_internal_scratch_init(int device);
int32_t *kv_internal_scratch(int device);
hipError_t kv_launch_saa_shifts(int np, int nplanes, float sid, float mF, float voxel_dy, const float *src_y,
                                const float *plane_z, int32_t *shifts, hipStream_t s);
// Planes per chunk of the shift-and-add kernel for these shifts (shift(k, p)
// = row shift of plane k, view p): the largest of 8, 4, 2 whose staged rows
// fit the kernel's LDS, else 1.
int kv_saa_chunk(int nplanes, int np, const std::function<int(int, int)> &shift);
hipError_t kv_launch_saa(int nu, int nv, int np, const float *proj, const int32_t *shifts, int nplanes, int kc,
                         float *out, hipStream_t s);
hipError_t kv_launch_sum(int64_t n, const float *x, double *partials, int nblocks, hipStream_t s);
hipError_t kv_launch_loglik(const KvGeom &g, const float *data, const float *ax, double *partials, int nblocks,
                            hipStream_t s);

// ---- data path (kv_misc.hip) --------------------------------------------
// px/py/pz: the voxels' grid points per axis (device); paves/spheres: device
// copies of the shape lists.
hipError_t kv_launch_phantom(int nx, int ny, int nz, const float *px, const float *py, const float *pz,
                             float background, const kv_pave *paves, int npaves, const kv_sphere *spheres,
                             int nspheres, float *out, hipStream_t s);
hipError_t kv_launch_correct(int64_t n, int loop, float relax, const float *data, float *proj, hipStream_t s);
hipError_t kv_launch_update(int64_t n, int loop, float *vol, const float *bp, hipStream_t s);
hipError_t kv_launch_log_scale(int64_t n, const float *in, float *out, double c, hipStream_t s);

// ---- trajectory extrapolation (kv_traj.hip) ------------------------------
struct KvTraj {
    const float *x[KV_TRAJ_MAX + 1];  // iterates, oldest first
};
struct KvTrajCoef {
    double c[KV_TRAJ_MAX];
};
// Workgroups of the Gram pass (fixed for a volume size) and the number of
// 4x4 blocks of the upper triangle it writes per workgroup (16 doubles each).
int kv_traj_gram_blocks(int64_t n);
int kv_traj_pairs(int h);
hipError_t kv_launch_traj_gram(const KvTraj &tr, int h, int64_t n, const float *sens, double *partials,
                               hipStream_t s);
hipError_t kv_launch_traj_combine(const KvTraj &tr, const KvTrajCoef &cf, int h, int64_t n, float *out,
                                  hipStream_t s);
