// C ABI of libkevernals_amd: context, geometry marshalling and dispatch.
//
// Replaces the host half of KevLoq/Kevernals Projector.cu (:873-1207), which
// allocated, copied, launched, synchronised and freed on every call and
// exited the process on CUDA errors (CudaErrorHandler.h).  Here a context
// keeps the geometry, its crossing tables and work buffers resident, work is
// enqueued on one stream, and every failure comes back as a kv_status with a
// message (kv_last_error).
#pragma clang fp contract(off)

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "kv_internal.h"

namespace {
thread_local std::string g_last_error;

template <class T>
void free_dev(T *&p)
{
    if (p)
        (void)hipFree(p);
    p = nullptr;
}

void free_tables(KvTables &t)
{
    free_dev(t.x);
    free_dev(t.y);
    free_dev(t.z);
    free_dev(t.lx2);
    free_dev(t.ly2);
    free_dev(t.lz2);
    free_dev(t.zs);
    free_dev(t.yz);
}

kv_status dev_alloc(void **p, size_t bytes)
{
    const hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        kv_set_error("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? KV_ERR_NOMEM : KV_ERR_DEVICE;
    }
    return KV_OK;
}

// Grows the context's reduction buffer to at least n doubles.  The capacity
// is cleared before the reallocation, so a failed allocation leaves a null
// buffer of capacity 0 (the next call allocates again) rather than a null
// buffer that a later, smaller request would launch into.
kv_status ensure_red(kv_context *ctx, size_t n)
{
    if (ctx->red_cap >= n)
        return KV_OK;
    free_dev(ctx->d_red);
    ctx->red_cap = 0;
    if (kv_status st = dev_alloc((void **)&ctx->d_red, n * sizeof(double)))
        return st;
    ctx->red_cap = n;
    return KV_OK;
}

kv_status check_ctx(kv_context *ctx, bool need_geom)
{
    if (!ctx) {
        kv_set_error("null context");
        return KV_ERR_INVALID;
    }
    if (need_geom && !ctx->has_geom) {
        kv_set_error("no geometry set on this context (call kv_set_geometry first)");
        return KV_ERR_STATE;
    }
    const hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) {
        kv_set_error("hipSetDevice(%d) failed: %s", ctx->device, hipGetErrorString(e));
        return KV_ERR_DEVICE;
    }
    return KV_OK;
}

// Whether the forward walk may replace its run counters by end thresholds and
// use 24-bit index multiplies: nx*ny < 2^24, and per axis
// (n + 2) * (max|source| + n/2 + 2) < 2^22 in voxel units, which keeps the
// FP32 accumulation error of a run below a quarter of its increment.
bool walk_fast_ok(const kv_geometry &q)
{
    if ((long long)q.nx * q.ny >= (1LL << 24))
        return false;
    const int n[3] = {q.nx, q.ny, q.nz};
    const double sp[3] = {q.sx, q.sy, q.sz};
    for (int a = 0; a < 3; ++a) {
        double smax = 0.0;
        for (int p = 0; p < q.np; ++p)
            smax = std::max(smax, std::fabs((double)q.src[3 * p + a] / sp[a]));
        if (((double)n[a] + 2.0) * (smax + (double)n[a] / 2.0 + 2.0) >= 4194304.0)
            return false;
    }
    return true;
}

// Every z plane of the volume crosses every ray of every view at alpha >=
// 1e-5 (evaluated in FP64 from the plane crossings (k - nz/2 - s) / d of the
// two outer planes; the others lie between them), well above the 1e-6 the
// backprojector's clamped weights need (kv_bp.hip, FW) whatever the FP32
// rounding of the tables.
bool bp_posz_ok(const kv_geometry &q)
{
    for (int p = 0; p < q.np; ++p) {
        const double s = (double)(q.src[3 * p + 2] / q.sz), dz = (double)(q.det[3 * p + 2] / q.sz) - s;
        if (!(std::fabs(dz) > 1e-20))
            return false;
        for (int k = 0; k <= q.nz; k += q.nz) {
            const double a = (((double)k - (double)q.nz / 2.0) - s) / dz;
            if (!(a >= 1e-5))
                return false;
        }
    }
    return true;
}

// KvGeom::rc on the host: the same FP32 operations, in the same order, as
// the per-ray set-up of the forward kernels (Projector.cu:303-319 and the
// z terms of :54-61 and :186), so every value is bit-identical.
std::vector<float> ray_constants(const kv_geometry &q)
{
    std::vector<float> rc((size_t)q.np * (KV_RC + 4 * q.nu + 4 * q.nv), 0.f);
    float *qx = rc.data() + (size_t)q.np * KV_RC, *qy = qx + (size_t)q.np * q.nu * 4;
    const float hx = (float)q.nx / 2.f, hy = (float)q.ny / 2.f, hz = (float)q.nz / 2.f;
    // per detector column (row): centre coordinate, the reference's face
    // alphas of the axis (Projector.cu:54-61, 0 unless |d| > 1e-6) and 1/d
    // (0 when d = 0; the mode-R walk uses it only when |d| > 1e-6)
    auto axis = [](float *o, float q, float s, float h) {
        const float d = q - s;
        const bool ok = std::fabs(d) > 1e-6f;
        o[0] = q;
        o[1] = ok ? (-h - s) / d : 0.f;
        o[2] = ok ? (h - s) / d : 0.f;
        o[3] = d != 0.f ? 1.f / d : 0.f;
    };
    for (int p = 0; p < q.np; ++p) {
        const float *S = q.src + 3 * p, *D = q.det + 3 * p;
        const float sz = S[2] / q.sz, dz = D[2] / q.sz - sz;
        const bool okz = std::fabs(dz) > 1e-6f;
        float *c = rc.data() + (size_t)KV_RC * p;
        c[0] = S[0] / q.sx;
        c[1] = S[1] / q.sy;
        c[2] = sz;
        c[3] = dz;
        c[4] = okz ? (-hz - sz) / dz : 0.f;
        c[5] = okz ? (hz - sz) / dz : 0.f;
        c[6] = okz ? 1.f / dz : 0.f;
        c[8] = D[0] / q.sx;
        c[9] = (D[0] + (float)q.nu * q.pu) / q.sx;
        for (int u = 0; u < q.nu; ++u)
            axis(qx + ((size_t)p * q.nu + u) * 4, (D[0] + ((float)u + 0.5f) * q.pu) / q.sx, c[0], hx);
        for (int v = 0; v < q.nv; ++v)
            axis(qy + ((size_t)p * q.nv + v) * 4, (D[1] + ((float)v + 0.5f) * q.pv) / q.sy, c[1], hy);
    }
    return rc;
}

kv_status check_mode(int mode)
{
    if (mode != KV_MODE_R && mode != KV_MODE_A) {
        kv_set_error("unknown operator mode %d", mode);
        return KV_ERR_INVALID;
    }
    return KV_OK;
}

kv_status ensure_tables(kv_context *ctx, int mode)
{
    const KvGeom &g = ctx->g;
    const size_t nxe = (size_t)g.np * g.nu * (g.nx + 1), nye = (size_t)g.np * g.nv * (g.ny + 1),
                 nze = (size_t)g.np * (g.nz + 1), nyz = (size_t)g.np * g.nz * g.ny * KV_YZ;
    if (mode == KV_MODE_R) {
        if (ctx->tab_r_ok)
            return KV_OK;
        KvTables &t = ctx->tab_r;
        kv_status st;
        if ((st = dev_alloc((void **)&t.x, nxe * 4)) || (st = dev_alloc((void **)&t.y, nye * 4 + 16)) ||
            (st = dev_alloc((void **)&t.z, nze * 4)) || (st = dev_alloc((void **)&t.zs, nze * KV_ZS * 4)) ||
            (st = dev_alloc((void **)&t.yz, nyz * 4))) {
            free_tables(t);
            return st;
        }
        KV_TRY_HIP(kv_launch_tables_r(g, t, ctx->stream));
        ctx->tab_r_ok = true;
        return KV_OK;
    }
    if (ctx->tab_a_ok)
        return KV_OK;
    KvTables &t = ctx->tab_a;
    kv_status st;
    if ((st = dev_alloc((void **)&t.x, nxe * 4)) || (st = dev_alloc((void **)&t.y, nye * 4 + 16)) ||
        (st = dev_alloc((void **)&t.z, nze * 4)) || (st = dev_alloc((void **)&t.lx2, (size_t)g.np * g.nu * 4)) ||
        (st = dev_alloc((void **)&t.ly2, (size_t)g.np * g.nv * 4)) ||
        (st = dev_alloc((void **)&t.lz2, (size_t)g.np * 4)) || (st = dev_alloc((void **)&t.zs, nze * KV_ZS * 4)) ||
        (st = dev_alloc((void **)&t.yz, nyz * 4))) {
        free_tables(t);
        return st;
    }
    KV_TRY_HIP(kv_launch_tables_a(g, t, ctx->stream));
    ctx->tab_a_ok = true;
    return KV_OK;
}

kv_status ensure_sens(kv_context *ctx)
{
    if (ctx->sens_ok)
        return KV_OK;
    kv_status st = ensure_tables(ctx, KV_MODE_A);
    if (st)
        return st;
    if (!ctx->d_sens && (st = dev_alloc((void **)&ctx->d_sens, (size_t)kv_nvox(ctx->g) * 4)))
        return st;
    KV_TRY_HIP(kv_launch_bp_a(ctx->g, ctx->tab_a, nullptr, ctx->d_sens, KV_EP_PLAIN, nullptr, true, ctx->stream));
    ctx->sens_ok = true;
    return KV_OK;
}

void release_geometry(kv_context *ctx)
{
    free_dev(ctx->d_det);
    free_dev(ctx->d_src);
    free_dev(ctx->d_rc);
    free_tables(ctx->tab_r);
    free_tables(ctx->tab_a);
    free_dev(ctx->d_ratio);
    free_dev(ctx->d_sens);
    free_dev(ctx->d_ll);
    free_dev(ctx->d_pack);
    ctx->tab_r_ok = ctx->tab_a_ok = ctx->sens_ok = false;
    ctx->has_geom = false;
}

bool finite_pos(float v) { return std::isfinite(v) && v > 0.f; }

// AlmostEqualRelative of commons/Maths.h:25-34 (FP32, epsilon default).
bool almost_equal_relative(float a, float b)
{
    const float diff = std::fabs(a - b);
    a = std::fabs(a);
    b = std::fabs(b);
    const float m = (a < b) ? b : a;
    return diff <= FLT_EPSILON * m;
}
}  // namespace

kv_status kv_internal_ensure_tables(kv_context *ctx, int mode) { return ensure_tables(ctx, mode); }
kv_status kv_internal_ensure_sens(kv_context *ctx) { return ensure_sens(ctx); }

void kv_set_error(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

extern "C" {

const char *kv_last_error(void) { return g_last_error.c_str(); }

const char *kv_version(void) { return "kevernals_amd 0.1 (gfx950)"; }

kv_status kv_create(int device, kv_context **out)
{
    if (!out) {
        kv_set_error("null output pointer");
        return KV_ERR_INVALID;
    }
    *out = nullptr;
    int count = 0;
    KV_TRY_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) {
        kv_set_error("device %d out of range (%d devices)", device, count);
        return KV_ERR_INVALID;
    }
    KV_TRY_HIP(hipSetDevice(device));
    KV_TRY_HIP(kv_internal_scratch_init(device));
    auto *ctx = new kv_context();
    ctx->device = device;
    const hipError_t e = hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        kv_set_error("hipStreamCreate failed: %s", hipGetErrorString(e));
        delete ctx;
        return KV_ERR_DEVICE;
    }
    ctx->stream = ctx->own_stream;
    *out = ctx;
    return KV_OK;
}

kv_status kv_destroy(kv_context *ctx)
{
    if (!ctx)
        return KV_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    release_geometry(ctx);
    free_dev(ctx->d_red);
    if (ctx->own_stream)
        (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return KV_OK;
}

kv_status kv_set_stream(kv_context *ctx, void *hip_stream)
{
    if (kv_status st = check_ctx(ctx, false))
        return st;
    ctx->stream = (hipStream_t)hip_stream;  // NULL selects the null stream
    return KV_OK;
}

void *kv_get_stream(kv_context *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

kv_status kv_synchronize(kv_context *ctx)
{
    if (kv_status st = check_ctx(ctx, false))
        return st;
    KV_TRY_HIP(hipStreamSynchronize(ctx->stream));
    return KV_OK;
}

kv_status kv_set_geometry(kv_context *ctx, const kv_geometry *geom)
{
    if (kv_status st = check_ctx(ctx, false))
        return st;
    if (!geom || !geom->det || !geom->src) {
        kv_set_error("null geometry or per-view arrays");
        return KV_ERR_INVALID;
    }
    const kv_geometry &q = *geom;
    if (q.nx <= 0 || q.ny <= 0 || q.nz <= 0 || q.nu <= 0 || q.nv <= 0 || q.np <= 0) {
        kv_set_error("non-positive size (volume %d x %d x %d, detector %d x %d x %d)", q.nx, q.ny, q.nz, q.nu, q.nv,
                     q.np);
        return KV_ERR_INVALID;
    }
    if (!finite_pos(q.sx) || !finite_pos(q.sy) || !finite_pos(q.sz) || !finite_pos(q.pu) || !finite_pos(q.pv)) {
        kv_set_error("voxel spacing and pixel pitch must be finite and positive");
        return KV_ERR_INVALID;
    }
    // voxel gathers use 32-bit byte offsets (at most 2^30 voxels); ray ids
    // are 31-bit; the backprojector addresses one view's plane (two rows of
    // slack included) and one view's x and y crossing tables with 31-bit byte
    // offsets
    const int64_t nvox = (int64_t)q.nx * q.ny * q.nz, nray = (int64_t)q.np * q.nu * q.nv;
    const int64_t plane_bytes = std::max({(int64_t)q.nu * ((int64_t)q.nv + 2) * 4 + 16,
                                          ((int64_t)q.nx + 1) * q.nu * 4, ((int64_t)q.ny + 1) * q.nv * 4});
    if (nvox > (1LL << 30) || nray > 0x7FFFFFFFLL || q.nx > (1 << 24) || q.ny > (1 << 24) || q.nz > (1 << 24) ||
        plane_bytes > 0x7FFFFFFFLL) {
        kv_set_error("problem too large: %lld voxels, %lld rays", (long long)nvox, (long long)nray);
        return KV_ERR_INVALID;
    }
    for (int i = 0; i < 3 * q.np; ++i)
        if (!std::isfinite(q.det[i]) || !std::isfinite(q.src[i])) {
            kv_set_error("non-finite source or detector coordinate at index %d", i);
            return KV_ERR_INVALID;
        }
    KV_TRY_HIP(hipStreamSynchronize(ctx->stream));
    release_geometry(ctx);
    kv_status st;
    if ((st = dev_alloc((void **)&ctx->d_det, (size_t)q.np * 12)) ||
        (st = dev_alloc((void **)&ctx->d_src, (size_t)q.np * 12)) ||
        (st = dev_alloc((void **)&ctx->d_rc, (size_t)q.np * (KV_RC + 4 * q.nu + 4 * q.nv) * 4)) ||
        (st = dev_alloc((void **)&ctx->d_ratio, (size_t)nray * 4))) {
        release_geometry(ctx);
        return st;
    }
    KV_TRY_HIP(hipMemcpy(ctx->d_det, q.det, (size_t)q.np * 12, hipMemcpyHostToDevice));
    KV_TRY_HIP(hipMemcpy(ctx->d_src, q.src, (size_t)q.np * 12, hipMemcpyHostToDevice));
    {
        const std::vector<float> rc = ray_constants(q);
        KV_TRY_HIP(hipMemcpy(ctx->d_rc, rc.data(), rc.size() * 4, hipMemcpyHostToDevice));
    }
    KvGeom &g = ctx->g;
    g.nx = q.nx; g.ny = q.ny; g.nz = q.nz;
    g.nu = q.nu; g.nv = q.nv; g.np = q.np;
    g.sx = q.sx; g.sy = q.sy; g.sz = q.sz;
    g.pu = q.pu; g.pv = q.pv;
    g.det = ctx->d_det;
    g.src = ctx->d_src;
    g.rc = ctx->d_rc;
    g.walk_fast = walk_fast_ok(q) ? 1 : 0;
    g.bp_posz = bp_posz_ok(q) ? 1 : 0;
    ctx->has_geom = true;
    return KV_OK;
}

kv_status kv_malloc(kv_context *ctx, size_t bytes, void **d_ptr)
{
    if (kv_status st = check_ctx(ctx, false))
        return st;
    if (!d_ptr) {
        kv_set_error("null output pointer");
        return KV_ERR_INVALID;
    }
    return dev_alloc(d_ptr, bytes ? bytes : 1);
}

kv_status kv_free(kv_context *ctx, void *d_ptr)
{
    if (kv_status st = check_ctx(ctx, false))
        return st;
    if (d_ptr)
        KV_TRY_HIP(hipFree(d_ptr));
    return KV_OK;
}

kv_status kv_memcpy_h2d(kv_context *ctx, void *d_dst, const void *h_src, size_t bytes)
{
    if (kv_status st = check_ctx(ctx, false))
        return st;
    KV_TRY_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    KV_TRY_HIP(hipStreamSynchronize(ctx->stream));
    return KV_OK;
}

kv_status kv_memcpy_d2h(kv_context *ctx, void *h_dst, const void *d_src, size_t bytes)
{
    if (kv_status st = check_ctx(ctx, false))
        return st;
    KV_TRY_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    KV_TRY_HIP(hipStreamSynchronize(ctx->stream));
    return KV_OK;
}

kv_status kv_memcpy_d2d(kv_context *ctx, void *d_dst, const void *d_src, size_t bytes)
{
    if (kv_status st = check_ctx(ctx, false))
        return st;
    if (!d_dst || !d_src) {
        kv_set_error("null buffer");
        return KV_ERR_INVALID;
    }
    KV_TRY_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return KV_OK;
}

kv_status kv_memset(kv_context *ctx, void *d_dst, int value, size_t bytes)
{
    if (kv_status st = check_ctx(ctx, false))
        return st;
    KV_TRY_HIP(hipMemsetAsync(d_dst, value, bytes, ctx->stream));
    return KV_OK;
}

kv_status kv_forward_project(kv_context *ctx, const float *d_vol, float *d_proj, int mode)
{
    if (kv_status st = check_ctx(ctx, true))
        return st;
    if (kv_status st = check_mode(mode))
        return st;
    if (!d_vol || !d_proj) {
        kv_set_error("null buffer");
        return KV_ERR_INVALID;
    }
    if (mode == KV_MODE_R)
        KV_TRY_HIP(kv_launch_fp_r(ctx->g, d_vol, d_proj, KV_EP_PLAIN, nullptr, 0.f, ctx->stream));
    else
        KV_TRY_HIP(kv_launch_fp_a(ctx->g, ctx->tab_a, d_vol, d_proj, KV_EP_PLAIN, nullptr, ctx->stream));
    return KV_OK;
}

kv_status kv_back_project(kv_context *ctx, const float *d_proj, float *d_vol, int mode)
{
    if (kv_status st = check_ctx(ctx, true))
        return st;
    if (kv_status st = check_mode(mode))
        return st;
    if (!d_vol || !d_proj) {
        kv_set_error("null buffer");
        return KV_ERR_INVALID;
    }
    if (kv_status st = ensure_tables(ctx, mode))
        return st;
    if (mode == KV_MODE_R) {
        KV_TRY_HIP(kv_launch_bp_r(ctx->g, ctx->tab_r, d_proj, d_vol, KV_EP_PLAIN, ctx->stream));
    } else {
        // each ray's length factor is applied once per pixel, into the
        // projection-sized work buffer, not once per (voxel, pixel) pair
        KV_TRY_HIP(kv_launch_scale_a(ctx->g, ctx->tab_a, d_proj, ctx->d_ratio, ctx->stream));
        KV_TRY_HIP(kv_launch_bp_a(ctx->g, ctx->tab_a, ctx->d_ratio, d_vol, KV_EP_PLAIN, nullptr, false, ctx->stream));
    }
    return KV_OK;
}

kv_status kv_mlem_forward(kv_context *ctx, const float *d_vol, const float *d_data, float relax, int mode)
{
    if (kv_status st = check_ctx(ctx, true))
        return st;
    if (kv_status st = check_mode(mode))
        return st;
    if (!d_vol || !d_data) {
        kv_set_error("null buffer");
        return KV_ERR_INVALID;
    }
    // tables (and the mode-A sensitivity) are built here so that the
    // backward half enqueues a single kernel
    if (kv_status st = (mode == KV_MODE_R) ? ensure_tables(ctx, KV_MODE_R) : ensure_sens(ctx))
        return st;
    if (mode == KV_MODE_R)
        KV_TRY_HIP(kv_launch_fp_r(ctx->g, d_vol, ctx->d_ratio, KV_EP_MLEM, d_data, relax, ctx->stream));
    else
        KV_TRY_HIP(kv_launch_fp_a(ctx->g, ctx->tab_a, d_vol, ctx->d_ratio, KV_EP_MLEM, d_data, ctx->stream));
    return KV_OK;
}

kv_status kv_mlem_backward(kv_context *ctx, float *d_vol, int mode)
{
    if (kv_status st = check_ctx(ctx, true))
        return st;
    if (kv_status st = check_mode(mode))
        return st;
    if (!d_vol) {
        kv_set_error("null buffer");
        return KV_ERR_INVALID;
    }
    if (kv_status st = (mode == KV_MODE_R) ? ensure_tables(ctx, KV_MODE_R) : ensure_sens(ctx))
        return st;
    if (mode == KV_MODE_R)
        KV_TRY_HIP(kv_launch_bp_r(ctx->g, ctx->tab_r, ctx->d_ratio, d_vol, KV_EP_MLEM, ctx->stream));
    else
        KV_TRY_HIP(
            kv_launch_bp_a(ctx->g, ctx->tab_a, ctx->d_ratio, d_vol, KV_EP_MLEM, ctx->d_sens, false, ctx->stream));
    return KV_OK;
}

kv_status kv_mlem_iteration(kv_context *ctx, float *d_vol, const float *d_data, float relax, int mode)
{
    if (kv_status st = kv_mlem_forward(ctx, d_vol, d_data, relax, mode))
        return st;
    return kv_mlem_backward(ctx, d_vol, mode);
}

kv_status kv_count_pairs(kv_context *ctx, int mode, int64_t *fp_pairs, int64_t *bp_pairs)
{
    if (kv_status st = check_ctx(ctx, true))
        return st;
    if (kv_status st = check_mode(mode))
        return st;
    if (!fp_pairs || !bp_pairs) {
        kv_set_error("null output pointer");
        return KV_ERR_INVALID;
    }
    if (kv_status st = ensure_tables(ctx, mode))
        return st;
    const int nb = 1024;
    if (kv_status st = ensure_red(ctx, (size_t)nb))
        return st;
    float *d_cnt = nullptr;
    if (kv_status st = dev_alloc((void **)&d_cnt, (size_t)kv_nvox(ctx->g) * 4))
        return st;
    std::vector<double> part(nb);
    double sums[2] = {0.0, 0.0};
    hipError_t e = hipSuccess;
    for (int pass = 0; pass < 2 && e == hipSuccess; ++pass) {
        float *buf = pass == 0 ? ctx->d_ratio : d_cnt;
        const int64_t n = pass == 0 ? kv_nray(ctx->g) : kv_nvox(ctx->g);
        if (pass == 0)
            e = mode == KV_MODE_R ? kv_launch_fp_r(ctx->g, nullptr, buf, KV_EP_COUNT, nullptr, 0.f, ctx->stream)
                                  : kv_launch_fp_a(ctx->g, ctx->tab_a, nullptr, buf, KV_EP_COUNT, nullptr, ctx->stream);
        else
            e = mode == KV_MODE_R ? kv_launch_bp_r(ctx->g, ctx->tab_r, nullptr, buf, KV_EP_COUNT, ctx->stream)
                                  : kv_launch_bp_a(ctx->g, ctx->tab_a, nullptr, buf, KV_EP_COUNT, nullptr, true,
                                                   ctx->stream);
        if (e == hipSuccess)
            e = kv_launch_sum(n, buf, ctx->d_red, nb, ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(part.data(), ctx->d_red, nb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(ctx->stream);
        for (int i = 0; i < nb && e == hipSuccess; ++i)
            sums[pass] += part[i];
    }
    free_dev(d_cnt);
    if (e != hipSuccess) {
        kv_set_error("pair count failed: %s", hipGetErrorString(e));
        return KV_ERR_DEVICE;
    }
    *fp_pairs = (int64_t)sums[0];
    *bp_pairs = (int64_t)sums[1];
    return KV_OK;
}

kv_status kv_sart_iteration(kv_context *ctx, float *d_vol, const float *d_data, float relax)
{
    if (kv_status st = check_ctx(ctx, true))
        return st;
    if (!d_vol || !d_data) {
        kv_set_error("null buffer");
        return KV_ERR_INVALID;
    }
    if (kv_status st = ensure_tables(ctx, KV_MODE_R))
        return st;
    KV_TRY_HIP(kv_launch_fp_r(ctx->g, d_vol, ctx->d_ratio, KV_EP_SART, d_data, relax, ctx->stream));
    KV_TRY_HIP(kv_launch_bp_r(ctx->g, ctx->tab_r, ctx->d_ratio, d_vol, KV_EP_SART, ctx->stream));
    return KV_OK;
}

kv_status kv_correct_projections(kv_context *ctx, int loop, float relax, const float *d_data, float *d_proj)
{
    if (kv_status st = check_ctx(ctx, true))
        return st;
    if (!d_data || !d_proj || (loop != KV_LOOP_ART && loop != KV_LOOP_MLEM)) {
        kv_set_error("bad correction arguments");
        return KV_ERR_INVALID;
    }
    KV_TRY_HIP(kv_launch_correct(kv_nray(ctx->g), loop, relax, d_data, d_proj, ctx->stream));
    return KV_OK;
}

kv_status kv_apply_update(kv_context *ctx, int loop, float *d_vol, const float *d_bp)
{
    if (kv_status st = check_ctx(ctx, true))
        return st;
    if (!d_vol || !d_bp || (loop != KV_LOOP_ART && loop != KV_LOOP_MLEM)) {
        kv_set_error("bad update arguments");
        return KV_ERR_INVALID;
    }
    KV_TRY_HIP(kv_launch_update(kv_nvox(ctx->g), loop, d_vol, d_bp, ctx->stream));
    return KV_OK;
}

kv_status kv_sensitivity(kv_context *ctx, float *d_vol)
{
    if (kv_status st = check_ctx(ctx, true))
        return st;
    if (!d_vol) {
        kv_set_error("null buffer");
        return KV_ERR_INVALID;
    }
    if (kv_status st = ensure_tables(ctx, KV_MODE_A))
        return st;
    KV_TRY_HIP(kv_launch_bp_a(ctx->g, ctx->tab_a, nullptr, d_vol, KV_EP_PLAIN, nullptr, true, ctx->stream));
    return KV_OK;
}

namespace {
// Arguments and preconditions of recons::ShiftAndAdd (Reconstructors.h:
// 274-305,319-323) and the row shifts t[k][p] = (int)(y_p (mZ - mF) / dy) of
// Reconstructors.h:308-331, written to `shifts` when it is not null (the
// no-allocation variant evaluates them on the device instead).
kv_status saa_check(kv_context *ctx, const float *d_proj, const float *src_y, int nplanes, const float *plane_z,
                    float sid, float fulcrum_z, float voxel_dy, const float *d_out, int32_t *shifts)
{
    if (kv_status st = check_ctx(ctx, true))
        return st;
    if (!d_proj || !d_out || !src_y || (nplanes > 0 && !plane_z)) {
        kv_set_error("null buffer");
        return KV_ERR_INVALID;
    }
    const KvGeom &g = ctx->g;
    if (nplanes <= 1) {
        kv_set_error("ShiftAndAdd: required nb of slices in reconstruction %d <= 1", nplanes);
        return KV_ERR_PRECONDITION;
    }
    if (g.np <= 1) {
        kv_set_error("ShiftAndAdd: required nb of projections %d <= 1", g.np);
        return KV_ERR_PRECONDITION;
    }
    if (almost_equal_relative(fulcrum_z, sid)) {
        kv_set_error("ShiftAndAdd: Fulcrum and imager are at the same height: impossible");
        return KV_ERR_PRECONDITION;
    }
    if (!finite_pos(voxel_dy)) {
        kv_set_error("ShiftAndAdd: voxel y spacing must be finite and positive");
        return KV_ERR_INVALID;
    }
    const float mF = sid / (sid - fulcrum_z);
    for (int k = 0; k < nplanes; ++k) {
        const float z = plane_z[k];
        if (almost_equal_relative(z, sid)) {
            kv_set_error("ShiftAndAdd: reconstruction slice and imager are at the same height: impossible");
            return KV_ERR_PRECONDITION;
        }
        const float mZ = sid / (sid - z);
        const float coeff = mZ - mF;
        for (int p = 0; p < g.np; ++p) {
            const float world = src_y[p] * coeff;
            const float t = world / voxel_dy;
            if (!(std::fabs(t) < 2.0e9f)) {
                kv_set_error("ShiftAndAdd: shift of plane %d view %d out of range", k, p);
                return KV_ERR_INVALID;
            }
            if (shifts)
                shifts[(size_t)k * g.np + p] = (int32_t)t;
        }
    }
    return KV_OK;
}
}  // namespace

kv_status kv_shift_and_add(kv_context *ctx, const float *d_proj, const float *src_y, int nplanes,
                           const float *plane_z, float sid, float fulcrum_z, float voxel_dy, float *d_out)
{
    if (kv_status st = check_ctx(ctx, true))
        return st;
    const KvGeom &g = ctx->g;
    std::vector<int32_t> shifts((size_t)std::max(nplanes, 0) * g.np);
    if (kv_status st = saa_check(ctx, d_proj, src_y, nplanes, plane_z, sid, fulcrum_z, voxel_dy, d_out,
                                 shifts.data()))
        return st;
    const int kc = kv_saa_chunk(nplanes, g.np, [&](int k, int p) { return shifts[(size_t)k * g.np + p]; });
    int32_t *d_shifts = nullptr;
    if (kv_status st = dev_alloc((void **)&d_shifts, shifts.size() * 4))
        return st;
    hipError_t e = hipMemcpyAsync(d_shifts, shifts.data(), shifts.size() * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = kv_launch_saa(g.nu, g.nv, g.np, d_proj, d_shifts, nplanes, kc, d_out, ctx->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d_shifts);
    if (e != hipSuccess) {
        kv_set_error("shift-and-add failed: %s", hipGetErrorString(e));
        return KV_ERR_DEVICE;
    }
    return KV_OK;
}

kv_status kv_shift_and_add_noalloc(kv_context *ctx, const float *d_proj, const float *src_y, int nplanes,
                                   const float *plane_z, float sid, float fulcrum_z, float voxel_dy, float *d_out)
{
    if (kv_status st = saa_check(ctx, d_proj, src_y, nplanes, plane_z, sid, fulcrum_z, voxel_dy, d_out, nullptr))
        return st;
    const KvGeom &g = ctx->g;
    if (g.np > KV_SAA_NOALLOC_MAX_VIEWS || nplanes > KV_SAA_NOALLOC_MAX_PLANES ||
        (int64_t)g.np * nplanes > KV_SAA_NOALLOC_MAX_SHIFTS) {
        kv_set_error("shift-and-add without allocation: %d views x %d planes exceeds the staging limits "
                     "(%d views, %d planes, %d view-plane pairs)",
                     g.np, nplanes, KV_SAA_NOALLOC_MAX_VIEWS, KV_SAA_NOALLOC_MAX_PLANES, KV_SAA_NOALLOC_MAX_SHIFTS);
        return KV_ERR_INVALID;
    }
    int32_t *scr = kv_internal_scratch(ctx->device);
    if (!scr) {
        kv_set_error("shift-and-add without allocation: staging array of device %d not resolved", ctx->device);
        return KV_ERR_STATE;
    }
    float *d_sy = (float *)scr;
    float *d_pz = (float *)(scr + KV_SAA_NOALLOC_MAX_VIEWS);
    int32_t *d_shifts = scr + KV_SAA_NOALLOC_MAX_VIEWS + KV_SAA_NOALLOC_MAX_PLANES;
    const float mF = sid / (sid - fulcrum_z);
    // the shifts the device evaluates, recomputed here (same expressions) to
    // pick the kernel's plane chunk
    const int kc = kv_saa_chunk(nplanes, g.np, [&](int k, int p) {
        const float mZ = sid / (sid - plane_z[k]);
        const float coeff = mZ - mF;
        const float world = src_y[p] * coeff;
        return (int)(int32_t)(world / voxel_dy);
    });
    hipError_t e = hipMemcpyAsync(d_sy, src_y, (size_t)g.np * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_pz, plane_z, (size_t)nplanes * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = kv_launch_saa_shifts(g.np, nplanes, sid, mF, voxel_dy, d_sy, d_pz, d_shifts, ctx->stream);
    if (e == hipSuccess)
        e = kv_launch_saa(g.nu, g.nv, g.np, d_proj, d_shifts, nplanes, kc, d_out, ctx->stream);
    // the caller's host arrays are read by the copies: return when they are done
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        kv_set_error("shift-and-add failed: %s", hipGetErrorString(e));
        return KV_ERR_DEVICE;
    }
    return KV_OK;
}

kv_status kv_log_likelihood(kv_context *ctx, const float *d_data, const float *d_vol, double *out)
{
    if (kv_status st = check_ctx(ctx, true))
        return st;
    if (!d_data || !d_vol || !out) {
        kv_set_error("null buffer");
        return KV_ERR_INVALID;
    }
    const int nb = 1024;
    if (kv_status st = ensure_red(ctx, (size_t)nb))
        return st;
    KV_TRY_HIP(kv_launch_fp_a(ctx->g, ctx->tab_a, d_vol, ctx->d_ratio, KV_EP_PLAIN, nullptr, ctx->stream));
    KV_TRY_HIP(kv_launch_loglik(ctx->g, d_data, ctx->d_ratio, ctx->d_red, nb, ctx->stream));
    std::vector<double> part(nb);
    KV_TRY_HIP(hipMemcpyAsync(part.data(), ctx->d_red, nb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    KV_TRY_HIP(hipStreamSynchronize(ctx->stream));
    double s = 0.0;
    for (int i = 0; i < nb; ++i)
        s += part[i];
    *out = s;
    return KV_OK;
}

kv_status kv_log_likelihoods(kv_context *ctx, const float *d_data, const float *const *d_vols, int n, double *out)
{
    if (kv_status st = check_ctx(ctx, true))
        return st;
    if (!d_data || !d_vols || !out || n < 1 || n > KV_LL_MAX) {
        kv_set_error("log-likelihoods need 1..%d volumes", KV_LL_MAX);
        return KV_ERR_INVALID;
    }
    for (int i = 0; i < n; ++i) {
        if (!d_vols[i]) {
            kv_set_error("null volume");
            return KV_ERR_INVALID;
        }
    }
    const int nb = 1024;
    if (kv_status st = ensure_red(ctx, (size_t)nb))
        return st;
    const size_t nray = (size_t)kv_nray(ctx->g);
    if (n > 1 && !ctx->d_ll) {
        if (kv_status st = dev_alloc((void **)&ctx->d_ll, (KV_LL_MAX - 1) * nray * sizeof(float)))
            return st;
    }
    if (n > 1 && !ctx->d_pack) {
        if (kv_status st = dev_alloc((void **)&ctx->d_pack, 4 * (size_t)kv_nvox(ctx->g) * sizeof(float)))
            return st;
    }
    float *outs[KV_LL_MAX] = {ctx->d_ratio, ctx->d_ll, ctx->d_ll ? ctx->d_ll + nray : nullptr};
    KV_TRY_HIP(kv_launch_fp_a_multi(ctx->g, n, d_vols, ctx->d_pack, outs, ctx->stream));
    std::vector<double> part(nb);
    for (int i = 0; i < n; ++i) {
        // the same reduction as kv_log_likelihood
        KV_TRY_HIP(kv_launch_loglik(ctx->g, d_data, outs[i], ctx->d_red, nb, ctx->stream));
        KV_TRY_HIP(hipMemcpyAsync(part.data(), ctx->d_red, nb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        KV_TRY_HIP(hipStreamSynchronize(ctx->stream));
        double sum = 0.0;
        for (int b = 0; b < nb; ++b)
            sum += part[b];
        out[i] = sum;
    }
    return KV_OK;
}

kv_status kv_phantom(kv_context *ctx, float background, const kv_pave *paves, int npaves, const kv_sphere *spheres,
                     int nspheres, float *d_vol)
{
    if (kv_status st = check_ctx(ctx, true))
        return st;
    if (!d_vol || npaves < 0 || nspheres < 0 || (npaves > 0 && !paves) || (nspheres > 0 && !spheres)) {
        kv_set_error("bad phantom arguments");
        return KV_ERR_INVALID;
    }
    const KvGeom &g = ctx->g;
    // grid points: the bottom-left-front corner -0.5 * (n * spacing)
    // (TomoGeometry.cpp:121-127, WSize3D.cpp:7-9) and sequential FP32 sums
    // of the spacing from it (TomoVolume.cpp:193-215)
    std::vector<float> pos((size_t)g.nx + g.ny + g.nz);
    auto axis = [&](int n, float sp, float *dst) {
        const float w = (float)n * sp;
        dst[0] = (float)(-0.5 * (double)w);
        for (int i = 1; i < n; ++i)
            dst[i] = dst[i - 1] + sp;
    };
    axis(g.nx, g.sx, pos.data());
    axis(g.ny, g.sy, pos.data() + g.nx);
    axis(g.nz, g.sz, pos.data() + g.nx + g.ny);
    const size_t bpos = pos.size() * sizeof(float), bpav = (size_t)npaves * sizeof(kv_pave),
                 bsph = (size_t)nspheres * sizeof(kv_sphere);
    const size_t opav = (bpos + 255) & ~(size_t)255, osph = (opav + bpav + 255) & ~(size_t)255;
    char *d_buf = nullptr;
    if (kv_status st = dev_alloc((void **)&d_buf, osph + bsph + 4))
        return st;
    hipError_t e = hipMemcpyAsync(d_buf, pos.data(), bpos, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && bpav)
        e = hipMemcpyAsync(d_buf + opav, paves, bpav, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && bsph)
        e = hipMemcpyAsync(d_buf + osph, spheres, bsph, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        const float *px = (const float *)d_buf;
        e = kv_launch_phantom(g.nx, g.ny, g.nz, px, px + g.nx, px + g.nx + g.ny, background,
                              (const kv_pave *)(d_buf + opav), npaves, (const kv_sphere *)(d_buf + osph), nspheres,
                              d_vol, ctx->stream);
    }
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d_buf);
    if (e != hipSuccess) {
        kv_set_error("phantom failed: %s", hipGetErrorString(e));
        return KV_ERR_DEVICE;
    }
    return KV_OK;
}

kv_status kv_log_scale(kv_context *ctx, const float *d_in, float *d_out, int64_t n, double c)
{
    if (kv_status st = check_ctx(ctx, false))
        return st;
    if (!d_in || !d_out || n < 0 || !std::isfinite(c)) {
        kv_set_error("bad log-scale arguments");
        return KV_ERR_INVALID;
    }
    if (n > 0)
        KV_TRY_HIP(kv_launch_log_scale(n, d_in, d_out, c, ctx->stream));
    return KV_OK;
}

namespace {
kv_status traj_args(kv_context *ctx, const float *const *d_iterates, int h, KvTraj &tr)
{
    if (kv_status st = check_ctx(ctx, true))
        return st;
    if (!d_iterates || h < 1 || h > KV_TRAJ_MAX) {
        kv_set_error("trajectory needs 1..%d log-ratios (h = %d)", KV_TRAJ_MAX, h);
        return KV_ERR_INVALID;
    }
    for (int i = 0; i <= h; ++i) {
        if (!d_iterates[i]) {
            kv_set_error("null iterate %d", i);
            return KV_ERR_INVALID;
        }
        tr.x[i] = d_iterates[i];
    }
    for (int i = h + 1; i <= KV_TRAJ_MAX; ++i)
        tr.x[i] = nullptr;
    return KV_OK;
}
}  // namespace

kv_status kv_traj_gram(kv_context *ctx, const float *const *d_iterates, int h, double *gram)
{
    KvTraj tr;
    if (kv_status st = traj_args(ctx, d_iterates, h, tr))
        return st;
    if (!gram) {
        kv_set_error("null gram");
        return KV_ERR_INVALID;
    }
    if (kv_status st = ensure_sens(ctx))
        return st;
    const int64_t n = kv_nvox(ctx->g);
    const int nb = kv_traj_gram_blocks(n), np = kv_traj_pairs(h);
    const int need = nb * np * 16;
    if (kv_status st = ensure_red(ctx, (size_t)need))
        return st;
    KV_TRY_HIP(kv_launch_traj_gram(tr, h, n, ctx->d_sens, ctx->d_red, ctx->stream));
    std::vector<double> part((size_t)need);
    KV_TRY_HIP(hipMemcpyAsync(part.data(), ctx->d_red, (size_t)need * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
    KV_TRY_HIP(hipStreamSynchronize(ctx->stream));
    // workgroup partials in workgroup order; block pair p = (bi, bj), bi <= bj,
    // row-major over the upper triangle of 4x4 blocks
    const int nbk = (h + 3) / 4;
    std::vector<double> blk((size_t)np * 16, 0.0);
    for (int b = 0; b < nb; ++b)
        for (int o = 0; o < np * 16; ++o)
            blk[o] += part[(size_t)b * np * 16 + o];
    int p = 0;
    for (int bi = 0; bi < nbk; ++bi)
        for (int bj = bi; bj < nbk; ++bj, ++p)
            for (int r = 0; r < 4; ++r)
                for (int c = 0; c < 4; ++c) {
                    const int i = 4 * bi + r, j = 4 * bj + c;
                    if (i >= h || j >= h)
                        continue;
                    const double v = blk[(size_t)p * 16 + r * 4 + c];
                    gram[i * h + j] = v;
                    gram[j * h + i] = v;
                }
    return KV_OK;
}

kv_status kv_traj_combine(kv_context *ctx, const float *const *d_iterates, int h, const double *coef, float *d_out)
{
    KvTraj tr;
    if (kv_status st = traj_args(ctx, d_iterates, h, tr))
        return st;
    if (!coef || !d_out) {
        kv_set_error("null coefficients or output");
        return KV_ERR_INVALID;
    }
    KvTrajCoef cf;
    for (int i = 0; i < KV_TRAJ_MAX; ++i) {
        cf.c[i] = i < h ? coef[i] : 0.0;
        if (!std::isfinite(cf.c[i])) {
            kv_set_error("coefficient %d is not finite", i);
            return KV_ERR_INVALID;
        }
    }
    KV_TRY_HIP(kv_launch_traj_combine(tr, cf, h, kv_nvox(ctx->g), d_out, ctx->stream));
    return KV_OK;
}

kv_status kv_trace_rays(kv_context *ctx, int mode, int nrays, const int32_t *rays, int cap, int64_t *idx, float *w,
                        int32_t *counts)
{
    if (kv_status st = check_ctx(ctx, true))
        return st;
    if (kv_status st = check_mode(mode))
        return st;
    if (nrays <= 0 || cap <= 0 || !rays || !idx || !w || !counts) {
        kv_set_error("bad trace arguments");
        return KV_ERR_INVALID;
    }
    const int64_t nr = kv_nray(ctx->g);
    for (int i = 0; i < nrays; ++i)
        if (rays[i] < 0 || rays[i] >= nr) {
            kv_set_error("ray id %d out of range", rays[i]);
            return KV_ERR_INVALID;
        }
    int32_t *d_rays = nullptr, *d_counts = nullptr;
    int64_t *d_idx = nullptr;
    float *d_w = nullptr;
    kv_status st = KV_OK;
    const size_t slots = (size_t)nrays * cap;
    if ((st = dev_alloc((void **)&d_rays, (size_t)nrays * 4)) || (st = dev_alloc((void **)&d_counts, (size_t)nrays * 4)) ||
        (st = dev_alloc((void **)&d_idx, slots * 8)) || (st = dev_alloc((void **)&d_w, slots * 4))) {
        free_dev(d_rays);
        free_dev(d_counts);
        free_dev(d_idx);
        free_dev(d_w);
        return st;
    }
    hipError_t e = hipMemcpyAsync(d_rays, rays, (size_t)nrays * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = kv_launch_trace(ctx->g, mode, nrays, d_rays, cap, d_idx, d_w, d_counts, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(idx, d_idx, slots * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(w, d_w, slots * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(counts, d_counts, (size_t)nrays * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    free_dev(d_rays);
    free_dev(d_counts);
    free_dev(d_idx);
    free_dev(d_w);
    if (e != hipSuccess) {
        kv_set_error("trace failed: %s", hipGetErrorString(e));
        return KV_ERR_DEVICE;
    }
    return KV_OK;
}

static kv_status host_roundtrip(kv_context *ctx, const float *h_in, size_t in_n, float *h_out, size_t out_n, int mode,
                                bool forward)
{
    float *d_in = nullptr, *d_out = nullptr;
    kv_status st;
    if ((st = dev_alloc((void **)&d_in, in_n * 4)) || (st = dev_alloc((void **)&d_out, out_n * 4))) {
        free_dev(d_in);
        free_dev(d_out);
        return st;
    }
    hipError_t e = hipMemcpyAsync(d_in, h_in, in_n * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
        kv_set_error("copy to device failed: %s", hipGetErrorString(e));
        st = KV_ERR_DEVICE;
    } else {
        st = forward ? kv_forward_project(ctx, d_in, d_out, mode) : kv_back_project(ctx, d_in, d_out, mode);
    }
    if (st == KV_OK) {
        e = hipMemcpyAsync(h_out, d_out, out_n * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            kv_set_error("copy to host failed: %s", hipGetErrorString(e));
            st = KV_ERR_DEVICE;
        }
    } else {
        (void)hipStreamSynchronize(ctx->stream);
    }
    free_dev(d_in);
    free_dev(d_out);
    return st;
}

kv_status kv_forward_project_host(kv_context *ctx, const float *h_vol, float *h_proj, int mode)
{
    if (kv_status st = check_ctx(ctx, true))
        return st;
    if (!h_vol || !h_proj) {
        kv_set_error("null buffer");
        return KV_ERR_INVALID;
    }
    return host_roundtrip(ctx, h_vol, (size_t)kv_nvox(ctx->g), h_proj, (size_t)kv_nray(ctx->g), mode, true);
}

kv_status kv_back_project_host(kv_context *ctx, const float *h_proj, float *h_vol, int mode)
{
    if (kv_status st = check_ctx(ctx, true))
        return st;
    if (!h_vol || !h_proj) {
        kv_set_error("null buffer");
        return KV_ERR_INVALID;
    }
    return host_roundtrip(ctx, h_proj, (size_t)kv_nray(ctx->g), h_vol, (size_t)kv_nvox(ctx->g), mode, false);
}

}  // extern "C"
