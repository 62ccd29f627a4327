// Backprojectors for gfx950: voxel-driven gathers.  Every voxel sums its own
// contributions in a fixed order (views, detector rows, detector columns, as
// Projector.cu:452-729), so results are deterministic without atomics.
//
// The reference evaluates, per voxel and view, every pixel of an 11x11 window
// with two FP32 divisions each (Projector.cu:656-669).  Three observations make
// this cheap without changing a single bit of the result:
//  * alpha1.x / alpha2.x depend only on (view, detector column, voxel x), and
//    alpha2.x of voxel x equals alpha1.x of voxel x+1 (k - nx/2 + 1 is exact
//    in FP32): they are one table per view, x[p][k][u] for planes k = 0..nx.
//    The same holds for y (per row) and z (per view).  The tables are built
//    once per geometry.
//  * A pixel contributes only when its ray crosses the voxel box with positive
//    alpha-length (weight > 1e-24, Projector.cu:710-725), i.e. when the pixel
//    centre lies inside the box's shadow on the detector.  Each lane loops
//    over the window intersected with that shadow (widened by 0.02 pixel);
//    the skipped pixels are exactly those whose weight test fails, and the
//    accumulation order of the others is unchanged.
//  * The detector rows of a voxel's shadow and window depend only on (y, z),
//    so the 64 lanes of a wavefront (64 consecutive x of one (y, z) row) walk
//    the same rows: the y and z terms are wave-uniform.  They are tabulated
//    once per geometry (KvTables::yz) and, with the y-table rows and the
//    per-view constants, read by scalar loads, so the scalar unit does that
//    share of the work beside the vector unit.
//
// A 256-thread workgroup owns a 64 x 4 x ZT brick (one wavefront per y row,
// looping over ZT slices).  Per view it stages the bands of the x table its
// brick needs in LDS, with an odd row stride so that the lanes' reads, which
// sit at similar offsets of consecutive rows, fall in different banks.  Table
// entries outside a band are read from the global table, so the band width
// only affects speed.
#pragma clang fp contract(off)

#include "kv_internal.h"

namespace {

constexpr float kInf = __builtin_inff();

// ------------------------------------------------------------ tables -----
// Layout: x[p][k][u] (k = 0..nx), y[p][j][v] (j = 0..ny), z[p][l] (l = 0..nz).
// Mode R entries, Projector.cu:651-669.
__global__ void tables_r_kernel(KvGeom g, KvTables t)
{
    const int64_t nxe = (int64_t)g.np * (g.nx + 1) * g.nu;
    const int64_t nye = (int64_t)g.np * (g.ny + 1) * g.nv;
    const int64_t nze = (int64_t)g.np * (g.nz + 1);
    const float tol = 1e-24f;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nxe + nye + nze;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (i < nxe) {
            const int u = (int)(i % g.nu);
            const int64_t r = i / g.nu;
            const int k = (int)(r % (g.nx + 1)), p = (int)(r / (g.nx + 1));
            const float sfx = g.src[3 * p] / g.sx;
            const float ppx = (g.det[3 * p] + ((float)u + 0.5f) * g.pu) / g.sx;
            const float rx = ppx - sfx;
            const float cv = (float)k - (float)g.nx / 2.f;
            t.x[i] = fabsf(rx) > tol ? (cv - sfx) / rx : 0.f;
        } else if (i < nxe + nye) {
            const int64_t j = i - nxe;
            const int v = (int)(j % g.nv);
            const int64_t r = j / g.nv;
            const int k = (int)(r % (g.ny + 1)), p = (int)(r / (g.ny + 1));
            const float sfy = g.src[3 * p + 1] / g.sy;
            const float ppy = (g.det[3 * p + 1] + ((float)v + 0.5f) * g.pv) / g.sy;
            const float ry = ppy - sfy;
            const float cv = (float)k - (float)g.ny / 2.f;
            t.y[j] = fabsf(ry) > tol ? (cv - sfy) / ry : 0.f;
        } else {
            const int64_t j = i - nxe - nye;
            const int k = (int)(j % (g.nz + 1));
            const int p = (int)(j / (g.nz + 1));
            const float sfz = g.src[3 * p + 2] / g.sz;
            const float ppz = g.det[3 * p + 2] / g.sz;
            const float rz = ppz - sfz;
            const float cv = (float)k - (float)g.nz / 2.f;
            t.z[j] = fabsf(rz) > tol ? (cv - sfz) / rz : 0.f;
        }
    }
}

// Mode A entries: ((k - n/2) - s) * (1/d); a ray parallel to the planes of an
// axis pins one voxel index f: entries k <= f are -INF, k > f are +INF.
__device__ __forceinline__ float plane_a_entry(float s, float q, float h, int n, int k)
{
    const float d = q - s;
    if (d != 0.f) {
        const float inv = 1.f / d;
        return (((float)k - h) - s) * inv;
    }
    const float pos = s + h;
    const int f = (pos > 0.f && pos < (float)n) ? (int)floorf(pos) : -1;
    return k <= f ? -kInf : kInf;
}

__global__ void tables_a_kernel(KvGeom g, KvTables t)
{
    const int64_t nxe = (int64_t)g.np * (g.nx + 1) * g.nu;
    const int64_t nye = (int64_t)g.np * (g.ny + 1) * g.nv;
    const int64_t nze = (int64_t)g.np * (g.nz + 1);
    const float hx = (float)g.nx / 2.f, hy = (float)g.ny / 2.f, hz = (float)g.nz / 2.f;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nxe + nye + nze;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (i < nxe) {
            const int u = (int)(i % g.nu);
            const int64_t r = i / g.nu;
            const int k = (int)(r % (g.nx + 1)), p = (int)(r / (g.nx + 1));
            const float s = g.src[3 * p] / g.sx;
            const float q = (g.det[3 * p] + ((float)u + 0.5f) * g.pu) / g.sx;
            t.x[i] = plane_a_entry(s, q, hx, g.nx, k);
            if (k == 0) {
                const float l = (q - s) * g.sx;
                t.lx2[(size_t)p * g.nu + u] = l * l;
            }
        } else if (i < nxe + nye) {
            const int64_t j = i - nxe;
            const int v = (int)(j % g.nv);
            const int64_t r = j / g.nv;
            const int k = (int)(r % (g.ny + 1)), p = (int)(r / (g.ny + 1));
            const float s = g.src[3 * p + 1] / g.sy;
            const float q = (g.det[3 * p + 1] + ((float)v + 0.5f) * g.pv) / g.sy;
            t.y[j] = plane_a_entry(s, q, hy, g.ny, k);
            if (k == 0) {
                const float l = (q - s) * g.sy;
                t.ly2[(size_t)p * g.nv + v] = l * l;
            }
        } else {
            const int64_t j = i - nxe - nye;
            const int k = (int)(j % (g.nz + 1));
            const int p = (int)(j / (g.nz + 1));
            const float s = g.src[3 * p + 2] / g.sz;
            const float q = g.det[3 * p + 2] / g.sz;
            t.z[j] = plane_a_entry(s, q, hz, g.nz, k);
            if (k == 0) {
                const float l = (q - s) * g.sz;
                t.lz2[p] = l * l;
            }
        }
    }
}

// Per view and plane: the (view, slice)-uniform terms of the backprojector,
// computed with the kernel's own expressions so that T and AZ are the same
// FP32 values (see KvTables::zs).
__global__ void zs_kernel(KvGeom g, float *__restrict__ zs)
{
    const int n = g.np * (g.nz + 1);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int l = i % (g.nz + 1), p = i / (g.nz + 1);
    const float *S = g.src + 3 * p;
    const float *D = g.det + 3 * p;
    const float sfx = S[0] / g.sx, sfy = S[1] / g.sy, sfz = S[2] / g.sz;
    const float ofx = D[0] / g.sx, ofy = D[1] / g.sy, ofz = D[2] / g.sz;
    const float dd = ofz - sfz;
    const float cvz = (float)l - (float)g.nz / 2.f;
    const float T = dd / (cvz - sfz);
    float az = 0.f;
    if (l < g.nz) {
        const float rdz = (cvz + 0.5f) - sfz;  // Projector.cu:473-480
        az = fabsf(rdz) > 1e-24f ? (ofz - sfz) / rdz : 0.f;
    }
    const float scx = g.sx / g.pu, scy = g.sy / g.pv;
    float *o = zs + (size_t)i * KV_ZS;
    o[0] = T;
    o[1] = az;
    o[2] = T * scx;
    o[3] = ((sfx - sfx * T) - ofx) * scx - 0.5f;
    o[4] = T * scy;
    o[5] = ((sfy - sfy * T) - ofy) * scy - 0.5f;
    o[6] = 0.f;
    o[7] = 0.f;
}

// ------------------------------------------------------ tiled kernel -----
constexpr int TX = 64;           // voxels along x per workgroup (one per lane)
constexpr int TY = 4;            // y rows per workgroup (one per wavefront)
// x band width (detector columns per plane): it must hold the shadows of a
// plane's voxels over the brick's slices (beyond it, table reads go to
// global memory); 24 measured 2 % faster than 32 at full size, 16 26 %
// slower (its shadows no longer fit)
constexpr int XBW = 24;
constexpr int XBS = XBW + 1;     // its LDS row stride (odd: conflict-free)

// lo < hi and v not NaN: the median of the three is the clamp
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return __builtin_amdgcn_fmed3f(v, lo, hi); }

// min/max as signed-integer min/max of the IEEE bit patterns.  They order two
// floats exactly as fminf/fmaxf when both are >= +0, and put every negative
// value (including -0) below every non-negative one.  The fast path uses them
// only in w = min(H, Hy) - max(L, Ly) with L >= +0 (L carries the clip at
// alpha = 0 as +0): w > 0 needs all of H, Hy and min(H, Hy) positive, where
// the results are the FP32 ones bit for bit; when any of them is negative both
// forms give w < 0.  No NaN reaches these operands (entries are finite or
// +-INF).
__device__ __forceinline__ float bmin(float a, float b) { return __int_as_float(min(__float_as_int(a), __float_as_int(b))); }
__device__ __forceinline__ float bmax(float a, float b) { return __int_as_float(max(__float_as_int(a), __float_as_int(b))); }

// Detector index range [lo, hi] whose pixel centres can fall in the shadow of
// the lateral interval [c, c+1] between the depths of two planes, given each
// plane's affine map u = K0 + c * K1 (K1 > 0) from the per-view table.  The
// map is increasing in c, so the shadow spans [min over the planes at c, max
// over the planes at c+1].  The range only has to contain every pixel whose
// ray crosses the voxel with positive weight (the weight test itself decides
// acceptance), so it is widened by 0.02 pixel, far above the FP32 rounding
// of these bounds and of the crossing tables (~1e-3 pixel).
__device__ __forceinline__ void shadow_affine(float c, float k1a, float k0a, float k1b, float k0b, int n, int &lo,
                                              int &hi)
{
    const float pad = 0.02f;
    const float c1 = c + 1.f;
    const float mn = fminf(__builtin_fmaf(c, k1a, k0a), __builtin_fmaf(c, k1b, k0b));
    const float mx = fmaxf(__builtin_fmaf(c1, k1a, k0a), __builtin_fmaf(c1, k1b, k0b));
    const float a = clampf(mn - pad, -2.f, (float)n + 1.f);
    const float b = clampf(mx + pad, -2.f, (float)n + 1.f);
    lo = max((int)ceilf(a), 0);
    hi = min((int)floorf(b), n - 1);
}

// Per view, y row and z slice: the terms of the backprojector that the
// voxels of one (y, z) row share (KvTables::yz), with the expressions the
// backprojector evaluated per row before they were tabulated, so every value
// is the same: the detector-row range of the voxel box's shadow between the
// slice's planes (shadow_affine on the zs table), in mode R intersected with
// the +-5 window around the row of the voxel centre's projection and empty
// when that centre lies off the detector widened by 5 pixels
// (Projector.cu:473-492, 607-646), and the z-plane alpha interval.
template <int MODE>
__global__ void yz_kernel(KvGeom g, KvTables t)
{
    const int64_t n = (int64_t)g.np * g.nz * g.ny;
    const int nz1 = g.nz + 1;
    const float hy = (float)g.ny / 2.f;
    const int half = 5;
    const float wy = (float)half * g.pv / g.sy;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(i % g.nz);
        const int64_t r = i / g.nz;
        const int y = (int)(r % g.ny), p = (int)(r / g.ny);
        const float *S = g.src + 3 * p;
        const float *D = g.det + 3 * p;
        const float sfy = S[1] / g.sy, ofy = D[1] / g.sy;
        const float cvy = (float)y - hy;  // Projector.cu:437-440
        const float rdy = (cvy + 0.5f) - sfy;
        const float efy = (D[1] + (float)g.nv * g.pv) / g.sy;
        const float *Z = t.zs + ((size_t)p * nz1 + z) * KV_ZS;
        const float t0 = Z[0], t1 = Z[KV_ZS];
        const bool hs = t0 > 0.f && t1 > 0.f && t0 < 1e30f && t1 < 1e30f;
        int f0 = 0, f1 = g.nv - 1;
        if (hs)
            shadow_affine(cvy, Z[4], Z[5], Z[KV_ZS + 4], Z[KV_ZS + 5], g.nv, f0, f1);
        if (MODE == KV_MODE_R) {
            const float iy = sfy + rdy * Z[1];
            const bool yin = iy > ofy - wy && iy < efy + wy;
            const int pcy = (int)(((iy - ofy) * g.sy) / g.pv);
            int wly = pcy - half, why = pcy + half;
            wly = wly < 0 ? 0 : (wly >= g.nv ? g.nv - 1 : wly);
            why = why < 0 ? 0 : (why >= g.nv ? g.nv - 1 : why);
            f0 = max(f0, wly);
            f1 = min(f1, why);
            if (!yin) {
                f0 = 0;
                f1 = -1;
            }
        }
        const float a1z = t.z[p * nz1 + z], a2z = t.z[p * nz1 + z + 1];
        float *o = t.yz + i * KV_YZ;
        o[0] = __int_as_float(f0);
        o[1] = __int_as_float(f1);
        o[2] = __int_as_float(hs ? 1 : 0);
        o[3] = 0.f;
        o[4] = fmaxf(0.f, fminf(a1z, a2z));
        o[5] = fminf(1.f, fmaxf(a1z, a2z));
        o[6] = bmax(0.f, fminf(a1z, a2z));  // lz with a +0 floor in bit order
        o[7] = 0.f;
    }
}

// LDS of one brick: the x-table bands and the per-voxel running sums.
template <int ZT>
struct BpShared {
    float xb[(TX + 1) * XBS];
    int xbase[TX + 1];
    float accs[ZT][2][256];  // (value, weight) per slice of the brick
};

// Read-only device memory seen through the constant address space: loads at
// wavefront-uniform addresses become scalar loads (and scalar arithmetic).
// Only for memory no kernel running at the same time writes.
typedef const __attribute__((address_space(4))) float KvConstF;
__device__ __forceinline__ KvConstF *kv_const(const float *p) { return (KvConstF *)(size_t)p; }


// IEEE min/max of two floats that are results of arithmetic (never
// signalling NaNs), without the operand quieting fminf/fmaxf add.
__device__ __forceinline__ float vmin(float a, float b)
{
    float r;
    asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ float vmax(float a, float b)
{
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// floor(x) as an integer in one instruction; x must lie well inside the int
// range (callers clamp it first).
__device__ __forceinline__ int cvt_flr(float x)
{
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// Backprojection of brick (bxi, byi, bzi) of depth zt <= ZT slices.  Called
// by all 256 threads of a workgroup (it synchronises).  SR band loads per
// thread are issued together before their LDS writes.
// FW (mode R): the geometry puts every z plane at alpha >= 1e-6 on every
// ray (KvGeom::bp_posz), so a positive weight w = H - L has L >= 1e-6 and is
// at least one FP32 ulp of 1e-6 (~1e-13): w > 1e-24 is then w > 0, and the
// accumulation takes the clamped weight max(w, +0) with no select (mode A,
// whose test is w > 0, always does).  The product is a legacy multiply,
// which gives 0 for a zero weight whatever the pixel value (a skipped
// pixel's value may be any float, INF and NaN included, as in the
// reference, which never reads it) and the IEEE product otherwise; adding
// +0 leaves the sums bit-identical (they are never -0).  The clamp is the
// subtraction's own [0, 1] output clamp: H <= 1 and L >= +0, so w <= 1 and
// clamping to [0, 1] is max(w, +0).
template <int MODE, int EP, int ZT, bool UNIT, int SR, bool FW>
__device__ __forceinline__ void bp_brick(const KvGeom &g, const KvTables &t, const float *__restrict__ proj,
                                         float *__restrict__ vol, const float *__restrict__ sens, int bxi, int byi,
                                         int bzi, int zt, BpShared<ZT> &sm)
{
    float *xb = sm.xb;
    int *xbase = sm.xbase;
    auto &accs = sm.accs;

    const int tid = threadIdx.x;
    // the wavefront index is uniform: its y row, table rows and detector
    // rows are scalar values
    const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int x0 = bxi * TX, y0 = byi * TY, z0 = bzi * zt;
    const int x = x0 + lane, y = y0 + wv;
    const int nzt = min(zt, g.nz - z0);
    const bool active = x < g.nx && y < g.ny;
    const float hx = (float)g.nx / 2.f;
    const float cvx = (float)x - hx;  // Projector.cu:437-440
    const int nx1 = g.nx + 1, ny1 = g.ny + 1, nz1 = g.nz + 1;
    const float scx = g.sx / g.pu;
    const float tol = MODE == KV_MODE_R ? 1e-24f : 0.f;
    const int half = 5;
    const float wx = (float)half * g.pu / g.sx;
    const float ipu = 1.f / g.pu;
    const bool fast_window = g.nu < (1 << 22) && g.nv < (1 << 22);
    const float pad = 0.02f;
    const float nuf = (float)g.nu;

    for (int i = 0; i < ZT; ++i)
        accs[i][0][tid] = accs[i][1][tid] = 0.f;

    for (int p = 0; p < g.np; ++p) {
        // Per-view terms, uniform over the workgroup and evaluated once per
        // geometry with the same FP32 operations (scalar loads): source x
        // over sx, detector corner and far edge x over sx, and the lateral
        // scales T of the brick's top and bottom z planes (the zs table).
        const KvConstF *C = kv_const(g.rc + KV_RC * p);
        const float sfx = C[0], ofx = C[8], efx = C[9];
        const KvConstF *Zg = kv_const(t.zs + ((size_t)p * nz1 + z0) * KV_ZS);
        const float ta = Zg[0], tb = Zg[nzt * KV_ZS];
        const float *tx = t.x + (size_t)p * nx1 * g.nu;
        const float *ty = t.y + (size_t)p * ny1 * g.nv;

        __syncthreads();  // the previous view's readers are done with the bands
        if (tid <= TX) {
            // band of plane k: shadows of planes k-1 and k+1 at both depths
            const int k = x0 + tid;
            float mn = kInf;
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const float c = ((float)(k - 1 + 2 * q) - hx) - sfx;
                mn = fminf(mn, fminf(c * ta, c * tb));
            }
            const float u = clampf((sfx + mn - ofx) * scx - 0.5f, -1e6f, 1e6f);
            xbase[tid] = (int)floorf(u) - 2;
        }
        __syncthreads();
        // Table bands through a bounds-checked buffer resource over this
        // view's x table: entries past the table read 0, and entries whose
        // detector column lies off the detector are never used (the band
        // slots read below always belong to columns in [0, nu)), so no range
        // test is needed.  A thread stages eight consecutive entries of one
        // band with two 16-byte loads issued together.
        {
            const __amdgpu_buffer_rsrc_t txs =
                __builtin_amdgcn_make_buffer_rsrc((void *)tx, (short)0, nx1 * g.nu * 4, 0x00020000);
            for (int e = tid; e < (TX + 1) * (XBW / 8); e += 256) {
                const int i = e / (XBW / 8), j = (e % (XBW / 8)) * 8;
                const unsigned k = (unsigned)(x0 + i), px = (unsigned)(xbase[i] + j);
                const int off = (int)((k * (unsigned)g.nu + px) * 4u);
                const auto v0 = __builtin_amdgcn_raw_buffer_load_b128(txs, off, 0, 0);
                const auto v1 = __builtin_amdgcn_raw_buffer_load_b128(txs, off + 16, 0, 0);
                float *d = xb + i * XBS + j;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    d[q] = __uint_as_float(v0[q]);
                    d[4 + q] = __uint_as_float(v1[q]);
                }
            }
        }
        __syncthreads();
        if (!active)
            continue;

        const float rdx = (cvx + 0.5f) - sfx;
        const int xb0 = xbase[lane], xb1 = xbase[lane + 1];
        const float *pr = proj + (size_t)p * g.nu * g.nv;
        const __amdgpu_buffer_rsrc_t prs =
            __builtin_amdgcn_make_buffer_rsrc((void *)pr, (short)0, g.nu * g.nv * 4, 0x00020000);
        const float *lx2 = MODE == KV_MODE_A ? t.lx2 + (size_t)p * g.nu : nullptr;
        const float *ly2 = MODE == KV_MODE_A ? t.ly2 + (size_t)p * g.nv : nullptr;
        const float lz2 = MODE == KV_MODE_A ? t.lz2[p] : 0.f;
        const float *xr0 = xb + lane * XBS, *xr1 = xb + (lane + 1) * XBS;
        // this wavefront's two y-plane rows of the view's y table and its
        // (y, z)-row terms: uniform addresses in memory this kernel does not
        // write, read through the constant address space (scalar loads)
        const KvConstF *ty0 = kv_const(ty + (size_t)y * g.nv), *ty1 = kv_const(ty + (size_t)(y + 1) * g.nv);
        const KvConstF *YZ = kv_const(t.yz + (((size_t)p * g.ny + y) * g.nz + z0) * KV_YZ);
        // Detector column of the voxel's lateral edges x = cvx and cvx + 1
        // on the plane at the top of slice zi (affine map of the per-view
        // table, bounds only); the bottom plane of slice zi is the top plane
        // of slice zi + 1, so each plane's pair is evaluated once.
        float ua0 = __builtin_fmaf(cvx, Zg[2], Zg[3]), ua1 = __builtin_fmaf(cvx + 1.f, Zg[2], Zg[3]);

        for (int zi = 0; zi < nzt; ++zi) {
            const KvConstF *Z = Zg + zi * KV_ZS;  // plane z0+zi; Z + KV_ZS is the next plane
            const float ub0 = __builtin_fmaf(cvx, Z[KV_ZS + 2], Z[KV_ZS + 3]);
            const float ub1 = __builtin_fmaf(cvx + 1.f, Z[KV_ZS + 2], Z[KV_ZS + 3]);
            const float mnx = vmin(ua0, ub0), mxx = vmax(ua1, ub1);
            ua0 = ub0;
            ua1 = ub1;
            // the wavefront's detector rows (empty when mode R's voxel
            // centre row lies off the widened detector) and z-plane terms
            const KvConstF *E = YZ + zi * KV_YZ;
            const int ly = __float_as_int(E[0]), hyw = __float_as_int(E[1]);
            const bool has_shadow = __float_as_int(E[2]) != 0;
            const float lz = E[4], uz = E[5], lzp = E[6];  // one 8-word scalar load with the above
            if (ly > hyw)
                continue;
            // Shadow of the voxel box on the detector columns: the columns
            // whose centres lie in [mnx - pad, mxx + pad] (see shadow_affine;
            // floor(a + 1) for ceil(a) only drops a column lying exactly on
            // the widened edge, which the padding keeps clear of every ray
            // that crosses the box).  rlo/rhi before the clamp to the
            // detector tell whether the shadow lies inside it.
            int fx0 = 0, fx1 = g.nu - 1;
            bool interior = false;
            if (has_shadow) {
                const int rlo = cvt_flr(clampf(mnx + (1.f - pad), -1.f, nuf + 2.f));
                const int rhi = cvt_flr(clampf(mxx + pad, -2.f, nuf + 1.f));
                fx0 = max(rlo, 0);
                fx1 = min(rhi, g.nu - 1);
                interior = rlo >= 0 && rhi <= g.nu - 1;
            }
            int lx = fx0, hxw = fx1;
            if (MODE == KV_MODE_R && !(interior && hxw - lx < 4)) {
                // Voxel centre on the detector and the +-5 window,
                // Projector.cu:473-492, 607-646 (az is the reference's
                // (ofz - sfz) / rdz, from the per-view table).  Skipped for
                // a shadow of at most four columns inside the detector: the
                // voxel centre projects into its shadow (a point of the box
                // projects inside the box's projection), so the centre lies
                // within the shadow's width (< 5 pixels) of every shadow
                // column, the window [pc - 5, pc + 5] contains the shadow and
                // the centre lies on the detector: both tests pass.
                const float ix = sfx + rdx * Z[1];
                if (!(ix > ofx - wx && ix < efx + wx))
                    continue;
                const float qx = (ix - ofx) * g.sx;
                // The window is [pc-5, pc+5] with pc = (int)(qx / pu).  When
                // the shadow fits inside it for every pc within one of an
                // estimate, the window changes nothing and the division is
                // skipped: |qx * (1/pu) - qx / pu| < 1/2 for the
                // |qx / pu| <= nu + 5 < 2^22 that pass the test above, so the
                // truncations differ by at most one.
                bool exact = !has_shadow || !fast_window;
                if (!exact) {
                    const int ex = (int)(qx * ipu);
                    exact = !(fx0 >= ex - 4 && fx1 <= ex + 4);
                }
                if (exact) {
                    const int pcx = (int)(qx / g.pu);
                    int wlx = pcx - half, whx = pcx + half;
                    wlx = wlx < 0 ? 0 : (wlx >= g.nu ? g.nu - 1 : wlx);
                    whx = whx < 0 ? 0 : (whx >= g.nu ? g.nu - 1 : whx);
                    lx = max(wlx, fx0);
                    hxw = min(whx, fx1);
                }
            }
            if (lx > hxw)
                continue;
            // one running sum per voxel across all views, as the reference
            // accumulates (Projector.cu:727-728)
            float a = accs[zi][0][tid], wsm = accs[zi][1][tid];
            const int ncol = hxw - lx + 1;
            // lx - xb >= 0 and hxw - xb < XBW for both planes' bands, as two
            // unsigned compares (meaningful for the ncol <= 4 it is used with)
            const bool in_band = (unsigned)(lx - xb0) <= (unsigned)(XBW - ncol) &&
                                 (unsigned)(lx - xb1) <= (unsigned)(XBW - ncol);
            // Columns lx .. lx + 3 all in both bands and on the detector, and
            // [lx, hxw] the whole shadow (not cut by the window or the
            // detector edge): the columns after hxw lie outside the widened
            // shadow, so their weights fail the acceptance test in every row
            // (w <= 0, exactly the skipped pixels of the shadow argument),
            // and they serve as the spare columns of a narrower range with no
            // per-column tests.
            const bool simple = interior && ncol <= 4 && lx + 3 <= g.nu - 1 &&
                                (unsigned)(lx - xb0) <= (unsigned)(XBW - 4) &&
                                (unsigned)(lx - xb1) <= (unsigned)(XBW - 4);
            // x intervals of the (at most NC) columns stay in registers; NC is
            // the wavefront's widest column range (2, 3 or 4), so narrow
            // shadows do not pay for four columns per row
            auto rows = [&](auto nc_tag, auto simple_tag) {
                constexpr int NC = decltype(nc_tag)::value;
                constexpr bool SIMPLE = decltype(simple_tag)::value;
                float Lx[NC], Hx[NC], L2[NC];
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    Lx[c] = 1.f;
                    Hx[c] = 0.f;  // empty: any weight <= -1
                    L2[c] = 0.f;
                    if (SIMPLE || c < ncol) {
                        const float a1x = xr0[lx + c - xb0], a2x = xr1[lx + c - xb1];
                        Lx[c] = bmax(lzp, bmin(a1x, a2x));
                        Hx[c] = bmin(uz, bmax(a1x, a2x));
                        if (MODE == KV_MODE_A && UNIT)
                            L2[c] = lx2[lx + c];
                    }
                }
                // Projection values through a bounds-checked buffer resource
                // over this view's plane: the columns of a row are one 32-bit
                // offset plus immediate offsets (one wide load), columns past
                // the row end read the next row (never used: their weight is
                // not positive) and reads past the plane return 0 instead of
                // faulting.  The next row's values are requested before the
                // current row is reduced, into two register sets that
                // alternate between rows (no copies).  (A resource rebuilt
                // per row on the scalar unit, to save the offset's vector
                // add, measured 13 % slower: the scalar unit issues at the
                // vector unit's rate and was the busier of the two.)
                float va[NC], vb[NC];
#pragma unroll
                for (int c = 0; c < NC; ++c)
                    va[c] = vb[c] = 0.f;
                constexpr bool LOADS = EP != KV_EP_COUNT && !UNIT;
                int off = (ly * g.nu + lx) * 4 - 4 * g.nu;
                auto load_row = [&](int py, float (&dst)[NC]) {
                    (void)py;
                    off += 4 * g.nu;
                    if (LOADS) {
#pragma unroll
                        for (int c = 0; c < NC; ++c)
                            dst[c] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(prs, off + 4 * c, 0, 0));
                    }
                };
                // ly_, hy_: the row's y interval from the wavefront's two
                // y-plane crossings (uniform: scalar min/max; exact, as the
                // reference's grouping fmax(fmax(0,mx),fmax(my,mz)) regroups
                // freely)
                auto row = [&](int py, const float (&v)[NC], float ly_, float hy_) {
                    const float lyy = (MODE == KV_MODE_A && UNIT) ? ly2[py] : 0.f;
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        const float hmin = bmin(Hx[c], hy_), lmax = bmax(Lx[c], ly_);
                        if (EP == KV_EP_COUNT) {
                            const float w = hmin - lmax;
                            wsm = w > tol ? wsm + 1.f : wsm;
                        } else if (MODE == KV_MODE_R && FW) {
                            const float wp = __builtin_amdgcn_fmed3f(hmin - lmax, 0.f, 1.f);
                            wsm = wsm + wp;
                            a = a + kv_mul_legacy(wp, v[c]);
                        } else if (MODE == KV_MODE_R) {
                            const float w = hmin - lmax;
                            const bool ok = w > tol;
                            wsm = ok ? wsm + w : wsm;
                            a = ok ? a + w * v[c] : a;
                        } else if (UNIT) {
                            const float wp = __builtin_amdgcn_fmed3f(hmin - lmax, 0.f, 1.f);
                            const float L = sqrtf((L2[c] + lyy) + lz2);
                            a = a + wp * L;
                        } else {
                            // input already multiplied by the ray's length
                            // factor (kv_launch_scale_a)
                            const float wp = __builtin_amdgcn_fmed3f(hmin - lmax, 0.f, 1.f);
                            a = a + kv_mul_legacy(wp, v[c]);
                        }
                    }
                };
                // The y-plane crossings of a row pair are loaded one pair
                // ahead (row indices clamped to the last row, so every load
                // stays in the table): the pair's scalar loads complete
                // while the previous pair is reduced, and the wait that the
                // first row of a pair needs finds them done.
                // The y intervals of a row pair come from scalar loads
                // issued one pair ahead, at the top of the previous pair's
                // turn, and reduced to (min, max) at its end, so the loads
                // complete while that pair is reduced (kept in place by
                // scheduling barriers: the loads of constant memory would
                // otherwise be re-issued where their values are used).  Rows
                // up to hyw + 2 <= nv + 1 may be read: they lie in the table,
                // or in the 16 bytes of padding after it, and are never
                // used.
                auto yint = [&](int r, float &lo, float &hi) {
                    const float a = ty0[r], b = ty1[r];
                    lo = bmin(a, b);
                    hi = bmax(a, b);
                };
                load_row(ly, va);
                int py = ly;
                float lya, hya, lyb, hyb;
                yint(py, lya, hya);
                yint(py + 1, lyb, hyb);
                for (; py < hyw; py += 2) {
                    const float n0a = ty0[py + 2], n1a = ty1[py + 2], n0b = ty0[py + 3], n1b = ty1[py + 3];
                    __builtin_amdgcn_sched_barrier(0);
                    load_row(py + 1, vb);
                    row(py, va, lya, hya);
                    load_row(py + 2, va);
                    row(py + 1, vb, lyb, hyb);
                    __builtin_amdgcn_sched_barrier(0);
                    lya = bmin(n0a, n1a);
                    hya = bmax(n0a, n1a);
                    lyb = bmin(n0b, n1b);
                    hyb = bmax(n0b, n1b);
                }
                if (py == hyw)
                    row(py, va, lya, hya);
            };
            auto rows_nc = [&](auto simple_tag) {
                if (__all(ncol <= 2))
                    rows(std::integral_constant<int, 2>{}, simple_tag);
                else if (__all(ncol <= 3))
                    rows(std::integral_constant<int, 3>{}, simple_tag);
                else
                    rows(std::integral_constant<int, 4>{}, simple_tag);
            };
            if (__all(simple)) {
                rows_nc(std::true_type{});
            } else if (ncol <= 4 && in_band) {
                rows_nc(std::false_type{});
            } else {
                for (int py = ly; py <= hyw; ++py) {
                    const float a1y = ty0[py], a2y = ty1[py];
                    const float lo = fmaxf(lz, fminf(a1y, a2y));
                    const float up = fminf(uz, fmaxf(a1y, a2y));
                    const float *prow = pr + (size_t)py * g.nu;
                    const float lyy = (MODE == KV_MODE_A && UNIT) ? ly2[py] : 0.f;
                    for (int px = lx; px <= hxw; ++px) {
                        const int j0 = px - xb0, j1 = px - xb1;
                        const float a1x = (j0 >= 0 && j0 < XBW) ? xr0[j0] : tx[(size_t)x * g.nu + px];
                        const float a2x = (j1 >= 0 && j1 < XBW) ? xr1[j1] : tx[(size_t)(x + 1) * g.nu + px];
                        const float w = fminf(up, fmaxf(a1x, a2x)) - fmaxf(lo, fminf(a1x, a2x));
                        if (w > tol) {
                            if (EP == KV_EP_COUNT) {
                                wsm += 1.f;
                            } else if (MODE == KV_MODE_R) {
                                wsm += w;
                                a += w * prow[px];
                            } else if (UNIT) {
                                a += w * sqrtf((lx2[px] + lyy) + lz2);
                            } else {
                                a += w * prow[px];
                            }
                        }
                    }
                }
            }
            accs[zi][0][tid] = a;
            accs[zi][1][tid] = wsm;
        }
    }
    if (!active)
        return;
    for (int zi = 0; zi < nzt; ++zi) {
        const size_t o = ((size_t)(z0 + zi) * g.ny + y) * g.nx + x;
        const float a = accs[zi][0][tid], wsm = accs[zi][1][tid];
        if (EP == KV_EP_COUNT) {
            vol[o] = wsm;
        } else if (MODE == KV_MODE_R) {
            const float bp = wsm > 1e-24f ? a / wsm : 0.f;  // Projector.cu:773-776
            if (EP == KV_EP_PLAIN)
                vol[o] = bp;
            else if (EP == KV_EP_MLEM)
                vol[o] = vol[o] * bp;  // Reconstructors.h:238-241
            else
                vol[o] = vol[o] + bp;  // Reconstructors.h:118-121
        } else {
            if (EP == KV_EP_PLAIN) {
                vol[o] = a;
            } else {
                const float s = sens[o];
                vol[o] = s > 0.f ? (vol[o] * a) / s : vol[o];
            }
        }
    }
}

constexpr int kZT = 8;

// out[p][v][u] = proj[p][v][u] * sqrt((lx2[p][u] + ly2[p][v]) + lz2[p]): each
// ray's length factor, applied once per pixel.  The mode-A backprojector's
// weight of pixel (u, v) for a voxel is the alpha-length times this factor
// times the pixel value; the product factor * value is the same FP32 value
// whether formed here or per (voxel, pixel) pair.
__global__ __launch_bounds__(256) void scale_a_kernel(KvGeom g, KvTables t, const float *__restrict__ proj,
                                                      float *__restrict__ out)
{
    const int64_t plane = (int64_t)g.nu * g.nv;
    const int64_t n = plane * g.np;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int p = (int)(i / plane);
        const int64_t r = i - p * plane;
        const int v = (int)(r / g.nu), u = (int)(r - (int64_t)v * g.nu);
        const float L = sqrtf((t.lx2[(size_t)p * g.nu + u] + t.ly2[(size_t)p * g.nv + v]) + t.lz2[p]);
        out[i] = L * proj[i];
    }
}

// One brick per workgroup.
template <int MODE, int EP, int ZT, bool UNIT, bool FW>
__global__ __launch_bounds__(256, 5) void bp_tiled_kernel(KvGeom g, KvTables t, const float *__restrict__ proj,
                                                       float *__restrict__ vol, const float *__restrict__ sens)
{
    __shared__ BpShared<ZT> sm;
    // rounds of 3 loads: mode A measured 3 % faster; mode R's registers would
    // spill with them
    bp_brick<MODE, EP, ZT, UNIT, MODE == KV_MODE_A ? 3 : 1, FW>(g, t, proj, vol, sens, blockIdx.x, blockIdx.y,
                                                               blockIdx.z, ZT, sm);
}

// Persistent backprojection + MLEM update over a batch of reconstructions:
// workgroups claim bricks from a global counter until the list is exhausted.
// Mode A divides by each reconstruction's sensitivity A^T 1.
template <int MODE>
__global__ __launch_bounds__(256, 5) void bp_batch_kernel(const KvProblem *__restrict__ probs, int nprob,
                                                       long long total, unsigned long long *counter, int iteration)
{
    __shared__ BpShared<kZT> sm;
    __shared__ long long item_s;
    for (;;) {
        if (threadIdx.x == 0)
            item_s = (long long)atomicAdd(counter, 1ull);
        __syncthreads();
        const long long item = item_s;
        __syncthreads();
        if (item >= total)
            break;
        int pi = 0;
        while (pi + 1 < nprob && probs[pi + 1].bp_first <= item)
            ++pi;
        const KvProblem &P = probs[pi];
        if (iteration >= P.iters)
            continue;
        const long long local = item - P.bp_first;
        const int per_z = P.bp_bx * P.bp_by;
        const int bz = (int)(local / per_z);
        const int rem = (int)(local - (long long)bz * per_z);
        const int by = rem / P.bp_bx, bx = rem - by * P.bp_bx;
        bp_brick<MODE, KV_EP_MLEM, kZT, false, 1, false>(P.g, P.t, P.ratio, P.vol, P.sens, bx, by, bz, P.bp_zt,
                                                         sm);
    }
}

__global__ void zero_u64_kernel(unsigned long long *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0ull;
}

template <int MODE, int EP, bool UNIT>
hipError_t launch_tiled(const KvGeom &g, const KvTables &t, const float *proj, float *vol, const float *sens,
                        hipStream_t s)
{
    const dim3 grid((g.nx + TX - 1) / TX, (g.ny + TY - 1) / TY, (g.nz + kZT - 1) / kZT), block(256);
    if (MODE == KV_MODE_R && (EP == KV_EP_MLEM || EP == KV_EP_SART || EP == KV_EP_PLAIN) && g.bp_posz)
        bp_tiled_kernel<MODE, EP, kZT, UNIT, true><<<grid, block, 0, s>>>(g, t, proj, vol, sens);
    else
        bp_tiled_kernel<MODE, EP, kZT, UNIT, false><<<grid, block, 0, s>>>(g, t, proj, vol, sens);
    return hipGetLastError();
}

}  // namespace

hipError_t kv_launch_bp_batch(int mode, const KvProblem *probs, int nprob, long long total,
                              unsigned long long *counter, int iteration, int grid, hipStream_t s)
{
    if (mode == KV_MODE_A)
        bp_batch_kernel<KV_MODE_A><<<grid, 256, 0, s>>>(probs, nprob, total, counter, iteration);
    else
        bp_batch_kernel<KV_MODE_R><<<grid, 256, 0, s>>>(probs, nprob, total, counter, iteration);
    return hipGetLastError();
}

hipError_t kv_launch_zero_u64(unsigned long long *p, int n, hipStream_t s)
{
    zero_u64_kernel<<<(n + 255) / 256, 256, 0, s>>>(p, n);
    return hipGetLastError();
}

int kv_bp_batch_blocks_per_cu(int mode)
{
    int n = 0;
    const hipError_t e =
        mode == KV_MODE_A ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, bp_batch_kernel<KV_MODE_A>, 256, 0)
                          : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, bp_batch_kernel<KV_MODE_R>, 256, 0);
    if (e != hipSuccess)
        return 2;
    return n > 0 ? n : 1;
}

hipError_t kv_launch_tables_r(const KvGeom &g, const KvTables &t, hipStream_t s)
{
    tables_r_kernel<<<2048, 256, 0, s>>>(g, t);
    const int n = g.np * (g.nz + 1);
    zs_kernel<<<(n + 255) / 256, 256, 0, s>>>(g, t.zs);
    yz_kernel<KV_MODE_R><<<1024, 256, 0, s>>>(g, t);
    return hipGetLastError();
}

hipError_t kv_launch_tables_a(const KvGeom &g, const KvTables &t, hipStream_t s)
{
    tables_a_kernel<<<2048, 256, 0, s>>>(g, t);
    const int n = g.np * (g.nz + 1);
    zs_kernel<<<(n + 255) / 256, 256, 0, s>>>(g, t.zs);
    yz_kernel<KV_MODE_A><<<1024, 256, 0, s>>>(g, t);
    return hipGetLastError();
}

hipError_t kv_launch_bp_r(const KvGeom &g, const KvTables &t, const float *proj, float *vol, int epilogue,
                          hipStream_t s)
{
    switch (epilogue) {
    case KV_EP_PLAIN: return launch_tiled<KV_MODE_R, KV_EP_PLAIN, false>(g, t, proj, vol, nullptr, s);
    case KV_EP_MLEM: return launch_tiled<KV_MODE_R, KV_EP_MLEM, false>(g, t, proj, vol, nullptr, s);
    case KV_EP_COUNT: return launch_tiled<KV_MODE_R, KV_EP_COUNT, true>(g, t, nullptr, vol, nullptr, s);
    default: return launch_tiled<KV_MODE_R, KV_EP_SART, false>(g, t, proj, vol, nullptr, s);
    }
}

hipError_t kv_launch_scale_a(const KvGeom &g, const KvTables &t, const float *proj, float *out, hipStream_t s)
{
    const int64_t n = kv_nray(g);
    const int nb = (int)std::min<int64_t>(8192, (n + 255) / 256);
    scale_a_kernel<<<std::max(nb, 1), 256, 0, s>>>(g, t, proj, out);
    return hipGetLastError();
}

hipError_t kv_launch_bp_a(const KvGeom &g, const KvTables &t, const float *proj, float *vol, int epilogue,
                          const float *sens, bool unit_input, hipStream_t s)
{
    if (epilogue == KV_EP_COUNT)
        return launch_tiled<KV_MODE_A, KV_EP_COUNT, true>(g, t, nullptr, vol, nullptr, s);
    if (epilogue == KV_EP_PLAIN && unit_input)
        return launch_tiled<KV_MODE_A, KV_EP_PLAIN, true>(g, t, nullptr, vol, nullptr, s);
    if (epilogue == KV_EP_PLAIN)
        return launch_tiled<KV_MODE_A, KV_EP_PLAIN, false>(g, t, proj, vol, nullptr, s);
    return launch_tiled<KV_MODE_A, KV_EP_MLEM, false>(g, t, proj, vol, sens, s);
}
