// Forward projectors for gfx950.
//
// One ray per lane; a 256-thread workgroup is four wavefronts, each covering
// a (64/L) x L tile of detector pixels (8 x 8 for the plain walk; 16 x 4 and
// 32 x 2 for the walks that gather two and three volumes per step), so the
// rays of a wavefront stay close in u and v and their voxel gathers at each
// step fall on nearby lines of the x-fastest volume.  Workgroup tiles are
// launched in a one-dimensional order that puts stripes of 32 detector
// columns on the eight XCDs in turn (stripe i on XCD i % 8, each XCD walking
// its stripes column-first), so every XCD's L2 holds one band of the volume.
//
// Mode R reproduces CudaPerformProjection (KevLoq/Kevernals
// source/modules/reconstruction/Projector.cu:227-372) bit for bit.  The
// reference fills a per-thread 2000-float array with 1.0, writes three
// arithmetic runs of plane crossings into it and selection-sorts it
// (FindProjectionAlphas_d, Projector.cu:45-223).  Each run is non-decreasing
// (a[i] = a[i-1] + |1/d| in FP32), so the sorted list is exactly the merge of
// the x run, the y run, the (truncated) z run and the leftover 1.0 fillers:
// here the four runs are merged on the fly in registers, with no scratch
// array and O(n) work instead of O(n^2).
//
// All arithmetic that decides voxel indices is evaluated as the reference
// writes it, with contraction off.
#pragma clang fp contract(off)

#include <algorithm>

#include "kv_internal.h"

// The structured-buffer load intrinsic (no clang builtin spells it):
// rsrc, record index, byte offset within the record, scalar offset, cache
// policy.
typedef int KvRsrc __attribute__((ext_vector_type(4)));
typedef float KvF32x2 __attribute__((ext_vector_type(2)));
typedef float KvF32x4 __attribute__((ext_vector_type(4)));
__device__ float kv_struct_load_f32(KvRsrc rsrc, int vindex, int voffset, int soffset, int aux) __asm(
    "llvm.amdgcn.struct.buffer.load.f32");
__device__ KvF32x2 kv_struct_load_f32x2(KvRsrc rsrc, int vindex, int voffset, int soffset, int aux) __asm(
    "llvm.amdgcn.struct.buffer.load.v2f32");
__device__ KvF32x4 kv_struct_load_f32x4(KvRsrc rsrc, int vindex, int voffset, int soffset, int aux) __asm(
    "llvm.amdgcn.struct.buffer.load.v4f32");

namespace {

constexpr float kInf = __builtin_inff();

// Projector.cu:22-42
__device__ __forceinline__ bool feq(float a, float b, float t)
{
    if (t < 0)
        return (a - b) < -t && (a - b) > t;
    return (a - b) < t && (a - b) > -t;
}
__device__ __forceinline__ bool fsupeq(float a, float b, float t) { return a > b || feq(a, b, t); }

struct RayF {
    float sx, sy, sz;    // source, floating-voxel units
    float dx, dy, dz;    // source -> pixel centre
    float hx, hy, hz;    // half sizes
    float a1x, aNx, ivx; // x-face alphas (0 when |dx| <= 1e-6) and 1/dx (0 when dx = 0)
    float a1y, aNy, ivy; // the same for y
    float a1z, aNz, stz; // z-face alphas and 1/dz (0 when |dz| <= 1e-6)
};

// Projector.cu:303-323.  The terms that depend on the view alone, and those
// of the pixel's detector column and row (its centre coordinate, the face
// alphas and the reciprocal of the direction along that axis), come from the
// geometry's ray constants: the same FP32 operations evaluated once on the
// host.
__device__ __forceinline__ void ray_setup(const KvGeom &g, int p, int u, int v, RayF &r)
{
    const float *C = g.rc + KV_RC * p;
    r.sx = C[0];
    r.sy = C[1];
    r.sz = C[2];
    const float4 cx = *(const float4 *)(g.rc + (size_t)g.np * KV_RC + ((size_t)p * g.nu + u) * 4);
    const float4 cy = *(const float4 *)(g.rc + (size_t)g.np * (KV_RC + 4 * g.nu) + ((size_t)p * g.nv + v) * 4);
    r.dx = cx.x - r.sx;
    r.dy = cy.x - r.sy;
    r.a1x = cx.y;
    r.aNx = cx.z;
    r.ivx = cx.w;
    r.a1y = cy.y;
    r.aNy = cy.z;
    r.ivy = cy.w;
    r.dz = C[3];
    r.hx = (float)g.nx / 2.f;
    r.hy = (float)g.ny / 2.f;
    r.hz = (float)g.nz / 2.f;
    r.a1z = C[4];
    r.aNz = C[5];
    r.stz = C[6];
}

// Projector.cu:88-120
__device__ __forceinline__ void axis_range(float d, float s, float h, int n, float amin, float amax, int &lo, int &hi)
{
    int mx, mn;
    if (d > 0) {
        mx = max(0, min(n - 1, (int)(h + s + amax * d)));
        mn = max(0, min(mx, n - (int)(h - s - amin * d)));
    } else {
        mx = max(0, min(n - 1, (int)(h + s + amin * d)));
        mn = max(0, min(mx, n - (int)(h - s - amax * d)));
    }
    lo = mn;
    hi = mx;
}

// a * b + c on the low 24 bits of a and b, one instruction (the compiler
// would otherwise form two products and a three-way add); b is
// wavefront-uniform (a scalar operand).
__device__ __forceinline__ unsigned mad_u24(unsigned a, unsigned b, unsigned c)
{
    unsigned r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
    return r;
}

// a in the lanes whose bit of mask is set, else b: one v_cndmask with a
// scalar mask (lanes outside the mask's active set are not written).
__device__ __forceinline__ float selm(unsigned long long mask, float a, float b)
{
    float r;
    asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(b), "v"(a), "s"(mask));
    return r;
}

// min(1, a) as one instruction (a is never NaN where it is used).
__device__ __forceinline__ float min1(float a)
{
    float r;
    asm("v_min_f32 %0, 1.0, %1" : "=v"(r) : "v"(a));
    return r;
}

// Minimum of three run heads as one v_min3_f32.  The heads are finite or
// +INF (never NaN), so the IEEE-mode operand canonicalisation the compiler
// would add for fminf is not needed.
__device__ __forceinline__ float min3f(float a, float b, float c)
{
    float m;
    asm("v_min3_f32 %0, %1, %2, %3" : "=v"(m) : "v"(a), "v"(b), "v"(c));
    return m;
}

// Volume gathers by voxel index: a structured buffer of 4-byte records,
// bounds-checked by index (an index at or past the voxel count reads 0), so
// the gather takes the linear index with no byte-offset shift.
__device__ __forceinline__ KvRsrc voxel_rsrc(const KvGeom &g, const float *vol, unsigned stride = 4)
{
    const unsigned long long a = (unsigned long long)vol;
    KvRsrc r;
    r.x = (int)(unsigned)a;
    r.y = (int)(((unsigned)(a >> 32) & 0xFFFFu) | (stride << 16));  // record stride in bytes
    r.z = g.nx * g.ny * g.nz;                                        // records (at most 2^30)
    r.w = 0x00020000;
    return r;
}

// Mode-R walk of one ray.  visit(ok, voxel_index, weight) is called for every
// segment of the sorted list in order; ok marks the segments the reference
// accepts (Projector.cu:356-366).
// FAST (the geometry's walk_fast flag, set by kv_set_geometry): nx * ny <
// 2^24, so an accepted segment's index vz*(nx*ny) + vy*nx + vx uses full-rate
// 24-bit multiplies (a rejected segment's index is never dereferenced
// outside a bounds-checked buffer access); and for every axis
// (n + 2) * (max|source| + n/2 + 2) < 2^22 in voxel units.  A run's values are
// a0 + j*inc accumulated in FP32 with |a| <= (|s| + n/2) * inc, so after j <= n+1
// additions the accumulated rounding error stays below inc/4: the value
// T = a0 + (c - 1/2) * inc (rounded once to FP32) then lies strictly
// between the run's last value and the one after it, and "exhausted" is one
// comparison with T instead of a count.  Without FAST each run counts down.
template <bool FAST, class Visit>
__device__ __forceinline__ void walk_r(const KvGeom &g, const RayF &r, Visit &&visit)
{
    const float tol = 1e-6f;
    const bool okx = fabsf(r.dx) > tol, oky = fabsf(r.dy) > tol, okz = fabsf(r.dz) > tol;
    // okx ? (-hx - sx) / dx : 0 and the like, from the ray constants
    const float a1x = r.a1x, a1y = r.a1y, a1z = r.a1z;
    const float aNx = r.aNx, aNy = r.aNy, aNz = r.aNz;
    const float amin = fmaxf(fmaxf(0.f, fminf(a1x, aNx)), fmaxf(fminf(a1y, aNy), fminf(a1z, aNz)));
    const float amax = fminf(fminf(1.f, fmaxf(a1x, aNx)), fminf(fmaxf(a1y, aNy), fmaxf(a1z, aNz)));
    if (fsupeq(amin, amax, 1e-7f))
        return;  // the reference's early exit (Projector.cu:79-85): empty list

    int lox, hix, loy, hiy, loz, hiz;
    axis_range(r.dx, r.sx, r.hx, g.nx, amin, amax, lox, hix);
    axis_range(r.dy, r.sy, r.hy, g.ny, amin, amax, loy, hiy);
    axis_range(r.dz, r.sz, r.hz, g.nz, amin, amax, loz, hiz);
    const int n = (hix - lox) + (hiy - loy) + (hiz - loz) + 3;

    // Run heads, lengths and increments (Projector.cu:139-201).  A run whose
    // axis fails |d| > tol leaves its slots at the 1.0 filler.
    int cx = 0, cy = 0, cz = 0;
    float ax = 0.f, ay = 0.f, az = 0.f, ix = 0.f, iy = 0.f, iz = 0.f;
    if (okx) {
        const float st = r.ivx;  // 1 / dx
        ax = st * (-r.hx - r.sx + (float)(r.dx > 0 ? lox : hix));
        ix = fabsf(st);
        cx = hix - lox + 1;
    }
    if (oky) {
        const float st = r.ivy;  // 1 / dy
        ay = st * (-r.hy - r.sy + (float)(r.dy > 0 ? loy : hiy));
        iy = fabsf(st);
        cy = hiy - loy + 1;
    }
    if (okz) {
        // z run written over (hiz-loz)-1 slots, at least its first entry
        // (Projector.cu:181-201).
        const float st = r.stz;  // 1 / dz
        az = st * (-r.hz - r.sz + (float)(r.dz > 0 ? loz : hiz));
        iz = fabsf(st);
        const int dzr = hiz - loz;
        cz = dzr >= 2 ? dzr - 1 : 1;
    }
    const int cf = n - cx - cy - cz;  // 1.0 fillers left unsorted-over
    // T = a0 + (c - 1/2) * inc rounded once to FP32 (one fma: c - 1/2 is
    // exact): |T| < 2^22 / (n + 2) * inc under FAST, so the rounding moves T
    // by less than inc / (4 (n + 2)), far inside the inc / 4 margins above
    // (any T in the band separates the same run values).
    const float tx = FAST ? __builtin_fmaf((float)cx - 0.5f, ix, ax) : 0.f;
    const float ty = FAST ? __builtin_fmaf((float)cy - 0.5f, iy, ay) : 0.f;
    const float tz = FAST ? __builtin_fmaf((float)cz - 0.5f, iz, az) : 0.f;

    // A run's head holds +INF once the run is exhausted, so the next list
    // entry is the plain minimum of the heads.  All heads equal to the
    // minimum are consumed together: equal entries of the sorted list only
    // bound zero-width segments, which Projector.cu:358-359 rejects (w must
    // exceed 1e-6), so the accepted segments and their order are those of
    // the full list.  The 1.0 fillers form one entry for the same reason,
    // merged in between two loops over the runs (values below 1.0, then the
    // rare values above it).
    if (cx == 0) ax = kInf;
    if (cy == 0) ay = kInf;
    if (cz == 0) az = kInf;
    // Steps of the main loops, where m is the minimum head (FAST): only the
    // first run in x, y, z order whose head equals m advances.  A tie
    // between runs is then consumed over several steps, each extra step a
    // zero-width segment that Projector.cu:358-359 rejects, so the accepted
    // segments are unchanged and one run's increment and end threshold are
    // selected instead of stepping all three.
    auto advance1 = [&](float m) {
        // Two vector compares of bit patterns give the lane masks of "x
        // holds m" and "y holds m"; the scalar unit combines them into the
        // three run masks and the selects take them directly (left to
        // itself the compiler negates a mask with a third vector compare).
        // m is one of the three heads, so its bits equal those of the first
        // run that holds it; heads that equal m as floats with other bits
        // (+0 and -0) only order the consumption of equal list entries,
        // which bound zero-width segments.
        const unsigned mb = __float_as_uint(m);
        const unsigned long long mx = __builtin_amdgcn_ballot_w64(__float_as_uint(ax) == mb);
        const unsigned long long my = __builtin_amdgcn_ballot_w64(__float_as_uint(ay) == mb);
        const unsigned long long mpy = my & ~mx, mpz = ~(mx | my);
        const float inc = selm(mx, ix, selm(my, iy, iz));
        const float thr = selm(mx, tx, selm(my, ty, tz));
        float nv = m + inc;
        nv = nv > thr ? kInf : nv;
        ax = selm(mx, nv, ax);
        ay = selm(mpy, nv, ay);
        az = selm(mpz, nv, az);
    };
    // a[i] = a[i-1] + |1/d| in FP32 (Projector.cu:156,177,199)
    auto advance = [&](float m) {
        const bool px = ax == m, py = ay == m, pz = az == m;
        if (FAST) {
            const float nx_ = ax + ix, ny_ = ay + iy, nz_ = az + iz;
            ax = px ? (nx_ > tx ? kInf : nx_) : ax;
            ay = py ? (ny_ > ty ? kInf : ny_) : ay;
            az = pz ? (nz_ > tz ? kInf : nz_) : az;
        } else {
            cx -= px;
            cy -= py;
            cz -= pz;
            ax = px ? (cx == 0 ? kInf : ax + ix) : ax;
            ay = py ? (cy == 0 ? kInf : ay + iy) : ay;
            az = pz ? (cz == 0 ? kInf : az + iz) : az;
        }
    };

    // Segment (prev, cur), Projector.cu:348-367.  The midpoint is the FP32
    // sum divided by 2.0 in FP64.  (double)float * (double)float is exact,
    // and so is the halving, so each index expression equals one FP64 fma of
    // the FP32 sum with the halved direction.  visit(ok, k, w): ok marks an
    // accepted segment; k is its voxel index when ok and meaningless
    // (possibly out of range) otherwise.
    const double bx = (double)(r.hx + r.sx);
    const double by = (double)(r.hy + r.sy);
    const double bz = (double)(r.hz + r.sz);
    const double hdx = 0.5 * (double)r.dx, hdy = 0.5 * (double)r.dy, hdz = 0.5 * (double)r.dz;
    const unsigned nx = (unsigned)g.nx, ny = (unsigned)g.ny, nz = (unsigned)g.nz, nxy = nx * ny;
    auto segment = [&](float prev, float cur) {
        const float w = cur - prev;
        const double sum = (double)(cur + prev);
        const int vx = (int)__builtin_fma(sum, hdx, bx);
        const int vy = (int)__builtin_fma(sum, hdy, by);
        const int vz = (int)__builtin_fma(sum, hdz, bz);
        const bool ok = (unsigned)vx < nx && (unsigned)vy < ny && (unsigned)vz < nz && w > tol;
        const unsigned k = FAST ? mad_u24((unsigned)vz, nxy, mad_u24((unsigned)vy, nx, (unsigned)vx))
                                : ((unsigned)vz * ny + (unsigned)vy) * nx + (unsigned)vx;
        visit(ok, k, w);
    };

    const float fill = cf > 0 ? 1.0f : kInf;
    // first entry
    float prev = min3f(ax, ay, az);
    bool filler_left = cf > 0;
    if (fill <= prev) {
        prev = fill;
        filler_left = false;
    }
    // entries below the fillers' value (all of them when there is no filler);
    // two entries per turn, so that the previous entry needs no copy
    const float lim = filler_left ? fill : kInf;
    // The entries below lim are known in advance when every run of two or
    // more entries ends below lim (its values lie below its end threshold):
    // a run of one entry contributes it if it lies below lim.  Each step of
    // advance1 then takes exactly one of them, so the loop below runs a fixed
    // count of steps: those entries, less the ones the first entry takes
    // (every run head equal to it).  Run entries at or above lim stay for
    // the part after the fillers.
    auto ends_below = [&](int c, float t) { return c <= 1 || t < lim; };
    auto below = [&](int c, float a) { return c == 1 ? (int)(a < lim) : c; };
    const bool counted = FAST && ends_below(cx, tx) && ends_below(cy, ty) && ends_below(cz, tz);
    const int left = (below(cx, ax) + below(cy, ay) + below(cz, az)) - ((ax == prev) + (ay == prev) + (az == prev));
    advance(prev);
    if (counted) {
        // pairs of steps under a wavefront-uniform turn counter (the
        // per-lane count only masks lanes), then the odd step
        const int pairs = left >> 1;
        for (int it = 0;; ++it) {
            const bool act = it < pairs;
            if (!__any(act))
                break;
            if (act) {
                const float c1 = min3f(ax, ay, az);
                advance1(c1);
                segment(prev, c1);
                const float c2 = min3f(ax, ay, az);
                advance1(c2);
                segment(c1, c2);
                prev = c2;
            }
        }
        if (left & 1) {
            const float c1 = min3f(ax, ay, az);
            advance1(c1);
            segment(prev, c1);
            prev = c1;
        }
    } else if (FAST) {
        for (;;) {
            const float c1 = min3f(ax, ay, az);
            if (!(c1 < lim))
                break;
            advance1(c1);
            segment(prev, c1);
            const float c2 = min3f(ax, ay, az);
            if (!(c2 < lim)) {
                prev = c1;
                break;
            }
            advance1(c2);
            segment(c1, c2);
            prev = c2;
        }
    } else {
        for (;;) {
            const float cur = min3f(ax, ay, az);
            if (!(cur < lim))
                break;
            advance(cur);
            segment(prev, cur);
            prev = cur;
        }
    }
    if (filler_left) {
        // the 1.0 entry (run values equal to it are the same entry), then
        // run values above it
        advance(fill);
        segment(prev, fill);
        prev = fill;
        for (;;) {
            const float cur = min3f(ax, ay, az);
            if (cur == kInf)
                break;
            advance(cur);
            segment(prev, cur);
            prev = cur;
        }
    }
}

// AlmostEqualRelative(v, 0) of commons/Maths.h:25-34 in FP32.
// One ray of the mode-R forward pass with its epilogue.
template <int EP>
__device__ __forceinline__ void fp_r_ray(const KvGeom &g, const float *__restrict__ vol, float *__restrict__ out,
                                         const float *__restrict__ data, float relax, int p, int u, int v)
{
    RayF r;
    ray_setup(g, p, u, v, r);
    const size_t o = ((size_t)p * g.nv + v) * g.nu + u;
    if (EP == KV_EP_COUNT) {
        int c = 0;
        walk_r<false>(g, r, [&](bool ok, unsigned, float) { c += ok; });
        out[o] = (float)c;
        return;
    }
    float tot = 0.f, tw = 0.f;
    // Branch-free accumulation: every segment gathers through a
    // bounds-checked buffer resource over the volume (an offset outside it
    // reads 0; a rejected segment's index may lie outside it).  Each segment is added one step after its gather is issued, so the
    // load is in flight while the next segment is walked; the sums take
    // the segments in the same order.  A rejected segment carries weight +0
    // and adds +0 to tw and kv_mul_legacy(+0, value) = +-0 to tot (for any
    // value, INF and NaN included); neither sum is ever -0 (they start at
    // +0 and x + (-x) rounds to +0), so both are left bit-identical.
    const KvRsrc vrs = voxel_rsrc(g, vol);
    float pw = 0.f, pv = 0.f;
    auto acc = [&](bool ok, unsigned k, float w) {
        const float val = kv_struct_load_f32(vrs, (int)k, 0, 0, 0);
        tw = tw + pw;
        tot = tot + kv_mul_legacy(pw, pv);
        pw = ok ? w : 0.f;
        pv = val;
    };
    if (g.walk_fast)
        walk_r<true>(g, r, acc);
    else
        walk_r<false>(g, r, acc);
    tw = tw + pw;
    tot = tot + kv_mul_legacy(pw, pv);
    const float fp = tw > 1e-6f ? tot / tw : 0.f;
    if (EP == KV_EP_PLAIN) {
        out[o] = fp;
    } else if (EP == KV_EP_MLEM) {
        // Reconstructors.h:207-213: the ratio overwrites the projection
        // unless it is (almost) zero, where the projection value stays.
        out[o] = almost_equal_zero(fp) ? fp : relax * (data[o] / fp);
    } else {
        // Reconstructors.h:90-93
        out[o] = relax * (data[o] - fp);
    }
}

// Pixel of lane `tid` in forward tile (tx, ty) for tile layout L (1, 2, 4 or
// 8): each wavefront covers 64/L consecutive detector columns of L rows, the
// four wavefronts of a workgroup are stacked along v; a tile is (64/L) x 4L
// pixels.  64/L is a power of two: the lane splits by mask and shift.
__device__ __forceinline__ void fp_tile_pixel(int layout, int tx, int ty, int tid, int &u, int &v)
{
    const int lane = tid & 63, wv = tid >> 6;
    const int wu = 64 / layout;  // uniform
    const int lw = __builtin_ctz((unsigned)wu);
    u = tx * wu + (lane & (wu - 1));
    v = ty * (4 * layout) + wv * layout + (lane >> lw);
}

// n / d for n < 2^31 with the divisor's precomputed multiplier (KvUDiv); on
// wavefront-uniform operands it runs on the scalar unit.
__device__ __forceinline__ unsigned udiv(unsigned n, KvUDiv d) { return (__umulhi(n, d.m) + n) >> d.l; }

// Forward tile of this workgroup in an XCD-aware order.  Workgroups are
// handed to the eight XCDs round-robin by linear id.  The launch is
// one-dimensional over gxp x gy x np tiles; tile columns are grouped in
// stripes of `chunk` tiles, and each XCD walks its stripes column-fastest,
// then rows, then views.  In tile row r (rows and views counted together)
// stripe i goes to XCD (i - r) mod 8.  A stripe keeps an XCD's L2 on a few
// neighbouring detector columns (round-robin single tiles of an 8-column
// layout made every L2 fetch the same volume lines, 43.5 GB per full-size
// launch), and the interleaving balances the eight XCDs' shares of the
// uneven ray lengths (one contiguous band per XCD left the outer bands
// idle: 11.4 ms against 8.8 ms).  Rotating the assignment by one stripe per
// tile row gives every XCD every stripe equally often: with a fixed
// i mod 8 two XCDs carried 7 % more than the mean at full size, as the
// detector's columns hold ray lengths in a V profile cut off at the edges
// of the volume's shadow.  gxp = tile columns rounded up to a multiple of
// 8 x chunk; tiles of padding columns (u past the detector) exit at the
// kernel's bounds test.  A launch covers views p0, p0 + 1, ...  The
// divisions by chunk, the stripe count and gy use precomputed multipliers
// (they were a tenth of the vector instructions a wavefront issues before
// its walk).
__device__ __forceinline__ void fp_xcd_tile(const KvFpMap &fm, int &tx, int &ty, int &p)
{
    constexpr unsigned X = KV_XCDS;
    const unsigned w = blockIdx.x;
    const unsigned j = w / X;
    const unsigned jc = udiv(j, fm.dchunk), t = j - jc * (unsigned)fm.chunk;
    const unsigned r = udiv(jc, fm.dnch), ci = jc - r * (unsigned)fm.nch;
    tx = (int)((ci * X + (w + r) % X) * (unsigned)fm.chunk + t);
    const unsigned q = udiv(r, fm.dgy);
    ty = (int)(r - q * (unsigned)fm.gy);
    p = fm.p0 + (int)q;
}

template <int EP>
__global__ __launch_bounds__(256) void fp_r_kernel(KvGeom g, const float *__restrict__ vol, float *__restrict__ out,
                                                   const float *__restrict__ data, float relax, KvFpMap fm)
{
    int tx, ty, p, u, v;
    fp_xcd_tile(fm, tx, ty, p);
    fp_tile_pixel(fm.layout, tx, ty, threadIdx.x, u, v);
    if (u >= g.nu || v >= g.nv)
        return;
    fp_r_ray<EP>(g, vol, out, data, relax, p, u, v);
}

// Persistent forward pass over a batch of reconstructions.  The tiles sit in
// KV_XCDS lists (KvProblem), list k holding the stripes an XCD-ordered
// launch gives XCD k (fp_xcd_tile); `total` items per list, one counter per
// list.  A workgroup claims tiles from the list of the XCD a one-dimensional
// launch places it on (blockIdx.x % KV_XCDS), then from the other lists in
// turn, until every list is exhausted: every tile is processed once whatever
// the placement, which only decides the caches' locality.  Problems whose
// iteration count is reached skip their tiles; padding columns (u past the
// detector) are skipped at the bounds test.
template <typename Ray>
__device__ __forceinline__ void fp_batch_tiles(const KvProblem *__restrict__ probs, int nprob, long long total,
                                               unsigned long long *counters, int iteration, Ray &&ray)
{
    __shared__ long long item_s;
    const int own = (int)(blockIdx.x % (unsigned)KV_XCDS);
    for (int c = 0; c < KV_XCDS; ++c) {
        const int k = (own + c) % KV_XCDS;
        for (;;) {
            if (threadIdx.x == 0)
                item_s = (long long)atomicAdd(counters + k, 1ull);
            __syncthreads();
            const long long item = item_s;
            __syncthreads();
            if (item >= total)
                break;
            int pi = 0;
            while (pi + 1 < nprob && probs[pi + 1].fp_first <= item)
                ++pi;
            const KvProblem &P = probs[pi];
            if (iteration >= P.iters)
                continue;
            // item indices within a reconstruction stay far below 2^31 (it
            // has fewer than 2^31 rays, 256 per tile): 32-bit divisions by
            // the precomputed multipliers
            const unsigned local = (unsigned)(item - P.fp_first);
            const unsigned ch = (unsigned)P.fp_chunk, nch = (unsigned)P.fp_nch;
            const unsigned jc = udiv(local, P.fp_dchunk), r = udiv(jc, P.fp_dnch);
            const int tx = (int)(((jc - r * nch) * KV_XCDS + (unsigned)k) * ch + (local - jc * ch));
            const unsigned q = udiv(r, P.fp_dty);
            const int ty = (int)(r - q * (unsigned)P.fp_ty), p = (int)q;
            int u, v;
            fp_tile_pixel(P.fp_layout, tx, ty, threadIdx.x, u, v);
            if (u < P.g.nu && v < P.g.nv)
                ray(P, p, u, v);
        }
    }
}

template <int EP>
__global__ __launch_bounds__(256) void fp_r_batch_kernel(const KvProblem *__restrict__ probs, int nprob,
                                                         long long total, unsigned long long *counters, int iteration,
                                                         float relax)
{
    fp_batch_tiles(probs, nprob, total, counters, iteration, [&](const KvProblem &P, int p, int u, int v) {
        fp_r_ray<EP>(P.g, P.vol, P.ratio, P.data, relax, p, u, v);
    });
}

// ---------------------------------------------------------------- mode A --
struct RayA {
    float s[3], d[3], inv[3], h[3];
    int n[3];
    int fixed[3];
    float L;
};

__device__ __forceinline__ void ray_setup_a(const KvGeom &g, int p, int u, int v, RayA &r)
{
    RayF f;
    ray_setup(g, p, u, v, f);
    r.s[0] = f.sx; r.s[1] = f.sy; r.s[2] = f.sz;
    r.d[0] = f.dx; r.d[1] = f.dy; r.d[2] = f.dz;
    r.h[0] = f.hx; r.h[1] = f.hy; r.h[2] = f.hz;
    r.n[0] = g.nx; r.n[1] = g.ny; r.n[2] = g.nz;
    r.inv[0] = f.ivx;
    r.inv[1] = f.ivy;
    r.inv[2] = r.d[2] != 0.f ? 1.f / r.d[2] : 0.f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        r.fixed[c] = 0;
        if (r.d[c] == 0.f) {
            const float pos = r.s[c] + r.h[c];
            r.fixed[c] = (pos > 0.f && pos < (float)r.n[c]) ? (int)floorf(pos) : -1;
        }
    }
    const float lx = r.d[0] * g.sx, ly = r.d[1] * g.sy, lz = r.d[2] * g.sz;
    r.L = sqrtf(lx * lx + ly * ly + lz * lz);
}

__device__ __forceinline__ float plane_a(const RayA &r, int c, int k)
{
    return (((float)k - r.h[c]) - r.s[c]) * r.inv[c];
}

// Maximum of three values as one v_max3_f32 (operands finite or -INF, never
// NaN: no canonicalisation needed).
__device__ __forceinline__ float max3f(float a, float b, float c)
{
    float m;
    asm("v_max3_f32 %0, %1, %2, %3" : "=v"(m) : "v"(a), "v"(b), "v"(c));
    return m;
}

// Mode-A Siddon walk: visit(ok, voxel, len) for every voxel the ray's
// parameter interval [ain, aout] passes through, ok = len > 0 (voxel is the
// linear index; in range for every call).
// Each axis keeps the crossing of its next plane, hi = ((k - N/2) - s) / d
// evaluated as plane_a does, with k - N/2 carried as an exact float (an
// integer or half-integer far below 2^24), the crossing of the last plane
// passed (lo) and the voxel index folded into one linear index.  The axis
// that exits first (first minimum on ties, as the oracle) is found from the
// minimum of the three heads; an axis without motion holds hi = +INF.  The
// walk ends when the nearest exit reaches aout: the exit through an axis's
// boundary plane is that plane's crossing, which is where aout comes from,
// so no index is ever stepped outside the volume.
template <class Visit>
__device__ __forceinline__ void walk_a(const RayA &r, Visit &&visit)
{
    float ain = 0.f, aout = 1.f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (r.d[c] == 0.f) {
            if (r.fixed[c] < 0)
                return;
            continue;
        }
        const float a0 = plane_a(r, c, 0), aN = plane_a(r, c, r.n[c]);
        ain = fmaxf(ain, fminf(a0, aN));
        aout = fminf(aout, fmaxf(a0, aN));
    }
    if (!(ain < aout))
        return;
    int ii[3];
    float lo[3], hi[3], kf[3], st[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (r.d[c] == 0.f) {
            ii[c] = r.fixed[c];
            lo[c] = -kInf;
            hi[c] = kInf;
            kf[c] = 0.f;
            st[c] = 0.f;
            continue;
        }
        const float pos = r.s[c] + r.h[c] + ain * r.d[c];
        int i = (int)floorf(pos);
        i = max(0, min(r.n[c] - 1, i));
        const int stp = r.d[c] > 0.f ? 1 : -1;
        // lo/hi of voxel i on axis c
        auto lohi = [&](int k, float &l, float &h) {
            const float a0 = plane_a(r, c, k), a1 = plane_a(r, c, k + 1);
            l = fminf(a0, a1);
            h = fmaxf(a0, a1);
        };
        float l, h;
        for (;;) {
            lohi(i, l, h);
            const int j = i - stp;
            if (l > ain && j >= 0 && j < r.n[c]) i = j; else break;
        }
        for (;;) {
            lohi(i, l, h);
            const int j = i + stp;
            if (h <= ain && j >= 0 && j < r.n[c]) i = j; else break;
        }
        ii[c] = i;
        lo[c] = l;
        hi[c] = h;
        // the plane whose crossing is hi: k = i + 1 going up, i going down
        kf[c] = (float)(stp > 0 ? i + 1 : i) - r.h[c];
        st[c] = (float)stp;
    }
    const int nx = r.n[0], nxy = r.n[0] * r.n[1];
    const int sx = (int)st[0], sy = (int)st[1] * nx, sz = (int)st[2] * nxy;
    unsigned idx = (unsigned)(ii[2] * nxy + ii[1] * nx + ii[0]);
    const int guard = r.n[0] + r.n[1] + r.n[2] + 3;
    // max(0, lo[0], lo[1], lo[2]) of the current voxel.  The exiting axis's
    // lo becomes its old hi, which is the step's minimum head and at least
    // its old lo, so the new maximum is max(old maximum, that minimum): one
    // running max, the same value as the maximum over the three axes (IEEE
    // max of non-NaN values is associative).
    float ent = fmaxf(0.f, max3f(lo[0], lo[1], lo[2]));
    // One step: visit the segment through the current voxel, then true at
    // the walk's end, else move to the next voxel.  The exit axis's lane
    // masks come from two vector compares (the scalar unit forms the first-
    // minimum masks), and the selects take them directly.  Two steps per
    // turn of the guard loop: the guard only bounds the walk (a ray crosses
    // fewer planes than it counts), so one test per two steps will do.
    auto step = [&]() {
        const float best = min3f(hi[0], hi[1], hi[2]);
        const float len = min1(best) - ent;
        visit(len > 0.f, idx, len);
        if (best >= aout)
            return true;
        ent = max3f(ent, best, best);
        const unsigned long long m0 = __builtin_amdgcn_ballot_w64(hi[0] == best);
        const unsigned long long m1 = __builtin_amdgcn_ballot_w64(hi[1] == best);
        const unsigned long long p1 = m1 & ~m0, p2 = ~(m0 | m1);
        idx += (unsigned)__float_as_int(
            selm(m0, __int_as_float(sx), selm(p1, __int_as_float(sy), __int_as_float(sz))));
        // the exiting axis: its next plane (same operation sequence as plane_a)
        const float k0 = kf[0] + st[0], k1 = kf[1] + st[1], k2 = kf[2] + st[2];
        hi[0] = selm(m0, (k0 - r.s[0]) * r.inv[0], hi[0]);
        hi[1] = selm(p1, (k1 - r.s[1]) * r.inv[1], hi[1]);
        hi[2] = selm(p2, (k2 - r.s[2]) * r.inv[2], hi[2]);
        kf[0] = selm(m0, k0, kf[0]);
        kf[1] = selm(p1, k1, kf[1]);
        kf[2] = selm(p2, k2, kf[2]);
        return false;
    };
    for (int n = 0; n < guard; n += 2) {
        if (step())
            break;
        if (step())
            break;
    }
}

// Sum over a mode-A walk of len * vol[voxel] for the accepted segments, the
// value of acc = ok ? acc + len * v : acc, with no branch around the load:
// every step gathers its voxel (the walk's indices are always in range) and
// a rejected segment carries weight +0, whose legacy product is +-0 for any
// voxel value (INF and NaN included) and leaves acc unchanged (acc starts
// at +0 and never becomes -0); an accepted weight is positive, where the
// legacy product is the IEEE one.  Each product is added one step later, so
// that the load is in flight while the next step is walked.
struct AccA {
    KvRsrc rs;
    float acc = 0.f, pw = 0.f, pv = 0.f;
    __device__ __forceinline__ void init(const KvGeom &g, const float *vol) { rs = voxel_rsrc(g, vol); }
    __device__ __forceinline__ void visit(bool ok, unsigned k, float len)
    {
        const float v = kv_struct_load_f32(rs, (int)k, 0, 0, 0);
        acc = acc + kv_mul_legacy(pw, pv);
        pw = ok ? len : 0.f;
        pv = v;
    }
    __device__ __forceinline__ float sum() const { return acc + kv_mul_legacy(pw, pv); }
};

// One ray of the mode-A forward pass with its epilogue.
template <int EP>
__device__ __forceinline__ void fp_a_ray(const KvGeom &g, const KvTables &t, const float *__restrict__ vol,
                                         float *__restrict__ out, const float *__restrict__ data, int p, int u, int v)
{
    RayA r;
    ray_setup_a(g, p, u, v, r);
    const size_t o = ((size_t)p * g.nv + v) * g.nu + u;
    if (EP == KV_EP_COUNT) {
        int c = 0;
        walk_a(r, [&](bool ok, unsigned, float) { c += ok; });
        out[o] = (float)c;
        return;
    }
    AccA a;
    a.init(g, vol);
    walk_a(r, [&](bool ok, unsigned k, float len) { a.visit(ok, k, len); });
    const float ax = r.L * a.sum();
    if (EP == KV_EP_PLAIN) {
        out[o] = ax;
    } else {
        // MLEM ratio on line-integral data, stored premultiplied by the
        // ray's length factor as the mode-A backprojector weighs it (its
        // tables' sqrt((lx2 + ly2) + lz2), the same FP32 product it would
        // form per pixel)
        const float L = sqrtf((t.lx2[(size_t)p * g.nu + u] + t.ly2[(size_t)p * g.nv + v]) + t.lz2[p]);
        out[o] = ax > 0.f ? L * (data[o] / ax) : 0.f;
    }
}

template <int EP>
__global__ __launch_bounds__(256) void fp_a_kernel(KvGeom g, KvTables t, const float *__restrict__ vol,
                                                   float *__restrict__ out, const float *__restrict__ data, KvFpMap fm)
{
    int tx, ty, p, u, v;
    fp_xcd_tile(fm, tx, ty, p);
    fp_tile_pixel(fm.layout, tx, ty, threadIdx.x, u, v);
    if (u >= g.nu || v >= g.nv)
        return;
    fp_a_ray<EP>(g, t, vol, out, data, p, u, v);
}

// Mode-A forward projections of NV <= 3 volumes in one walk per ray: one
// volume gathered as fp_a_ray does, two or three interleaved per voxel
// (pack_kernel: records of 2 or 4 floats) so that one gather per step reads
// them all.  Each volume's sum takes the same products and additions in the
// same order as fp_a_ray, so every projection is bit-identical to a
// single-volume pass.
struct KvMulti {
    const float *vol[KV_LL_MAX];
    float *out[KV_LL_MAX];
};

// The sums of a walk over NV = 2 or 3 interleaved volumes, as AccA's.
template <int NV>
struct AccPack {
    static constexpr int W = NV == 2 ? 2 : 4;  // floats per record
    KvRsrc rs;
    float acc[NV], pv[NV];
    float pw = 0.f;
    __device__ __forceinline__ void init(const KvGeom &g, const float *pack)
    {
        rs = voxel_rsrc(g, pack, 4 * W);
#pragma unroll
        for (int i = 0; i < NV; ++i)
            acc[i] = pv[i] = 0.f;
    }
    __device__ __forceinline__ void visit(bool ok, unsigned k, float len)
    {
        float v[NV];
        if constexpr (NV == 2) {
            const KvF32x2 r = kv_struct_load_f32x2(rs, (int)k, 0, 0, 0);
            v[0] = r.x, v[1] = r.y;
        } else {
            const KvF32x4 r = kv_struct_load_f32x4(rs, (int)k, 0, 0, 0);
            v[0] = r.x, v[1] = r.y, v[2] = r.z;
        }
#pragma unroll
        for (int i = 0; i < NV; ++i)
            acc[i] = acc[i] + kv_mul_legacy(pw, pv[i]);
        pw = ok ? len : 0.f;
#pragma unroll
        for (int i = 0; i < NV; ++i)
            pv[i] = v[i];
    }
    __device__ __forceinline__ float sum(int i) const { return acc[i] + kv_mul_legacy(pw, pv[i]); }
};

// Volumes m.vol[0..NV) interleaved per voxel: float2 records for two, float4
// records (fourth float 0) for three.
template <int NV>
__global__ __launch_bounds__(256) void pack_kernel(KvMulti m, int64_t n, float *__restrict__ pack)
{
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        if constexpr (NV == 2) {
            const float2 r = make_float2(m.vol[0][v], m.vol[1][v]);
            reinterpret_cast<float2 *>(pack)[v] = r;
        } else {
            const float4 r = make_float4(m.vol[0][v], m.vol[1][v], m.vol[2][v], 0.f);
            reinterpret_cast<float4 *>(pack)[v] = r;
        }
    }
}

template <int NV>
__global__ __launch_bounds__(256) void fp_a_multi_kernel(KvGeom g, const float *__restrict__ pack, KvMulti m,
                                                         KvFpMap fm)
{
    int tx, ty, p, u, v;
    fp_xcd_tile(fm, tx, ty, p);
    fp_tile_pixel(fm.layout, tx, ty, threadIdx.x, u, v);
    if (u >= g.nu || v >= g.nv)
        return;
    RayA r;
    ray_setup_a(g, p, u, v, r);
    const size_t o = ((size_t)p * g.nv + v) * g.nu + u;
    if constexpr (NV == 1) {
        AccA a;
        a.init(g, m.vol[0]);
        walk_a(r, [&](bool ok, unsigned k, float len) { a.visit(ok, k, len); });
        m.out[0][o] = r.L * a.sum();
    } else {
        AccPack<NV> a;
        a.init(g, pack);
        walk_a(r, [&](bool ok, unsigned k, float len) { a.visit(ok, k, len); });
#pragma unroll
        for (int i = 0; i < NV; ++i)
            m.out[i][o] = r.L * a.sum(i);
    }
}

// Persistent mode-A forward pass (MLEM ratio) over a batch, as
// fp_r_batch_kernel.
__global__ __launch_bounds__(256) void fp_a_batch_kernel(const KvProblem *__restrict__ probs, int nprob,
                                                         long long total, unsigned long long *counters, int iteration)
{
    fp_batch_tiles(probs, nprob, total, counters, iteration, [&](const KvProblem &P, int p, int u, int v) {
        fp_a_ray<KV_EP_MLEM>(P.g, P.t, P.vol, P.ratio, P.data, p, u, v);
    });
}

// Records the accepted pairs of sampled rays through the same walks.
__global__ void trace_kernel(KvGeom g, int mode, int nrays, const int32_t *__restrict__ rays, int cap,
                             int64_t *__restrict__ idx, float *__restrict__ w, int32_t *__restrict__ counts)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrays)
        return;
    const int id = rays[t];
    const int per = g.nu * g.nv;
    const int p = id / per, rem = id - p * per, v = rem / g.nu, u = rem - v * g.nu;
    int c = 0;
    auto rec = [&](bool ok, unsigned k, float wt) {
        if (!ok)
            return;
        if (c < cap) {
            idx[(size_t)t * cap + c] = (int64_t)k;
            w[(size_t)t * cap + c] = wt;
        }
        ++c;
    };
    if (mode == KV_MODE_R) {
        RayF r;
        ray_setup(g, p, u, v, r);
        walk_r<false>(g, r, rec);
    } else {
        RayA r;
        ray_setup_a(g, p, u, v, r);
        walk_a(r, rec);
    }
    counts[t] = c;
}

}  // namespace

// One-dimensional launches of the single-reconstruction forward kernels over
// gxp x gy tiles per view (fp_xcd_tile), as many views per launch as keep a
// launch within 2^32 threads.
struct FpGrid {
    int gxp, gy, chunk, views;
    KvFpMap fm{};
    FpGrid(const KvGeom &g, int layout)
    {
        const int wu = 64 / layout, wv = 4 * layout;
        chunk = std::max(1, KV_FP_STRIPE / wu);
        gxp = ((g.nu + wu - 1) / wu + KV_XCDS * chunk - 1) / (KV_XCDS * chunk) * (KV_XCDS * chunk);
        gy = (g.nv + wv - 1) / wv;
        const unsigned long long per_view = (unsigned long long)gxp * (unsigned long long)gy * 256ull;
        views = (int)std::min<unsigned long long>((unsigned long long)g.np, 0xFFFFFFFFull / per_view);
        fm.layout = layout;
        fm.chunk = chunk;
        fm.nch = gxp / (KV_XCDS * chunk);
        fm.gy = gy;
        fm.dchunk = kv_udiv(chunk);
        fm.dnch = kv_udiv(fm.nch);
        fm.dgy = kv_udiv(gy);
    }
    KvFpMap map(int p0) const
    {
        KvFpMap m = fm;
        m.p0 = p0;
        return m;
    }
    // Runs launch(blocks, p0) for each range of views; a view too large for
    // one launch is an invalid configuration.
    template <typename F>
    hipError_t run(const KvGeom &g, F &&launch) const
    {
        if (views < 1)
            return hipErrorInvalidConfiguration;
        for (int p0 = 0; p0 < g.np; p0 += views) {
            const int nv = std::min(views, g.np - p0);
            launch((unsigned)gxp * (unsigned)gy * (unsigned)nv, p0);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess)
                return e;
        }
        return hipSuccess;
    }
};

hipError_t kv_launch_fp_r(const KvGeom &g, const float *vol, float *out, int epilogue, const float *data, float relax,
                          hipStream_t s)
{
    const int L = KV_FP_LAYOUT;
    const FpGrid fg(g, L);
    return fg.run(g, [&](unsigned nb, int p0) {
        switch (epilogue) {
        case KV_EP_PLAIN:
            fp_r_kernel<KV_EP_PLAIN><<<nb, 256, 0, s>>>(g, vol, out, data, relax, fg.map(p0));
            break;
        case KV_EP_MLEM:
            fp_r_kernel<KV_EP_MLEM><<<nb, 256, 0, s>>>(g, vol, out, data, relax, fg.map(p0));
            break;
        case KV_EP_COUNT:
            fp_r_kernel<KV_EP_COUNT><<<nb, 256, 0, s>>>(g, vol, out, data, relax, fg.map(p0));
            break;
        default:
            fp_r_kernel<KV_EP_SART><<<nb, 256, 0, s>>>(g, vol, out, data, relax, fg.map(p0));
            break;
        }
    });
}

hipError_t kv_launch_fp_a_multi(const KvGeom &g, int n, const float *const *vols, float *pack, float *const *outs,
                                hipStream_t s)
{
    KvMulti m{};
    for (int i = 0; i < n; ++i)
        m.vol[i] = vols[i], m.out[i] = outs[i];
    if (n >= 2) {
        const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
        const unsigned nb = (unsigned)std::max<int64_t>(1, std::min<int64_t>(8192, (nvox + 255) / 256));
        if (n == 2)
            pack_kernel<2><<<nb, 256, 0, s>>>(m, nvox, pack);
        else
            pack_kernel<3><<<nb, 256, 0, s>>>(m, nvox, pack);
        if (hipError_t e = hipGetLastError())
            return e;
    }
    const int L = KV_FP_LAYOUT_MULTI[std::min(std::max(n, 1), KV_LL_MAX) - 1];
    const FpGrid fg(g, L);
    return fg.run(g, [&](unsigned nb, int p0) {
        switch (n) {
        case 1: fp_a_multi_kernel<1><<<nb, 256, 0, s>>>(g, pack, m, fg.map(p0)); break;
        case 2: fp_a_multi_kernel<2><<<nb, 256, 0, s>>>(g, pack, m, fg.map(p0)); break;
        default: fp_a_multi_kernel<3><<<nb, 256, 0, s>>>(g, pack, m, fg.map(p0)); break;
        }
    });
}

hipError_t kv_launch_fp_a(const KvGeom &g, const KvTables &t, const float *vol, float *out, int epilogue,
                          const float *data, hipStream_t s)
{
    const int L = KV_FP_LAYOUT;
    const FpGrid fg(g, L);
    return fg.run(g, [&](unsigned nb, int p0) {
        if (epilogue == KV_EP_PLAIN)
            fp_a_kernel<KV_EP_PLAIN><<<nb, 256, 0, s>>>(g, t, vol, out, data, fg.map(p0));
        else if (epilogue == KV_EP_COUNT)
            fp_a_kernel<KV_EP_COUNT><<<nb, 256, 0, s>>>(g, t, vol, out, data, fg.map(p0));
        else
            fp_a_kernel<KV_EP_MLEM><<<nb, 256, 0, s>>>(g, t, vol, out, data, fg.map(p0));
    });
}

hipError_t kv_launch_trace(const KvGeom &g, int mode, int nrays, const int32_t *rays, int cap, int64_t *idx, float *w,
                           int32_t *counts, hipStream_t s)
{
    trace_kernel<<<(nrays + 63) / 64, 64, 0, s>>>(g, mode, nrays, rays, cap, idx, w, counts);
    return hipGetLastError();
}

hipError_t kv_launch_fp_batch(int mode, const KvProblem *probs, int nprob, long long total,
                              unsigned long long *counters, int iteration, float relax, int grid, hipStream_t s)
{
    if (mode == KV_MODE_A)
        fp_a_batch_kernel<<<grid, 256, 0, s>>>(probs, nprob, total, counters, iteration);
    else
        fp_r_batch_kernel<KV_EP_MLEM><<<grid, 256, 0, s>>>(probs, nprob, total, counters, iteration, relax);
    return hipGetLastError();
}

int kv_fp_batch_blocks_per_cu(int mode)
{
    int n = 0;
    const hipError_t e = mode == KV_MODE_A
                             ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fp_a_batch_kernel, 256, 0)
                             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fp_r_batch_kernel<KV_EP_MLEM>, 256, 0);
    if (e != hipSuccess)
        return 4;
    return n > 0 ? n : 1;
}
