// Backprojectors for gfx950: voxel-driven gathers, one voxel per lane, a
// wavefront covering 64 consecutive x of one (y, z) row.  Every voxel sums its
// own contributions in a fixed order (views, detector rows, detector columns,
// as Projector.cu:452-729), so results are deterministic without atomics.
//
// The reference evaluates, per voxel and view, all pixels of an 11x11 window
// with two FP32 divisions each (Projector.cu:656-669).  Two observations make
// this cheap without changing a single bit of the result:
//  * alpha1.x / alpha2.x depend only on (view, detector column, voxel x), and
//    alpha2.x of voxel x equals alpha1.x of voxel x+1 (k - nx/2 + 1 is exact in
//    FP32), so they are one table per view: x[p][u][k] with k = 0..nx.  The
//    same holds for y (per row) and z (per view).
//  * A window pixel contributes only when its ray crosses the voxel box with
//    positive alpha-length (weight > 1e-24, Projector.cu:710-725).  The loop
//    therefore runs over the window intersected with the box's projected
//    footprint (corner shadow plus one pixel of margin); the skipped pixels are
//    exactly those whose weight test fails, and the accumulation order of the
//    remaining ones is unchanged.
#pragma clang fp contract(off)

#include "kv_internal.h"

namespace {

constexpr float kInf = __builtin_inff();

// ------------------------------------------------------------ tables -----
// Mode R entries, Projector.cu:651-669.
__global__ void tables_r_kernel(KvGeom g, KvTables t)
{
    const int64_t nxe = (int64_t)g.np * g.nu * (g.nx + 1);
    const int64_t nye = (int64_t)g.np * g.nv * (g.ny + 1);
    const int64_t nze = (int64_t)g.np * (g.nz + 1);
    const float tol = 1e-24f;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nxe + nye + nze;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (i < nxe) {
            const int k = (int)(i % (g.nx + 1));
            const int64_t r = i / (g.nx + 1);
            const int u = (int)(r % g.nu), p = (int)(r / g.nu);
            const float sfx = g.src[3 * p] / g.sx;
            const float ppx = (g.det[3 * p] + ((float)u + 0.5f) * g.pu) / g.sx;
            const float rx = ppx - sfx;
            const float cv = (float)k - (float)g.nx / 2.f;
            t.x[i] = fabsf(rx) > tol ? (cv - sfx) / rx : 0.f;
        } else if (i < nxe + nye) {
            const int64_t j = i - nxe;
            const int k = (int)(j % (g.ny + 1));
            const int64_t r = j / (g.ny + 1);
            const int v = (int)(r % g.nv), p = (int)(r / g.nv);
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
    const int64_t nxe = (int64_t)g.np * g.nu * (g.nx + 1);
    const int64_t nye = (int64_t)g.np * g.nv * (g.ny + 1);
    const int64_t nze = (int64_t)g.np * (g.nz + 1);
    const float hx = (float)g.nx / 2.f, hy = (float)g.ny / 2.f, hz = (float)g.nz / 2.f;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nxe + nye + nze;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (i < nxe) {
            const int k = (int)(i % (g.nx + 1));
            const int64_t r = i / (g.nx + 1);
            const int u = (int)(r % g.nu), p = (int)(r / g.nu);
            const float s = g.src[3 * p] / g.sx;
            const float q = (g.det[3 * p] + ((float)u + 0.5f) * g.pu) / g.sx;
            t.x[i] = plane_a_entry(s, q, hx, g.nx, k);
            if (k == 0) {
                const float l = (q - s) * g.sx;
                t.lx2[r] = l * l;
            }
        } else if (i < nxe + nye) {
            const int64_t j = i - nxe;
            const int k = (int)(j % (g.ny + 1));
            const int64_t r = j / (g.ny + 1);
            const int v = (int)(r % g.nv), p = (int)(r / g.nv);
            const float s = g.src[3 * p + 1] / g.sy;
            const float q = (g.det[3 * p + 1] + ((float)v + 0.5f) * g.pv) / g.sy;
            t.y[j] = plane_a_entry(s, q, hy, g.ny, k);
            if (k == 0) {
                const float l = (q - s) * g.sy;
                t.ly2[r] = l * l;
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

// Pixel range [lo, hi] whose ray can cross voxel box [c, c+1] x [cz, cz+1]
// of one axis pair (lateral axis, z), all in floating-voxel units: shadow of
// the four corners on the detector plane at depth `oz`, one pixel of margin.
// `scale` = voxel spacing / pixel pitch.  Returns false when the box is not
// strictly between source and detector plane (caller uses the full range).
__device__ __forceinline__ bool shadow_range(float c, float cz, float s, float sz, float o, float oz, float scale, int n,
                                             int &lo, int &hi)
{
    const float d0 = cz - sz, d1 = (cz + 1.f) - sz, dd = oz - sz;
    const float t0 = dd / d0, t1 = dd / d1;
    if (!(t0 > 0.f && t1 > 0.f && t0 < 1e30f && t1 < 1e30f))
        return false;
    const float e0 = s + (c - s) * t0, e1 = s + ((c + 1.f) - s) * t0;
    const float e2 = s + (c - s) * t1, e3 = s + ((c + 1.f) - s) * t1;
    const float mn = fminf(fminf(e0, e1), fminf(e2, e3));
    const float mx = fmaxf(fmaxf(e0, e1), fmaxf(e2, e3));
    const float a = (mn - o) * scale - 0.5f, b = (mx - o) * scale - 0.5f;
    lo = (int)fminf(fmaxf(floorf(a) - 1.f, -1.f), (float)n);
    hi = (int)fmaxf(fminf(ceilf(b) + 1.f, (float)n), -1.f);
    lo = max(lo, 0);
    hi = min(hi, n - 1);
    return true;
}

// ------------------------------------------------------------ mode R -----
template <int EP>
__global__ __launch_bounds__(256) void bp_r_kernel(KvGeom g, KvTables t, const float *__restrict__ proj,
                                                   float *__restrict__ vol)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    if (x >= g.nx || y >= g.ny)
        return;
    const float tol = 1e-24f;
    const int half = 5;
    // Projector.cu:437-440
    const float cvx = (float)x - (float)g.nx / 2.f;
    const float cvy = (float)y - (float)g.ny / 2.f;
    const float cvz = (float)z - (float)g.nz / 2.f;
    const int nx1 = g.nx + 1, ny1 = g.ny + 1, nz1 = g.nz + 1;
    const float wx = (float)half * g.pu / g.sx, wy = (float)half * g.pv / g.sy;
    const float scx = g.sx / g.pu, scy = g.sy / g.pv;
    float tot = 0.f, tw = 0.f;
    for (int p = 0; p < g.np; ++p) {
        const float *S = g.src + 3 * p;
        const float *D = g.det + 3 * p;
        const float sfx = S[0] / g.sx, sfy = S[1] / g.sy, sfz = S[2] / g.sz;
        const float ofx = D[0] / g.sx, ofy = D[1] / g.sy, ofz = D[2] / g.sz;
        const float efx = (D[0] + (float)g.nu * g.pu) / g.sx;
        const float efy = (D[1] + (float)g.nv * g.pv) / g.sy;
        // voxel centre on the detector plane, Projector.cu:473-492
        const float rdx = (cvx + 0.5f) - sfx, rdy = (cvy + 0.5f) - sfy, rdz = (cvz + 0.5f) - sfz;
        const float az = fabsf(rdz) > tol ? (ofz - sfz) / rdz : 0.f;
        const float ix = sfx + rdx * az;
        const float iy = sfy + rdy * az;
        if (!(ix > ofx - wx && ix < efx + wx && iy > ofy - wy && iy < efy + wy))
            continue;
        // +-5 window, Projector.cu:607-646
        const int pcx = (int)((ix - ofx) * g.sx / g.pu);
        const int pcy = (int)((iy - ofy) * g.sy / g.pv);
        int lx = pcx - half, hx = pcx + half, ly = pcy - half, hy = pcy + half;
        lx = lx < 0 ? 0 : (lx >= g.nu ? g.nu - 1 : lx);
        hx = hx < 0 ? 0 : (hx >= g.nu ? g.nu - 1 : hx);
        ly = ly < 0 ? 0 : (ly >= g.nv ? g.nv - 1 : ly);
        hy = hy < 0 ? 0 : (hy >= g.nv ? g.nv - 1 : hy);
        // footprint of the voxel box
        int fx0, fx1, fy0, fy1;
        if (shadow_range(cvx, cvz, sfx, sfz, ofx, ofz, scx, g.nu, fx0, fx1) &&
            shadow_range(cvy, cvz, sfy, sfz, ofy, ofz, scy, g.nv, fy0, fy1)) {
            lx = max(lx, fx0);
            hx = min(hx, fx1);
            ly = max(ly, fy0);
            hy = min(hy, fy1);
        }
        const float a1z = t.z[p * nz1 + z], a2z = t.z[p * nz1 + z + 1];
        const float mnz = fminf(a1z, a2z), mxz = fmaxf(a1z, a2z);
        const float *ty = t.y + (size_t)p * g.nv * ny1 + y;
        const float *tx = t.x + (size_t)p * g.nu * nx1 + x;
        const float *pr = proj + (size_t)p * g.nu * g.nv;
        for (int py = ly; py <= hy; ++py) {
            const float a1y = ty[(size_t)py * ny1], a2y = ty[(size_t)py * ny1 + 1];
            // max/min are exact, so the reference's grouping
            // fmax(fmax(0, mx), fmax(my, mz)) may be regrouped freely.
            const float lo = fmaxf(fmaxf(0.f, fminf(a1y, a2y)), mnz);
            const float up = fminf(fminf(1.f, fmaxf(a1y, a2y)), mxz);
            const float *prow = pr + (size_t)py * g.nu;
            for (int px = lx; px <= hx; ++px) {
                const float a1x = tx[(size_t)px * nx1], a2x = tx[(size_t)px * nx1 + 1];
                const float amin = fmaxf(lo, fminf(a1x, a2x));
                const float amax = fminf(up, fmaxf(a1x, a2x));
                // floatSupEq(amax, amin) && amax - amin > 1e-24 reduces to
                // the second test alone (Projector.cu:710-725)
                const float w = amax - amin;
                if (w > tol) {
                    if (EP == KV_EP_COUNT) {
                        tw += 1.f;
                    } else {
                        tw += w;
                        tot += w * prow[px];
                    }
                }
            }
        }
    }
    const size_t o = ((size_t)z * g.ny + y) * g.nx + x;
    if (EP == KV_EP_COUNT) {
        vol[o] = tw;
        return;
    }
    const float bp = tw > tol ? tot / tw : 0.f;  // Projector.cu:773-776
    if (EP == KV_EP_PLAIN)
        vol[o] = bp;
    else if (EP == KV_EP_MLEM)
        vol[o] = vol[o] * bp;  // Reconstructors.h:238-241
    else
        vol[o] = vol[o] + bp;  // Reconstructors.h:118-121
}

// ------------------------------------------------------------ mode A -----
// UNIT: backproject a projection of ones (the sensitivity A^T 1) without
// reading `proj` at all.
template <int EP, bool UNIT>
__global__ __launch_bounds__(256) void bp_a_kernel(KvGeom g, KvTables t, const float *__restrict__ proj,
                                                   float *__restrict__ vol, const float *__restrict__ sens)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    if (x >= g.nx || y >= g.ny)
        return;
    const float cvx = (float)x - (float)g.nx / 2.f;
    const float cvy = (float)y - (float)g.ny / 2.f;
    const float cvz = (float)z - (float)g.nz / 2.f;
    const int nx1 = g.nx + 1, ny1 = g.ny + 1, nz1 = g.nz + 1;
    const float scx = g.sx / g.pu, scy = g.sy / g.pv;
    float acc = 0.f;
    for (int p = 0; p < g.np; ++p) {
        const float *S = g.src + 3 * p;
        const float *D = g.det + 3 * p;
        const float sfx = S[0] / g.sx, sfy = S[1] / g.sy, sfz = S[2] / g.sz;
        const float ofx = D[0] / g.sx, ofy = D[1] / g.sy, ofz = D[2] / g.sz;
        int lx = 0, hx = g.nu - 1, ly = 0, hy = g.nv - 1;
        int fx0, fx1, fy0, fy1;
        if (shadow_range(cvx, cvz, sfx, sfz, ofx, ofz, scx, g.nu, fx0, fx1) &&
            shadow_range(cvy, cvz, sfy, sfz, ofy, ofz, scy, g.nv, fy0, fy1)) {
            lx = fx0;
            hx = fx1;
            ly = fy0;
            hy = fy1;
        }
        const float a1z = t.z[p * nz1 + z], a2z = t.z[p * nz1 + z + 1];
        const float mnz = fminf(a1z, a2z), mxz = fmaxf(a1z, a2z);
        const float lz2 = t.lz2[p];
        const float *ty = t.y + (size_t)p * g.nv * ny1 + y;
        const float *tx = t.x + (size_t)p * g.nu * nx1 + x;
        const float *lxr = t.lx2 + (size_t)p * g.nu;
        const float *lyr = t.ly2 + (size_t)p * g.nv;
        const size_t pbase = (size_t)p * g.nu * g.nv;
        for (int py = ly; py <= hy; ++py) {
            const float a1y = ty[(size_t)py * ny1], a2y = ty[(size_t)py * ny1 + 1];
            const float lo = fmaxf(fmaxf(0.f, fminf(a1y, a2y)), mnz);
            const float up = fminf(fminf(1.f, fmaxf(a1y, a2y)), mxz);
            const float ly2 = lyr[py];
            const size_t rbase = pbase + (size_t)py * g.nu;
            for (int px = lx; px <= hx; ++px) {
                const float a1x = tx[(size_t)px * nx1], a2x = tx[(size_t)px * nx1 + 1];
                const float len = fminf(up, fmaxf(a1x, a2x)) - fmaxf(lo, fminf(a1x, a2x));
                if (len > 0.f) {
                    if (EP == KV_EP_COUNT) {
                        acc += 1.f;
                    } else {
                        const float L = sqrtf((lxr[px] + ly2) + lz2);
                        if (UNIT)
                            acc += len * L;  // len * (L * 1)
                        else
                            acc += len * (L * proj[rbase + px]);
                    }
                }
            }
        }
    }
    const size_t o = ((size_t)z * g.ny + y) * g.nx + x;
    if (EP == KV_EP_PLAIN || EP == KV_EP_COUNT) {
        vol[o] = acc;
    } else {
        const float s = sens[o];
        vol[o] = s > 0.f ? (vol[o] * acc) / s : vol[o];
    }
}

}  // namespace

hipError_t kv_launch_tables_r(const KvGeom &g, const KvTables &t, hipStream_t s)
{
    tables_r_kernel<<<2048, 256, 0, s>>>(g, t);
    return hipGetLastError();
}

hipError_t kv_launch_tables_a(const KvGeom &g, const KvTables &t, hipStream_t s)
{
    tables_a_kernel<<<2048, 256, 0, s>>>(g, t);
    return hipGetLastError();
}

hipError_t kv_launch_bp_r(const KvGeom &g, const KvTables &t, const float *proj, float *vol, int epilogue,
                          hipStream_t s)
{
    const dim3 grid((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz), block(256);
    switch (epilogue) {
    case KV_EP_PLAIN: bp_r_kernel<KV_EP_PLAIN><<<grid, block, 0, s>>>(g, t, proj, vol); break;
    case KV_EP_MLEM: bp_r_kernel<KV_EP_MLEM><<<grid, block, 0, s>>>(g, t, proj, vol); break;
    case KV_EP_COUNT: bp_r_kernel<KV_EP_COUNT><<<grid, block, 0, s>>>(g, t, proj, vol); break;
    default: bp_r_kernel<KV_EP_SART><<<grid, block, 0, s>>>(g, t, proj, vol); break;
    }
    return hipGetLastError();
}

hipError_t kv_launch_bp_a(const KvGeom &g, const KvTables &t, const float *proj, float *vol, int epilogue,
                          const float *sens, bool unit_input, hipStream_t s)
{
    const dim3 grid((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz), block(256);
    if (epilogue == KV_EP_COUNT)
        bp_a_kernel<KV_EP_COUNT, true><<<grid, block, 0, s>>>(g, t, nullptr, vol, nullptr);
    else if (epilogue == KV_EP_PLAIN && unit_input)
        bp_a_kernel<KV_EP_PLAIN, true><<<grid, block, 0, s>>>(g, t, nullptr, vol, nullptr);
    else if (epilogue == KV_EP_PLAIN)
        bp_a_kernel<KV_EP_PLAIN, false><<<grid, block, 0, s>>>(g, t, proj, vol, nullptr);
    else
        bp_a_kernel<KV_EP_MLEM, false><<<grid, block, 0, s>>>(g, t, proj, vol, sens);
    return hipGetLastError();
}
