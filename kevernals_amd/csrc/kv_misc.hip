// Shift-and-add and the deterministic log-likelihood reduction.
#pragma clang fp contract(off)

#include <algorithm>
#include <functional>
#include <mutex>

#include "kv_internal.h"

namespace {

// recons::ShiftAndAdd (Reconstructors.h:314-378): for plane k,
// out(u, v) = sum_p (1/P) * proj_p(u, v + t[k][p]), zero outside the
// detector (vtkImageConstantPad), summed in view order in FP64 and stored as
// FP32.
//
// A workgroup owns a strip of SAA_TU detector columns by SAA_RV rows and
// walks the planes in chunks of kc (<= SAA_KC).  For each chunk it visits the
// views in order; the rows of view p that the chunk's planes read,
// [v0 + min_k t, v0 + SAA_RV - 1 + max_k t], are staged once in LDS as their
// FP64 products x * (1/P) (0 outside the detector), and every (plane, row)
// of the chunk adds its staged product: each projection row is read from
// memory once per chunk instead of once per plane, and the product is formed
// once per staged element.  The sums keep the view order and the
// per-element arithmetic of the one-lane-per-pixel form, so the result is
// bit-identical.  The host picks kc so that SAA_RV plus the chunk's largest
// shift spread fits the SAA_SR staged rows (kc = 1 always fits).
constexpr int SAA_TU = 64, SAA_RV = 16, SAA_KC = 8, SAA_SR = 64;
constexpr int SAA_RPT = SAA_RV / 4;  // output rows per thread (4 wavefronts)

__global__ __launch_bounds__(256) void saa_kernel(int nu, int nv, int np, const float *__restrict__ proj,
                                                  const int32_t *__restrict__ shifts, int nplanes, int kc,
                                                  float *__restrict__ out)
{
    __shared__ double st[SAA_SR][SAA_TU];
    const int lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
    const int u = blockIdx.x * SAA_TU + lane;
    const int v0 = blockIdx.y * SAA_RV;
    const double wgt = 1.0 / (double)np;
    const size_t plane = (size_t)nu * nv;
    for (int k0 = 0; k0 < nplanes; k0 += kc) {
        const int nk = min(kc, nplanes - k0);
        double acc[SAA_KC][SAA_RPT];
#pragma unroll
        for (int j = 0; j < SAA_KC; ++j)
#pragma unroll
            for (int i = 0; i < SAA_RPT; ++i)
                acc[j][i] = 0.0;
        for (int p = 0; p < np; ++p) {
            // the chunk's shifts for view p (wave-uniform)
            int t[SAA_KC];
            int tlo = shifts[k0 * np + p], thi = tlo;
#pragma unroll
            for (int j = 0; j < SAA_KC; ++j) {
                t[j] = j < nk ? shifts[(k0 + j) * np + p] : t[0];
                tlo = min(tlo, t[j]);
                thi = max(thi, t[j]);
            }
            // <= SAA_SR by the host's choice of kc; the clamps below keep
            // every LDS access inside the array whatever the shifts
            const int nrows = min(SAA_RV + thi - tlo, SAA_SR);
            __syncthreads();  // the previous view's readers are done
            const float *src = proj + (size_t)p * plane;
            for (int r = grp; r < nrows; r += 4) {
                const int vv = v0 + tlo + r;
                double x = 0.0;
                if (vv >= 0 && vv < nv && u < nu)
                    x = (double)src[(size_t)vv * nu + u] * wgt;
                st[r][lane] = x;
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < SAA_KC; ++j) {
                if (j < nk) {
                    const int r0 = min(grp * SAA_RPT + t[j] - tlo, SAA_SR - SAA_RPT);
#pragma unroll
                    for (int i = 0; i < SAA_RPT; ++i)
                        acc[j][i] += st[r0 + i][lane];
                }
            }
        }
        if (u < nu) {
#pragma unroll
            for (int j = 0; j < SAA_KC; ++j) {
                if (j < nk) {
#pragma unroll
                    for (int i = 0; i < SAA_RPT; ++i) {
                        const int v = v0 + grp * SAA_RPT + i;
                        if (v < nv)
                            out[(size_t)(k0 + j) * plane + (size_t)v * nu + u] = (float)acc[j][i];
                    }
                }
            }
        }
    }
}

// Poisson log-likelihood terms y*log(ax) - ax (ax > 0) reduced per block in a
// fixed tree order; the host adds the block partials in index order.
__global__ __launch_bounds__(256) void loglik_kernel(int64_t n, const float *__restrict__ y,
                                                     const float *__restrict__ ax, double *__restrict__ partials)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double a = ax[i];
        if (a > 0.0)
            acc += (double)y[i] * log(a) - a;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        partials[blockIdx.x] = red[0];
}

// Plain FP64 sum with the same fixed-order block reduction.
__global__ __launch_bounds__(256) void sum_kernel(int64_t n, const float *__restrict__ x, double *__restrict__ partials)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        acc += (double)x[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        partials[blockIdx.x] = red[0];
}

// AlmostEqualRelative(a, b) of commons/Maths.h:25-34 in FP32.
__device__ __forceinline__ bool almost_equal_relative(float a, float b)
{
    const float diff = fabsf(a - b);
    const float A = fabsf(a), B = fabsf(b);
    const float m = (A < B) ? B : A;  // std::max
    return diff <= __FLT_EPSILON__ * m;
}

// PhantomMaker::GetPhantom's per-voxel filler (PhantomMaker.cpp:13-29): the
// background, then each containing box, then each containing sphere, added
// in FP32 in list order.  The grid point of voxel (x, y, z) is
// (px[x], py[y], pz[z]) (its lower corner, TomoVolume.cpp:193-229).  The
// shape lists are staged in LDS by chunks.
__global__ __launch_bounds__(256) void phantom_kernel(int nx, int ny, int nz, const float *__restrict__ px,
                                                      const float *__restrict__ py, const float *__restrict__ pz,
                                                      float background, const kv_pave *__restrict__ paves, int npaves,
                                                      const kv_sphere *__restrict__ spheres, int nspheres,
                                                      float *__restrict__ out)
{
    constexpr int kChunk = 256;
    __shared__ kv_pave sp[kChunk];
    __shared__ kv_sphere ss[kChunk];
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t first = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    // every workgroup walks the same number of voxel rounds, so all of its
    // threads reach every barrier
    const int64_t rounds = (n + stride - 1) / stride;
    for (int64_t r = 0; r < rounds; ++r) {
        const int64_t i = first + r * stride;
        const bool live = i < n;
        float gx = 0.f, gy = 0.f, gz = 0.f;
        if (live) {
            const int x = (int)(i % nx);
            const int64_t q = i / nx;
            const int y = (int)(q % ny), z = (int)(q / ny);
            gx = px[x];
            gy = py[y];
            gz = pz[z];
        }
        float d = background;
        for (int c0 = 0; c0 < npaves; c0 += kChunk) {
            const int m = min(kChunk, npaves - c0);
            __syncthreads();
            if ((int)threadIdx.x < m)
                sp[threadIdx.x] = paves[c0 + threadIdx.x];
            __syncthreads();
            for (int k = 0; k < m; ++k) {
                const kv_pave P = sp[k];
                // Pave::Contains (Pave.h:51-59)
                const float ex = P.x + P.wx, ey = P.y + P.wy, ez = P.z + P.wz;
                const bool in = (gx > P.x || almost_equal_relative(gx, P.x)) &&
                                (gx < ex || almost_equal_relative(gx, ex)) &&
                                (gy > P.y || almost_equal_relative(gy, P.y)) &&
                                (gy < ey || almost_equal_relative(gy, ey)) &&
                                (gz > P.z || almost_equal_relative(gz, P.z)) &&
                                (gz < ez || almost_equal_relative(gz, ez));
                d = in ? d + P.density : d;
            }
        }
        for (int c0 = 0; c0 < nspheres; c0 += kChunk) {
            const int m = min(kChunk, nspheres - c0);
            __syncthreads();
            if ((int)threadIdx.x < m)
                ss[threadIdx.x] = spheres[c0 + threadIdx.x];
            __syncthreads();
            for (int k = 0; k < m; ++k) {
                const kv_sphere S = ss[k];
                // Sphere::Contains (Sphere.h:37-40): |centre - point| < radius,
                // the norm as Vector3D::NormL2 (Vector3D.cpp:12-15)
                const float dx = S.cx - gx, dy = S.cy - gy, dz = S.cz - gz;
                const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
                d = dist < S.radius ? d + S.density : d;
            }
        }
        if (live)
            out[i] = d;
    }
}

// vtkImageLogarithmicScale (VTK 9.0.1): c*log(1 + x) for x > 0,
// -c*log(1 - x) otherwise, in FP64, stored as FP32.
__global__ __launch_bounds__(256) void log_scale_kernel(int64_t n, const float *__restrict__ in,
                                                        float *__restrict__ out, double c)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double x = (double)in[i];
        out[i] = (float)(x > 0.0 ? c * log(x + 1.0) : -c * log(1.0 - x));
    }
}

// Step 2 of recons::ART (Reconstructors.h:87-93) or recons::MLEM (:204-214)
// on a forward projection: the forward projector's fused epilogues taken
// apart.
__global__ __launch_bounds__(256) void correct_kernel(int64_t n, int loop, float relax,
                                                      const float *__restrict__ data, float *__restrict__ proj)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float fp = proj[i];
        if (loop == KV_LOOP_MLEM)
            proj[i] = almost_equal_zero(fp) ? fp : relax * (data[i] / fp);
        else
            proj[i] = relax * (data[i] - fp);
    }
}

// Step 4: x += b (ART, Reconstructors.h:113-121) or x *= b (MLEM, :235-241).
__global__ __launch_bounds__(256) void update_kernel(int64_t n, int loop, float *__restrict__ vol,
                                                     const float *__restrict__ bp)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        vol[i] = loop == KV_LOOP_MLEM ? vol[i] * bp[i] : vol[i] + bp[i];
}

// Row-shift table of plane k, view p (Reconstructors.h:308-331), evaluated
// from the staged source y positions and plane heights with the host's FP32
// expressions (IEEE division: the library is built with correctly rounded
// FP32 division); the host has already checked every |t| < 2e9.
__global__ __launch_bounds__(256) void saa_shift_kernel(int np, int nplanes, float sid, float mF, float voxel_dy,
                                                        const float *__restrict__ src_y,
                                                        const float *__restrict__ plane_z, int32_t *__restrict__ shifts)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= np * nplanes)
        return;
    const int k = i / np, p = i - k * np;
    const float mZ = sid / (sid - plane_z[k]);
    const float coeff = mZ - mF;
    const float world = src_y[p] * coeff;
    shifts[i] = (int32_t)(world / voxel_dy);
}

}  // namespace

// Staging of kv_shift_and_add_noalloc: the source y positions, the plane
// heights and the shift table, in one file-scope array (each GPU holds its
// own copy until the process ends).  Host code reaches it only through the
// per-device address that hipGetSymbolAddress returns.
static __device__ __align__(256) int32_t scratch[KV_SAA_NOALLOC_MAX_VIEWS + KV_SAA_NOALLOC_MAX_PLANES +
                                                 KV_SAA_NOALLOC_MAX_SHIFTS];

namespace {
constexpr int KV_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
int32_t *g_scratch[KV_MAX_DEVICES] = {};
}  // namespace

hipError_t kv_internal_scratch_init(int device)
{
    if (device < 0 || device >= KV_MAX_DEVICES)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device])
        return hipSuccess;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));  // the current device's copy
    if (e != hipSuccess)
        return e;
    g_scratch[device] = (int32_t *)ptr;
    return hipSuccess;
}

int32_t *kv_internal_scratch(int device)
{
    if (device < 0 || device >= KV_MAX_DEVICES)
        return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch[device];
}

hipError_t kv_launch_saa_shifts(int np, int nplanes, float sid, float mF, float voxel_dy, const float *src_y,
                                const float *plane_z, int32_t *shifts, hipStream_t s)
{
    const int n = np * nplanes;
    saa_shift_kernel<<<(n + 255) / 256, 256, 0, s>>>(np, nplanes, sid, mF, voxel_dy, src_y, plane_z, shifts);
    return hipGetLastError();
}

hipError_t kv_launch_correct(int64_t n, int loop, float relax, const float *data, float *proj, hipStream_t s)
{
    const int nb = (int)std::min<int64_t>(8192, std::max<int64_t>(1, (n + 255) / 256));
    correct_kernel<<<nb, 256, 0, s>>>(n, loop, relax, data, proj);
    return hipGetLastError();
}

hipError_t kv_launch_update(int64_t n, int loop, float *vol, const float *bp, hipStream_t s)
{
    const int nb = (int)std::min<int64_t>(8192, std::max<int64_t>(1, (n + 255) / 256));
    update_kernel<<<nb, 256, 0, s>>>(n, loop, vol, bp);
    return hipGetLastError();
}

hipError_t kv_launch_phantom(int nx, int ny, int nz, const float *px, const float *py, const float *pz,
                             float background, const kv_pave *paves, int npaves, const kv_sphere *spheres,
                             int nspheres, float *out, hipStream_t s)
{
    const int64_t n = (int64_t)nx * ny * nz;
    const int nb = (int)std::min<int64_t>(16384, std::max<int64_t>(1, (n + 255) / 256));
    phantom_kernel<<<nb, 256, 0, s>>>(nx, ny, nz, px, py, pz, background, paves, npaves, spheres, nspheres, out);
    return hipGetLastError();
}

hipError_t kv_launch_log_scale(int64_t n, const float *in, float *out, double c, hipStream_t s)
{
    const int nb = (int)std::min<int64_t>(8192, std::max<int64_t>(1, (n + 255) / 256));
    log_scale_kernel<<<nb, 256, 0, s>>>(n, in, out, c);
    return hipGetLastError();
}

hipError_t kv_launch_sum(int64_t n, const float *x, double *partials, int nblocks, hipStream_t s)
{
    sum_kernel<<<nblocks, 256, 0, s>>>(n, x, partials);
    return hipGetLastError();
}

int kv_saa_chunk(int nplanes, int np, const std::function<int(int, int)> &shift)
{
    for (int kc = SAA_KC; kc > 1; kc /= 2) {
        int spread = 0;
        for (int k0 = 0; k0 < nplanes; k0 += kc)
            for (int p = 0; p < np; ++p) {
                int lo = shift(k0, p), hi = lo;
                for (int k = k0 + 1; k < std::min(nplanes, k0 + kc); ++k) {
                    const int t = shift(k, p);
                    lo = std::min(lo, t);
                    hi = std::max(hi, t);
                }
                spread = std::max(spread, hi - lo);
            }
        if (SAA_RV + spread <= SAA_SR)
            return kc;
    }
    return 1;
}

hipError_t kv_launch_saa(int nu, int nv, int np, const float *proj, const int32_t *shifts, int nplanes, int kc,
                         float *out, hipStream_t s)
{
    if (kc < 1 || kc > SAA_KC)
        return hipErrorInvalidValue;
    const dim3 grid((nu + SAA_TU - 1) / SAA_TU, (nv + SAA_RV - 1) / SAA_RV), block(256);
    saa_kernel<<<grid, block, 0, s>>>(nu, nv, np, proj, shifts, nplanes, kc, out);
    return hipGetLastError();
}

hipError_t kv_launch_loglik(const KvGeom &g, const float *data, const float *ax, double *partials, int nblocks,
                            hipStream_t s)
{
    loglik_kernel<<<nblocks, 256, 0, s>>>(kv_nray(g), data, ax, partials);
    return hipGetLastError();
}
