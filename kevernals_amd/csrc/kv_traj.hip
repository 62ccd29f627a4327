// Trajectory extrapolation of mode-A MLEM: the two volume-sized passes.
//
// After k iterations the last h+1 iterates x_0..x_h define the log-ratios
// d_i = log(x_{i+1}/x_i), i = 0..h-1.  Near convergence MLEM is a linear
// iteration d_{i+1} = J d_i whose Jacobian is self-adjoint in the inner
// product weighted by w = sens * x (log domain), so the small Gram matrix
// G_ij = sum_v w d_i d_j gives the Rayleigh-Ritz projection of J onto the
// span of the history; the host turns it into coefficients c_i, and the
// extrapolated estimate is y = x_h * exp(sum_i c_i d_i).
//
// traj_gram_kernel: one pass over the h+1 iterates.  A workgroup takes
// 256-voxel chunks (a fixed chunk -> workgroup map); each thread loads its
// voxel's h+1 iterates at once (the kernel is instantiated for h up to 8,
// 16, 24 and 32, so the loads are unrolled and in flight together), stages
// a_i = sqrt(w) d_i (FP32, component-major) in LDS, and then accumulates one
// 4x4 block of the symmetric Gram matrix over a fixed subset of the chunk's
// voxels: FP32 x FP32 products are exact in FP64 and accumulate in FP64.
// The per-workgroup partials are summed on the host in workgroup order, so
// the result is deterministic.
// traj_combine_kernel: y = x_h * exp(sum_i c_i d_i), the sum in FP64, with
// the same unrolled loads.
//
// A voxel where an iterate is not positive has d_i = 0 for the ratios that
// involve it, and weight 0 unless both sens and x_h are positive.
// This pass has no counterpart in the reference: it belongs to the
// accelerated reconstruction built on the mode-A MLEM update.
#include "kv_internal.h"

namespace {

constexpr int kChunk = 256;

// log(b / a) for positive a and b (0 otherwise): log1p of the exact-or-
// nearly difference over a, accurate relative to the (small) result.  The
// passes take it where log_ratio_fast cannot (extreme magnitudes).
__device__ __forceinline__ float log_ratio_ref(float a, float b)
{
    return (a > 0.f && b > 0.f) ? log1pf((b - a) / a) : 0.f;
}

// log(b / a) for positive a and b (0 otherwise), a few units in the last
// place of the result, which near convergence is far smaller than either
// log.  With s = (b - a) / (b + a) in [-1/4, 1/4] (b / a in [3/5, 5/3],
// where b - a is exact) it is 2 atanh(s) = 2s (1 + s^2/3 + s^4/5 + ... +
// s^12/13), the series cut where the rest is below 4e-10 of the sum
// (s^2 <= 1/16); elsewhere |log(b / a)| > 0.51 and the hardware log2 of the
// quotient is accurate relative to it.  Quotients use the hardware
// reciprocal with one Newton correction.  Where a reciprocal or the
// quotient leaves the normal range (a + b below 2^-126 or near the top of
// the range, b / a beyond it) the result is not usable: *redo is set and
// the caller takes log_ratio_ref.
__device__ __forceinline__ float log_ratio_fast(float a, float b, bool &redo)
{
    const float n = b - a, m = b + a;
    const float r = __builtin_amdgcn_rcpf(m);
    float s = n * r;
    s = __builtin_fmaf(__builtin_fmaf(-m, s, n), r, s);
    const float q = s * s;
    float p = 1.f / 13.f;
    p = __builtin_fmaf(p, q, 1.f / 11.f);
    p = __builtin_fmaf(p, q, 1.f / 9.f);
    p = __builtin_fmaf(p, q, 1.f / 7.f);
    p = __builtin_fmaf(p, q, 1.f / 5.f);
    p = __builtin_fmaf(p, q, 1.f / 3.f);
    const float s2 = s + s;
    float d = __builtin_fmaf(s2 * q, p, s2);
    bool ok = __builtin_isnormal(r);
    if (!(fabsf(s) <= 0.25f)) {
        const float ra = __builtin_amdgcn_rcpf(a);
        float rho = b * ra;
        rho = __builtin_fmaf(__builtin_fmaf(-a, rho, b), ra, rho);
        d = __builtin_amdgcn_logf(rho) * 0x1.62e430p-1f;  // log2(rho) * ln 2
        ok = ok && __builtin_isnormal(ra) && __builtin_isnormal(rho);
    }
    const bool pos = a > 0.f && b > 0.f;
    redo = pos && !ok;
    return pos && ok ? d : 0.f;
}

// The iterates x_0..x_HB of one voxel (those past x_h read as x_h, from the
// cache); every load is issued before the first is used.
template <int HB>
__device__ __forceinline__ void load_history(const KvTraj &tr, int h, int64_t v, float (&x)[HB + 1])
{
#pragma unroll
    for (int i = 0; i <= HB; ++i)
        x[i] = tr.x[i < h ? i : h][v];
}

template <int HB>
__global__ __launch_bounds__(256) void traj_gram_kernel(KvTraj tr, int h, int64_t n, const float *__restrict__ sens,
                                                        double *__restrict__ partials)
{
    static_assert(HB % 4 == 0 && HB <= KV_TRAJ_MAX, "bucket");
    __shared__ __align__(16) float a[KV_TRAJ_MAX * kChunk];  // 32 KiB; reused for the phase reduction
    const int nbk = (h + 3) / 4;
    const int npair = nbk * (nbk + 1) / 2;
    const int nph = 256 / npair;  // voxel phases per block pair
    const int t = threadIdx.x;
    const bool active = t < npair * nph;
    const int pair = active ? t / nph : 0;
    const int phase = active ? t % nph : 0;
    // pair -> (bi, bj), bi <= bj, row-major over the upper triangle
    int bi = 0, rem = pair;
    while (rem >= nbk - bi) {
        rem -= nbk - bi;
        ++bi;
    }
    const int bj = bi + rem;

    double acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q)
        acc[q] = 0.0;

    const int64_t nchunk = (n + kChunk - 1) / kChunk;
    for (int64_t c = blockIdx.x; c < nchunk; c += gridDim.x) {
        const int64_t v = c * kChunk + t;
        // component-major: a[i * kChunk + voxel], conflict-free writes
        if (v < n) {
            float x[HB + 1];
            load_history<HB>(tr, h, v, x);
            const float xh = x[HB];
            const float s = sens[v];
            // sqrt(w) = sqrt(s) sqrt(x_h): no overflow of the product near the top of the range
            const float sw = (s > 0.f && xh > 0.f) ? sqrtf(s) * sqrtf(xh) : 0.f;
            bool any = false;
#pragma unroll
            for (int i = 0; i < HB; ++i) {
                bool redo = false;
                const float d = i < h ? log_ratio_fast(x[i], x[i + 1], redo) : 0.f;
                a[i * kChunk + t] = sw * d;
                any = any || redo;
            }
            if (any) {
                for (int i = 0; i < h; ++i)
                    a[i * kChunk + t] = sw * log_ratio_ref(tr.x[i][v], tr.x[i + 1][v]);
            }
        } else {
#pragma unroll
            for (int i = 0; i < HB; ++i)
                a[i * kChunk + t] = 0.f;
        }
        __syncthreads();
        if (active) {
            // the pairs sharing a wavefront start at different voxels, so
            // their reads of different rows fall on different LDS banks
            for (int j = phase; j < kChunk; j += nph) {
                const int u = (j + 37 * pair) & (kChunk - 1);
                double pr[4], qc[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    pr[r] = (double)a[(4 * bi + r) * kChunk + u];
                    qc[r] = (double)a[(4 * bj + r) * kChunk + u];
                }
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int cc = 0; cc < 4; ++cc)
                        acc[r * 4 + cc] = __builtin_fma(pr[r], qc[cc], acc[r * 4 + cc]);
            }
        }
        __syncthreads();
    }
    // sum the phases of each pair in phase order
    double *red = reinterpret_cast<double *>(a);  // 256 threads x 16 doubles = 32 KiB
#pragma unroll
    for (int q = 0; q < 16; ++q)
        red[t * 16 + q] = acc[q];
    __syncthreads();
    for (int o = t; o < npair * 16; o += 256) {
        const int pp = o / 16, q = o % 16;
        double s = 0.0;
        for (int ph = 0; ph < nph; ++ph)
            s += red[(pp * nph + ph) * 16 + q];
        partials[(size_t)blockIdx.x * npair * 16 + o] = s;
    }
}

template <int HB>
__global__ __launch_bounds__(256) void traj_combine_kernel(KvTraj tr, KvTrajCoef cf, int h, int64_t n,
                                                           float *__restrict__ out)
{
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        float x[HB + 1];
        load_history<HB>(tr, h, v, x);
        double e = 0.0;
        bool any = false;
#pragma unroll
        for (int i = 0; i < HB; ++i) {
            if (i < h) {
                bool redo = false;
                e = __builtin_fma(cf.c[i], (double)log_ratio_fast(x[i], x[i + 1], redo), e);
                any = any || redo;
            }
        }
        if (any) {
            e = 0.0;
            for (int i = 0; i < h; ++i)
                e = __builtin_fma(cf.c[i], (double)log_ratio_ref(tr.x[i][v], tr.x[i + 1][v]), e);
        }
        const float xh = x[HB];
        out[v] = xh > 0.f ? xh * expf((float)e) : xh;
    }
}

// the smallest instantiated history length (8, 16, 24, 32) holding h
int bucket(int h)
{
    return h <= 8 ? 8 : h <= 16 ? 16 : h <= 24 ? 24 : 32;
}

}  // namespace

int kv_traj_gram_blocks(int64_t n)
{
    const int64_t nchunk = (n + kChunk - 1) / kChunk;
    return (int)std::min<int64_t>(1024, std::max<int64_t>(1, nchunk));
}

int kv_traj_pairs(int h)
{
    const int nbk = (h + 3) / 4;
    return nbk * (nbk + 1) / 2;
}

hipError_t kv_launch_traj_gram(const KvTraj &tr, int h, int64_t n, const float *sens, double *partials,
                               hipStream_t s)
{
    const unsigned nb = kv_traj_gram_blocks(n);
    switch (bucket(h)) {
    case 8: traj_gram_kernel<8><<<nb, 256, 0, s>>>(tr, h, n, sens, partials); break;
    case 16: traj_gram_kernel<16><<<nb, 256, 0, s>>>(tr, h, n, sens, partials); break;
    case 24: traj_gram_kernel<24><<<nb, 256, 0, s>>>(tr, h, n, sens, partials); break;
    default: traj_gram_kernel<32><<<nb, 256, 0, s>>>(tr, h, n, sens, partials); break;
    }
    return hipGetLastError();
}

hipError_t kv_launch_traj_combine(const KvTraj &tr, const KvTrajCoef &cf, int h, int64_t n, float *out,
                                  hipStream_t s)
{
    const int64_t nb = std::min<int64_t>(4096, (n + 255) / 256);
    const unsigned g = (unsigned)std::max<int64_t>(1, nb);
    switch (bucket(h)) {
    case 8: traj_combine_kernel<8><<<g, 256, 0, s>>>(tr, cf, h, n, out); break;
    case 16: traj_combine_kernel<16><<<g, 256, 0, s>>>(tr, cf, h, n, out); break;
    case 24: traj_combine_kernel<24><<<g, 256, 0, s>>>(tr, cf, h, n, out); break;
    default: traj_combine_kernel<32><<<g, 256, 0, s>>>(tr, cf, h, n, out); break;
    }
    return hipGetLastError();
}
