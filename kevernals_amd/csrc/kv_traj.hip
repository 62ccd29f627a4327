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
// 256-voxel chunks (a fixed chunk -> workgroup map), stages a_i = sqrt(w) d_i
// (FP32, component-major) in LDS, and each thread accumulates one 4x4 block
// of the symmetric Gram matrix over a fixed subset of the chunk's voxels: FP32 x FP32 products
// are exact in FP64 and accumulate in FP64.  The per-workgroup partials are
// summed on the host in workgroup order, so the result is deterministic.
// traj_combine_kernel: y = x_h * exp(sum_i c_i d_i), the sum in FP64.
//
// A voxel where an iterate is not positive has d_i = 0 for the ratios that
// involve it, and weight 0 unless both sens and x_h are positive.
// This pass has no counterpart in the reference: it belongs to the
// accelerated reconstruction built on the mode-A MLEM update.
#include "kv_internal.h"

namespace {

constexpr int kChunk = 256;

__device__ __forceinline__ float log_ratio(float a, float b)
{
    // log(b / a) from the exact-or-nearly difference: accurate relative to
    // the (small) ratio itself when b is close to a.
    return (a > 0.f && b > 0.f) ? log1pf((b - a) / a) : 0.f;
}

__global__ __launch_bounds__(256) void traj_gram_kernel(KvTraj tr, int h, int64_t n, const float *__restrict__ sens,
                                                        double *__restrict__ partials)
{
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
            const float xh = tr.x[h][v];
            const float s = sens[v];
            const float sw = (s > 0.f && xh > 0.f) ? sqrtf(s * xh) : 0.f;
            float prev = tr.x[0][v];
            for (int i = 0; i < h; ++i) {
                const float cur = tr.x[i + 1][v];
                a[i * kChunk + t] = sw * log_ratio(prev, cur);
                prev = cur;
            }
        } else {
            for (int i = 0; i < h; ++i)
                a[i * kChunk + t] = 0.f;
        }
        for (int i = h; i < 4 * nbk; ++i)
            a[i * kChunk + t] = 0.f;
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

__global__ __launch_bounds__(256) void traj_combine_kernel(KvTraj tr, KvTrajCoef cf, int h, int64_t n,
                                                           float *__restrict__ out)
{
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        double e = 0.0;
        float prev = tr.x[0][v];
        for (int i = 0; i < h; ++i) {
            const float cur = tr.x[i + 1][v];
            e = __builtin_fma(cf.c[i], (double)log_ratio(prev, cur), e);
            prev = cur;
        }
        out[v] = prev > 0.f ? prev * expf((float)e) : prev;
    }
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
    traj_gram_kernel<<<kv_traj_gram_blocks(n), 256, 0, s>>>(tr, h, n, sens, partials);
    return hipGetLastError();
}

hipError_t kv_launch_traj_combine(const KvTraj &tr, const KvTrajCoef &cf, int h, int64_t n, float *out,
                                  hipStream_t s)
{
    const int64_t nb = std::min<int64_t>(4096, (n + 255) / 256);
    traj_combine_kernel<<<(int)std::max<int64_t>(1, nb), 256, 0, s>>>(tr, cf, h, n, out);
    return hipGetLastError();
}
