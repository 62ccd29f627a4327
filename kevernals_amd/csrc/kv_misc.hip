// Shift-and-add and the deterministic log-likelihood reduction.
#pragma clang fp contract(off)

#include "kv_internal.h"

namespace {

// recons::ShiftAndAdd (Reconstructors.h:314-378): for plane k,
// out(u, v) = sum_p (1/P) * proj_p(u, v + t[k][p]), zero outside the
// detector (vtkImageConstantPad), summed in view order in FP64 and stored as
// FP32.  One lane per (u, v); the plane loop keeps the rows a lane reads
// close together (t changes slowly with k).
__global__ __launch_bounds__(256) void saa_kernel(int nu, int nv, int np, const float *__restrict__ proj,
                                                  const int32_t *__restrict__ shifts, int nplanes,
                                                  float *__restrict__ out)
{
    const int u = blockIdx.x * 64 + (threadIdx.x & 63);
    const int v = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (u >= nu || v >= nv)
        return;
    const double wgt = 1.0 / (double)np;
    const size_t plane = (size_t)nu * nv;
    for (int k = 0; k < nplanes; ++k) {
        double acc = 0.0;
        for (int p = 0; p < np; ++p) {
            const int vv = v + shifts[k * np + p];
            const double x = (vv >= 0 && vv < nv) ? (double)proj[p * plane + (size_t)vv * nu + u] : 0.0;
            acc += x * wgt;
        }
        out[k * plane + (size_t)v * nu + u] = (float)acc;
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

}  // namespace

hipError_t kv_launch_sum(int64_t n, const float *x, double *partials, int nblocks, hipStream_t s)
{
    sum_kernel<<<nblocks, 256, 0, s>>>(n, x, partials);
    return hipGetLastError();
}

hipError_t kv_launch_saa(int nu, int nv, int np, const float *proj, const int32_t *shifts, int nplanes, float *out,
                         hipStream_t s)
{
    const dim3 grid((nu + 63) / 64, (nv + 3) / 4), block(256);
    saa_kernel<<<grid, block, 0, s>>>(nu, nv, np, proj, shifts, nplanes, out);
    return hipGetLastError();
}

hipError_t kv_launch_loglik(const KvGeom &g, const float *data, const float *ax, double *partials, int nblocks,
                            hipStream_t s)
{
    loglik_kernel<<<nblocks, 256, 0, s>>>(kv_nray(g), data, ax, partials);
    return hipGetLastError();
}
