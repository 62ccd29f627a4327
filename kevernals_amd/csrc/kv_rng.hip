// Seeded transmission data: Poisson counts drawn on the device from the
// project's own counter-based generator, with no dependence on a library's
// random streams.
//
// The reference simulates no data (it reads DICOM projections,
// mainReconstruct.cpp:88-124); the synthetic workloads need
//   I ~ Poisson(I0 exp(-l)),  l_hat = -ln(clamp(I, 1, I0) / I0)
// for every detector ray, l the mode-A line integral in mm.
//
// * Generator: Philox4x64-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
//   numbers: as easy as 1, 2, 3", SC 2011): a 10-round bijection of a 256-bit
//   counter under a 128-bit key.  Element i of a draw with seed s reads the
//   blocks counter = (i, j, 0, 0), key = (s, 0x6b76) for j = 0, 1, ...; each
//   block gives four uniforms u = (w >> 11) * 2^-53 in [0, 1).
// * Sampler: for mean < 10, inversion by sequential search from k = 0;
//   otherwise PTRS, the transformed rejection with squeeze of W. Hoermann,
//   "The transformed rejection method for generating Poisson random
//   variables", Insurance: Mathematics and Economics 12 (1993) 39-45.
// * exp, log and log k! are evaluated here with IEEE +, -, *, / only (no
//   library transcendental), so the host restatement of the same steps
//   draws the same counts bit for bit.
#pragma clang fp contract(off)

#include <cmath>
#include <cstdint>

#include "kv_internal.h"

namespace {

struct U4 {
    uint64_t w[4];
};

__device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) { return __umul64hi(a, b); }

__device__ U4 philox4x64_10(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3, uint64_t k0, uint64_t k1)
{
    constexpr uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
    constexpr uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
    for (int r = 0; r < 10; ++r) {
        const uint64_t lo0 = M0 * c0, hi0 = mulhi64(M0, c0);
        const uint64_t lo1 = M1 * c2, hi1 = mulhi64(M1, c2);
        const uint64_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += W0;
        k1 += W1;
    }
    return U4{{c0, c1, c2, c3}};
}

// Uniform stream of one element: four doubles per Philox block.
struct Stream {
    uint64_t i, key, block = 0;
    int used = 4;
    U4 buf{};
    __device__ double next()
    {
        if (used == 4) {
            buf = philox4x64_10(i, block++, 0, 0, key, 0x6b76ull);
            used = 0;
        }
        return (double)(buf.w[used++] >> 11) * 0x1p-53;
    }
};

constexpr double LN2 = 0.6931471805599453094;  // ln 2 rounded to double
// ln 2 = LN2_HI + LN2_LO with LN2_HI holding 24 significant bits, so that
// n * LN2_HI is exact for the |n| < 1100 of kv_exp
constexpr double LN2_HI = 0x1.62e42fp-1, LN2_LO = 0x1.df473de6af279p-26;

// log x for finite x > 0: x = m 2^e with m in [sqrt(1/2), sqrt(2)), then
// log m = 2 atanh(s), s = (m - 1) / (m + 1), |s| <= 0.1716, summed to s^23.
__device__ double kv_log(double x)
{
    int e = 0;
    double m = frexp(x, &e);  // x = m 2^e, m in [0.5, 1): exact
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        e -= 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double t = 1.0 / 23.0;
    for (int k = 21; k >= 1; k -= 2)
        t = 1.0 / (double)k + s2 * t;
    return (double)e * LN2 + 2.0 * s * t;
}

// exp x (0 below -745, +inf above 709): x = n ln2 + r with |r| <= ln2 / 2
// (n the nearest integer; r by the two-part ln 2), exp r by its Taylor series
// to r^18 / 18!, scaled by 2^n exactly.
__device__ double kv_exp(double x)
{
    if (x < -745.0)
        return 0.0;
    if (x > 709.0)
        return INFINITY;
    const double n = rint(x / LN2);
    const double r = (x - n * LN2_HI) - n * LN2_LO;
    double t = 1.0;
    for (int k = 18; k >= 1; --k)
        t = 1.0 + r * t / (double)k;
    return ldexp(t, (int)n);
}

// log k! : the sum of logs below 32, Stirling's series for lgamma(k + 1)
// (terms to 1/(1680 n^7), n = k + 1) from there.
__device__ double kv_log_factorial(double k)
{
    if (k < 32.0) {
        double s = 0.0;
        for (double i = 2.0; i <= k; i += 1.0)
            s += kv_log(i);
        return s;
    }
    const double n = k + 1.0;
    const double n2 = n * n;
    const double series = (1.0 / 12.0 - (1.0 / 360.0 - (1.0 / 1260.0 - (1.0 / 1680.0) / n2) / n2) / n2) / n;
    return (n - 0.5) * kv_log(n) - n + 0.91893853320467274178 + series;  // 0.5 log(2 pi)
}

__device__ double kv_poisson(double lam, Stream &st)
{
    if (!(lam > 0.0))
        return 0.0;
    if (lam < 10.0) {
        // inversion: smallest k with F(k) >= u
        const double u = st.next();
        double p = kv_exp(-lam), f = p, k = 0.0;
        while (u > f && k < 1000.0) {
            k += 1.0;
            p = p * lam / k;
            f = f + p;
        }
        return k;
    }
    // PTRS (Hoermann 1993, algorithm PTRS)
    const double loglam = kv_log(lam);
    const double slam = kv_exp(0.5 * loglam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (int it = 0; it < 1000; ++it) {
        const double U = st.next() - 0.5;
        const double V = st.next();
        const double us = 0.5 - fabs(U);
        const double k = floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr)
            return k;
        if (k < 0.0 || (us < 0.013 && V > us))
            continue;
        if (kv_log(V) + kv_log(invalpha) - kv_log(a / (us * us) + b) <= -lam + k * loglam - kv_log_factorial(k))
            return k;
    }
    return floor(lam);  // not reached in practice (acceptance probability > 0.9 per round)
}

__global__ __launch_bounds__(256) void poisson_data_kernel(int64_t n, const float *__restrict__ ell, double i0,
                                                           uint64_t seed, float *__restrict__ out)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        Stream st{(uint64_t)i, seed};
        const double lam = i0 * kv_exp(-(double)ell[i]);
        const double cnt = kv_poisson(lam, st);
        const double c = fmin(fmax(cnt, 1.0), i0);
        out[i] = (float)(-kv_log(c / i0));
    }
}

}  // namespace

kv_status kv_poisson_data(kv_context *ctx, const float *d_ell, int64_t n, float i0, uint64_t seed, float *d_out)
{
    if (!ctx) {
        kv_set_error("null context");
        return KV_ERR_INVALID;
    }
    if (!d_ell || !d_out || n < 0 || !(i0 >= 1.f) || !std::isfinite(i0)) {
        kv_set_error("poisson data: need buffers, n >= 0 and a finite I0 >= 1");
        return KV_ERR_INVALID;
    }
    if (n == 0)
        return KV_OK;
    const int nb = (int)std::min<int64_t>(16384, (n + 255) / 256);
    poisson_data_kernel<<<nb, 256, 0, ctx->stream>>>(n, d_ell, (double)i0, seed, d_out);
    KV_TRY_HIP(hipGetLastError());
    return KV_OK;
}
