// The learned step: a 3D residual CNN on the matrix cores (no reference
// counterpart).
//
//   a0 = s*x            a1 = relu(conv(W1, a0) + b1)     W1[16][1][27]
//   a2 = relu(conv(W2, a1) + b2)     a3 = relu(conv(W3, a2) + b3)
//   z4 = conv(W4, a3) + b4           y  = x + z4 / s   W4[1][16][27]
// conv(W, a)[v][o] = sum_{t,i} W[o][i][t] a[v + off(t)][i], zero padding,
// tap t = 9(dz+1) + 3(dy+1) + (dx+1).  Activations are bfloat16,
// channels-last ([voxel][16], 32 B per voxel); every product is a
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation.
//
// Convolutions are implicit GEMMs with the output channels as the 16 MFMA
// rows and 16 consecutive x-voxels as the 16 columns; K runs over (tap,
// input channel): 27 taps x 16 channels, padded to 28 taps = 14 MFMAs per
// row of 16 voxels.  A workgroup stages a 18x10x6-voxel halo of its
// 16x8x4 output brick in LDS (two 16-byte channel halves per voxel, stored
// as separate planes so that 16 lanes read 256 contiguous bytes) and keeps
// the weights in registers; it walks 16 such bricks down z, keeping the
// halo's z-planes in a ring so that each brick after the first loads only
// its four new planes.  The input gradient is the same kernel with the
// weights transposed and mirrored and a ReLU-mask epilogue.  The weight
// gradient is a GEMM over voxels (K) of the output gradient (rows) and the
// shifted input (columns), accumulated per workgroup over 32x4x4-voxel
// bricks and reduced in a fixed order; an all-ones column yields the bias
// gradient.  All reductions are deterministic.
#include <cmath>
#include <cstring>

#include "kv_internal.h"

namespace {

constexpr int BX = 16, BY = 8, BZ = 4;
constexpr int HX = BX + 2, HY = BY + 2, HZ = BZ + 2;
constexpr int HYX = HX * HY, HVOX = HYX * HZ;

constexpr int WX = 32, WY = 4, WZ = 4, WVOX = WX * WY * WZ;
constexpr int WHX = WX + 2, WHY = WY + 2, WHZ = WZ + 2, WHV = WHX * WHY * WHZ;
constexpr int WGRID = 512;  // weight-gradient workgroups (fixed: results do not depend on the device)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// float -> bfloat16 bits, round to nearest even (inputs are finite): the
// hardware conversion (v_cvt_pk_bf16_f32), one or two values at a time
typedef __bf16 kv_bf16x2 __attribute__((ext_vector_type(2)));
typedef float kv_f32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned int f2bf(float f)
{
    return __builtin_bit_cast(unsigned short, (__bf16)f);
}
// (lo, hi) -> the bfloat16 pair lo | hi << 16
__device__ __forceinline__ unsigned int f2bf2(float lo, float hi)
{
    return __builtin_bit_cast(unsigned int, __builtin_convertvector((kv_f32x2){lo, hi}, kv_bf16x2));
}

__device__ __forceinline__ float bf2f(unsigned int h) { return __uint_as_float(h << 16); }

__device__ __forceinline__ u32x4 pack8(const unsigned int e[8])
{
    u32x4 r;
    r[0] = e[0] | (e[1] << 16);
    r[1] = e[2] | (e[3] << 16);
    r[2] = e[4] | (e[5] << 16);
    r[3] = e[6] | (e[7] << 16);
    return r;
}

__device__ __forceinline__ f32x4 mfma(const u32x4 &a, const u32x4 &b, const f32x4 &c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}

__device__ __forceinline__ int tap_dz(int t) { return t / 9 - 1; }
__device__ __forceinline__ int tap_dy(int t) { return (t / 3) % 3 - 1; }
__device__ __forceinline__ int tap_dx(int t) { return t % 3 - 1; }

// Element (o, i, t) of the operator a layer applies: the layer's own
// W[o][i][t] (cout x cin), or for the input gradient its transpose with
// mirrored taps W[i][o][26 - t].  Zero outside the layer's channels.
__device__ __forceinline__ float w_elem(const float *w, int cout, int cin, int tr, int o, int i, int t)
{
    if (t >= 27)
        return 0.f;
    if (!tr)
        return (o < cout && i < cin) ? w[(o * cin + i) * 27 + t] : 0.f;
    return (i < cout && o < cin) ? w[(i * cin + o) * 27 + (26 - t)] : 0.f;
}

enum { EP_RELU = 0, EP_MASK = 1, EP_OUT = 2, EP_LOSS = 3 };

__device__ __forceinline__ size_t vidx(int x, int y, int z, int nx, int ny)
{
    return ((size_t)z * ny + y) * nx + x;
}

// Stores four output channels (4g..4g+3) of voxel v: bias+ReLU or ReLU mask.
template <int EP>
__device__ __forceinline__ void store4(const f32x4 &acc, int g, size_t v, const float *bias,
                                       unsigned short *out16, const unsigned short *mask)
{
    float o[4];
    if (EP == EP_RELU) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
            o[i] = fmaxf(acc[i] + bias[4 * g + i], 0.f);
    } else {
        const uint2 m = *reinterpret_cast<const uint2 *>(mask + v * 16 + 4 * g);
        const unsigned int mb[4] = {m.x & 0xFFFFu, m.x >> 16, m.y & 0xFFFFu, m.y >> 16};
#pragma unroll
        for (int i = 0; i < 4; ++i)
            o[i] = bf2f(mb[i]) > 0.f ? acc[i] : 0.f;
    }
    *reinterpret_cast<uint2 *>(out16 + v * 16 + 4 * g) = make_uint2(f2bf2(o[0], o[1]), f2bf2(o[2], o[3]));
}

// Weight operands of every convolution, packed once per call as the MFMA A
// operands (bf16): variant v, operand m, lane l at wpk[(v * WV_OPS + m) * 64 + l].
// conv16 variants: m = 2 (3 dz + dy) + p for the taps (dz, dy, dx) of the
// pair p: p = 0 holds dx = 0 in k-groups 0-1 and dx = 1 in k-groups 2-3,
// p = 1 holds dx = 2 in k-groups 0-1 and zeros in 2-3; k-group g covers
// channels 8 (g & 1) .. +7.  conv1 variants use m = 0 (k-group g holds taps
// 8g .. 8g+7).
enum { WV_L2 = 0, WV_L3 = 1, WV_L4 = 2, WV_L3T = 3, WV_L2T = 4, WV_L1 = 5, WV_L4T = 6, WV_COUNT = 7 };
constexpr int WV_OPS = 18;

__global__ void pack_weights_kernel(const float *__restrict__ p, u32x4 *__restrict__ wpk, int ow1, int ow2, int ow3,
                                    int ow4)
{
    const int v = blockIdx.x, lane = threadIdx.x, col = lane & 15, g = lane >> 4;
    const float *w;
    int cout, cin, tr;
    switch (v) {
    case WV_L2: w = p + ow2; cout = 16; cin = 16; tr = 0; break;
    case WV_L3: w = p + ow3; cout = 16; cin = 16; tr = 0; break;
    case WV_L4: w = p + ow4; cout = 1; cin = 16; tr = 0; break;
    case WV_L3T: w = p + ow3; cout = 16; cin = 16; tr = 1; break;
    case WV_L2T: w = p + ow2; cout = 16; cin = 16; tr = 1; break;
    case WV_L1: w = p + ow1; cout = 16; cin = 1; tr = 0; break;
    default: w = p + ow4; cout = 1; cin = 16; tr = 1; break;
    }
    for (int m = 0; m < WV_OPS; ++m) {
        const int pair = m & 1, dzdy = m >> 1, dx = pair ? 2 : (g >> 1);
        unsigned int e[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            float val;
            if (v >= WV_L1)
                val = m == 0 ? w_elem(w, cout, cin, tr, col, 0, 8 * g + j) : 0.f;
            else
                val = pair && g >= 2 ? 0.f : w_elem(w, cout, cin, tr, col, 8 * (g & 1) + j, 3 * dzdy + dx);
            e[j] = f2bf(val);
        }
        wpk[(v * WV_OPS + m) * 64 + lane] = pack8(e);
    }
}

// z-chunks of BZ slices per workgroup: weights loaded once per workgroup,
// and conv16's halo planes shared by consecutive chunks.  Measured at
// 1024x1024x64 (one CNN application): 4 / 8 / 16 chunks 4.82 / 4.65 /
// 4.59 ms with the shared planes, 5.06 ms for 4 chunks without.
constexpr int ZPB = 16;
constexpr int SPT = (HVOX + 255) / 256;  // halo voxels staged per thread
constexpr int SPT1 = 3;                   // of which loaded before the chunk barrier
constexpr int SPTN = (BZ * HYX + 255) / 256;  // voxels per thread of a later chunk's BZ new planes
static_assert(SPT - SPT1 <= SPT1, "the second staging round reuses the first round's registers");
static_assert(SPTN <= SPT1, "a later chunk's loads use the first round's registers");

// 16 -> 16 (or 16 -> 1) channel convolution of a bf16 [V][16] input.
// A workgroup computes a 16x8x4 brick per z-chunk; wavefront w owns the
// output rows y = 2w, 2w+1 (all four z).  The halo sits in LDS as two planes
// of 16-byte channel halves.  One B operand holds an input row segment at x
// shift dx (k-groups 0-1) and dx + 1 (k-groups 2-3): it is read once and
// feeds every output row whose 3x3x3 window holds that row segment, up to six
// MFMAs, with the weights of the tap pair from registers.  Each output row
// accumulates its taps in tap order.
template <int EP>
__global__ __launch_bounds__(256) void conv16_kernel(int nx, int ny, int nz, const u32x4 *__restrict__ in,
                                                     const u32x4 *__restrict__ wpk, const float *__restrict__ bias,
                                                     unsigned short *__restrict__ out16,
                                                     const unsigned short *__restrict__ mask,
                                                     const float *__restrict__ x, const float *__restrict__ target,
                                                     float *__restrict__ outf, float s, float gscale,
                                                     double *__restrict__ partial)
{
    __shared__ u32x4 lds[2 * HVOX];
    __shared__ double red[4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int col = lane & 15, g = lane >> 4;
    const int x0 = blockIdx.x * BX, y0 = blockIdx.y * BY;

    u32x4 A[WV_OPS];
#pragma unroll
    for (int m = 0; m < WV_OPS; ++m)
        A[m] = wpk[m * 64 + lane];
    // this lane's B address for halo plane 0, halo row 2w, pair 0
    const int bbase = (g & 1) * HVOX + (2 * wv) * HX + col + (g >> 1);
    const int shift2 = 2 - (g >> 1);  // pair 1 reads dx = 2 in every k-group

    double lsum = 0.0;
    const int zc0 = (int)blockIdx.z * ZPB;
    const int zc_end = min(zc0 + ZPB, (nz + BZ - 1) / BZ);
    for (int zc = zc0; zc < zc_end; ++zc) {
        const int z0 = zc * BZ;
        // The halo's HZ planes sit in a ring of HZ plane slots: halo plane zp
        // of this chunk in slot (rot + zp) mod HZ.  The next chunk's planes 0
        // and 1 are this chunk's planes BZ and BZ + 1, so after the first
        // chunk only the BZ new planes (zp = 2 .. HZ - 1) are loaded.
        const bool first = zc == zc0;
        const int rot = (BZ * (zc - zc0)) % HZ;
        const int zp0 = first ? 0 : HZ - BZ;
        const int nnew = (HZ - zp0) * HYX;
        // The new planes' z-slabs as one buffer resource (at most six slabs;
        // the launcher checks they span under 4 GiB): a voxel outside the
        // volume reads at the resource's end and loads zeros.  A chunk's
        // loads are issued before the barrier, so they overlap the other
        // wavefronts' work on the previous chunk.
        const int zlo = max(z0 - 1 + zp0, 0), zhi = min(z0 + BZ, nz - 1);
        const unsigned span =
            zhi >= zlo ? (unsigned)((unsigned long long)(zhi - zlo + 1) * nx * ny * 32ull) : 0u;
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc((void *)(in + (size_t)zlo * nx * ny * 2), (short)0, (int)span, 0x00020000);
        auto load = [&](int k, u32x4 &a, u32x4 &b) {
            const int q = tid + 256 * k;
            const int lx = q % HX, ly = (q / HX) % HY, lz = zp0 + q / HYX;
            const int gx = x0 + lx - 1, gy = y0 + ly - 1, gz = z0 + lz - 1;
            const bool ok = q < nnew && gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz;
            const unsigned off =
                ok ? (((unsigned)(gz - zlo) * (unsigned)ny + (unsigned)gy) * (unsigned)nx + (unsigned)gx) * 32u : span;
            a = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(rs, (int)off, 0, 0));
            b = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(rs, (int)(off + 16u), 0, 0));
        };
        auto store = [&](int k, const u32x4 &a, const u32x4 &b) {
            const int q = tid + 256 * k;
            if (q < nnew) {
                int slot = rot + zp0 + q / HYX;
                slot = slot >= HZ ? slot - HZ : slot;
                const int d = slot * HYX + q % HYX;
                lds[d] = a;
                lds[HVOX + d] = b;
            }
        };
        u32x4 h0[SPT1], h1[SPT1];
        if (first) {
            // the first SPT1 voxels' loads are in flight across the barrier;
            // the rest follow (fewer registers held)
#pragma unroll
            for (int k = 0; k < SPT1; ++k)
                load(k, h0[k], h1[k]);
            __syncthreads();  // (no earlier chunk here; keeps one staging sequence)
#pragma unroll
            for (int k = 0; k < SPT1; ++k)
                store(k, h0[k], h1[k]);
#pragma unroll
            for (int k = SPT1; k < SPT; ++k)
                load(k, h0[k - SPT1], h1[k - SPT1]);
#pragma unroll
            for (int k = SPT1; k < SPT; ++k)
                store(k, h0[k - SPT1], h1[k - SPT1]);
        } else {
            // BZ planes: every load in flight across the barrier
#pragma unroll
            for (int k = 0; k < SPTN; ++k)
                load(k, h0[k], h1[k]);
            __syncthreads();  // the previous chunk's operands are consumed
#pragma unroll
            for (int k = 0; k < SPTN; ++k)
                store(k, h0[k], h1[k]);
        }
        __syncthreads();
        f32x4 acc[2][BZ];
#pragma unroll
        for (int jy = 0; jy < 2; ++jy)
#pragma unroll
            for (int jz = 0; jz < BZ; ++jz)
                acc[jy][jz] = f32x4{0.f, 0.f, 0.f, 0.f};
        // halo plane zp (output z = zp - 1 + dz), halo row 2w + yp
#pragma unroll
        for (int zp = 0; zp < BZ + 2; ++zp) {
            const int sl = rot + zp >= HZ ? rot + zp - HZ : rot + zp;  // ring slot of halo plane zp
#pragma unroll
            for (int yp = 0; yp < 4; ++yp)
#pragma unroll
                for (int pair = 0; pair < 2; ++pair) {
                    const u32x4 b = lds[bbase + sl * HYX + yp * HX + (pair ? shift2 : 0)];
#pragma unroll
                    for (int jz = 0; jz < BZ; ++jz)
#pragma unroll
                        for (int jy = 0; jy < 2; ++jy) {
                            const int tz = zp - jz, ty = yp - jy;
                            if (tz >= 0 && tz <= 2 && ty >= 0 && ty <= 2)
                                acc[jy][jz] = mfma(A[2 * (3 * tz + ty) + pair], b, acc[jy][jz]);
                        }
                }
        }
#pragma unroll
        for (int jz = 0; jz < BZ; ++jz)
#pragma unroll
            for (int jy = 0; jy < 2; ++jy) {
                const int gx = x0 + col, gy = y0 + 2 * wv + jy, gz = z0 + jz;
                if (gx >= nx || gy >= ny || gz >= nz)
                    continue;
                const size_t v = vidx(gx, gy, gz, nx, ny);
                if (EP == EP_RELU || EP == EP_MASK) {
                    store4<EP>(acc[jy][jz], g, v, bias, out16, mask);
                } else if (g == 0) {
                    const float z4 = acc[jy][jz][0] + bias[0];
                    if (EP == EP_OUT) {
                        outf[v] = x[v] + z4 / s;
                    } else {
                        const float res = (s * x[v] + z4) - s * target[v];
                        outf[v] = res * gscale;
                        lsum += (double)res * (double)res;
                    }
                }
            }
    }
    if (EP == EP_LOSS) {
        // fixed-order block reduction of the squared residuals
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1)
            lsum += __shfl_xor(lsum, d);
        if (lane == 0)
            red[wv] = lsum;
        __syncthreads();
        if (tid == 0)
            partial[((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] =
                ((red[0] + red[1]) + red[2]) + red[3];
    }
}

// 1 -> 16 channel convolution of an FP32 scalar field (rounded to bf16 after
// scaling by sc): the first layer (input s*x), or the input gradient of the
// last layer (input dz4, ReLU mask of a3).
template <int EP>
__global__ __launch_bounds__(256) void conv1_kernel(int nx, int ny, int nz, const float *__restrict__ in, float sc,
                                                    const u32x4 *__restrict__ wpk, const float *__restrict__ bias,
                                                    unsigned short *__restrict__ out16,
                                                    const unsigned short *__restrict__ mask)
{
    __shared__ unsigned short lds[HVOX];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int col = lane & 15, g = lane >> 4;
    const int x0 = blockIdx.x * BX, y0 = blockIdx.y * BY;
    const u32x4 A = wpk[lane];
    int off[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int t = 8 * g + j;
        const int tt = t < 27 ? t : 13;
        off[j] = (tap_dz(tt) + 1) * HYX + (tap_dy(tt) + 1) * HX + (tap_dx(tt) + 1) + col;
    }
    const int zc_end = min((int)(blockIdx.z + 1) * ZPB, (nz + BZ - 1) / BZ);
    for (int zc = blockIdx.z * ZPB; zc < zc_end; ++zc) {
        const int z0 = zc * BZ;
        // the halo's z-slabs as one buffer resource (reads outside the
        // volume load zeros); every load is issued before the barrier
        const int zlo = max(z0 - 1, 0), zhi = min(z0 + BZ, nz - 1);
        const unsigned span = (unsigned)((unsigned long long)(zhi - zlo + 1) * nx * ny * 4ull);
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc((void *)(in + (size_t)zlo * nx * ny), (short)0, (int)span, 0x00020000);
        float hv[SPT];
        bool hok[SPT];
#pragma unroll
        for (int k = 0; k < SPT; ++k) {
            const int q = tid + 256 * k;
            const int lx = q % HX, ly = (q / HX) % HY, lz = q / HYX;
            const int gx = x0 + lx - 1, gy = y0 + ly - 1, gz = z0 + lz - 1;
            hok[k] = q < HVOX && gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz;
            const unsigned off =
                hok[k] ? (((unsigned)(gz - zlo) * (unsigned)ny + (unsigned)gy) * (unsigned)nx + (unsigned)gx) * 4u : span;
            hv[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, (int)off, 0, 0));
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < SPT; ++k) {
            const int q = tid + 256 * k;
            if (q < HVOX)
                lds[q] = (unsigned short)(hok[k] ? f2bf(sc * hv[k]) : 0u);
        }
        __syncthreads();
        for (int r = wv; r < BY * BZ; r += 4) {
            const int ry = r % BY, rz = r / BY;
            const int base = rz * HYX + ry * HX;
            unsigned int b[8];
#pragma unroll
            for (int j = 0; j < 8; ++j)
                b[j] = (8 * g + j < 27) ? (unsigned int)lds[base + off[j]] : 0u;
            const f32x4 acc = mfma(A, pack8(b), f32x4{0.f, 0.f, 0.f, 0.f});
            const int gx = x0 + col, gy = y0 + ry, gz = z0 + rz;
            if (gx < nx && gy < ny && gz < nz)
                store4<EP>(acc, g, vidx(gx, gy, gz, nx, ny), bias, out16, mask);
        }
    }
}

// Weight (and bias) gradient of one layer: dW[o][i][t] = sum_v gz[v][o] a[v + off(t)][i],
// db[o] = sum_v gz[v][o].  Rows: o (COUT real of 16); columns: (tap, i) in
// blocks of 16 -- CIN 16: block t = tap t (i = column), block 27 = ones
// column; CIN 1: column = tap, column 27 = ones.  GSCALAR: gz is FP32
// [V] (rounded to bf16); ASCALAR: a is FP32 [V] scaled by as (rounded).
template <int CIN, int COUT, bool GSCALAR, bool ASCALAR>
__global__ __launch_bounds__(256) void wgrad_kernel(int nx, int ny, int nz, const void *__restrict__ gsrc,
                                                    const void *__restrict__ asrc, float as,
                                                    float *__restrict__ partial, int nbx, int nby, int nbricks)
{
    constexpr int NB = CIN == 16 ? 28 : 2;
    __shared__ __attribute__((aligned(16))) unsigned short gT[COUT][WVOX];
    __shared__ unsigned short aT[CIN][WHV];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int col = lane & 15, g = lane >> 4;
    f32x4 acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
        acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int brick = blockIdx.x; brick < nbricks; brick += gridDim.x) {
        const int bx = brick % nbx, by = (brick / nbx) % nby, bz = brick / (nbx * nby);
        const int x0 = bx * WX, y0 = by * WY, z0 = bz * WZ;
        __syncthreads();  // the previous brick's operands are consumed
        if (GSCALAR) {
            const float *gf = static_cast<const float *>(gsrc);
            for (int q = tid; q < WVOX; q += 256) {
                const int gx = x0 + (q & 31), gy = y0 + ((q >> 5) & 3), gz = z0 + (q >> 7);
                unsigned int v = 0u;
                if (gx < nx && gy < ny && gz < nz)
                    v = f2bf(gf[vidx(gx, gy, gz, nx, ny)]);
                gT[0][q] = (unsigned short)v;
            }
        } else {
            const u32x4 *g16 = static_cast<const u32x4 *>(gsrc);
            for (int c = tid; c < 2 * WVOX; c += 256) {
                const int q = c >> 1, h = c & 1;
                const int gx = x0 + (q & 31), gy = y0 + ((q >> 5) & 3), gz = z0 + (q >> 7);
                u32x4 v = {0u, 0u, 0u, 0u};
                if (gx < nx && gy < ny && gz < nz)
                    v = g16[vidx(gx, gy, gz, nx, ny) * 2 + h];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    gT[8 * h + 2 * j][q] = (unsigned short)(v[j] & 0xFFFFu);
                    gT[8 * h + 2 * j + 1][q] = (unsigned short)(v[j] >> 16);
                }
            }
        }
        if (ASCALAR) {
            const float *af = static_cast<const float *>(asrc);
            for (int q = tid; q < WHV; q += 256) {
                const int lx = q % WHX, ly = (q / WHX) % WHY, lz = q / (WHX * WHY);
                const int gx = x0 + lx - 1, gy = y0 + ly - 1, gz = z0 + lz - 1;
                unsigned int v = 0u;
                if (gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz)
                    v = f2bf(as * af[vidx(gx, gy, gz, nx, ny)]);
                aT[0][q] = (unsigned short)v;
            }
        } else {
            const u32x4 *a16 = static_cast<const u32x4 *>(asrc);
            for (int c = tid; c < 2 * WHV; c += 256) {
                const int q = c >> 1, h = c & 1;
                const int lx = q % WHX, ly = (q / WHX) % WHY, lz = q / (WHX * WHY);
                const int gx = x0 + lx - 1, gy = y0 + ly - 1, gz = z0 + lz - 1;
                u32x4 v = {0u, 0u, 0u, 0u};
                if (gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz)
                    v = a16[vidx(gx, gy, gz, nx, ny) * 2 + h];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    aT[8 * h + 2 * j][q] = (unsigned short)(v[j] & 0xFFFFu);
                    aT[8 * h + 2 * j + 1][q] = (unsigned short)(v[j] >> 16);
                }
            }
        }
        __syncthreads();
        // each wave takes 4 of the 16 k-chunks (rows of 32 voxels)
        for (int kc = wv; kc < 16; kc += 4) {
            const int ky = kc & 3, kz = kc >> 2;
            u32x4 Aop = {0u, 0u, 0u, 0u};
            if (col < COUT)
                Aop = *reinterpret_cast<const u32x4 *>(&gT[col][kc * 32 + 8 * g]);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const int t = CIN == 16 ? nb : 16 * nb + col;
                const int ci = CIN == 16 ? col : 0;
                unsigned int e[8];
                if (t < 27) {
                    const int base = ((kz + tap_dz(t) + 1) * WHY + (ky + tap_dy(t) + 1)) * WHX + 8 * g + tap_dx(t) + 1;
#pragma unroll
                    for (int j = 0; j < 8; ++j)
                        e[j] = aT[ci][base + j];
                } else {
                    const unsigned int one = (t == 27 && (CIN == 1 || col == 0)) ? 0x3F80u : 0u;
#pragma unroll
                    for (int j = 0; j < 8; ++j)
                        e[j] = one;
                }
                acc[nb] = mfma(Aop, pack8(e), acc[nb]);
            }
        }
    }
    // the four waves hold partial sums over different k-chunks: add them in
    // wave order through LDS, then write this workgroup's partial
    __syncthreads();
    float *stage = reinterpret_cast<float *>(&aT[0][0]);  // reused: needs 16*NB*16 floats
    constexpr int NCOL = NB * 16;
    static_assert(sizeof(aT) >= 16 * NCOL * 4 || CIN == 1, "staging space");
    float *out = partial + (size_t)blockIdx.x * 16 * NCOL;
    if (CIN == 16) {
        for (int w = 0; w < 4; ++w) {
            if (wv == w) {
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        float &dst = stage[(4 * g + i) * NCOL + nb * 16 + col];
                        dst = (w == 0 ? 0.f : dst) + acc[nb][i];
                    }
            }
            __syncthreads();
        }
        for (int k = tid; k < 16 * NCOL; k += 256)
            out[k] = stage[k];
    } else {
        __shared__ float st1[16 * 32];
        for (int w = 0; w < 4; ++w) {
            if (wv == w) {
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        float &dst = st1[(4 * g + i) * NCOL + nb * 16 + col];
                        dst = (w == 0 ? 0.f : dst) + acc[nb][i];
                    }
            }
            __syncthreads();
        }
        for (int k = tid; k < 16 * NCOL; k += 256)
            out[k] = st1[k];
    }
}

// Sums the workgroup partials (fixed order, FP64) and scatters them into the
// gradient blob: W[o][i][t] at wofs, b[o] at wofs + cout*cin*27.
__global__ void wgrad_reduce_kernel(const float *__restrict__ partial, int nblk, int cin, int cout, int ncol,
                                    float *__restrict__ grad)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 16 * ncol)
        return;
    const int o = k / ncol, c = k % ncol;
    if (o >= cout)
        return;
    int dst;
    if (cin == 16) {
        const int nb = c / 16, i = c % 16;
        if (nb < 27)
            dst = (o * cin + i) * 27 + nb;
        else if (nb == 27 && i == 0)
            dst = cout * cin * 27 + o;
        else
            return;
    } else {
        if (c < 27)
            dst = o * 27 + c;
        else if (c == 27)
            dst = cout * 27 + o;
        else
            return;
    }
    double s = 0.0;
    for (int b = 0; b < nblk; ++b)
        s += (double)partial[(size_t)b * 16 * ncol + k];
    grad[dst] = (float)s;
}

__global__ void sum_f64_kernel(const double *__restrict__ p, int n, double *__restrict__ out)
{
    __shared__ double sm[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256)
        s += p[i];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int d = 128; d >= 1; d >>= 1) {
        if ((int)threadIdx.x < d)
            sm[threadIdx.x] += sm[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *out = sm[0];
}

__global__ void adam_kernel(float *__restrict__ p, const float *__restrict__ gr, float *__restrict__ m,
                            float *__restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float g = gr[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * (g * g);
    m[i] = mi;
    v[i] = vi;
    p[i] = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---- layout of the parameter blob and the workspace -------------------------
constexpr int P_W1 = 0, P_B1 = 432, P_W2 = 448, P_B2 = 448 + 6912, P_W3 = P_B2 + 16, P_B3 = P_W3 + 6912,
              P_W4 = P_B3 + 16, P_B4 = P_W4 + 432, P_END = P_B4 + 1;
static_assert(P_END == 14737, "parameter count");

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

struct Layout {
    size_t wpk, a1, a2, a3, dz4, dz3, dz2, dz1, wpart, lpart, total;
    int gx, gy, gz, nconv, nbx, nby, nbricks, wgrid;
};

Layout layout(int nx, int ny, int nz, bool training)
{
    Layout L{};
    const size_t V = (size_t)nx * ny * nz;
    L.gx = (nx + BX - 1) / BX;
    L.gy = (ny + BY - 1) / BY;
    L.gz = ((nz + BZ - 1) / BZ + ZPB - 1) / ZPB;
    L.nconv = L.gx * L.gy * L.gz;
    L.nbx = (nx + WX - 1) / WX;
    L.nby = (ny + WY - 1) / WY;
    L.nbricks = L.nbx * L.nby * ((nz + WZ - 1) / WZ);
    L.wgrid = L.nbricks < WGRID ? L.nbricks : WGRID;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t r = o;
        o += align256(bytes);
        return r;
    };
    L.wpk = take((size_t)WV_COUNT * WV_OPS * 64 * 16);
    L.a1 = take(V * 32);
    L.a2 = take(V * 32);
    if (training) {
        L.a3 = take(V * 32);
        L.dz4 = take(V * 4);
        L.dz3 = take(V * 32);
        L.dz2 = take(V * 32);
        L.dz1 = take(V * 32);
        L.wpart = take((size_t)L.wgrid * 16 * 28 * 16 * 4);
        L.lpart = take((size_t)L.nconv * 8 + 8);
    } else {
        L.a3 = L.a1;
        L.dz4 = L.dz3 = L.dz2 = L.dz1 = L.wpart = L.lpart = 0;
    }
    L.total = o;
    return L;
}

kv_status check_dims(int nx, int ny, int nz)
{
    if (nx < 1 || ny < 1 || nz < 1 || nx > (1 << 24) || ny > (1 << 24) || nz > (1 << 24) ||
        (long long)nx * ny * nz > (1LL << 30)) {
        kv_set_error("bad CNN volume size %d x %d x %d", nx, ny, nz);
        return KV_ERR_INVALID;
    }
    // conv16_kernel addresses the BZ + 2 slices of a chunk's halo (32 bytes
    // per voxel) through one buffer resource with 32-bit offsets
    if ((unsigned long long)nx * ny * 32ull * (BZ + 2) + 32ull > 0xFFFFFFFFull) {
        kv_set_error("CNN slices of %d x %d voxels exceed the %d-voxel limit", nx, ny, (int)(0xFFFFFFFFull / 192));
        return KV_ERR_INVALID;
    }
    return KV_OK;
}

kv_status check_common(kv_context *ctx, int nx, int ny, int nz, bool training, void *work, size_t work_bytes)
{
    if (!ctx) {
        kv_set_error("null context");
        return KV_ERR_INVALID;
    }
    if (kv_status st = check_dims(nx, ny, nz))
        return st;
    if (!work || work_bytes < layout(nx, ny, nz, training).total) {
        kv_set_error("CNN workspace too small (%zu bytes, need %zu)", work_bytes, layout(nx, ny, nz, training).total);
        return KV_ERR_INVALID;
    }
    if (reinterpret_cast<uintptr_t>(work) % 256) {
        kv_set_error("CNN workspace must be 256-byte aligned");
        return KV_ERR_INVALID;
    }
    const hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) {
        kv_set_error("hipSetDevice(%d) failed: %s", ctx->device, hipGetErrorString(e));
        return KV_ERR_DEVICE;
    }
    return KV_OK;
}

template <class T>
T *at(void *base, size_t off)
{
    return reinterpret_cast<T *>(static_cast<char *>(base) + off);
}

template <int CIN, int COUT, bool GS, bool AS>
hipError_t launch_wgrad(const Layout &L, int nx, int ny, int nz, const void *gsrc, const void *asrc, float as,
                        float *partial, float *grad, hipStream_t s)
{
    wgrad_kernel<CIN, COUT, GS, AS><<<L.wgrid, 256, 0, s>>>(nx, ny, nz, gsrc, asrc, as, partial, L.nbx, L.nby,
                                                           L.nbricks);
    const int ncol = (CIN == 16 ? 28 : 2) * 16;
    wgrad_reduce_kernel<<<(16 * ncol + 255) / 256, 256, 0, s>>>(partial, L.wgrid, CIN, COUT, ncol, grad);
    return hipGetLastError();
}

}  // namespace

extern "C" {

kv_status kv_cnn_workspace_size(int nx, int ny, int nz, int training, size_t *bytes)
{
    if (!bytes) {
        kv_set_error("null output");
        return KV_ERR_INVALID;
    }
    if (kv_status st = check_dims(nx, ny, nz))
        return st;
    *bytes = layout(nx, ny, nz, training != 0).total;
    return KV_OK;
}

kv_status kv_cnn_workspace_offsets(int nx, int ny, int nz, size_t *offsets)
{
    if (!offsets) {
        kv_set_error("null output");
        return KV_ERR_INVALID;
    }
    if (kv_status st = check_dims(nx, ny, nz))
        return st;
    const Layout L = layout(nx, ny, nz, true);
    const size_t o[7] = {L.a1, L.a2, L.a3, L.dz4, L.dz3, L.dz2, L.dz1};
    memcpy(offsets, o, sizeof(o));
    return KV_OK;
}

kv_status kv_cnn_forward(kv_context *ctx, const float *d_params, float scale, const float *d_x, float *d_y, int nx,
                         int ny, int nz, void *d_work, size_t work_bytes)
{
    if (kv_status st = check_common(ctx, nx, ny, nz, false, d_work, work_bytes))
        return st;
    if (!d_params || !d_x || !d_y || !(scale > 0.f) || !std::isfinite(scale)) {
        kv_set_error("bad CNN arguments");
        return KV_ERR_INVALID;
    }
    if (d_x == d_y) {
        kv_set_error("kv_cnn_forward: output must not alias the input (the last layer reads x while writing y)");
        return KV_ERR_INVALID;
    }
    const Layout L = layout(nx, ny, nz, false);
    const dim3 grid(L.gx, L.gy, L.gz);
    hipStream_t s = ctx->stream;
    auto *a1 = at<unsigned short>(d_work, L.a1), *a2 = at<unsigned short>(d_work, L.a2);
    u32x4 *wpk = at<u32x4>(d_work, L.wpk);
    auto W = [&](int v) { return wpk + v * WV_OPS * 64; };
    auto c16 = [](const unsigned short *q) { return reinterpret_cast<const u32x4 *>(q); };
    const float *p = d_params;
    pack_weights_kernel<<<WV_COUNT, 64, 0, s>>>(p, wpk, P_W1, P_W2, P_W3, P_W4);
    conv1_kernel<EP_RELU><<<grid, 256, 0, s>>>(nx, ny, nz, d_x, scale, W(WV_L1), p + P_B1, a1, nullptr);
    conv16_kernel<EP_RELU><<<grid, 256, 0, s>>>(nx, ny, nz, c16(a1), W(WV_L2), p + P_B2, a2, nullptr, nullptr,
                                                nullptr, nullptr, 0.f, 0.f, nullptr);
    conv16_kernel<EP_RELU><<<grid, 256, 0, s>>>(nx, ny, nz, c16(a2), W(WV_L3), p + P_B3, a1, nullptr, nullptr,
                                                nullptr, nullptr, 0.f, 0.f, nullptr);
    conv16_kernel<EP_OUT><<<grid, 256, 0, s>>>(nx, ny, nz, c16(a1), W(WV_L4), p + P_B4, nullptr, nullptr, d_x,
                                               nullptr, d_y, scale, 0.f, nullptr);
    KV_TRY_HIP(hipGetLastError());
    return KV_OK;
}

kv_status kv_cnn_gradient(kv_context *ctx, const float *d_params, float scale, const float *d_x,
                          const float *d_target, int nx, int ny, int nz, void *d_work, size_t work_bytes,
                          float *d_grad, double *loss)
{
    if (kv_status st = check_common(ctx, nx, ny, nz, true, d_work, work_bytes))
        return st;
    if (!d_params || !d_x || !d_target || !d_grad || !(scale > 0.f) || !std::isfinite(scale)) {
        kv_set_error("bad CNN arguments");
        return KV_ERR_INVALID;
    }
    const Layout L = layout(nx, ny, nz, true);
    const dim3 grid(L.gx, L.gy, L.gz);
    hipStream_t s = ctx->stream;
    auto *a1 = at<unsigned short>(d_work, L.a1), *a2 = at<unsigned short>(d_work, L.a2),
         *a3 = at<unsigned short>(d_work, L.a3);
    auto *dz3 = at<unsigned short>(d_work, L.dz3), *dz2 = at<unsigned short>(d_work, L.dz2),
         *dz1 = at<unsigned short>(d_work, L.dz1);
    float *dz4 = at<float>(d_work, L.dz4), *wpart = at<float>(d_work, L.wpart);
    double *lpart = at<double>(d_work, L.lpart), *lsum = lpart + L.nconv;
    const float *p = d_params;
    const double nvox = (double)nx * ny * nz;
    const float gscale = (float)(2.0 / nvox);
    auto c16 = [](const unsigned short *q) { return reinterpret_cast<const u32x4 *>(q); };
    u32x4 *wpk = at<u32x4>(d_work, L.wpk);
    auto W = [&](int v) { return wpk + v * WV_OPS * 64; };

    pack_weights_kernel<<<WV_COUNT, 64, 0, s>>>(p, wpk, P_W1, P_W2, P_W3, P_W4);
    conv1_kernel<EP_RELU><<<grid, 256, 0, s>>>(nx, ny, nz, d_x, scale, W(WV_L1), p + P_B1, a1, nullptr);
    conv16_kernel<EP_RELU><<<grid, 256, 0, s>>>(nx, ny, nz, c16(a1), W(WV_L2), p + P_B2, a2, nullptr, nullptr,
                                                nullptr, nullptr, 0.f, 0.f, nullptr);
    conv16_kernel<EP_RELU><<<grid, 256, 0, s>>>(nx, ny, nz, c16(a2), W(WV_L3), p + P_B3, a3, nullptr, nullptr,
                                                nullptr, nullptr, 0.f, 0.f, nullptr);
    conv16_kernel<EP_LOSS><<<grid, 256, 0, s>>>(nx, ny, nz, c16(a3), W(WV_L4), p + P_B4, nullptr, nullptr, d_x,
                                                d_target, dz4, scale, gscale, lpart);
    sum_f64_kernel<<<1, 256, 0, s>>>(lpart, L.nconv, lsum);
    // layer 4: weights, then the gradient into a3 (mask a3 > 0)
    KV_TRY_HIP((launch_wgrad<16, 1, true, false>(L, nx, ny, nz, dz4, a3, 1.f, wpart, d_grad + P_W4, s)));
    conv1_kernel<EP_MASK><<<grid, 256, 0, s>>>(nx, ny, nz, dz4, 1.f, W(WV_L4T), nullptr, dz3, a3);
    // layer 3
    KV_TRY_HIP((launch_wgrad<16, 16, false, false>(L, nx, ny, nz, dz3, a2, 1.f, wpart, d_grad + P_W3, s)));
    conv16_kernel<EP_MASK><<<grid, 256, 0, s>>>(nx, ny, nz, c16(dz3), W(WV_L3T), nullptr, dz2, a2, nullptr, nullptr,
                                                nullptr, 0.f, 0.f, nullptr);
    // layer 2
    KV_TRY_HIP((launch_wgrad<16, 16, false, false>(L, nx, ny, nz, dz2, a1, 1.f, wpart, d_grad + P_W2, s)));
    conv16_kernel<EP_MASK><<<grid, 256, 0, s>>>(nx, ny, nz, c16(dz2), W(WV_L2T), nullptr, dz1, a1, nullptr, nullptr,
                                                nullptr, 0.f, 0.f, nullptr);
    // layer 1 (input s*x)
    KV_TRY_HIP((launch_wgrad<1, 16, false, true>(L, nx, ny, nz, dz1, d_x, scale, wpart, d_grad + P_W1, s)));
    KV_TRY_HIP(hipGetLastError());
    if (loss) {
        double h = 0.0;
        KV_TRY_HIP(hipMemcpyAsync(&h, lsum, sizeof(double), hipMemcpyDeviceToHost, s));
        KV_TRY_HIP(hipStreamSynchronize(s));
        *loss = h / nvox;
    }
    return KV_OK;
}

kv_status kv_cnn_adam(kv_context *ctx, float *d_params, const float *d_grad, float *d_m, float *d_v, int step,
                      float lr)
{
    if (!ctx || !d_params || !d_grad || !d_m || !d_v || step < 1 || !(lr > 0.f)) {
        kv_set_error("bad Adam arguments");
        return KV_ERR_INVALID;
    }
    KV_TRY_HIP(hipSetDevice(ctx->device));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = (float)(1.0 - std::pow(0.9, step)), c2 = (float)(1.0 - std::pow(0.999, step));
    adam_kernel<<<(P_END + 255) / 256, 256, 0, ctx->stream>>>(d_params, d_grad, d_m, d_v, P_END, lr, b1, b2, eps, c1,
                                                              c2);
    KV_TRY_HIP(hipGetLastError());
    return KV_OK;
}

}  // extern "C"
