// The learned initializer: a small 3-D CNN that predicts an affine start for
// the registration from a downsampled image pair, trained and run by this
// library's own kernels. No counterpart in the reference, whose initializers
// are geometric (src/registration/initializer.rs: CenteredTransformInitializer).
//
// Network (input 64^3 x 2 channels, fixed and moving, z-scored, BF16):
//   conv1  2 -> 16, 3^3, stride 1, ReLU        64^3
//   conv2 16 -> 32, 3^3, stride 2, ReLU        32^3
//   conv3 32 -> 64, 3^3, stride 2, ReLU        16^3
//   conv4 64 -> 64, 3^3, stride 2, ReLU         8^3
//   conv5 64 ->128, 3^3, stride 2, ReLU         4^3
//   global average pool -> 128
//   fc1 128 -> 128, ReLU ; fc2 128 -> 12
// Zero 'same' padding. Outputs: the 9 entries of (A - I) and t / 10 mm of the
// affine x -> A (x - c) + t + c about the fixed image's centre c.
//
// Layout: activations channels-last [n][z][y][x][c] in BF16 (input channels
// padded to 8 with zeros so every 8-wide K group is one 16-byte load); conv
// weights [co][tap][ci] (tap = kz*9 + ky*3 + kx). Every convolution -- forward,
// input gradient, weight gradient -- is one implicit GEMM on the gfx950 BF16
// matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation): a workgroup of 4
// waves computes a (32 WM) x (32 WN) tile, each wave 2 x 2 MFMA tiles, with the
// A and B fragments fetched by per-GEMM gather functors. The weight gradient
// is split over K (the output positions) into fixed partial rows that a second
// kernel adds in order, so training is deterministic. Master weights, their
// gradients and Adam's moments are FP32; the MLP head (0.01% of the FLOPs) runs
// in FP32 on the vector units.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "image.hpp"
#include "metric.hpp"

namespace sitkhip {
namespace cnn {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr int kLayers = 5;
constexpr int kEdge = 64;          // input edge
constexpr int kCin0 = 8;           // padded input channels (2 used)
constexpr int kFeat = 128;         // GAP features
constexpr int kHidden = 128;
constexpr int kOut = 12;
constexpr int kCh[kLayers + 1] = {kCin0, 16, 32, 64, 64, 128};
constexpr int kStride[kLayers] = {1, 2, 2, 2, 2};

// ------------------------------------------------------------------ BF16
__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);                       // round to nearest even
    return static_cast<uint16_t>(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
    const uint32_t u = static_cast<uint32_t>(h) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

union Frag {
    uint4 u;
    bf16x8 v;
    uint16_t h[8];
};

__device__ __forceinline__ bf16x8 zero_frag() {
    Frag f;
    f.u = make_uint4(0, 0, 0, 0);
    return f.v;
}
__device__ __forceinline__ bf16x8 load_frag(const uint16_t *p) {
    Frag f;
    f.u = *reinterpret_cast<const uint4 *>(p);
    return f.v;
}

// One convolution's geometry (all edges and channel counts powers of two).
struct ConvGeom {
    int n;                      // batch
    int lg_di, lg_do;           // log2 input / output edge
    int lg_ci, lg_co;           // log2 input / output channels
    int stride;                 // 1 or 2
    int ci, co;
    int di, dout;
};

__device__ __forceinline__ void split_pos(int m, int lg, int &n, int &z, int &y, int &x) {
    const int mask = (1 << lg) - 1;
    x = m & mask;
    y = (m >> lg) & mask;
    z = (m >> (2 * lg)) & mask;
    n = m >> (3 * lg);
}

__device__ __forceinline__ size_t pos_index(int n, int z, int y, int x, int lg) {
    return (((((static_cast<size_t>(n) << lg) + z) << lg) + y) << lg) + x;
}

__device__ __forceinline__ void split_tap(int tap, int &kz, int &ky, int &kx) {
    kz = tap / 9;
    const int r = tap - 9 * kz;
    ky = r / 3;
    kx = r - 3 * ky;
}

// ------------------------------------------------------- gather functors
// Forward, A: im2col of the input. Row m = output position, k = tap * ci + c.
struct FwdA {
    const uint16_t *x;
    ConvGeom g;
    __device__ __forceinline__ bf16x8 operator()(int m, int k) const {
        const int M = g.n << (3 * g.lg_do);
        const int K = 27 << g.lg_ci;
        if (m >= M || k >= K) return zero_frag();
        const int tap = k >> g.lg_ci, c0 = k & (g.ci - 1);
        int kz, ky, kx, n, zo, yo, xo;
        split_tap(tap, kz, ky, kx);
        split_pos(m, g.lg_do, n, zo, yo, xo);
        const int zi = zo * g.stride + kz - 1, yi = yo * g.stride + ky - 1, xi = xo * g.stride + kx - 1;
        if (zi < 0 || yi < 0 || xi < 0 || zi >= g.di || yi >= g.di || xi >= g.di) return zero_frag();
        const size_t pos = pos_index(n, zi, yi, xi, g.lg_di);
        return load_frag(x + (pos << g.lg_ci) + c0);
    }
};

// A dense row-major BF16 matrix [rows][ld] read as 8-wide K groups.
struct RowsB {
    const uint16_t *w;
    int rows, K, ld;
    __device__ __forceinline__ bf16x8 operator()(int r, int k) const {
        if (r >= rows || k >= K) return zero_frag();
        return load_frag(w + static_cast<size_t>(r) * ld + k);
    }
};

// Input gradient, A: row m = input position, k = tap * co + c. The output
// position that tap reads this input from: zo = (zi + 1 - kz) / stride when
// that divides and lands inside.
struct DgradA {
    const uint16_t *dy;  // masked output gradient, channels-last
    ConvGeom g;
    __device__ __forceinline__ bf16x8 operator()(int m, int k) const {
        const int M = g.n << (3 * g.lg_di);
        const int K = 27 << g.lg_co;
        if (m >= M || k >= K) return zero_frag();
        const int tap = k >> g.lg_co, c0 = k & (g.co - 1);
        int kz, ky, kx, n, zi, yi, xi;
        split_tap(tap, kz, ky, kx);
        split_pos(m, g.lg_di, n, zi, yi, xi);
        int zo = zi + 1 - kz, yo = yi + 1 - ky, xo = xi + 1 - kx;
        if (g.stride == 2) {
            if ((zo | yo | xo) & 1) return zero_frag();
            zo >>= 1;
            yo >>= 1;
            xo >>= 1;
        }
        if (zo < 0 || yo < 0 || xo < 0 || zo >= g.dout || yo >= g.dout || xo >= g.dout) return zero_frag();
        const size_t pos = pos_index(n, zo, yo, xo, g.lg_do);
        return load_frag(dy + (pos << g.lg_co) + c0);
    }
};

// Weight gradient, A: row = output channel co, k = 8 consecutive output
// positions (gathered with stride co).
struct WgradA {
    const uint16_t *dy;
    ConvGeom g;
    __device__ __forceinline__ bf16x8 operator()(int co, int k) const {
        const int M = g.n << (3 * g.lg_do);
        Frag f;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            f.h[j] = (co < g.co && k + j < M) ? dy[(static_cast<size_t>(k + j) << g.lg_co) + co] : 0;
        return f.v;
    }
};

// Weight gradient, B: row = (tap, ci), k = 8 consecutive output positions:
// the input value that each position reads through that tap.
struct WgradB {
    const uint16_t *x;
    ConvGeom g;
    __device__ __forceinline__ bf16x8 operator()(int r, int k) const {
        const int M = g.n << (3 * g.lg_do);
        const int R = 27 << g.lg_ci;
        Frag f;
        if (r >= R) {
            f.u = make_uint4(0, 0, 0, 0);
            return f.v;
        }
        const int tap = r >> g.lg_ci, c = r & (g.ci - 1);
        int kz, ky, kx;
        split_tap(tap, kz, ky, kx);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            uint16_t v = 0;
            const int m = k + j;
            if (m < M) {
                int n, zo, yo, xo;
                split_pos(m, g.lg_do, n, zo, yo, xo);
                const int zi = zo * g.stride + kz - 1, yi = yo * g.stride + ky - 1,
                          xi = xo * g.stride + kx - 1;
                if (zi >= 0 && yi >= 0 && xi >= 0 && zi < g.di && yi < g.di && xi < g.di) {
                    const size_t pos = pos_index(n, zi, yi, xi, g.lg_di);
                    v = x[(pos << g.lg_ci) + c];
                }
            }
            f.h[j] = v;
        }
        return f.v;
    }
};

// ------------------------------------------------------------ epilogues
struct EpiBiasRelu {  // y = bf16(relu(acc + bias))
    uint16_t *y;
    const float *bias;
    int ldc;
    __device__ __forceinline__ void operator()(int, int m, int n, float acc) const {
        const float v = fmaxf(acc + bias[n], 0.0f);
        y[static_cast<size_t>(m) * ldc + n] = f2bf(v);
    }
};
struct EpiStoreBf16 {
    uint16_t *y;
    int ldc;
    __device__ __forceinline__ void operator()(int, int m, int n, float acc) const {
        y[static_cast<size_t>(m) * ldc + n] = f2bf(acc);
    }
};
struct EpiPartial {  // partial[split][m][n]
    float *p;
    int M, N;
    __device__ __forceinline__ void operator()(int split, int m, int n, float acc) const {
        p[(static_cast<size_t>(split) * M + m) * N + n] = acc;
    }
};

// ------------------------------------------------------------------ GEMM
// C[m][n] = sum_k A(m, k) B(n, k) over k in this workgroup's split
// [z * kchunk, min(K, (z + 1) * kchunk)); kchunk is a multiple of 32.
// MFMA 16x16x32 BF16 lane layout: lane l supplies A row (l mod 16) and B
// column (l mod 16) with K group 8 * (l / 16) .. + 7, and receives D rows
// 4 * (l / 16) .. + 3 of column (l mod 16).
template <int WM, int WN, bool PIPE, class FA, class FB, class EP>
__global__ __launch_bounds__(256) void mfma_gemm(FA fa, FB fb, EP ep, int M, int N, int K,
                                                 int kchunk) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wm = w % WM, wn = w / WM;
    const int m0 = blockIdx.x * (32 * WM) + wm * 32;
    const int n0 = blockIdx.y * (32 * WN) + wn * 32;
    const int kb = blockIdx.z * kchunk;
    const int ke = min(K, kb + kchunk);
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const int r = lane & 15, kg = (lane >> 4) * 8;
    // PIPE: the next K step's four fragments are fetched before the current
    // step's MFMAs issue, so their gathers overlap the math (measured: the
    // forward layers 1-4 139 -> 93 us each for 32 pairs; the input-gradient
    // GEMMs, whose gathers are already spread, 371 -> 418 us, so they do not)
    bf16x8 a0, a1, b0, b1;
    if (!PIPE) {
        for (int k = kb; k < ke; k += 32) {
            a0 = fa(m0 + r, k + kg);
            a1 = fa(m0 + 16 + r, k + kg);
            b0 = fb(n0 + r, k + kg);
            b1 = fb(n0 + 16 + r, k + kg);
            acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc[1][1], 0, 0, 0);
        }
    }
    if (PIPE && kb < ke) {
        a0 = fa(m0 + r, kb + kg);
        a1 = fa(m0 + 16 + r, kb + kg);
        b0 = fb(n0 + r, kb + kg);
        b1 = fb(n0 + 16 + r, kb + kg);
    }
    for (int k = kb; PIPE && k < ke; k += 32) {
        const int kn = k + 32;
        bf16x8 na0 = zero_frag(), na1 = zero_frag(), nb0 = zero_frag(), nb1 = zero_frag();
        if (kn < ke) {
            na0 = fa(m0 + r, kn + kg);
            na1 = fa(m0 + 16 + r, kn + kg);
            nb0 = fb(n0 + r, kn + kg);
            nb1 = fb(n0 + 16 + r, kn + kg);
        }
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc[1][1], 0, 0, 0);
        a0 = na0;
        a1 = na1;
        b0 = nb0;
        b1 = nb1;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int row = m0 + 16 * i + 4 * (lane >> 4) + e;
                const int col = n0 + 16 * j + (lane & 15);
                if (row < M && col < N) ep(blockIdx.z, row, col, acc[i][j][e]);
            }
}

template <bool PIPE = false, class FA, class FB, class EP>
static void gemm(const char *name, FA fa, FB fb, EP ep, int M, int N, int K, int splits = 1) {
    if (M <= 0 || N <= 0 || K <= 0) return;
    int kchunk = ((K + splits - 1) / splits + 31) / 32 * 32;
    splits = (K + kchunk - 1) / kchunk;
    LaunchScope scope(name);
    hipStream_t st = backend().stream;
    if (N <= 32) {
        dim3 grid((M + 127) / 128, (N + 31) / 32, splits);
        hipLaunchKernelGGL((mfma_gemm<4, 1, PIPE, FA, FB, EP>), grid, dim3(256), 0, st, fa, fb, ep, M, N, K,
                           kchunk);
    } else if (M <= 32) {
        dim3 grid((M + 31) / 32, (N + 127) / 128, splits);
        hipLaunchKernelGGL((mfma_gemm<1, 4, PIPE, FA, FB, EP>), grid, dim3(256), 0, st, fa, fb, ep, M, N, K,
                           kchunk);
    } else {
        dim3 grid((M + 63) / 64, (N + 63) / 64, splits);
        hipLaunchKernelGGL((mfma_gemm<2, 2, PIPE, FA, FB, EP>), grid, dim3(256), 0, st, fa, fb, ep, M, N, K,
                           kchunk);
    }
    check_launch(name);
}

// ------------------------------------------------- first layer, forward
// Layer 0 (2 -> 16 channels, 3^3, stride 1, 64^3). As a GEMM over the padded
// input it has K = 27 taps x 8 channels = 216, three quarters of it zeros;
// here K = 27 taps x the 2 used channels = 54, zero-padded to 64 (two MFMA K
// steps). A workgroup takes four output rows (along x) of one (n, z): it
// stages the 3 x 6 x 66 input voxels they read -- the 4 bytes of each
// voxel's two channels, zero outside the volume -- in LDS with coalesced
// loads, and each wave builds its B fragments from there. The product is
// taken transposed (weights x patches: D[co][position]), so a lane ends with 4
// consecutive output channels of one position, stored as one 8-byte word.
constexpr int kC0Rows = 4;  // output rows (waves) per workgroup

__global__ __launch_bounds__(256) void conv0_forward(const uint16_t *__restrict__ x,
                                                     const uint16_t *__restrict__ wb,
                                                     const float *__restrict__ bias,
                                                     uint16_t *__restrict__ y, int n) {
    constexpr int E = kEdge, LG = 6, CI = kCin0, CO = 16;
    constexpr int TY = kC0Rows + 2, TX = E + 2;
    __shared__ uint32_t tile[3][TY][TX];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int groups = E / kC0Rows;
    const int y0 = (blockIdx.x % groups) * kC0Rows;
    const int zz = (blockIdx.x / groups) % E;
    const int nn = blockIdx.x / (groups * E);
    for (int i = threadIdx.x; i < 3 * TY * TX; i += 256) {
        const int tx = i % TX, ty = (i / TX) % TY, tz = i / (TX * TY);
        const int zi = zz + tz - 1, yi = y0 + ty - 1, xi = tx - 1;
        uint32_t v = 0;
        if (nn < n && zi >= 0 && yi >= 0 && xi >= 0 && zi < E && yi < E && xi < E)
            v = *reinterpret_cast<const uint32_t *>(x + pos_index(nn, zi, yi, xi, LG) * CI);
        tile[tz][ty][tx] = v;
    }
    __syncthreads();
    if (nn >= n) return;
    const int yy = y0 + w;
    const int r = lane & 15, kg = (lane >> 4) * 8;
    bf16x8 a[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        Frag f;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = 32 * s + kg + j, tap = k >> 1, c = k & 1;
            f.h[j] = tap < 27 ? wb[(r * 27 + tap) * CI + c] : static_cast<uint16_t>(0);
        }
        a[s] = f.v;
    }
    f32x4 acc[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        acc[b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        const int xx = 16 * b + r;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            Frag f;
#pragma unroll
            for (int j4 = 0; j4 < 4; ++j4) {
                const int tap = 16 * s + (kg >> 1) + j4;
                uint32_t v = 0;
                if (tap < 27) {
                    int kz, ky, kx;
                    split_tap(tap, kz, ky, kx);
                    v = tile[kz][w + ky][xx + kx];
                }
                f.h[2 * j4] = static_cast<uint16_t>(v & 0xffffu);
                f.h[2 * j4 + 1] = static_cast<uint16_t>(v >> 16);
            }
            acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[s], f.v, acc[b], 0, 0, 0);
        }
    }
    const int co0 = 4 * (lane >> 4);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int xx = 16 * b + r;
        uint16_t h[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) h[e] = f2bf(fmaxf(acc[b][e] + bias[co0 + e], 0.0f));
        uint2 word;
        word.x = static_cast<uint32_t>(h[0]) | (static_cast<uint32_t>(h[1]) << 16);
        word.y = static_cast<uint32_t>(h[2]) | (static_cast<uint32_t>(h[3]) << 16);
        *reinterpret_cast<uint2 *>(y + pos_index(nn, zz, yy, xx, LG) * CO + co0) = word;
    }
}

// ------------------------------------------------------- small kernels
// out[i] = sum over `splits` rows of partial[s][i], in split order.
__global__ void sum_partials(const float *__restrict__ p, int splits, size_t len,
                             float *__restrict__ out) {
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < len;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        float s = 0.0f;
        for (int k = 0; k < splits; ++k) s += p[static_cast<size_t>(k) * len + i];
        out[i] = s;
    }
}

// dy_masked = dy * (y > 0), BF16 in place on dy.
__global__ void relu_mask(uint16_t *__restrict__ dy, const uint16_t *__restrict__ y, size_t n) {
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        if (!(bf2f(y[i]) > 0.0f)) dy[i] = 0;
}

// Bias gradient, first stage: partial[s][c] = sum of dy[m][c] over the
// positions of split s (workgroup (s), lanes over channels then positions,
// fixed tree). The second stage is sum_partials over the splits.
constexpr int kBiasSplits = 256;
__global__ __launch_bounds__(256) void bias_grad_partials(const uint16_t *__restrict__ dy, int M,
                                                          int C, float *__restrict__ partial) {
    __shared__ float red[256];
    const int s = blockIdx.x;
    const int per = (M + kBiasSplits - 1) / kBiasSplits;
    const int m0 = s * per, m1 = min(M, m0 + per);
    const int lanes_per_c = 256 / C;  // C divides 256 (16..128)
    const int c = threadIdx.x % C, part = threadIdx.x / C;
    float acc = 0.0f;
    for (int m = m0 + part; m < m1; m += lanes_per_c) acc += bf2f(dy[static_cast<size_t>(m) * C + c]);
    red[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < C) {
        float t = 0.0f;
        for (int p = 0; p < lanes_per_c; ++p) t += red[p * C + threadIdx.x];
        partial[static_cast<size_t>(s) * C + threadIdx.x] = t;
    }
}

// Global average pool of the last conv output [n][64][128] -> feat [n][128].
__global__ void gap_forward(const uint16_t *__restrict__ y, int n, float *__restrict__ feat) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * kFeat) return;
    const int b = i / kFeat, c = i - b * kFeat;
    float s = 0.0f;
    for (int p = 0; p < 64; ++p) s += bf2f(y[(static_cast<size_t>(b) * 64 + p) * kFeat + c]);
    feat[i] = s * (1.0f / 64.0f);
}

__global__ void gap_backward(const float *__restrict__ dfeat, int n, uint16_t *__restrict__ dy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * 64 * kFeat) return;
    const int c = i % kFeat, b = i / (64 * kFeat);
    dy[i] = f2bf(dfeat[b * kFeat + c] * (1.0f / 64.0f));
}

// MLP head forward, one workgroup (128 lanes) per sample.
__global__ __launch_bounds__(128) void mlp_forward(const float *__restrict__ feat,
                                                   const float *__restrict__ w1,
                                                   const float *__restrict__ b1,
                                                   const float *__restrict__ w2,
                                                   const float *__restrict__ b2,
                                                   float *__restrict__ hid, float *__restrict__ out) {
    __shared__ float f[kFeat], h[kHidden];
    const int b = blockIdx.x, t = threadIdx.x;
    f[t] = feat[b * kFeat + t];
    __syncthreads();
    float s = b1[t];
    for (int i = 0; i < kFeat; ++i) s += w1[t * kFeat + i] * f[i];
    s = fmaxf(s, 0.0f);
    h[t] = s;
    hid[b * kHidden + t] = s;
    __syncthreads();
    if (t < kOut) {
        float o = b2[t];
        for (int j = 0; j < kHidden; ++j) o += w2[t * kHidden + j] * h[j];
        out[b * kOut + t] = o;
    }
}

// MLP head backward over the whole batch (one workgroup, fixed loop orders):
// gradients of w2, b2, w1, b1 and of the pooled features.
__global__ __launch_bounds__(128) void mlp_backward(int n, const float *__restrict__ feat,
                                                    const float *__restrict__ hid,
                                                    const float *__restrict__ dout,
                                                    const float *__restrict__ w1,
                                                    const float *__restrict__ w2,
                                                    float *__restrict__ gw1, float *__restrict__ gb1,
                                                    float *__restrict__ gw2, float *__restrict__ gb2,
                                                    float *__restrict__ dhid,
                                                    float *__restrict__ dfeat) {
    const int t = threadIdx.x;
    // dhid[b][t] = relu'(hid) * sum_o w2[o][t] dout[b][o]
    for (int b = 0; b < n; ++b) {
        float s = 0.0f;
        for (int o = 0; o < kOut; ++o) s += w2[o * kHidden + t] * dout[b * kOut + o];
        dhid[b * kHidden + t] = hid[b * kHidden + t] > 0.0f ? s : 0.0f;
    }
    __syncthreads();
    // gw2[o][t] = sum_b dout[b][o] hid[b][t]; gb2[o] = sum_b dout[b][o]
    for (int o = 0; o < kOut; ++o) {
        float s = 0.0f;
        for (int b = 0; b < n; ++b) s += dout[b * kOut + o] * hid[b * kHidden + t];
        gw2[o * kHidden + t] = s;
    }
    if (t < kOut) {
        float s = 0.0f;
        for (int b = 0; b < n; ++b) s += dout[b * kOut + t];
        gb2[t] = s;
    }
    // gw1[t][i] = sum_b dhid[b][t] feat[b][i]; gb1[t] = sum_b dhid[b][t]
    for (int i = 0; i < kFeat; ++i) {
        float s = 0.0f;
        for (int b = 0; b < n; ++b) s += dhid[b * kHidden + t] * feat[b * kFeat + i];
        gw1[t * kFeat + i] = s;
    }
    {
        float s = 0.0f;
        for (int b = 0; b < n; ++b) s += dhid[b * kHidden + t];
        gb1[t] = s;
    }
    // dfeat[b][t] = sum_j w1[j][t] dhid[b][j]
    for (int b = 0; b < n; ++b) {
        float s = 0.0f;
        for (int j = 0; j < kHidden; ++j) s += w1[j * kFeat + t] * dhid[b * kHidden + j];
        dfeat[b * kFeat + t] = s;
    }
}

// Adam (Kingma & Ba 2015) on the flat FP32 parameters.
__global__ void adam_step(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                          float *__restrict__ v, size_t n, float lr, float b1, float b2, float eps,
                          float c1, float c2) {
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// BF16 copies of one conv layer's weights: w [co][27][ci] and the transposed
// wt [ci][27][co] the input-gradient GEMM reads.
__global__ void cast_weights(const float *__restrict__ w, int co, int ci, uint16_t *__restrict__ wb,
                             uint16_t *__restrict__ wt) {
    const int n = co * 27 * ci;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int c = i % ci, tap = (i / ci) % 27, o = i / (27 * ci);
        const uint16_t h = f2bf(w[i]);
        wb[i] = h;
        wt[(static_cast<size_t>(c) * 27 + tap) * co + o] = h;
    }
}

// Sum and sum of squares of a volume (FP64 partials per workgroup); volume
// blockIdx.y of consecutive volumes of n voxels, its partials 2 * gridDim.x
// apart.
__global__ __launch_bounds__(256) void moments_partials(const float *__restrict__ v, size_t n,
                                                        double *__restrict__ out) {
    __shared__ double s1[256], s2[256];
    v += static_cast<size_t>(blockIdx.y) * n;
    out += static_cast<size_t>(blockIdx.y) * 2 * gridDim.x;
    double a = 0.0, b = 0.0;
    for (size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * 256) {
        const double x = v[i];
        a += x;
        b += x * x;
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s1[threadIdx.x] += s1[threadIdx.x + w];
            s2[threadIdx.x] += s2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = s1[0];
        out[2 * blockIdx.x + 1] = s2[0];
    }
}

// One input slot: channel `ch` of [64^3][8] = (v - mean) * inv_std as BF16;
// channels 2..7 are zero.
// Each volume's normalisation from its 256 partial (sum, sum of squares)
// pairs, added in order by one lane (lane t: volume t of nvol): norm[2t..] =
// {mean, 1 / std} as floats.
__global__ void moments_finish(const double *__restrict__ part, size_t nv, float *__restrict__ norm,
                               int nvol) {
    const int t = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= nvol) return;
    part += static_cast<size_t>(t) * 512;
    norm += 2 * t;
    double s1 = 0.0, s2 = 0.0;
    for (int b = 0; b < 256; ++b) {
        s1 += part[2 * b];
        s2 += part[2 * b + 1];
    }
    const double mean = s1 / static_cast<double>(nv);
    const double var = fmax(s2 / static_cast<double>(nv) - mean * mean, 1e-12);
    norm[0] = static_cast<float>(mean);
    norm[1] = static_cast<float>(1.0 / sqrt(var));
}

// Slots first .. first + gridDim.y - 1 from their two volumes (fixed, moving:
// volumes 2b and 2b + 1 of `v`, normalisations 4b .. 4b + 3): voxel i of
// slot b is {(f - mean_f) / std_f, (m - mean_m) / std_m, 0, ..., 0} as BF16.
__global__ void pack_slots(const float *__restrict__ v, const float *__restrict__ norm,
                           uint16_t *__restrict__ slots) {
    constexpr int nv = kEdge * kEdge * kEdge;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    const int b = blockIdx.y;
    const float *f = v + static_cast<size_t>(2 * b) * nv;
    const float *m = f + nv;
    const float *q = norm + 4 * b;
    uint16_t *o = slots + (static_cast<size_t>(b) * nv + i) * kCin0;
    o[0] = f2bf((f[i] - q[0]) * q[1]);
    o[1] = f2bf((m[i] - q[2]) * q[3]);
#pragma unroll
    for (int c = 2; c < kCin0; ++c) o[c] = 0;
}

// --------------------------------------------------------------- the net
struct Net {
    int max_batch = 0;
    // parameter offsets (floats) in the flat buffers
    size_t off_w[kLayers], off_b[kLayers], off_w1, off_b1, off_w2, off_b2, nparams;
    DeviceBuffer<float> params, grads, adam_m, adam_v;
    long long adam_t = 0;
    DeviceBuffer<uint16_t> wb[kLayers], wt[kLayers];
    DeviceBuffer<uint16_t> act[kLayers + 1];   // act[0] = input slots, act[l+1] = conv l output
    DeviceBuffer<uint16_t> dact[kLayers + 1];  // gradients w.r.t. act (masked in place)
    DeviceBuffer<float> partials;
    DeviceBuffer<float> feat, hid, out, dout, dhid, dfeat;
    bool weights_dirty = true;
};

static int out_edge(int layer) {  // edge of conv `layer`'s output (layer 0..4)
    int e = kEdge;
    for (int l = 0; l <= layer; ++l) e /= kStride[l];
    return e;
}
static int in_edge(int layer) { return layer == 0 ? kEdge : out_edge(layer - 1); }
static int lg2(int v) {
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}

static ConvGeom geom_of(int layer, int n) {
    ConvGeom g;
    g.n = n;
    g.ci = kCh[layer];
    g.co = kCh[layer + 1];
    g.di = in_edge(layer);
    g.dout = out_edge(layer);
    g.stride = kStride[layer];
    g.lg_ci = lg2(g.ci);
    g.lg_co = lg2(g.co);
    g.lg_di = lg2(g.di);
    g.lg_do = lg2(g.dout);
    return g;
}

static size_t act_elems(int idx, int n) {  // idx 0 = input, idx l+1 = conv l output
    const int e = idx == 0 ? kEdge : out_edge(idx - 1);
    const int c = idx == 0 ? kCin0 : kCh[idx];
    return static_cast<size_t>(n) * e * e * e * c;
}

static void init_params(Net &net, uint64_t seed) {
    std::vector<float> h(net.nparams, 0.0f);
    uint64_t s = seed;
    auto next = [&]() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    auto normal = [&]() {
        const double u1 = (static_cast<double>(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        const double u2 = (static_cast<double>(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    for (int l = 0; l < kLayers; ++l) {
        const int ci = kCh[l], co = kCh[l + 1];
        const int ci_used = l == 0 ? 2 : ci;
        const double sd = std::sqrt(2.0 / (27.0 * ci_used));  // He initialisation
        for (int o = 0; o < co; ++o)
            for (int t = 0; t < 27; ++t)
                for (int c = 0; c < ci; ++c)
                    h[net.off_w[l] + (static_cast<size_t>(o) * 27 + t) * ci + c] =
                        c < ci_used ? static_cast<float>(sd * normal()) : 0.0f;
    }
    const double sd1 = std::sqrt(2.0 / kFeat), sd2 = std::sqrt(1.0 / kHidden) * 0.1;
    for (int i = 0; i < kHidden * kFeat; ++i) h[net.off_w1 + i] = static_cast<float>(sd1 * normal());
    for (int i = 0; i < kOut * kHidden; ++i) h[net.off_w2 + i] = static_cast<float>(sd2 * normal());
    net.params.upload(h.data(), h.size());
    sync();
    net.weights_dirty = true;
}

Net *create(int max_batch, uint64_t seed) {
    if (max_batch < 1 || max_batch > 256) fail(SITKHIP_ERR_INVALID_ARGUMENT, "batch must be 1..256");
    auto *net = new Net();
    try {
        net->max_batch = max_batch;
        size_t o = 0;
        for (int l = 0; l < kLayers; ++l) {
            net->off_w[l] = o;
            o += static_cast<size_t>(kCh[l + 1]) * 27 * kCh[l];
            net->off_b[l] = o;
            o += kCh[l + 1];
        }
        net->off_w1 = o; o += kHidden * kFeat;
        net->off_b1 = o; o += kHidden;
        net->off_w2 = o; o += kOut * kHidden;
        net->off_b2 = o; o += kOut;
        net->nparams = o;
        net->params.alloc(o);
        net->grads.alloc(o);
        net->adam_m.alloc(o);
        net->adam_v.alloc(o);
        HIP_OK(hipMemsetAsync(net->grads.ptr, 0, o * sizeof(float), backend().stream));
        HIP_OK(hipMemsetAsync(net->adam_m.ptr, 0, o * sizeof(float), backend().stream));
        HIP_OK(hipMemsetAsync(net->adam_v.ptr, 0, o * sizeof(float), backend().stream));
        for (int l = 0; l < kLayers; ++l) {
            const size_t n = static_cast<size_t>(kCh[l + 1]) * 27 * kCh[l];
            net->wb[l].alloc(n);
            net->wt[l].alloc(n);
        }
        for (int i = 0; i <= kLayers; ++i) {
            net->act[i].alloc(act_elems(i, max_batch));
            net->dact[i].alloc(act_elems(i, max_batch));
        }
        HIP_OK(hipMemsetAsync(net->act[0].ptr, 0, net->act[0].bytes(), backend().stream));
        // weight-gradient partials: the largest (splits x Cout x 27 Cin)
        size_t most = 0;
        for (int l = 0; l < kLayers; ++l) {
            const ConvGeom g = geom_of(l, max_batch);
            const size_t M = static_cast<size_t>(g.n) << (3 * g.lg_do);
            const int splits = static_cast<int>(std::min<size_t>(256, std::max<size_t>(1, M / 4096)));
            most = std::max(most, static_cast<size_t>(splits) * g.co * 27 * g.ci);
        }
        net->partials.alloc(std::max<size_t>(most, static_cast<size_t>(kBiasSplits) * 128));
        net->feat.alloc(static_cast<size_t>(max_batch) * kFeat);
        net->hid.alloc(static_cast<size_t>(max_batch) * kHidden);
        net->out.alloc(static_cast<size_t>(max_batch) * kOut);
        net->dout.alloc(static_cast<size_t>(max_batch) * kOut);
        net->dhid.alloc(static_cast<size_t>(max_batch) * kHidden);
        net->dfeat.alloc(static_cast<size_t>(max_batch) * kFeat);
        init_params(*net, seed);
    } catch (...) {
        delete net;
        throw;
    }
    return net;
}

static void refresh_weights(Net &net) {
    if (!net.weights_dirty) return;
    for (int l = 0; l < kLayers; ++l) {
        LaunchScope scope("cnn_cast_weights");
        const int n = kCh[l + 1] * 27 * kCh[l];
        hipLaunchKernelGGL(cast_weights, dim3((n + 255) / 256), dim3(256), 0, backend().stream,
                           net.params.ptr + net.off_w[l], kCh[l + 1], kCh[l], net.wb[l].ptr,
                           net.wt[l].ptr);
        check_launch("cast_weights");
    }
    net.weights_dirty = false;
}

static void check_batch(const Net &net, int n) {
    if (n < 1 || n > net.max_batch)
        fail(SITKHIP_ERR_INVALID_ARGUMENT, "batch must be 1.." + std::to_string(net.max_batch));
}

void forward(Net &net, int n, float *out_host) {
    check_batch(net, n);
    refresh_weights(net);
    {
        LaunchScope scope("cnn_conv0_fwd");
        hipLaunchKernelGGL(conv0_forward, dim3(n * kEdge * (kEdge / kC0Rows)), dim3(256), 0,
                           backend().stream, net.act[0].ptr, net.wb[0].ptr,
                           net.params.ptr + net.off_b[0], net.act[1].ptr, n);
        check_launch("conv0_forward");
    }
    for (int l = 1; l < kLayers; ++l) {
        const ConvGeom g = geom_of(l, n);
        const int M = g.n << (3 * g.lg_do), N = g.co, K = 27 * g.ci;
        gemm<true>("cnn_conv_fwd", FwdA{net.act[l].ptr, g}, RowsB{net.wb[l].ptr, g.co, K, K},
             EpiBiasRelu{net.act[l + 1].ptr, net.params.ptr + net.off_b[l], g.co}, M, N, K);
    }
    {
        LaunchScope scope("cnn_head");
        hipLaunchKernelGGL(gap_forward, dim3((n * kFeat + 255) / 256), dim3(256), 0, backend().stream,
                           net.act[kLayers].ptr, n, net.feat.ptr);
        hipLaunchKernelGGL(mlp_forward, dim3(n), dim3(128), 0, backend().stream, net.feat.ptr,
                           net.params.ptr + net.off_w1, net.params.ptr + net.off_b1,
                           net.params.ptr + net.off_w2, net.params.ptr + net.off_b2, net.hid.ptr,
                           net.out.ptr);
        check_launch("cnn_head");
    }
    if (out_host) {
        net.out.download(out_host, static_cast<size_t>(n) * kOut);
    }
    sync();
}

// Backward from dL/dout (host, n x 12) after forward() of the same batch:
// every parameter gradient into net.grads.
void backward(Net &net, int n, const float *dout_host) {
    check_batch(net, n);
    hipStream_t st = backend().stream;
    net.dout.upload(dout_host, static_cast<size_t>(n) * kOut);
    {
        LaunchScope scope("cnn_head_bwd");
        hipLaunchKernelGGL(mlp_backward, dim3(1), dim3(128), 0, st, n, net.feat.ptr, net.hid.ptr,
                           net.dout.ptr, net.params.ptr + net.off_w1, net.params.ptr + net.off_w2,
                           net.grads.ptr + net.off_w1, net.grads.ptr + net.off_b1,
                           net.grads.ptr + net.off_w2, net.grads.ptr + net.off_b2, net.dhid.ptr,
                           net.dfeat.ptr);
        hipLaunchKernelGGL(gap_backward, dim3((n * 64 * kFeat + 255) / 256), dim3(256), 0, st,
                           net.dfeat.ptr, n, net.dact[kLayers].ptr);
        check_launch("cnn_head_bwd");
    }
    for (int l = kLayers - 1; l >= 0; --l) {
        const ConvGeom g = geom_of(l, n);
        const size_t ny = act_elems(l + 1, n);
        const int M = g.n << (3 * g.lg_do);
        {
            LaunchScope scope("cnn_relu_mask");
            hipLaunchKernelGGL(relu_mask, dim3(grid_for(ny, 256, 8192)), dim3(256), 0, st,
                               net.dact[l + 1].ptr, net.act[l + 1].ptr, ny);
            check_launch("cnn_relu_mask");
        }
        {
            LaunchScope scope("cnn_bias_grad");
            hipLaunchKernelGGL(bias_grad_partials, dim3(kBiasSplits), dim3(256), 0, st,
                               net.dact[l + 1].ptr, M, g.co, net.partials.ptr);
            hipLaunchKernelGGL(sum_partials, dim3(1), dim3(256), 0, st, net.partials.ptr, kBiasSplits,
                               static_cast<size_t>(g.co), net.grads.ptr + net.off_b[l]);
            check_launch("cnn_bias_grad");
        }
        // weight gradient: [co] x [27 ci] over K = M output positions, split
        const int R = 27 * g.ci;
        const int splits = static_cast<int>(std::min<long long>(256, std::max<long long>(1, M / 4096)));
        int kchunk = ((M + splits - 1) / splits + 31) / 32 * 32;
        const int nsplit = (M + kchunk - 1) / kchunk;
        gemm("cnn_conv_wgrad", WgradA{net.dact[l + 1].ptr, g}, WgradB{net.act[l].ptr, g},
             EpiPartial{net.partials.ptr, g.co, R}, g.co, R, M, splits);
        {
            LaunchScope scope("cnn_sum_partials");
            const size_t len = static_cast<size_t>(g.co) * R;
            hipLaunchKernelGGL(sum_partials, dim3(grid_for(len, 256, 1024)), dim3(256), 0, st,
                               net.partials.ptr, nsplit, len, net.grads.ptr + net.off_w[l]);
            check_launch("sum_partials");
        }
        if (l > 0) {
            // input gradient: M' = input positions, N' = ci, K' = 27 co
            const int Mi = g.n << (3 * g.lg_di);
            const int Kd = 27 * g.co;
            gemm("cnn_conv_dgrad", DgradA{net.dact[l + 1].ptr, g}, RowsB{net.wt[l].ptr, g.ci, Kd, Kd},
                 EpiStoreBf16{net.dact[l].ptr, g.ci}, Mi, g.ci, Kd);
        }
    }
    sync();
}

void adam(Net &net, float lr, float b1, float b2, float eps) {
    ++net.adam_t;
    const float c1 = 1.0f - std::pow(b1, static_cast<float>(net.adam_t));
    const float c2 = 1.0f - std::pow(b2, static_cast<float>(net.adam_t));
    LaunchScope scope("cnn_adam");
    hipLaunchKernelGGL(adam_step, dim3(grid_for(net.nparams, 256, 1024)), dim3(256), 0,
                       backend().stream, net.params.ptr, net.grads.ptr, net.adam_m.ptr,
                       net.adam_v.ptr, net.nparams, lr, b1, b2, eps, c1, c2);
    check_launch("adam_step");
    net.weights_dirty = true;
    sync();
}

// Input slots first .. first + n - 1 from 2n 64^3 float volumes already on
// the device, consecutive: slot b's fixed volume is volume 2b, its moving
// volume 2b + 1. The normalisations are computed on the device (no host round
// trip), every volume's in one launch.
static void pack(Net &net, int first, int n, const float *vols) {
    if (first < 0 || n < 1 || first + n > net.max_batch)
        fail(SITKHIP_ERR_INVALID_ARGUMENT, "slot out of range");
    const size_t nv = static_cast<size_t>(kEdge) * kEdge * kEdge;
    uint16_t *dst = net.act[0].ptr + static_cast<size_t>(first) * nv * kCin0;
    DeviceBuffer<double> part(static_cast<size_t>(2 * n) * 2 * 256);
    DeviceBuffer<float> norm(static_cast<size_t>(2 * n) * 2);
    hipLaunchKernelGGL(moments_partials, dim3(256, static_cast<unsigned>(2 * n)), dim3(256), 0,
                       backend().stream, vols, nv, part.ptr);
    check_launch("moments_partials");
    hipLaunchKernelGGL(moments_finish, dim3(static_cast<unsigned>((2 * n + 63) / 64)), dim3(64), 0,
                       backend().stream, part.ptr, nv, norm.ptr, 2 * n);
    check_launch("moments_finish");
    hipLaunchKernelGGL(pack_slots, dim3(static_cast<unsigned>((nv + 255) / 256), static_cast<unsigned>(n)),
                       dim3(256), 0, backend().stream, vols, norm.ptr, dst);
    check_launch("pack_slots");
}

// The 64^3 grid spanning an image's physical extent (same origin corner and
// direction, spacing size * spacing / 64): each output voxel is the linear
// interpolation at the centre of the input block it covers.
static sitkhip_geometry grid64(const sitkhip_geometry &g) {
    sitkhip_geometry o = g;
    for (int d = 0; d < 3; ++d) {
        o.size[d] = kEdge;
        o.spacing[d] = g.spacing[d] * static_cast<double>(g.size[d]) / kEdge;
    }
    // keep the physical box: origin moves by half the spacing change
    for (int r = 0; r < 3; ++r) {
        double acc = g.origin[r];
        for (int c = 0; c < 3; ++c) acc += g.direction[r * 3 + c] * 0.5 * (o.spacing[c] - g.spacing[c]);
        o.origin[r] = acc;
    }
    return o;
}

// Slots first .. first + n - 1 from n image pairs: every image linearly
// resampled (identity map) onto the 64^3 grid spanning its own extent -- all
// 2n placements in one launch -- then z-scored and packed.
void set_inputs(Net &net, int first, int n, const Image *const *fixed, const Image *const *moving) {
    if (first < 0 || n < 1 || first + n > net.max_batch)
        fail(SITKHIP_ERR_INVALID_ARGUMENT, "slot out of range");
    const size_t nv = static_cast<size_t>(kEdge) * kEdge * kEdge;
    DeviceBuffer<float> vols(static_cast<size_t>(2 * n) * nv);
    PrepBatch prep;
    for (int b = 0; b < n; ++b) {
        for (int ch = 0; ch < 2; ++ch) {
            const Image *im = ch ? moving[b] : fixed[b];
            if (!im) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null image");
            const sitkhip_geometry og = grid64(im->geom);
            prep.places.push_back({im->buf.ptr, im->geom, vols.ptr + static_cast<size_t>(2 * b + ch) * nv, og});
        }
    }
    enqueue_placements(prep);
    pack(net, first, n, vols.ptr);
    sync();  // the placement descriptors' host copies live as long as `prep`
}

void set_input(Net &net, int slot, const Image &fixed, const Image &moving) {
    const Image *f = &fixed, *m = &moving;
    set_inputs(net, slot, 1, &f, &m);
}

void set_input_raw(Net &net, int slot, const float *fixed_host, const float *moving_host) {
    const size_t nv = static_cast<size_t>(kEdge) * kEdge * kEdge;
    DeviceBuffer<float> v(2 * nv);
    HIP_OK(hipMemcpyAsync(v.ptr, fixed_host, nv * sizeof(float), hipMemcpyHostToDevice, backend().stream));
    HIP_OK(hipMemcpyAsync(v.ptr + nv, moving_host, nv * sizeof(float), hipMemcpyHostToDevice,
                          backend().stream));
    pack(net, slot, 1, v.ptr);
    sync();  // the caller's host arrays may go once this returns
}

}  // namespace cnn
}  // namespace sitkhip

// =================================================================== C ABI
using namespace sitkhip;

extern "C" {

int sitkhip_cnn_create(int max_batch, uint64_t seed, sitkhip_cnn *out) {
    return run_guarded([&] {
        if (!out) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null out-pointer");
        *out = reinterpret_cast<sitkhip_cnn>(cnn::create(max_batch, seed));
    });
}

void sitkhip_cnn_free(sitkhip_cnn h) { delete reinterpret_cast<cnn::Net *>(h); }

static cnn::Net &net_ref(sitkhip_cnn h) {
    auto n = reinterpret_cast<cnn::Net *>(h);
    if (!n) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null network");
    return *n;
}

int sitkhip_cnn_param_count(sitkhip_cnn h, size_t *n) {
    return run_guarded([&] {
        if (!n) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null out-pointer");
        *n = net_ref(h).nparams;
    });
}

int sitkhip_cnn_get_params(sitkhip_cnn h, float *out, size_t n) {
    return run_guarded([&] {
        cnn::Net &net = net_ref(h);
        if (!out || n != net.nparams) fail(SITKHIP_ERR_INVALID_ARGUMENT, "length mismatch");
        net.params.download(out, n);
        sync();
    });
}

int sitkhip_cnn_set_params(sitkhip_cnn h, const float *in, size_t n) {
    return run_guarded([&] {
        cnn::Net &net = net_ref(h);
        if (!in || n != net.nparams) fail(SITKHIP_ERR_INVALID_ARGUMENT, "length mismatch");
        net.params.upload(in, n);
        sync();
        net.weights_dirty = true;
    });
}

int sitkhip_cnn_get_grads(sitkhip_cnn h, float *out, size_t n) {
    return run_guarded([&] {
        cnn::Net &net = net_ref(h);
        if (!out || n != net.nparams) fail(SITKHIP_ERR_INVALID_ARGUMENT, "length mismatch");
        net.grads.download(out, n);
        sync();
    });
}

int sitkhip_cnn_set_input(sitkhip_cnn h, int slot, sitkhip_image fixed, sitkhip_image moving) {
    return run_guarded([&] {
        auto f = reinterpret_cast<Image *>(fixed);
        auto m = reinterpret_cast<Image *>(moving);
        if (!f || !m) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null image");
        cnn::set_input(net_ref(h), slot, *f, *m);
    });
}

int sitkhip_cnn_set_inputs(sitkhip_cnn h, int first_slot, int n, const sitkhip_image *fixed,
                           const sitkhip_image *moving) {
    return run_guarded([&] {
        if (!fixed || !moving) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null image array");
        cnn::set_inputs(net_ref(h), first_slot, n, reinterpret_cast<const Image *const *>(fixed),
                        reinterpret_cast<const Image *const *>(moving));
    });
}

int sitkhip_cnn_set_input_raw(sitkhip_cnn h, int slot, const float *fixed64, const float *moving64) {
    return run_guarded([&] {
        if (!fixed64 || !moving64) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null volume");
        cnn::set_input_raw(net_ref(h), slot, fixed64, moving64);
    });
}

int sitkhip_cnn_forward(sitkhip_cnn h, int batch, float *out) {
    return run_guarded([&] { cnn::forward(net_ref(h), batch, out); });
}

int sitkhip_cnn_backward(sitkhip_cnn h, int batch, const float *dout) {
    return run_guarded([&] {
        if (!dout) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null gradient");
        cnn::backward(net_ref(h), batch, dout);
    });
}

int sitkhip_cnn_adam(sitkhip_cnn h, float lr, float beta1, float beta2, float eps) {
    return run_guarded([&] { cnn::adam(net_ref(h), lr, beta1, beta2, eps); });
}

int sitkhip_cnn_train_step(sitkhip_cnn h, int batch, const float *targets, float lr, float *loss) {
    return run_guarded([&] {
        cnn::Net &net = net_ref(h);
        if (!targets) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null targets");
        std::vector<float> out(static_cast<size_t>(batch) * cnn::kOut), d(out.size());
        cnn::forward(net, batch, out.data());
        double l = 0.0;
        const float scale = 2.0f / static_cast<float>(out.size());
        for (size_t i = 0; i < out.size(); ++i) {
            const float e = out[i] - targets[i];
            l += static_cast<double>(e) * e;
            d[i] = scale * e;
        }
        if (loss) *loss = static_cast<float>(l / static_cast<double>(out.size()));
        cnn::backward(net, batch, d.data());
        cnn::adam(net, lr, 0.9f, 0.999f, 1e-8f);
    });
}

// Activation `index` (0 = packed input, l + 1 = conv l output) of the last
// forward pass, or with `gradient` != 0 the gradient w.r.t. it from the last
// backward pass (conv outputs: after the ReLU mask), as floats.
int sitkhip_cnn_activation(sitkhip_cnn h, int index, int gradient, int batch, float *out, size_t n) {
    return run_guarded([&] {
        cnn::Net &net = net_ref(h);
        if (index < 0 || index > cnn::kLayers) fail(SITKHIP_ERR_INVALID_ARGUMENT, "bad index");
        cnn::check_batch(net, batch);
        const size_t len = cnn::act_elems(index, batch);
        if (!out || n != len) fail(SITKHIP_ERR_INVALID_ARGUMENT, "length mismatch");
        std::vector<uint16_t> hb(len);
        (gradient ? net.dact[index] : net.act[index]).download(hb.data(), len);
        sync();
        for (size_t i = 0; i < len; ++i) out[i] = cnn::bf2f(hb[i]);
    });
}

}  // extern "C"
