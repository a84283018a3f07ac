// Mattes mutual information through a cubic B-spline transform: the value and
// the sparse-support derivative (crates/sitk/src/registration/mattes.rs:722-796,
// the path the reference's device backend refuses, registration/device.rs:79-100).
//
// Point map (src/transform/bspline.rs:485-583): the fixed point's continuous
// index in the control grid, the valid-region test with the far-boundary snap,
// the per-axis support start floor(c - 1) and cubic B-spline weights, and the
// displacement sum over the 64 support points (axis 0 fastest) -- all FP64 in
// the reference's order. A point outside the valid region maps to itself and
// has a zero Jacobian.
//
// Derivative, deterministic and without atomics: the samples are grouped once
// (at creation) by their support cell -- the start triple of their 4^3
// support, which depends only on the fixed point. Per evaluation one workgroup
// per cell contracts its samples into the 3 x 64 partial derivatives of the
// cell's support (sum over samples of (pRatio-weighted B3' sum) * dM/dx_d *
// weight, as FP64 matrix products in a fixed order); then every control point
// gathers the partials of the (up to 64) cells whose support holds it, in
// cell order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <exception>
#include <functional>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "host_pool.hpp"
#include "lbfgsb.hpp"

#include "common.hpp"
#include "image.hpp"
#include "metric.hpp"
#include "metric_core.hpp"

namespace sitkhip {

struct BSplineMap {
    int gs[3];          // control grid size (mesh + 3)
    int ncp;            // control points
    double origin[3];   // grid origin
    double p2i[9];      // physical -> grid continuous index
    const double *coef; // [3][ncp]
    // A moving-initial transform applied after the B-spline, as one
    // matrix-offset stage (registration/method.rs:198-240: moving point =
    // moving_initial(optimized(x))); post == 0: none.
    int post;
    double stage[12];   // matrix row-major [9], offset [3]
};

template <>
struct MapTraits<BSplineMap> {
    static constexpr bool affine = false;
    static constexpr bool points = false;
};

__host__ __device__ inline double bspline_b3(double u) {
    const double a = fabs(u);
    if (a < 1.0) {
        const double sq = a * a;
        return (4.0 - 6.0 * sq + 3.0 * sq * a) / 6.0;
    } else if (a < 2.0) {
        const double sq = a * a;
        return (8.0 - 12.0 * a + 6.0 * sq - sq * a) / 6.0;
    }
    return 0.0;
}

// Support start and the 1-D weights of fixed point x; false outside the
// valid region.
__host__ __device__ inline bool bspline_support(const BSplineMap &B, const double x[3], int start[3],
                                                double w[3][4]) {
    double c[3];
    for (int r = 0; r < 3; ++r) {
        double acc = -0.0;
        acc += B.p2i[r * 3 + 0] * (x[0] - B.origin[0]);
        acc += B.p2i[r * 3 + 1] * (x[1] - B.origin[1]);
        acc += B.p2i[r * 3 + 2] * (x[2] - B.origin[2]);
        c[r] = acc;
    }
    const double min_limit = 1.0;
    for (int d = 0; d < 3; ++d) {
        const double max_limit = static_cast<double>(B.gs[d]) - 1.0 - 1.0;
        const double m = fmax(fabs(max_limit), 1.0);
        const double ulp = 4.0 * DBL_EPSILON * m;
        if (fabs(c[d] - max_limit) <= ulp) {
            c[d] = max_limit - 6.0 * DBL_EPSILON * m;
        } else if (c[d] >= max_limit || c[d] < min_limit) {
            return false;
        }
    }
    for (int d = 0; d < 3; ++d) {
        const double s = floor(c[d] + 0.5 - 1.5);
        start[d] = static_cast<int>(s);
        double u = c[d] - s;
        for (int k = 0; k < 4; ++k) {
            w[d][k] = bspline_b3(u);
            u -= 1.0;
        }
    }
    return true;
}

// p = stage(q), one matrix-offset stage as the stage replay rounds it.
__device__ __forceinline__ void apply_post_stage(const BSplineMap &B, double p[3]) {
    if (!B.post) return;
    const double *m = B.stage;
    const double q0 = p[0], q1 = p[1], q2 = p[2];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    a0 += m[0] * q0; a0 += m[1] * q1; a0 += m[2] * q2;
    a1 += m[3] * q0; a1 += m[4] * q1; a1 += m[5] * q2;
    a2 += m[6] * q0; a2 += m[7] * q1; a2 += m[8] * q2;
    p[0] = a0 + m[9];
    p[1] = a1 + m[10];
    p[2] = a2 + m[11];
}

__device__ __forceinline__ void map_point(const BSplineMap &B, const double x[3], double p[3]) {
    int start[3];
    double w[3][4];
    if (!bspline_support(B, x, start, w)) {
        p[0] = x[0]; p[1] = x[1]; p[2] = x[2];
        apply_post_stage(B, p);
        return;
    }
    double d0 = 0.0, d1 = 0.0, d2 = 0.0;
    const int g0 = B.gs[0], g01 = B.gs[0] * B.gs[1];
    for (int k = 0; k < 4; ++k)
        for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 4; ++i) {
                const double wt = (w[0][i] * w[1][j]) * w[2][k];
                const int cp = (start[0] + i) + g0 * (start[1] + j) + g01 * (start[2] + k);
                d0 += wt * B.coef[cp];
                d1 += wt * B.coef[B.ncp + cp];
                d2 += wt * B.coef[2 * B.ncp + cp];
            }
    p[0] = x[0] + d0;
    p[1] = x[1] + d1;
    p[2] = x[2] + d2;
    apply_post_stage(B, p);
}

// ----------------------------------------------------------------- passes
// One launch serves any number of starts (independent optimisations of one
// pair that share the sample set and the images): blockIdx.x picks a start
// from the active list (so consecutive workgroups share a cell's samples); each start has its own coefficients, moving-initial
// stage, histogram copies, pRatio, tail, records, cell partials and
// derivative. A single evaluation is the one-start case.
//
// The start's map: coefficients coef + s * 3 ncp and stage s (when given; else
// the base map's own).
__device__ __forceinline__ BSplineMap start_map(const BSplineMap &B0, const double *__restrict__ coef,
                                                const double *__restrict__ stage, int s) {
    BSplineMap B = B0;
    if (coef) B.coef = coef + static_cast<size_t>(s) * 3 * B0.ncp;
    if (stage) {
        B.post = 1;
#pragma unroll
        for (int i = 0; i < 12; ++i) B.stage[i] = stage[s * 12 + i];
    }
    return B;
}

// Cells per value-pass workgroup: 4 when many starts fill the chip, else 1.
constexpr int kCellsPerBlockMany = 4;

// The 3 x 64 coefficients of a cell's support, s[d * 64 + k * 16 + j * 4 + i].
__device__ __forceinline__ void stage_cell_coefs(const BSplineMap &B, int cell, double *s) {
    const int m0 = B.gs[0] - 3, m1 = B.gs[1] - 3;
    const int sx = cell % m0, sy = (cell / m0) % m1, sz = cell / (m0 * m1);
    for (int t = threadIdx.x; t < 192; t += blockDim.x) {
        const int d = t >> 6, o = t & 63;
        const int i = o & 3, j = (o >> 2) & 3, k = o >> 4;
        const int cp = (sx + i) + B.gs[0] * ((sy + j) + B.gs[1] * (sz + k));
        s[t] = B.coef[static_cast<size_t>(d) * B.ncp + cp];
    }
}

// map_point for a sample of a cell: the stored support weights ws (w0, w1,
// w2; 12 doubles) and the cell's coefficients from LDS, the same sums in the
// same order as the general map; a sample of the outside bucket maps to
// itself. Then the stage, if any.
__device__ __forceinline__ void map_point_cell(const BSplineMap &B, const double x[3], bool in_cell,
                                               const double *__restrict__ ws, const double *sc,
                                               double p[3]) {
    if (!in_cell) {
        p[0] = x[0]; p[1] = x[1]; p[2] = x[2];
        apply_post_stage(B, p);
        return;
    }
    double w2[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) w2[k] = ws[8 + k];
    // (w0[i] w1[j]) once per sample, then one k-slab (48 LDS reads) unrolled
    // so the reads are issued ahead of the dependent sums
    double w01[16];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) w01[j * 4 + i] = ws[i] * ws[4 + j];
    double d0 = 0.0, d1 = 0.0, d2 = 0.0;
#pragma unroll 1
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const double wt = w01[j * 4 + i] * w2[k];
                const int o = k * 16 + j * 4 + i;
                d0 += wt * sc[o];
                d1 += wt * sc[64 + o];
                d2 += wt * sc[128 + o];
            }
    p[0] = x[0] + d0;
    p[1] = x[1] + d1;
    p[2] = x[2] + d2;
    apply_post_stage(B, p);
}

// The support weights of every sample of a cell (they depend only on the
// fixed point): w[12] per sample in the cell-grouped order, computed once.
// A sample whose support on the device is not its cell's (the grouping was
// made on the host) raises *mismatch.
__global__ void bs_weights(FixedGrid F, BSplineMap B, const ChainConst *__restrict__ C,
                           const uint32_t *__restrict__ idx, const int *__restrict__ cell_begin,
                           int ncells, double *__restrict__ wts, int *__restrict__ mismatch) {
    const int cell = blockIdx.y;
    const int m0 = B.gs[0] - 3, m1 = B.gs[1] - 3;
    const int sx = cell % m0, sy = (cell / m0) % m1, sz = cell / (m0 * m1);
    const int e = cell_begin[cell + 1];
    for (int s = cell_begin[cell] + blockIdx.x * blockDim.x + threadIdx.x; s < e;
         s += gridDim.x * blockDim.x) {
        double x[3], w[3][4];
        int start[3];
        fixed_point(F, *C, idx[s], x);
        const bool in = bspline_support(B, x, start, w);
        if (!in || start[0] != sx || start[1] != sy || start[2] != sz) *mismatch = 1;
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int k = 0; k < 4; ++k) wts[static_cast<size_t>(s) * 12 + d * 4 + k] = w[d][k];
    }
}

// Value pass, `cpb` cells per workgroup (the last bucket, cell == ncells,
// holds the samples outside the valid region): the cell's support
// coefficients in LDS, the joint histogram in LDS, one record per sample (the
// moving value and its index gradient, NaN value = dropped) for the
// derivative pass. Any grouping gives the same integer histogram.
template <typename R>
__global__ __launch_bounds__(kBlock) void bs_value(FixedGrid F, MovingVol M, BSplineMap B0,
                                                   const double *__restrict__ coef,
                                                   const double *__restrict__ stage,
                                                   const int *__restrict__ active,
                                                   const ChainConst *__restrict__ C,
                                                   const uint32_t *__restrict__ idx,
                                                   const uint8_t *__restrict__ fbin,
                                                   const int *__restrict__ cell_begin,
                                                   const double *__restrict__ wts, int ncells,
                                                   int cpb, MattesParams P,
                                                   unsigned long long *__restrict__ g_joint,
                                                   unsigned long long *__restrict__ g_marg,
                                                   SampleRec<R> *__restrict__ rec, long long n) {
    extern __shared__ unsigned long long s_hist[];
    __shared__ double s_coef[192];
    const int st = active[blockIdx.x];
    const BSplineMap B = start_map(B0, coef, stage, st);
    const int nb = P.bins * P.bins + P.bins;
    unsigned long long *h_joint = s_hist, *h_marg = s_hist + P.bins * P.bins;
    SampleRec<R> *r = rec + static_cast<size_t>(st) * n;
    for (int i = threadIdx.x; i < nb; i += kBlock) s_hist[i] = 0ull;
    for (int q = 0; q < cpb; ++q) {
        const int cell = blockIdx.y * cpb + q;
        if (cell > ncells) break;  // the same for every lane
        __syncthreads();           // the previous cell's coefficients are no longer read
        if (cell < ncells) stage_cell_coefs(B, cell, s_coef);
        __syncthreads();
        const int e = cell_begin[cell + 1];
        for (int s = cell_begin[cell] + threadIdx.x; s < e; s += kBlock) {
            const int fb = fbin[s];
            Cell cl;
            bool ok = false;
            if (fb != kNoBin) {
                double x[3], p[3];
                fixed_point(F, *C, idx[s], x);
                map_point_cell(B, x, cell < ncells, wts + static_cast<size_t>(s) * 12, s_coef, p);
                ok = locate(M, *C, p, cl);
            }
            if (!ok) {
                cl.base[0] = cl.base[1] = cl.base[2] = 0;
                cl.frac[0] = cl.frac[1] = cl.frac[2] = 0.0;
            }
            R fr[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) fr[d] = static_cast<R>(cl.frac[d]);
            float v[8];
            corners(M, cl, v);
            hist_sample<R, true>(P, h_joint, h_marg, r, s, ok, fb, fr, v);
        }
    }
    __syncthreads();
    const unsigned copy = (blockIdx.y * gridDim.x + blockIdx.x) % kHistCopies;
    g_joint += (static_cast<size_t>(st) * kHistCopies + copy) * P.bins * P.bins;
    g_marg += (static_cast<size_t>(st) * kHistCopies + copy) * P.bins;
    for (int i = threadIdx.x; i < nb; i += kBlock) {
        const unsigned long long v = s_hist[i];
        if (v) atomicAdd(i < P.bins * P.bins ? g_joint + i : g_marg + (i - P.bins * P.bins), v);
    }
}

// One workgroup per active start: the Mattes tail on its histogram copies
// (cleared for the next evaluation), pRatio in double and float.
__global__ __launch_bounds__(kTail) void bs_tail(const int *__restrict__ active, int B, double mbs,
                                                 unsigned long long *__restrict__ g_joint,
                                                 unsigned long long *__restrict__ g_marg,
                                                 double *__restrict__ pratio,
                                                 float *__restrict__ pratio_f,
                                                 TailOut *__restrict__ tail) {
    extern __shared__ unsigned long long s_fold[];
    const int st = active[blockIdx.x];
    const size_t nb = static_cast<size_t>(B) * B;
    mattes_tail_block(g_joint + static_cast<size_t>(st) * kHistCopies * nb,
                      g_marg + static_cast<size_t>(st) * kHistCopies * B, B, mbs,
                      pratio + st * nb, pratio_f + st * nb, tail + st, s_fold);
}

// Derivative pass, one workgroup per (cell, start), from the value pass's
// records: per valid sample the pRatio-weighted B3' sum a times the moving
// gradient (through the stage's transposed matrix when there is one), g_d;
// then out[cell][16 n + m] = sum over the cell's samples of u_m t_n with
// u_m = w0[m % 4] w1[m / 4] and t_n = g_{n / 4} w2[n % 4] (n < 12), i.e.
// out[d * 64 + k * 16 + j * 4 + i] = sum of g_d w0[i] w1[j] w2[k]. The sum
// is a 16 x 4 by 4 x 16 FP64 matrix product per four samples
// (v_mfma_f64_16x16x4f64): each lane stages its sample's weights and g in
// LDS, then lane l feeds A[l % 16][l / 16] = u and B[l / 16][l % 16] = t of
// sample 4 r + l / 16. The four waves' tiles are added in a fixed order.
typedef double f64x4 __attribute__((ext_vector_type(4)));
static_assert(kBlock == 256, "the derivative pass adds four waves' tiles");

template <typename R>
__global__ __launch_bounds__(kBlock) void bs_deriv(FixedGrid F, MovingVol M, BSplineMap B0,
                                                   const double *__restrict__ stage,
                                                   const int *__restrict__ active,
                                                   const uint32_t *__restrict__ idx,
                                                   const int *__restrict__ cell_begin,
                                                   const double *__restrict__ wts,
                                                   const uint8_t *__restrict__ fbin,
                                                   MattesParams P, const R *__restrict__ pratio,
                                                   const SampleRec<R> *__restrict__ rec, long long n,
                                                   int ncells, double *__restrict__ out) {
    __shared__ double s_x[kBlock / 64][64][16];  // per lane: w[3][4], g[3], pad
    const int st = active[blockIdx.x];
    const BSplineMap B = start_map(B0, nullptr, stage, st);
    const SampleRec<R> *r = rec + static_cast<size_t>(st) * n;
    pratio += static_cast<size_t>(st) * P.bins * P.bins;
    const int cell = blockIdx.y;
    const int b = cell_begin[cell], e = cell_begin[cell + 1];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int c = lane & 15, g4 = lane >> 4;
    f64x4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int s0 = b; s0 < e; s0 += kBlock) {
        const int s = s0 + threadIdx.x;
        double g[3] = {0.0, 0.0, 0.0};
        double w[3][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
        if (s < e) {
            const SampleRec<R> rr = r[s];
            if (rr.mv == rr.mv) {  // a sample the value pass kept
#pragma unroll
                for (int d = 0; d < 3; ++d)
#pragma unroll
                    for (int k = 0; k < 4; ++k) w[d][k] = wts[static_cast<size_t>(s) * 12 + d * 4 + k];
                R gp[3];
                physical_gradient<R>(M, rr.gi, gp);
                R arg[4], dv[4];
                bool mid;
                const int m0 = Parzen<R>::taps(P, rr.mv, arg, mid);
                parzen_derivs<R>(arg, mid, dv);
                const R *row = pratio + fbin[s] * P.bins + m0;
                R a = R(0);
#pragma unroll
                for (int k = 0; k < 4; ++k) a += row[k] * dv[k];
                double gq[3] = {static_cast<double>(gp[0]), static_cast<double>(gp[1]),
                                static_cast<double>(gp[2])};
                if (B.post) {
                    // the sparse Jacobian's columns left-multiplied by the
                    // stage's matrix: the gradient enters as M' grad
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        double acc2 = 0.0;
#pragma unroll
                        for (int q = 0; q < 3; ++q)
                            acc2 += B.stage[q * 3 + d] * static_cast<double>(gp[q]);
                        gq[d] = acc2;
                    }
                }
#pragma unroll
                for (int d = 0; d < 3; ++d) g[d] = static_cast<double>(a) * gq[d];
            }
        }
        double *x = s_x[wv][lane];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            x[k] = w[0][k];
            x[4 + k] = w[1][k];
            x[8 + k] = w[2][k];
        }
        x[12] = g[0];
        x[13] = g[1];
        x[14] = g[2];
        __syncthreads();
#pragma unroll 4
        for (int q = 0; q < 16; ++q) {
            const double *y = s_x[wv][4 * q + g4];
            const double ua = y[c & 3] * y[4 + (c >> 2)];
            const double tb = c < 12 ? y[12 + (c >> 2)] * y[8 + (c & 3)] : 0.0;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ua, tb, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // accumulator i of lane l is tile element (row l / 16 + 4 i, column l % 16);
    // wave w's tile goes to red[w][row][column]
    double *red = &s_x[0][0][0];
#pragma unroll
    for (int i = 0; i < 4; ++i) red[wv * 256 + (g4 + 4 * i) * 16 + c] = acc[i];
    __syncthreads();
    if (threadIdx.x < 192) {
        const int t = threadIdx.x;
        const int o = (t & 15) * 16 + (t >> 4);
        out[(static_cast<size_t>(st) * ncells + cell) * 192 + t] =
            (red[o] + red[256 + o]) + (red[512 + o] + red[768 + o]);
    }
}

// deriv[d][cp] = sum over the cells whose support holds cp (cell order).
__global__ void bs_gather(BSplineMap B, const int *__restrict__ active, int ncells,
                          const double *__restrict__ cellT, double *__restrict__ deriv) {
    const int st = active[blockIdx.y];
    cellT += static_cast<size_t>(st) * ncells * 192;
    deriv += static_cast<size_t>(st) * 3 * B.ncp;
    const int cp = blockIdx.x * blockDim.x + threadIdx.x;
    if (cp >= B.ncp) return;
    const int m0 = B.gs[0] - 3, m1 = B.gs[1] - 3, m2 = B.gs[2] - 3;
    const int px = cp % B.gs[0], py = (cp / B.gs[0]) % B.gs[1], pz = cp / (B.gs[0] * B.gs[1]);
    double sum[3] = {0.0, 0.0, 0.0};
    for (int sz = max(0, pz - 3); sz <= min(m2 - 1, pz); ++sz)
        for (int sy = max(0, py - 3); sy <= min(m1 - 1, py); ++sy)
            for (int sx = max(0, px - 3); sx <= min(m0 - 1, px); ++sx) {
                const int cell = sx + m0 * (sy + m1 * sz);
                const int o = (px - sx) + 4 * (py - sy) + 16 * (pz - sz);
#pragma unroll
                for (int d = 0; d < 3; ++d) sum[d] += cellT[static_cast<size_t>(cell) * 192 + d * 64 + o];
            }
#pragma unroll
    for (int d = 0; d < 3; ++d) deriv[static_cast<size_t>(d) * B.ncp + cp] = sum[d];
}

// ------------------------------------------------------------------ host side
struct BSplineMulti;
void bspline_multi_free(BSplineMulti *u);

struct BSplineMetric {
    Metric *m = nullptr;        // the Mattes metric (bins set); not owned
    BSplineMap B{};             // coef unused: every evaluation brings its coefficients
    int ncells = 0;
    SampleSet SS{};             // cell-grouped samples
    DeviceBuffer<uint32_t> idx;
    DeviceBuffer<uint8_t> fbin;
    DeviceBuffer<int> cell_begin;  // ncells + 2 bounds: the cells, then the outside bucket
    DeviceBuffer<double> wts;      // support weights w[3][4] per sample of a cell
    BSplineMulti *one = nullptr;   // the single-start evaluation (its stage: B's own)
    ~BSplineMetric() { bspline_multi_free(one); }
};

BSplineMetric *bspline_create(Metric &m, const size_t gs[3], const double origin[3],
                              const double spacing[3], const double direction[9]) {
    if (m.kind != SITKHIP_METRIC_MATTES || !m.bins_set)
        fail(SITKHIP_ERR_HISTOGRAM_SHAPE, "a B-spline metric needs a Mattes metric with bins set");
    if (metric_uses_bspline_interpolation(m))
        fail(SITKHIP_ERR_INVALID_ARGUMENT,
             "the B-spline transform metric interpolates the moving image linearly");
    for (int d = 0; d < 3; ++d) {
        if (gs[d] < 4 || gs[d] > 4096)
            fail(SITKHIP_ERR_INVALID_ARGUMENT, "control grid size must be 4..4096 per axis");
        if (!(spacing[d] > 0.0)) fail(SITKHIP_ERR_INVALID_ARGUMENT, "grid spacing must be positive");
    }
    auto *b = new BSplineMetric();
    try {
        b->m = &m;
        for (int d = 0; d < 3; ++d) {
            b->B.gs[d] = static_cast<int>(gs[d]);
            b->B.origin[d] = origin[d];
        }
        const size_t ncp = gs[0] * gs[1] * gs[2];
        if (ncp > (1u << 26)) fail(SITKHIP_ERR_INVALID_ARGUMENT, "too many control points");
        b->B.ncp = static_cast<int>(ncp);
        double i2p[9];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) i2p[r * 3 + c] = direction[r * 3 + c] * spacing[c];
        invert3_host(i2p, b->B.p2i);
        const int m0 = b->B.gs[0] - 3, m1 = b->B.gs[1] - 3, m2 = b->B.gs[2] - 3;
        b->ncells = m0 * m1 * m2;

        // Group the samples by support cell (outside-valid-region samples last).
        const size_t n = m.nsamples;
        std::vector<uint32_t> flat(n);
        std::vector<uint8_t> fb(n);
        if (m.SS.idx) m.idx.download(flat.data(), n);
        else for (size_t s = 0; s < n; ++s) flat[s] = static_cast<uint32_t>(s);
        m.fbin.download(fb.data(), n);
        sync();
        std::vector<int> cell(n);
        std::vector<int> count(b->ncells + 1, 0);
        for (size_t s = 0; s < n; ++s) {
            const uint32_t f = flat[s];
            const uint32_t i0 = f % m.F.size[0];
            const uint32_t q = f / m.F.size[0];
            const uint32_t i1 = q % m.F.size[1], i2 = q / m.F.size[1];
            double x[3];
            for (int r = 0; r < 3; ++r) {
                double acc = m.F.origin[r];
                acc += m.F.i2p[r * 3 + 0] * static_cast<double>(i0);
                acc += m.F.i2p[r * 3 + 1] * static_cast<double>(i1);
                acc += m.F.i2p[r * 3 + 2] * static_cast<double>(i2);
                x[r] = acc;
            }
            int st[3];
            double w[3][4];
            int c = b->ncells;
            if (bspline_support(b->B, x, st, w)) c = st[0] + m0 * (st[1] + m1 * st[2]);
            cell[s] = c;
            ++count[c];
        }
        std::vector<int> begin(b->ncells + 2, 0);
        for (int c = 0; c <= b->ncells; ++c) begin[c + 1] = begin[c] + count[c];
        std::vector<uint32_t> sflat(n);
        std::vector<uint8_t> sfb(n);
        std::vector<int> fill(begin.begin(), begin.end() - 1);
        for (size_t s = 0; s < n; ++s) {
            const int pos = fill[cell[s]]++;
            sflat[pos] = flat[s];
            sfb[pos] = fb[s];
        }
        b->idx.alloc(n);
        b->fbin.alloc(n);
        if (n) {
            b->idx.upload(sflat.data(), n);
            b->fbin.upload(sfb.data(), n);
        }
        b->cell_begin.alloc(b->ncells + 2);
        b->cell_begin.upload(begin.data(), b->ncells + 2);
        b->SS.idx = b->idx.ptr;
        b->SS.n = static_cast<long long>(n);
        b->B.coef = nullptr;
        b->wts.alloc(n * 12);
        if (b->ncells > 0 && n > 0) {
            DeviceBuffer<int> bad(1);
            HIP_OK(hipMemsetAsync(bad.ptr, 0, sizeof(int), backend().stream));
            hipLaunchKernelGGL(bs_weights, dim3(8, b->ncells), dim3(256), 0, backend().stream, m.F,
                               b->B, m.chain.ptr, b->idx.ptr, b->cell_begin.ptr, b->ncells,
                               b->wts.ptr, bad.ptr);
            check_launch("bs_weights");
            int flag = 0;
            HIP_OK(hipMemcpyAsync(&flag, bad.ptr, sizeof(int), hipMemcpyDeviceToHost, backend().stream));
            sync();
            if (flag)
                fail(SITKHIP_ERR_DRIVER, "a sample's B-spline support on the device differs from its cell");
        }
        sync();
    } catch (...) {
        delete b;
        throw;
    }
    return b;
}

void bspline_evaluate(BSplineMetric &b, const double *coeffs, bool derivative, double *value,
                      double *deriv, size_t *valid);

}  // namespace sitkhip

using namespace sitkhip;

extern "C" {

int sitkhip_bspline_create(sitkhip_metric metric, const size_t grid_size[3],
                           const double grid_origin[3], const double grid_spacing[3],
                           const double grid_direction[9], sitkhip_bspline *out) {
    return run_guarded([&] {
        auto m = reinterpret_cast<Metric *>(metric);
        if (!m || !grid_size || !grid_origin || !grid_spacing || !grid_direction || !out)
            fail(SITKHIP_ERR_INVALID_ARGUMENT, "null argument");
        if (metric_has_explicit_points(*m))
            fail(SITKHIP_ERR_INVALID_ARGUMENT, "the B-spline metric groups samples by grid voxel; "
                                               "explicit points have none");
        *out = reinterpret_cast<sitkhip_bspline>(
            bspline_create(*m, grid_size, grid_origin, grid_spacing, grid_direction));
    });
}

void sitkhip_bspline_free(sitkhip_bspline b) { delete reinterpret_cast<BSplineMetric *>(b); }

int sitkhip_bspline_parameter_count(sitkhip_bspline h, size_t *n) {
    return run_guarded([&] {
        auto b = reinterpret_cast<BSplineMetric *>(h);
        if (!b || !n) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null argument");
        *n = 3 * static_cast<size_t>(b->B.ncp);
    });
}

int sitkhip_bspline_mattes_value(sitkhip_bspline h, const double *coeffs, double *value,
                                 size_t *valid) {
    return run_guarded([&] {
        auto b = reinterpret_cast<BSplineMetric *>(h);
        if (!b) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null handle");
        bspline_evaluate(*b, coeffs, false, value, nullptr, valid);
    });
}

int sitkhip_bspline_mattes_value_and_derivative(sitkhip_bspline h, const double *coeffs,
                                                double *value, double *deriv, size_t *valid) {
    return run_guarded([&] {
        auto b = reinterpret_cast<BSplineMetric *>(h);
        if (!b) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null handle");
        bspline_evaluate(*b, coeffs, true, value, deriv, valid);
    });
}

int sitkhip_bspline_set_moving_initial(sitkhip_bspline h, const double stage[12]) {
    return run_guarded([&] {
        auto b = reinterpret_cast<BSplineMetric *>(h);
        if (!b) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null handle");
        b->B.post = stage ? 1 : 0;
        for (int i = 0; i < 12; ++i) b->B.stage[i] = stage ? stage[i] : 0.0;
    });
}

}  // extern "C"

// ============================================================ many starts
namespace sitkhip {

struct BSplineMulti {
    BSplineMetric *b = nullptr;  // sample grouping and images; not owned
    int S = 0;
    size_t np = 0;               // parameters per start (3 ncp)
    bool staged = false;         // per-start stages (else the metric's own)
    DeviceBuffer<double> coef, stage, pratio, cellT, deriv;
    DeviceBuffer<float> pratio_f;
    DeviceBuffer<unsigned long long> joint, marg;
    DeviceBuffer<TailOut> tail;
    DeviceBuffer<int> active;
    DeviceBuffer<unsigned char> rec;  // S x n sample records
    double *h_coef = nullptr, *h_deriv = nullptr;  // pinned, S x np
    TailOut *h_tail = nullptr;
    int *h_active = nullptr;
    ~BSplineMulti() {
        if (h_coef) (void)hipHostFree(h_coef);
        if (h_deriv) (void)hipHostFree(h_deriv);
        if (h_tail) (void)hipHostFree(h_tail);
        if (h_active) (void)hipHostFree(h_active);
    }
};

void bspline_multi_free(BSplineMulti *u) { delete u; }

// stages: S x 12, or null for the metric's own moving-initial stage.
BSplineMulti *bspline_multi_create(BSplineMetric &b, int S, const double *stages) {
    if (S < 1 || S > 4096) fail(SITKHIP_ERR_INVALID_ARGUMENT, "start count must be 1..4096");
    const int B = b.m->P.bins;
    if (B > kLdsMaxBins) fail(SITKHIP_ERR_TOO_MANY_BINS, "the B-spline value pass holds at most 90 bins in LDS");
    auto *u = new BSplineMulti();
    try {
        u->b = &b;
        u->S = S;
        u->np = 3 * static_cast<size_t>(b.B.ncp);
        u->staged = stages != nullptr;
        const size_t nb = static_cast<size_t>(B) * B;
        const size_t recb = b.m->precision == SITKHIP_PRECISION_F64 ? sizeof(SampleRec<double>)
                                                                    : sizeof(SampleRec<float>);
        u->coef.alloc(S * u->np);
        u->stage.alloc(static_cast<size_t>(S) * 12);
        u->pratio.alloc(S * nb);
        u->pratio_f.alloc(S * nb);
        u->cellT.alloc(static_cast<size_t>(S) * std::max(1, b.ncells) * 192);
        u->deriv.alloc(S * u->np);
        u->joint.alloc(static_cast<size_t>(S) * kHistCopies * nb);
        u->marg.alloc(static_cast<size_t>(S) * kHistCopies * B);
        u->tail.alloc(S);
        u->active.alloc(S);
        u->rec.alloc(static_cast<size_t>(S) * static_cast<size_t>(b.SS.n) * recb);
        HIP_OK(hipMemsetAsync(u->joint.ptr, 0, u->joint.bytes(), backend().stream));
        HIP_OK(hipMemsetAsync(u->marg.ptr, 0, u->marg.bytes(), backend().stream));
        HIP_OK(hipHostMalloc(reinterpret_cast<void **>(&u->h_coef), S * u->np * sizeof(double), hipHostMallocDefault));
        HIP_OK(hipHostMalloc(reinterpret_cast<void **>(&u->h_deriv), S * u->np * sizeof(double), hipHostMallocDefault));
        HIP_OK(hipHostMalloc(reinterpret_cast<void **>(&u->h_tail), S * sizeof(TailOut), hipHostMallocDefault));
        HIP_OK(hipHostMalloc(reinterpret_cast<void **>(&u->h_active), S * sizeof(int), hipHostMallocDefault));
        if (stages) u->stage.upload(stages, static_cast<size_t>(S) * 12);
        std::memset(u->h_coef, 0, S * u->np * sizeof(double));
        sync();
    } catch (...) {
        delete u;
        throw;
    }
    return u;
}

template <typename R>
static void multi_passes(BSplineMulti &u, const int *active, int nact, bool derivative) {
    BSplineMetric &b = *u.b;
    Metric &m = *b.m;
    const int B = m.P.bins;
    const hipStream_t st = backend().stream;
    const size_t shm = static_cast<size_t>(B * B + B) * sizeof(unsigned long long);
    const double *stage = u.staged ? u.stage.ptr : nullptr;
    auto *rec = reinterpret_cast<SampleRec<R> *>(u.rec.ptr);
    const int cpb = nact >= 16 ? kCellsPerBlockMany : 1;
    const unsigned groups = static_cast<unsigned>((b.ncells + 1 + cpb - 1) / cpb);
    {
        LaunchScope scope("bs_value");
        hipLaunchKernelGGL(bs_value<R>, dim3(nact, groups), dim3(kBlock), shm, st, m.F, m.Mv, b.B,
                           u.coef.ptr, stage, active, m.chain.ptr, b.idx.ptr, b.fbin.ptr,
                           b.cell_begin.ptr, b.wts.ptr, b.ncells, cpb, m.P, u.joint.ptr, u.marg.ptr,
                           rec, b.SS.n);
        check_launch("bs_value");
    }
    {
        LaunchScope scope("bs_tail");
        hipLaunchKernelGGL(bs_tail, dim3(nact), dim3(kTail), shm, st, active, B, m.P.mbs,
                           u.joint.ptr, u.marg.ptr, u.pratio.ptr, u.pratio_f.ptr, u.tail.ptr);
        check_launch("bs_tail");
    }
    if (!derivative) return;
    const R *pr = sizeof(R) == 4 ? reinterpret_cast<const R *>(u.pratio_f.ptr)
                                 : reinterpret_cast<const R *>(u.pratio.ptr);
    if (b.ncells > 0) {
        LaunchScope scope("bs_deriv");
        hipLaunchKernelGGL(bs_deriv<R>, dim3(nact, b.ncells), dim3(kBlock), 0, st, m.F, m.Mv, b.B,
                           stage, active, b.idx.ptr, b.cell_begin.ptr, b.wts.ptr, b.fbin.ptr, m.P,
                           pr, rec,
                           b.SS.n, b.ncells, u.cellT.ptr);
        check_launch("bs_deriv");
    }
    LaunchScope scope("bs_gather");
    hipLaunchKernelGGL(bs_gather, dim3((b.B.ncp + 255) / 256, nact), dim3(256), 0, st, b.B,
                       active, b.ncells, u.cellT.ptr, u.deriv.ptr);
    check_launch("bs_gather");
}

// Evaluates the listed starts at the coefficients already in u.h_coef (start
// s at h_coef + s * np). Results land in u.h_tail and, with the derivative,
// u.h_deriv (zero for a start without valid samples).
void bspline_multi_evaluate(BSplineMulti &u, const int *act, int nact, bool derivative = true) {
    if (nact <= 0) return;
    if (nact > u.S) fail(SITKHIP_ERR_INVALID_ARGUMENT, "more active starts than starts");
    std::vector<char> seen(u.S, 0);
    for (int i = 0; i < nact; ++i) {
        if (act[i] < 0 || act[i] >= u.S) fail(SITKHIP_ERR_INVALID_ARGUMENT, "active start out of range");
        // a start listed twice would add its samples to its histogram twice
        if (seen[act[i]]) fail(SITKHIP_ERR_INVALID_ARGUMENT, "a start is listed twice");
        seen[act[i]] = 1;
        u.h_active[i] = act[i];
    }
    const hipStream_t st = backend().stream;
    const size_t bytes = u.np * sizeof(double);
    // one copy each way over the span of listed starts (the rows of starts in
    // between that are not listed travel too and are not used)
    const int lo = *std::min_element(act, act + nact), hi = *std::max_element(act, act + nact);
    const size_t span = static_cast<size_t>(hi - lo + 1);
    HIP_OK(hipMemcpyAsync(u.coef.ptr + lo * u.np, u.h_coef + lo * u.np, span * bytes,
                          hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(u.active.ptr, u.h_active, nact * sizeof(int), hipMemcpyHostToDevice, st));
    if (u.b->m->precision == SITKHIP_PRECISION_F64) multi_passes<double>(u, u.active.ptr, nact, derivative);
    else multi_passes<float>(u, u.active.ptr, nact, derivative);
    if (derivative)
        HIP_OK(hipMemcpyAsync(u.h_deriv + lo * u.np, u.deriv.ptr + lo * u.np, span * bytes,
                              hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(u.h_tail + lo, u.tail.ptr + lo, span * sizeof(TailOut),
                          hipMemcpyDeviceToHost, st));
    sync();
    if (derivative)
        for (int i = 0; i < nact; ++i)
            if (u.h_tail[act[i]].valid == 0.0) std::memset(u.h_deriv + act[i] * u.np, 0, bytes);
}

void bspline_evaluate(BSplineMetric &b, const double *coeffs, bool derivative, double *value,
                      double *deriv, size_t *valid) {
    if (!coeffs) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null coefficients");
    if (!b.one) b.one = bspline_multi_create(b, 1, nullptr);
    BSplineMulti &u = *b.one;
    std::memcpy(u.h_coef, coeffs, u.np * sizeof(double));
    const int act = 0;
    bspline_multi_evaluate(u, &act, 1, derivative);
    if (value) *value = u.h_tail[0].value;
    if (valid) *valid = static_cast<size_t>(u.h_tail[0].valid);
    if (derivative && deriv) std::memcpy(deriv, u.h_deriv, u.np * sizeof(double));
}

// A contiguous range of starts evaluated on its own (register_starts keeps
// two, so the host updates one range's optimizers while the device evaluates
// the other's). The ranges are disjoint, so each evaluation touches only its
// own rows of the coefficient, derivative and tail buffers.
struct StartLane {
    int lo = 0, hi = 0;        // starts [lo, hi)
    std::vector<int> act;      // the starts wanting an evaluation
    DeviceBuffer<int> active;
    int *h_active = nullptr;   // pinned
    hipEvent_t done = nullptr;
    bool pending = false;
    StartLane() = default;
    StartLane(const StartLane &) = delete;
    StartLane &operator=(const StartLane &) = delete;
    ~StartLane() {
        if (h_active) (void)hipHostFree(h_active);
        if (done) (void)hipEventDestroy(done);
    }
};

static void lane_launch(BSplineMulti &u, StartLane &L) {
    const int nact = static_cast<int>(L.act.size());
    const hipStream_t st = backend().stream;
    const size_t bytes = u.np * sizeof(double);
    const int lo = *std::min_element(L.act.begin(), L.act.end());
    const int hi = *std::max_element(L.act.begin(), L.act.end());
    const size_t span = static_cast<size_t>(hi - lo + 1);
    std::memcpy(L.h_active, L.act.data(), nact * sizeof(int));
    HIP_OK(hipMemcpyAsync(u.coef.ptr + lo * u.np, u.h_coef + lo * u.np, span * bytes,
                          hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(L.active.ptr, L.h_active, nact * sizeof(int), hipMemcpyHostToDevice, st));
    if (u.b->m->precision == SITKHIP_PRECISION_F64) multi_passes<double>(u, L.active.ptr, nact, true);
    else multi_passes<float>(u, L.active.ptr, nact, true);
    HIP_OK(hipMemcpyAsync(u.h_deriv + lo * u.np, u.deriv.ptr + lo * u.np, span * bytes,
                          hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(u.h_tail + lo, u.tail.ptr + lo, span * sizeof(TailOut),
                          hipMemcpyDeviceToHost, st));
    HIP_OK(hipEventRecord(L.done, st));
    L.pending = true;
}

static void lane_wait(BSplineMulti &u, StartLane &L) {
    HIP_OK(hipEventSynchronize(L.done));
    L.pending = false;
    for (int s : L.act)
        if (u.h_tail[s].valid == 0.0) std::memset(u.h_deriv + s * u.np, 0, u.np * sizeof(double));
}

// Every start optimised by its own L-BFGS-B (lbfgsb.rs:1590-1746); each round
// evaluates all starts whose optimizer asks for a point in one batched set of
// launches. x: S x np, in (starts) and out (best points).
static void register_starts(BSplineMulti &u, const sitkhip_bspline_config &cfg, double *x,
                            const int *nbd, const double *lo, const double *hi,
                            sitkhip_bspline_result *res) {
    LbfgsbOptions o;
    o.corrections = cfg.corrections;
    o.factr = cfg.factr;
    o.pgtol = cfg.pgtol;
    o.max_iterations = cfg.iterations;
    o.max_evaluations = cfg.max_evaluations;
    const int S = u.S;
    const size_t np = u.np;
    std::vector<std::unique_ptr<Lbfgsb>> opt(S);
    std::vector<int> act, next;
    std::vector<double> fbest(S, 0.0);
    std::vector<size_t> vbest(S, 0);
    std::vector<int> nev(S, 0);
    std::vector<char> want(S, 0);
    for (int s = 0; s < S; ++s) {
        opt[s].reset(new Lbfgsb(np, o, lo, hi, nbd));
        if (opt[s]->start(x + s * np)) {
            act.push_back(s);
            std::memcpy(u.h_coef + s * np, opt[s]->point(), np * sizeof(double));
        }
    }
    HostPool pool(S);
    // two lanes of contiguous starts once there are enough to share the
    // device: while one lane's evaluation runs, the other's optimizers step
    const int nlanes = S >= 8 ? 2 : 1;
    StartLane lanes[2];
    for (int l = 0; l < nlanes; ++l) {
        StartLane &L = lanes[l];
        L.lo = S * l / nlanes;
        L.hi = S * (l + 1) / nlanes;
        for (int s : act)
            if (s >= L.lo && s < L.hi) L.act.push_back(s);
        L.active.alloc(static_cast<size_t>(L.hi - L.lo));
        HIP_OK(hipHostMalloc(reinterpret_cast<void **>(&L.h_active), (L.hi - L.lo) * sizeof(int),
                             hipHostMallocDefault));
        HIP_OK(hipEventCreateWithFlags(&L.done, hipEventDisableTiming));
    }
    // on an error, let the evaluations in flight finish before the buffers go
    struct DrainOnError {
        ~DrainOnError() {
            if (std::uncaught_exceptions() > 0) (void)hipStreamSynchronize(backend().stream);
        }
    } drain;
    for (int l = 0; l < nlanes; ++l)
        if (!lanes[l].act.empty()) lane_launch(u, lanes[l]);
    for (bool busy = true; busy;) {
        busy = false;
        for (int l = 0; l < nlanes; ++l) {
            StartLane &L = lanes[l];
            if (!L.pending) continue;
            lane_wait(u, L);
            // the optimizers' updates are independent: one start per task
            pool.run(static_cast<int>(L.act.size()), [&](int i) {
                const int s = L.act[i];
                const double f = u.h_tail[s].value;
                if (nev[s] == 0 || f < fbest[s]) {
                    fbest[s] = f;
                    vbest[s] = static_cast<size_t>(u.h_tail[s].valid);
                }
                ++nev[s];
                want[s] = opt[s]->tell(f, u.h_deriv + s * np) ? 1 : 0;
                if (want[s]) std::memcpy(u.h_coef + s * np, opt[s]->point(), np * sizeof(double));
            });
            next.clear();
            for (int s : L.act)
                if (want[s]) next.push_back(s);
            L.act.swap(next);
            if (!L.act.empty()) {
                lane_launch(u, L);
                busy = true;
            }
        }
    }
    for (int s = 0; s < S; ++s) {
        std::memcpy(x + s * np, opt[s]->best(), np * sizeof(double));
        res[s].value = opt[s]->best_value();
        res[s].valid_points = vbest[s];
        res[s].iterations = opt[s]->iterations();
        res[s].evaluations = opt[s]->evaluations();
        res[s].stop = opt[s]->stop();
    }
}

// The Mattes metric and B-spline grid of a deformable registration: the
// transform's domain is the fixed image's extent (bspline.rs:324-339) with
// `mesh` cells per axis.
struct BSplineSetup {
    std::unique_ptr<Metric, void (*)(Metric *)> metric{nullptr, metric_destroy};
    std::unique_ptr<BSplineMetric> bs;
};

static void bspline_setup(Image &f, Image &mv, const sitkhip_bspline_config &cfg, size_t ncoeffs,
                          BSplineSetup &out) {
    const sitkhip_geometry &g = f.geom;
    size_t gs[3];
    double origin[3], spacing[3], shift[3];
    for (int d = 0; d < 3; ++d) {
        if (cfg.mesh[d] < 1) fail(SITKHIP_ERR_INVALID_ARGUMENT, "mesh must be >= 1 per axis");
        gs[d] = cfg.mesh[d] + 3;
        spacing[d] = static_cast<double>(g.size[d]) * g.spacing[d] / static_cast<double>(cfg.mesh[d]);
        shift[d] = -0.5 * spacing[d] * (3.0 - 1.0);
    }
    for (int r = 0; r < 3; ++r) {
        double acc = 0.0;
        for (int c = 0; c < 3; ++c) acc += g.direction[r * 3 + c] * shift[c];
        origin[r] = g.origin[r] + acc;
    }
    const size_t ncp = gs[0] * gs[1] * gs[2];
    if (ncoeffs != 3 * ncp) fail(SITKHIP_ERR_INVALID_ARGUMENT, "coefficient count must be 3 x (mesh+3)^3");
    Metric *m = nullptr;
    if (cfg.sampling != 0 && cfg.sampling != 2)
        fail(SITKHIP_ERR_INVALID_ARGUMENT, "B-spline registration samples every voxel or draws at random");
    check_sampling_fraction(cfg.sampling, cfg.sampling_percentage, f.len());
    if (cfg.sampling == 2) {
        const size_t cnt = static_cast<size_t>(static_cast<double>(f.len()) * cfg.sampling_percentage);
        m = metric_create_drawn(SITKHIP_METRIC_MATTES, f, cnt, cfg.seed, mv, cfg.precision);
    } else {
        m = metric_create(SITKHIP_METRIC_MATTES, f, nullptr, 0, nullptr, mv, nullptr, cfg.precision);
    }
    out.metric.reset(m);
    double fr[2], mr[2];
    metric_value_ranges(*m, fr, mr);
    sitkhip_mattes_bins bins;
    mattes_bins_from_ranges(fr, mr, static_cast<size_t>(cfg.bins), &bins);
    mattes_set_bins(*m, bins);
    out.bs.reset(bspline_create(*m, gs, origin, spacing, g.direction));
}

}  // namespace sitkhip

extern "C" {

int sitkhip_bspline_multi_create(sitkhip_bspline h, size_t nstarts, const double *stages,
                                 sitkhip_bspline_multi *out) {
    return run_guarded([&] {
        auto b = reinterpret_cast<BSplineMetric *>(h);
        if (!b || !out) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null argument");
        if (nstarts < 1 || nstarts > 4096) fail(SITKHIP_ERR_INVALID_ARGUMENT, "start count must be 1..4096");
        *out = reinterpret_cast<sitkhip_bspline_multi>(
            bspline_multi_create(*b, static_cast<int>(nstarts), stages));
    });
}

void sitkhip_bspline_multi_free(sitkhip_bspline_multi h) { delete reinterpret_cast<BSplineMulti *>(h); }

int sitkhip_bspline_multi_value_and_derivative(sitkhip_bspline_multi h, size_t nactive,
                                               const int *active, const double *coeffs,
                                               double *values, double *derivs, size_t *valid) {
    return run_guarded([&] {
        auto u = reinterpret_cast<BSplineMulti *>(h);
        if (!u || !active || !coeffs) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null argument");
        if (nactive > static_cast<size_t>(u->S)) fail(SITKHIP_ERR_INVALID_ARGUMENT, "more active starts than starts");
        std::vector<int> act(active, active + nactive);
        for (int s : act) {
            if (s < 0 || s >= u->S) fail(SITKHIP_ERR_INVALID_ARGUMENT, "active start out of range");
            std::memcpy(u->h_coef + s * u->np, coeffs + s * u->np, u->np * sizeof(double));
        }
        bspline_multi_evaluate(*u, act.data(), static_cast<int>(act.size()));
        for (int s : act) {
            if (values) values[s] = u->h_tail[s].value;
            if (valid) valid[s] = static_cast<size_t>(u->h_tail[s].valid);
            if (derivs) std::memcpy(derivs + s * u->np, u->h_deriv + s * u->np, u->np * sizeof(double));
        }
    });
}

int sitkhip_register_bspline_multistart(sitkhip_image fixed, sitkhip_image moving,
                                        const sitkhip_bspline_config *cfg, size_t nstarts,
                                        const double *stages, double *coeffs, size_t ncoeffs,
                                        const int *bound_selection, const double *lower,
                                        const double *upper, sitkhip_bspline_result *results,
                                        size_t *best) {
    return run_guarded([&] {
        auto f = reinterpret_cast<Image *>(fixed);
        auto mv = reinterpret_cast<Image *>(moving);
        if (!f || !mv || !cfg || !coeffs || !results) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null argument");
        if (nstarts < 1 || nstarts > 4096) fail(SITKHIP_ERR_INVALID_ARGUMENT, "start count must be 1..4096");
        const auto t0 = std::chrono::steady_clock::now();
        BSplineSetup su;
        bspline_setup(*f, *mv, *cfg, ncoeffs, su);
        std::unique_ptr<BSplineMulti> u(bspline_multi_create(*su.bs, static_cast<int>(nstarts), stages));
        register_starts(*u, *cfg, coeffs, bound_selection, lower, upper, results);
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        size_t bi = 0;
        for (size_t s = 0; s < nstarts; ++s) {
            results[s].seconds = secs;
            if (results[s].value < results[bi].value) bi = s;
        }
        if (best) *best = bi;
    });
}

int sitkhip_register_bspline(sitkhip_image fixed, sitkhip_image moving, const sitkhip_bspline_config *cfg,
                             double *coeffs, size_t ncoeffs, sitkhip_bspline_result *result) {
    return sitkhip_register_bspline_multistart(fixed, moving, cfg, 1, nullptr, coeffs, ncoeffs, nullptr,
                                               nullptr, nullptr, result, nullptr);
}

}  // extern "C"
