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
// weight), in sample order; then every control point gathers the partials of
// the (up to 64) cells whose support holds it, in cell order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

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

// Value pass over the cell-grouped sample list (any order gives the same
// integer histogram).
template <typename R>
__global__ __launch_bounds__(kBlock) void bs_hist(FixedGrid F, MovingVol M, BSplineMap B,
                                                  const ChainConst *__restrict__ C, SampleSet SS, const uint8_t *__restrict__ fbin,
                                                  MattesParams P, unsigned long long *__restrict__ g_joint,
                                                  unsigned long long *__restrict__ g_marg) {
    extern __shared__ unsigned long long s_hist[];
    const int nb = P.bins * P.bins + P.bins;
    const unsigned copy = blockIdx.x % kHistCopies;
    g_joint += static_cast<size_t>(copy) * P.bins * P.bins;
    g_marg += copy * P.bins;
    for (int i = threadIdx.x; i < nb; i += kBlock) s_hist[i] = 0ull;
    __syncthreads();
    long long begin, end;
    unsigned logical;
    block_slice(SS.n, begin, end, logical);
    const IndexMap none{};
    hist_range<R, false>(F, M, B, none, C, SS, fbin, P, s_hist, s_hist + P.bins * P.bins,
                         static_cast<SampleRec<R> *>(nullptr), begin, end);
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += kBlock) {
        const unsigned long long v = s_hist[i];
        if (v) atomicAdd(i < P.bins * P.bins ? g_joint + i : g_marg + (i - P.bins * P.bins), v);
    }
}

constexpr int kCellChunk = 256;

// One workgroup per cell: out[cell][d * 64 + (k * 16 + j * 4 + i)] =
// sum over the cell's samples of g_d * w0[i] w1[j] w2[k].
template <typename R>
__device__ __forceinline__ void bs_cells_body(const FixedGrid &F, const MovingVol &M, const BSplineMap &B,
                                              const uint32_t *__restrict__ idx,
                                              const uint8_t *__restrict__ fbin,
                                              const int *__restrict__ cell_begin,
                                              const MattesParams &P, const R *__restrict__ pratio,
                                              double *__restrict__ out) {
    __shared__ double s_g[kCellChunk][3];
    __shared__ double s_w[kCellChunk][12];
    const int cell = blockIdx.x;
    const int b = cell_begin[cell], e = cell_begin[cell + 1];
    const int t = threadIdx.x;
    double acc = 0.0;
    const int od = t / 64, oi = t & 3, oj = (t >> 2) & 3, ok = (t >> 4) & 3;
    for (int s0 = b; s0 < e; s0 += kCellChunk) {
        const int s = s0 + t;
        double g[3] = {0.0, 0.0, 0.0};
        double w[3][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
        if (s < e && fbin[s] != kNoBin) {
            double x[3], p[3];
            fixed_point(F, idx[s], x);
            int start[3];
            if (bspline_support(B, x, start, w)) {
                map_point(B, x, p);
                Cell cl;
                if (locate(M, p, cl)) {
                    float v[8];
                    corners(M, cl, v);
                    R fr[3] = {static_cast<R>(cl.frac[0]), static_cast<R>(cl.frac[1]),
                               static_cast<R>(cl.frac[2])};
                    R mv, gi[3], gp[3];
                    Trilinear<R>::template eval<true>(fr, v, mv, gi);
                    if (Parzen<R>::in_range(P, mv)) {
                        physical_gradient<R>(M, gi, gp);
                        R arg[4], dv[4];
                        bool mid;
                        const int m0 = Parzen<R>::taps(P, mv, arg, mid);
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
                                double acc = 0.0;
#pragma unroll
                                for (int r = 0; r < 3; ++r)
                                    acc += B.stage[r * 3 + d] * static_cast<double>(gp[r]);
                                gq[d] = acc;
                            }
                        }
#pragma unroll
                        for (int d = 0; d < 3; ++d) g[d] = static_cast<double>(a) * gq[d];
                    }
                }
            }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) s_g[t][d] = g[d];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int k = 0; k < 4; ++k) s_w[t][d * 4 + k] = w[d][k];
        __syncthreads();
        if (t < 192) {
            const int n = min(kCellChunk, e - s0);
            for (int q = 0; q < n; ++q)
                acc += s_g[q][od] * ((s_w[q][oi] * s_w[q][4 + oj]) * s_w[q][8 + ok]);
        }
        __syncthreads();
    }
    if (t < 192) out[static_cast<size_t>(cell) * 192 + t] = acc;
}

template <typename R>
__global__ __launch_bounds__(kCellChunk) void bs_cells(FixedGrid F, MovingVol M, BSplineMap B,
                                                       const uint32_t *__restrict__ idx,
                                                       const uint8_t *__restrict__ fbin,
                                                       const int *__restrict__ cell_begin,
                                                       MattesParams P, const R *__restrict__ pratio,
                                                       double *__restrict__ out) {
    bs_cells_body<R>(F, M, B, idx, fbin, cell_begin, P, pratio, out);
}

// deriv[d][cp] = sum over the cells whose support holds cp (cell order).
__device__ __forceinline__ void bs_gather_body(const BSplineMap &B, const double *__restrict__ cellT,
                                               double *__restrict__ deriv) {
    const int cp = blockIdx.x * blockDim.x + threadIdx.x;
    if (cp >= B.ncp) return;
    const int m0 = B.gs[0] - 3, m1 = B.gs[1] - 3, m2 = B.gs[2] - 3;
    const int px = cp % B.gs[0], py = (cp / B.gs[0]) % B.gs[1], pz = cp / (B.gs[0] * B.gs[1]);
    double s[3] = {0.0, 0.0, 0.0};
    for (int sz = max(0, pz - 3); sz <= min(m2 - 1, pz); ++sz)
        for (int sy = max(0, py - 3); sy <= min(m1 - 1, py); ++sy)
            for (int sx = max(0, px - 3); sx <= min(m0 - 1, px); ++sx) {
                const int cell = sx + m0 * (sy + m1 * sz);
                const int o = (px - sx) + 4 * (py - sy) + 16 * (pz - sz);
#pragma unroll
                for (int d = 0; d < 3; ++d) s[d] += cellT[static_cast<size_t>(cell) * 192 + d * 64 + o];
            }
#pragma unroll
    for (int d = 0; d < 3; ++d) deriv[static_cast<size_t>(d) * B.ncp + cp] = s[d];
}

__global__ void bs_gather(BSplineMap B, const double *__restrict__ cellT, double *__restrict__ deriv) {
    bs_gather_body(B, cellT, deriv);
}

// ------------------------------------------------------- many starts at once
// The same passes for S independent starts that share the sample set and the
// images: blockIdx.y picks a start from the active list; each start has its
// own coefficients, moving-initial stage, histogram copies, pRatio, tail,
// cell partials and derivative. Per start the arithmetic is the single-start
// passes' (bit-equal results).
__device__ __forceinline__ BSplineMap start_map(const BSplineMap &B0, const double *__restrict__ coef,
                                                const double *__restrict__ stage, int s) {
    BSplineMap B = B0;
    B.coef = coef + static_cast<size_t>(s) * 3 * B0.ncp;
    B.post = 1;
#pragma unroll
    for (int i = 0; i < 12; ++i) B.stage[i] = stage[s * 12 + i];
    return B;
}

template <typename R>
__global__ __launch_bounds__(kBlock) void bs_hist_multi(FixedGrid F, MovingVol M, BSplineMap B0,
                                                        const double *__restrict__ coef,
                                                        const double *__restrict__ stage,
                                                        const int *__restrict__ active,
                                                        const ChainConst *__restrict__ C, SampleSet SS,
                                                        const uint8_t *__restrict__ fbin, MattesParams P,
                                                        unsigned long long *__restrict__ g_joint,
                                                        unsigned long long *__restrict__ g_marg) {
    extern __shared__ unsigned long long s_hist[];
    const int st = active[blockIdx.y];
    const BSplineMap B = start_map(B0, coef, stage, st);
    const int nb = P.bins * P.bins + P.bins;
    const unsigned copy = blockIdx.x % kHistCopies;
    g_joint += (static_cast<size_t>(st) * kHistCopies + copy) * P.bins * P.bins;
    g_marg += (static_cast<size_t>(st) * kHistCopies + copy) * P.bins;
    for (int i = threadIdx.x; i < nb; i += kBlock) s_hist[i] = 0ull;
    __syncthreads();
    long long begin, end;
    unsigned logical;
    block_slice(SS.n, begin, end, logical);
    const IndexMap none{};
    hist_range<R, false>(F, M, B, none, C, SS, fbin, P, s_hist, s_hist + P.bins * P.bins,
                         static_cast<SampleRec<R> *>(nullptr), begin, end);
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += kBlock) {
        const unsigned long long v = s_hist[i];
        if (v) atomicAdd(i < P.bins * P.bins ? g_joint + i : g_marg + (i - P.bins * P.bins), v);
    }
}

// One workgroup per active start: the Mattes tail on its histogram copies
// (cleared for the next evaluation), pRatio in double and float.
__global__ __launch_bounds__(kTail) void bs_tail_multi(const int *__restrict__ active, int B, double mbs,
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

template <typename R>
__global__ __launch_bounds__(kCellChunk) void bs_cells_multi(FixedGrid F, MovingVol M, BSplineMap B0,
                                                             const double *__restrict__ coef,
                                                             const double *__restrict__ stage,
                                                             const int *__restrict__ active,
                                                             const uint32_t *__restrict__ idx,
                                                             const uint8_t *__restrict__ fbin,
                                                             const int *__restrict__ cell_begin,
                                                             MattesParams P, const R *__restrict__ pratio,
                                                             int ncells, double *__restrict__ out) {
    const int st = active[blockIdx.y];
    const BSplineMap B = start_map(B0, coef, stage, st);
    bs_cells_body<R>(F, M, B, idx, fbin, cell_begin, P,
                     pratio + static_cast<size_t>(st) * P.bins * P.bins,
                     out + static_cast<size_t>(st) * ncells * 192);
}

__global__ void bs_gather_multi(BSplineMap B, const int *__restrict__ active, int ncells,
                                const double *__restrict__ cellT, double *__restrict__ deriv) {
    const int st = active[blockIdx.y];
    bs_gather_body(B, cellT + static_cast<size_t>(st) * ncells * 192,
                   deriv + static_cast<size_t>(st) * 3 * B.ncp);
}

// ------------------------------------------------------------------ host side
struct BSplineMetric {
    Metric *m = nullptr;        // the Mattes metric (bins set); not owned
    BSplineMap B{};
    int ncells = 0;
    SampleSet SS{};             // cell-grouped samples
    DeviceBuffer<uint32_t> idx;
    DeviceBuffer<uint8_t> fbin;
    DeviceBuffer<int> cell_begin;
    DeviceBuffer<double> coef, cellT, deriv;
    HostResult *res = nullptr, *res_dev = nullptr;
    ~BSplineMetric() { pinned_slot_free(res); }
};

BSplineMetric *bspline_create(Metric &m, const size_t gs[3], const double origin[3],
                              const double spacing[3], const double direction[9]) {
    if (m.kind != SITKHIP_METRIC_MATTES || !m.bins_set)
        fail(SITKHIP_ERR_HISTOGRAM_SHAPE, "a B-spline metric needs a Mattes metric with bins set");
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
        b->cell_begin.alloc(b->ncells + 1);
        b->cell_begin.upload(begin.data(), b->ncells + 1);
        b->SS.idx = b->idx.ptr;
        b->SS.n = static_cast<long long>(n);
        b->coef.alloc(3 * ncp);
        b->B.coef = b->coef.ptr;
        b->cellT.alloc(static_cast<size_t>(b->ncells) * 192);
        b->deriv.alloc(3 * ncp);
        void *dev = nullptr;
        b->res = static_cast<HostResult *>(pinned_slot_alloc(sizeof(HostResult), &dev));
        b->res_dev = static_cast<HostResult *>(dev);
        sync();
    } catch (...) {
        delete b;
        throw;
    }
    return b;
}

template <typename R>
static void bs_value_pass(BSplineMetric &b) {
    Metric &m = *b.m;
    const int B = m.P.bins;
    if (B > kLdsMaxBins) fail(SITKHIP_ERR_TOO_MANY_BINS, "the B-spline value pass holds at most 90 bins in LDS");
    const size_t shm = static_cast<size_t>(B * B + B) * sizeof(unsigned long long);
    {
        LaunchScope scope("bs_hist");
        hipLaunchKernelGGL(bs_hist<R>, dim3(kGrid), dim3(kBlock), shm, backend().stream, m.F, m.Mv,
                           b.B, m.chain.ptr, b.SS, b.fbin.ptr, m.P, m.joint.ptr, m.marg.ptr);
        check_launch("bs_hist");
    }
    launch_tail(m, &b.res_dev->tail);
}

template <typename R>
static void bs_deriv_pass(BSplineMetric &b) {
    Metric &m = *b.m;
    const R *pr = sizeof(R) == 4 ? reinterpret_cast<const R *>(m.pratio_f.ptr)
                                 : reinterpret_cast<const R *>(m.pratio.ptr);
    if (b.ncells > 0) {
        LaunchScope scope("bs_cells");
        hipLaunchKernelGGL(bs_cells<R>, dim3(b.ncells), dim3(kCellChunk), 0, backend().stream, m.F,
                           m.Mv, b.B, b.idx.ptr, b.fbin.ptr, b.cell_begin.ptr, m.P, pr, b.cellT.ptr);
        check_launch("bs_cells");
    }
    LaunchScope scope("bs_gather");
    hipLaunchKernelGGL(bs_gather, dim3((b.B.ncp + 255) / 256), dim3(256), 0, backend().stream, b.B,
                       b.cellT.ptr, b.deriv.ptr);
    check_launch("bs_gather");
}

void bspline_evaluate(BSplineMetric &b, const double *coeffs, bool derivative, double *value,
                      double *deriv, size_t *valid) {
    if (!coeffs) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null coefficients");
    const size_t n = 3 * static_cast<size_t>(b.B.ncp);
    b.coef.upload(coeffs, n);
    const bool f64 = b.m->precision == SITKHIP_PRECISION_F64;
    if (f64) bs_value_pass<double>(b);
    else bs_value_pass<float>(b);
    if (derivative) {
        if (f64) bs_deriv_pass<double>(b);
        else bs_deriv_pass<float>(b);
        if (deriv) b.deriv.download(deriv, n);
    }
    sync();
    if (value) *value = b.res->tail.value;
    if (valid) *valid = static_cast<size_t>(b.res->tail.valid);
    if (derivative && deriv && b.res->tail.valid == 0.0)
        for (size_t i = 0; i < n; ++i) deriv[i] = 0.0;
}

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
    unsigned G = 0;              // value-pass workgroups per start
    DeviceBuffer<double> coef, stage, pratio, cellT, deriv;
    DeviceBuffer<float> pratio_f;
    DeviceBuffer<unsigned long long> joint, marg;
    DeviceBuffer<TailOut> tail;
    DeviceBuffer<int> active;
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

static const double kIdentityStage[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};

// stages: S x 12 (null: every start without a moving-initial transform).
BSplineMulti *bspline_multi_create(BSplineMetric &b, int S, const double *stages) {
    if (S < 1 || S > 4096) fail(SITKHIP_ERR_INVALID_ARGUMENT, "start count must be 1..4096");
    const int B = b.m->P.bins;
    if (B > kLdsMaxBins) fail(SITKHIP_ERR_TOO_MANY_BINS, "the B-spline value pass holds at most 90 bins in LDS");
    auto *u = new BSplineMulti();
    try {
        u->b = &b;
        u->S = S;
        u->np = 3 * static_cast<size_t>(b.B.ncp);
        // about 4096 samples per workgroup, a multiple of the 8 histogram copies
        size_t g = (static_cast<size_t>(b.SS.n) + 4095) / 4096;
        g = std::max<size_t>(8, std::min<size_t>(1024, (g + 7) / 8 * 8));
        u->G = static_cast<unsigned>(g);
        const size_t nb = static_cast<size_t>(B) * B;
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
        HIP_OK(hipMemsetAsync(u->joint.ptr, 0, u->joint.bytes(), backend().stream));
        HIP_OK(hipMemsetAsync(u->marg.ptr, 0, u->marg.bytes(), backend().stream));
        HIP_OK(hipHostMalloc(reinterpret_cast<void **>(&u->h_coef), S * u->np * sizeof(double), hipHostMallocDefault));
        HIP_OK(hipHostMalloc(reinterpret_cast<void **>(&u->h_deriv), S * u->np * sizeof(double), hipHostMallocDefault));
        HIP_OK(hipHostMalloc(reinterpret_cast<void **>(&u->h_tail), S * sizeof(TailOut), hipHostMallocDefault));
        HIP_OK(hipHostMalloc(reinterpret_cast<void **>(&u->h_active), S * sizeof(int), hipHostMallocDefault));
        std::vector<double> st(static_cast<size_t>(S) * 12);
        for (int i = 0; i < S; ++i)
            std::memcpy(st.data() + 12 * i, stages ? stages + 12 * i : kIdentityStage, 12 * sizeof(double));
        u->stage.upload(st.data(), st.size());
        std::memset(u->h_coef, 0, S * u->np * sizeof(double));
        sync();
    } catch (...) {
        delete u;
        throw;
    }
    return u;
}

template <typename R>
static void multi_passes(BSplineMulti &u, int nact) {
    BSplineMetric &b = *u.b;
    Metric &m = *b.m;
    const int B = m.P.bins;
    const hipStream_t st = backend().stream;
    const size_t shm = static_cast<size_t>(B * B + B) * sizeof(unsigned long long);
    {
        LaunchScope scope("bs_hist_multi");
        hipLaunchKernelGGL(bs_hist_multi<R>, dim3(u.G, nact), dim3(kBlock), shm, st, m.F, m.Mv, b.B,
                           u.coef.ptr, u.stage.ptr, u.active.ptr, m.chain.ptr, b.SS, b.fbin.ptr, m.P,
                           u.joint.ptr, u.marg.ptr);
        check_launch("bs_hist_multi");
    }
    {
        LaunchScope scope("bs_tail_multi");
        hipLaunchKernelGGL(bs_tail_multi, dim3(nact), dim3(kTail), shm, st, u.active.ptr, B, m.P.mbs,
                           u.joint.ptr, u.marg.ptr, u.pratio.ptr, u.pratio_f.ptr, u.tail.ptr);
        check_launch("bs_tail_multi");
    }
    const R *pr = sizeof(R) == 4 ? reinterpret_cast<const R *>(u.pratio_f.ptr)
                                 : reinterpret_cast<const R *>(u.pratio.ptr);
    if (b.ncells > 0) {
        LaunchScope scope("bs_cells_multi");
        hipLaunchKernelGGL(bs_cells_multi<R>, dim3(b.ncells, nact), dim3(kCellChunk), 0, st, m.F, m.Mv,
                           b.B, u.coef.ptr, u.stage.ptr, u.active.ptr, b.idx.ptr, b.fbin.ptr,
                           b.cell_begin.ptr, m.P, pr, b.ncells, u.cellT.ptr);
        check_launch("bs_cells_multi");
    }
    LaunchScope scope("bs_gather_multi");
    hipLaunchKernelGGL(bs_gather_multi, dim3((b.B.ncp + 255) / 256, nact), dim3(256), 0, st, b.B,
                       u.active.ptr, b.ncells, u.cellT.ptr, u.deriv.ptr);
    check_launch("bs_gather_multi");
}

// Evaluates value and derivative of the listed starts at the coefficients
// already in u.h_coef (start s at h_coef + s * np). Results land in u.h_tail
// and u.h_deriv (zero derivative for a start without valid samples).
void bspline_multi_evaluate(BSplineMulti &u, const int *act, int nact) {
    if (nact <= 0) return;
    if (nact > u.S) fail(SITKHIP_ERR_INVALID_ARGUMENT, "more active starts than starts");
    for (int i = 0; i < nact; ++i) {
        if (act[i] < 0 || act[i] >= u.S) fail(SITKHIP_ERR_INVALID_ARGUMENT, "active start out of range");
        u.h_active[i] = act[i];
    }
    const hipStream_t st = backend().stream;
    const size_t bytes = u.np * sizeof(double);
    for (int i = 0; i < nact; ++i)
        HIP_OK(hipMemcpyAsync(u.coef.ptr + act[i] * u.np, u.h_coef + act[i] * u.np, bytes,
                              hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(u.active.ptr, u.h_active, nact * sizeof(int), hipMemcpyHostToDevice, st));
    if (u.b->m->precision == SITKHIP_PRECISION_F64) multi_passes<double>(u, nact);
    else multi_passes<float>(u, nact);
    for (int i = 0; i < nact; ++i) {
        HIP_OK(hipMemcpyAsync(u.h_deriv + act[i] * u.np, u.deriv.ptr + act[i] * u.np, bytes,
                              hipMemcpyDeviceToHost, st));
        HIP_OK(hipMemcpyAsync(u.h_tail + act[i], u.tail.ptr + act[i], sizeof(TailOut),
                              hipMemcpyDeviceToHost, st));
    }
    sync();
    for (int i = 0; i < nact; ++i)
        if (u.h_tail[act[i]].valid == 0.0) std::memset(u.h_deriv + act[i] * u.np, 0, bytes);
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
    for (int s = 0; s < S; ++s) {
        opt[s].reset(new Lbfgsb(np, o, lo, hi, nbd));
        if (opt[s]->start(x + s * np)) act.push_back(s);
    }
    while (!act.empty()) {
        for (int s : act) std::memcpy(u.h_coef + s * np, opt[s]->point(), np * sizeof(double));
        bspline_multi_evaluate(u, act.data(), static_cast<int>(act.size()));
        next.clear();
        for (int s : act) {
            const double f = u.h_tail[s].value;
            if (nev[s] == 0 || f < fbest[s]) {
                fbest[s] = f;
                vbest[s] = static_cast<size_t>(u.h_tail[s].valid);
            }
            ++nev[s];
            if (opt[s]->tell(f, u.h_deriv + s * np)) next.push_back(s);
        }
        act.swap(next);
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
