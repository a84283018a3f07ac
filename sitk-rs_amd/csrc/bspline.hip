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

__device__ __forceinline__ void map_point(const BSplineMap &B, const double x[3], double p[3]) {
    int start[3];
    double w[3][4];
    if (!bspline_support(B, x, start, w)) {
        p[0] = x[0]; p[1] = x[1]; p[2] = x[2];
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
__global__ __launch_bounds__(kCellChunk) void bs_cells(FixedGrid F, MovingVol M, BSplineMap B,
                                                       const uint32_t *__restrict__ idx,
                                                       const uint8_t *__restrict__ fbin,
                                                       const int *__restrict__ cell_begin,
                                                       MattesParams P, const R *__restrict__ pratio,
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
#pragma unroll
                        for (int d = 0; d < 3; ++d)
                            g[d] = static_cast<double>(a) * static_cast<double>(gp[d]);
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

// deriv[d][cp] = sum over the cells whose support holds cp (cell order).
__global__ void bs_gather(BSplineMap B, const double *__restrict__ cellT, double *__restrict__ deriv) {
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

}  // extern "C"

// ============================================================ registration
// A B-spline registration on the device metric, driven by a limited-memory
// BFGS optimizer on the host (LBFGSBOptimizer with every variable unbounded,
// src/registration/lbfgsb.rs:1590-1674: the case the reference documents as
// reducing to unconstrained L-BFGS). Its stopping rules are ITK's: relative
// reduction (f_k - f_{k+1}) / max(|f_k|, |f_{k+1}|, 1) <= factr * eps, the
// gradient's infinity norm <= pgtol, the iteration and evaluation caps; the
// result is the best point evaluated. The step search is a backtracking
// Armijo search from the unit step (the reference uses the netlib Moré-Thuente
// search; recorded as a deviation).
namespace sitkhip {

struct LbfgsResult {
    double value;
    size_t valid;
    int iterations;
    int evaluations;
    int stop;  // 0 max iterations, 1 gradient, 2 relative reduction, 3 evaluations, 4 no descent
};

static double dot(const std::vector<double> &a, const std::vector<double> &b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

static LbfgsResult lbfgs(BSplineMetric &b, std::vector<double> &x, int max_it, int max_eval, int m,
                         double factr, double pgtol) {
    const size_t n = x.size();
    std::vector<double> g(n), gn(n), xn(n), d(n), q(n);
    std::vector<std::vector<double>> S, Y;
    std::vector<double> rho;
    LbfgsResult r{DBL_MAX, 0, 0, 0, 0};
    size_t valid = 0;
    double f = 0.0;
    bspline_evaluate(b, x.data(), true, &f, g.data(), &valid);
    r.evaluations = 1;
    std::vector<double> best = x;
    double fbest = f;
    size_t vbest = valid;
    for (;;) {
        double ginf = 0.0;
        for (double v : g) ginf = std::max(ginf, std::fabs(v));
        if (ginf <= pgtol) { r.stop = 1; break; }
        if (r.iterations >= max_it) { r.stop = 0; break; }
        // two-loop recursion (Nocedal 1980) for d = -H g
        q = g;
        const int k = static_cast<int>(S.size());
        std::vector<double> alpha(k);
        for (int i = k - 1; i >= 0; --i) {
            alpha[i] = rho[i] * dot(S[i], q);
            for (size_t j = 0; j < n; ++j) q[j] -= alpha[i] * Y[i][j];
        }
        double gamma = 1.0 / std::max(ginf, 1e-300);  // first step: unit infinity-norm move
        if (k > 0) gamma = dot(S[k - 1], Y[k - 1]) / dot(Y[k - 1], Y[k - 1]);
        for (size_t j = 0; j < n; ++j) q[j] *= gamma;
        for (int i = 0; i < k; ++i) {
            const double beta = rho[i] * dot(Y[i], q);
            for (size_t j = 0; j < n; ++j) q[j] += S[i][j] * (alpha[i] - beta);
        }
        for (size_t j = 0; j < n; ++j) d[j] = -q[j];
        double gd = dot(g, d);
        if (!(gd < 0.0)) {  // not a descent direction: restart from steepest descent
            S.clear(); Y.clear(); rho.clear();
            for (size_t j = 0; j < n; ++j) d[j] = -g[j] / std::max(ginf, 1e-300);
            gd = dot(g, d);
        }
        double t = 1.0, fn = 0.0;
        size_t vn = 0;
        bool accepted = false;
        for (int tries = 0; tries < 30 && r.evaluations < max_eval; ++tries) {
            for (size_t j = 0; j < n; ++j) xn[j] = x[j] + t * d[j];
            bspline_evaluate(b, xn.data(), true, &fn, gn.data(), &vn);
            ++r.evaluations;
            if (fn <= f + 1e-4 * t * gd) { accepted = true; break; }
            t *= 0.5;
        }
        if (!accepted) { r.stop = r.evaluations >= max_eval ? 3 : 4; break; }
        ++r.iterations;
        std::vector<double> s(n), y(n);
        for (size_t j = 0; j < n; ++j) { s[j] = xn[j] - x[j]; y[j] = gn[j] - g[j]; }
        const double sy = dot(s, y);
        if (sy > DBL_EPSILON * dot(y, y)) {
            if (static_cast<int>(S.size()) == m) {
                S.erase(S.begin()); Y.erase(Y.begin()); rho.erase(rho.begin());
            }
            S.push_back(s); Y.push_back(y); rho.push_back(1.0 / sy);
        }
        const double fprev = f;
        x = xn; g = gn; f = fn; valid = vn;
        if (f < fbest) { fbest = f; best = x; vbest = valid; }
        if ((fprev - f) / std::max({std::fabs(fprev), std::fabs(f), 1.0}) <= factr * DBL_EPSILON) {
            r.stop = 2;
            break;
        }
        if (r.evaluations >= max_eval) { r.stop = 3; break; }
    }
    x = best;
    r.value = fbest;
    r.valid = vbest;
    return r;
}

}  // namespace sitkhip

extern "C" int sitkhip_register_bspline(sitkhip_image fixed, sitkhip_image moving,
                                        const sitkhip_bspline_config *cfg, double *coeffs,
                                        size_t ncoeffs, sitkhip_bspline_result *result) {
    return run_guarded([&] {
        auto f = reinterpret_cast<Image *>(fixed);
        auto mv = reinterpret_cast<Image *>(moving);
        if (!f || !mv || !cfg || !coeffs || !result) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null argument");
        const auto t0 = std::chrono::steady_clock::now();
        const sitkhip_geometry &g = f->geom;
        size_t gs[3];
        double origin[3], spacing[3], shift[3];
        for (int d = 0; d < 3; ++d) {
            if (cfg->mesh[d] < 1) fail(SITKHIP_ERR_INVALID_ARGUMENT, "mesh must be >= 1 per axis");
            gs[d] = cfg->mesh[d] + 3;
            // the transform domain is the fixed image's extent (bspline.rs:324-339)
            spacing[d] = static_cast<double>(g.size[d]) * g.spacing[d] / static_cast<double>(cfg->mesh[d]);
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
        if (cfg->sampling == 2) {
            const size_t cnt = static_cast<size_t>(static_cast<double>(f->len()) * cfg->sampling_percentage);
            m = metric_create_drawn(SITKHIP_METRIC_MATTES, *f, cnt, cfg->seed, *mv, cfg->precision);
        } else {
            m = metric_create(SITKHIP_METRIC_MATTES, *f, nullptr, 0, nullptr, *mv, nullptr, cfg->precision);
        }
        std::unique_ptr<Metric, void (*)(Metric *)> mh(m, metric_destroy);
        double fr[2], mr[2];
        metric_value_ranges(*m, fr, mr);
        sitkhip_mattes_bins bins;
        mattes_bins_from_ranges(fr, mr, static_cast<size_t>(cfg->bins), &bins);
        mattes_set_bins(*m, bins);
        std::unique_ptr<BSplineMetric> b(bspline_create(*m, gs, origin, spacing, g.direction));
        std::vector<double> x(coeffs, coeffs + ncoeffs);
        LbfgsResult r = lbfgs(*b, x, cfg->iterations, cfg->max_evaluations, cfg->corrections,
                              cfg->factr, cfg->pgtol);
        std::memcpy(coeffs, x.data(), ncoeffs * sizeof(double));
        result->value = r.value;
        result->valid_points = r.valid;
        result->iterations = r.iterations;
        result->evaluations = r.evaluations;
        result->stop = r.stop;
        result->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    });
}
