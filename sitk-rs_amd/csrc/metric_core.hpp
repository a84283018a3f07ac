// Device building blocks shared by the single-pair metric kernels
// (metric.hip) and the batched multi-pair kernels (batch.hip): sample sets, the
// Parzen window, the 64-bit fixed-point histogram, the per-range sample loops of
// the three passes, the Mattes tail, the FP64 reduction tree and the host-side
// state of one resident metric.
//
// Reference (crates/sitk/): src/registration/mattes.rs:92-328, 442-676,
// src/registration/metric.rs:1065-1223, src/cuda/ops/mattes.rs:75-91,
// src/cuda/ops/mean_squares.rs:101-145.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#include "common.hpp"
#include "sampler.hpp"

namespace sitkhip {

#ifndef SITKHIP_BLOCK
#define SITKHIP_BLOCK 256
#endif
constexpr int kBlock = SITKHIP_BLOCK;  // lanes per workgroup of the sample passes
#ifndef SITKHIP_GRID
#define SITKHIP_GRID 1024
#endif
constexpr unsigned kGrid = SITKHIP_GRID;  // fixed: the reduction tree must not depend on n
// The value pass's launch shape (its integer histogram does not depend on
// it): 512 workgroups (64 per XCD) of 512 lanes, two per CU. Measured on the
// 2M-sample level (value pass, us): 256 lanes x 1024 workgroups 26.5, x 2048
// 30.0, x 512 32.7; 512 x 1024 30.0, x 512 25.8; 1024 x 512 30.3, x 256 26.1.
// More workgroups add end-of-workgroup flushes (64-bit atomics into the
// XCD's histogram copy); fewer waves per CU leave gathers exposed.
#ifndef SITKHIP_HIST_GRID
#define SITKHIP_HIST_GRID 512
#endif
constexpr unsigned kHistGrid = SITKHIP_HIST_GRID;
#ifndef SITKHIP_HIST_BLOCK
#define SITKHIP_HIST_BLOCK 512
#endif
constexpr int kHistBlock = SITKHIP_HIST_BLOCK;
constexpr int kLdsMaxBins = 90;       // bins^2 * 8 B of LDS per workgroup (<= 64 KB)
constexpr double kFix = 4294967296.0; // 2^32 fixed-point units per histogram unit
constexpr double kInvFix = 1.0 / 4294967296.0;
constexpr uint8_t kNoBin = 255;
// The global joint histogram is kept as one copy per XCD (a workgroup folds
// its LDS histogram into copy blockIdx mod 8, the XCD it runs on), so the
// 64-bit atomics of different XCDs do not contend on one set of addresses; the
// tail adds the copies (exact integers).
#ifndef SITKHIP_HIST_COPIES
#define SITKHIP_HIST_COPIES 8
#endif
constexpr int kHistCopies = SITKHIP_HIST_COPIES;

struct SampleSet {
    const uint32_t *idx;  // nullptr: the whole fixed grid in order
    long long n;
};

struct MattesParams {
    int bins;
    double mbs, mnmin, inv_mbs;  // moving bin size, normalised min, 1/bin size
    double mmin, mmax;           // moving true range (reject outside)
    float mmin_f, mmax_f, inv_mbs_f;
};

// ------------------------------------------------------------- cubic B-spline
// itk::BSplineKernelFunction<3> and its derivative in the reference's
// expression order (registration/mattes.rs:96-124), piece by piece: the value
// on |u| < 1 (inner) and 1 <= |u| < 2 (outer); the derivative on [0, 1) (d1),
// (-1, 0) (d2), [1, 2) (d3) and (-2, -1] (d4).
template <typename R>
struct B3 {
    __device__ __forceinline__ static R inner(R a) {
        const R sq = a * a;
        return (R(4) - R(6) * sq + R(3) * sq * a) / R(6);
    }
    __device__ __forceinline__ static R outer(R a) {
        const R sq = a * a;
        return (R(8) - R(12) * a + R(6) * sq - sq * a) / R(6);
    }
    __device__ __forceinline__ static R d1(R u) { return R(-2) * u + R(1.5) * u * u; }
    __device__ __forceinline__ static R d2(R u) { return R(-2) * u - R(1.5) * u * u; }
    __device__ __forceinline__ static R d3(R u) { return R(-2) + R(2) * u - R(0.5) * u * u; }
    __device__ __forceinline__ static R d4(R u) { return R(2) + R(2) * u + R(0.5) * u * u; }
};
// FP32: the same polynomials times 6 ((2-a)^3 for the outer piece); the
// division by 6 is folded into the fixed-point scale of to_fixed<float>, in
// FP64. A float 1/6 would scale every weight by the same 1 + 3e-8: the joint
// mass would carry it into the fixed marginal's normalisation
// (mattes.rs:283-293) and shift the value by 3e-8 (2e-6 of an MI of 0.015).
template <>
__device__ __forceinline__ float B3<float>::inner(float a) {
    const float sq = a * a;
    return fmaf(fmaf(3.0f, a, -6.0f), sq, 4.0f);
}
template <>
__device__ __forceinline__ float B3<float>::outer(float a) {
    const float t = 2.0f - a;
    return t * t * t;
}

template <typename R>
__device__ __forceinline__ R bspline3(R u) {
    const R a = fabs(u);
    if (a < R(1)) return B3<R>::inner(a);
    if (a < R(2)) return B3<R>::outer(a);
    return R(0);
}

template <typename R>
__device__ __forceinline__ R bspline3_deriv(R u) {
    if (u >= R(0) && u < R(1)) return B3<R>::d1(u);
    if (u > R(-1) && u < R(0)) return B3<R>::d2(u);
    if (u >= R(1) && u < R(2)) return B3<R>::d3(u);
    if (u > R(-2) && u <= R(-1)) return B3<R>::d4(u);
    return R(0);
}

// The four taps of a Parzen window. When the bin index was not clamped
// (`mid`), arg = (-1-f, -f, 1-f, 2-f) with f in [0, 1): each tap lies in a
// known piece, and where two pieces meet (f = 0) they give the same bits, so
// these equal bspline3 / bspline3_deriv tap by tap. Otherwise the generic
// forms.
template <typename R>
__device__ __forceinline__ void parzen_values(const R arg[4], bool mid, R v[4]) {
    if (mid) {
        v[0] = B3<R>::outer(fabs(arg[0]));
        v[1] = B3<R>::inner(fabs(arg[1]));
        v[2] = B3<R>::inner(fabs(arg[2]));
        v[3] = arg[3] < R(2) ? B3<R>::outer(fabs(arg[3])) : R(0);
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = bspline3<R>(arg[k]);
    }
}
template <typename R>
__device__ __forceinline__ void parzen_derivs(const R arg[4], bool mid, R d[4]) {
    if (mid) {
        d[0] = B3<R>::d4(arg[0]);
        d[1] = arg[1] >= R(0) ? B3<R>::d1(arg[1]) : B3<R>::d2(arg[1]);
        d[2] = arg[2] < R(1) ? B3<R>::d1(arg[2]) : B3<R>::d3(arg[2]);
        d[3] = arg[3] < R(2) ? B3<R>::d3(arg[3]) : R(0);
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) d[k] = bspline3_deriv<R>(arg[k]);
    }
}

// Workgroup slice of the samples: contiguous, XCD-grouped (see file comment).
__device__ __forceinline__ void block_slice(long long n, long long &begin, long long &end,
                                            unsigned &logical) {
    const unsigned G = gridDim.x;
    const unsigned b = blockIdx.x;
    logical = (b % 8u) * (G / 8u) + b / 8u;
    const long long per = (n + G - 1) / G;
    begin = static_cast<long long>(logical) * per;
    end = begin + per;
    if (end > n) end = n;
}

// The value pass's walk of the samples: XCD x (workgroup mod 8) takes the
// x-th eighth of the samples and its G/8 workgroups sweep that range
// together, BLK (the workgroup size) consecutive samples each per step; a
// lane's samples are first + lane + k * step. The moving voxels one XCD reads at a time then lie
// in a thin band of slices that stays in its L2 (an integer histogram does
// not depend on which workgroup adds a sample).
template <int BLK = kBlock>
__device__ __forceinline__ void xcd_walk(long long n, long long &first, long long &step,
                                         long long &end) {
    const unsigned G = gridDim.x, b = blockIdx.x;
    const long long chunk = (n + 7) / 8;
    const long long xb = static_cast<long long>(b % 8u) * chunk;
    end = xb + chunk < n ? xb + chunk : n;
    first = xb + static_cast<long long>(b / 8u) * BLK;
    step = static_cast<long long>(G / 8u) * BLK;
}

__device__ __forceinline__ uint32_t sample_flat(const SampleSet &S, long long s) {
    return S.idx ? S.idx[s] : static_cast<uint32_t>(s);
}
// Parzen window of the moving value: the clamped bin index and the four tap
// arguments (m - term) for m = idx-1 .. idx+2 (mattes.rs:193-206, 464-481).
template <typename R>
struct Parzen;

template <>
struct Parzen<double> {
    __device__ __forceinline__ static int taps(const MattesParams &P, double mv, double arg[4],
                                               bool &mid) {
        const double term = mv / P.mbs - P.mnmin;
        long long idx = static_cast<long long>(term);
        mid = idx >= 2 && idx <= P.bins - 3;
        if (idx < 2) idx = 2;
        else if (idx > P.bins - 3) idx = P.bins - 3;
        const int start = static_cast<int>(idx) - 1;
#pragma unroll
        for (int k = 0; k < 4; ++k) arg[k] = static_cast<double>(start + k) - term;
        return start;
    }
    __device__ __forceinline__ static bool in_range(const MattesParams &P, double mv) {
        return !(mv < P.mmin || mv > P.mmax);
    }
};

template <>
struct Parzen<float> {
    // term = mv/binSize - normalizedMin = (mv - trueMin)/binSize + padding,
    // in FP64 (mv - trueMin is exact there), and each tap argument rounded to
    // float once. A float term carries the relative error of the rounded
    // 1/binSize into every sample alike -- a stretch of the moving bin axis
    // that does not average out (2e-6 of the value at 2M samples where MI is
    // small); rounding only the arguments leaves independent errors.
    // The arguments (start + k) - term = (k - 1) - u with u = term - idx: the
    // FP64 difference is exact (term and idx are within a factor of two) and
    // rounded to float once per sample, then each argument is one float
    // subtraction -- still independent roundings per sample.
    __device__ __forceinline__ static int taps(const MattesParams &P, float mv, float arg[4],
                                               bool &mid) {
        const double term = fma(static_cast<double>(mv) - P.mmin, P.inv_mbs, 2.0);
        int idx = static_cast<int>(term);
        mid = idx >= 2 && idx <= P.bins - 3;
        idx = idx < 2 ? 2 : (idx > P.bins - 3 ? P.bins - 3 : idx);
        const float u = static_cast<float>(term - static_cast<double>(idx));
        arg[0] = -1.0f - u;
        arg[1] = -u;
        arg[2] = 1.0f - u;
        arg[3] = 2.0f - u;
        return idx - 1;
    }
    __device__ __forceinline__ static bool in_range(const MattesParams &P, float mv) {
        return !(mv < P.mmin_f || mv > P.mmax_f);
    }
};

// B3 in [0, 2/3] (a rounding residue below zero becomes 0), scaled to 2^-32
// units: every contribution fits in 32 bits. The FP32 weights arrive times 6
// (B3<float>) and are scaled by 2^32 / 6 split into a float and its float
// remainder, so the product is rounded once from a scale exact to 2^-48: the
// rounding errors are then those of the individual weights, with no shared
// bias.
template <typename R>
__device__ __forceinline__ unsigned long long to_fixed(R v) {
    const double d = static_cast<double>(v);
    return d > 0.0 ? static_cast<unsigned long long>(llrint(d * kFix)) : 0ull;
}
constexpr double kFixSixth = 4294967296.0 / 6.0;
constexpr float kFixSixthHi = static_cast<float>(kFixSixth);
constexpr float kFixSixthLo = static_cast<float>(kFixSixth - static_cast<double>(kFixSixthHi));
// (a B3 weight is never negative; the max keeps the conversion defined)
template <>
__device__ __forceinline__ unsigned long long to_fixed<float>(float v) {
    v = fmaxf(v, 0.0f);
    return static_cast<unsigned long long>(__float2uint_rn(fmaf(v, kFixSixthHi, v * kFixSixthLo)));
}
// ------------------------------------------------------------ FP64 reduction
// A double moved across lanes by DPP (two 32-bit halves); lanes with no
// source lane, or in rows the row mask leaves out, read 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v) {
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<unsigned>(b)), CTRL,
                                               ROW_MASK, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<unsigned>(b >> 32)),
                                               CTRL, ROW_MASK, 0xf, true);
    return __longlong_as_double(static_cast<long long>(
        (static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) | static_cast<unsigned>(lo)));
}
// Wave64 sum in a fixed order, all in VALU: an inclusive scan inside each row
// of 16 lanes (row_shr 1, 2, 4, 8), then across rows (row_bcast 15 into rows
// 1 and 3, row_bcast 31 into rows 2 and 3). The total is in lane 63.
__device__ __forceinline__ double wave_total_f64(double v) {
    v += dpp_f64<0x111, 0xf>(v);
    v += dpp_f64<0x112, 0xf>(v);
    v += dpp_f64<0x114, 0xf>(v);
    v += dpp_f64<0x118, 0xf>(v);
    v += dpp_f64<0x142, 0xa>(v);
    v += dpp_f64<0x143, 0xc>(v);
    return v;
}

// Per-workgroup reduction of NSLOT per-lane FP64 sums (fixed DPP pattern +
// LDS) into partials[logical * NSLOT + k]. A separate one-workgroup launch
// folds the partials in logical-workgroup (= sample) order. Both trees are
// fixed, so the result is identical on every run. (No cross-workgroup
// handshake inside a pass: on a multi-XCD device an agent-scope fence in every
// workgroup writes back and invalidates its XCD's L2 and costs more than the
// extra launch -- measured on the record-driven derivative pass: a last-
// workgroup fold took it from 13 to 38 us.)
template <int NSLOT>
__device__ __forceinline__ void block_reduce_store(double (&acc)[NSLOT], double *partials,
                                                   unsigned logical) {
    __shared__ double red[kBlock / 64][NSLOT];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NSLOT; ++k) {
        const double v = wave_total_f64(acc[k]);
        if (lane == 63) red[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < NSLOT) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) t += red[w][threadIdx.x];
        partials[static_cast<size_t>(logical) * NSLOT + threadIdx.x] = t;
    }
}

// out[k] = sum over G partials of slot k: lane t sums rows t, t+256, ... (all
// loads issued first), then a fixed DPP + LDS tree. One workgroup.
template <int NSLOT>
__device__ __forceinline__ void fold_rows(const double *__restrict__ partials, unsigned G,
                                          double *__restrict__ out) {
    __shared__ double fold[NSLOT][kBlock / 64];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    double t[NSLOT];
#pragma unroll
    for (int k = 0; k < NSLOT; ++k) t[k] = 0.0;
    // four rows per lane per round, all loaded before any is added
    constexpr int kRows = 4;
    for (unsigned b0 = 0; b0 < G; b0 += kBlock * kRows) {
        double v[kRows][NSLOT];
#pragma unroll
        for (int r = 0; r < kRows; ++r) {
            const unsigned row = b0 + r * kBlock + threadIdx.x;
            const double *p = partials + static_cast<size_t>(row < G ? row : 0u) * NSLOT;
#pragma unroll
            for (int k = 0; k < NSLOT; ++k) {
                const double x = p[k];
                v[r][k] = row < G ? x : 0.0;
            }
        }
#pragma unroll
        for (int r = 0; r < kRows; ++r)
#pragma unroll
            for (int k = 0; k < NSLOT; ++k) t[k] += v[r][k];
    }
#pragma unroll
    for (int k = 0; k < NSLOT; ++k) {
        const double x = wave_total_f64(t[k]);
        if (lane == 63) fold[k][wave] = x;
    }
    __syncthreads();
    if (threadIdx.x < NSLOT) {
        double x = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) x += fold[threadIdx.x][w];
        out[threadIdx.x] = x;
    }
}

// The fold launch, which ends an evaluation: once the sums are out, lane 0
// (whose wave wrote them) publishes `seq` to the host-visible `done` after a
// system-scope fence, so the host can wait on that word (wait_done) instead of
// on the stream.
template <int NSLOT>
__global__ __launch_bounds__(kBlock) void fold_partials(const double *__restrict__ partials,
                                                        unsigned G, double *__restrict__ out,
                                                        unsigned long long *done,
                                                        unsigned long long seq) {
    fold_rows<NSLOT>(partials, G, out);
    if (done && threadIdx.x == 0) {
        __threadfence_system();
        *reinterpret_cast<volatile unsigned long long *>(done) = seq;
    }
}


// Samples per lane per iteration: all U samples' loads are issued before any of
// their values is used, so each lane keeps U independent gathers in flight. A
// lane visits samples begin + lane, + kBlock, + 2 kBlock, ... for any U, so U
// changes no sum. One is fastest on gfx950 with the 8 corner gathers of a
// sample already in flight together (measured: U = 2 and 4 are 3 us slower on
// a 2M-sample value pass).
#ifndef SITKHIP_BATCH
#define SITKHIP_BATCH 1
#endif
constexpr int kU = SITKHIP_BATCH;

// Phase A of the batched loop: which voxel each of the U samples is and
// whether it exists (s < end and, for Mattes, not an excluded fixed bin).
struct Batch {
    uint32_t flat[kU];
    int fb[kU];
    bool live[kU];
};

// Phase B: located samples -- the in-cell fraction in the interpolation type R
// and the 8 corner values. Loads are unconditional with clamped addresses, so
// every lane issues the same loads; the fixed point is recomputed later rather
// than held across the gathers.
template <typename R>
struct Located {
    R fr[kU][3];
    int base[kU][3];  // the cell (the B-spline interpolant's taps)
    float v[kU][8];
    bool ok[kU];
    bool unsure[kU];  // DEFER: the composed map could not decide; not located
};

// c = A idx + b for the fixed grid F, the point-map stages (nstage x 12) and
// the moving geometry M, composed in FP64 on the host (180 ns per map in
// x87 extended precision was most of a batched launch's host time; FP64
// keeps c within 5e-13 voxel of the exact composition over a 256^3 grid,
// against the 1e-6 voxel guard the kernels apply).
inline IndexMap make_index_map(const FixedGrid &F, const MovingVol &M, const double *stages,
                               int nstage) {
    double Mt[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, t[3] = {0, 0, 0};
    for (int s = 0; s < nstage; ++s) {
        const double *m = stages + 12 * s;
        double nM[3][3], nt[3];
        for (int r = 0; r < 3; ++r) {
            nt[r] = m[9 + r];
            for (int c = 0; c < 3; ++c) {
                nM[r][c] = 0;
                for (int k = 0; k < 3; ++k) nM[r][c] += m[r * 3 + k] * Mt[k][c];
                nt[r] += m[r * 3 + c] * t[c];
            }
        }
        for (int r = 0; r < 3; ++r) {
            t[r] = nt[r];
            for (int c = 0; c < 3; ++c) Mt[r][c] = nM[r][c];
        }
    }
    // q = Mt o_f + t - o_m ; A = P Mt I2P ; b = P q
    double q[3], MI[3][3];
    for (int r = 0; r < 3; ++r) {
        q[r] = t[r] - M.origin[r];
        for (int k = 0; k < 3; ++k) q[r] += Mt[r][k] * F.origin[k];
        for (int c = 0; c < 3; ++c) {
            MI[r][c] = 0;
            for (int k = 0; k < 3; ++k) MI[r][c] += Mt[r][k] * F.i2p[k * 3 + c];
        }
    }
    IndexMap X;
    for (int r = 0; r < 3; ++r) {
        double br = 0;
        for (int k = 0; k < 3; ++k) br += M.p2i[r * 3 + k] * q[k];
        X.b[r] = br;
        for (int c = 0; c < 3; ++c) {
            double a = 0;
            for (int k = 0; k < 3; ++k) a += M.p2i[r * 3 + k] * MI[k][c];
            X.A[r * 3 + c] = a;
        }
    }
    return X;
}

// The point the moment sums weight by: FP64 path, the fixed physical point
// (the reference's); FP32 path, the fixed index itself (exact small integers,
// no geometry in registers) -- its moments B'[d][c] become the physical ones
// on the host, B[d][e] = origin_e A_d + sum_c I2P[e][c] B'[d][c]
// (moments_to_physical).
// PTS: the metric's samples are explicit points (the physical point is
// itself the moment point: the grid is the identity); a separate
// instantiation, so the grid kernels carry no such branch.
template <typename R, bool PTS = false>
__device__ __forceinline__ void moment_point(const FixedGrid &F, const ChainConst *__restrict__ C,
                                             uint32_t flat, double x[3]) {
    if constexpr (PTS) {
        sample_point<true>(F, *C, flat, x);
    } else if constexpr (sizeof(R) == 4) {
        uint32_t i0, i1, i2;
        split_index(F, flat, i0, i1, i2);
        x[0] = static_cast<double>(i0);
        x[1] = static_cast<double>(i1);
        x[2] = static_cast<double>(i2);
    } else {
        fixed_point(F, *C, flat, x);
    }
}

// Index-space first/second moments -> physical: `a` the 3 first moments, `b`
// the 3 x 3 second moments [d][c] (rewritten in place as [d][e]).
inline void moments_to_physical(const FixedGrid &F, const double *a, double *b) {
    double out[9];
    for (int d = 0; d < 3; ++d)
        for (int e = 0; e < 3; ++e) {
            double acc = F.origin[e] * a[d];
            for (int c = 0; c < 3; ++c) acc += F.i2p[e * 3 + c] * b[d * 3 + c];
            out[d * 3 + e] = acc;
        }
    for (int i = 0; i < 9; ++i) b[i] = out[i];
}

// Whether a point map is affine (the composed index map applies to it).
template <typename Map>
struct MapTraits {
    static constexpr bool affine = true;
    static constexpr bool points = false;
};

// The stages of a metric over explicit points (and their LDS view): the
// samples have no grid index to compose the map with, so every sample takes
// the exact chain.
template <>
struct MapTraits<PointStages> {
    static constexpr bool affine = false;
    static constexpr bool points = true;
};
template <>
struct MapTraits<PointStageView> {
    static constexpr bool affine = false;
    static constexpr bool points = true;
};

// R = float: the composed index map decides, and the reference's FP64 chain
// runs only for samples within kGuard of a decision boundary. R = double: the
// chain always (the F64 path follows the reference term for term).
// DEFER (FP32, affine maps): a sample the composed map cannot decide is not
// located here but flagged `unsure`, for the caller to run the chain later
// (deferred_sample), which keeps the chain's registers out of the loop.
template <typename R, typename Map, bool DEFER = false, bool BSP = false>
__device__ __forceinline__ void locate_batch(const FixedGrid &F, const MovingVol &M,
                                             const Map &S, const IndexMap &X,
                                             const ChainConst *__restrict__ C, const Batch &b,
                                             Located<R> &L) {
    Cell cell[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
        bool done = false;
        L.unsure[u] = false;
        if constexpr (sizeof(R) == 4 && MapTraits<Map>::affine) {
            uint32_t i0, i1, i2;
            split_index(F, b.flat[u], i0, i1, i2);
            bool ok = false;
            if (locate_composed(M, X, i0, i1, i2, ok, cell[u])) {
                L.ok[u] = b.live[u] && ok;
                done = true;
            } else if constexpr (DEFER) {
                L.unsure[u] = b.live[u];
                L.ok[u] = false;
                done = true;
            }
        }
        if (!done) {
            // The exact chain's geometry is read through a pointer held in
            // vector registers: vector loads, so the rarely taken chain does
            // not claim the scalar registers the fast path keeps its
            // constants in.
            const ChainConst *Cv = C;
            asm volatile("" : "+v"(Cv));
            double x[3], p[3];
            sample_point<MapTraits<Map>::points>(F, *Cv, b.flat[u], x);
            map_point(S, x, p);
            L.ok[u] = b.live[u] && locate(M, *Cv, p, cell[u]);
        }
        if (!L.ok[u]) {
            cell[u].base[0] = cell[u].base[1] = cell[u].base[2] = 0;
            cell[u].frac[0] = cell[u].frac[1] = cell[u].frac[2] = 0.0;
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            L.fr[u][d] = static_cast<R>(cell[u].frac[d]);
            L.base[u][d] = cell[u].base[d];
        }
    }
    // (the cubic interpolant reads its 64 taps when it is evaluated)
    if constexpr (!BSP) {
#pragma unroll
        for (int u = 0; u < kU; ++u) corners(M, cell[u], L.v[u]);
    }
}

// ------------------------------------------------ per-range sample loops
// Each processes the samples [begin, end) of one sample set with the kBlock
// lanes of the calling workgroup, kU samples per lane per step (every lane
// issues its kU gathers before using any of them). `Map` is the point map:
// Stages (1..8 stages, kernel argument) or OneStage (a single stage).

// Mean squares: [sq, S0[3], S1[3][3], count] with diff = M - F
// (mean_squares.rs:101-145; contraction registration/cuda.rs:381-412), or
// [sq, count] without MOMENTS.
template <typename R, bool MOMENTS, typename Map, bool BSP = false>
__device__ __forceinline__ void ms_range(const FixedGrid &F, const MovingVol &M, const Map &S,
                                         const IndexMap &X, const ChainConst *__restrict__ C,
                                         const SampleSet &SS, const float *__restrict__ fval,
                                         long long begin, long long end,
                                         double (&acc)[MOMENTS ? 14 : 2]) {
    for (long long s0 = begin + threadIdx.x; s0 < end; s0 += kBlock * kU) {
        Batch b;
        float fv[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const long long s = s0 + static_cast<long long>(u) * kBlock;
            b.live[u] = s < end;
            const long long sc = b.live[u] ? s : begin;
            b.flat[u] = sample_flat(SS, sc);
            fv[u] = fval[sc];
        }
        Located<R> L;
        locate_batch<R, Map, false, BSP>(F, M, S, X, C, b, L);
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (!L.ok[u]) continue;
            R mv, gi[3], g[3];
            Interp<R, BSP>::template eval<MOMENTS>(M, L.base[u], L.fr[u], L.v[u], mv, gi);
            const double diff = static_cast<double>(mv) - static_cast<double>(fv[u]);
            acc[0] = fma(diff, diff, acc[0]);
            if constexpr (MOMENTS) {
                double x[3];
                moment_point<R, MapTraits<Map>::points>(F, C, b.flat[u], x);
                physical_gradient<R>(M, gi, g);
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double dg = diff * static_cast<double>(g[d]);
                    acc[1 + d] += dg;
#pragma unroll
                    for (int e = 0; e < 3; ++e)
                        acc[4 + d * 3 + e] = fma(dg, x[e], acc[4 + d * 3 + e]);
                }
                acc[13] += 1.0;
            } else {
                acc[1] += 1.0;
            }
        }
    }
}

// What the value pass leaves per sample for a derivative pass that follows it
// at the same transform: the moving value and its index-space gradient, or a
// NaN value where the sample is dropped (outside the moving buffer or mask,
// or outside the Parzen range). Written for every sample of the set.
template <typename R>
struct alignas(4 * sizeof(R)) SampleRec {
    R mv;
    R gi[3];
};

// One located sample of the Mattes value pass (mattes.rs:442-486): the four
// cubic B-spline taps of the moving Parzen window into row `fixed bin` of
// `joint` (64-bit fixed point), one count into marg[fixed bin], and with REC
// the sample's record (the value is the same: the interpolation computes it
// the same way with or without the gradient).
template <typename R, bool REC, bool BSP = false>
__device__ __forceinline__ void hist_sample(const MattesParams &P, unsigned long long *joint,
                                            unsigned long long *marg, SampleRec<R> *__restrict__ rec,
                                            long long s, bool ok, int fb, const R fr[3],
                                            const float v[8], const MovingVol *M = nullptr,
                                            const int *base = nullptr) {
    R mv = R(0), gi[3] = {R(0), R(0), R(0)};
    bool use = ok;
    if (use) {
        if constexpr (BSP) CubicBSpline<R>::template eval<REC>(*M, base, fr, mv, gi);
        else Trilinear<R>::template eval<REC>(fr, v, mv, gi);
        use = Parzen<R>::in_range(P, mv);
    }
    if (REC) {
        SampleRec<R> r;
        r.mv = use ? mv : static_cast<R>(NAN);
        r.gi[0] = gi[0];
        r.gi[1] = gi[1];
        r.gi[2] = gi[2];
        rec[s] = r;
    }
    if (!use) return;
    R arg[4], wv[4];
    bool mid;
    const int m0 = Parzen<R>::taps(P, mv, arg, mid);
    parzen_values<R>(arg, mid, wv);
    unsigned long long *row = joint + fb * P.bins + m0;
#pragma unroll
    for (int k = 0; k < 4; ++k) atomicAdd(row + k, to_fixed(wv[k]));
    atomicAdd(marg + fb, 1ull);
}

// The value pass over samples begin + lane + k * step (< end). DEFER (FP32,
// affine maps): a sample the composed index map cannot decide is appended to
// the workgroup's list `defer` (count in *defer_n, LDS) instead, for
// deferred_sample after the loop.
template <typename R, bool REC, typename Map, bool DEFER = false, bool BSP = false>
__device__ __forceinline__ void hist_range(const FixedGrid &F, const MovingVol &M, const Map &S,
                                           const IndexMap &X, const ChainConst *__restrict__ C,
                                           const SampleSet &SS, const uint8_t *__restrict__ fbin,
                                           const MattesParams &P, unsigned long long *joint,
                                           unsigned long long *marg, SampleRec<R> *__restrict__ rec,
                                           long long begin, long long end, long long step = kBlock,
                                           uint32_t *__restrict__ defer = nullptr,
                                           unsigned *defer_n = nullptr) {
    for (long long s0 = begin + threadIdx.x; s0 < end; s0 += step * kU) {
        Batch b;
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const long long s = s0 + static_cast<long long>(u) * step;
            const bool in = s < end;
            const long long sc = in ? s : begin;
            b.fb[u] = fbin[sc];
            b.flat[u] = sample_flat(SS, sc);
            b.live[u] = in && b.fb[u] != kNoBin;
        }
        Located<R> L;
        locate_batch<R, Map, DEFER, BSP>(F, M, S, X, C, b, L);
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const long long s = s0 + static_cast<long long>(u) * step;
            if (s >= end) continue;
            if (DEFER && L.unsure[u]) {
                defer[atomicAdd(defer_n, 1u)] = static_cast<uint32_t>(s);
                continue;
            }
            hist_sample<R, REC, BSP>(P, joint, marg, rec, s, L.ok[u], b.fb[u], L.fr[u], L.v[u], &M,
                                     L.base[u]);
        }
    }
}

// A deferred sample of the value pass: the reference's FP64 chain (fixed
// point, point map, continuous index), then the same accumulation. Runs after
// the loop, where the loop's registers are free.
template <typename R, bool REC, typename Map, bool BSP = false>
__device__ __forceinline__ void deferred_sample(const FixedGrid &F, const MovingVol &M, const Map &S,
                                                const ChainConst *__restrict__ C,
                                                const SampleSet &SS, const uint8_t *__restrict__ fbin,
                                                const MattesParams &P, unsigned long long *joint,
                                                unsigned long long *marg,
                                                SampleRec<R> *__restrict__ rec, long long s) {
    const int fb = fbin[s];
    double x[3], p[3];
    fixed_point(F, *C, sample_flat(SS, s), x);
    map_point(S, x, p);
    Cell cell;
    const bool ok = locate(M, *C, p, cell);
    if (!ok) {
        cell.base[0] = cell.base[1] = cell.base[2] = 0;
        cell.frac[0] = cell.frac[1] = cell.frac[2] = 0.0;
    }
    R fr[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) fr[d] = static_cast<R>(cell.frac[d]);
    float v[8];
    if constexpr (!BSP) corners(M, cell, v);
    hist_sample<R, REC, BSP>(P, joint, marg, rec, s, ok, fb, fr, v, &M, cell.base);
}

// sum_k row[k] * dv[k] over the four taps: FP64 in the reference's order,
// every product and sum rounded; FP32 as fused multiply-adds.
template <typename R>
__device__ __forceinline__ R tap_sum(const R *row, const R dv[4]) {
    R w = R(0);
#pragma unroll
    for (int k = 0; k < 4; ++k) w += row[k] * dv[k];
    return w;
}
template <>
__device__ __forceinline__ float tap_sum<float>(const float *row, const float dv[4]) {
    return fmaf(row[3], dv[3], fmaf(row[2], dv[2], fmaf(row[1], dv[1], row[0] * dv[0])));
}

// Mattes derivative pass: per valid sample w = sum_taps pRatio * B3'(arg)
// and the 12 moments A[d] += w dM_d, B[d][e] += w dM_d x_e
// (src/cuda/ops/mattes.rs:75-91).
template <typename R, typename Map, bool BSP = false>
__device__ __forceinline__ void deriv_range(const FixedGrid &F, const MovingVol &M, const Map &S,
                                            const IndexMap &X, const ChainConst *__restrict__ C,
                                           const SampleSet &SS, const uint8_t *__restrict__ fbin,
                                            const MattesParams &P, const R *pr, long long begin,
                                            long long end, double (&acc)[12]) {
    const int B = P.bins;
    for (long long s0 = begin + threadIdx.x; s0 < end; s0 += kBlock * kU) {
        Batch b;
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const long long s = s0 + static_cast<long long>(u) * kBlock;
            const bool in = s < end;
            const long long sc = in ? s : begin;
            b.fb[u] = fbin[sc];
            b.flat[u] = sample_flat(SS, sc);
            b.live[u] = in && b.fb[u] != kNoBin;
        }
        Located<R> L;
        locate_batch<R, Map, false, BSP>(F, M, S, X, C, b, L);
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (!L.ok[u]) continue;
            R mv, gi[3], g[3];
            Interp<R, BSP>::template eval<true>(M, L.base[u], L.fr[u], L.v[u], mv, gi);
            if (!Parzen<R>::in_range(P, mv)) continue;
            physical_gradient<R>(M, gi, g);
            R arg[4], dv[4];
            bool mid;
            const int m0 = Parzen<R>::taps(P, mv, arg, mid);
            parzen_derivs<R>(arg, mid, dv);
            const R *row = pr + b.fb[u] * B + m0;
            const R w = tap_sum<R>(row, dv);
            const double wd = static_cast<double>(w);
            double x[3];
            moment_point<R, MapTraits<Map>::points>(F, C, b.flat[u], x);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double wg = wd * static_cast<double>(g[d]);
                acc[d] += wg;
#pragma unroll
                for (int e = 0; e < 3; ++e) acc[3 + d * 3 + e] = fma(wg, x[e], acc[3 + d * 3 + e]);
            }
        }
    }
}

// The derivative pass from the value pass's records: the same samples, the
// same per-lane order and the same arithmetic as deriv_range, without the
// gather and interpolation (mattes.rs:75-91 as above). A lane visits samples
// begin + lane, + kBlock, + 2 kBlock, ... whatever the unroll, so kRecU
// samples' loads are issued at a time without changing the sums. `first`
// holds the first round's loads, issued by the caller.
#ifndef SITKHIP_RECU
#define SITKHIP_RECU 4
#endif
constexpr int kRecU = SITKHIP_RECU;

template <typename R>
struct RecRound {
    SampleRec<R> r[kRecU];
    uint32_t flat[kRecU];
    int fb[kRecU];
};

// Samples s0, s0 + kBlock, ...; those at or past `end` read sample `begin`
// instead (so begin < end is required) and are marked absent.
template <typename R>
__device__ __forceinline__ void load_rec_round(const SampleSet &SS, const uint8_t *__restrict__ fbin,
                                               const SampleRec<R> *__restrict__ rec, long long s0,
                                               long long begin, long long end, RecRound<R> &q) {
#pragma unroll
    for (int u = 0; u < kRecU; ++u) {
        const long long s = s0 + static_cast<long long>(u) * kBlock;
        const long long sc = s < end ? s : begin;
        q.fb[u] = s < end ? static_cast<int>(fbin[sc]) : static_cast<int>(kNoBin);
        q.r[u] = rec[sc];
        q.flat[u] = sample_flat(SS, sc);
    }
}

template <typename R>
__device__ __forceinline__ void deriv_records(const FixedGrid &F, const MovingVol &M,
                                              const ChainConst *__restrict__ C,
                                              const SampleSet &SS, const uint8_t *__restrict__ fbin,
                                              const MattesParams &P, const R *pr,
                                              const SampleRec<R> *__restrict__ rec,
                                              long long begin, long long end,
                                              const RecRound<R> &first, double (&acc)[12]) {
    const int B = P.bins;
    RecRound<R> q = first;
    for (long long s0 = begin + threadIdx.x; s0 < end; s0 += kBlock * kRecU) {
        if (s0 != begin + threadIdx.x) load_rec_round<R>(SS, fbin, rec, s0, begin, end, q);
#pragma unroll
        for (int u = 0; u < kRecU; ++u) {
            if (q.fb[u] == kNoBin || isnan(q.r[u].mv)) continue;
            R g[3];
            physical_gradient<R>(M, q.r[u].gi, g);
            R arg[4], dv[4];
            bool mid;
            const int m0 = Parzen<R>::taps(P, q.r[u].mv, arg, mid);
            parzen_derivs<R>(arg, mid, dv);
            const R *row = pr + q.fb[u] * B + m0;
            const R w = tap_sum<R>(row, dv);
            const double wd = static_cast<double>(w);
            double x[3];
            if (F.pts)
                moment_point<R, true>(F, C, q.flat[u], x);
            else
                moment_point<R>(F, C, q.flat[u], x);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double wg = wd * static_cast<double>(g[d]);
                acc[d] += wg;
#pragma unroll
                for (int e = 0; e < 3; ++e) acc[3 + d * 3 + e] = fma(wg, x[e], acc[3 + d * 3 + e]);
            }
        }
    }
}

// ------------------------------------------------------------------- Mattes
struct TailOut {
    double value;
    double valid;  // as double: the host reads one block of doubles
    double joint_sum;
    double pad;
};

// The Mattes tail (mattes.rs:273-328) on the finished integer histogram: -MI,
// valid, and the pRatio table (pRatio * n_factor; zero where the value sum
// skips the bin). The joint mass is summed exactly in integers and rounded
// once; the value terms are summed exactly as 128-bit fixed-point integers
// (2^-kValShift units) and rounded once, so the value does not depend on how
// the bins are split over lanes and workgroups: the one-workgroup tail of the
// batched engine and the split tail below give the same bits.
#ifndef SITKHIP_TAIL
#define SITKHIP_TAIL 512
#endif
constexpr int kTail = SITKHIP_TAIL;  // lanes of a tail workgroup
constexpr int kValShift = 100;
constexpr int kMaxTailSplit = 16;

__device__ __forceinline__ __int128 wave_sum_i128(__int128 v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long lo = __shfl_xor(static_cast<unsigned long long>(v), off, 64);
        const long long hi = __shfl_xor(static_cast<long long>(v >> 64), off, 64);
        v += static_cast<__int128>((static_cast<unsigned __int128>(static_cast<unsigned long long>(hi)) << 64) | lo);
    }
    return v;
}
// t * 2^kValShift truncated toward minus infinity; exact for |t| >= 2^-47.
// |t| < 2^8 here (a value term is jp * log(ratio) with jp <= 1 and the ratio
// within [eps, 1 / eps^2]), so the product fits in 108 bits.
__device__ __forceinline__ __int128 value_to_fixed(double t) {
    if (t == 0.0) return 0;
    int ex;
    const double f = frexp(t, &ex);  // t = f * 2^ex, 0.5 <= |f| < 1
    const long long mant = static_cast<long long>(ldexp(f, 53));  // exact
    const int sh = ex - 53 + kValShift;
    const __int128 q = static_cast<__int128>(mant);
    return sh >= 0 ? (q << sh) : (sh > -127 ? (q >> (-sh)) : (mant < 0 ? -1 : 0));
}

// Folds the per-XCD copies into s_hist (LDS) when given, else into copy 0 (one
// workgroup only). A lane's bins of all copies are loaded before any is summed;
// every load is unconditional (an index past the end reads element 0 and is
// dropped), so all of a lane's loads are in flight together. Each bin is read
// and written by one lane only, so the loop needs no barrier.
__device__ __forceinline__ void tail_fold(unsigned long long *g_joint, unsigned long long *g_marg,
                                          int B, unsigned long long *s_hist) {
    const int t = threadIdx.x, nbin = B * B, ntot = nbin + B;
    constexpr int kPer = 4;
    for (int i0 = t; i0 < ntot; i0 += kTail * kPer) {
        const unsigned long long *src[kPer];
        int stride[kPer];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int i = i0 + j * kTail;
            const bool marg = i >= nbin && i < ntot;
            src[j] = marg ? g_marg + (i - nbin) : g_joint + (i < nbin ? i : 0);
            stride[j] = marg ? B : nbin;
        }
        unsigned long long v[kPer] = {0, 0, 0, 0};
#pragma unroll
        for (int x = 0; x < kHistCopies; ++x)
#pragma unroll
            for (int j = 0; j < kPer; ++j) v[j] += src[j][x * stride[j]];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int i = i0 + j * kTail;
            if (i >= ntot) continue;
            if (s_hist) s_hist[i] = v[j];
            else if (i < nbin) g_joint[i] = v[j];
            else g_marg[i - nbin] = v[j];
        }
    }
}

// Whole-histogram statistics every tail lane needs, from the folded
// histogram: exact integer joint mass, valid count (fixed marginal) and
// moving-bin column sums, then the marginal pdfs and the fixed marginal's log.
struct TailStats {
    unsigned long long valid;
    double joint_sum, inv_sum, n_factor;
    bool degenerate;
};
struct TailShared {
    unsigned long long wu[kTail / 64][2];
    unsigned long long colsum[256];
    unsigned long long fcount[256];
    double mm[256], fm[256], logfm[256];
    long long part[kTail / 64][2];
};

// A 64-bit integer moved across lanes by DPP, as dpp_f64.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long v) {
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<unsigned>(v)), CTRL,
                                               ROW_MASK, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<unsigned>(v >> 32)),
                                               CTRL, ROW_MASK, 0xf, true);
    return (static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) | static_cast<unsigned>(lo);
}
// Wave64 integer sum (exact); the total is in lane 63.
__device__ __forceinline__ unsigned long long wave_total_u64(unsigned long long v) {
    v += dpp_u64<0x111, 0xf>(v);
    v += dpp_u64<0x112, 0xf>(v);
    v += dpp_u64<0x114, 0xf>(v);
    v += dpp_u64<0x118, 0xf>(v);
    v += dpp_u64<0x142, 0xa>(v);
    v += dpp_u64<0x143, 0xc>(v);
    return v;
}

// Column sums: wave w takes moving bins w, w + 16, ..., its lanes the fixed
// bins of a column, summed across the wave; the joint mass is the sum of the
// column sums and the valid count that of the fixed marginal (all exact
// integers, so the order does not matter).
__device__ __forceinline__ TailStats tail_stats(const unsigned long long *hj,
                                                const unsigned long long *hm, int B, double mbs,
                                                TailShared &sh) {
    constexpr int kWaves = kTail / 64;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    unsigned long long mass = 0;
    for (int m = wave; m < B; m += kWaves) {
        unsigned long long c = 0;
        for (int f = lane; f < B; f += 64) c += hj[f * B + m];
        c = wave_total_u64(c);
        if (lane == 63) sh.colsum[m] = c;
        mass += c;
    }
    unsigned long long cnt = 0;
    if (wave == kWaves - 1) {
        for (int f = lane; f < B; f += 64) cnt += hm[f];
        cnt = wave_total_u64(cnt);
    }
    if (lane == 63) {
        sh.wu[wave][0] = mass;
        sh.wu[wave][1] = cnt;
    }
    __syncthreads();
    unsigned long long tot = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) tot += sh.wu[w][0];
    const unsigned long long valid = sh.wu[kWaves - 1][1];
    TailStats st;
    st.valid = valid;
    st.joint_sum = static_cast<double>(tot) * kInvFix;
    st.degenerate = valid == 0 || st.joint_sum < DBL_EPSILON;
    st.inv_sum = st.degenerate ? 0.0 : 1.0 / st.joint_sum;
    st.n_factor = st.degenerate ? 0.0 : 1.0 / (mbs * static_cast<double>(valid));
    if (t < B) {
        const double fm = static_cast<double>(hm[t]) * st.inv_sum;
        sh.fm[t] = fm;
        sh.logfm[t] = fm > DBL_EPSILON ? log(fm) : 0.0;
        sh.mm[t] = static_cast<double>(sh.colsum[t]) * kInvFix * st.inv_sum;
    }
    __syncthreads();
    return st;
}

// The statistics of tail_stats from the marginal copies alone: the fixed
// marginal (valid count) and the moving-bin column sums (joint mass), each
// summed over the kHistCopies copies (exact integers: the same numbers as
// from the folded histogram).
__device__ __forceinline__ TailStats tail_stats_marginals(const unsigned long long *g_marg,
                                                          const unsigned long long *g_col, int B,
                                                          double mbs, TailShared &sh) {
    const int t = threadIdx.x;
    if (t < B) {
        unsigned long long fc = 0, cc = 0;
#pragma unroll
        for (int x = 0; x < kHistCopies; ++x) {
            fc += g_marg[x * B + t];
            cc += g_col[x * B + t];
        }
        sh.fcount[t] = fc;
        sh.colsum[t] = cc;
    }
    __syncthreads();
    // every wave adds the B entries across its lanes (exact integers)
    unsigned long long tot = 0, valid = 0;
    for (int i = t & 63; i < B; i += 64) {
        tot += sh.colsum[i];
        valid += sh.fcount[i];
    }
    tot = wave_total_u64(tot);
    valid = wave_total_u64(valid);
    tot = (static_cast<unsigned long long>(static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(tot >> 32), 63))) << 32) |
          static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<unsigned>(tot)), 63));
    valid = (static_cast<unsigned long long>(static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(valid >> 32), 63))) << 32) |
            static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<unsigned>(valid)), 63));
    TailStats st;
    st.valid = valid;
    st.joint_sum = static_cast<double>(tot) * kInvFix;
    st.degenerate = valid == 0 || st.joint_sum < DBL_EPSILON;
    st.inv_sum = st.degenerate ? 0.0 : 1.0 / st.joint_sum;
    st.n_factor = st.degenerate ? 0.0 : 1.0 / (mbs * static_cast<double>(valid));
    if (t < B) {
        const double fm = static_cast<double>(sh.fcount[t]) * st.inv_sum;
        sh.fm[t] = fm;
        sh.logfm[t] = fm > DBL_EPSILON ? log(fm) : 0.0;
        sh.mm[t] = static_cast<double>(sh.colsum[t]) * kInvFix * st.inv_sum;
    }
    __syncthreads();
    return st;
}

// pRatio of bins [i0, i1) (lane t: i0 + t, i0 + t + kTail, ...) and the
// workgroup's exact value sum over them (valid in lane 0 of wave 0 after
// the barrier inside).
// COPIES: hj holds the kHistCopies unfolded copies (B^2 apart), summed here
// bin by bin.
// `first`, when given, is the lane's first bin (i0 + lane) already summed by
// the caller.
template <bool COPIES = false>
__device__ __forceinline__ __int128 tail_bins(const unsigned long long *hj, int B, int i0, int i1,
                                              const TailStats &st, double *pratio, float *pratio_f,
                                              TailShared &sh, const unsigned long long *first = nullptr) {
    constexpr int kWaves = kTail / 64;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    __int128 part = 0;
    for (int i = i0 + t; i < i1; i += kTail) {
        double pr_out = 0.0;
        unsigned long long q = 0;
        if (first && i == i0 + t) {
            q = *first;
        } else if constexpr (COPIES) {
#pragma unroll
            for (int x = 0; x < kHistCopies; ++x) q += hj[static_cast<size_t>(x) * B * B + i];
        } else {
            q = hj[i];
        }
        if (q && !st.degenerate) {
            const int f = i / B, m = i - f * B;
            const double jp = static_cast<double>(q) * kInvFix * st.inv_sum;
            if (sh.fm[f] > DBL_EPSILON && sh.mm[m] > DBL_EPSILON && jp > DBL_EPSILON) {
                const double pr = log(jp / sh.mm[m]);
                part += value_to_fixed(jp * (pr - sh.logfm[f]));
                pr_out = pr * st.n_factor;
            }
        }
        pratio[i] = pr_out;
        pratio_f[i] = static_cast<float>(pr_out);
    }
    part = wave_sum_i128(part);
    if (lane == 0) {
        sh.part[wave][0] = static_cast<long long>(static_cast<unsigned long long>(part));
        sh.part[wave][1] = static_cast<long long>(part >> 64);
    }
    __syncthreads();
    __int128 sum = 0;
    if (t == 0)
#pragma unroll
        for (int w = 0; w < kWaves; ++w)
            sum += static_cast<__int128>((static_cast<unsigned __int128>(static_cast<unsigned long long>(sh.part[w][1])) << 64) |
                                         static_cast<unsigned long long>(sh.part[w][0]));
    return sum;
}

__device__ __forceinline__ __int128 tail_bins_copies(const unsigned long long *g_joint, int B, int i0,
                                                     int i1, const TailStats &st, double *pratio,
                                                     float *pratio_f, TailShared &sh,
                                                     const unsigned long long *first = nullptr) {
    return tail_bins<true>(g_joint, B, i0, i1, st, pratio, pratio_f, sh, first);
}
// The copies of the lane's first bin of [i0, i1) (zeros past i1), loaded
// only: the caller sums them after issuing its other loads.
__device__ __forceinline__ void first_bin_copies(const unsigned long long *g_joint, int B, int i0, int i1,
                                                 unsigned long long (&q)[kHistCopies]) {
    const int i = i0 + static_cast<int>(threadIdx.x);
#pragma unroll
    for (int x = 0; x < kHistCopies; ++x) q[x] = 0;
    if (i < i1) {
#pragma unroll
        for (int x = 0; x < kHistCopies; ++x) q[x] = g_joint[static_cast<size_t>(x) * B * B + i];
    }
}

__device__ __forceinline__ double tail_value(__int128 sum, bool degenerate) {
    return degenerate ? DBL_MAX : -ldexp(static_cast<double>(sum), -kValShift);
}
__device__ __forceinline__ void tail_finish(__int128 sum, const TailStats &st, TailOut *out) {
    out->value = tail_value(sum, st.degenerate);
    out->valid = static_cast<double>(st.valid);
    out->joint_sum = st.joint_sum;
}
// The value of a tail split over T workgroups, from their exact partial sums
// (part[2b], part[2b + 1]: low and high words) and the degenerate flag
// (part[2T]); one lane, in a later kernel.
__device__ __forceinline__ void finish_split_tail(const unsigned long long *part, int T,
                                                  TailOut *out) {
    __int128 total = 0;
    for (int b = 0; b < T; ++b)
        total += static_cast<__int128>((static_cast<unsigned __int128>(part[2 * b + 1]) << 64) | part[2 * b]);
    out->value = tail_value(total, part[2 * T] != 0);
}

// The whole tail in one workgroup of kTail lanes; clears the histogram copies
// for the next evaluation when asked.
__device__ __forceinline__ void mattes_tail_block(unsigned long long *g_joint, unsigned long long *g_marg, int B,
                                  double mbs, double *pratio, float *pratio_f, TailOut *out,
                                  unsigned long long *s_hist, bool clear = true) {
    __shared__ TailShared sh;
    const int t = threadIdx.x, nbin = B * B;
    tail_fold(g_joint, g_marg, B, s_hist);
    __syncthreads();
    const unsigned long long *hj = s_hist ? s_hist : g_joint;
    const unsigned long long *hm = s_hist ? s_hist + nbin : g_marg;
    const TailStats st = tail_stats(hj, hm, B, mbs, sh);
    const __int128 sum = tail_bins(hj, B, 0, nbin, st, pratio, pratio_f, sh);
    // every read of the histogram is done once all lanes passed the barrier
    // inside tail_bins: clear it for the next evaluation
    if (clear) {
        for (int i = t; i < kHistCopies * nbin; i += kTail) g_joint[i] = 0ull;
        for (int i = t; i < kHistCopies * B; i += kTail) g_marg[i] = 0ull;
    }
    if (t == 0) tail_finish(sum, st, out);
}
// ------------------------------------------------------------------ host side
struct HostResult {
    TailOut tail;
    double moments[14];
    unsigned long long done;  // the sequence number of the last finished evaluation
    double ranges[6];         // at creation: fixed (min, max, count), moving (min, max, count)
};

struct Metric {
    int kind = 0;
    int precision = 0;
    FixedGrid F{};
    MovingVol Mv{};
    SampleSet SS{};
    DeviceBuffer<ChainConst> chain;  // the exact chain's geometry (device copy)
    DeviceBuffer<float> moving;
    DeviceBuffer<uint8_t> mmask;
    DeviceBuffer<uint32_t> idx;
    DeviceBuffer<double> pts;     // explicit sample points (n x 3), else empty
    DeviceBuffer<float> fval;     // per-sample fixed values (MS, and the range/bins source)
    DeviceBuffer<uint8_t> fbin;   // per-sample fixed Parzen bins (Mattes)
    DeviceBuffer<double> partials;
    // Results land in pinned, device-mapped host memory: no copy per evaluation.
    HostResult *res = nullptr;        // host view
    HostResult *res_dev = nullptr;    // device view of the same memory
    // Mattes
    bool bins_set = false;
    sitkhip_mattes_bins bins{};
    MattesParams P{};
    DeviceBuffer<unsigned long long> joint, marg;
    DeviceBuffer<double> pratio;
    DeviceBuffer<float> pratio_f;
    DeviceBuffer<unsigned long long> tail_part;  // the split tail's partial value sums + flag
    DeviceBuffer<unsigned char> recs;            // per-sample records of the value pass
    DeviceBuffer<uint32_t> defer;                // the value pass's deferred-sample lists
    // Cubic B-spline interpolation of the moving image (metric_set_interpolator):
    // the coefficients Mv.coef / Mv.coef64 point into
    bool bspline = false;
    DeviceBuffer<float> coef32;
    DeviceBuffer<double> coef64;
    unsigned long long seq = 0;                  // evaluations issued (see HostResult::done)
    DeviceBuffer<double> range_part;             // value-range partials (creation only)
    bool ranges_pending = false;                 // ranges enqueued, not yet read on the host
    bool ranges_known = false;                   // fixed_range / moving_range hold them
    long long moving_voxels = 0;
    void *chain_host = nullptr;                  // pinned source of the chain's upload
    ~Metric() {
        pinned_slot_free(res);
        pinned_slot_free(chain_host);
    }
    double fixed_range[2] = {0, 0};
    double moving_range[2] = {0, 0};
    size_t nsamples = 0;
    size_t voxels_fixed = 0;
};

// The Mattes tail of metric m into `out` (one workgroup; metric.hip).
void launch_tail(Metric &m, TailOut *out, bool clear = true);

}  // namespace sitkhip
