// Resident mean-squares and Mattes mutual-information metrics.
//
// Reference (crates/sitk/):
//   host metrics   src/registration/metric.rs:1065-1223 (mean squares),
//                  src/registration/mattes.rs:152-328, 442-676 (Mattes)
//   device surface src/cuda/ops/mean_squares.rs:186-439 (ResidentMetric),
//                  src/cuda/ops/mattes.rs:500-934 (ResidentMattes),
//                  src/cuda/ops/resident.rs:420-773 (sample sets, partials)
//
// Design (one process, one GPU, everything resident in HBM):
//   * The sample set is built once per level: a grid (no list) or a sorted list
//     of flat fixed indices (uint32). A fixed mask is applied at build time, as
//     the host's FixedSamples does, so the kernels never see it.
//   * Per sample, the fixed side is pre-reduced to what the metric reads: the
//     float32 value (mean squares) or the FP64-computed Parzen bin as uint8
//     (Mattes), in sample order, so the per-iteration passes stream them.
//   * Mattes value pass: every workgroup scatters the four cubic-B-spline taps of
//     each sample into an LDS joint histogram of 64-bit fixed-point integers
//     (2^-32 units) and integer marginal counts, then folds its non-zero bins
//     into one global histogram with 64-bit integer atomics. Integer addition is
//     associative, so the histogram is identical on every run and launch shape.
//   * Mattes tail: one workgroup turns the integer histogram into -MI and the
//     pRatio table in FP64 on the device, and clears the histogram for the next
//     evaluation.
//   * Mattes derivative pass: pRatio in LDS; each sample's weight
//     w = sum_taps pRatio * B3'(arg) scales 12 FP64 moments A[d] = sum w dM_d,
//     B[d][e] = sum w dM_d x_e (the moment form of src/cuda/ops/mattes.rs:75-91).
//   * Mean squares: 14 FP64 moments in one pass (mean_squares.rs:101-145).
//   * FP64 moments are reduced per lane, per wavefront, per workgroup into a
//     fixed number of partials, then folded in workgroup order by one
//     workgroup: a fixed reduction tree, so results are identical run to run.
//   * Workgroup b handles a contiguous slice of the samples; slices are
//     assigned so that the workgroups one XCD runs (b mod 8) hold adjacent
//     slices and share that XCD's L2 for the moving-volume gathers.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "image.hpp"
#include "metric.hpp"
#include "sampler.hpp"
#include "metric_core.hpp"

namespace sitkhip {

// Occupancy requests (waves per SIMD) for the two sample passes; 0: none.
#ifndef SITKHIP_HIST_WAVES
#define SITKHIP_HIST_WAVES 0
#endif
#ifndef SITKHIP_DERIV_WAVES
#define SITKHIP_DERIV_WAVES 0
#endif
#if SITKHIP_HIST_WAVES > 0
#define SITKHIP_HIST_OCC __attribute__((amdgpu_waves_per_eu(SITKHIP_HIST_WAVES)))
#else
#define SITKHIP_HIST_OCC
#endif
#if SITKHIP_DERIV_WAVES > 0
#define SITKHIP_DERIV_OCC __attribute__((amdgpu_waves_per_eu(SITKHIP_DERIV_WAVES)))
#else
#define SITKHIP_DERIV_OCC
#endif

// -------------------------------------------------------------- mean squares
template <typename R, bool MOMENTS, typename Map, bool BSP = false>
__global__ __launch_bounds__(kBlock) void ms_pass(FixedGrid F, MovingVol M, Map S, IndexMap X,
                                                  const ChainConst *__restrict__ C, SampleSet SS,
                                                  const float *__restrict__ fval,
                                                  double *__restrict__ partials) {
    constexpr int NS = MOMENTS ? 14 : 2;
    double acc[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) acc[k] = 0.0;
    __shared__ double s_stage[kMaxStages * 12];
    const auto V = stage_to_lds(S, s_stage);
    long long begin, end;
    unsigned logical;
    block_slice(SS.n, begin, end, logical);
    ms_range<R, MOMENTS, std::decay_t<decltype(V)>, BSP>(F, M, V, X, C, SS, fval, begin, end, acc);
    block_reduce_store<NS>(acc, partials, logical);
}

// Pass 1: joint histogram (fixed point) + fixed-marginal counts; with REC
// also the per-sample records for the derivative pass that follows.
template <typename R, bool LDS, bool REC, typename Map, bool BSP = false>
__global__ __launch_bounds__(kHistBlock) SITKHIP_HIST_OCC void mattes_hist(FixedGrid F, MovingVol M, Map S,
                                                      IndexMap X, const ChainConst *__restrict__ C,
                                                      SampleSet SS, const uint8_t *__restrict__ fbin,
                                                      MattesParams P,
                                                      unsigned long long *__restrict__ g_joint,
                                                      unsigned long long *__restrict__ g_marg,
                                                      unsigned long long *__restrict__ g_col,
                                                      SampleRec<R> *__restrict__ rec,
                                                      uint32_t *__restrict__ defer,
                                                      unsigned defer_cap) {
    extern __shared__ unsigned long long s_hist[];  // bins^2 joint, then bins marginal
    const int B = P.bins;
    const int nb = B * B + B;
    const unsigned copy = blockIdx.x % kHistCopies;
    g_joint += static_cast<size_t>(copy) * B * B;
    g_marg += copy * B;
    unsigned long long *joint = LDS ? s_hist : g_joint;
    unsigned long long *marg = LDS ? s_hist + B * B : g_marg;
    if (LDS) {
        for (int i = threadIdx.x; i < nb; i += kHistBlock) s_hist[i] = 0ull;
        __syncthreads();
    }
    // FP32 with an affine map: samples the composed index map cannot decide
    // are listed (this workgroup's slice of `defer`) and run through the FP64
    // chain after the loop
    constexpr bool DEFER = sizeof(R) == 4 && MapTraits<Map>::affine;
    __shared__ unsigned s_defer_n;
    if (threadIdx.x == 0) s_defer_n = 0u;
    __shared__ double s_stage[kMaxStages * 12];
    const auto V = stage_to_lds(S, s_stage);  // ends with a barrier
    long long first, step, end;
    xcd_walk<kHistBlock>(SS.n, first, step, end);
    uint32_t *dq = DEFER ? defer + static_cast<size_t>(blockIdx.x) * defer_cap : nullptr;
    hist_range<R, REC, std::decay_t<decltype(V)>, DEFER, BSP>(F, M, V, X, C, SS, fbin, P, joint, marg, rec, first, end, step, dq,
                                         &s_defer_n);
    if (DEFER) {
        __syncthreads();
        const unsigned nq = s_defer_n;
        for (unsigned i = threadIdx.x; i < nq; i += kHistBlock)
            deferred_sample<R, REC, std::decay_t<decltype(V)>, BSP>(F, M, V, C, SS, fbin, P, joint, marg, rec,
                                                                     dq[i]);
    }
    if (LDS) {
        __syncthreads();
        // The workgroups of one XCD finish together; each starts its flush at
        // its own bin so that they do not queue on the same addresses.
        const int rot = static_cast<int>((blockIdx.x / kHistCopies) * 193u % static_cast<unsigned>(nb));
        for (int k = threadIdx.x; k < nb; k += kHistBlock) {
            const int i = k + rot < nb ? k + rot : k + rot - nb;
            const unsigned long long v = s_hist[i];
            if (v) atomicAdd(i < B * B ? g_joint + i : g_marg + (i - B * B), v);
        }
        // for the split tail: this workgroup's moving-bin column sums into
        // the XCD's copy (integers: the same totals as summing the folded
        // histogram's columns)
        if (g_col) {
            for (int m = threadIdx.x; m < B; m += kHistBlock) {
                unsigned long long c = 0;
                for (int f = 0; f < B; ++f) c += s_hist[f * B + m];
                if (c) atomicAdd(g_col + copy * B + m, c);
            }
        }
    }
}

// The tail in one workgroup (value-only evaluations, and bin counts whose
// histogram does not fit in LDS: then it reads global memory and folds the
// copies into copy 0).
template <bool STAGE>
__global__ __launch_bounds__(kTail) void mattes_tail_kernel(unsigned long long *g_joint,
                                                             unsigned long long *g_marg, int B,
                                                             double mbs, double *pratio,
                                                             float *pratio_f, TailOut *out,
                                                             int clear) {
    extern __shared__ unsigned long long s_stage[];
    mattes_tail_block(g_joint, g_marg, B, mbs, pratio, pratio_f, out, STAGE ? s_stage : nullptr,
                      clear != 0);
}

// The tail of a value+derivative evaluation split over T = gridDim.x
// workgroups. The statistics come from the marginals alone -- the fixed
// marginal and the moving-bin column sums the value pass added per XCD copy
// (joint mass = sum of the column sums, valid = sum of the fixed marginal) --
// so a workgroup folds the histogram copies only for its own share of the
// bins, then computes their pRatio and its exact partial value sum
// (part[2b], part[2b + 1]); workgroup 0 writes the valid count and joint mass,
// and part[2T] = 1 for a degenerate histogram. The derivative pass that
// follows adds the partials (finish_split_tail) and clears the copies.
__global__ __launch_bounds__(kTail) void mattes_tail_split(const unsigned long long *g_joint,
                                                            const unsigned long long *g_marg,
                                                            const unsigned long long *g_col, int B,
                                                            double mbs, double *pratio,
                                                            float *pratio_f, TailOut *out,
                                                            unsigned long long *part) {
    __shared__ TailShared sh;
    const int t = threadIdx.x, nbin = B * B, T = static_cast<int>(gridDim.x);
    const int per = (nbin + T - 1) / T;
    const int i0 = static_cast<int>(blockIdx.x) * per;
    const int i1 = i0 + per < nbin ? i0 + per : nbin;
    // the lane's first bin is loaded with the marginals: one round trip to
    // memory instead of two
    unsigned long long qx[kHistCopies];
    first_bin_copies(g_joint, B, i0, i1, qx);
    const TailStats st = tail_stats_marginals(g_marg, g_col, B, mbs, sh);
    unsigned long long q0 = 0;
#pragma unroll
    for (int x = 0; x < kHistCopies; ++x) q0 += qx[x];
    const __int128 sum = tail_bins_copies(g_joint, B, i0, i1, st, pratio, pratio_f, sh, &q0);
    if (t == 0) {
        part[2 * blockIdx.x] = static_cast<unsigned long long>(sum);
        part[2 * blockIdx.x + 1] = static_cast<unsigned long long>(sum >> 64);
        if (blockIdx.x == 0) {
            part[2 * T] = st.degenerate ? 1ull : 0ull;
            out->valid = static_cast<double>(st.valid);
            out->joint_sum = st.joint_sum;
        }
    }
}

// After the split tail of a value-only evaluation: the value from the
// workgroups' exact partial sums (one lane), and the histogram copies, the
// marginal copies and the column-sum copies cleared for the next evaluation.
__global__ __launch_bounds__(256) void mattes_tail_finish(const unsigned long long *__restrict__ part,
                                                          int T, TailOut *out,
                                                          unsigned long long *__restrict__ joint,
                                                          unsigned n_joint,
                                                          unsigned long long *__restrict__ marg,
                                                          unsigned n_marg) {
    if (blockIdx.x == 0 && threadIdx.x == 0) finish_split_tail(part, T, out);
    const unsigned stride = gridDim.x * 256u;
    for (unsigned i = blockIdx.x * 256u + threadIdx.x; i < n_joint; i += stride) joint[i] = 0ull;
    for (unsigned i = blockIdx.x * 256u + threadIdx.x; i < n_marg; i += stride) marg[i] = 0ull;
}

// Pass 2: the 12 pRatio-weighted moments, folded by the last workgroup.
template <typename R, bool LDS, typename Map, bool BSP = false>
__global__ __launch_bounds__(kBlock) void mattes_deriv(FixedGrid F, MovingVol M, Map S,
                                                       IndexMap X, const ChainConst *__restrict__ C,
                                                       SampleSet SS, const uint8_t *__restrict__ fbin,
                                                       MattesParams P,
                                                       const R *__restrict__ g_pratio,
                                                       double *__restrict__ partials,
                                                       unsigned long long *__restrict__ clr,
                                                       unsigned clr_n,
                                                       unsigned long long *__restrict__ clr2,
                                                       unsigned clr2_n) {
    extern __shared__ unsigned char s_raw[];
    // the value pass's histogram copies, read by the tail, are cleared here
    // for the next evaluation (every workgroup a slice)
    for (unsigned i = blockIdx.x * kBlock + threadIdx.x; i < clr_n; i += gridDim.x * kBlock)
        clr[i] = 0ull;
    for (unsigned i = blockIdx.x * kBlock + threadIdx.x; i < clr2_n; i += gridDim.x * kBlock)
        clr2[i] = 0ull;
    R *s_pr = reinterpret_cast<R *>(s_raw);
    const int B = P.bins;
    const R *pr = LDS ? s_pr : g_pratio;
    if (LDS) {
        for (int i = threadIdx.x; i < B * B; i += kBlock) s_pr[i] = g_pratio[i];
        __syncthreads();
    }
    double acc[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = 0.0;
    __shared__ double s_stage[kMaxStages * 12];
    const auto V = stage_to_lds(S, s_stage);
    long long begin, end;
    unsigned logical;
    block_slice(SS.n, begin, end, logical);
    deriv_range<R, std::decay_t<decltype(V)>, BSP>(F, M, V, X, C, SS, fbin, P, pr, begin, end, acc);
    block_reduce_store<12>(acc, partials, logical);
}

// Pass 2 from the value pass's records (same transform): the same moments
// as mattes_deriv, in the same order. The pRatio table is read from global
// memory: a few KB read four times per sample stays in the vector cache, and
// not staging it spares the workgroup a load round trip and a barrier.
template <typename R>
__global__ __launch_bounds__(kBlock) SITKHIP_DERIV_OCC void mattes_deriv_rec(FixedGrid F, MovingVol M,
                                                           const ChainConst *__restrict__ C,
                                                           SampleSet SS, const uint8_t *__restrict__ fbin,
                                                           MattesParams P,
                                                           const R *__restrict__ pratio,
                                                           const SampleRec<R> *__restrict__ rec,
                                                           double *__restrict__ partials,
                                                           unsigned long long *__restrict__ clr,
                                                           unsigned clr_n,
                                                           unsigned long long *__restrict__ clr2,
                                                           unsigned clr2_n,
                                                           const unsigned long long *__restrict__ tail_part,
                                                           int tail_T, TailOut *tail_out) {
    long long begin, end;
    unsigned logical;
    block_slice(SS.n, begin, end, logical);
    // (a workgroup whose slice is empty -- begin >= end when the slices
    // overrun n -- loads nothing: begin is not a sample then)
    RecRound<R> first;
    if (begin < end) load_rec_round<R>(SS, fbin, rec, begin + threadIdx.x, begin, end, first);
    if (tail_T > 0 && blockIdx.x == 0 && threadIdx.x == 0) finish_split_tail(tail_part, tail_T, tail_out);
    // the value pass's histogram copies, read by the tail, are cleared here
    // for the next evaluation (every workgroup a slice)
    for (unsigned i = blockIdx.x * kBlock + threadIdx.x; i < clr_n; i += gridDim.x * kBlock)
        clr[i] = 0ull;
    for (unsigned i = blockIdx.x * kBlock + threadIdx.x; i < clr2_n; i += gridDim.x * kBlock)
        clr2[i] = 0ull;
    double acc[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = 0.0;
    deriv_records<R>(F, M, C, SS, fbin, P, pratio, rec, begin, end, first, acc);
    block_reduce_store<12>(acc, partials, logical);
}

// Fixed Parzen bin per sample, FP64 (mattes.rs:193-206 with the fixed axis'
// geometry): idx = trunc(v / binSize - normalizedMin) clamped to [2, B-3].
__global__ void fixed_bins(SampleSet SS, const float *__restrict__ fval_per_sample, int B,
                           double fbs, double fnmin, uint8_t *__restrict__ out) {
    size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t s = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         s < static_cast<size_t>(SS.n); s += stride) {
        const double v = static_cast<double>(fval_per_sample[s]);
        const double term = v / fbs - fnmin;
        long long idx = static_cast<long long>(term);
        if (idx < 2) idx = 2;
        else if (idx > B - 3) idx = B - 3;
        out[s] = static_cast<uint8_t>(idx);
    }
}

__global__ void gather_values(const uint32_t *__restrict__ idx, long long n,
                              const float *__restrict__ vol, float *__restrict__ out) {
    size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t s = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         s < static_cast<size_t>(n); s += stride)
        out[s] = vol[idx[s]];
}

// (min, max, count) partials over values (optionally masked), folded on host
// in block order. min/max are selections, exact in any order.
__global__ __launch_bounds__(kBlock) void range_partials(const float *__restrict__ v,
                                                         const uint8_t *__restrict__ mask, long long n,
                                                         double *__restrict__ partials) {
    __shared__ double smin[kBlock], smax[kBlock], scnt[kBlock];
    double lo = INFINITY, hi = -INFINITY, c = 0.0;
    size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         i < static_cast<size_t>(n); i += stride) {
        if (mask && !mask[i]) continue;
        const double x = static_cast<double>(v[i]);
        lo = fmin(lo, x);
        hi = fmax(hi, x);
        c += 1.0;
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    scnt[threadIdx.x] = c;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
            scnt[threadIdx.x] += scnt[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[blockIdx.x * 3 + 0] = smin[0];
        partials[blockIdx.x * 3 + 1] = smax[0];
        partials[blockIdx.x * 3 + 2] = scnt[0];
    }
}


// The two sets' partials folded on the device by a tree (min and max
// are selections and the counts exact integers, so any order gives the same
// result): out = fixed (min, max, count), moving (min, max, count).
__global__ __launch_bounds__(kBlock) void range_fold(const double *__restrict__ part, unsigned G,
                                                     double *__restrict__ out) {
    __shared__ double smin[kBlock], smax[kBlock], scnt[kBlock];
    for (int k = 0; k < 2; ++k) {
        const double *p = part + static_cast<size_t>(k) * G * 3;
        double lo = INFINITY, hi = -INFINITY, c = 0.0;
        for (unsigned b = threadIdx.x; b < G; b += kBlock) {
            lo = fmin(lo, p[b * 3 + 0]);
            hi = fmax(hi, p[b * 3 + 1]);
            c += p[b * 3 + 2];
        }
        smin[threadIdx.x] = lo;
        smax[threadIdx.x] = hi;
        scnt[threadIdx.x] = c;
        __syncthreads();
        for (int w = kBlock / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) {
                smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
                smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
                scnt[threadIdx.x] += scnt[threadIdx.x + w];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            out[k * 3 + 0] = smin[0];
            out[k * 3 + 1] = smax[0];
            out[k * 3 + 2] = scnt[0];
        }
        __syncthreads();
    }
}

// The value ranges of the metric's two sets (ops/mattes.rs:587-589: fixed over
// the sample set, moving over its mask), enqueued into the pinned result
// block; read with finish_ranges once the stream has drained.
static void enqueue_ranges(Metric &m, long long nm) {
    const unsigned G = 512;
    m.range_part.alloc(2 * G * 3);
    LaunchScope scope("range_partials");
    hipLaunchKernelGGL(range_partials, dim3(G), dim3(kBlock), 0, backend().stream, m.fval.ptr,
                       static_cast<const uint8_t *>(nullptr), m.SS.n, m.range_part.ptr);
    hipLaunchKernelGGL(range_partials, dim3(G), dim3(kBlock), 0, backend().stream, m.moving.ptr,
                       m.Mv.mask, nm, m.range_part.ptr + G * 3);
    hipLaunchKernelGGL(range_fold, dim3(1), dim3(kBlock), 0, backend().stream, m.range_part.ptr, G,
                       m.res_dev->ranges);
    check_launch("range_partials");
    m.ranges_pending = true;
}

static void finish_ranges(Metric &m) {
    for (int k = 0; k < 2; ++k) {
        const double *r = m.res->ranges + 3 * k;
        double *out = k == 0 ? m.fixed_range : m.moving_range;
        if (r[2] == 0.0) { out[0] = 0.0; out[1] = 0.0; }
        else { out[0] = r[0]; out[1] = r[1]; }
    }
    m.ranges_pending = false;
    m.ranges_known = true;
    m.range_part.release();
}

void invert3_host(const double m[9], double out[9]) {
    // Gauss-Jordan with partial pivoting, dividing rows by the pivot
    // (src/core/matrix.rs:44-102), so a diagonal geometry gives diag(1/spacing).
    double a[3][6];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) a[r][c] = m[r * 3 + c];
        for (int c = 0; c < 3; ++c) a[r][3 + c] = r == c ? 1.0 : 0.0;
    }
    for (int col = 0; col < 3; ++col) {
        int piv = col;
        double best = std::fabs(a[col][col]);
        for (int r = col + 1; r < 3; ++r)
            if (std::fabs(a[r][col]) > best) { best = std::fabs(a[r][col]); piv = r; }
        if (best < 1e-12) fail(SITKHIP_ERR_UNSUPPORTED_GEOMETRY, "singular direction * spacing matrix");
        if (piv != col)
            for (int c = 0; c < 6; ++c) std::swap(a[col][c], a[piv][c]);
        const double diag = a[col][col];
        for (int c = 0; c < 6; ++c) a[col][c] /= diag;
        for (int r = 0; r < 3; ++r) {
            if (r == col) continue;
            const double f = a[r][col];
            if (f == 0.0) continue;
            for (int c = 0; c < 6; ++c) a[r][c] -= f * a[col][c];
        }
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) out[r * 3 + c] = a[r][3 + c];
}

static void index_to_physical(const sitkhip_geometry &g, double out[9]) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) out[r * 3 + c] = g.direction[r * 3 + c] * g.spacing[c];
}

// Random sampling on the device: draw i is SplitMix64(seed + (i + 1) gamma)
// mod voxels (metric.rs:101-122, 419-437: the same draws as the host's), and
// the multiset is laid out in ascending voxel order by a counting sort
// (counts, exclusive scan, fill) -- the order the host path sorts into.
__device__ __forceinline__ uint64_t draw_voxel(uint64_t seed, uint64_t voxels, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return z % voxels;
}

__global__ void draw_counts(uint64_t seed, uint64_t voxels, size_t n, uint32_t *__restrict__ counts) {
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        atomicAdd(&counts[draw_voxel(seed, voxels, i)], 1u);
}

__global__ void fill_sorted(const uint32_t *__restrict__ counts, const uint32_t *__restrict__ offsets,
                            size_t voxels, uint32_t *__restrict__ out) {
    for (size_t v = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < voxels;
         v += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const uint32_t c = counts[v], o = offsets[v];
        for (uint32_t k = 0; k < c; ++k) out[o + k] = static_cast<uint32_t>(v);
    }
}

static void device_draws(Metric &m, size_t voxels, size_t n, uint64_t seed) {
    hipStream_t st = backend().stream;
    DeviceBuffer<uint32_t> counts(voxels), offsets(voxels);
    HIP_OK(hipMemsetAsync(counts.ptr, 0, voxels * sizeof(uint32_t), st));
    {
        LaunchScope scope("draw_counts");
        hipLaunchKernelGGL(draw_counts, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st, seed,
                           static_cast<uint64_t>(voxels), n, counts.ptr);
        check_launch("draw_counts");
    }
    size_t temp = 0;
    HIP_OK(hipcub::DeviceScan::ExclusiveSum(nullptr, temp, counts.ptr, offsets.ptr,
                                            static_cast<int>(voxels), st));
    DeviceBuffer<unsigned char> scratch(temp);
    HIP_OK(hipcub::DeviceScan::ExclusiveSum(scratch.ptr, temp, counts.ptr, offsets.ptr,
                                            static_cast<int>(voxels), st));
    m.idx.alloc(n);
    {
        LaunchScope scope("fill_sorted");
        hipLaunchKernelGGL(fill_sorted, dim3(grid_for(voxels, 256, 8192)), dim3(256), 0, st, counts.ptr,
                           offsets.ptr, voxels, m.idx.ptr);
        check_launch("fill_sorted");
    }
}

// ------------------------------------------------------- batched level set-up
// Each kernel below serves many jobs of one kind in one launch: workgroup b
// (256 lanes) takes job job_of[b] and the (b - first)-th chunk of its items;
// per item the arithmetic is that of the one-job kernel above it.
constexpr int kPrepLanes = 256;
constexpr int kPrepItems = 8;  // items per lane and workgroup chunk
constexpr long long kPrepChunk = static_cast<long long>(kPrepLanes) * kPrepItems;
constexpr long long kRangeChunk = static_cast<long long>(kPrepLanes) * 64;

struct CopyJob {  // src == nullptr: fill with zeros
    const uint32_t *src;
    uint32_t *dst;
    const uint8_t *src_tail;
    uint8_t *dst_tail;
    long long words, tail;
    unsigned first;
};
struct DrawJob {
    uint64_t seed, voxels, n;
    uint32_t *counts;
    unsigned first;
};
struct FillJob {
    const uint32_t *counts, *offsets;
    uint64_t voxels;
    uint32_t base;  // the scan's value at the job's first voxel
    uint32_t *out;
    unsigned first;
};
struct GatherJob {
    const uint32_t *idx;
    long long n;
    const float *src;
    float *out;
    unsigned first;
};
struct RangeJob {
    const float *v;
    const uint8_t *mask;
    long long n;
    double *out;
    unsigned first, nblocks;
};
struct BinsJob {
    const float *fval;
    long long n;
    int B;
    double fbs, fnmin;
    uint8_t *out;
    unsigned first;
};

__global__ __launch_bounds__(kPrepLanes) void copy_batch(const CopyJob *__restrict__ jobs,
                                                         const int *__restrict__ job_of) {
    const CopyJob &J = jobs[job_of[blockIdx.x]];
    const long long c = static_cast<long long>(blockIdx.x - J.first) * kPrepChunk;
#pragma unroll
    for (int k = 0; k < kPrepItems; ++k) {
        const long long w = c + k * kPrepLanes + threadIdx.x;
        if (w < J.words) J.dst[w] = J.src ? J.src[w] : 0u;
    }
    if (c == 0 && threadIdx.x < J.tail) J.dst_tail[threadIdx.x] = J.src ? J.src_tail[threadIdx.x] : 0u;
}

__global__ __launch_bounds__(kPrepLanes) void draw_counts_batch(const DrawJob *__restrict__ jobs,
                                                                const int *__restrict__ job_of) {
    const DrawJob &J = jobs[job_of[blockIdx.x]];
    const long long c = static_cast<long long>(blockIdx.x - J.first) * kPrepChunk;
#pragma unroll
    for (int k = 0; k < kPrepItems; ++k) {
        const long long i = c + k * kPrepLanes + threadIdx.x;
        if (i < static_cast<long long>(J.n))
            atomicAdd(&J.counts[draw_voxel(J.seed, J.voxels, static_cast<uint64_t>(i))], 1u);
    }
}

__global__ __launch_bounds__(kPrepLanes) void fill_sorted_batch(const FillJob *__restrict__ jobs,
                                                                const int *__restrict__ job_of) {
    const FillJob &J = jobs[job_of[blockIdx.x]];
    const long long c = static_cast<long long>(blockIdx.x - J.first) * kPrepChunk;
#pragma unroll
    for (int k = 0; k < kPrepItems; ++k) {
        const long long v = c + k * kPrepLanes + threadIdx.x;
        if (v < static_cast<long long>(J.voxels)) {
            const uint32_t n = J.counts[v], o = J.offsets[v] - J.base;
            for (uint32_t q = 0; q < n; ++q) J.out[o + q] = static_cast<uint32_t>(v);
        }
    }
}

__global__ __launch_bounds__(kPrepLanes) void gather_batch(const GatherJob *__restrict__ jobs,
                                                           const int *__restrict__ job_of) {
    const GatherJob &J = jobs[job_of[blockIdx.x]];
    const long long c = static_cast<long long>(blockIdx.x - J.first) * kPrepChunk;
#pragma unroll
    for (int k = 0; k < kPrepItems; ++k) {
        const long long s = c + k * kPrepLanes + threadIdx.x;
        if (s < J.n) J.out[s] = J.src[J.idx[s]];
    }
}

// (min, max, count) of one chunk of kRangeChunk values per workgroup, into
// partials[3 * b]; range_fold_batch folds a job's chunks in order. min and
// max are selections and the count an integer: exact in any order.
__global__ __launch_bounds__(kPrepLanes) void range_partials_batch(const RangeJob *__restrict__ jobs,
                                                                   const int *__restrict__ job_of,
                                                                   double *__restrict__ partials) {
    __shared__ double smin[kPrepLanes], smax[kPrepLanes], scnt[kPrepLanes];
    const RangeJob &J = jobs[job_of[blockIdx.x]];
    const long long c = static_cast<long long>(blockIdx.x - J.first) * kRangeChunk;
    const long long e = c + kRangeChunk < J.n ? c + kRangeChunk : J.n;
    double lo = INFINITY, hi = -INFINITY, cnt = 0.0;
    for (long long i = c + threadIdx.x; i < e; i += kPrepLanes) {
        if (J.mask && !J.mask[i]) continue;
        const double x = static_cast<double>(J.v[i]);
        lo = fmin(lo, x);
        hi = fmax(hi, x);
        cnt += 1.0;
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    scnt[threadIdx.x] = cnt;
    __syncthreads();
    for (int w = kPrepLanes / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
            scnt[threadIdx.x] += scnt[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[3 * blockIdx.x + 0] = smin[0];
        partials[3 * blockIdx.x + 1] = smax[0];
        partials[3 * blockIdx.x + 2] = scnt[0];
    }
}

__global__ void range_fold_batch(const RangeJob *__restrict__ jobs, const double *__restrict__ partials) {
    const RangeJob &J = jobs[blockIdx.x];
    if (threadIdx.x != 0) return;
    double lo = INFINITY, hi = -INFINITY, cnt = 0.0;
    for (unsigned b = 0; b < J.nblocks; ++b) {
        const double *p = partials + 3 * static_cast<size_t>(J.first + b);
        lo = fmin(lo, p[0]);
        hi = fmax(hi, p[1]);
        cnt += p[2];
    }
    J.out[0] = lo;
    J.out[1] = hi;
    J.out[2] = cnt;
}

__global__ __launch_bounds__(kPrepLanes) void fixed_bins_batch(const BinsJob *__restrict__ jobs,
                                                               const int *__restrict__ job_of) {
    const BinsJob &J = jobs[job_of[blockIdx.x]];
    const long long c = static_cast<long long>(blockIdx.x - J.first) * kPrepChunk;
#pragma unroll
    for (int k = 0; k < kPrepItems; ++k) {
        const long long s = c + k * kPrepLanes + threadIdx.x;
        if (s < J.n) {
            const double v = static_cast<double>(J.fval[s]);
            const double term = v / J.fbs - J.fnmin;
            long long idx = static_cast<long long>(term);
            if (idx < 2) idx = 2;
            else if (idx > J.B - 3) idx = J.B - 3;
            J.out[s] = static_cast<uint8_t>(idx);
        }
    }
}

// Appends job j's chunks to the workgroup map; returns its first workgroup.
static unsigned map_chunks(std::vector<int> &job_of, long long items, long long chunk, int j) {
    const unsigned first = static_cast<unsigned>(job_of.size());
    const long long nb = (items + chunk - 1) / chunk;
    job_of.insert(job_of.end(), static_cast<size_t>(nb), j);
    if (job_of.size() >= (1ull << 31)) fail(SITKHIP_ERR_INVALID_ARGUMENT, "batched set-up too large");
    return first;
}

template <typename Job, typename Kernel>
static void launch_jobs(PrepBatch &pb, const char *name, Kernel kernel, const std::vector<Job> &jobs,
                        const std::vector<int> &job_of) {
    if (job_of.empty()) return;
    const Job *dj = pb.upload(jobs);
    const int *dm = pb.upload(job_of);
    launch_kernel(name, kernel, dim3(static_cast<unsigned>(job_of.size())), dim3(kPrepLanes), 0, dj, dm);
    check_launch(name);
}

static void flush_copies(PrepBatch &pb) {
    std::vector<CopyJob> jobs;
    std::vector<int> job_of;
    auto add = [&](const void *src, void *dst, size_t bytes) {
        if (bytes == 0) return;
        if ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) % 4 != 0)
            fail(SITKHIP_ERR_DRIVER, "batched set-up copy of unaligned memory");
        CopyJob J{};
        J.src = static_cast<const uint32_t *>(src);
        J.dst = static_cast<uint32_t *>(dst);
        J.words = static_cast<long long>(bytes / 4);
        J.tail = static_cast<long long>(bytes % 4);
        J.src_tail = src ? static_cast<const uint8_t *>(src) + 4 * J.words : nullptr;
        J.dst_tail = static_cast<uint8_t *>(dst) + 4 * J.words;
        J.first = map_chunks(job_of, J.words > 0 ? J.words : 1, kPrepChunk, static_cast<int>(jobs.size()));
        jobs.push_back(J);
    };
    for (const auto &c : pb.copies) add(c.src, c.dst, c.bytes);
    for (const auto &z : pb.zeros) add(nullptr, z.dst, z.bytes);
    launch_jobs(pb, "copy_batch", copy_batch, jobs, job_of);
    pb.copies.clear();
    pb.zeros.clear();
}

// The draws of every job, laid out like device_draws: counts per voxel over
// all jobs' grids end to end, one exclusive scan, then each job's samples in
// ascending voxel order. Groups keep the scan's int item count and the
// offsets' 32 bits in range.
static void flush_draws(PrepBatch &pb) {
    size_t i0 = 0;
    const auto &D = pb.draws;
    while (i0 < D.size()) {
        size_t i1 = i0;
        uint64_t vox = 0, draws = 0;
        while (i1 < D.size() && (i1 == i0 || (vox + D[i1].voxels < (1ull << 31) &&
                                             draws + D[i1].n < (1ull << 32)))) {
            vox += D[i1].voxels;
            draws += D[i1].n;
            ++i1;
        }
        if (vox >= (1ull << 31) || draws >= (1ull << 32))
            fail(SITKHIP_ERR_UNSUPPORTED_GEOMETRY, "device draws need a fixed grid below 2^31 voxels");
        pb.keep.emplace_back(static_cast<size_t>(vox) * 2 * sizeof(uint32_t));
        uint32_t *counts = reinterpret_cast<uint32_t *>(pb.keep.back().ptr);
        uint32_t *offsets = counts + vox;
        std::vector<DrawJob> dj;
        std::vector<FillJob> fj;
        std::vector<int> dmap, fmap;
        uint64_t vbase = 0, nbase = 0;
        for (size_t i = i0; i < i1; ++i) {
            DrawJob d{D[i].seed, D[i].voxels, D[i].n, counts + vbase, 0};
            d.first = map_chunks(dmap, static_cast<long long>(D[i].n), kPrepChunk, static_cast<int>(dj.size()));
            if (D[i].n > 0) dj.push_back(d);
            else dmap.resize(d.first);
            FillJob f{counts + vbase, offsets + vbase, D[i].voxels, static_cast<uint32_t>(nbase), D[i].out, 0};
            f.first = map_chunks(fmap, static_cast<long long>(D[i].voxels), kPrepChunk, static_cast<int>(fj.size()));
            fj.push_back(f);
            vbase += D[i].voxels;
            nbase += D[i].n;
        }
        PrepBatch::Zero z{counts, static_cast<size_t>(vox) * sizeof(uint32_t)};
        pb.zeros.push_back(z);
        flush_copies(pb);
        launch_jobs(pb, "draw_counts_batch", draw_counts_batch, dj, dmap);
        size_t temp = 0;
        HIP_OK(hipcub::DeviceScan::ExclusiveSum(nullptr, temp, counts, offsets, static_cast<int>(vox),
                                                backend().stream));
        pb.keep.emplace_back(temp);
        HIP_OK(hipcub::DeviceScan::ExclusiveSum(pb.keep.back().ptr, temp, counts, offsets,
                                                static_cast<int>(vox), backend().stream));
        launch_jobs(pb, "fill_sorted_batch", fill_sorted_batch, fj, fmap);
        i0 = i1;
    }
    pb.draws.clear();
}

void PrepBatch::flush() {
    enqueue_placements(*this);
    flush_copies(*this);
    flush_draws(*this);
    {
        std::vector<GatherJob> jobs;
        std::vector<int> job_of;
        for (const auto &g : gathers) {
            if (g.n <= 0) continue;
            GatherJob J{g.idx, g.n, g.src, g.out, 0};
            J.first = map_chunks(job_of, g.n, kPrepChunk, static_cast<int>(jobs.size()));
            jobs.push_back(J);
        }
        launch_jobs(*this, "gather_batch", gather_batch, jobs, job_of);
        gathers.clear();
    }
    if (!ranges.empty()) {
        std::vector<RangeJob> jobs;
        std::vector<int> job_of;
        for (const auto &r : ranges) {
            RangeJob J{r.v, r.mask, r.n, r.out, 0, 0};
            J.first = map_chunks(job_of, r.n > 0 ? r.n : 1, kRangeChunk, static_cast<int>(jobs.size()));
            J.nblocks = static_cast<unsigned>(job_of.size()) - J.first;
            jobs.push_back(J);
        }
        const RangeJob *dj = upload(jobs);
        const int *dm = upload(job_of);
        keep.emplace_back(job_of.size() * 3 * sizeof(double));
        double *part = reinterpret_cast<double *>(keep.back().ptr);
        launch_kernel("range_partials_batch", range_partials_batch, dim3(static_cast<unsigned>(job_of.size())),
                      dim3(kPrepLanes), 0, dj, dm, part);
        check_launch("range_partials_batch");
        launch_kernel("range_fold_batch", range_fold_batch, dim3(static_cast<unsigned>(jobs.size())), dim3(64),
                      0, dj, static_cast<const double *>(part));
        check_launch("range_fold_batch");
        ranges.clear();
    }
    {
        std::vector<BinsJob> jobs;
        std::vector<int> job_of;
        for (const auto &b : bins) {
            if (b.n <= 0) continue;
            BinsJob J{b.fval, b.n, b.B, b.fbs, b.fnmin, b.out, 0};
            J.first = map_chunks(job_of, b.n, kPrepChunk, static_cast<int>(jobs.size()));
            jobs.push_back(J);
        }
        launch_jobs(*this, "fixed_bins_batch", fixed_bins_batch, jobs, job_of);
        bins.clear();
    }
}

static Metric *metric_create_impl(int kind, const Image &fixed, const int64_t *indices,
                                  size_t n_indices, const Mask *fmask, const Image &moving,
                                  const Mask *mmask, int precision, bool draw_on_device,
                                  size_t n_draws, uint64_t draw_seed,
                                  const CreateOptions &opt = CreateOptions{}) {
    if (kind != SITKHIP_METRIC_MEAN_SQUARES && kind != SITKHIP_METRIC_MATTES)
        fail(SITKHIP_ERR_INVALID_ARGUMENT, "unknown metric kind");
    if (precision != SITKHIP_PRECISION_F32 && precision != SITKHIP_PRECISION_F64)
        fail(SITKHIP_ERR_INVALID_ARGUMENT, "unknown precision");
    const sitkhip_geometry &fg = fixed.geom;
    const sitkhip_geometry &mg = moving.geom;
    for (int d = 0; d < 3; ++d) {
        if (fg.size[d] >= (1u << 31) || mg.size[d] >= (1u << 31))
            fail(SITKHIP_ERR_UNSUPPORTED_GEOMETRY, "axis longer than 2^31 voxels");
    }
    if (moving.len() >= (1ull << 31))
        fail(SITKHIP_ERR_UNSUPPORTED_GEOMETRY, "moving volume above 2^31 voxels");
    const size_t nf = fixed.len();
    if (nf >= (1ull << 32)) fail(SITKHIP_ERR_UNSUPPORTED_GEOMETRY, "fixed grid above 2^32 voxels");
    if (fmask && !same_geometry(fmask->geom, fg))
        fail(SITKHIP_ERR_UNSUPPORTED_GEOMETRY, "fixed mask geometry differs from the fixed image");
    if (mmask && !same_geometry(mmask->geom, mg))
        fail(SITKHIP_ERR_UNSUPPORTED_GEOMETRY, "moving mask geometry differs from the moving image");
    PrepBatch *prep = opt.prep;
    if (prep && (fmask || opt.points || !opt.defer_ranges))
        fail(SITKHIP_ERR_INVALID_ARGUMENT,
             "a batched set-up takes no fixed mask or explicit points and defers the ranges");

    auto *m = new Metric();
    try {
        m->kind = kind;
        m->precision = precision;
        m->voxels_fixed = nf;
        // Fixed grid geometry.
        for (int d = 0; d < 3; ++d) {
            m->F.size[d] = static_cast<uint32_t>(fg.size[d]);
            m->F.origin[d] = fg.origin[d];
        }
        m->F.div0 = make_fastdiv(m->F.size[0]);
        m->F.div1 = make_fastdiv(m->F.size[1]);
        index_to_physical(fg, m->F.i2p);
        if (opt.f_i2p) std::memcpy(m->F.i2p, opt.f_i2p, sizeof(m->F.i2p));
        if (opt.points) {
            // n x 1 x 1 grid of gathered values, origin 0, identity i2p: the
            // index-space moments are then the physical ones
            m->pts.alloc(3 * nf);
            m->pts.upload(opt.points, 3 * nf);
            m->F.pts = m->pts.ptr;
        }

        // Sample list (host-side validation and mask filter, as FixedSamples does).
        std::vector<uint8_t> hmask;
        if (fmask) {
            hmask.resize(nf);
            fmask->buf.download(hmask.data(), nf);
            sync();
        }
        std::vector<uint32_t> list;
        bool grid = indices == nullptr && !draw_on_device;
        if (draw_on_device) {
            if (fmask) fail(SITKHIP_ERR_INVALID_ARGUMENT, "device draws take no fixed mask");
            if (n_draws >= (1ull << 31)) fail(SITKHIP_ERR_INVALID_ARGUMENT, "too many draws");
            // the counting sort scans one count per voxel with an int item count
            if (nf >= (1ull << 31))
                fail(SITKHIP_ERR_UNSUPPORTED_GEOMETRY,
                     "device draws need a fixed grid below 2^31 voxels");
            if (prep) {
                m->idx.alloc(n_draws);
                prep->draws.push_back({static_cast<uint64_t>(nf), static_cast<uint64_t>(n_draws), draw_seed,
                                       m->idx.ptr});
            } else {
                device_draws(*m, nf, n_draws, draw_seed);
            }
        } else if (!grid) {
            list.reserve(n_indices);
            for (size_t s = 0; s < n_indices; ++s) {
                const int64_t v = indices[s];
                if (v < 0 || static_cast<uint64_t>(v) >= nf)
                    fail(SITKHIP_ERR_SAMPLE_INDEX_OUT_OF_GRID,
                         "fixed sample index " + std::to_string(v) + " is outside the fixed grid (" +
                             std::to_string(nf) + " voxels)");
                if (fmask && !hmask[static_cast<size_t>(v)]) continue;
                list.push_back(static_cast<uint32_t>(v));
            }
            // Ascending order: the metric is a sum over the sample multiset, so
            // the order only changes rounding of the FP64 moments (the histogram
            // is integer) and it turns the gathers into near-streaming reads.
            // Dense draws are counted per voxel (linear time), sparse ones sorted.
            if (list.size() >= nf / 16) {
                std::vector<uint32_t> count(nf, 0);
                for (uint32_t v : list) ++count[v];
                size_t o = 0;
                for (size_t v = 0; v < nf; ++v)
                    for (uint32_t k = count[v]; k > 0; --k) list[o++] = static_cast<uint32_t>(v);
            } else {
                std::sort(list.begin(), list.end());
            }
        } else if (fmask) {
            for (size_t s = 0; s < nf; ++s)
                if (hmask[s]) list.push_back(static_cast<uint32_t>(s));
            grid = false;
        }
        if (grid) {
            m->SS.idx = nullptr;
            m->SS.n = static_cast<long long>(nf);
        } else if (draw_on_device) {
            m->SS.idx = m->idx.ptr;
            m->SS.n = static_cast<long long>(n_draws);
        } else {
            m->idx.alloc(list.size());
            m->idx.upload(list.data(), list.size());
            m->SS.idx = m->idx.ptr;
            m->SS.n = static_cast<long long>(list.size());
        }
        m->nsamples = static_cast<size_t>(m->SS.n);

        // Per-sample fixed values (grid: a private copy of the fixed volume,
        // or the volume itself when the caller hands it over).
        if (grid && opt.adopt_fixed == &fixed) {
            m->fval = std::move(opt.adopt_fixed->buf);
        } else if (grid) {
            m->fval.alloc(m->nsamples);
            if (prep)
                prep->copies.push_back({fixed.buf.ptr, m->fval.ptr, nf * sizeof(float)});
            else
                HIP_OK(hipMemcpyAsync(m->fval.ptr, fixed.buf.ptr, nf * sizeof(float),
                                      hipMemcpyDeviceToDevice, backend().stream));
        } else if (m->nsamples > 0 && prep) {
            m->fval.alloc(m->nsamples);
            prep->gathers.push_back({m->idx.ptr, m->SS.n, fixed.buf.ptr, m->fval.ptr});
        } else if (m->nsamples > 0) {
            m->fval.alloc(m->nsamples);
            LaunchScope scope("gather_values");
            hipLaunchKernelGGL(gather_values, dim3(grid_for(m->nsamples, 256, 4096)), dim3(256), 0,
                               backend().stream, m->idx.ptr, m->SS.n, fixed.buf.ptr, m->fval.ptr);
            check_launch("gather_values");
        }

        // Moving volume and mask: private device copies (the volume itself when
        // the caller hands it over).
        const size_t nm = moving.len();
        if (opt.adopt_moving == &moving) {
            m->moving = std::move(opt.adopt_moving->buf);
        } else if (prep) {
            m->moving.alloc(nm);
            prep->copies.push_back({moving.buf.ptr, m->moving.ptr, nm * sizeof(float)});
        } else {
            m->moving.alloc(nm);
            HIP_OK(hipMemcpyAsync(m->moving.ptr, moving.buf.ptr, nm * sizeof(float),
                                  hipMemcpyDeviceToDevice, backend().stream));
        }
        if (mmask) {
            m->mmask.alloc(nm);
            if (prep)
                prep->copies.push_back({mmask->buf.ptr, m->mmask.ptr, nm});
            else
                HIP_OK(hipMemcpyAsync(m->mmask.ptr, mmask->buf.ptr, nm, hipMemcpyDeviceToDevice,
                                      backend().stream));
        }
        m->Mv.vol = m->moving.ptr;
        m->Mv.mask = mmask ? m->mmask.ptr : nullptr;
        for (int d = 0; d < 3; ++d) {
            m->Mv.size[d] = static_cast<int>(mg.size[d]);
            m->Mv.origin[d] = mg.origin[d];
        }
        m->Mv.stride1 = static_cast<long long>(mg.size[0]);
        m->Mv.stride2 = static_cast<long long>(mg.size[0] * mg.size[1]);
        double mi2p[9];
        index_to_physical(mg, mi2p);
        invert3_host(mi2p, m->Mv.p2i);
        if (opt.m_p2i) std::memcpy(m->Mv.p2i, opt.m_p2i, sizeof(m->Mv.p2i));
        for (int i = 0; i < 9; ++i) m->Mv.p2i_f[i] = static_cast<float>(m->Mv.p2i[i]);
        {
            ChainConst cc;
            for (int d = 0; d < 3; ++d) {
                cc.f_origin[d] = m->F.origin[d];
                cc.m_origin[d] = m->Mv.origin[d];
            }
            for (int i = 0; i < 9; ++i) {
                cc.f_i2p[i] = m->F.i2p[i];
                cc.m_p2i[i] = m->Mv.p2i[i];
            }
            // from a pinned slot: a copy from pageable memory would wait for
            // the stream to reach it
            void *chain_dev = nullptr;
            m->chain_host = pinned_slot_alloc(sizeof(ChainConst), &chain_dev);
            std::memcpy(m->chain_host, &cc, sizeof(ChainConst));
            m->chain.alloc(1);
            if (prep)  // the kernel reads the pinned slot through its device view
                prep->copies.push_back({chain_dev, m->chain.ptr, sizeof(ChainConst)});
            else
                HIP_OK(hipMemcpyAsync(m->chain.ptr, m->chain_host, sizeof(ChainConst),
                                      hipMemcpyHostToDevice, backend().stream));
        }

        m->partials.alloc(kGrid * 14);
        void *dev = nullptr;
        m->res = static_cast<HostResult *>(pinned_slot_alloc(sizeof(HostResult), &dev));
        m->res_dev = static_cast<HostResult *>(dev);
        m->res->done = 0;

        m->moving_voxels = static_cast<long long>(nm);
        if (opt.ranges && prep) {
            prep->ranges.push_back({m->fval.ptr, nullptr, m->SS.n, m->res_dev->ranges});
            prep->ranges.push_back({m->moving.ptr, m->Mv.mask, static_cast<long long>(nm),
                                    m->res_dev->ranges + 3});
            m->ranges_pending = true;
        } else if (opt.ranges) {
            enqueue_ranges(*m, static_cast<long long>(nm));
            if (!opt.defer_ranges) {
                sync();
                finish_ranges(*m);
            }
        }
    } catch (...) {
        delete m;
        throw;
    }
    return m;
}

Metric *metric_create(int kind, const Image &fixed, const int64_t *indices, size_t n_indices,
                      const Mask *fmask, const Image &moving, const Mask *mmask, int precision) {
    return metric_create_impl(kind, fixed, indices, n_indices, fmask, moving, mmask, precision,
                              false, 0, 0);
}

Metric *metric_create_drawn(int kind, const Image &fixed, size_t n_draws, uint64_t seed,
                            const Image &moving, int precision, const CreateOptions &opt) {
    return metric_create_impl(kind, fixed, nullptr, 0, opt.fixed_mask, moving, opt.moving_mask,
                              precision, true, n_draws, seed, opt);
}

Metric *metric_create_with(int kind, const Image &fixed, const int64_t *indices, size_t n_indices,
                           const Image &moving, int precision, const CreateOptions &opt) {
    return metric_create_impl(kind, fixed, indices, n_indices, opt.fixed_mask, moving,
                              opt.moving_mask, precision, false, 0, 0, opt);
}

std::vector<int64_t> host_draws(size_t voxels, int strategy, double percentage, uint64_t seed) {
    std::vector<int64_t> out;
    if (voxels == 0) fail(SITKHIP_ERR_INVALID_ARGUMENT, "no voxels to draw from");
    check_sampling_fraction(strategy, percentage, voxels);
    if (strategy == 1) {
        size_t stride = static_cast<size_t>(std::ceil(1.0 / percentage));
        if (stride < 1) stride = 1;
        out.reserve(voxels / stride + 1);
        for (size_t v = 0; v < voxels; v += stride) out.push_back(static_cast<int64_t>(v));
        return out;
    }
    if (strategy != 2) fail(SITKHIP_ERR_INVALID_ARGUMENT, "unknown sampling strategy");
    const size_t count = static_cast<size_t>(static_cast<double>(voxels) * percentage);
    out.resize(count);
    uint64_t state = seed;
    for (size_t i = 0; i < count; ++i) {
        state += 0x9E3779B97F4A7C15ull;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        out[i] = static_cast<int64_t>(z % static_cast<uint64_t>(voxels));
    }
    return out;
}

Metric *metric_create_host(int kind, size_t n, const double *fixed_values,
                           const sitkhip_fixed_points &pts, const uint8_t *fixed_mask,
                           const sitkhip_moving_geometry &mg, const double *moving_values,
                           int precision) {
    // The checks of Resident::build (resident.rs:530-660), in its order.
    if (n == 0) fail(SITKHIP_ERR_DEGENERATE_INPUT, "no fixed samples");
    if (!fixed_values || !moving_values) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null value array");
    const size_t mlen = mg.size[0] * mg.size[1] * mg.size[2];
    if (mg.len != mlen || mlen == 0)
        fail(SITKHIP_ERR_DEGENERATE_INPUT, "moving length does not match its size");
    if (mg.strides[0] != 1 || mg.strides[1] != mg.size[0] || mg.strides[2] != mg.size[0] * mg.size[1])
        fail(SITKHIP_ERR_UNSUPPORTED_GEOMETRY, "the moving volume must be dense, dimension 0 fastest");
    sitkhip_geometry fg{}, mgeo{};
    for (int d = 0; d < 3; ++d) {
        mgeo.size[d] = mg.size[d];
        mgeo.spacing[d] = 1.0;
        mgeo.origin[d] = mg.origin[d];
        mgeo.direction[d * 4] = 1.0;
        fg.spacing[d] = 1.0;
        fg.direction[d * 4] = 1.0;
    }
    CreateOptions opt;
    opt.m_p2i = mg.phys_to_index;
    const int64_t *indices = nullptr;
    size_t n_indices = 0, fixed_len = 0;
    switch (pts.mode) {
        case SITKHIP_POINTS_EXPLICIT:
            if (!pts.points) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null explicit points");
            fg.size[0] = n;
            fg.size[1] = fg.size[2] = 1;
            fixed_len = n;
            opt.points = pts.points;
            break;
        case SITKHIP_POINTS_GRID:
        case SITKHIP_POINTS_INDICES:
            for (int d = 0; d < 3; ++d) {
                fg.size[d] = pts.size[d];
                fg.origin[d] = pts.origin[d];
            }
            fixed_len = pts.size[0] * pts.size[1] * pts.size[2];
            if (fixed_len == 0) fail(SITKHIP_ERR_DEGENERATE_INPUT, "empty fixed grid");
            if (pts.mode == SITKHIP_POINTS_GRID && fixed_len != n)
                fail(SITKHIP_ERR_DEGENERATE_INPUT, "a grid sample set has one sample per voxel");
            if (pts.mode == SITKHIP_POINTS_INDICES) {
                if (!pts.indices) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null sample indices");
                indices = pts.indices;
                n_indices = n;
            }
            opt.f_i2p = pts.idx_to_phys;
            break;
        default:
            fail(SITKHIP_ERR_INVALID_ARGUMENT, "unknown fixed point mode");
    }
    if (fixed_mask && pts.mode == SITKHIP_POINTS_EXPLICIT)
        fail(SITKHIP_ERR_MASKED_EXPLICIT_POINTS,
             "a fixed mask gates samples by grid voxel; explicit points have none");
    // A grid sample set whose index list would be the grid itself: no list
    // (the list form is only needed with a subset or repeats).
    std::unique_ptr<Image> fimg(image_upload(SITKHIP_FLOAT64, fixed_values, fixed_len, fg));
    std::unique_ptr<Image> mimg(image_upload(SITKHIP_FLOAT64, moving_values, mlen, mgeo));
    std::unique_ptr<Mask> fmask, mmask;
    if (fixed_mask) fmask.reset(mask_upload(fixed_mask, fixed_len, fg));
    if (mg.mask) mmask.reset(mask_upload(mg.mask, mlen, mgeo));
    return metric_create_impl(kind, *fimg, indices, n_indices, fmask.get(), *mimg, mmask.get(),
                              precision, false, 0, 0, opt);
}

bool metric_has_explicit_points(const Metric &m) { return m.F.pts != nullptr; }

void metric_finish_ranges(Metric &m) {
    if (m.ranges_pending) finish_ranges(m);
}

void metric_destroy(Metric *m) { delete m; }
size_t metric_sample_count(const Metric &m) { return m.nsamples; }
void metric_value_ranges(Metric &m, double fr[2], double mr[2]) {
    if (!m.ranges_pending && !m.ranges_known) enqueue_ranges(m, m.moving_voxels);
    if (m.ranges_pending) {
        sync();
        finish_ranges(m);
    }
    fr[0] = m.fixed_range[0]; fr[1] = m.fixed_range[1];
    mr[0] = m.moving_range[0]; mr[1] = m.moving_range[1];
}

static void check_stages(int nstage, const double *stages) {
    if (nstage < 1 || nstage > kMaxStages)
        fail(SITKHIP_ERR_POINT_MAP_STAGE_COUNT, "the point map has " + std::to_string(nstage) +
                                                    " stages; the device replays 1..=8");
    if (!stages) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null stages");
}

static Stages make_stages(const double *stages, int nstage) {
    check_stages(nstage, stages);
    Stages S;
    S.n = nstage;
    std::memset(S.s, 0, sizeof(S.s));
    std::memcpy(S.s, stages, sizeof(double) * 12 * nstage);
    return S;
}

void mattes_bins_from_ranges(const double fr[2], const double mr[2], size_t bins,
                             sitkhip_mattes_bins *out) {
    // MattesGeometry::new (mattes.rs:152-186).
    const size_t padding = 2;
    if (bins < 2 * padding + 1)
        fail(SITKHIP_ERR_TOO_FEW_BINS, "Mattes needs at least 5 histogram bins, got " +
                                           std::to_string(bins));
    if (fr[1] - fr[0] <= DBL_EPSILON)
        fail(SITKHIP_ERR_CONSTANT_INTENSITY, "the fixed image is constant over the sample set");
    if (mr[1] - mr[0] <= DBL_EPSILON)
        fail(SITKHIP_ERR_CONSTANT_INTENSITY, "the moving image is constant over its mask");
    const double denom = static_cast<double>(bins - 2 * padding);
    out->bins = bins;
    out->padding = padding;
    out->fixed_bin_size = (fr[1] - fr[0]) / denom;
    out->moving_bin_size = (mr[1] - mr[0]) / denom;
    out->fixed_normalized_min = fr[0] / out->fixed_bin_size - static_cast<double>(padding);
    out->moving_normalized_min = mr[0] / out->moving_bin_size - static_cast<double>(padding);
    out->moving_true_min = mr[0];
    out->moving_true_max = mr[1];
}

void mattes_set_bins(Metric &m, const sitkhip_mattes_bins &b, PrepBatch *prep) {
    if (m.kind != SITKHIP_METRIC_MATTES) fail(SITKHIP_ERR_INVALID_ARGUMENT, "not a Mattes metric");
    if (b.bins < 5) fail(SITKHIP_ERR_TOO_FEW_BINS, "Mattes needs at least 5 histogram bins");
    if (b.bins > SITKHIP_MAX_BINS)
        fail(SITKHIP_ERR_TOO_MANY_BINS, "at most " + std::to_string(SITKHIP_MAX_BINS) + " bins");
    if (b.padding != 2) fail(SITKHIP_ERR_HISTOGRAM_SHAPE, "padding must be 2");
    if (!(b.moving_bin_size > 0.0) || !(b.fixed_bin_size > 0.0))
        fail(SITKHIP_ERR_HISTOGRAM_SHAPE, "bin sizes must be positive");
    m.bins = b;
    const int B = static_cast<int>(b.bins);
    m.P.bins = B;
    m.P.mbs = b.moving_bin_size;
    m.P.inv_mbs = 1.0 / b.moving_bin_size;
    m.P.mnmin = b.moving_normalized_min;
    m.P.mmin = b.moving_true_min;
    m.P.mmax = b.moving_true_max;
    m.P.mmin_f = static_cast<float>(b.moving_true_min);
    m.P.mmax_f = static_cast<float>(b.moving_true_max);
    m.P.inv_mbs_f = static_cast<float>(1.0 / b.moving_bin_size);
    // float bounds must not shrink the accepted range
    if (static_cast<double>(m.P.mmin_f) > b.moving_true_min)
        m.P.mmin_f = std::nextafter(m.P.mmin_f, -INFINITY);
    if (static_cast<double>(m.P.mmax_f) < b.moving_true_max)
        m.P.mmax_f = std::nextafter(m.P.mmax_f, INFINITY);
    m.fbin.alloc(m.nsamples);
    if (prep) {
        prep->bins.push_back({m.fval.ptr, m.SS.n, B, b.fixed_bin_size, b.fixed_normalized_min, m.fbin.ptr});
    } else if (m.nsamples > 0) {
        LaunchScope scope("fixed_bins");
        hipLaunchKernelGGL(fixed_bins, dim3(grid_for(m.nsamples, 256, 4096)), dim3(256), 0,
                           backend().stream, m.SS, m.fval.ptr, B, b.fixed_bin_size,
                           b.fixed_normalized_min, m.fbin.ptr);
        check_launch("fixed_bins");
    }
    m.joint.alloc(static_cast<size_t>(kHistCopies) * B * B);
    // the fixed-marginal copies, then the moving-bin column-sum copies (the
    // split tail's statistics); cleared together
    m.marg.alloc(static_cast<size_t>(kHistCopies) * B * 2);
    m.pratio.alloc(static_cast<size_t>(B) * B);
    m.pratio_f.alloc(static_cast<size_t>(B) * B);
    m.tail_part.alloc(2 * kMaxTailSplit + 1);
    if (prep) {
        prep->zeros.push_back({m.joint.ptr, m.joint.bytes()});
        prep->zeros.push_back({m.marg.ptr, m.marg.bytes()});
    } else {
        HIP_OK(hipMemsetAsync(m.joint.ptr, 0, m.joint.bytes(), backend().stream));
        HIP_OK(hipMemsetAsync(m.marg.ptr, 0, m.marg.bytes(), backend().stream));
    }
    m.bins_set = true;
}

static void require_bins(const Metric &m) {
    if (m.kind != SITKHIP_METRIC_MATTES) fail(SITKHIP_ERR_INVALID_ARGUMENT, "not a Mattes metric");
    if (!m.bins_set)
        fail(SITKHIP_ERR_HISTOGRAM_SHAPE, "the Mattes histogram geometry has not been set");
}

// Calls f with the point map as the kernels take it: one stage in registers
// (the common case, 12 doubles of kernel argument) or the general replay.
// Metrics over explicit points take PointStages (every sample on the exact
// chain).
template <typename Fn>
static void with_map(const Metric &m, const Stages &S, Fn &&f) {
    if (m.F.pts) {
        PointStages P;
        static_cast<Stages &>(P) = S;
        f(P);
    } else if (S.n == 1) {
        OneStage one;
        std::memcpy(one.m, S.s, sizeof(one.m));
        f(one);
    } else {
        f(S);
    }
}

void launch_tail(Metric &m, TailOut *out, bool clear) {
    const int B = m.P.bins;
    // staged histogram plus the tail's own ~15 KB of LDS within 64 KB
    const bool lds = B <= 76;
    const size_t shm = lds ? static_cast<size_t>(B * B + B) * sizeof(unsigned long long) : 0;
    launch_kernel("mattes_tail", lds ? mattes_tail_kernel<true> : mattes_tail_kernel<false>, dim3(1),
                  dim3(kTail), shm, m.joint.ptr, m.marg.ptr, B, m.P.mbs, m.pratio.ptr, m.pratio_f.ptr, out,
                  clear ? 1 : 0);
    check_launch("mattes_tail");
}

// The split tail when the histogram fits in LDS (returns its workgroup count
// for the derivative pass to finish), else the one-workgroup tail leaving the
// histogram for the derivative pass to clear (returns 0).
static int launch_tail_for_derivative(Metric &m, TailOut *out) {
    const int B = m.P.bins;
    if (B > 76) {
        launch_tail(m, out, false);
        return 0;
    }
    int T = (B * B + kTail - 1) / kTail;  // about one bin per lane
    if (T > kMaxTailSplit) T = kMaxTailSplit;
    launch_kernel("mattes_tail", mattes_tail_split, dim3(T), dim3(kTail), 0, m.joint.ptr, m.marg.ptr,
                  m.marg.ptr + static_cast<size_t>(kHistCopies) * B, B, m.P.mbs, m.pratio.ptr, m.pratio_f.ptr,
                  out, m.tail_part.ptr);
    check_launch("mattes_tail");
    return T;
}

// tail: run the tail after the value pass; clear: the tail clears the
// histogram (otherwise the derivative pass that follows does).
template <typename R>
static void launch_hist(Metric &m, const Stages &S, bool tail, bool clear = true,
                        SampleRec<R> *rec = nullptr) {
    const int B = m.P.bins;
    const bool lds = B <= kLdsMaxBins;
    const size_t shm = lds ? static_cast<size_t>(B * B + B) * sizeof(unsigned long long) : 0;
    // the deferred-sample lists (FP32): one slice per workgroup, as long as
    // the most samples xcd_walk gives one workgroup
    uint32_t *dq = nullptr;
    unsigned cap = 0;
    if (sizeof(R) == 4) {
        const long long chunk = (m.SS.n + 7) / 8;
        const long long stride = static_cast<long long>(kHistGrid / 8) * kHistBlock;
        cap = static_cast<unsigned>((chunk + stride - 1) / stride * kHistBlock);
        const size_t need = static_cast<size_t>(kHistGrid) * cap;
        if (m.defer.len < need) m.defer.alloc(need);
        dq = m.defer.ptr;
    }
    // the split tail follows a value pass with records and a value-only pass
    // that clears after itself (B <= 76): it reads the column sums this pass
    // adds after the marginal copies
    const bool split = B <= 76 && ((rec && !tail) || (tail && clear));
    unsigned long long *col = split ? m.marg.ptr + static_cast<size_t>(kHistCopies) * B : nullptr;
    {
        const IndexMap X = make_index_map(m.F, m.Mv, S.s, S.n);
        with_map(m, S, [&](const auto &map) {
            using Map = std::decay_t<decltype(map)>;
            auto go = [&](auto kernel, size_t bytes) {
                launch_kernel("mattes_hist", kernel, dim3(kHistGrid), dim3(kHistBlock), bytes, m.F, m.Mv, map,
                              X, m.chain.ptr, m.SS, m.fbin.ptr, m.P, m.joint.ptr, m.marg.ptr, col, rec, dq,
                              cap);
            };
            if (m.bspline) {
                if (lds) {
                    if (rec) go(mattes_hist<R, true, true, Map, true>, shm);
                    else go(mattes_hist<R, true, false, Map, true>, shm);
                } else {
                    if (rec) go(mattes_hist<R, false, true, Map, true>, 0);
                    else go(mattes_hist<R, false, false, Map, true>, 0);
                }
            } else if (lds) {
                if (rec) go(mattes_hist<R, true, true, Map>, shm);
                else go(mattes_hist<R, true, false, Map>, shm);
            } else {
                if (rec) go(mattes_hist<R, false, true, Map>, 0);
                else go(mattes_hist<R, false, false, Map>, 0);
            }
        });
        check_launch("mattes_hist");
    }
    if (!tail) return;
    if (split) {
        const int T = launch_tail_for_derivative(m, &m.res_dev->tail);
        launch_kernel("mattes_tail_finish", mattes_tail_finish, dim3(16), dim3(256), 0, m.tail_part.ptr, T,
                      &m.res_dev->tail, m.joint.ptr, static_cast<unsigned>(m.joint.len), m.marg.ptr,
                      static_cast<unsigned>(m.marg.len));
        check_launch("mattes_tail_finish");
        return;
    }
    launch_tail(m, &m.res_dev->tail, clear);
}

// Completion numbers are unique across metrics: a result slot reused by a
// new metric can never show a number the new metric waits for.
static std::atomic<unsigned long long> g_eval_seq{0};

template <int NS>
static void launch_fold(Metric &m) {
    m.seq = g_eval_seq.fetch_add(1) + 1;
    launch_kernel("fold_partials", fold_partials<NS>, dim3(1), dim3(kBlock), 0, m.partials.ptr, kGrid,
                  m.res_dev->moments, &m.res_dev->done, m.seq);
    check_launch("fold_partials");
}

// Waits for the evaluation that ended with launch_fold: spins on the result
// block's sequence word (a few microseconds sooner than a stream wait), and
// after 2 ms falls back to the stream's own wait, which also reports any
// error of the launches.
static void wait_done(Metric &m) {
    const volatile unsigned long long *done = &m.res->done;
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned n = 1;; ++n) {
        if (*done == m.seq) {
            std::atomic_thread_fence(std::memory_order_acquire);
            return;
        }
        if ((n & 255u) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) break;
    }
    sync();
    if (*done != m.seq) fail(SITKHIP_ERR_DRIVER, "evaluation finished without its completion word");
}

// clear_hist: the derivative pass also clears the histogram copies (after a
// tail that left them).
template <typename R>
static void launch_deriv(Metric &m, const Stages &S, const R *pratio, bool clear_hist = false) {
    const int B = m.P.bins;
    unsigned long long *clr = clear_hist ? m.joint.ptr : nullptr;
    const unsigned clr_n = clear_hist ? static_cast<unsigned>(m.joint.len) : 0u;
    unsigned long long *clr2 = clear_hist ? m.marg.ptr : nullptr;
    const unsigned clr2_n = clear_hist ? static_cast<unsigned>(m.marg.len) : 0u;
    const size_t shm = static_cast<size_t>(B) * B * sizeof(R);
    const bool lds = shm + 2048 <= 64 * 1024;  // plus the reduction's static LDS
    {
        const IndexMap X = make_index_map(m.F, m.Mv, S.s, S.n);
        with_map(m, S, [&](const auto &map) {
            using Map = std::decay_t<decltype(map)>;
            auto go = [&](auto kernel, size_t bytes) {
                launch_kernel("mattes_deriv", kernel, dim3(kGrid), dim3(kBlock), bytes, m.F, m.Mv, map, X,
                              m.chain.ptr, m.SS, m.fbin.ptr, m.P, pratio, m.partials.ptr, clr, clr_n, clr2,
                              clr2_n);
            };
            if (m.bspline) {
                if (lds) go(mattes_deriv<R, true, Map, true>, shm);
                else go(mattes_deriv<R, false, Map, true>, 0);
            } else if (lds) {
                go(mattes_deriv<R, true, Map>, shm);
            } else {
                go(mattes_deriv<R, false, Map>, 0);
            }
        });
        check_launch("mattes_deriv");
    }
    launch_fold<12>(m);
}

// The derivative pass from the records of the value pass just run; it also
// clears the histogram copies the tail left.
template <typename R>
static void launch_deriv_rec(Metric &m, const R *pratio, const SampleRec<R> *rec, int tail_T) {
    {
        launch_kernel("mattes_deriv", mattes_deriv_rec<R>, dim3(kGrid), dim3(kBlock), 0, m.F, m.Mv,
                      m.chain.ptr, m.SS, m.fbin.ptr, m.P, pratio, rec, m.partials.ptr, m.joint.ptr,
                      static_cast<unsigned>(m.joint.len), m.marg.ptr, static_cast<unsigned>(m.marg.len),
                      m.tail_part.ptr, tail_T, &m.res_dev->tail);
        check_launch("mattes_deriv");
    }
    launch_fold<12>(m);
}

template <typename R>
static SampleRec<R> *sample_records(Metric &m) {
    const size_t bytes = static_cast<size_t>(m.SS.n) * sizeof(SampleRec<R>);
    if (m.recs.len < bytes) m.recs.alloc(bytes);
    return reinterpret_cast<SampleRec<R> *>(m.recs.ptr);
}

void *mattes_sample_records(Metric &m) {
    if (m.precision == SITKHIP_PRECISION_F64) return sample_records<double>(m);
    return sample_records<float>(m);
}

void mattes_joint_histogram(Metric &m, const double *stages, int nstage, double *joint,
                            double *marginal, size_t *valid) {
    require_bins(m);
    Stages S = make_stages(stages, nstage);
    const int B = m.P.bins;
    if (m.precision == SITKHIP_PRECISION_F64) launch_hist<double>(m, S, false);
    else launch_hist<float>(m, S, false);
    std::vector<unsigned long long> cj(static_cast<size_t>(kHistCopies) * B * B),
        cm(static_cast<size_t>(kHistCopies) * B);
    m.joint.download(cj.data(), cj.size());
    m.marg.download(cm.data(), cm.size());
    HIP_OK(hipMemsetAsync(m.joint.ptr, 0, m.joint.bytes(), backend().stream));
    HIP_OK(hipMemsetAsync(m.marg.ptr, 0, m.marg.bytes(), backend().stream));
    sync();
    std::vector<unsigned long long> hj(static_cast<size_t>(B) * B, 0), hm(B, 0);
    for (int x = 0; x < kHistCopies; ++x) {
        for (int i = 0; i < B * B; ++i) hj[i] += cj[static_cast<size_t>(x) * B * B + i];
        for (int i = 0; i < B; ++i) hm[i] += cm[static_cast<size_t>(x) * B + i];
    }
    unsigned long long v = 0;
    for (int i = 0; i < B * B; ++i)
        if (joint) joint[i] = static_cast<double>(hj[i]) * kInvFix;
    for (int i = 0; i < B; ++i) {
        if (marginal) marginal[i] = static_cast<double>(hm[i]);
        v += hm[i];
    }
    if (valid) *valid = static_cast<size_t>(v);
}

void mattes_derivative_moments(Metric &m, const double *stages, int nstage, const double *pratio,
                               double moments[12]) {
    require_bins(m);
    if (!pratio) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null pratio");
    Stages S = make_stages(stages, nstage);
    const int B = m.P.bins;
    m.pratio.upload(pratio, static_cast<size_t>(B) * B);
    if (m.precision == SITKHIP_PRECISION_F64) {
        launch_deriv<double>(m, S, m.pratio.ptr);
    } else {
        std::vector<float> pf(static_cast<size_t>(B) * B);
        for (size_t i = 0; i < pf.size(); ++i) pf[i] = static_cast<float>(pratio[i]);
        m.pratio_f.upload(pf.data(), pf.size());
        launch_deriv<float>(m, S, m.pratio_f.ptr);
    }
    sync();
    std::memcpy(moments, m.res->moments, 12 * sizeof(double));
    if (m.precision == SITKHIP_PRECISION_F32) moments_to_physical(m.F, moments, moments + 3);
}

static void value_pass(Metric &m, const Stages &S) {
    if (m.precision == SITKHIP_PRECISION_F64) launch_hist<double>(m, S, true);
    else launch_hist<float>(m, S, true);
}

void mattes_value(Metric &m, const double *stages, int nstage, double *value, size_t *valid) {
    require_bins(m);
    Stages S = make_stages(stages, nstage);
    value_pass(m, S);
    sync();
    const TailOut &t = m.res->tail;
    if (value) *value = t.value;
    if (valid) *valid = static_cast<size_t>(t.valid);
}

void mattes_value_and_moments(Metric &m, const double *stages, int nstage, double *value,
                              double moments[12], size_t *valid) {
    require_bins(m);
    Stages S = make_stages(stages, nstage);
    if (m.precision == SITKHIP_PRECISION_F64) {
        SampleRec<double> *rec = sample_records<double>(m);
        launch_hist<double>(m, S, false, false, rec);
        const int T = launch_tail_for_derivative(m, &m.res_dev->tail);
        launch_deriv_rec<double>(m, m.pratio.ptr, rec, T);
    } else {
        SampleRec<float> *rec = sample_records<float>(m);
        launch_hist<float>(m, S, false, false, rec);
        const int T = launch_tail_for_derivative(m, &m.res_dev->tail);
        launch_deriv_rec<float>(m, m.pratio_f.ptr, rec, T);
    }
    wait_done(m);
    const TailOut &t = m.res->tail;
    if (moments) {
        std::memcpy(moments, m.res->moments, 12 * sizeof(double));
        if (m.precision == SITKHIP_PRECISION_F32) moments_to_physical(m.F, moments, moments + 3);
    }
    if (value) *value = t.value;
    if (valid) *valid = static_cast<size_t>(t.valid);
}

template <typename R, bool MOM>
static void launch_ms(Metric &m, const Stages &S) {
    {
        const char *name = MOM ? "ms_moments" : "ms_value";
        const IndexMap X = make_index_map(m.F, m.Mv, S.s, S.n);
        with_map(m, S, [&](const auto &map) {
            using Map = std::decay_t<decltype(map)>;
            if (m.bspline)
                launch_kernel(name, ms_pass<R, MOM, Map, true>, dim3(kGrid), dim3(kBlock), 0, m.F, m.Mv, map,
                              X, m.chain.ptr, m.SS, m.fval.ptr, m.partials.ptr);
            else
                launch_kernel(name, ms_pass<R, MOM, Map>, dim3(kGrid), dim3(kBlock), 0, m.F, m.Mv, map, X,
                              m.chain.ptr, m.SS, m.fval.ptr, m.partials.ptr);
        });
        check_launch("ms_pass");
    }
    launch_fold<MOM ? 14 : 2>(m);
}

void ms_moments(Metric &m, const double *stages, int nstage, double out[14]) {
    if (m.kind != SITKHIP_METRIC_MEAN_SQUARES)
        fail(SITKHIP_ERR_INVALID_ARGUMENT, "not a mean-squares metric");
    Stages S = make_stages(stages, nstage);
    if (m.precision == SITKHIP_PRECISION_F64) launch_ms<double, true>(m, S);
    else launch_ms<float, true>(m, S);
    wait_done(m);
    std::memcpy(out, m.res->moments, 14 * sizeof(double));
    if (m.precision == SITKHIP_PRECISION_F32) moments_to_physical(m.F, out + 1, out + 4);
}

void ms_value(Metric &m, const double *stages, int nstage, double *sq, size_t *valid) {
    if (m.kind != SITKHIP_METRIC_MEAN_SQUARES)
        fail(SITKHIP_ERR_INVALID_ARGUMENT, "not a mean-squares metric");
    Stages S = make_stages(stages, nstage);
    if (m.precision == SITKHIP_PRECISION_F64) launch_ms<double, false>(m, S);
    else launch_ms<float, false>(m, S);
    wait_done(m);
    if (sq) *sq = m.res->moments[0];
    if (valid) *valid = static_cast<size_t>(m.res->moments[1]);
}

// ------------------------------------------- cubic B-spline interpolation

// compiler-rt's __powidf2, the routine Rust's f64::powi lowers to: binary
// exponentiation, one rounding per multiply (b >= 0 here).
__device__ __forceinline__ double powi_nonneg(double a, int b) {
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b >>= 1;
        if (b == 0) break;
        a *= a;
    }
    return r;
}

__global__ void widen_f32(const float *__restrict__ s, double *__restrict__ d, size_t n) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        d[i] = static_cast<double>(s[i]);
}

__global__ void narrow_f64(const double *__restrict__ s, float *__restrict__ d, size_t n) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        d[i] = static_cast<float>(s[i]);
}

// bspline3_filter_1d (interpolator.rs:236-270) on every line of one axis, in
// place, one line (n >= 2 voxels, `stride` apart) per lane: the gain (applied
// as each value is first read: the same one rounding), the mirror-boundary
// causal initial value by the full sum, the causal recursion, the
// anti-causal initial value and recursion. FP64, one rounding per operation
// (the file is compiled without contraction). Lanes take lines in the order
// of the `inner` axis (x for the axes 1 and 2: coalesced).
__global__ void bspline_prefilter(double *vol, int n, long long stride, long long nlines,
                                  long long inner, long long inner_stride, long long outer_stride,
                                  double z) {
    const long long line = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (line >= nlines) return;
    double *p = vol + (line % inner) * inner_stride + (line / inner) * outer_stride;
    const double gain = (1.0 - z) * (1.0 - 1.0 / z);
    const double iz = 1.0 / z;
    double zn = z;
    double z2n = powi_nonneg(z, n - 1);
    double sum = p[0] * gain + z2n * (p[(n - 1) * stride] * gain);
    z2n = z2n * z2n * iz;
    for (int k = 1; k + 1 < n; ++k) {
        sum += (zn + z2n) * (p[k * stride] * gain);
        zn *= z;
        z2n *= iz;
    }
    double prev = sum / (1.0 - zn * zn);
    p[0] = prev;
    for (int k = 1; k < n; ++k) {
        prev = p[k * stride] * gain + z * prev;
        p[k * stride] = prev;
    }
    double next = (z / (z * z - 1.0)) * (z * p[(n - 2) * stride] + prev);
    p[(n - 1) * stride] = next;
    for (int k = n - 2; k >= 0; --k) {
        next = z * (next - p[k * stride]);
        p[k * stride] = next;
    }
}

// MovingImage::from_image_with_interpolator (registration/metric.rs:720-750)
// for the resident metric: the cubic B-spline coefficients of its moving
// volume (bspline_coefficients, interpolator.rs:170-190), FP64 on the
// device, then narrowed to float for the FP32 path. Linear drops them.
void metric_set_interpolator(Metric &m, int interpolator) {
    if (interpolator == SITKHIP_INTERP_LINEAR) {
        m.bspline = false;
        m.Mv.coef = nullptr;
        m.Mv.coef64 = nullptr;
        m.coef32.release();
        m.coef64.release();
        return;
    }
    if (interpolator != SITKHIP_INTERP_BSPLINE)
        fail(SITKHIP_ERR_INVALID_ARGUMENT, "unknown interpolator " + std::to_string(interpolator));
    for (int d = 0; d < 3; ++d)
        if (m.Mv.size[d] == 2)
            fail(SITKHIP_ERR_UNSUPPORTED_GEOMETRY,
                 "cubic B-spline interpolation: moving axis " + std::to_string(d) +
                     " has 2 voxels, where the mirror fold of the taps leaves the volume");
    const size_t n = static_cast<size_t>(m.Mv.size[0]) * m.Mv.size[1] * m.Mv.size[2];
    DeviceBuffer<double> c64(n);
    {
        LaunchScope scope("bspline_prefilter");
        hipLaunchKernelGGL(widen_f32, dim3(grid_for(n, 256, 4096)), dim3(256), 0, backend().stream,
                           m.Mv.vol, c64.ptr, n);
        check_launch("widen_f32");
        const double z = std::sqrt(3.0) - 2.0;
        const long long st[3] = {1, m.Mv.stride1, m.Mv.stride2};
        for (int a = 0; a < 3; ++a) {
            const int len = m.Mv.size[a];
            if (len == 1) continue;
            const int o1 = a == 0 ? 1 : 0, o2 = a == 2 ? 1 : 2;
            const long long inner = m.Mv.size[o1];
            const long long nlines = inner * m.Mv.size[o2];
            const unsigned blocks = static_cast<unsigned>((nlines + 255) / 256);
            hipLaunchKernelGGL(bspline_prefilter, dim3(blocks), dim3(256), 0, backend().stream,
                               c64.ptr, len, st[a], nlines, inner, st[o1], st[o2], z);
            check_launch("bspline_prefilter");
        }
    }
    if (m.precision == SITKHIP_PRECISION_F64) {
        m.coef32.release();
        m.coef64 = std::move(c64);
        m.Mv.coef64 = m.coef64.ptr;
        m.Mv.coef = nullptr;
    } else {
        m.coef32.alloc(n);
        hipLaunchKernelGGL(narrow_f64, dim3(grid_for(n, 256, 4096)), dim3(256), 0, backend().stream,
                           c64.ptr, m.coef32.ptr, n);
        check_launch("narrow_f64");
        m.coef64.release();
        m.Mv.coef = m.coef32.ptr;
        m.Mv.coef64 = nullptr;
    }
    sync();
    m.bspline = true;
}

bool metric_uses_bspline_interpolation(const Metric &m) { return m.bspline; }

}  // namespace sitkhip

// =================================================================== C ABI
using namespace sitkhip;

extern "C" {

int sitkhip_metric_create(int kind, sitkhip_image fixed, const int64_t *fixed_indices,
                          size_t n_indices, sitkhip_mask fixed_mask, sitkhip_image moving,
                          sitkhip_mask moving_mask, int precision, sitkhip_metric *out) {
    return run_guarded([&] {
        auto f = reinterpret_cast<Image *>(fixed);
        auto mv = reinterpret_cast<Image *>(moving);
        if (!f || !mv || !out) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null image or out-pointer");
        *out = reinterpret_cast<sitkhip_metric>(
            metric_create(kind, *f, fixed_indices, n_indices, reinterpret_cast<Mask *>(fixed_mask),
                          *mv, reinterpret_cast<Mask *>(moving_mask), precision));
    });
}

int sitkhip_metric_create_drawn(int kind, sitkhip_image fixed, size_t n_draws, uint64_t seed,
                                sitkhip_image moving, int precision, sitkhip_metric *out) {
    return run_guarded([&] {
        auto f = reinterpret_cast<Image *>(fixed);
        auto mv = reinterpret_cast<Image *>(moving);
        if (!f || !mv || !out) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null image or out-pointer");
        *out = reinterpret_cast<sitkhip_metric>(
            metric_create_drawn(kind, *f, n_draws, seed, *mv, precision));
    });
}

int sitkhip_metric_create_host(int kind, size_t n, const double *fixed_values,
                               const sitkhip_fixed_points *points, const uint8_t *fixed_mask,
                               const sitkhip_moving_geometry *moving, const double *moving_values,
                               int precision, sitkhip_metric *out) {
    return run_guarded([&] {
        if (!points || !moving || !out) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null argument");
        *out = reinterpret_cast<sitkhip_metric>(metric_create_host(
            kind, n, fixed_values, *points, fixed_mask, *moving, moving_values, precision));
    });
}

void sitkhip_metric_free(sitkhip_metric m) { delete reinterpret_cast<Metric *>(m); }

static Metric &metric_ref(sitkhip_metric h) {
    auto m = reinterpret_cast<Metric *>(h);
    if (!m) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null metric");
    return *m;
}

int sitkhip_metric_set_interpolator(sitkhip_metric h, int interpolator) {
    return run_guarded([&] { metric_set_interpolator(metric_ref(h), interpolator); });
}

int sitkhip_metric_bspline_coefficients(sitkhip_metric h, double *out, size_t len) {
    return run_guarded([&] {
        Metric &m = metric_ref(h);
        if (!m.bspline) fail(SITKHIP_ERR_INVALID_ARGUMENT, "the metric interpolates linearly");
        const size_t n = static_cast<size_t>(m.Mv.size[0]) * m.Mv.size[1] * m.Mv.size[2];
        if (!out || len != n)
            fail(SITKHIP_ERR_INVALID_ARGUMENT, "output length " + std::to_string(len) +
                                                   " is not the moving voxel count " +
                                                   std::to_string(n));
        if (m.Mv.coef64) {
            m.coef64.download(out, n);
            sync();
        } else {
            std::vector<float> tmp(n);
            m.coef32.download(tmp.data(), n);
            sync();
            for (size_t i = 0; i < n; ++i) out[i] = static_cast<double>(tmp[i]);
        }
    });
}




int sitkhip_metric_sample_count(sitkhip_metric h, size_t *out) {
    return run_guarded([&] {
        if (!out) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null out-pointer");
        *out = metric_ref(h).nsamples;
    });
}

int sitkhip_metric_volume_bytes(sitkhip_metric h, size_t *out) {
    return run_guarded([&] {
        if (!out) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null out-pointer");
        Metric &m = metric_ref(h);
        *out = m.moving.bytes() + m.mmask.bytes() + m.idx.bytes() + m.fval.bytes() + m.fbin.bytes();
    });
}

int sitkhip_metric_value_ranges(sitkhip_metric h, double fixed_range[2], double moving_range[2]) {
    return run_guarded([&] {
        double fr[2], mr[2];
        metric_value_ranges(metric_ref(h), fr, mr);
        if (fixed_range) { fixed_range[0] = fr[0]; fixed_range[1] = fr[1]; }
        if (moving_range) { moving_range[0] = mr[0]; moving_range[1] = mr[1]; }
    });
}

int sitkhip_mattes_bins_from_ranges(const double fixed_range[2], const double moving_range[2],
                                    size_t bins, sitkhip_mattes_bins *out) {
    return run_guarded([&] {
        if (!fixed_range || !moving_range || !out)
            fail(SITKHIP_ERR_INVALID_ARGUMENT, "null argument");
        mattes_bins_from_ranges(fixed_range, moving_range, bins, out);
    });
}

int sitkhip_mattes_set_bins(sitkhip_metric h, const sitkhip_mattes_bins *bins) {
    return run_guarded([&] {
        if (!bins) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null bins");
        mattes_set_bins(metric_ref(h), *bins);
    });
}

int sitkhip_mattes_joint_histogram(sitkhip_metric h, const double *stages, int nstage,
                                   double *joint, double *marginal, size_t *valid) {
    return run_guarded([&] { mattes_joint_histogram(metric_ref(h), stages, nstage, joint, marginal, valid); });
}

int sitkhip_mattes_derivative_moments(sitkhip_metric h, const double *stages, int nstage,
                                      const double *pratio, double moments[12]) {
    return run_guarded([&] {
        if (!moments) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null moments");
        mattes_derivative_moments(metric_ref(h), stages, nstage, pratio, moments);
    });
}

int sitkhip_mattes_value(sitkhip_metric h, const double *stages, int nstage, double *value,
                         size_t *valid) {
    return run_guarded([&] { mattes_value(metric_ref(h), stages, nstage, value, valid); });
}

int sitkhip_mattes_value_and_moments(sitkhip_metric h, const double *stages, int nstage,
                                     double *value, double moments[12], size_t *valid) {
    return run_guarded(
        [&] { mattes_value_and_moments(metric_ref(h), stages, nstage, value, moments, valid); });
}

int sitkhip_mattes_value_and_moments_repeat(sitkhip_metric h, const double *stages, int nstage,
                                            int repeats, double *value, double moments[12],
                                            size_t *valid) {
    return run_guarded([&] {
        if (repeats < 1) fail(SITKHIP_ERR_INVALID_ARGUMENT, "repeats must be >= 1");
        Metric &m = metric_ref(h);
        for (int r = 0; r < repeats; ++r)
            mattes_value_and_moments(m, stages, nstage, value, moments, valid);
    });
}

int sitkhip_ms_moments(sitkhip_metric h, const double *stages, int nstage, double moments[14]) {
    return run_guarded([&] {
        if (!moments) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null moments");
        ms_moments(metric_ref(h), stages, nstage, moments);
    });
}

int sitkhip_ms_value(sitkhip_metric h, const double *stages, int nstage, double *sq, size_t *valid) {
    return run_guarded([&] { ms_value(metric_ref(h), stages, nstage, sq, valid); });
}

// contract (registration/cuda.rs:381-412).
int sitkhip_contract_ms(const double mom[14], const double *j0, const double *c, int nparams,
                        double *value, double *deriv, size_t *valid) {
    return run_guarded([&] {
        if (!mom || !j0 || !c || nparams <= 0) fail(SITKHIP_ERR_INVALID_ARGUMENT, "bad arguments");
        const double count = mom[13];
        if (valid) *valid = static_cast<size_t>(count);
        if (count == 0.0) {
            if (value) *value = DBL_MAX;
            if (deriv)
                for (int k = 0; k < nparams; ++k) deriv[k] = 0.0;
            return;
        }
        const double inv = 1.0 / count;
        if (deriv) {
            for (int k = 0; k < nparams; ++k) {
                double g = 0.0;
                for (int d = 0; d < 3; ++d) {
                    g += j0[d * nparams + k] * mom[1 + d];
                    for (int e = 0; e < 3; ++e) g += c[(e * 3 + d) * nparams + k] * mom[4 + d * 3 + e];
                }
                deriv[k] = 2.0 * g * inv;
            }
        }
        if (value) *value = mom[0] * inv;
    });
}

// contract_mattes (registration/cuda.rs:505-524).
int sitkhip_contract_mattes(const double mom[12], const double *j0, const double *c, int nparams,
                            double *deriv) {
    return run_guarded([&] {
        if (!mom || !j0 || !c || !deriv || nparams <= 0)
            fail(SITKHIP_ERR_INVALID_ARGUMENT, "bad arguments");
        for (int k = 0; k < nparams; ++k) {
            double g = 0.0;
            for (int d = 0; d < 3; ++d) {
                g += j0[d * nparams + k] * mom[d];
                for (int e = 0; e < 3; ++e) g += c[(e * 3 + d) * nparams + k] * mom[3 + d * 3 + e];
            }
            deriv[k] = g;
        }
    });
}

}  // extern "C"

// ------------------------------------------------------------------ sampling
// draw_samples (crates/sitk/src/registration/metric.rs:101-122, 419-437): the
// host draws, the device is handed the list. strategy 0 None (returns
// SIZE_MAX in *n: the identity selection needs no list), 1 Regular (stride
// ceil(1/pct)), 2 Random (floor(voxels*pct) SplitMix64 draws, next_below =
// u64 % voxels, with replacement). out may be NULL to query the count.
extern "C" int sitkhip_draw_samples(size_t voxels, int strategy, double percentage, uint64_t seed,
                                    int64_t *out, size_t *n) {
    return sitkhip::run_guarded([&] {
        if (!n) sitkhip::fail(SITKHIP_ERR_INVALID_ARGUMENT, "null count out-pointer");
        if (strategy == 0) {
            *n = SIZE_MAX;
            return;
        }
        // the count alone costs no list (strategy 1: ceil(voxels / stride))
        if (!out && strategy == 1 && voxels > 0) {
            sitkhip::check_sampling_fraction(strategy, percentage, voxels);
            size_t stride = static_cast<size_t>(std::ceil(1.0 / percentage));
            if (stride < 1) stride = 1;
            *n = (voxels + stride - 1) / stride;
            return;
        }
        if (!out && strategy == 2 && voxels > 0) {
            sitkhip::check_sampling_fraction(strategy, percentage, voxels);
            *n = static_cast<size_t>(static_cast<double>(voxels) * percentage);
            return;
        }
        const std::vector<int64_t> d = sitkhip::host_draws(voxels, strategy, percentage, seed);
        if (out) std::memcpy(out, d.data(), d.size() * sizeof(int64_t));
        *n = d.size();
    });
}
