// Batched evaluation of many resident metrics: one evaluation of every pair of
// a batch -- mean squares and Mattes mixed, pairs of uneven sizes -- as four
// persistent launches (value pass, Mattes tails, derivative pass, per-pair
// fold) captured once into a HIP graph and replayed per optimizer iteration.
//
// This has no counterpart in the reference, whose device metrics evaluate one
// pair per call (src/cuda/ops/mattes.rs:697-841, mean_squares.rs:329-439). Per
// pair the arithmetic is the single-pair kernels' own (the same per-range
// sample loops, metric_core.hpp); only the workgroup-to-samples map differs:
//   * every pair's samples are cut into tiles of `tile` samples; a tile is the
//     unit of work and owns one FP64 partial row, so the per-pair fold over its
//     tiles is a fixed tree (deterministic for a given tiling) and the Mattes
//     histogram is exact (integer atomics) whatever the tiling;
//   * the tile list is cut into 8 contiguous segments, one per XCD (workgroup
//     b runs on XCD b mod 8), so a pair's tiles -- and its moving volume --
//     stay in one XCD's L2;
//   * `grid` persistent workgroups walk their segment with a fixed stride, so
//     one launch covers any number of tiles.
// The point-map stage and an active flag per pair are copied from pinned host
// memory by the graph's first node; results land in pinned, mapped host
// memory. Pairs marked inactive cost one flag test per tile.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "batch.hpp"
#include "common.hpp"
#include "image.hpp"
#include "metric.hpp"
#include "metric_core.hpp"

namespace sitkhip {

constexpr int kBatchMaxBins = 64;  // LDS histogram (B^2 + B) * 8 B <= 33 KB

struct PairDesc {
    const ChainConst *C;
    FixedGrid F;
    MovingVol M;
    SampleSet SS;
    const uint8_t *fbin;
    const float *fval;
    MattesParams P;
    unsigned long long *joint, *marg;
    const double *pratio;
    const float *pratio_f;
    double *pratio_w;
    float *pratio_fw;
    double *partials;  // ntiles rows of 12 (Mattes) or 14 (mean squares)
    void *rec;         // Mattes: per-sample records of the value pass (SampleRec<R>)
    long long tile;
    int kind;
    int ntiles;
};

struct PairControl {
    double stage[12];
    IndexMap X;
    int active;
    int pad;
};

struct Tile {
    int pair;
    int index;
};

struct Segments {
    int off[9];
};

// The tiles of XCD (blockIdx mod 8), walked with stride gridDim/8.
#define FOR_EACH_TILE(seg, t)                                                        \
    for (int t = (seg).off[blockIdx.x % 8u] + static_cast<int>(blockIdx.x / 8u);      \
         t < (seg).off[blockIdx.x % 8u + 1]; t += static_cast<int>(gridDim.x / 8u))

__device__ __forceinline__ void tile_range(const PairDesc &d, const Tile &tl, long long &begin,
                                           long long &end) {
    begin = static_cast<long long>(tl.index) * d.tile;
    end = begin + d.tile;
    if (end > d.SS.n) end = d.SS.n;
}

__device__ __forceinline__ OneStage load_stage(const PairControl &c) {
    OneStage S;
#pragma unroll
    for (int k = 0; k < 12; ++k) S.m[k] = c.stage[k];
    return S;
}

// Pass 1: Mattes tiles scatter into an LDS histogram and fold it into their
// pair's global histogram; mean-squares tiles reduce their 14 moments.
#ifndef SITKHIP_BATCH_WAVES
#define SITKHIP_BATCH_WAVES 0
#endif
#if SITKHIP_BATCH_WAVES > 0
#define SITKHIP_BATCH_OCC __attribute__((amdgpu_waves_per_eu(SITKHIP_BATCH_WAVES)))
#else
#define SITKHIP_BATCH_OCC
#endif
template <typename R>
__global__ __launch_bounds__(kBlock) SITKHIP_BATCH_OCC void batch_pass1(const PairDesc *__restrict__ D,
                                                      const Tile *__restrict__ T, Segments seg,
                                                      const PairControl *__restrict__ C) {
    extern __shared__ unsigned long long s_hist[];
    FOR_EACH_TILE(seg, t) {
        const Tile tl = T[t];
        const PairControl &c = C[tl.pair];
        if (!c.active) continue;
        const PairDesc &d = D[tl.pair];
        const OneStage S = load_stage(c);
        long long begin, end;
        tile_range(d, tl, begin, end);
        if (d.kind == SITKHIP_METRIC_MATTES) {
            const int B = d.P.bins;
            const int nb = B * B + B;
            for (int i = threadIdx.x; i < nb; i += kBlock) s_hist[i] = 0ull;
            __syncthreads();
            hist_range<R, true>(d.F, d.M, S, c.X, d.C, d.SS, d.fbin, d.P, s_hist, s_hist + B * B,
                                static_cast<SampleRec<R> *>(d.rec), begin, end);
            __syncthreads();
            const unsigned copy = blockIdx.x % kHistCopies;
            unsigned long long *gj = d.joint + static_cast<size_t>(copy) * B * B;
            unsigned long long *gm = d.marg + copy * B;
            for (int i = threadIdx.x; i < nb; i += kBlock) {
                const unsigned long long v = s_hist[i];
                if (v) atomicAdd(i < B * B ? gj + i : gm + (i - B * B), v);
            }
        } else {
            double acc[14];
#pragma unroll
            for (int k = 0; k < 14; ++k) acc[k] = 0.0;
            ms_range<R, true>(d.F, d.M, S, c.X, d.C, d.SS, d.fval, begin, end, acc);
            block_reduce_store<14>(acc, d.partials, static_cast<unsigned>(tl.index));
        }
        __syncthreads();  // the next tile reuses the LDS
    }
}

// The Mattes tail of every active Mattes pair, one workgroup each.
__global__ __launch_bounds__(kTail) void batch_tail(const PairDesc *__restrict__ D,
                                                    const int *__restrict__ mattes_pairs,
                                                    const PairControl *__restrict__ C,
                                                    HostResult *__restrict__ res) {
    extern __shared__ unsigned long long s_stage[];
    const int pi = mattes_pairs[blockIdx.x];
    if (!C[pi].active) return;
    const PairDesc &d = D[pi];
    mattes_tail_block(d.joint, d.marg, d.P.bins, d.P.mbs, d.pratio_w, d.pratio_fw, &res[pi].tail,
                      s_stage);
}

// Pass 2: the Mattes derivative moments per tile from the value pass's
// per-sample records, pRatio staged in LDS (kept across consecutive tiles of
// the same pair).
template <typename R>
__global__ __launch_bounds__(kBlock) void batch_pass2(const PairDesc *__restrict__ D,
                                                      const Tile *__restrict__ T, Segments seg,
                                                      const PairControl *__restrict__ C) {
    extern __shared__ unsigned char s_raw[];
    R *s_pr = reinterpret_cast<R *>(s_raw);
    int staged = -1;
    FOR_EACH_TILE(seg, t) {
        const Tile tl = T[t];
        const PairControl &c = C[tl.pair];
        if (!c.active) continue;
        const PairDesc &d = D[tl.pair];
        const int B = d.P.bins;
        if (staged != tl.pair) {
            const R *g = sizeof(R) == 4 ? reinterpret_cast<const R *>(d.pratio_f)
                                        : reinterpret_cast<const R *>(d.pratio);
            for (int i = threadIdx.x; i < B * B; i += kBlock) s_pr[i] = g[i];
            staged = tl.pair;
            __syncthreads();
        }
        long long begin, end;
        tile_range(d, tl, begin, end);
        const SampleRec<R> *rec = static_cast<const SampleRec<R> *>(d.rec);
        RecRound<R> first;
        load_rec_round<R>(d.SS, d.fbin, rec, begin + threadIdx.x, begin, end, first);
        double acc[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) acc[k] = 0.0;
        deriv_records<R>(d.F, d.M, d.C, d.SS, d.fbin, d.P, s_pr, rec, begin, end, first, acc);
        block_reduce_store<12>(acc, d.partials, static_cast<unsigned>(tl.index));
        __syncthreads();
    }
}

// Per pair: its tile partials folded in tile order (lane t sums rows t,
// t + 256, ..., then a fixed shuffle + LDS tree) into the pinned result.
__global__ __launch_bounds__(kBlock) void batch_fold(const PairDesc *__restrict__ D,
                                                     const PairControl *__restrict__ C,
                                                     HostResult *__restrict__ res) {
    __shared__ double fold[14][kBlock / 64];
    const int pi = blockIdx.x;
    if (!C[pi].active) return;
    const PairDesc &d = D[pi];
    const int ns = d.kind == SITKHIP_METRIC_MATTES ? 12 : 14;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    double t[14];
#pragma unroll
    for (int k = 0; k < 14; ++k) t[k] = 0.0;
    for (int r = threadIdx.x; r < d.ntiles; r += kBlock) {
        const double *row = d.partials + static_cast<size_t>(r) * ns;
#pragma unroll
        for (int k = 0; k < 14; ++k)
            if (k < ns) t[k] += row[k];
    }
#pragma unroll
    for (int k = 0; k < 14; ++k) {
        const double x = wave_total_f64(t[k]);
        if (lane == 63) fold[k][wave] = x;
    }
    __syncthreads();
    if (threadIdx.x < 14) {
        double x = 0.0;
        if (threadIdx.x < ns) {
#pragma unroll
            for (int w = 0; w < kBlock / 64; ++w) x += fold[threadIdx.x][w];
        }
        res[pi].moments[threadIdx.x] = x;
    }
}

// ------------------------------------------------------------------ host side
struct BatchEngine {
    std::vector<Metric *> metrics;
    size_t n = 0;
    int precision = SITKHIP_PRECISION_F32;
    BatchTiling tiling;
    bool use_graph = true;

    std::vector<PairDesc> hdesc;
    DeviceBuffer<PairDesc> desc;
    DeviceBuffer<Tile> tiles_all, tiles_mattes;
    Segments seg_all{}, seg_mattes{};
    int ntiles_all = 0, ntiles_mattes = 0;
    DeviceBuffer<int> mattes_pairs;
    int n_mattes = 0;
    DeviceBuffer<double> partials;
    DeviceBuffer<PairControl> ctl;
    PairControl *ctl_host = nullptr;
    HostResult *res = nullptr, *res_dev = nullptr;
    size_t lds_hist = 8, lds_pr = 8, lds_tail = 8;

    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;  // recorded after each launched evaluation

    ~BatchEngine() {
        drop_graph();
        if (done) (void)hipEventDestroy(done);
        if (ctl_host) (void)hipHostFree(ctl_host);
        if (res) (void)hipHostFree(res);
    }
    void drop_graph() {
        if (exec) { (void)hipGraphExecDestroy(exec); exec = nullptr; }
        if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
    }
};

static BatchTiling default_tiling() {
    BatchTiling t;
    t.tile = 4096;
    t.grid = static_cast<unsigned>(std::max(8, backend().cu_count * 4));
    t.grid -= t.grid % 8u;
    return t;
}

// Split `tiles` (pair-major) into 8 contiguous segments of near-equal length.
static Segments segment(int ntiles) {
    Segments s;
    for (int x = 0; x <= 8; ++x) s.off[x] = static_cast<int>((static_cast<long long>(ntiles) * x) / 8);
    return s;
}

static void build_tiles(BatchEngine &b) {
    const size_t T = b.tiling.tile;
    std::vector<Tile> all, mattes;
    size_t rows = 0;
    std::vector<size_t> row_off(b.n);
    for (size_t i = 0; i < b.n; ++i) {
        const Metric &m = *b.metrics[i];
        const int nt = static_cast<int>((m.nsamples + T - 1) / T);
        b.hdesc[i].tile = static_cast<long long>(T);
        b.hdesc[i].ntiles = nt;
        row_off[i] = rows;
        rows += static_cast<size_t>(nt);
        for (int k = 0; k < nt; ++k) {
            all.push_back({static_cast<int>(i), k});
            if (m.kind == SITKHIP_METRIC_MATTES) mattes.push_back({static_cast<int>(i), k});
        }
    }
    if (all.size() >= (1u << 30)) fail(SITKHIP_ERR_INVALID_ARGUMENT, "too many tiles");
    b.partials.alloc(rows * 14);
    for (size_t i = 0; i < b.n; ++i) b.hdesc[i].partials = b.partials.ptr + row_off[i] * 14;
    b.ntiles_all = static_cast<int>(all.size());
    b.ntiles_mattes = static_cast<int>(mattes.size());
    b.tiles_all.alloc(all.size());
    b.tiles_mattes.alloc(mattes.size());
    if (!all.empty()) b.tiles_all.upload(all.data(), all.size());
    if (!mattes.empty()) b.tiles_mattes.upload(mattes.data(), mattes.size());
    b.seg_all = segment(b.ntiles_all);
    b.seg_mattes = segment(b.ntiles_mattes);
    b.desc.upload(b.hdesc.data(), b.n);
    sync();
    b.drop_graph();
}

BatchEngine *batch_create(Metric *const *metrics, size_t n) {
    if (n == 0 || !metrics) fail(SITKHIP_ERR_INVALID_ARGUMENT, "an empty batch");
    if (n > (1u << 20)) fail(SITKHIP_ERR_INVALID_ARGUMENT, "at most 2^20 pairs per batch");
    auto *b = new BatchEngine();
    try {
        b->n = n;
        b->metrics.assign(metrics, metrics + n);
        b->precision = metrics[0] ? metrics[0]->precision : 0;
        b->hdesc.resize(n);
        b->desc.alloc(n);
        std::vector<int> mp;
        int maxB = 0;
        for (size_t i = 0; i < n; ++i) {
            Metric *m = metrics[i];
            if (!m) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null metric in batch");
            if (metric_uses_bspline_interpolation(*m))
                fail(SITKHIP_ERR_INVALID_ARGUMENT,
                     "batched evaluation interpolates linearly; metric " + std::to_string(i) +
                         " uses the cubic B-spline interpolator");
            if (m->precision != b->precision)
                fail(SITKHIP_ERR_INVALID_ARGUMENT, "every metric of a batch has the same precision");
            if (m->SS.n >= (1ll << 31))
                fail(SITKHIP_ERR_INVALID_ARGUMENT, "a batched pair holds fewer than 2^31 samples");
            PairDesc &d = b->hdesc[i];
            std::memset(&d, 0, sizeof(d));
            d.C = m->chain.ptr;
            d.F = m->F;
            d.M = m->Mv;
            d.SS = m->SS;
            d.fval = m->fval.ptr;
            d.kind = m->kind;
            if (m->kind == SITKHIP_METRIC_MATTES) {
                if (!m->bins_set)
                    fail(SITKHIP_ERR_HISTOGRAM_SHAPE,
                         "pair " + std::to_string(i) + ": the Mattes histogram geometry has not been set");
                if (m->P.bins > kBatchMaxBins)
                    fail(SITKHIP_ERR_TOO_MANY_BINS, "a batched Mattes pair holds its histogram in LDS: at most " +
                                                        std::to_string(kBatchMaxBins) + " bins");
                d.fbin = m->fbin.ptr;
                d.P = m->P;
                d.joint = m->joint.ptr;
                d.marg = m->marg.ptr;
                d.pratio = m->pratio.ptr;
                d.pratio_f = m->pratio_f.ptr;
                d.pratio_w = m->pratio.ptr;
                d.pratio_fw = m->pratio_f.ptr;
                d.rec = mattes_sample_records(*m);
                mp.push_back(static_cast<int>(i));
                maxB = std::max(maxB, m->P.bins);
            }
        }
        b->n_mattes = static_cast<int>(mp.size());
        b->mattes_pairs.alloc(mp.size());
        if (!mp.empty()) b->mattes_pairs.upload(mp.data(), mp.size());
        const size_t rsz = b->precision == SITKHIP_PRECISION_F64 ? sizeof(double) : sizeof(float);
        b->lds_hist = std::max<size_t>(8, static_cast<size_t>(maxB * maxB + maxB) * 8);
        b->lds_tail = b->lds_hist;
        b->lds_pr = std::max<size_t>(8, static_cast<size_t>(maxB) * maxB * rsz);
        b->ctl.alloc(n);
        HIP_OK(hipHostMalloc(reinterpret_cast<void **>(&b->ctl_host), n * sizeof(PairControl),
                             hipHostMallocDefault));
        std::memset(b->ctl_host, 0, n * sizeof(PairControl));
        HIP_OK(hipHostMalloc(reinterpret_cast<void **>(&b->res), n * sizeof(HostResult),
                             hipHostMallocMapped | hipHostMallocCoherent));
        std::memset(b->res, 0, n * sizeof(HostResult));
        HIP_OK(hipHostGetDevicePointer(reinterpret_cast<void **>(&b->res_dev), b->res, 0));
        b->tiling = default_tiling();
        build_tiles(*b);
    } catch (...) {
        delete b;
        throw;
    }
    return b;
}

void batch_destroy(BatchEngine *b) { delete b; }
size_t batch_size(const BatchEngine &b) { return b.n; }
BatchTiling batch_tiling(const BatchEngine &b) { return b.tiling; }

void batch_set_tiling(BatchEngine &b, const BatchTiling &t) {
    BatchTiling d = default_tiling();
    BatchTiling nt;
    nt.tile = t.tile ? t.tile : d.tile;
    nt.grid = t.grid ? t.grid : d.grid;
    if (nt.tile < kBlock || nt.tile > (1u << 24))
        fail(SITKHIP_ERR_INVALID_ARGUMENT, "tile must be 256 .. 2^24 samples");
    if (nt.grid < 8 || nt.grid % 8 != 0 || nt.grid > 65536)
        fail(SITKHIP_ERR_INVALID_ARGUMENT, "grid must be a multiple of 8 in 8 .. 65536");
    b.tiling = nt;
    build_tiles(b);
}

// The four launches of one evaluation (plus the control copy), in stream
// order. `profile`: time each launch with events (never under capture).
template <typename R>
static void enqueue(BatchEngine &b, bool profile) {
    hipStream_t st = backend().stream;
    HIP_OK(hipMemcpyAsync(b.ctl.ptr, b.ctl_host, b.n * sizeof(PairControl), hipMemcpyHostToDevice,
                          st));
    // profiled launches carry their events on the dispatch (launch_kernel)
    auto go = [&](const char *name, auto kernel, dim3 grid, dim3 block, size_t shm, auto... args) {
        if (profile) launch_kernel(name, kernel, grid, block, shm, args...);
        else hipLaunchKernelGGL(kernel, grid, block, shm, st, args...);
        check_launch(name);
    };
    const unsigned G = b.tiling.grid;
    if (b.ntiles_all > 0) {
        const unsigned g = std::min<unsigned>(G, static_cast<unsigned>((b.ntiles_all + 7) / 8 * 8));
        go("batch_pass1", batch_pass1<R>, dim3(g), dim3(kBlock), b.lds_hist, b.desc.ptr, b.tiles_all.ptr,
           b.seg_all, b.ctl.ptr);
    }
    if (b.n_mattes > 0)
        go("batch_tail", batch_tail, dim3(b.n_mattes), dim3(kTail), b.lds_tail, b.desc.ptr,
           b.mattes_pairs.ptr, b.ctl.ptr, b.res_dev);
    if (b.ntiles_mattes > 0) {
        const unsigned g = std::min<unsigned>(G, static_cast<unsigned>((b.ntiles_mattes + 7) / 8 * 8));
        go("batch_pass2", batch_pass2<R>, dim3(g), dim3(kBlock), b.lds_pr, b.desc.ptr, b.tiles_mattes.ptr,
           b.seg_mattes, b.ctl.ptr);
    }
    go("batch_fold", batch_fold, dim3(static_cast<unsigned>(b.n)), dim3(kBlock), 0, b.desc.ptr, b.ctl.ptr,
       b.res_dev);
}

static void enqueue_any(BatchEngine &b, bool profile) {
    if (b.precision == SITKHIP_PRECISION_F64) enqueue<double>(b, profile);
    else enqueue<float>(b, profile);
}

static void ensure_graph(BatchEngine &b) {
    if (b.exec) return;
    hipStream_t st = backend().stream;
    HIP_OK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        enqueue_any(b, false);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(st, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    HIP_OK(hipStreamEndCapture(st, &b.graph));
    HIP_OK(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
}

// Enqueues one evaluation of the active pairs (results land in b.res) and
// records b.done after it.
static void enqueue_once(BatchEngine &b) {
    if (b.use_graph && !profiling_enabled()) {
        ensure_graph(b);
        HIP_OK(hipGraphLaunch(b.exec, backend().stream));
    } else if (b.use_graph) {
        ensure_graph(b);
        LaunchScope s("batch_graph");
        HIP_OK(hipGraphLaunch(b.exec, backend().stream));
    } else {
        enqueue_any(b, profiling_enabled());
    }
    if (!b.done) HIP_OK(hipEventCreateWithFlags(&b.done, hipEventDisableTiming));
    HIP_OK(hipEventRecord(b.done, backend().stream));
}

// One evaluation of the active pairs, results left in b.res.
static void run_once(BatchEngine &b) {
    enqueue_once(b);
    sync();
}

static void set_control(BatchEngine &b, const double *stages, const uint8_t *active) {
    if (!stages) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null stages");
    for (size_t i = 0; i < b.n; ++i) {
        const int on = active ? (active[i] ? 1 : 0) : 1;
        b.ctl_host[i].active = on;
        if (!on) continue;  // the kernels skip the pair: its stage is not read
        std::memcpy(b.ctl_host[i].stage, stages + 12 * i, 12 * sizeof(double));
        b.ctl_host[i].X = make_index_map(b.hdesc[i].F, b.hdesc[i].M, stages + 12 * i, 1);
    }
}

void batch_prepare(BatchEngine &b) {
    if (b.use_graph) ensure_graph(b);
}

void batch_launch(BatchEngine &b, const double *stages, const uint8_t *active) {
    set_control(b, stages, active);
    enqueue_once(b);
}

static void read_outputs(BatchEngine &b, std::vector<BatchOutput> &out);

void batch_collect(BatchEngine &b, std::vector<BatchOutput> &out) {
    if (!b.done) fail(SITKHIP_ERR_INVALID_ARGUMENT, "no evaluation was launched");
    HIP_OK(hipEventSynchronize(b.done));
    read_outputs(b, out);
}

void batch_evaluate(BatchEngine &b, const double *stages, const uint8_t *active,
                    std::vector<BatchOutput> &out) {
    set_control(b, stages, active);
    run_once(b);
    read_outputs(b, out);
}

static void read_outputs(BatchEngine &b, std::vector<BatchOutput> &out) {
    out.resize(b.n);
    for (size_t i = 0; i < b.n; ++i) {
        if (!b.ctl_host[i].active) continue;
        const HostResult &r = b.res[i];
        BatchOutput &o = out[i];
        std::memcpy(o.moments, r.moments, sizeof(o.moments));
        const bool f32 = b.precision == SITKHIP_PRECISION_F32;
        if (b.metrics[i]->kind == SITKHIP_METRIC_MATTES) {
            o.value = r.tail.value;
            o.valid = static_cast<size_t>(r.tail.valid);
            o.moments[12] = o.moments[13] = 0.0;
            if (f32) moments_to_physical(b.metrics[i]->F, o.moments, o.moments + 3);
        } else {
            if (f32) moments_to_physical(b.metrics[i]->F, o.moments + 1, o.moments + 4);
            const double count = r.moments[13];
            o.valid = static_cast<size_t>(count);
            o.value = count == 0.0 ? DBL_MAX : r.moments[0] / count;
        }
    }
}

BatchTiling batch_autotune(BatchEngine &b, const double *stages, double *best_ms) {
    LaunchScope scope("batch_autotune");
    set_control(b, stages, nullptr);
    const int cu = std::max(8, backend().cu_count);
    const size_t tiles[] = {2048, 8192};
    const unsigned grids[] = {static_cast<unsigned>(cu * 4), static_cast<unsigned>(cu * 8)};
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    BatchTiling best = b.tiling;
    double best_t = INFINITY;
    try {
        for (size_t tile : tiles) {
            for (unsigned g : grids) {
                BatchTiling t;
                t.tile = tile;
                t.grid = g - g % 8u;
                batch_set_tiling(b, t);
                ensure_graph(b);
                HIP_OK(hipGraphLaunch(b.exec, backend().stream));
                const int reps = 3;
                HIP_OK(hipEventRecord(e0, backend().stream));
                for (int r = 0; r < reps; ++r) HIP_OK(hipGraphLaunch(b.exec, backend().stream));
                HIP_OK(hipEventRecord(e1, backend().stream));
                HIP_OK(hipEventSynchronize(e1));
                float ms = 0.0f;
                HIP_OK(hipEventElapsedTime(&ms, e0, e1));
                const double per = static_cast<double>(ms) / reps;
                if (per < best_t) { best_t = per; best = b.tiling; }
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    batch_set_tiling(b, best);
    if (best_ms) *best_ms = best_t;
    return best;
}

}  // namespace sitkhip

// =================================================================== C ABI
using namespace sitkhip;

extern "C" {

int sitkhip_batch_create(const sitkhip_metric *metrics, size_t n, sitkhip_batch *out) {
    return run_guarded([&] {
        if (!metrics || !out) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null argument");
        for (size_t i = 0; i < n; ++i)
            if (metrics[i] && metric_has_explicit_points(*reinterpret_cast<Metric *>(metrics[i])))
                fail(SITKHIP_ERR_INVALID_ARGUMENT, "metrics over explicit points are evaluated one at a time");
        *out = reinterpret_cast<sitkhip_batch>(
            batch_create(reinterpret_cast<Metric *const *>(metrics), n));
    });
}

void sitkhip_batch_free(sitkhip_batch b) { batch_destroy(reinterpret_cast<BatchEngine *>(b)); }

static BatchEngine &batch_ref(sitkhip_batch h) {
    auto b = reinterpret_cast<BatchEngine *>(h);
    if (!b) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null batch");
    return *b;
}

int sitkhip_batch_set_tiling(sitkhip_batch h, size_t tile_samples, unsigned grid, int use_graph) {
    return run_guarded([&] {
        BatchEngine &b = batch_ref(h);
        BatchTiling t;
        t.tile = tile_samples;
        t.grid = grid;
        batch_set_tiling(b, t);
        b.use_graph = use_graph != 0;
    });
}

int sitkhip_batch_get_tiling(sitkhip_batch h, size_t *tile_samples, unsigned *grid) {
    return run_guarded([&] {
        const BatchTiling t = batch_tiling(batch_ref(h));
        if (tile_samples) *tile_samples = t.tile;
        if (grid) *grid = t.grid;
    });
}

int sitkhip_batch_autotune(sitkhip_batch h, const double *stages, size_t *tile_samples,
                           unsigned *grid, double *ms_per_evaluation) {
    return run_guarded([&] {
        const BatchTiling t = batch_autotune(batch_ref(h), stages, ms_per_evaluation);
        if (tile_samples) *tile_samples = t.tile;
        if (grid) *grid = t.grid;
    });
}

int sitkhip_batch_evaluate(sitkhip_batch h, const double *stages, const uint8_t *active,
                           double *values, double *moments, size_t *valid) {
    return run_guarded([&] {
        BatchEngine &b = batch_ref(h);
        std::vector<BatchOutput> out(b.n);
        batch_evaluate(b, stages, active, out);
        for (size_t i = 0; i < b.n; ++i) {
            if (active && !active[i]) continue;
            if (values) values[i] = out[i].value;
            if (valid) valid[i] = out[i].valid;
            if (moments) std::memcpy(moments + 14 * i, out[i].moments, 14 * sizeof(double));
        }
    });
}

}  // extern "C"
