// Recursive (Deriche/Farneback 4th-order IIR) Gaussian smoothing on the device,
// bit-exact against the reference's host filter.
//
// Reference: crates/sitk/src/filters/recursive_gaussian.rs:171-204 (axis loop),
// :472-528 (filter_line), crates/sitk/src/core/deriche.rs:187-347 (the 20
// coefficients), crates/sitk/src/cuda/ops/pyramid.rs:217-293 (device entry).
//
// Bit-exactness rests on three things, all kept here:
//   1. The recursion runs in FP64 with every multiply and add rounded separately
//      (this file is compiled with -ffp-contract=off) and in the reference's
//      left-to-right term order, one line per lane.
//   2. The coefficients are computed on the host with the platform libm, in the
//      reference's expression order, and handed to the kernel.
//   3. Intermediate axes stay FP64; the only narrowing is (float) of the last
//      smoothed axis' output, which equals the reference's single `as f32`.
//
// Layout: volumes are dimension-0 fastest. A smoothed axis is one launch. Lines
// along axes 1 and 2 are assigned to lanes by their dimension-0 coordinate, so
// every step of the recursion is one coalesced access across a wavefront; axis 0
// lines are contiguous, so there a step touches one cache line per lane and the
// following steps hit it in cache.
#include <hip/hip_runtime.h>

#include <array>
#include <algorithm>
#include <vector>

#include <cmath>
#include <mutex>

#include "common.hpp"
#include "image.hpp"

namespace sitkhip {

// The 20 coefficients in to_array order (deriche.rs:290-296):
// [n0..n3, d1..d4, m1..m4, bn1..bn4, bm1..bm4].
struct IirCoef {
    double c[20];
};

// ----------------------------------------------------------- host: coefficients
// Poles and the ZeroOrder Farneback column (deriche.rs:112-117, 135-140).
namespace {
constexpr double kW1 = 0.6681, kL1 = -1.3932, kW2 = 2.0787, kL2 = -1.3732;
constexpr double kA1 = 1.3530, kB1 = 1.8151, kA2 = -0.3531, kB2 = 0.0902;
}  // namespace

// ZeroOrder coefficients for index-space sigma `sigmad` (deriche.rs:187-280
// with compute_d_coefficients :304-319 and compute_n_coefficients :325-347).
// The same products and sums as the reference, in the same order.
IirCoef zero_order_coefficients(double sigmad) {
    const double cos1 = std::cos(kW1 / sigmad), cos2 = std::cos(kW2 / sigmad);
    const double sin1 = std::sin(kW1 / sigmad), sin2 = std::sin(kW2 / sigmad);
    const double e1 = std::exp(kL1 / sigmad), e2 = std::exp(kL2 / sigmad);

    const double d4 = e1 * e1 * e2 * e2;
    const double d3 = -2.0 * cos1 * e1 * e2 * e2 - 2.0 * cos2 * e2 * e1 * e1;
    const double d2 = 4.0 * cos2 * cos1 * e1 * e2 + e1 * e1 + e2 * e2;
    const double d1 = -2.0 * (e2 * cos2 + e1 * cos1);
    const double sd = 1.0 + d1 + d2 + d3 + d4;

    double n0 = kA1 + kA2;
    double n1 = e2 * (kB2 * sin2 - (kA2 + 2.0 * kA1) * cos2) +
                e1 * (kB1 * sin1 - (kA1 + 2.0 * kA2) * cos1);
    double n2 = ((kA1 + kA2) * cos2 * cos1 - kB1 * cos2 * sin1 - kB2 * cos1 * sin2) * 2.0 * e1 * e2 +
                kA2 * e1 * e1 + kA1 * e2 * e2;
    double n3 = e2 * e1 * e1 * (kB2 * sin2 - kA2 * cos2) + e1 * e2 * e2 * (kB1 * sin1 - kA1 * cos1);
    const double sn_raw = n0 + n1 + n2 + n3;

    // Unity DC gain.
    const double alpha0 = 2.0 * sn_raw / sd - n0;
    n0 = n0 / alpha0;
    n1 = n1 / alpha0;
    n2 = n2 / alpha0;
    n3 = n3 / alpha0;

    // Symmetric anti-causal numerator and the boundary coefficients.
    const double m1 = n1 - d1 * n0, m2 = n2 - d2 * n0, m3 = n3 - d3 * n0, m4 = -d4 * n0;
    const double sn = n0 + n1 + n2 + n3;
    const double sm = m1 + m2 + m3 + m4;

    IirCoef k;
    double *c = k.c;
    c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
    c[4] = d1; c[5] = d2; c[6] = d3; c[7] = d4;
    c[8] = m1; c[9] = m2; c[10] = m3; c[11] = m4;
    c[12] = d1 * sn / sd; c[13] = d2 * sn / sd; c[14] = d3 * sn / sd; c[15] = d4 * sn / sd;
    c[16] = d1 * sm / sd; c[17] = d2 * sm / sd; c[18] = d3 * sm / sd; c[19] = d4 * sm / sd;
    return k;
}

// ------------------------------------------------------------------- device
template <typename T>
__device__ __forceinline__ double widen(T v) { return static_cast<double>(v); }

template <typename T>
__device__ __forceinline__ T narrow(double v);
template <>
__device__ __forceinline__ double narrow<double>(double v) { return v; }
template <>
__device__ __forceinline__ float narrow<float>(double v) { return static_cast<float>(v); }

// The line geometry of one smoothed axis.
struct LineSet {
    long long nlines;
    int len;             // elements per line
    long long stride;    // element stride along the line
    long long n0, n01;   // for decoding a line index
    int axis;
};

__device__ __forceinline__ long long line_start(const LineSet &L, long long l) {
    if (L.axis == 0) return l * L.n0;                                 // l = i1 + i2*n1
    if (L.axis == 1) return (l % L.n0) + (l / L.n0) * L.n01;          // l = i0 + i2*n0
    return l;                                                         // l = i0 + i1*n0
}

// One line per lane, straight from global memory (axes 1 and 2: adjacent lanes
// touch adjacent addresses at every step).
//
// Causal pass: writes outs[k] to tmp. Anti-causal pass: walks back, and writes
// dst[k] = narrow(outs[k] + scratch[k]) as soon as scratch[k] exists -- the
// same single addition the reference's final loop performs (:525-527).
// tmp may alias dst when DstT is double (each element is read before it is
// overwritten, by the same lane).
template <typename SrcT, typename DstT>
__global__ __launch_bounds__(256) void iir_lines(const SrcT *__restrict__ src, double *tmp,
                                                 DstT *dst, LineSet L, IirCoef k) {
    const long long l = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (l >= L.nlines) return;
    const long long s0 = line_start(L, l);
    const long long st = L.stride;
    const int ln = L.len;
    const double *c = k.c;
    const double n0 = c[0], n1 = c[1], n2 = c[2], n3 = c[3];
    const double d1 = c[4], d2 = c[5], d3 = c[6], d4 = c[7];
    const double m1 = c[8], m2 = c[9], m3 = c[10], m4 = c[11];

    // ---- causal (recursive_gaussian.rs:476-493)
    const double v1 = widen(src[s0]);
    const double x1 = widen(src[s0 + st]);
    const double x2 = widen(src[s0 + 2 * st]);
    const double x3 = widen(src[s0 + 3 * st]);
    double o0 = v1 * n0 + v1 * n1 + v1 * n2 + v1 * n3;
    double o1 = x1 * n0 + v1 * n1 + v1 * n2 + v1 * n3;
    double o2 = x2 * n0 + x1 * n1 + v1 * n2 + v1 * n3;
    double o3 = x3 * n0 + x2 * n1 + x1 * n2 + v1 * n3;
    o0 -= v1 * c[12] + v1 * c[13] + v1 * c[14] + v1 * c[15];
    o1 -= o0 * d1 + v1 * c[13] + v1 * c[14] + v1 * c[15];
    o2 -= o1 * d1 + o0 * d2 + v1 * c[14] + v1 * c[15];
    o3 -= o2 * d1 + o1 * d2 + o0 * d3 + v1 * c[15];
    tmp[s0] = o0;
    tmp[s0 + st] = o1;
    tmp[s0 + 2 * st] = o2;
    tmp[s0 + 3 * st] = o3;
    // window: a = data[i-1], b = data[i-2], cc = data[i-3]; p1..p4 = outs[i-1..i-4]
    double a = x3, b = x2, cc = x1;
    double p1 = o3, p2 = o2, p3 = o1, p4 = o0;
    for (int i = 4; i < ln; ++i) {
        const double x = widen(src[s0 + i * st]);
        double o = x * n0 + a * n1 + b * n2 + cc * n3;
        o -= p1 * d1 + p2 * d2 + p3 * d3 + p4 * d4;
        tmp[s0 + i * st] = o;
        cc = b; b = a; a = x;
        p4 = p3; p3 = p2; p2 = p1; p1 = o;
    }

    // ---- anti-causal (recursive_gaussian.rs:495-522) and the roll-in (:525-527)
    const long long e = s0 + static_cast<long long>(ln - 1) * st;
    const double v2 = widen(src[e]);
    const double y1 = widen(src[e - st]);      // data[ln-2]
    const double y2 = widen(src[e - 2 * st]);  // data[ln-3]
    double s_1 = v2 * m1 + v2 * m2 + v2 * m3 + v2 * m4;          // scratch[ln-1]
    double s_2 = v2 * m1 + v2 * m2 + v2 * m3 + v2 * m4;          // data[ln-1] is v2
    double s_3 = y1 * m1 + v2 * m2 + v2 * m3 + v2 * m4;          // scratch[ln-3]
    double s_4 = y2 * m1 + y1 * m2 + v2 * m3 + v2 * m4;          // scratch[ln-4]
    s_1 -= v2 * c[16] + v2 * c[17] + v2 * c[18] + v2 * c[19];
    s_2 -= s_1 * d1 + v2 * c[17] + v2 * c[18] + v2 * c[19];
    s_3 -= s_2 * d1 + s_1 * d2 + v2 * c[18] + v2 * c[19];
    s_4 -= s_3 * d1 + s_2 * d2 + s_1 * d3 + v2 * c[19];
    dst[e] = narrow<DstT>(tmp[e] + s_1);
    dst[e - st] = narrow<DstT>(tmp[e - st] + s_2);
    dst[e - 2 * st] = narrow<DstT>(tmp[e - 2 * st] + s_3);
    dst[e - 3 * st] = narrow<DstT>(tmp[e - 3 * st] + s_4);
    // window for index j: data[j+1..j+4] = (w1, w2, w3, w4); scratch[j+1..j+4] = (q1..q4)
    double w2 = y2, w3 = y1, w4 = v2;    // data[ln-3], data[ln-2], data[ln-1]
    double q1 = s_4, q2 = s_3, q3 = s_2, q4 = s_1;
    for (int j = ln - 5; j >= 0; --j) {
        const double w1 = widen(src[s0 + (j + 1) * st]);  // data[j+1]
        double s = w1 * m1 + w2 * m2 + w3 * m3 + w4 * m4;
        s -= q1 * d1 + q2 * d2 + q3 * d3 + q4 * d4;
        dst[s0 + j * st] = narrow<DstT>(tmp[s0 + j * st] + s);
        w4 = w3; w3 = w2; w2 = w1;
        q4 = q3; q3 = q2; q2 = q1; q1 = s;
    }
}

// Two lanes per line, 32 lines per wavefront: lane l < 32 runs the causal
// recursion of line l forward while lane l + 32 runs the anti-causal recursion
// of the same line backward, in lock-step. At step i the causal lane produces
// outs[i] and the anti-causal lane scratch[n-1-i]. Until the middle (i < H =
// ceil(n/2)) each lane parks its value in LDS; after it, each lane reads its
// partner's value for the same element (produced n-1-i steps earlier) and
// writes the finished output outs[k] + scratch[k]. The causal outputs never
// leave the chip, so a line costs one read of its input per lane and one write
// of its output. Both recursions have the same shape,
//     v = (((x*c0 + a*c1) + b*c2) + cc*c3) - (((p1*d1 + p2*d2) + p3*d3) + p4*d4),
// with (c0..c3) = (n0..n3) for the causal lane and (m1..m4) for the anti-causal
// lane (whose x is data[j+1] for output j), so the loop is lane-uniform.
// Loads run one to two blocks of kPD steps ahead of their use.
constexpr int kPairLines = 32;
#ifndef SITKHIP_IIR_PD
#define SITKHIP_IIR_PD 8
#endif
constexpr int kPD = SITKHIP_IIR_PD;

template <typename SrcT, typename DstT>
__device__ __forceinline__ void iir_pair_body(const SrcT *__restrict__ src, DstT *__restrict__ dst,
                                              const LineSet &L, const IirCoef &k, int H,
                                              unsigned block, double *lds) {
    const int lane = threadIdx.x;
    const int role = lane >> 5;  // 0 causal, 1 anti-causal
    const int li = lane & 31;
    long long line = static_cast<long long>(block) * kPairLines + li;
    const bool act = line < L.nlines;
    if (!act) line = L.nlines - 1;  // mirror a valid line; stores are suppressed
    const long long s0 = line_start(L, line);
    const long long st = L.stride;
    const int n = L.len;
    const double *c = k.c;
    const double d1 = c[4], d2 = c[5], d3 = c[6], d4 = c[7];
    const double c0 = role ? c[8] : c[0];
    const double c1 = role ? c[9] : c[1];
    const double c2 = role ? c[10] : c[2];
    const double c3 = role ? c[11] : c[3];
    double *mine = lds + role * H * kPairLines;
    const double *partner = lds + (1 - role) * H * kPairLines;

    double a, b, cc, p1, p2, p3, p4;
    double v[4];
    if (role == 0) {
        // causal boundary (recursive_gaussian.rs:476-487)
        const double n0 = c[0], n1 = c[1], n2 = c[2], n3 = c[3];
        const double v1 = widen(src[s0]);
        const double x1 = widen(src[s0 + st]);
        const double x2 = widen(src[s0 + 2 * st]);
        const double x3 = widen(src[s0 + 3 * st]);
        double o0 = v1 * n0 + v1 * n1 + v1 * n2 + v1 * n3;
        double o1 = x1 * n0 + v1 * n1 + v1 * n2 + v1 * n3;
        double o2 = x2 * n0 + x1 * n1 + v1 * n2 + v1 * n3;
        double o3 = x3 * n0 + x2 * n1 + x1 * n2 + v1 * n3;
        o0 -= v1 * c[12] + v1 * c[13] + v1 * c[14] + v1 * c[15];
        o1 -= o0 * d1 + v1 * c[13] + v1 * c[14] + v1 * c[15];
        o2 -= o1 * d1 + o0 * d2 + v1 * c[14] + v1 * c[15];
        o3 -= o2 * d1 + o1 * d2 + o0 * d3 + v1 * c[15];
        v[0] = o0; v[1] = o1; v[2] = o2; v[3] = o3;
        a = x3; b = x2; cc = x1;
        p1 = o3; p2 = o2; p3 = o1; p4 = o0;
    } else {
        // anti-causal boundary (recursive_gaussian.rs:496-510)
        const double m1 = c[8], m2 = c[9], m3 = c[10], m4 = c[11];
        const long long e = s0 + static_cast<long long>(n - 1) * st;
        const double v2 = widen(src[e]);
        const double y1 = widen(src[e - st]);
        const double y2 = widen(src[e - 2 * st]);
        double q1 = v2 * m1 + v2 * m2 + v2 * m3 + v2 * m4;
        double q2 = v2 * m1 + v2 * m2 + v2 * m3 + v2 * m4;
        double q3 = y1 * m1 + v2 * m2 + v2 * m3 + v2 * m4;
        double q4 = y2 * m1 + y1 * m2 + v2 * m3 + v2 * m4;
        q1 -= v2 * c[16] + v2 * c[17] + v2 * c[18] + v2 * c[19];
        q2 -= q1 * d1 + v2 * c[17] + v2 * c[18] + v2 * c[19];
        q3 -= q2 * d1 + q1 * d2 + v2 * c[18] + v2 * c[19];
        q4 -= q3 * d1 + q2 * d2 + q1 * d3 + v2 * c[19];
        v[0] = q1; v[1] = q2; v[2] = q3; v[3] = q4;  // scratch[n-1 .. n-4]
        a = y2; b = y1; cc = v2;                     // data[n-3], data[n-2], data[n-1]
        p1 = q4; p2 = q3; p3 = q2; p4 = q1;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) mine[i * kPairLines + li] = v[i];

    // value read at step i sits at A + i*S; output of step i goes to W + i*S
    const long long S = role ? -st : st;
    const long long A = role ? s0 + static_cast<long long>(n) * st : s0;
    const long long W = role ? s0 + static_cast<long long>(n - 1) * st : s0;

    // Loads run one block of kPD steps ahead, into two register blocks used
    // in turn (no copies between them, which would wait for the loads). Every
    // load is unconditional: a step past the line re-reads step n-1's value,
    // which is never used.
    auto load_block = [&](SrcT (&buf)[kPD], int base) {
#pragma unroll
        for (int j = 0; j < kPD; ++j) {
            const int i = base + j < n ? base + j : n - 1;
            buf[j] = src[A + i * S];
        }
    };
    auto recur = [&](SrcT v) {
        const double x = widen(v);
        double o = x * c0 + a * c1 + b * c2 + cc * c3;
        o -= p1 * d1 + p2 * d2 + p3 * d3 + p4 * d4;
        cc = b; b = a; a = x;
        p4 = p3; p3 = p2; p2 = p1; p1 = o;
        return o;
    };
    // the workgroup is this one wave: before the first read of the partner's
    // parked values only this wave's LDS writes need to have landed
    auto lds_sync = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    };
    auto run_block = [&](const SrcT (&buf)[kPD], int base) {
        if (base + kPD <= H) {  // every step parks its value
#pragma unroll
            for (int j = 0; j < kPD; ++j) mine[(base + j) * kPairLines + li] = recur(buf[j]);
        } else if (base >= H && base + kPD <= n) {  // every step finishes an output
            if (base == H) lds_sync();
            double q[kPD];
#pragma unroll
            for (int j = 0; j < kPD; ++j) q[j] = partner[(n - 1 - base - j) * kPairLines + li];
#pragma unroll
            for (int j = 0; j < kPD; ++j) {
                const double o = recur(buf[j]);
                if (act) dst[W + (base + j) * S] = narrow<DstT>(o + q[j]);
            }
        } else {
#pragma unroll
            for (int j = 0; j < kPD; ++j) {
                const int i = base + j;
                if (i < n) {
                    const double o = recur(buf[j]);
                    if (i < H) {
                        mine[i * kPairLines + li] = o;
                    } else {
                        if (i == H) lds_sync();
                        const double q = partner[(n - 1 - i) * kPairLines + li];
                        if (act) dst[W + i * S] = narrow<DstT>(o + q);
                    }
                }
            }
        }
    };
    SrcT bufA[kPD], bufB[kPD];
    load_block(bufA, 4);
    for (int base = 4; base < n; base += 2 * kPD) {
        load_block(bufB, base + kPD);
        run_block(bufA, base);
        load_block(bufA, base + 2 * kPD);
        run_block(bufB, base + kPD);
    }
    // odd n: the middle element was parked by both lanes at step H-1
    if ((n & 1) && role == 0 && act) {
        const int mid = (n - 1) >> 1;
        dst[s0 + static_cast<long long>(mid) * st] =
            narrow<DstT>(mine[mid * kPairLines + li] + partner[mid * kPairLines + li]);
    }
}

template <typename SrcT, typename DstT>
__global__ __launch_bounds__(64) void iir_pair_lines(const SrcT *__restrict__ src,
                                                     DstT *__restrict__ dst, LineSet L, IirCoef k,
                                                     int H) {
    extern __shared__ double lds[];  // [2][H][32]: role, step, line
    iir_pair_body<SrcT, DstT>(src, dst, L, k, H, blockIdx.x, lds);
}

// Many volumes' lines along one axis in one launch: workgroup b takes
// volume vol_of[b], lines (b - first_block) * 32 ... of it, with that
// volume's line geometry and coefficients; the per-line arithmetic is
// iir_pair_lines'.
struct BatchLines {
    const void *src;
    void *dst;
    LineSet L;
    IirCoef k;
    int H;
    unsigned first_block;
};

template <typename SrcT, typename DstT>
__global__ __launch_bounds__(64) void iir_pair_lines_batch(const BatchLines *__restrict__ vols,
                                                           const int *__restrict__ vol_of) {
    extern __shared__ double lds[];
    const BatchLines &v = vols[vol_of[blockIdx.x]];
    iir_pair_body<SrcT, DstT>(static_cast<const SrcT *>(v.src), static_cast<DstT *>(v.dst), v.L, v.k,
                              v.H, blockIdx.x - v.first_block, lds);
}

static LineSet lines_for(const sitkhip_geometry &g, int axis) {
    LineSet L;
    long long n0 = static_cast<long long>(g.size[0]);
    long long n1 = static_cast<long long>(g.size[1]);
    long long n2 = static_cast<long long>(g.size[2]);
    long long n = n0 * n1 * n2;
    L.axis = axis;
    L.len = static_cast<int>(g.size[axis]);
    L.nlines = n / static_cast<long long>(g.size[axis]);
    L.n0 = n0;
    L.n01 = n0 * n1;
    L.stride = axis == 0 ? 1 : (axis == 1 ? n0 : n0 * n1);
    return L;
}

static size_t max_lds_per_block() {
    static size_t v = 0;
    if (!v) {
        int a = 0;
        HIP_OK(hipDeviceGetAttribute(&a, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                     backend().device));
        v = static_cast<size_t>(a);
    }
    return v;
}

// tmp: an FP64 scratch volume the single-lane kernel needs for its causal
// outputs (it may alias dst when DstT is double); the two-lane kernel, used
// whenever its LDS fits, needs none.
template <typename SrcT, typename DstT>
static void launch_axis(const SrcT *src, double *tmp, DstT *dst, const sitkhip_geometry &g, int axis,
                        const IirCoef &k) {
    LineSet L = lines_for(g, axis);
    static const char *names[3] = {"gauss_iir_axis0", "gauss_iir_axis1", "gauss_iir_axis2"};
    const int H = (L.len + 1) / 2;
    const size_t lds = static_cast<size_t>(2) * H * kPairLines * sizeof(double);
    if (L.len >= 8 && lds <= max_lds_per_block()) {
        unsigned blocks = static_cast<unsigned>((L.nlines + kPairLines - 1) / kPairLines);
        if (lds > 64 * 1024)
            HIP_OK(hipFuncSetAttribute(reinterpret_cast<const void *>(&iir_pair_lines<SrcT, DstT>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       static_cast<int>(lds)));
        launch_kernel(names[axis], iir_pair_lines<SrcT, DstT>, dim3(blocks), dim3(64), lds, src, dst, L, k, H);
    } else {
        unsigned blocks = static_cast<unsigned>((L.nlines + 255) / 256);
        launch_kernel(names[axis], iir_lines<SrcT, DstT>, dim3(blocks), dim3(256), 0, src, tmp, dst, L, k);
    }
    check_launch(names[axis]);
}

__global__ void copy_f32(const float *__restrict__ s, float *__restrict__ d, size_t n) {
    size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        d[i] = s[i];
}

// Refusals of recursive_gaussian (recursive_gaussian.rs:171-204): sigma >= 0
// on every axis, and a smoothed axis (sigma > 0) needs at least 4 voxels.
// Returns the smoothed axes in order.
static int smoothed_axes(const sitkhip_geometry &g, const double sigma[3], int axes[3]) {
    for (int d = 0; d < 3; ++d) {
        if (!(sigma[d] >= 0.0))
            fail(SITKHIP_ERR_INVALID_ARGUMENT, "recursive_gaussian: sigma must be >= 0");
    }
    int na = 0;
    for (int d = 0; d < 3; ++d) {
        if (sigma[d] <= 0.0) continue;
        if (g.size[d] < 4)
            fail(SITKHIP_ERR_UNSUPPORTED_GEOMETRY,
                 "recursive_gaussian: axis " + std::to_string(d) + " has " +
                     std::to_string(g.size[d]) + " voxels; the recursion needs at least 4");
        axes[na++] = d;
    }
    return na;
}

// Enqueues the smoothing of `src` (geometry g) into `dst` on the library
// stream. A and B are FP64 volumes of g's voxel count: A is used whenever an
// axis is smoothed, B when two or more are.
static void enqueue_smoothing(const float *src, float *dst, const sitkhip_geometry &g,
                              const double sigma[3], const int axes[3], int na, double *A,
                              double *B) {
    const size_t n = geometry_len(g);
    if (na == 0) {
        launch_kernel("copy_f32", copy_f32, dim3(grid_for(n, 256, 4096)), dim3(256), 0, src, dst, n);
        check_launch("copy_f32");
        return;
    }
    if (na == 1) {
        IirCoef k = zero_order_coefficients(sigma[axes[0]] / g.spacing[axes[0]]);
        launch_axis<float, float>(src, A, dst, g, axes[0], k);
        return;
    }
    // first smoothed axis: f32 -> A (causal outputs land in A itself)
    IirCoef k0 = zero_order_coefficients(sigma[axes[0]] / g.spacing[axes[0]]);
    launch_axis<float, double>(src, A, A, g, axes[0], k0);
    double *cur = A, *other = B;
    for (int a = 1; a < na - 1; ++a) {
        IirCoef k = zero_order_coefficients(sigma[axes[a]] / g.spacing[axes[a]]);
        launch_axis<double, double>(cur, other, other, g, axes[a], k);
        double *t = cur; cur = other; other = t;
    }
    IirCoef kl = zero_order_coefficients(sigma[axes[na - 1]] / g.spacing[axes[na - 1]]);
    launch_axis<double, float>(cur, other, dst, g, axes[na - 1], kl);
}

// recursive_gaussian (pyramid.rs:217-293 / recursive_gaussian.rs:171-204).
Image *recursive_gaussian(const Image &in, const double sigma[3], bool synchronous) {
    int axes[3];
    const int na = smoothed_axes(in.geom, sigma, axes);
    const sitkhip_geometry &g = in.geom;
    size_t n = in.len();
    Image *out = image_like(g);
    try {
        // FP64 intermediates of this call, from the caching allocator: freed
        // at return while the launches are still queued, a block is reused
        // only by later work on the same stream (or held back while work is
        // spread over several streams), so calls on different streams never
        // share one.
        DeviceBuffer<double> A(na >= 1 ? n : 1), B(na >= 2 ? n : 1);
        enqueue_smoothing(in.buf.ptr, out->buf.ptr, g, sigma, axes, na, A.ptr, B.ptr);
        if (synchronous) sync();
    } catch (...) {
        delete out;
        throw;
    }
    return out;
}

// ------------------------------------------------ smoothing without allocation
// The FP64 intermediates of sitkhip_recursive_gaussian_into: two volumes of at
// most kIntoMaxVoxels voxels (2 x 64 MiB), a static device array that every
// device the library runs on holds until the process ends. Host code never
// uses the array's name as an address: the device address of each device's
// copy is looked up once, when the backend is initialised on that device, and
// kept in g_scratch_addr.
constexpr size_t kIntoMaxVoxels = SITKHIP_GAUSSIAN_INTO_MAX_VOXELS;
static __device__ __align__(256) double scratch[2 * kIntoMaxVoxels];
constexpr int kMaxDevices = 64;
static double *g_scratch_addr[kMaxDevices] = {};

void gaussian_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices)
        fail(SITKHIP_ERR_NO_DEVICE, "device index " + std::to_string(device) +
                                        " beyond the scratch table of " +
                                        std::to_string(kMaxDevices) + " devices");
    void *p = nullptr;
    HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_addr[device] = static_cast<double *>(p);
}

// recursive_gaussian into a caller's image through the static scratch: no
// device allocation. Every call on one device shares that device's scratch.
static void recursive_gaussian_into(const Image &in, const double sigma[3], Image &out) {
    int axes[3];
    const int na = smoothed_axes(in.geom, sigma, axes);
    if (!same_geometry(in.geom, out.geom))
        fail(SITKHIP_ERR_INVALID_ARGUMENT,
             "recursive_gaussian_into: the output image's geometry is not the input's");
    if (in.buf.ptr == out.buf.ptr)
        fail(SITKHIP_ERR_INVALID_ARGUMENT,
             "recursive_gaussian_into: the output image is the input image");
    const size_t n = in.len();
    if (n > kIntoMaxVoxels)
        fail(SITKHIP_ERR_INVALID_ARGUMENT,
             "recursive_gaussian_into: " + std::to_string(n) + " voxels exceed the " +
                 std::to_string(kIntoMaxVoxels) + " its scratch holds");
    const int dev = backend().device;
    double *base = (dev >= 0 && dev < kMaxDevices) ? g_scratch_addr[dev] : nullptr;
    if (!base)
        fail(SITKHIP_ERR_DRIVER, "recursive_gaussian_into: scratch not set up on device " +
                                     std::to_string(dev));
    enqueue_smoothing(in.buf.ptr, out.buf.ptr, in.geom, sigma, axes, na, base,
                      base + kIntoMaxVoxels);
    sync();
}

// ----------------------------------------------------------- batched volumes
template <typename SrcT, typename DstT>
static void launch_axis_batch(const std::vector<const SrcT *> &src, const std::vector<DstT *> &dst,
                              const std::vector<const sitkhip_geometry *> &g,
                              const std::vector<IirCoef> &k, int axis) {
    const size_t n = src.size();
    std::vector<BatchLines> hv(n);
    std::vector<int> map;
    int Hmax = 1;
    unsigned blocks = 0;
    for (size_t i = 0; i < n; ++i) {
        BatchLines &v = hv[i];
        v.src = src[i];
        v.dst = dst[i];
        v.L = lines_for(*g[i], axis);
        v.k = k[i];
        v.H = (v.L.len + 1) / 2;
        v.first_block = blocks;
        const unsigned nb = static_cast<unsigned>((v.L.nlines + kPairLines - 1) / kPairLines);
        map.insert(map.end(), nb, static_cast<int>(i));
        blocks += nb;
        Hmax = std::max(Hmax, v.H);
    }
    const size_t lds = static_cast<size_t>(2) * Hmax * kPairLines * sizeof(double);
    DeviceBuffer<BatchLines> dv(n);
    DeviceBuffer<int> dmap(map.size());
    dv.upload(hv.data(), n);
    dmap.upload(map.data(), map.size());
    static const char *names[3] = {"gauss_iir_axis0", "gauss_iir_axis1", "gauss_iir_axis2"};
    if (lds > 64 * 1024)
        HIP_OK(hipFuncSetAttribute(reinterpret_cast<const void *>(&iir_pair_lines_batch<SrcT, DstT>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    launch_kernel(names[axis], iir_pair_lines_batch<SrcT, DstT>, dim3(blocks), dim3(64), lds,
                  static_cast<const BatchLines *>(dv.ptr), static_cast<const int *>(dmap.ptr));
    check_launch(names[axis]);
}

// recursive_gaussian of many volumes, smoothing all three axes of each
// (sigma[i] > 0 on every axis): three launches per group of volumes whose
// FP64 intermediates fit a 1 GiB budget, the per-line arithmetic that of the
// one-volume path (so the results are the same bits). Volumes the paired-lane
// kernel cannot take (an axis shorter than 8 voxels, or lines whose LDS does
// not fit) go through recursive_gaussian one by one. Asynchronous.
std::vector<Image *> recursive_gaussian_batch(const std::vector<const Image *> &in,
                                              const std::vector<std::array<double, 3>> &sigma) {
    const size_t n = in.size();
    std::vector<Image *> out(n, nullptr);
    try {
        std::vector<size_t> group;
        size_t group_bytes = 0;
        auto flush = [&]() {
            if (group.empty()) return;
            std::vector<DeviceBuffer<double>> A(group.size()), B(group.size());
            std::vector<const sitkhip_geometry *> g;
            std::vector<IirCoef> k0, k1, k2;
            std::vector<const float *> s0;
            std::vector<double *> d0, d1;
            std::vector<const double *> s1, s2;
            std::vector<float *> d2;
            for (size_t j = 0; j < group.size(); ++j) {
                const size_t i = group[j];
                const sitkhip_geometry &gi = in[i]->geom;
                A[j].alloc(in[i]->len());
                B[j].alloc(in[i]->len());
                out[i] = image_like(gi);
                g.push_back(&gi);
                k0.push_back(zero_order_coefficients(sigma[i][0] / gi.spacing[0]));
                k1.push_back(zero_order_coefficients(sigma[i][1] / gi.spacing[1]));
                k2.push_back(zero_order_coefficients(sigma[i][2] / gi.spacing[2]));
                s0.push_back(in[i]->buf.ptr);
                d0.push_back(A[j].ptr);
                s1.push_back(A[j].ptr);
                d1.push_back(B[j].ptr);
                s2.push_back(B[j].ptr);
                d2.push_back(out[i]->buf.ptr);
            }
            launch_axis_batch<float, double>(s0, d0, g, k0, 0);
            launch_axis_batch<double, double>(s1, d1, g, k1, 1);
            launch_axis_batch<double, float>(s2, d2, g, k2, 2);
            group.clear();
            group_bytes = 0;
        };
        for (size_t i = 0; i < n; ++i) {
            const sitkhip_geometry &gi = in[i]->geom;
            bool ok = true;
            for (int d = 0; d < 3; ++d) {
                if (!(sigma[i][d] > 0.0) || gi.size[d] < 8) ok = false;
                const size_t lds = static_cast<size_t>(2) * ((gi.size[d] + 1) / 2) * kPairLines * sizeof(double);
                if (lds > max_lds_per_block()) ok = false;
            }
            if (!ok) {
                out[i] = recursive_gaussian(*in[i], sigma[i].data(), false);
                continue;
            }
            const size_t bytes = in[i]->len() * 2 * sizeof(double);
            if (!group.empty() && group_bytes + bytes > (static_cast<size_t>(1) << 30)) flush();
            group.push_back(i);
            group_bytes += bytes;
        }
        flush();
    } catch (...) {
        for (Image *p : out) delete p;
        throw;
    }
    return out;
}
}  // namespace sitkhip

extern "C" int sitkhip_recursive_gaussian_batch(size_t n, const sitkhip_image *in,
                                                const double *sigma, sitkhip_image *out) {
    return sitkhip::run_guarded([&] {
        if (n == 0) return;
        if (!in || !sigma || !out) sitkhip::fail(SITKHIP_ERR_INVALID_ARGUMENT, "null array");
        std::vector<const sitkhip::Image *> imgs(n);
        std::vector<std::array<double, 3>> sig(n);
        for (size_t i = 0; i < n; ++i) {
            imgs[i] = reinterpret_cast<const sitkhip::Image *>(in[i]);
            if (!imgs[i]) sitkhip::fail(SITKHIP_ERR_INVALID_ARGUMENT, "null image");
            for (int d = 0; d < 3; ++d) {
                sig[i][d] = sigma[3 * i + d];
                if (!(sig[i][d] >= 0.0))
                    sitkhip::fail(SITKHIP_ERR_INVALID_ARGUMENT, "recursive_gaussian: sigma must be >= 0");
            }
        }
        std::vector<sitkhip::Image *> res = sitkhip::recursive_gaussian_batch(imgs, sig);
        sitkhip::sync();
        for (size_t i = 0; i < n; ++i) out[i] = reinterpret_cast<sitkhip_image>(res[i]);
    });
}

extern "C" int sitkhip_recursive_gaussian_into(sitkhip_image in, const double sigma[3],
                                               sitkhip_image out) {
    return sitkhip::run_guarded([&] {
        auto src = reinterpret_cast<sitkhip::Image *>(in);
        auto dst = reinterpret_cast<sitkhip::Image *>(out);
        if (!src || !sigma || !dst)
            sitkhip::fail(SITKHIP_ERR_INVALID_ARGUMENT, "null image or sigma");
        sitkhip::recursive_gaussian_into(*src, sigma, *dst);
    });
}

extern "C" int sitkhip_recursive_gaussian(sitkhip_image in, const double sigma[3],
                                          sitkhip_image *out) {
    return sitkhip::run_guarded([&] {
        auto img = reinterpret_cast<sitkhip::Image *>(in);
        if (!img || !sigma || !out)
            sitkhip::fail(SITKHIP_ERR_INVALID_ARGUMENT, "null image, sigma or out-pointer");
        *out = reinterpret_cast<sitkhip_image>(sitkhip::recursive_gaussian(*img, sigma, true));
    });
}
