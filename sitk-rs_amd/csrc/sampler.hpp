// The per-sample chain shared by every metric kernel: fixed voxel -> physical
// point -> point map -> moving continuous index -> mask / inside tests ->
// trilinear value (and gradient) of the moving volume.
//
// Reference chain (crates/sitk/):
//   point      src/registration/scales.rs:259-271 (origin-first fold)
//   point map  src/transform/matrix_offset.rs:102-109, src/core/matrix.rs:9-21
//   index      src/registration/metric.rs:795-807
//   mask gate  src/registration/metric.rs:817-831, src/core/coord.rs:189-191
//   inside     src/transform/interpolator.rs:29-34
//   trilinear  src/transform/interpolator.rs:58-149
//   gradient   src/registration/metric.rs:848-864
//
// Everything up to and including the cell floor is FP64 in the reference's
// operation order (this code is compiled with -ffp-contract=off), so every
// discrete decision -- inside, mask, which cell -- is the reference's own. The
// interpolation itself is templated on the arithmetic type R: double follows the
// reference term for term; float computes weights, value and gradient in FP32
// from the FP64 fraction.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sitkhip {

// Unsigned division by a runtime constant via multiply-high (Granlund &
// Montgomery 1994): q = (umulhi(n, m) + n) >> l for 32-bit n and d >= 1.
struct FastDiv {
    uint32_t d, m, l;
};

inline FastDiv make_fastdiv(uint32_t d) {
    FastDiv f;
    f.d = d;
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    f.l = l;
    f.m = static_cast<uint32_t>(((1ull << 32) * ((1ull << l) - d)) / d + 1);
    return f;
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv &f) {
    uint64_t t = static_cast<uint64_t>(__umulhi(n, f.m)) + n;
    return static_cast<uint32_t>(t >> f.l);
}

constexpr int kMaxStages = 8;

struct Stages {
    int n;
    double s[kMaxStages * 12];  // per stage: matrix row-major [9], offset [3]
};

struct FixedGrid {
    uint32_t size[3];
    FastDiv div0, div1;  // flat -> (i0, i1, i2)
    double origin[3];
    double i2p[9];
    // Explicit points (FixedPoints::Explicit, resident.rs:464-484): sample s
    // sits at pts[3 s .. 3 s + 2] and has no grid voxel; nullptr for a grid.
    // The "grid" is then n x 1 x 1 with origin 0 and an identity i2p, so that
    // index-space moments are the physical ones.
    const double *pts;
};

struct MovingVol {
    const float *vol;
    const uint8_t *mask;  // nullptr: no moving mask
    int size[3];
    long long stride1, stride2;
    double origin[3];
    double p2i[9];
    float p2i_f[9];       // p2i rounded to float (the FP32 path's gradient)
    // Cubic B-spline coefficients of the volume (the moving image's cubic
    // interpolator, interpolator.rs:170-190): float for the FP32 path, double
    // for the FP64 path; nullptr with linear interpolation.
    const float *coef;
    const double *coef64;
};

// The FP64 geometry of the exact chain in device memory, for kernels whose
// fast path does not need it: read there only when a sample falls back to the
// chain, instead of occupying scalar registers for the whole kernel.
struct ChainConst {
    double f_origin[3];
    double f_i2p[9];
    double m_origin[3];
    double m_p2i[9];
};

// VirtualGrid::write_point for flat voxel `flat` (origin first, then the three
// columns in order).
__device__ __forceinline__ void fixed_point(const FixedGrid &F, uint32_t flat, double x[3]) {
    uint32_t q0 = fdiv(flat, F.div0);
    uint32_t i0 = flat - q0 * F.size[0];
    uint32_t i2 = fdiv(q0, F.div1);
    uint32_t i1 = q0 - i2 * F.size[1];
    const double id0 = static_cast<double>(i0), id1 = static_cast<double>(i1),
                 id2 = static_cast<double>(i2);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        double acc = F.origin[r];
        acc += F.i2p[r * 3 + 0] * id0;
        acc += F.i2p[r * 3 + 1] * id1;
        acc += F.i2p[r * 3 + 2] * id2;
        x[r] = acc;
    }
}

// The same point with the geometry read from a ChainConst.
__device__ __forceinline__ void fixed_point(const FixedGrid &F, const ChainConst &C, uint32_t flat,
                                            double x[3]) {
    uint32_t q0 = fdiv(flat, F.div0);
    uint32_t i0 = flat - q0 * F.size[0];
    uint32_t i2 = fdiv(q0, F.div1);
    uint32_t i1 = q0 - i2 * F.size[1];
    const double id0 = static_cast<double>(i0), id1 = static_cast<double>(i1),
                 id2 = static_cast<double>(i2);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        double acc = C.f_origin[r];
        acc += C.f_i2p[r * 3 + 0] * id0;
        acc += C.f_i2p[r * 3 + 1] * id1;
        acc += C.f_i2p[r * 3 + 2] * id2;
        x[r] = acc;
    }
}

// The physical point of sample `flat`: PTS (explicit points), the stored
// point; else the grid voxel's point as above.
template <bool PTS>
__device__ __forceinline__ void sample_point(const FixedGrid &F, const ChainConst &C, uint32_t flat,
                                             double x[3]) {
    if constexpr (PTS) {
        x[0] = F.pts[3 * static_cast<size_t>(flat) + 0];
        x[1] = F.pts[3 * static_cast<size_t>(flat) + 1];
        x[2] = F.pts[3 * static_cast<size_t>(flat) + 2];
    } else {
        fixed_point(F, C, flat, x);
    }
}

// replay_stages: p <- mat_vec(M, p) + t per stage, mat_vec seeded at 0.0.
__device__ __forceinline__ void map_point(const Stages &S, const double x[3], double p[3]) {
    double q0 = x[0], q1 = x[1], q2 = x[2];
    for (int s = 0; s < S.n; ++s) {
        const double *m = S.s + 12 * s;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        a0 += m[0] * q0; a0 += m[1] * q1; a0 += m[2] * q2;
        a1 += m[3] * q0; a1 += m[4] * q1; a1 += m[5] * q2;
        a2 += m[6] * q0; a2 += m[7] * q1; a2 += m[8] * q2;
        q0 = a0 + m[9];
        q1 = a1 + m[10];
        q2 = a2 + m[11];
    }
    p[0] = q0; p[1] = q1; p[2] = q2;
}

// A single stage held in registers (the batched kernels, one stage per pair):
// the same expression as one step of the replay above.
struct OneStage {
    double m[12];
};

__device__ __forceinline__ void map_point(const OneStage &S, const double x[3], double p[3]) {
    const double *m = S.m;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    a0 += m[0] * x[0]; a0 += m[1] * x[1]; a0 += m[2] * x[2];
    a1 += m[3] * x[0]; a1 += m[4] * x[1]; a1 += m[5] * x[2];
    a2 += m[6] * x[0]; a2 += m[7] * x[1]; a2 += m[8] * x[2];
    p[0] = a0 + m[9];
    p[1] = a1 + m[10];
    p[2] = a2 + m[11];
}

// Stages staged in LDS by the calling workgroup (read only where a sample
// takes the exact chain, so the stage does not pin scalar registers for the
// whole kernel): the same replay as map_point(Stages).
struct StageView {
    const double *m;  // n x 12 doubles in LDS
    int n;
};

__device__ __forceinline__ void map_point(const StageView &S, const double x[3], double p[3]) {
    double q0 = x[0], q1 = x[1], q2 = x[2];
    for (int s = 0; s < S.n; ++s) {
        const double *m = S.m + 12 * s;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        a0 += m[0] * q0; a0 += m[1] * q1; a0 += m[2] * q2;
        a1 += m[3] * q0; a1 += m[4] * q1; a1 += m[5] * q2;
        a2 += m[6] * q0; a2 += m[7] * q1; a2 += m[8] * q2;
        q0 = a0 + m[9];
        q1 = a1 + m[10];
        q2 = a2 + m[11];
    }
    p[0] = q0; p[1] = q1; p[2] = q2;
}

// The stages of a metric whose samples are explicit points, and their LDS
// view: the same replay, a distinct type so that kernels are instantiated
// separately for such metrics (every sample on the exact chain).
struct PointStages : Stages {};
struct PointStageView : StageView {};

// Copies the stages into the workgroup's LDS (s_stage: kMaxStages * 12
// doubles); all lanes must call it, and it ends with a barrier.
__device__ __forceinline__ PointStageView stage_to_lds(const PointStages &S, double *s_stage) {
    for (int i = threadIdx.x; i < S.n * 12; i += blockDim.x) s_stage[i] = S.s[i];
    __syncthreads();
    PointStageView v;
    v.m = s_stage;
    v.n = S.n;
    return v;
}
__device__ __forceinline__ StageView stage_to_lds(const Stages &S, double *s_stage) {
    for (int i = threadIdx.x; i < S.n * 12; i += blockDim.x) s_stage[i] = S.s[i];
    __syncthreads();
    return StageView{s_stage, S.n};
}
__device__ __forceinline__ StageView stage_to_lds(const OneStage &S, double *s_stage) {
    for (int i = threadIdx.x; i < 12; i += blockDim.x) s_stage[i] = S.m[i];
    __syncthreads();
    return StageView{s_stage, 1};
}

// The located sample: which cell of the moving volume and where inside it.
struct Cell {
    int base[3];
    double frac[3];
};

// The whole chain fixed index -> physical point -> point map -> moving
// continuous index is affine: c = A idx + b. Composed once per evaluation on the
// host, it gives c in 9 FMAs instead of the chain's ~60 rounded operations. It
// is not bit-identical to the chain (the reference rounds every step), so it is
// trusted only where no discrete decision can differ: every component at least
// kGuard (in voxels) away from an integer (the cell floor) and from a
// half-integer (the inside test and the mask's round-half-up). The two
// evaluations of c differ by far less than kGuard for any geometry below 10^6
// voxels of extent (a few hundred roundings of 2^-53 relative each).
struct IndexMap {
    double A[9];
    double b[3];
};
constexpr double kGuard = 1e-6;

__device__ __forceinline__ void split_index(const FixedGrid &F, uint32_t flat, uint32_t &i0,
                                            uint32_t &i1, uint32_t &i2) {
    const uint32_t q0 = fdiv(flat, F.div0);
    i0 = flat - q0 * F.size[0];
    i2 = fdiv(q0, F.div1);
    i1 = q0 - i2 * F.size[1];
}

// continuous_index + mask_allows + is_inside + floor. Returns false for a
// sample the reference drops (outside the buffer or on a zero mask voxel).
__device__ __forceinline__ bool locate_index(const MovingVol &M, const double c[3], Cell &cell);

__device__ __forceinline__ bool locate(const MovingVol &M, const double p[3], Cell &cell) {
    double c[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        double acc = -0.0;  // Rust's f64 sum seeds at -0.0
        acc += M.p2i[r * 3 + 0] * (p[0] - M.origin[0]);
        acc += M.p2i[r * 3 + 1] * (p[1] - M.origin[1]);
        acc += M.p2i[r * 3 + 2] * (p[2] - M.origin[2]);
        c[r] = acc;
    }
    return locate_index(M, c, cell);
}

// The same with the moving geometry read from a ChainConst.
__device__ __forceinline__ bool locate(const MovingVol &M, const ChainConst &C, const double p[3],
                                       Cell &cell) {
    double c[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        double acc = -0.0;
        acc += C.m_p2i[r * 3 + 0] * (p[0] - C.m_origin[0]);
        acc += C.m_p2i[r * 3 + 1] * (p[1] - C.m_origin[1]);
        acc += C.m_p2i[r * 3 + 2] * (p[2] - C.m_origin[2]);
        c[r] = acc;
    }
    return locate_index(M, c, cell);
}

// mask_allows + is_inside + floor on the continuous index c (which may be
// modified: it is a copy).
__device__ __forceinline__ bool locate_index(const MovingVol &M, const double c_in[3], Cell &cell) {
    double c[3] = {c_in[0], c_in[1], c_in[2]};
    if (M.mask) {
        long long flat = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double r = floor(c[d] + 0.5);  // RoundHalfIntegerUp
            if (!(r >= 0.0) || r >= static_cast<double>(M.size[d])) return false;
            long long ri = static_cast<long long>(r);
            flat += d == 0 ? ri : (d == 1 ? ri * M.stride1 : ri * M.stride2);
        }
        if (M.mask[flat] == 0) return false;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (!(c[d] >= -0.5 && c[d] < static_cast<double>(M.size[d]) - 0.5)) return false;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double f = floor(c[d]);
        cell.base[d] = static_cast<int>(f);
        cell.frac[d] = c[d] - f;
    }
    return true;
}

// The composed-map locate. Returns false when some component of c lies within
// kGuard of a decision boundary: the caller then runs the exact chain.
// Otherwise `ok` and `cell` are what locate() would give (the fraction up to
// rounding, which the FP32 interpolation does not see).
__device__ __forceinline__ bool locate_composed(const MovingVol &M, const IndexMap &X, uint32_t i0,
                                                uint32_t i1, uint32_t i2, bool &ok, Cell &cell) {
    const double d0 = static_cast<double>(i0), d1 = static_cast<double>(i1),
                 d2 = static_cast<double>(i2);
    bool sure = true;
    bool inside = true;
    double c[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        c[r] = fma(X.A[r * 3 + 2], d2, fma(X.A[r * 3 + 1], d1, fma(X.A[r * 3 + 0], d0, X.b[r])));
        const double fl = floor(c[r]);
        const double fr = c[r] - fl;
        // kGuard from every integer and half-integer: 2c at least 2 kGuard
        // from every integer (2c, its rounding and their difference are exact)
        const double t = 2.0 * c[r];
        sure = sure && fabs(t - rint(t)) >= 2.0 * kGuard;
        inside = inside && c[r] >= -0.5 && c[r] < static_cast<double>(M.size[r]) - 0.5;
        cell.base[r] = static_cast<int>(fl);
        cell.frac[r] = fr;
    }
    if (!sure) return false;
    ok = inside;
    if (ok && M.mask) {
        // floor(c + 0.5) of an inside point is within [0, size)
        long long flat = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const long long ri = static_cast<long long>(cell.base[d]) + (cell.frac[d] >= 0.5 ? 1 : 0);
            flat += d == 0 ? ri : (d == 1 ? ri * M.stride1 : ri * M.stride2);
        }
        ok = M.mask[flat] != 0;
    }
    return true;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// The 8 corner values, in corner order (bit d of the corner selects base+1 on
// axis d), with neighbour indices clamped into the volume. A located cell has
// base in [-1, size - 1] on every axis (the inside test is -0.5 <= c <
// size - 0.5), so the clamp of base is max(base, 0) and of base + 1 is
// min(base + 1, size - 1).
// Along x the two corners are read as one 8-byte pair (p, p + 1) with
// p = clamp(base, 0, size - 2): at base = -1 both corners are voxel 0 (the
// pair's first), at base = size - 1 both are voxel size - 1 (its second).
// Offsets are 32-bit: a moving volume holds fewer than 2^31 voxels.
struct alignas(4) FloatPair {
    float a, b;
};

__device__ __forceinline__ void corners(const MovingVol &M, const Cell &cell, float v[8]) {
    const int y0 = max(cell.base[1], 0), y1 = min(cell.base[1] + 1, M.size[1] - 1);
    const int z0 = max(cell.base[2], 0), z1 = min(cell.base[2] + 1, M.size[2] - 1);
    const int s1 = static_cast<int>(M.stride1), s2 = static_cast<int>(M.stride2);
    const int b0 = cell.base[0], n0 = M.size[0];
    if (n0 >= 2) {
        const int p = min(max(b0, 0), n0 - 2);
        const bool both_hi = b0 == n0 - 1;
        const bool both_lo = b0 < 0;
        // non-negative (clamped) offsets below 2^31: unsigned, so the address
        // is base + zero-extended offset (no sign extension per load)
        const unsigned rows[4] = {static_cast<unsigned>(y0 * s1 + z0 * s2 + p),
                                  static_cast<unsigned>(y1 * s1 + z0 * s2 + p),
                                  static_cast<unsigned>(y0 * s1 + z1 * s2 + p),
                                  static_cast<unsigned>(y1 * s1 + z1 * s2 + p)};
        FloatPair q[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) q[k] = *reinterpret_cast<const FloatPair *>(M.vol + rows[k]);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            v[2 * k] = both_hi ? q[k].b : q[k].a;
            v[2 * k + 1] = both_lo ? q[k].a : q[k].b;
        }
    } else {
        // one voxel along x: both x corners are voxel 0
        const int rows[4] = {y0 * s1 + z0 * s2, y1 * s1 + z0 * s2, y0 * s1 + z1 * s2, y1 * s1 + z1 * s2};
#pragma unroll
        for (int k = 0; k < 4; ++k) v[2 * k] = v[2 * k + 1] = M.vol[rows[k]];
    }
}

// Trilinear value and index-space gradient from the in-cell fraction `fr` and
// the 8 corner values. double: the reference's order (interpolator.rs:103-149),
// weight = ((1*w0)*w1)*w2, value += weight*b, grad[j] += (sign*w_without_j)*b,
// every operation rounded. float: the separable form with FMAs.
template <typename R>
struct Trilinear;

template <>
struct Trilinear<double> {
    template <bool GRAD>
    __device__ __forceinline__ static void eval(const double fr[3], const float vf[8],
                                                double &value, double g[3]) {
        const double f0 = fr[0], f1 = fr[1], f2 = fr[2];
        const double u0 = 1.0 - f0, u1 = 1.0 - f1, u2 = 1.0 - f2;
        double v = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
#pragma unroll
        for (int corner = 0; corner < 8; ++corner) {
            const double w0 = (corner & 1) ? f0 : u0;
            const double w1 = (corner & 2) ? f1 : u1;
            const double w2 = (corner & 4) ? f2 : u2;
            const double b = static_cast<double>(vf[corner]);
            const double weight = (w0 * w1) * w2;
            v += weight * b;
            if (GRAD) {
                const double s0 = (corner & 1) ? 1.0 : -1.0;
                const double s1 = (corner & 2) ? 1.0 : -1.0;
                const double s2 = (corner & 4) ? 1.0 : -1.0;
                g0 += s0 * (w1 * w2) * b;
                g1 += s1 * (w0 * w2) * b;
                g2 += s2 * (w0 * w1) * b;
            }
        }
        value = v;
        if (GRAD) { g[0] = g0; g[1] = g1; g[2] = g2; }
    }
};

template <>
struct Trilinear<float> {
    template <bool GRAD>
    __device__ __forceinline__ static void eval(const float fr[3], const float vf[8], float &value,
                                                float g[3]) {
        const float f0 = fr[0], f1 = fr[1], f2 = fr[2];
        const float u0 = 1.0f - f0, u1 = 1.0f - f1, u2 = 1.0f - f2;
        // Separable form: interpolate along axis 0 first, then 1, then 2.
        const float a00 = fmaf(f0, vf[1], u0 * vf[0]);
        const float a10 = fmaf(f0, vf[3], u0 * vf[2]);
        const float a01 = fmaf(f0, vf[5], u0 * vf[4]);
        const float a11 = fmaf(f0, vf[7], u0 * vf[6]);
        const float b0 = fmaf(f1, a10, u1 * a00);
        const float b1 = fmaf(f1, a11, u1 * a01);
        value = fmaf(f2, b1, u2 * b0);
        if (GRAD) {
            // d/dc0: differences along axis 0, weighted over axes 1, 2.
            const float d00 = vf[1] - vf[0], d10 = vf[3] - vf[2];
            const float d01 = vf[5] - vf[4], d11 = vf[7] - vf[6];
            const float e0 = fmaf(f1, d10, u1 * d00);
            const float e1 = fmaf(f1, d11, u1 * d01);
            g[0] = fmaf(f2, e1, u2 * e0);
            // d/dc1: differences along axis 1 of the axis-0 interpolants.
            g[1] = fmaf(f2, a11 - a01, u2 * (a10 - a00));
            // d/dc2
            g[2] = b1 - b0;
        }
    }
};

// Cubic B-spline interpolation from the prefiltered coefficients
// (bspline_value_and_gradient, interpolator.rs:284-350): on each axis the
// taps base-1 .. base+2 of the cell base = floor(c), folded at the edges
// (i < 0 -> -i, then i > n-1 -> 2(n-1) - i; an axis of one voxel folds every
// tap onto 0; axes of two voxels are refused when the coefficients are made,
// the fold leaves them), value weights the cubic B-spline at the in-cell
// fraction, gradient weights the quadratic spline at fraction + 1/2.
__device__ __forceinline__ int bspline_fold(int i, int n) {
    if (n == 1) return 0;
    if (i < 0) i = -i;
    if (i > n - 1) i = 2 * (n - 1) - i;
    return i;
}

// The reference's weight expressions, one rounding per operation.
template <typename R>
__device__ __forceinline__ void bspline_weights(R f, R w[4], R dw[4]) {
    const R w3 = (R(1) / R(6)) * f * f * f;
    const R w0 = R(1) / R(6) + R(0.5) * f * (f - R(1)) - w3;
    const R w2 = f + w0 - R(2) * w3;
    const R w1 = R(1) - w0 - w2 - w3;
    w[0] = w0; w[1] = w1; w[2] = w2; w[3] = w3;
    const R a = f - R(0.5);
    const R d2 = R(0.75) - a * a;
    const R d3 = R(0.5) * (a - d2 + R(1));
    const R d1 = R(1) - d2 - d3;
    dw[0] = -d1; dw[1] = d1 - d2; dw[2] = d2 - d3; dw[3] = d3;
}

template <typename R>
struct CubicBSpline;

// FP64: the reference's 64-tap loop (tap index of axis 0 fastest), weight
// products ((w0 * w1) * w2), every operation rounded.
template <>
struct CubicBSpline<double> {
    template <bool GRAD>
    __device__ __forceinline__ static void eval(const MovingVol &M, const int base[3],
                                                const double fr[3], double &value, double g[3]) {
        double w[3][4], dw[3][4];
        int idx[3][4];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            bspline_weights<double>(fr[d], w[d], dw[d]);
#pragma unroll
            for (int k = 0; k < 4; ++k) idx[d][k] = bspline_fold(base[d] - 1 + k, M.size[d]);
        }
        const long long s1 = M.stride1, s2 = M.stride2;
        double v = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
        for (int t2 = 0; t2 < 4; ++t2) {
            for (int t1 = 0; t1 < 4; ++t1) {
                const double *row = M.coef64 + idx[2][t2] * s2 + idx[1][t1] * s1;
#pragma unroll
                for (int t0 = 0; t0 < 4; ++t0) {
                    const double c = row[idx[0][t0]];
                    const double wp = (w[0][t0] * w[1][t1]) * w[2][t2];
                    v += wp * c;
                    if (GRAD) {
                        g0 += ((dw[0][t0] * w[1][t1]) * w[2][t2]) * c;
                        g1 += ((w[0][t0] * dw[1][t1]) * w[2][t2]) * c;
                        g2 += ((w[0][t0] * w[1][t1]) * dw[2][t2]) * c;
                    }
                }
            }
        }
        value = v;
        if (GRAD) {
            g[0] = g0; g[1] = g1; g[2] = g2;
        }
    }
};

// FP32: the separable form with FMAs (axis 0, then 1, then 2), 16 rows of
// four taps.
template <>
struct CubicBSpline<float> {
    template <bool GRAD>
    __device__ __forceinline__ static void eval(const MovingVol &M, const int base[3],
                                                const float fr[3], float &value, float g[3]) {
        float w[3][4], dw[3][4];
        int ix[4], iy[4], iz[4];
#pragma unroll
        for (int d = 0; d < 3; ++d) bspline_weights<float>(fr[d], w[d], dw[d]);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            ix[k] = bspline_fold(base[0] - 1 + k, M.size[0]);
            iy[k] = bspline_fold(base[1] - 1 + k, M.size[1]) * static_cast<int>(M.stride1);
            iz[k] = bspline_fold(base[2] - 1 + k, M.size[2]) * static_cast<int>(M.stride2);
        }
        float v = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
#pragma unroll
        for (int t2 = 0; t2 < 4; ++t2) {
            float sv = 0.0f, sx = 0.0f, sy = 0.0f;
#pragma unroll
            for (int t1 = 0; t1 < 4; ++t1) {
                const float *row = M.coef + iz[t2] + iy[t1];
                const float c0 = row[ix[0]], c1 = row[ix[1]], c2 = row[ix[2]], c3 = row[ix[3]];
                const float r = fmaf(w[0][3], c3, fmaf(w[0][2], c2, fmaf(w[0][1], c1, w[0][0] * c0)));
                sv = fmaf(w[1][t1], r, sv);
                if (GRAD) {
                    const float rd =
                        fmaf(dw[0][3], c3, fmaf(dw[0][2], c2, fmaf(dw[0][1], c1, dw[0][0] * c0)));
                    sx = fmaf(w[1][t1], rd, sx);
                    sy = fmaf(dw[1][t1], r, sy);
                }
            }
            v = fmaf(w[2][t2], sv, v);
            if (GRAD) {
                gx = fmaf(w[2][t2], sx, gx);
                gy = fmaf(w[2][t2], sy, gy);
                gz = fmaf(dw[2][t2], sv, gz);
            }
        }
        value = v;
        if (GRAD) {
            g[0] = gx; g[1] = gy; g[2] = gz;
        }
    }
};

// The moving image's interpolant at a located cell: trilinear from the 8
// gathered corners, or the cubic B-spline from the coefficients (BSP).
template <typename R, bool BSP>
struct Interp {
    template <bool GRAD>
    __device__ __forceinline__ static void eval(const MovingVol &M, const int base[3], const R fr[3],
                                                const float v[8], R &value, R g[3]) {
        if constexpr (BSP) CubicBSpline<R>::template eval<GRAD>(M, base, fr, value, g);
        else Trilinear<R>::template eval<GRAD>(fr, v, value, g);
    }
};

// Index-space gradient -> physical gradient: g_d = sum_j gi_j * P[j][d].
template <typename R>
__device__ __forceinline__ void physical_gradient(const MovingVol &M, const R gi[3], R gp[3]) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        R acc = static_cast<R>(-0.0);
        acc += gi[0] * static_cast<R>(M.p2i[0 * 3 + d]);
        acc += gi[1] * static_cast<R>(M.p2i[1 * 3 + d]);
        acc += gi[2] * static_cast<R>(M.p2i[2 * 3 + d]);
        gp[d] = acc;
    }
}

// FP32: the same sum with fused multiply-adds (three instructions per axis)
template <>
__device__ __forceinline__ void physical_gradient<float>(const MovingVol &M, const float gi[3],
                                                         float gp[3]) {
#pragma unroll
    for (int d = 0; d < 3; ++d)
        gp[d] = fmaf(gi[2], M.p2i_f[2 * 3 + d], fmaf(gi[1], M.p2i_f[1 * 3 + d], gi[0] * M.p2i_f[0 * 3 + d]));
}

}  // namespace sitkhip
