// Pyramid placement on the device: shrink geometry, shrink gather, and linear /
// nearest-neighbour resampling through a point map.
//
// Reference (crates/sitk/): src/cuda/ops/pyramid.rs:322-342 (shrink3),
// :355-420 (shrink geometry, ITK ShrinkImageFilter::GenerateOutputInformation,
// src/filters/shrink.rs:51-91), :479-642 (cindex_of, resample_linear3,
// resample_nearest3); the callers are prepare_level / prepare_level_on_device
// (src/registration/method.rs:2416-2569, 2919-2963).
//
// Output voxel o -> output index -> physical point (origin first) -> point-map
// stages replayed in order -> input continuous index -> inside test -> linear
// (zero weights skipped) or round-half-up nearest, default value outside. FP64
// with separately rounded operations (compiled with -ffp-contract=off), so a
// floor or a round downstream sees the reference's bits.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "image.hpp"
#include "metric.hpp"
#include "sampler.hpp"

namespace sitkhip {

void shrunk_geometry(const sitkhip_geometry &g, const size_t f[3], sitkhip_geometry &out,
                     long long offset[3]) {
    double delta[3];
    for (int d = 0; d < 3; ++d) {
        if (f[d] == 0) fail(SITKHIP_ERR_UNSUPPORTED_GEOMETRY, "shrink factor of zero");
        size_t s = g.size[d] / f[d];
        out.size[d] = s < 1 ? 1 : s;
        out.spacing[d] = g.spacing[d] * static_cast<double>(f[d]);
        delta[d] = (static_cast<double>(g.size[d]) - 1.0) / 2.0 -
                   static_cast<double>(f[d]) * (static_cast<double>(out.size[d]) - 1.0) / 2.0;
        double o = std::floor(delta[d] + 0.5);
        offset[d] = o < 0.0 ? 0 : static_cast<long long>(o);
    }
    for (int i = 0; i < 3; ++i) {
        double acc = 0.0;
        for (int j = 0; j < 3; ++j) acc += g.direction[i * 3 + j] * g.spacing[j] * delta[j];
        out.origin[i] = g.origin[i] + acc;
    }
    std::memcpy(out.direction, g.direction, sizeof(out.direction));
}

struct GridMap {
    FixedGrid out;     // output grid: size, fast division, origin, index->physical
    int in_size[3];
    double in_origin[3];
    double in_p2i[9];
};

// cindex_of (pyramid.rs:479-560).
template <typename Map>
__device__ __forceinline__ bool resample_cindex(const GridMap &G, const Map &S, uint32_t o,
                                                double c[3]) {
    double x[3], p[3];
    fixed_point(G.out, o, x);
    map_point(S, x, p);
    double diff[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) diff[d] = p[d] - G.in_origin[d];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        double acc = 0.0;
        acc += G.in_p2i[r * 3 + 0] * diff[0];
        acc += G.in_p2i[r * 3 + 1] * diff[1];
        acc += G.in_p2i[r * 3 + 2] * diff[2];
        c[r] = acc;
    }
    bool inside = true;
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (!(c[d] >= -0.5 && c[d] < static_cast<double>(G.in_size[d]) - 0.5)) inside = false;
    return inside;
}

// Output voxel o of the linear resample (resample_linear3, pyramid.rs:562-600).
template <typename Map>
__device__ __forceinline__ float linear_voxel(const float *__restrict__ in, const GridMap &G,
                                              const Map &S, double dflt, uint32_t o) {
    double c[3];
    double v = dflt;
    if (resample_cindex(G, S, o, c)) {
        int base[3];
        double frac[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double f = floor(c[d]);
            base[d] = static_cast<int>(f);
            frac[d] = c[d] - f;
        }
        double acc = 0.0;
        const long long st1 = G.in_size[0], st2 = static_cast<long long>(G.in_size[0]) * G.in_size[1];
#pragma unroll
        for (int corner = 0; corner < 8; ++corner) {
            double weight = 1.0;
            long long off = 0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int bit = (corner >> d) & 1;
                weight *= bit ? frac[d] : (1.0 - frac[d]);
                const int idx = clampi(base[d] + bit, 0, G.in_size[d] - 1);
                off += d == 0 ? idx : (d == 1 ? idx * st1 : idx * st2);
            }
            if (weight != 0.0) acc += weight * static_cast<double>(in[off]);
        }
        v = acc;
    }
    return static_cast<float>(v);
}

__global__ void resample_linear3(const float *__restrict__ in, float *__restrict__ out, GridMap G,
                                 Stages S, double dflt, uint32_t n_out) {
    const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n_out) return;
    out[o] = linear_voxel(in, G, S, dflt, o);
}

// Many identity placements (prepare_level's resample through
// AffineTransform::identity, method.rs:2462-2466) in one launch: workgroup b
// takes placement job_of[b], output voxels (b - first) * 256 ... of it; the
// per-voxel arithmetic is resample_linear3's with the one identity stage.
struct PlaceJob {
    const float *in;
    float *out;
    GridMap G;
    uint32_t n_out;
    unsigned first_block;
};

__global__ void place_linear_batch(const PlaceJob *__restrict__ jobs, const int *__restrict__ job_of) {
    const PlaceJob &J = jobs[job_of[blockIdx.x]];
    const uint32_t o = (blockIdx.x - J.first_block) * blockDim.x + threadIdx.x;
    if (o >= J.n_out) return;
    OneStage I;
#pragma unroll
    for (int k = 0; k < 12; ++k) I.m[k] = (k == 0 || k == 4 || k == 8) ? 1.0 : 0.0;
    J.out[o] = linear_voxel(J.in, J.G, I, 0.0, o);
}

template <typename T>
__global__ void resample_nearest3(const T *__restrict__ in, T *__restrict__ out, GridMap G,
                                  Stages S, T dflt, uint32_t n_out) {
    const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n_out) return;
    double c[3];
    T v = dflt;
    if (resample_cindex(G, S, o, c)) {
        const long long st1 = G.in_size[0], st2 = static_cast<long long>(G.in_size[0]) * G.in_size[1];
        long long off = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            int i = static_cast<int>(floor(c[d] + 0.5));  // RoundHalfIntegerUp
            i = clampi(i, 0, G.in_size[d] - 1);
            off += d == 0 ? i : (d == 1 ? i * st1 : i * st2);
        }
        v = in[off];
    }
    out[o] = v;
}

__global__ void shrink3(const float *__restrict__ in, float *__restrict__ out, FixedGrid og,
                        long long off0, long long off1, long long off2, long long f0, long long f1,
                        long long f2, long long in0, long long in01, uint32_t n_out) {
    const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n_out) return;
    const uint32_t q0 = fdiv(o, og.div0);
    const uint32_t i0 = o - q0 * og.size[0];
    const uint32_t i2 = fdiv(q0, og.div1);
    const uint32_t i1 = q0 - i2 * og.size[1];
    out[o] = in[(i0 * f0 + off0) + (i1 * f1 + off1) * in0 + (i2 * f2 + off2) * in01];
}

static void index_to_physical(const sitkhip_geometry &g, double out[9]) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) out[r * 3 + c] = g.direction[r * 3 + c] * g.spacing[c];
}

static GridMap grid_map(const sitkhip_geometry &in, const sitkhip_geometry &out) {
    GridMap G{};
    for (int d = 0; d < 3; ++d) {
        if (out.size[d] >= (1u << 31) || in.size[d] >= (1u << 31))
            fail(SITKHIP_ERR_UNSUPPORTED_GEOMETRY, "axis longer than 2^31 voxels");
        G.out.size[d] = static_cast<uint32_t>(out.size[d]);
        G.out.origin[d] = out.origin[d];
        G.in_size[d] = static_cast<int>(in.size[d]);
        G.in_origin[d] = in.origin[d];
    }
    G.out.div0 = make_fastdiv(G.out.size[0]);
    G.out.div1 = make_fastdiv(G.out.size[1]);
    index_to_physical(out, G.out.i2p);
    double i2p[9];
    index_to_physical(in, i2p);
    invert3_host(i2p, G.in_p2i);
    return G;
}

static Stages stages_or_identity(const double *stages, int nstage) {
    Stages S;
    std::memset(&S, 0, sizeof(S));
    if (!stages || nstage == 0) {
        // prepare_level resamples through AffineTransform::identity
        // (method.rs:2462-2466): one I / 0 stage, replayed like any other.
        S.n = 1;
        S.s[0] = S.s[4] = S.s[8] = 1.0;
        return S;
    }
    if (nstage < 1 || nstage > kMaxStages)
        fail(SITKHIP_ERR_POINT_MAP_STAGE_COUNT,
             "the point map has " + std::to_string(nstage) + " stages; the device replays 1..=8");
    S.n = nstage;
    std::memcpy(S.s, stages, sizeof(double) * 12 * nstage);
    return S;
}

Image *resample_linear(const Image &in, const sitkhip_geometry &og, const double *stages,
                       int nstage, double dflt, bool synchronous) {
    validate_geometry(og, "resample output");
    const size_t n = geometry_len(og);
    if (n >= (1ull << 32)) fail(SITKHIP_ERR_UNSUPPORTED_GEOMETRY, "output above 2^32 voxels");
    Stages S = stages_or_identity(stages, nstage);
    GridMap G = grid_map(in.geom, og);
    Image *out = image_like(og);
    {
        LaunchScope scope("resample_linear3");
        hipLaunchKernelGGL(resample_linear3, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256),
                           0, backend().stream, in.buf.ptr, out->buf.ptr, G, S, dflt,
                           static_cast<uint32_t>(n));
        check_launch("resample_linear3");
    }
    if (synchronous) sync();
    return out;
}

Image *resample_linear_into(const Image &in, const sitkhip_geometry &og, PrepBatch &prep) {
    validate_geometry(og, "resample output");
    const size_t n = geometry_len(og);
    if (n >= (1ull << 32)) fail(SITKHIP_ERR_UNSUPPORTED_GEOMETRY, "output above 2^32 voxels");
    (void)grid_map(in.geom, og);  // its refusals, now
    Image *out = image_like(og);
    prep.places.push_back({in.buf.ptr, in.geom, out->buf.ptr, og});
    return out;
}

void enqueue_placements(PrepBatch &prep) {
    if (prep.places.empty()) return;
    std::vector<PlaceJob> jobs;
    std::vector<int> job_of;
    for (const PrepBatch::Place &p : prep.places) {
        PlaceJob J{};
        J.in = p.in;
        J.out = p.out;
        J.G = grid_map(p.in_geom, p.out_geom);
        J.n_out = static_cast<uint32_t>(geometry_len(p.out_geom));
        J.first_block = static_cast<unsigned>(job_of.size());
        job_of.insert(job_of.end(), (J.n_out + 255) / 256, static_cast<int>(jobs.size()));
        jobs.push_back(J);
    }
    if (job_of.empty()) return;
    const PlaceJob *dj = prep.upload(jobs);
    const int *dm = prep.upload(job_of);
    launch_kernel("place_linear_batch", place_linear_batch, dim3(static_cast<unsigned>(job_of.size())),
                  dim3(256), 0, dj, dm);
    check_launch("place_linear_batch");
    prep.places.clear();
}

Mask *resample_nearest_mask(const Mask &in, const sitkhip_geometry &og, const double *stages,
                            int nstage, bool synchronous) {
    validate_geometry(og, "resample output");
    const size_t n = geometry_len(og);
    if (n >= (1ull << 32)) fail(SITKHIP_ERR_UNSUPPORTED_GEOMETRY, "output above 2^32 voxels");
    Stages S = stages_or_identity(stages, nstage);
    GridMap G = grid_map(in.geom, og);
    auto *out = new Mask();
    out->geom = og;
    out->buf.alloc(n);
    {
        LaunchScope scope("resample_nearest3");
        hipLaunchKernelGGL((resample_nearest3<uint8_t>), dim3(static_cast<unsigned>((n + 255) / 256)),
                           dim3(256), 0, backend().stream, in.buf.ptr, out->buf.ptr, G, S,
                           static_cast<uint8_t>(0), static_cast<uint32_t>(n));
        check_launch("resample_nearest3");
    }
    if (synchronous) sync();
    return out;
}

Image *resample_nearest(const Image &in, const sitkhip_geometry &og, const double *stages,
                        int nstage, float dflt) {
    validate_geometry(og, "resample output");
    const size_t n = geometry_len(og);
    if (n >= (1ull << 32)) fail(SITKHIP_ERR_UNSUPPORTED_GEOMETRY, "output above 2^32 voxels");
    Stages S = stages_or_identity(stages, nstage);
    GridMap G = grid_map(in.geom, og);
    Image *out = image_like(og);
    {
        LaunchScope scope("resample_nearest3");
        hipLaunchKernelGGL((resample_nearest3<float>), dim3(static_cast<unsigned>((n + 255) / 256)),
                           dim3(256), 0, backend().stream, in.buf.ptr, out->buf.ptr, G, S, dflt,
                           static_cast<uint32_t>(n));
        check_launch("resample_nearest3");
    }
    sync();
    return out;
}

Image *shrink(const Image &in, const size_t f[3]) {
    sitkhip_geometry og;
    long long off[3];
    shrunk_geometry(in.geom, f, og, off);
    const size_t n = geometry_len(og);
    FixedGrid fg{};
    for (int d = 0; d < 3; ++d) fg.size[d] = static_cast<uint32_t>(og.size[d]);
    fg.div0 = make_fastdiv(fg.size[0]);
    fg.div1 = make_fastdiv(fg.size[1]);
    Image *out = image_like(og);
    {
        LaunchScope scope("shrink3");
        const long long in0 = static_cast<long long>(in.geom.size[0]);
        const long long in01 = in0 * static_cast<long long>(in.geom.size[1]);
        hipLaunchKernelGGL(shrink3, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                           backend().stream, in.buf.ptr, out->buf.ptr, fg, off[0], off[1], off[2],
                           static_cast<long long>(f[0]), static_cast<long long>(f[1]),
                           static_cast<long long>(f[2]), in0, in01, static_cast<uint32_t>(n));
        check_launch("shrink3");
    }
    sync();
    return out;
}

}  // namespace sitkhip

using namespace sitkhip;

extern "C" {

int sitkhip_shrunk_geometry(const sitkhip_geometry *in, const size_t factors[3],
                            sitkhip_geometry *out) {
    return run_guarded([&] {
        if (!in || !factors || !out) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null argument");
        long long off[3];
        shrunk_geometry(*in, factors, *out, off);
    });
}

int sitkhip_shrink(sitkhip_image in, const size_t factors[3], sitkhip_image *out) {
    return run_guarded([&] {
        auto img = reinterpret_cast<Image *>(in);
        if (!img || !factors || !out) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null argument");
        *out = reinterpret_cast<sitkhip_image>(shrink(*img, factors));
    });
}

int sitkhip_resample_linear(sitkhip_image in, const sitkhip_geometry *out_geometry,
                            const double *stages, int nstage, double default_value,
                            sitkhip_image *out) {
    return run_guarded([&] {
        auto img = reinterpret_cast<Image *>(in);
        if (!img || !out_geometry || !out) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null argument");
        *out = reinterpret_cast<sitkhip_image>(
            resample_linear(*img, *out_geometry, stages, nstage, default_value));
    });
}

int sitkhip_resample_nearest(sitkhip_image in, const sitkhip_geometry *out_geometry,
                             const double *stages, int nstage, double default_value,
                             sitkhip_image *out) {
    return run_guarded([&] {
        auto img = reinterpret_cast<Image *>(in);
        if (!img || !out_geometry || !out) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null argument");
        *out = reinterpret_cast<sitkhip_image>(resample_nearest(
            *img, *out_geometry, stages, nstage, static_cast<float>(default_value)));
    });
}

int sitkhip_mask_resample_nearest(sitkhip_mask in, const sitkhip_geometry *out_geometry,
                                  const double *stages, int nstage, sitkhip_mask *out) {
    return run_guarded([&] {
        auto m = reinterpret_cast<Mask *>(in);
        if (!m || !out_geometry || !out) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null argument");
        *out = reinterpret_cast<sitkhip_mask>(
            resample_nearest_mask(*m, *out_geometry, stages, nstage));
    });
}

}  // extern "C"
