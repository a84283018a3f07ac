// Seeded synthetic volumes on the device: a sum of 3-D Gaussian blobs plus
// voxel noise, the generator of the benchmark pairs (the reference has none).
// The blob parameters are drawn by the caller; the noise is a counter-based
// draw per voxel (SplitMix64 of seed and voxel index, Box-Muller), so a volume
// is reproducible bit for bit on one device and independent of launch shape.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "common.hpp"
#include "image.hpp"

namespace sitkhip {

struct Blob {
    float c[3];
    float amp;
    float inv2s2;  // 1 / (2 sigma^2)
    float reach;   // 4.5 sigma: no contribution beyond it along any axis
};

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void synth_blobs(float *__restrict__ out, unsigned n0, unsigned n1, unsigned n2,
                            double o0, double o1, double o2, double s0, double s1, double s2,
                            const Blob *__restrict__ blobs, int nblobs, float noise,
                            uint64_t seed) {
    extern __shared__ Blob s_blobs[];
    for (int i = threadIdx.x; i < nblobs; i += blockDim.x) s_blobs[i] = blobs[i];
    __syncthreads();
    const size_t n = static_cast<size_t>(n0) * n1 * n2;
    for (size_t v = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
         v += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const unsigned i0 = static_cast<unsigned>(v % n0);
        const unsigned i1 = static_cast<unsigned>((v / n0) % n1);
        const unsigned i2 = static_cast<unsigned>(v / (static_cast<size_t>(n0) * n1));
        const float x0 = static_cast<float>(o0 + s0 * i0);
        const float x1 = static_cast<float>(o1 + s1 * i1);
        const float x2 = static_cast<float>(o2 + s2 * i2);
        float acc = 0.0f;
        for (int b = 0; b < nblobs; ++b) {
            const Blob &B = s_blobs[b];
            const float d0 = x0 - B.c[0], d1 = x1 - B.c[1], d2 = x2 - B.c[2];
            if (fabsf(d0) > B.reach || fabsf(d1) > B.reach || fabsf(d2) > B.reach) continue;
            acc += B.amp * __expf(-(d0 * d0 + d1 * d1 + d2 * d2) * B.inv2s2);
        }
        if (noise > 0.0f) {
            const uint64_t h1 = mix64(seed ^ (2 * v));
            const uint64_t h2 = mix64(seed ^ (2 * v + 1));
            const float u1 = (static_cast<float>(h1 >> 40) + 0.5f) * (1.0f / 16777216.0f);
            const float u2 = static_cast<float>(h2 >> 40) * (1.0f / 16777216.0f);
            acc += noise * sqrtf(-2.0f * __logf(u1)) * __cosf(6.2831853f * u2);
        }
        out[v] = acc;
    }
}

}  // namespace sitkhip

using namespace sitkhip;

extern "C" int sitkhip_synth_blobs(const sitkhip_geometry *geometry, int nblobs,
                                   const double *centres, const double *amplitudes,
                                   const double *sigmas, double noise_sigma, uint64_t seed,
                                   sitkhip_image *out) {
    return run_guarded([&] {
        if (!geometry || !out || nblobs < 0 || (nblobs > 0 && (!centres || !amplitudes || !sigmas)))
            fail(SITKHIP_ERR_INVALID_ARGUMENT, "null argument");
        if (nblobs > 2048) fail(SITKHIP_ERR_INVALID_ARGUMENT, "at most 2048 blobs");
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                if (geometry->direction[r * 3 + c] != (r == c ? 1.0 : 0.0))
                    fail(SITKHIP_ERR_UNSUPPORTED_GEOMETRY, "the generator takes identity directions");
        std::vector<Blob> hb(nblobs > 0 ? nblobs : 1);
        for (int b = 0; b < nblobs; ++b) {
            for (int d = 0; d < 3; ++d) hb[b].c[d] = static_cast<float>(centres[3 * b + d]);
            hb[b].amp = static_cast<float>(amplitudes[b]);
            if (!(sigmas[b] > 0.0)) fail(SITKHIP_ERR_INVALID_ARGUMENT, "blob sigma must be positive");
            hb[b].inv2s2 = static_cast<float>(1.0 / (2.0 * sigmas[b] * sigmas[b]));
            hb[b].reach = static_cast<float>(4.5 * sigmas[b]);
        }
        Image *img = image_like(*geometry);
        try {
            DeviceBuffer<Blob> db(hb.size());
            db.upload(hb.data(), hb.size());
            const sitkhip_geometry &g = *geometry;
            const size_t n = img->len();
            LaunchScope scope("synth_blobs");
            hipLaunchKernelGGL(synth_blobs, dim3(grid_for(n, 256, 16384)), dim3(256),
                               sizeof(Blob) * static_cast<size_t>(nblobs > 0 ? nblobs : 1),
                               backend().stream, img->buf.ptr, static_cast<unsigned>(g.size[0]),
                               static_cast<unsigned>(g.size[1]), static_cast<unsigned>(g.size[2]),
                               g.origin[0], g.origin[1], g.origin[2], g.spacing[0], g.spacing[1],
                               g.spacing[2], db.ptr, nblobs, static_cast<float>(noise_sigma), seed);
            check_launch("synth_blobs");
            sync();
        } catch (...) {
            delete img;
            throw;
        }
        *out = reinterpret_cast<sitkhip_image>(img);
    });
}
