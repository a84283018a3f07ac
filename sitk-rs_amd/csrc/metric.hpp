// Host entry points of the resident metrics (defined in metric.hip), for the
// registration driver in the same library.
#pragma once

#include <array>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "image.hpp"

namespace sitkhip {

struct Metric;

// The device work of setting up many metrics, one launch per kind of work
// for all of them (the batched registration driver prepares a level of every
// pair this way instead of some ten launches per pair). Creation calls given a
// PrepBatch (CreateOptions::prep, resample_linear_into, mattes_set_bins) record
// their work here; flush() enqueues each kind for every recorded item at once,
// in dependency order: placements, copies and clears, draws (counted, scanned
// and laid out in ascending voxel order), gathers, value ranges, fixed bins.
// Records hold device addresses and geometry by value (the Image objects may
// be gone by the flush); the device memory they point to must stay allocated
// until the flushed work has run (device_defer_frees while recording). The
// PrepBatch keeps its own descriptors until it is destroyed, which must be
// after the stream has drained.
struct PrepBatch {
    struct Place {  // identity resample of in onto out's grid (the images' buffers and geometry)
        const float *in;
        sitkhip_geometry in_geom;
        float *out;
        sitkhip_geometry out_geom;
    };
    struct Copy { const void *src; void *dst; size_t bytes; };  // src device-readable
    struct Zero { void *dst; size_t bytes; };
    struct Draw { uint64_t voxels, n, seed; uint32_t *out; };
    struct Gather { const uint32_t *idx; long long n; const float *src; float *out; };
    struct Range { const float *v; const uint8_t *mask; long long n; double *out; };  // out[3]
    struct Bins { const float *fval; long long n; int B; double fbs, fnmin; uint8_t *out; };
    std::vector<Place> places;
    std::vector<Copy> copies;
    std::vector<Zero> zeros;
    std::vector<Draw> draws;
    std::vector<Gather> gathers;
    std::vector<Range> ranges;
    std::vector<Bins> bins;
    // enqueues everything recorded so far (on the library stream) and clears
    // the records
    void flush();
    // the flushed launches' descriptors, device copies and their host sources
    std::vector<DeviceBuffer<unsigned char>> keep;
    std::vector<std::vector<unsigned char>> keep_host;
    // a device copy of v, kept with the batch
    template <typename T>
    T *upload(const std::vector<T> &v) {
        std::vector<unsigned char> h(v.size() * sizeof(T));
        if (!h.empty()) std::memcpy(h.data(), v.data(), h.size());
        keep.emplace_back(h.size());
        keep_host.push_back(std::move(h));
        DeviceBuffer<unsigned char> &d = keep.back();
        if (!keep_host.back().empty()) d.upload(keep_host.back().data(), keep_host.back().size());
        return reinterpret_cast<T *>(d.ptr);
    }
};

Metric *metric_create(int kind, const Image &fixed, const int64_t *indices, size_t n_indices,
                      const Mask *fmask, const Image &moving, const Mask *mmask, int precision);
// A metric whose samples are `n_draws` random draws (SplitMix64 from `seed`,
// with replacement) over the fixed grid, drawn and sorted on the device.
// Creation options for the registration driver:
//   defer_ranges : the value ranges are only enqueued (no stream wait); call
//                  metric_finish_ranges after the stream has drained;
//   ranges       : compute them at all (else on the first metric_value_ranges);
//   adopt_fixed / adopt_moving : the caller hands over that image's buffer (it
//                  must be the `fixed` / `moving` argument) instead of a copy.
//   f_i2p / m_p2i : the fixed grid's index->physical matrix / the moving
//                  physical->index matrix as given (row-major), instead of the
//                  ones derived from the images' spacing and direction;
//   points       : explicit sample points (host, n x 3): the fixed image is then
//                  the n gathered sample values as an n x 1 x 1 grid.
struct CreateOptions {
    bool defer_ranges = false;
    bool ranges = true;
    Image *adopt_fixed = nullptr;
    Image *adopt_moving = nullptr;
    const double *f_i2p = nullptr;
    const double *m_p2i = nullptr;
    const double *points = nullptr;
    const Mask *fixed_mask = nullptr;   // metric_create_with: gates the sample list
    const Mask *moving_mask = nullptr;  // metric_create_with / _drawn
    PrepBatch *prep = nullptr;          // record the device work there (no fixed mask)
};
Metric *metric_create_drawn(int kind, const Image &fixed, size_t n_draws, uint64_t seed,
                            const Image &moving, int precision,
                            const CreateOptions &opt = CreateOptions{});
Metric *metric_create_with(int kind, const Image &fixed, const int64_t *indices, size_t n_indices,
                           const Image &moving, int precision, const CreateOptions &opt);
// ResidentMetric::new (mean_squares.rs:201-256): both volumes from host values
// (narrowed to float32 as an upload is), the sample set as FixedPoints.
Metric *metric_create_host(int kind, size_t n, const double *fixed_values,
                           const sitkhip_fixed_points &points, const uint8_t *fixed_mask,
                           const sitkhip_moving_geometry &moving, const double *moving_values,
                           int precision);
// Whether the metric's samples are explicit points (no grid index).
bool metric_has_explicit_points(const Metric &m);
// The moving image's interpolator (SITKHIP_INTERP_*): B-spline computes the
// cubic coefficients of the moving volume on the device.
void metric_set_interpolator(Metric &m, int interpolator);
bool metric_uses_bspline_interpolation(const Metric &m);
void metric_finish_ranges(Metric &m);
void metric_destroy(Metric *m);
size_t metric_sample_count(const Metric &m);
void metric_value_ranges(Metric &m, double fixed_range[2], double moving_range[2]);
void mattes_bins_from_ranges(const double fr[2], const double mr[2], size_t bins,
                             sitkhip_mattes_bins *out);
void mattes_set_bins(Metric &m, const sitkhip_mattes_bins &b, PrepBatch *prep = nullptr);
void mattes_value(Metric &m, const double *stages, int nstage, double *value, size_t *valid);
void mattes_value_and_moments(Metric &m, const double *stages, int nstage, double *value,
                              double moments[12], size_t *valid);
// The per-sample records a Mattes value pass leaves for the derivative pass
// (SampleRec<float> or <double> by the metric's precision), allocated on first use.
void *mattes_sample_records(Metric &m);
void ms_moments(Metric &m, const double *stages, int nstage, double out[14]);
void ms_value(Metric &m, const double *stages, int nstage, double *sq, size_t *valid);
void invert3_host(const double m[9], double out[9]);

// Pyramid placement (pyramid.hip).
void shrunk_geometry(const sitkhip_geometry &g, const size_t f[3], sitkhip_geometry &out,
                     long long offset[3]);
// `synchronous` = false leaves the work enqueued on the library stream (the
// C ABI entry points wait; internal callers that go on enqueueing need not).
Image *resample_linear(const Image &in, const sitkhip_geometry &og, const double *stages,
                       int nstage, double dflt, bool synchronous = true);
// The identity placement of `in` onto grid og (default 0) recorded in `prep`;
// the returned image holds the values once prep has been flushed and run.
Image *resample_linear_into(const Image &in, const sitkhip_geometry &og, PrepBatch &prep);
// Enqueues the recorded placements, one launch for all (pyramid.hip).
void enqueue_placements(PrepBatch &prep);
Image *recursive_gaussian(const Image &in, const double sigma[3], bool synchronous = true);
// A mask placed on grid og by the nearest-neighbour resample, default 0.
Mask *resample_nearest_mask(const Mask &in, const sitkhip_geometry &og, const double *stages,
                            int nstage, bool synchronous = true);
// draw_samples (metric.rs:419-437) on the host: strategy 1 regular, 2 random.
std::vector<int64_t> host_draws(size_t voxels, int strategy, double percentage, uint64_t seed);
std::vector<Image *> recursive_gaussian_batch(const std::vector<const Image *> &in,
                                              const std::vector<std::array<double, 3>> &sigma);

}  // namespace sitkhip
