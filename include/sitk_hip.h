/*
 * sitk_hip.h -- C ABI of libsitk_hip, an MI355X (gfx950) implementation of the
 * registration metric-evaluation path of sitk-rs (physwkim/sitk-rs).
 *
 * The surface mirrors the reference's device backend, `crate::cuda`
 * (crates/sitk/src/cuda/mod.rs:35-56): device-resident images and masks with an
 * upload / to_host residency contract, the bit-exact recursive Gaussian, and
 * resident mean-squares and Mattes metrics that are built once per pyramid level
 * and evaluated against any number of point maps. Citations are relative to
 * crates/sitk/ of the reference.
 *
 * Conventions (the reference's, src/cuda/mod.rs:3-18, buffer.rs:12-149):
 *   - Every function returns an int status (SITKHIP_OK = 0 or a sitkhip_status
 *     code mirroring a CudaError variant, src/cuda/error.rs:11-108); the message
 *     of the last failure on the calling thread is sitkhip_last_error().
 *   - Every operation is synchronous: it has finished on the device when it
 *     returns, as the reference syncs its one stream before returning.
 *   - Handles own device memory; *_free releases it. A resident metric takes
 *     private device-to-device copies of its images and masks, so the caller may
 *     free them after creation (buffer.rs:143-149, mean_squares.rs:252-253).
 *   - Inputs are validated on the host before any launch, and refusals are named
 *     (stage count 1..8, sample index in grid, bins >= 5, resident.rs:598-703).
 *   - Volumes are 3-D, dimension-0 fastest, float32 on the device
 *     (image.rs:28-56). Geometry is size/spacing/origin/direction (row-major).
 *   - One process-wide backend on one device (backend.rs:154-160).
 */
#ifndef SITK_HIP_H
#define SITK_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
/* Mirrors CudaError (src/cuda/error.rs:11-108) plus the registration
 * refusals the resident Mattes constructor raises (registration/mattes.rs:152-168). */
typedef enum {
    SITKHIP_OK = 0,
    SITKHIP_ERR_NO_DEVICE = 1,              /* CudaError::NoDevice */
    SITKHIP_ERR_DRIVER = 2,                 /* CudaError::Driver (HIP runtime) */
    SITKHIP_ERR_UNSUPPORTED_PIXEL_TYPE = 3, /* CudaError::UnsupportedPixelType */
    SITKHIP_ERR_UNSUPPORTED_GEOMETRY = 4,   /* CudaError::UnsupportedGeometry */
    SITKHIP_ERR_DEGENERATE_INPUT = 5,       /* CudaError::DegenerateInput */
    SITKHIP_ERR_MASKED_EXPLICIT_POINTS = 6, /* CudaError::MaskedExplicitPoints */
    SITKHIP_ERR_SAMPLE_INDEX_OUT_OF_GRID = 7, /* CudaError::SampleIndexOutOfGrid */
    SITKHIP_ERR_POINT_MAP_STAGE_COUNT = 8,  /* CudaError::PointMapStageCount */
    SITKHIP_ERR_HISTOGRAM_SHAPE = 9,        /* CudaError::HistogramShape */
    SITKHIP_ERR_NOT_SCALAR = 10,            /* CudaError::NotScalar */
    SITKHIP_ERR_INVALID_ARGUMENT = 11,      /* null handle/pointer, length mismatch */
    SITKHIP_ERR_TOO_FEW_BINS = 12,          /* RegistrationError::TooFewHistogramBins */
    SITKHIP_ERR_CONSTANT_INTENSITY = 13,    /* RegistrationError::ConstantIntensity */
    SITKHIP_ERR_TOO_MANY_BINS = 14,         /* bins above SITKHIP_MAX_BINS */
    SITKHIP_ERR_NOT_BUILT = 15,             /* library has no code object for this device */
    SITKHIP_ERR_MASK_SIZE_MISMATCH = 16     /* RegistrationError::MaskSizeMismatch */
} sitkhip_status;

const char *sitkhip_status_name(int status);
/* Message of the last failure on this thread ("" if none). */
const char *sitkhip_last_error(void);

/* --------------------------------------------------------------- backend */
/* backend() (src/cuda/backend.rs:154-160): select and initialise `device`
 * once per process. Idempotent; a second call with another device fails. */
int sitkhip_backend_init(int device);
/* The HIP stream every operation of this library is enqueued on, as a
 * hipStream_t cast to void* (for event timing by a caller). */
int sitkhip_backend_stream(void **stream);
int sitkhip_device_name(char *buf, size_t len);

/* ------------------------------------------------------------ pixel types */
/* The scalar PixelId variants the upload cast accepts (image.rs:93-106). */
typedef enum {
    SITKHIP_UINT8 = 0, SITKHIP_INT8 = 1, SITKHIP_UINT16 = 2, SITKHIP_INT16 = 3,
    SITKHIP_UINT32 = 4, SITKHIP_INT32 = 5, SITKHIP_UINT64 = 6, SITKHIP_INT64 = 7,
    SITKHIP_FLOAT32 = 8, SITKHIP_FLOAT64 = 9
} sitkhip_pixel_type;

/* Geometry {size, spacing, origin, direction} (src/cuda/image.rs:122-155). */
typedef struct {
    size_t size[3];
    double spacing[3];
    double origin[3];
    double direction[9];
} sitkhip_geometry;

/* ----------------------------------------------------------------- images */
typedef struct sitkhip_image_s *sitkhip_image;

/* DeviceImage::upload (image.rs:182-268): copy `len` = prod(size) voxels of
 * `pixel_type` and cast each on the device as (float)(double)x. */
int sitkhip_image_upload(int pixel_type, const void *data, size_t len,
                         const sitkhip_geometry *geometry, sitkhip_image *out);
/* DeviceImage::filled (image.rs:305-331). */
int sitkhip_image_filled(const sitkhip_geometry *geometry, float value, sitkhip_image *out);
/* DeviceImage::to_host (image.rs:269-284): `len` must equal the voxel count. */
int sitkhip_image_to_host(sitkhip_image img, float *out, size_t len);
int sitkhip_image_geometry(sitkhip_image img, sitkhip_geometry *out);
size_t sitkhip_image_len(sitkhip_image img);
/* A process-unique identity (image.rs:364-368). */
uint64_t sitkhip_image_id(sitkhip_image img);
/* Device address of the float32 voxels (valid until sitkhip_image_free). */
const float *sitkhip_image_device_ptr(sitkhip_image img);
void sitkhip_image_free(sitkhip_image img);

/* ------------------------------------------------------------------ masks */
typedef struct sitkhip_mask_s *sitkhip_mask;

/* DeviceMask::upload (mask.rs:93-117): nonzero is inside. */
int sitkhip_mask_upload(const uint8_t *data, size_t len, const sitkhip_geometry *geometry,
                        sitkhip_mask *out);
/* DeviceMask::from_device_image (mask.rs:118-150): voxel != 0.0. */
int sitkhip_mask_from_image(sitkhip_image img, sitkhip_mask *out);
/* DeviceMask::intersect (mask.rs:151-186): same geometry required. */
int sitkhip_mask_intersect(sitkhip_mask a, sitkhip_mask b, sitkhip_mask *out);
int sitkhip_mask_to_host(sitkhip_mask m, uint8_t *out, size_t len);
void sitkhip_mask_free(sitkhip_mask m);

/* ---------------------------------------------------------------- pyramid */
/* recursive_gaussian (src/cuda/ops/pyramid.rs:217-293): ZeroOrder Deriche
 * smoothing with physical `sigma` per axis, axes in order 0,1,2, sigma <= 0
 * skips an axis, a smoothed axis needs >= 4 voxels. Bit-exact against the
 * reference's host filter (src/filters/recursive_gaussian.rs:171-204,472-528):
 * FP64 recursion, one narrowing to float at the end. Output geometry = input. */
int sitkhip_recursive_gaussian(sitkhip_image in, const double sigma[3], sitkhip_image *out);

/* No counterpart in the reference (which smooths one image per call): the
 * recursive Gaussian of n images, sigma[3 i .. 3 i + 2] for image i, in a few
 * multi-image launches (one per axis per group of images). out[i] is the same
 * image, bit for bit, as sitkhip_recursive_gaussian(in[i], sigma + 3 i) gives.
 * For the batched registration's level preparation. */
int sitkhip_recursive_gaussian_batch(size_t n, const sitkhip_image *in, const double *sigma,
                                     sitkhip_image *out);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): sitkhip_recursive_gaussian into the caller's image `out`
 * (the input's geometry, not the input itself), bit for bit the same volume,
 * with no device allocation. The FP64 intermediates live in a static device
 * scratch of 2 x SITKHIP_GAUSSIAN_INTO_MAX_VOXELS doubles (128 MiB) that each
 * device the library runs on holds until the process ends; larger images are
 * refused (SITKHIP_ERR_INVALID_ARGUMENT). Every call on one device shares that
 * scratch, so calls that overlap on one device (several host threads) race. */
#define SITKHIP_GAUSSIAN_INTO_MAX_VOXELS 8388608
int sitkhip_recursive_gaussian_into(sitkhip_image in, const double sigma[3], sitkhip_image out);

/* ------------------------------------------------------ pyramid placement */
/* shrunk_geometry (src/cuda/ops/pyramid.rs:355-420, ITK ShrinkImageFilter): size
 * max(1, size / f), spacing * f, origin moved by Direction * diag(spacing) *
 * delta with delta = (size - 1)/2 - f (out - 1)/2. */
int sitkhip_shrunk_geometry(const sitkhip_geometry *in, const size_t factors[3],
                            sitkhip_geometry *out);
/* shrink (src/cuda/ops/pyramid.rs:422-470): gather every f-th voxel from the
 * rounded centre offset. */
int sitkhip_shrink(sitkhip_image in, const size_t factors[3], sitkhip_image *out);
/* resample_linear[_through] (src/cuda/ops/pyramid.rs:677-738): the input
 * sampled at the point-map image of every output voxel, linear, default value
 * outside. stages NULL / nstage 0: the identity map. */
int sitkhip_resample_linear(sitkhip_image in, const sitkhip_geometry *out_geometry,
                            const double *stages, int nstage, double default_value,
                            sitkhip_image *out);
/* resample_nearest[_through] (src/cuda/ops/pyramid.rs:739-790), round half up. */
int sitkhip_resample_nearest(sitkhip_image in, const sitkhip_geometry *out_geometry,
                             const double *stages, int nstage, double default_value,
                             sitkhip_image *out);
int sitkhip_mask_resample_nearest(sitkhip_mask in, const sitkhip_geometry *out_geometry,
                                  const double *stages, int nstage, sitkhip_mask *out);

/* ------------------------------------------------------- resident metrics */
#define SITKHIP_MAX_STAGES 8 /* MAX_STAGES (src/cuda/ops/resident.rs:74) */
#define SITKHIP_MAX_BINS 255 /* fixed Parzen bins are held as uint8 per sample */

typedef enum { SITKHIP_METRIC_MEAN_SQUARES = 0, SITKHIP_METRIC_MATTES = 1 } sitkhip_metric_kind;
/* Per-sample arithmetic of the sampler. F32 (default): point map and every
 * discrete decision (inside test, cell floor, mask round) in FP64, the
 * trilinear weights / value / gradient and the Parzen window in FP32.
 * F64: everything in FP64 in the host's operation order. Reductions are FP64
 * (moments) or 64-bit fixed point (histogram) in both. */
typedef enum { SITKHIP_PRECISION_F32 = 0, SITKHIP_PRECISION_F64 = 1 } sitkhip_precision;

/* MattesBins (src/cuda/ops/mattes.rs:129-143): derived by the host from the
 * value ranges, handed to the device. */
typedef struct {
    size_t bins;
    size_t padding;
    double fixed_bin_size;
    double moving_bin_size;
    double fixed_normalized_min;
    double moving_normalized_min;
    double moving_true_min;
    double moving_true_max;
} sitkhip_mattes_bins;

typedef struct sitkhip_metric_s *sitkhip_metric;

/* ResidentMetric::from_device_masked / ResidentMattes::from_device_masked
 * (mean_squares.rs:258-312, ops/mattes.rs:526-572).
 *   fixed_indices == NULL: FixedPoints::Grid (every voxel, dim-0-fastest);
 *   otherwise FixedPoints::Indices: sample s is fixed voxel fixed_indices[s]
 *     (duplicates count twice; every index must name a voxel, else
 *     SITKHIP_ERR_SAMPLE_INDEX_OUT_OF_GRID).
 *   fixed_mask / moving_mask may be NULL.
 * The metric copies everything it needs; the handles may be freed after. */
int sitkhip_metric_create(int kind, sitkhip_image fixed, const int64_t *fixed_indices,
                          size_t n_indices, sitkhip_mask fixed_mask, sitkhip_image moving,
                          sitkhip_mask moving_mask, int precision, sitkhip_metric *out);
/* The same with FixedPoints::Indices drawn on the device: `n_draws` random
 * draws with replacement, SplitMix64 from `seed` (metric.rs:101-122, 419-437;
 * the multiset sitkhip_draw_samples(voxels, 2, ...) returns), laid out in
 * ascending voxel order by a device counting sort. No fixed mask. */
int sitkhip_metric_create_drawn(int kind, sitkhip_image fixed, size_t n_draws, uint64_t seed,
                                sitkhip_image moving, int precision, sitkhip_metric *out);

/* FixedPoints (src/cuda/ops/resident.rs:464-484): where the fixed samples'
 * physical points come from.
 *   GRID     : every voxel of the grid (size, origin, idx_to_phys), dim-0
 *              fastest; the sample count is the voxel count;
 *   INDICES  : sample s is grid voxel indices[s] (duplicates count twice);
 *   EXPLICIT : sample s is the physical point points[3 s .. 3 s + 2] and has
 *              no grid voxel (the grid fields are unused). */
typedef enum {
    SITKHIP_POINTS_GRID = 0,
    SITKHIP_POINTS_EXPLICIT = 1,
    SITKHIP_POINTS_INDICES = 2
} sitkhip_points_mode;
typedef struct {
    int mode;                   /* sitkhip_points_mode */
    const double *points;       /* EXPLICIT: n x 3, row-major */
    const int64_t *indices;     /* INDICES: n flat grid indices */
    size_t size[3];             /* GRID / INDICES: the fixed grid */
    double origin[3];
    double idx_to_phys[9];      /* row-major, Direction * diag(spacing) */
} sitkhip_fixed_points;

/* MovingGeometry (src/cuda/ops/resident.rs:424-437). strides must be the
 * dense layout (1, size[0], size[0] * size[1]); mask may be NULL (len bytes,
 * non-zero = inside). phys_to_index is the inverse of the composed
 * Direction * diag(spacing), row-major. */
typedef struct {
    size_t len;
    size_t size[3];
    size_t strides[3];
    double origin[3];
    double phys_to_index[9];
    const uint8_t *mask;
} sitkhip_moving_geometry;

/* ResidentMetric::new (src/cuda/ops/mean_squares.rs:201-256), the constructor
 * CudaMetricBackend calls (src/registration/cuda.rs:195-200): both volumes
 * from host values instead of device images.
 *   fixed_values  : n values (EXPLICIT: the host's gathered per-sample values)
 *                   or one per grid voxel (GRID, INDICES);
 *   moving_values : moving->len values;
 *   fixed_mask    : optional, one byte per grid voxel; refused with EXPLICIT
 *                   points (SITKHIP_ERR_MASKED_EXPLICIT_POINTS, error.rs).
 * The reference's producers widen each pixel to f64; here the values are
 * narrowed to float32 on upload, (float)x, as every device volume of this
 * library is float32 (exact for float32 and integer pixels below 2^24).
 * kind may be mean squares or Mattes (then set the bins as for any metric).
 * Refusals in the order of Resident::build (resident.rs:530-660): n == 0 or a
 * length that does not match its size -> SITKHIP_ERR_DEGENERATE_INPUT, an
 * index outside the grid -> SITKHIP_ERR_SAMPLE_INDEX_OUT_OF_GRID. */
int sitkhip_metric_create_host(int kind, size_t n, const double *fixed_values,
                               const sitkhip_fixed_points *points, const uint8_t *fixed_mask,
                               const sitkhip_moving_geometry *moving, const double *moving_values,
                               int precision, sitkhip_metric *out);
void sitkhip_metric_free(sitkhip_metric m);
/* sample_count (mean_squares.rs:313-318): samples the metric walks (after the
 * fixed mask). */
int sitkhip_metric_sample_count(sitkhip_metric m, size_t *out);
/* volume_bytes (mean_squares.rs:319-328): device bytes the metric holds. */
int sitkhip_metric_volume_bytes(sitkhip_metric m, size_t *out);
/* value_ranges (ops/mattes.rs:587-600): (min, max) of the fixed sample set and
 * of the moving volume over its mask; (0,0) when empty. */
int sitkhip_metric_value_ranges(sitkhip_metric m, double fixed_range[2], double moving_range[2]);

/* The moving image's interpolator, MovingImage::from_image_with_interpolator
 * (src/registration/metric.rs:720-750; the reference's device path refuses
 * every interpolator but linear, registration/method.rs:2724-2728):
 * SITKHIP_INTERP_LINEAR (the default) or SITKHIP_INTERP_BSPLINE, the cubic
 * B-spline of src/transform/interpolator.rs:170-350 -- its coefficients (the
 * mirror-boundary prefilter, FP64 on the device, bit for bit the reference's;
 * narrowed to float on the FP32 path) are computed here from the metric's
 * moving volume, and every later evaluation interpolates with its 64 taps.
 * Value ranges stay those of the voxel values (metric.rs:913-919). Refused:
 * an unknown interpolator (SITKHIP_ERR_INVALID_ARGUMENT), a moving axis of two
 * voxels (SITKHIP_ERR_UNSUPPORTED_GEOMETRY: the reference's mirror fold reads
 * index -1 there). Batched evaluation and the B-spline transform take linear
 * metrics only (SITKHIP_ERR_INVALID_ARGUMENT). */
#define SITKHIP_INTERP_LINEAR 0
#define SITKHIP_INTERP_BSPLINE 1
int sitkhip_metric_set_interpolator(sitkhip_metric m, int interpolator);
/* The metric's cubic B-spline coefficients (len = moving voxels), as FP64
 * (the FP32 path's float values widened); SITKHIP_ERR_INVALID_ARGUMENT for a
 * linear metric. For checking. */
int sitkhip_metric_bspline_coefficients(sitkhip_metric m, double *out, size_t len);

/* draw_samples (src/registration/metric.rs:101-122, 419-437), on the host: the
 * device is handed the list and never draws. strategy 0 None (*n = SIZE_MAX,
 * no list: use FixedPoints::Grid), 1 Regular (stride ceil(1/percentage)),
 * 2 Random (floor(voxels * percentage) SplitMix64 draws with replacement).
 * `out` may be NULL to query the count. */
int sitkhip_draw_samples(size_t voxels, int strategy, double percentage, uint64_t seed,
                         int64_t *out, size_t *n);

/* MattesGeometry::new (registration/mattes.rs:152-186) on the host: refuses
 * bins < 5 and constant ranges by name. padding is 2. */
int sitkhip_mattes_bins_from_ranges(const double fixed_range[2], const double moving_range[2],
                                    size_t bins, sitkhip_mattes_bins *out);
/* Install the histogram geometry the Mattes evaluations below use (bins <=
 * SITKHIP_MAX_BINS). Precomputes every sample's fixed Parzen bin in FP64. */
int sitkhip_mattes_set_bins(sitkhip_metric m, const sitkhip_mattes_bins *bins);

/* Point maps: `stages` is nstage * 12 doubles, each a PointStage (matrix
 * row-major 3x3 then offset[3], resident.rs:60-65), applied in order
 * p <- M p + t (matrix_offset.rs:102-109). 1 <= nstage <= 8. */

/* ResidentMattes::joint_histogram (ops/mattes.rs:697-730): unnormalised joint
 * histogram (bins^2, row-major [fixed*bins + moving]), fixed marginal (bins)
 * and valid-sample count. Accumulated in 64-bit fixed point (2^-32 units),
 * returned as double. */
int sitkhip_mattes_joint_histogram(sitkhip_metric m, const double *stages, int nstage,
                                   double *joint, double *marginal, size_t *valid);
/* ResidentMattes::derivative_moments (ops/mattes.rs:737-841): the 12
 * pRatio-weighted moments A[3], B[3][3] (row-major [d][e]) of the second pass,
 * against the caller's pratio table (bins^2, pRatio * n_factor). */
int sitkhip_mattes_derivative_moments(sitkhip_metric m, const double *stages, int nstage,
                                      const double *pratio, double moments[12]);
/* Fused value: histogram pass + the Mattes tail (mattes.rs:273-328) on the
 * device. value = -MI, or DBL_MAX with valid = 0 for a degenerate histogram. */
int sitkhip_mattes_value(sitkhip_metric m, const double *stages, int nstage, double *value,
                         size_t *valid);
/* Fused value and derivative moments: histogram pass, device tail, derivative
 * pass, with no host round trip between them. */
int sitkhip_mattes_value_and_moments(sitkhip_metric m, const double *stages, int nstage,
                                     double *value, double moments[12], size_t *valid);
/* No counterpart in the reference: sitkhip_mattes_value_and_moments `repeats`
 * times in a row at the same stages, each evaluation waiting for its result
 * as an optimizer's would; the last result is returned. For timing an
 * evaluation the way compiled callers (the reference's optimizer loop) make
 * it, without a foreign-function call per evaluation. */
int sitkhip_mattes_value_and_moments_repeat(sitkhip_metric m, const double *stages, int nstage,
                                            int repeats, double *value, double moments[12],
                                            size_t *valid);

/* ResidentMetric::evaluate (mean_squares.rs:329-439): the 14 moments
 * [sq, S0[3], S1[3][3] row-major [d][e], count] of the mean-squares metric,
 * with diff = M - F. */
int sitkhip_ms_moments(sitkhip_metric m, const double *stages, int nstage, double moments[14]);
/* Value-only mean squares: sum of diff^2 and the valid count. */
int sitkhip_ms_value(sitkhip_metric m, const double *stages, int nstage, double *sq,
                     size_t *valid);

/* ------------------------------------------------- host contractions */
/* The Jacobian's affine decomposition J(x) = J0 + sum_e x_e C_e
 * (registration/cuda.rs:294-363): j0 is 3 x nparams row-major, c is
 * 3 (e) x 3 (d) x nparams. */
/* contract (registration/cuda.rs:381-412): value = sq/count, deriv_k =
 * (2/count) (J0 . S0 + C . S1). count 0 gives DBL_MAX and a zero derivative. */
int sitkhip_contract_ms(const double moments[14], const double *j0, const double *c,
                        int nparams, double *value, double *deriv, size_t *valid);
/* contract_mattes (registration/cuda.rs:505-524). */
int sitkhip_contract_mattes(const double moments[12], const double *j0, const double *c,
                            int nparams, double *deriv);

/* ----------------------------------------------------------- registration */
/* ImageRegistrationMethod::execute_on_device (src/registration/method.rs:
 * 2698-2879): a multi-level registration with every volume resident on the
 * device; the optimizers (src/registration/optimizer.rs:271-817) and the
 * physical-shift scales / learning-rate estimator (src/registration/
 * scales.rs:225-871) run on the host, one metric evaluation per probe. */
#define SITKHIP_MAX_LEVELS 8

typedef enum {
    SITKHIP_TRANSFORM_TRANSLATION = 0, /* t[3] */
    SITKHIP_TRANSFORM_EULER3D = 1,     /* angle x, y, z (rad), t[3]; about center */
    SITKHIP_TRANSFORM_AFFINE = 2       /* A row-major [9], t[3]; about center */
} sitkhip_transform_kind;

typedef struct {
    int kind;
    int compute_zyx; /* Euler3D composition Rz.Ry.Rx instead of Rz.Rx.Ry */
    double center[3];
    double params[12];
} sitkhip_transform;

typedef enum { SITKHIP_OPTIMIZER_REGULAR_STEP = 0, SITKHIP_OPTIMIZER_LINE_SEARCH = 1 } sitkhip_optimizer;
typedef enum { SITKHIP_LR_MANUAL = 0, SITKHIP_LR_ONCE = 1, SITKHIP_LR_EACH_ITERATION = 2 } sitkhip_lr_mode;
typedef enum { SITKHIP_SCALES_UNIT = 0, SITKHIP_SCALES_PHYSICAL_SHIFT = 1 } sitkhip_scales_mode;
typedef enum {
    SITKHIP_STOP_MAX_ITERATIONS = 0,
    SITKHIP_STOP_GRADIENT_CONVERGED = 1,
    SITKHIP_STOP_STEP_TOO_SMALL = 2,
    SITKHIP_STOP_CONVERGED = 3
} sitkhip_stop_reason;

typedef struct {
    int metric;                 /* sitkhip_metric_kind */
    int bins;                   /* Mattes histogram bins */
    int sampling;               /* 0 every voxel, 1 regular, 2 random (SplitMix64) */
    double sampling_percentage;
    uint64_t seed;
    int precision;              /* sitkhip_precision */
    int optimizer;              /* sitkhip_optimizer */
    double learning_rate;       /* manual rate (RSGD: initial step length) */
    double minimum_step;        /* RSGD stop */
    int iterations;             /* per level */
    double relaxation;          /* RSGD step relaxation on reversal (0.5) */
    double gradient_tolerance;  /* RSGD stationary-point stop (1e-4) */
    int learning_rate_mode;     /* sitkhip_lr_mode */
    int scales_mode;            /* sitkhip_scales_mode */
    int levels;
    size_t shrink[SITKHIP_MAX_LEVELS][3];
    double sigma[SITKHIP_MAX_LEVELS]; /* physical smoothing sigma per level, all axes */
    /* set_metric_fixed_mask / set_metric_moving_mask (method.rs:1555-1590),
     * NULL for none. The fixed mask lies on the fixed image's grid and is
     * placed on every level's grid by the nearest-neighbour resample with
     * default 0 (level_mask_on_device, method.rs:3068-3110); the moving mask
     * lies on the moving image's grid and is used as is at every level (the
     * moving volume is never shrunk). A mask whose size differs from its
     * image's is SITKHIP_ERR_MASK_SIZE_MISMATCH (method.rs:2070-2083). */
    sitkhip_mask fixed_mask;
    sitkhip_mask moving_mask;
    /* set_interpolator (method.rs:1507-1520): the moving image's interpolator
     * on every level, SITKHIP_INTERP_LINEAR (0, the default) or
     * SITKHIP_INTERP_BSPLINE (sitkhip_metric_set_interpolator). Batched
     * registration interpolates linearly and refuses the B-spline. */
    int interpolator;
} sitkhip_registration_config;

typedef struct {
    double value;               /* metric value at the result, last level */
    size_t valid_points;
    int iterations;             /* summed over levels */
    int level_iterations[SITKHIP_MAX_LEVELS];
    int level_stop[SITKHIP_MAX_LEVELS];
    double level_value[SITKHIP_MAX_LEVELS];
    size_t evaluations;         /* value + derivative evaluations */
    size_t value_evaluations;   /* value-only evaluations (line-search probes) */
    double seconds;             /* wall clock of the whole run */
} sitkhip_registration_result;

/* transform: the initial transform in, the registered one out. */
int sitkhip_register(sitkhip_image fixed, sitkhip_image moving,
                     const sitkhip_registration_config *config, sitkhip_transform *transform,
                     sitkhip_registration_result *result);

/* ---------------------------------------------------------------- batches */
/* Many resident metrics evaluated together. No counterpart in the reference,
 * whose device metrics evaluate one pair per call (src/cuda/ops/mattes.rs:
 * 697-841, mean_squares.rs:329-439); it is what a batched registration loop
 * calls in place of one evaluate per pair. One evaluation is one replay of a
 * captured HIP graph: a persistent value pass over the tiles of every pair
 * (Mattes histograms and mean-squares moments), the Mattes tails, a persistent
 * derivative pass over the Mattes tiles, and a per-pair fold. Per pair the
 * result equals the single-pair evaluation: the Mattes histogram (and so its
 * value) exactly, the FP64 moments up to the order of their sum. */
typedef struct sitkhip_batch_s *sitkhip_batch;

/* The metrics (same precision; Mattes ones with bins set, at most 64 bins)
 * must outlive the batch and must not be evaluated on their own while it is in
 * use (they share histogram buffers). */
int sitkhip_batch_create(const sitkhip_metric *metrics, size_t n, sitkhip_batch *out);
void sitkhip_batch_free(sitkhip_batch b);
/* tile_samples: samples per tile (0: 4096); grid: persistent workgroups, a
 * multiple of 8 (0: 4 per CU); use_graph 0 launches the passes directly. */
int sitkhip_batch_set_tiling(sitkhip_batch b, size_t tile_samples, unsigned grid, int use_graph);
int sitkhip_batch_get_tiling(sitkhip_batch b, size_t *tile_samples, unsigned *grid);
/* Times every candidate (tile, grid) on `stages` and keeps the fastest. */
int sitkhip_batch_autotune(sitkhip_batch b, const double *stages, size_t *tile_samples,
                           unsigned *grid, double *ms_per_evaluation);
/* stages: n x 12 doubles, one PointStage per pair; active: n flags or NULL
 * (all). Per active pair: values[i] (-MI or mean squares; DBL_MAX when no
 * sample is valid), valid[i], and moments[14 i ..] (12 Mattes moments then two
 * zeros, or the 14 mean-squares moments). Inactive pairs are not evaluated
 * and their outputs are left untouched. */
int sitkhip_batch_evaluate(sitkhip_batch b, const double *stages, const uint8_t *active,
                           double *values, double *moments, size_t *valid);

/* n registrations run together: per level every pair's level is prepared as
 * sitkhip_register prepares it, then each optimizer iteration evaluates every
 * still-running pair in one batched evaluation. RegularStepGradientDescent
 * only. configs, transforms and results are arrays of n. autotune != 0 picks
 * the tiling per level with sitkhip_batch_autotune. */
int sitkhip_register_batch(size_t n, const sitkhip_image *fixed, const sitkhip_image *moving,
                           const sitkhip_registration_config *configs,
                           sitkhip_transform *transforms, sitkhip_registration_result *results,
                           int autotune);

/* sitkhip_register_batch, also reporting each pair's metric value (-MI or mean
 * squares, as the optimizer sees it before its own evaluation) at a second
 * parameter set probe_params[12 i ..] (the transform kind's parameter count
 * used), evaluated on the first level's prepared images before optimizing:
 * probe_values[i]. For an initializer's fallback rule (keep the optimized
 * result only if its value beats the probe's) without preparing the level
 * twice. No counterpart in the reference. */
int sitkhip_register_batch_probe(size_t n, const sitkhip_image *fixed, const sitkhip_image *moving,
                                 const sitkhip_registration_config *configs,
                                 sitkhip_transform *transforms, sitkhip_registration_result *results,
                                 int autotune, const double *probe_params, double *probe_values);

/* ---------------------------------------------------------------- L-BFGS-B */
/* The bound-constrained limited-memory BFGS optimizer
 * (src/registration/lbfgsb.rs:1590-1746, LBFGSBOptimizer: Byrd, Lu, Nocedal
 * & Zhu with the Moré-Thuente line search and ITK's stopping rules), on the
 * host, driven by reverse communication instead of an `eval` closure: start
 * and tell write the point where they want the value and gradient to x (n
 * doubles) and set *want = 1, or *want = 0 once the optimizer has stopped;
 * result then gives the best point evaluated (the reference returns that
 * point, not the last iterate), its value, the iteration and evaluation counts
 * and the stop reason (sitkhip_bspline_result.stop numbering: 0 iterations,
 * 1 projected gradient, 2 relative reduction, 3 evaluations, 4 line search).
 * bound_selection (ITK BoundSelection, lbfgsb.rs:1632-1640): 0 unbounded,
 * 1 lower only, 2 both, 3 upper only, null = every variable unbounded;
 * lower / upper are read where a selection uses them. An infeasible pair
 * (lower > upper under selection 2) is SITKHIP_ERR_INVALID_ARGUMENT, as the
 * reference panics (lbfgsb.rs:1711-1720). Host code: no device is touched. */
typedef struct sitkhip_lbfgsb_s *sitkhip_lbfgsb;

typedef struct {
    int corrections;            /* m, stored corrections (5) */
    double factr;               /* relative-reduction factor (1e7) */
    double pgtol;               /* projected-gradient tolerance (1e-5) */
    int max_iterations;
    int max_evaluations;        /* (2000) */
} sitkhip_lbfgsb_options;

int sitkhip_lbfgsb_create(size_t n, const sitkhip_lbfgsb_options *options, const double *lower,
                          const double *upper, const int *bound_selection, sitkhip_lbfgsb *out);
void sitkhip_lbfgsb_free(sitkhip_lbfgsb optimizer);
int sitkhip_lbfgsb_start(sitkhip_lbfgsb optimizer, const double *initial, double *x, int *want);
int sitkhip_lbfgsb_tell(sitkhip_lbfgsb optimizer, double value, const double *gradient, double *x,
                        int *want);
int sitkhip_lbfgsb_result(sitkhip_lbfgsb optimizer, double *best, double *value, int *iterations,
                          int *evaluations, int *stop);

/* ---------------------------------------------------------------- B-spline */
/* Mattes MI through a cubic B-spline transform: the value and the
 * sparse-support derivative (src/registration/mattes.rs:722-796, transform
 * src/transform/bspline.rs:485-703), which the reference's device backend
 * refuses (registration/device.rs:79-100). Built on a Mattes metric with bins
 * set (the metric must outlive it and is used by it); the control grid is the
 * transform's fixed parameters (bspline.rs:229-322): grid_size = mesh + 3,
 * origin, spacing, direction. Coefficients are 3 x ncp doubles [d][cp], cp =
 * x + gs0 (y + gs1 z). The derivative is deterministic (no atomics). */
typedef struct sitkhip_bspline_s *sitkhip_bspline;

int sitkhip_bspline_create(sitkhip_metric metric, const size_t grid_size[3],
                           const double grid_origin[3], const double grid_spacing[3],
                           const double grid_direction[9], sitkhip_bspline *out);
void sitkhip_bspline_free(sitkhip_bspline b);
int sitkhip_bspline_parameter_count(sitkhip_bspline b, size_t *n);
int sitkhip_bspline_mattes_value(sitkhip_bspline b, const double *coeffs, double *value,
                                 size_t *valid);
/* deriv: 3 x ncp doubles, zero where no valid sample supports a point. */
int sitkhip_bspline_mattes_value_and_derivative(sitkhip_bspline b, const double *coeffs,
                                                double *value, double *deriv, size_t *valid);

/* Sets (stage non-null) or clears the moving-initial transform applied
 * after the B-spline: moving point = stage(B(x)), stage = matrix row-major
 * [9], offset [3] (ImageRegistrationMethod::set_moving_initial_transform and
 * the composition of registration/method.rs:198-340, whose sparse Jacobian
 * is left-multiplied by the stage's matrix). */
int sitkhip_bspline_set_moving_initial(sitkhip_bspline b, const double stage[12]);

/* Many starts evaluated together: one B-spline metric (its images, sample
 * set and grid) with nstarts coefficient vectors and moving-initial stages
 * (stages: nstarts x 12, null = none). value_and_derivative evaluates the
 * starts listed in active[0..nactive) at coeffs[s * 3 ncp ..] in one set of
 * launches and writes values[s], derivs[s * 3 ncp ..], valid[s] of those
 * starts; each start's results are bit-equal to a single-start evaluation
 * with the same stage. No counterpart in the reference (its device path
 * refuses B-spline, registration/device.rs:79-100). */
typedef struct sitkhip_bspline_multi_s *sitkhip_bspline_multi;

int sitkhip_bspline_multi_create(sitkhip_bspline b, size_t nstarts, const double *stages,
                                 sitkhip_bspline_multi *out);
void sitkhip_bspline_multi_free(sitkhip_bspline_multi m);
int sitkhip_bspline_multi_value_and_derivative(sitkhip_bspline_multi m, size_t nactive,
                                               const int *active, const double *coeffs,
                                               double *values, double *derivs, size_t *valid);

/* A deformable registration: Mattes MI through a cubic B-spline transform
 * whose domain is the fixed image's extent (bspline.rs:324-339) with `mesh`
 * cells per axis, optimised by L-BFGS-B on the host (LBFGSBOptimizer,
 * lbfgsb.rs:1590-1746, every variable unbounded). coeffs: the initial
 * coefficients in (3 x (mesh+3)^3, [d][cp]), the best ones out. */
typedef struct {
    int bins;
    int sampling;               /* 0 every voxel, 2 random (drawn on the device) */
    double sampling_percentage;
    uint64_t seed;
    int precision;
    int iterations;             /* maximum iterations */
    int max_evaluations;
    int corrections;            /* stored corrections m (5) */
    double factr;               /* relative-reduction factor (1e7) */
    double pgtol;               /* gradient infinity-norm tolerance (1e-5) */
    size_t mesh[3];
} sitkhip_bspline_config;

typedef struct {
    double value;
    size_t valid_points;
    int iterations;
    int evaluations;
    int stop;                   /* 0 iterations, 1 gradient, 2 reduction, 3 evaluations, 4 no descent */
    double seconds;
} sitkhip_bspline_result;

int sitkhip_register_bspline(sitkhip_image fixed, sitkhip_image moving,
                             const sitkhip_bspline_config *config, double *coeffs, size_t ncoeffs,
                             sitkhip_bspline_result *result);

/* nstarts registrations of one pair at once, start s from its own
 * moving-initial stage stages[12 s ..] (null: none) and coefficients
 * coeffs[s * ncoeffs ..] (the best ones out), each with its own L-BFGS-B;
 * every round, all optimizers that want an evaluation get it from one
 * batched set of launches. bound_selection / lower / upper: per coefficient,
 * shared by every start (null selection: unbounded). results[s] per start;
 * *best (may be null) = the start with the lowest final value. Each start's
 * result equals a one-start call's. No counterpart in the reference. */
int sitkhip_register_bspline_multistart(sitkhip_image fixed, sitkhip_image moving,
                                        const sitkhip_bspline_config *config, size_t nstarts,
                                        const double *stages, double *coeffs, size_t ncoeffs,
                                        const int *bound_selection, const double *lower,
                                        const double *upper, sitkhip_bspline_result *results,
                                        size_t *best);

/* ------------------------------------------------------------- generators */
/* Seeded synthetic volume (no counterpart in the reference; the benchmark
 * pairs' generator): the sum over `nblobs` Gaussian blobs (physical centres
 * [3 b ..], amplitudes, sigmas; each blob ends at 4.5 sigma along every axis)
 * plus N(0, noise_sigma) voxel noise drawn per voxel from SplitMix64(seed, voxel
 * index) by Box-Muller. Identity direction only. */
int sitkhip_synth_blobs(const sitkhip_geometry *geometry, int nblobs, const double *centres,
                        const double *amplitudes, const double *sigmas, double noise_sigma,
                        uint64_t seed, sitkhip_image *out);

/* ------------------------------------------------------ learned initializer */
/* A small 3-D CNN that predicts an affine start for a registration from the
 * image pair downsampled to 64^3. No counterpart in the reference (its
 * initializers are geometric, src/registration/initializer.rs); it is what a
 * registration loop calls before the optimizer. Architecture: five 3^3
 * convolutions (2-16-32-64-64-128 channels, stride 2 from the second, ReLU),
 * global average pool, fc 128-128 (ReLU), fc 128-12. Outputs: the 9 entries of
 * (A - I) and t / 10 mm of x -> A (x - c) + t + c about the fixed image's
 * centre c (fixed to moving, the registration's own direction). Convolutions
 * (forward and backward) run as BF16 implicit GEMMs on the matrix cores with
 * FP32 accumulation; weights and Adam state are FP32. */
typedef struct sitkhip_cnn_s *sitkhip_cnn;

/* He-initialised weights from `seed` (SplitMix64); batches of up to max_batch. */
int sitkhip_cnn_create(int max_batch, uint64_t seed, sitkhip_cnn *out);
void sitkhip_cnn_free(sitkhip_cnn net);
int sitkhip_cnn_param_count(sitkhip_cnn net, size_t *n);
/* Flat FP32 parameters: per conv layer W[co][27][ci] (ci of the first layer
 * padded to 8, the padding zero) then bias[co]; fc1 W[128][128], b[128]; fc2
 * W[12][128], b[12]. */
int sitkhip_cnn_get_params(sitkhip_cnn net, float *out, size_t n);
int sitkhip_cnn_set_params(sitkhip_cnn net, const float *in, size_t n);
/* Gradients of the last backward pass, same layout. */
int sitkhip_cnn_get_grads(sitkhip_cnn net, float *out, size_t n);
/* Input slot `slot`: both images linearly resampled onto 64^3 grids spanning
 * their own extents, z-scored, packed as BF16. */
int sitkhip_cnn_set_input(sitkhip_cnn net, int slot, sitkhip_image fixed, sitkhip_image moving);
/* The same from two host 64^3 float volumes (x fastest). */
int sitkhip_cnn_set_input_raw(sitkhip_cnn net, int slot, const float *fixed64,
                              const float *moving64);
/* Slots first_slot .. first_slot + n - 1 from n image pairs (fixed[b],
 * moving[b]) as sitkhip_cnn_set_input does for each, the whole set in four
 * launches (the 2n placements, the normalisations, the packing). */
int sitkhip_cnn_set_inputs(sitkhip_cnn net, int first_slot, int n, const sitkhip_image *fixed,
                           const sitkhip_image *moving);
/* Forward pass over slots 0..batch-1; out: batch x 12 (may be NULL). */
int sitkhip_cnn_forward(sitkhip_cnn net, int batch, float *out);
/* Backward pass of the last forward from dL/dout (batch x 12). */
int sitkhip_cnn_backward(sitkhip_cnn net, int batch, const float *dout);
/* One Adam step on the gradients of the last backward pass. */
int sitkhip_cnn_adam(sitkhip_cnn net, float lr, float beta1, float beta2, float eps);
/* forward + mean-squared error against targets (batch x 12) + backward + Adam
 * (beta 0.9 / 0.999, eps 1e-8). */
int sitkhip_cnn_train_step(sitkhip_cnn net, int batch, const float *targets, float lr,
                           float *loss);
/* Activation `index` (0 the packed input, l + 1 the output of conv l) of the
 * last forward, or with gradient != 0 the gradient w.r.t. it from the last
 * backward (after the ReLU mask), as floats; n must be its element count. */
int sitkhip_cnn_activation(sitkhip_cnn net, int index, int gradient, int batch, float *out,
                           size_t n);

/* ----------------------------------------------------------------- memory */
/* Device memory is cached by size class after a free (hipMalloc / hipFree
 * cost tens of microseconds and hipFree waits for the device); this returns
 * every cached block to the runtime. No counterpart in the reference, whose
 * buffers free on drop (src/cuda/buffer.rs:12-14). */
int sitkhip_empty_cache(void);

/* -------------------------------------------------------------- profiling */
/* When enabled, every kernel launch of this library is bracketed by HIP events
 * on the library stream and its duration accumulated per kernel name. */
int sitkhip_profile_enable(int on);
int sitkhip_profile_reset(void);
/* Accumulated (count, total milliseconds) of kernels whose name starts with
 * `prefix`; returns the number of distinct kernel names matched. */
int sitkhip_profile_query(const char *prefix, uint64_t *count, double *total_ms);
/* Writes "name count total_ms\n" lines into buf. */
int sitkhip_profile_dump(char *buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif
