/*
 * CPU oracle for the registration metric-evaluation path of sitk-rs.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C, FP64, serial restatement of the
 * reference's CPU algorithm, used as the checker for the HIP library. Only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark load it;
 * the product library never links, loads or calls it.
 *
 * Every function names the reference file:line it restates (paths relative to
 * crates/sitk/ of physwkim/sitk-rs). Expression order follows the reference
 * term by term: Rust evaluates `a*b + c*d + ...` left to right and never fuses
 * a multiply with an add, so this file is compiled with -ffp-contract=off.
 *
 * Parity pins (what checks each function; see tests/test_oracle_golden.py):
 *   - orc_synth / orc_fnv1a64: the bench generator's first five values and the
 *     input checksums held by tests/bench_correctness_gate.rs:162-211.
 *   - orc_deriche_coefficients / orc_filter_line / orc_recursive_gaussian_*:
 *     four golden FNV checksums (bench/results/rust-cuda.ndjson:5,53,101 and
 *     tests/bit_parity.rs:151-160), DC gain, constant and sigma=0 invariants.
 *   - Mattes (geometry, B3, B3', histogram, tail, value, global-support
 *     derivative): the reference-held value -1.063357897140305
 *     (doc/upstream-findings.md:273, fixture tests/mattes_device.rs:42-81),
 *     partition of unity, brute-force tiny cases, central finite differences.
 *   - Mean squares: zero at identity, hand-computable one-voxel shift, finite
 *     differences, moment-form equality.
 *   - Transforms: Euler3D Jacobian against finite differences; stage replay
 *     equals transform_point bitwise.
 *   - B-spline transform + sparse-support Mattes: finite differences only
 *     ("parity pinned by FD and by the dense path on the identity", no
 *     reference-held vector exists).
 */
#ifndef SITK_ORACLE_H
#define SITK_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- bench generator and checksum (benches/bench_ops/synth.rs:19-29,
 *      benches/bench_ops/checksum.rs:8-18) ---- */
void orc_synth(uint64_t seed, size_t n, float *out);
uint64_t orc_fnv1a64(const void *bytes, size_t len);

/* ---- Deriche recursive Gaussian (src/core/deriche.rs:187-347,
 *      src/filters/recursive_gaussian.rs:171-204, 472-528) ---- */
void orc_deriche_coefficients(int order, double sigmad, double sigma,
                              int normalize_across_scale, double out[20]);
void orc_filter_line(const double *data, size_t ln, const double c[20],
                     double *outs, double *scratch);
/* In place on a dim-0-fastest FP64 volume. Returns 0, or -1 if a smoothed
 * axis is shorter than 4 (AxisTooShortForRecursion), -2 for a negative sigma. */
int orc_recursive_gaussian_f64(double *buf, const size_t size[3],
                               const double spacing[3], const double sigma[3]);
/* smoothing_recursive_gaussian on a Float32 volume: widen, filter, narrow
 * once with `as f32` (src/filters/recursive_gaussian.rs:363-391). */
int orc_smoothing_recursive_gaussian_f32(const float *in, float *out,
                                         const size_t size[3],
                                         const double spacing[3],
                                         const double sigma[3]);
/* Same, with the per-line work spread over `threads` POSIX threads (lines are
 * independent, so the output is bit-identical to the serial form; mirrors
 * parallel::for_each_line_mut, src/core/parallel.rs:990). */
int orc_smoothing_recursive_gaussian_f32_mt(const float *in, float *out,
                                            const size_t size[3],
                                            const double spacing[3],
                                            const double sigma[3], int threads);

/* ---- geometry (src/core/coord.rs:50-88, src/core/matrix.rs:9-102) ---- */
void orc_index_to_physical_matrix(const double dir[9], const double spacing[3],
                                  double out[9]);
int orc_invert3(const double m[9], double out[9]); /* 0 ok, -1 singular */

/* ---- transforms (src/transform/parametric.rs) ---- */
enum { ORC_TRANSLATION = 0, ORC_EULER3D = 1, ORC_AFFINE = 2, ORC_BSPLINE = 3 };

typedef struct {
    int kind;
    int nparams;
    double params[12];  /* Translation: t[3]; Euler3D: ax,ay,az,tx,ty,tz;
                           Affine: A row-major[9], t[3] */
    double center[3];
    int compute_zyx;
    double matrix[9];   /* cached (Euler3D / Affine), identity for Translation */
    double offset[3];
    /* B-spline (src/transform/bspline.rs): grid geometry + coefficients */
    size_t grid_size[3];
    double grid_origin[3];
    double grid_spacing[3];
    double grid_direction[9];
    double grid_p2i[9];   /* inverse(direction * diag(grid_spacing)) */
    const double *coeffs; /* 3 * prod(grid_size), component-major */
} orc_transform;

/* BSplineTransform::from_image_domain (src/transform/bspline.rs:229-339):
 * cubic grid covering an image, mesh+3 control points per axis. */
int orc_bspline_from_image_domain(orc_transform *t, const size_t img_size[3],
                                  const double img_spacing[3], const double img_origin[3],
                                  const double img_dir[9], const size_t mesh[3],
                                  const double *coeffs);

void orc_transform_recompute(orc_transform *t);
void orc_transform_point(const orc_transform *t, const double p[3], double out[3]);
/* Dense Jacobian, row-major 3 x nparams (Translation/Euler3D/Affine). */
void orc_transform_jacobian(const orc_transform *t, const double p[3], double *jac);
/* The single point-map stage (matrix row-major, offset) the transform reports
 * (src/transform/matrix_offset.rs:83-109). */
void orc_transform_stage(const orc_transform *t, double stage[12]);
void orc_replay_stages(const double *stages, int nstage, const double p[3],
                       double out[3]);

/* ---- the metric problem: fixed samples + moving image ---- */
typedef struct {
    size_t fsize[3];
    double forigin[3];
    double fi2p[9];          /* index -> physical, Direction * diag(spacing) */
    const float *fvals;      /* whole fixed grid, dim-0-fastest (explicit points: one per sample) */
    const int64_t *idx;      /* sample list (flat indices, sample order), or NULL = grid */
    size_t nidx;             /* length of idx, or the number of explicit points */
    const double *pts;       /* explicit points (resident.rs:464-484), nidx x 3, or NULL */
    const uint8_t *fmask;    /* optional fixed mask on the fixed grid */
    size_t msize[3];
    double morigin[3];
    double mp2i[9];          /* physical -> index, inverse of the composed matrix */
    const float *mvals;
    const uint8_t *mmask;    /* optional moving mask on the moving grid */
    const double *mcoef;     /* cubic B-spline coefficients of mvals (interpolator.rs:170-190),
                                or NULL = linear interpolation */
} orc_problem;

/* Sampling (src/registration/metric.rs:101-122, 419-437). strategy: 0 None,
 * 1 Regular, 2 Random. Returns the number of indices written (out may be NULL
 * to query), or (size_t)-1 for None. */
size_t orc_draw_samples(size_t voxels, int strategy, double percentage,
                        uint64_t seed, int64_t *out);
/* SplitMix64 next_u64 stepping a caller-held state (metric.rs:110-116). */
uint64_t orc_splitmix64_next(uint64_t *state);

/* (min, max) over the sample set / masked moving volume; returns 0 if empty. */
/* bspline_coefficients (src/transform/interpolator.rs:170-190, 236-270): the
 * cubic B-spline prefilter (Unser, mirror boundary) of a float volume, axis by
 * axis (axes of one voxel skipped), into `out` (FP64, same layout). */
void orc_bspline_coefficients(const float *in, const size_t size[3], double *out);
/* compiler-rt's __powidf2 (the routine Rust's f64::powi lowers to): binary
 * exponentiation, one rounding per multiply. */
double orc_powi(double a, int b);
/* bspline_value_and_gradient (interpolator.rs:273-350) at continuous index c:
 * value and index-space gradient from the coefficients; 0 if outside. */
int orc_bspline_value_and_gradient(const double *coef, const size_t size[3], const double c[3],
                                   double *value, double grad[3]);
int orc_fixed_value_range(const orc_problem *pb, double out[2]);
int orc_moving_value_range(const orc_problem *pb, double out[2]);

/* Per-sample sampler: returns 1 and fills (value, physical gradient) if valid. */
int orc_moving_value_and_physical_gradient(const orc_problem *pb, const double p[3],
                                           double *value, double grad[3]);
int orc_moving_value_at(const orc_problem *pb, const double p[3], double *value);

/* ---- Mattes (src/registration/mattes.rs:92-328, 442-796) ---- */
typedef struct {
    size_t num_bins;
    double moving_true_min, moving_true_max;
    double fixed_bin_size, moving_bin_size;
    double fixed_normalized_min, moving_normalized_min;
} orc_mattes_geom;

/* 0 ok, -1 too few bins, -2 constant fixed, -3 constant moving */
int orc_mattes_geometry(const double fixed_range[2], const double moving_range[2],
                        size_t bins, orc_mattes_geom *g);
double orc_cubic_bspline(double u);
double orc_cubic_bspline_derivative(double u);
size_t orc_parzen_window_index(const orc_mattes_geom *g, double value,
                               double bin_size, double normalized_min);
double orc_compensated_sum(const double *v, size_t n);
/* 1 (default): Kahan joint-PDF mass as the current reference; 0: the naive
 * sum the reference used when its ledger value was recorded. */
void orc_set_joint_sum_compensated(int on);
/* build_histogram with transform_point of t. joint: bins^2, marginal: bins. */
void orc_mattes_histogram(const orc_problem *pb, const orc_mattes_geom *g,
                          const orc_transform *t, double *joint,
                          double *marginal, size_t *valid);
/* mattes_tail: returns 1 and fills value/pratio (bins^2, may be NULL), or 0
 * for the degenerate histogram (value f64::MAX). joint/marginal are consumed
 * (normalized in place). */
int orc_mattes_tail(double *joint, double *marginal, size_t valid,
                    const orc_mattes_geom *g, double *value, double *pratio);
double orc_mattes_value(const orc_problem *pb, const orc_mattes_geom *g,
                        const orc_transform *t);
/* evaluate(): global support for dense transforms, sparse for B-spline. */
double orc_mattes_evaluate(const orc_problem *pb, const orc_mattes_geom *g,
                           const orc_transform *t, double *deriv, size_t *valid);
/* The sparse (B-spline) evaluation with a moving-initial transform applied
 * after it: moving_point = init(t(x)) (registration/method.rs:198-340); init
 * must be linear (translation, Euler3D, affine) or NULL. Returns -1 for a
 * non-B-spline t. */
double orc_mattes_evaluate_bspline_composed(const orc_problem *pb, const orc_mattes_geom *g,
                                            const orc_transform *t, const orc_transform *init,
                                            double *deriv, size_t *valid);
/* The 12 pRatio-weighted moments A[3], B[3][3] (src/cuda/ops/mattes.rs:75-91)
 * in FP64, for pinning the moment form against the dense form. */
void orc_mattes_moments(const orc_problem *pb, const orc_mattes_geom *g,
                        const orc_transform *t, const double *pratio,
                        double moments[12]);

/* ---- mean squares (src/registration/metric.rs:1065-1223) ---- */
double orc_ms_evaluate(const orc_problem *pb, const orc_transform *t,
                       double *deriv, size_t *valid);
/* orc_ms_evaluate computed the way the reference threads it: rows in parallel
 * over `threads` threads, added in sample order on one (the same bits). */
double orc_ms_evaluate_threads(const orc_problem *pb, const orc_transform *t, double *deriv,
                               size_t *valid, int threads);
double orc_ms_value(const orc_problem *pb, const orc_transform *t, size_t *valid);

/* ---- pyramid placement (src/transform/resample.rs:226-290,
 *      src/filters/shrink.rs:51-91) ---- */
void orc_resample_linear(const float *in, const size_t in_size[3], const double in_spacing[3],
                         const double in_origin[3], const double in_dir[9],
                         const size_t out_size[3], const double out_spacing[3],
                         const double out_origin[3], const double out_dir[9],
                         const orc_transform *t, double dflt, float *out);
void orc_resample_nearest(const float *in, const size_t in_size[3], const double in_spacing[3],
                         const double in_origin[3], const double in_dir[9],
                         const size_t out_size[3], const double out_spacing[3],
                         const double out_origin[3], const double out_dir[9],
                         const orc_transform *t, double dflt, float *out);
void orc_shrunk_geometry(const size_t size[3], const double spacing[3], const double origin[3],
                         const double dir[9], const size_t f[3], size_t out_size[3],
                         double out_spacing[3], double out_origin[3]);

/* ---- fixtures of tests/mattes_device.rs:42-100 ---- */
void orc_fixture_volume(size_t n, const double shift[3], float *out);
void orc_fixture_inverted(size_t n, const double shift[3], float *out);
void orc_fixture_commensurate(size_t n, float *out);

#ifdef __cplusplus
}
#endif
#endif
