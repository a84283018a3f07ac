/*
 * CPU oracle (test infrastructure only) -- FP64 serial restatement of the
 * reference's registration metric path. See oracle.h for what pins each part.
 * Compile with -ffp-contract=off and without -ffast-math: every expression
 * below is evaluated in the reference's left-to-right order with separately
 * rounded multiplies and adds.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* benches/bench_ops/synth.rs:19-29                                           */
void orc_synth(uint64_t seed, size_t n, float *out) {
    uint64_t state = seed;
    for (size_t i = 0; i < n; ++i) {
        state ^= state >> 12;
        state ^= state << 25;
        state ^= state >> 27;
        out[i] = (float)(((state * 0x2545F4914F6CDD1DULL) >> 33) % 1000ULL);
    }
}

/* benches/bench_ops/checksum.rs:8-18 */
uint64_t orc_fnv1a64(const void *bytes, size_t len) {
    const uint8_t *b = (const uint8_t *)bytes;
    uint64_t h = 0xcbf29ce484222325ULL;
    for (size_t i = 0; i < len; ++i) {
        h ^= (uint64_t)b[i];
        h *= 0x00000100000001b3ULL;
    }
    return h;
}

/* ------------------------------------------------------------------------- */
/* src/core/deriche.rs:112-117 (poles), :133-154 (Farneback columns)          */
static const double W1 = 0.6681, L1 = -1.3932, W2 = 2.0787, L2 = -1.3732;
typedef struct { double a1, b1, a2, b2; } column_t;
static column_t column_of(int order) {
    column_t c;
    if (order == 0) { c.a1 = 1.3530; c.b1 = 1.8151; c.a2 = -0.3531; c.b2 = 0.0902; }
    else if (order == 1) { c.a1 = -0.6724; c.b1 = -3.4327; c.a2 = 0.6724; c.b2 = 0.6100; }
    else { c.a1 = -1.3563; c.b1 = 5.2318; c.a2 = 0.3446; c.b2 = -2.2355; }
    return c;
}

/* src/core/deriche.rs:304-319 */
static void compute_d(double sigmad, double *d1, double *d2, double *d3, double *d4,
                      double *sd, double *dd, double *ed) {
    double cos1 = cos(W1 / sigmad);
    double cos2 = cos(W2 / sigmad);
    double exp1 = exp(L1 / sigmad);
    double exp2 = exp(L2 / sigmad);
    *d4 = exp1 * exp1 * exp2 * exp2;
    *d3 = -2.0 * cos1 * exp1 * exp2 * exp2 - 2.0 * cos2 * exp2 * exp1 * exp1;
    *d2 = 4.0 * cos2 * cos1 * exp1 * exp2 + exp1 * exp1 + exp2 * exp2;
    *d1 = -2.0 * (exp2 * cos2 + exp1 * cos1);
    *sd = 1.0 + *d1 + *d2 + *d3 + *d4;
    *dd = *d1 + 2.0 * *d2 + 3.0 * *d3 + 4.0 * *d4;
    *ed = *d1 + 4.0 * *d2 + 9.0 * *d3 + 16.0 * *d4;
}

/* src/core/deriche.rs:325-347 */
static void compute_n(double sigmad, column_t c, double *n0, double *n1, double *n2,
                      double *n3, double *sn, double *dn, double *en) {
    double sin1 = sin(W1 / sigmad);
    double sin2 = sin(W2 / sigmad);
    double cos1 = cos(W1 / sigmad);
    double cos2 = cos(W2 / sigmad);
    double exp1 = exp(L1 / sigmad);
    double exp2 = exp(L2 / sigmad);
    *n0 = c.a1 + c.a2;
    *n1 = exp2 * (c.b2 * sin2 - (c.a2 + 2.0 * c.a1) * cos2) +
          exp1 * (c.b1 * sin1 - (c.a1 + 2.0 * c.a2) * cos1);
    *n2 = ((c.a1 + c.a2) * cos2 * cos1 - c.b1 * cos2 * sin1 - c.b2 * cos1 * sin2) * 2.0 *
              exp1 * exp2 +
          c.a2 * exp1 * exp1 + c.a1 * exp2 * exp2;
    *n3 = exp2 * exp1 * exp1 * (c.b2 * sin2 - c.a2 * cos2) +
          exp1 * exp2 * exp2 * (c.b1 * sin1 - c.a1 * cos1);
    *sn = *n0 + *n1 + *n2 + *n3;
    *dn = *n1 + 2.0 * *n2 + 3.0 * *n3;
    *en = *n1 + 4.0 * *n2 + 9.0 * *n3;
}

/* src/core/deriche.rs:187-296; output order is to_array (:290-296). */
void orc_deriche_coefficients(int order, double sigmad, double sigma,
                              int normalize_across_scale, double out[20]) {
    double d1, d2, d3, d4, sd, dd, ed;
    compute_d(sigmad, &d1, &d2, &d3, &d4, &sd, &dd, &ed);
    double n0, n1, n2, n3;
    int symmetric;
    if (order == 0) {
        double sn, dn_, en_;
        compute_n(sigmad, column_of(0), &n0, &n1, &n2, &n3, &sn, &dn_, &en_);
        double alpha0 = 2.0 * sn / sd - n0;
        n0 = n0 / alpha0; n1 = n1 / alpha0; n2 = n2 / alpha0; n3 = n3 / alpha0;
        symmetric = 1;
    } else if (order == 1) {
        double sn, dn, en_;
        compute_n(sigmad, column_of(1), &n0, &n1, &n2, &n3, &sn, &dn, &en_);
        double alpha1 = 2.0 * (sn * dd - dn * sd) / (sd * sd);
        double across = normalize_across_scale ? sigma : 1.0;
        double scale = across / alpha1;
        n0 = n0 * scale; n1 = n1 * scale; n2 = n2 * scale; n3 = n3 * scale;
        symmetric = 0;
    } else {
        double a0, a1_, a2_, a3, sn0, dn0, en0;
        double b0, b1_, b2_, b3, sn2, dn2, en2;
        compute_n(sigmad, column_of(0), &a0, &a1_, &a2_, &a3, &sn0, &dn0, &en0);
        compute_n(sigmad, column_of(2), &b0, &b1_, &b2_, &b3, &sn2, &dn2, &en2);
        double beta = -(2.0 * sn2 - sd * b0) / (2.0 * sn0 - sd * a0);
        n0 = b0 + beta * a0;
        n1 = b1_ + beta * a1_;
        n2 = b2_ + beta * a2_;
        n3 = b3 + beta * a3;
        double sn = sn2 + beta * sn0;
        double dn = dn2 + beta * dn0;
        double en = en2 + beta * en0;
        double alpha2 = (en * sd * sd - ed * sn * sd - 2.0 * dn * dd * sd + 2.0 * dd * dd * sn) /
                        (sd * sd * sd);
        double across = normalize_across_scale ? sigma * sigma : 1.0;
        double scale = across / alpha2;
        n0 = n0 * scale; n1 = n1 * scale; n2 = n2 * scale; n3 = n3 * scale;
        symmetric = 1;
    }
    double m1, m2, m3, m4;
    if (symmetric) {
        m1 = n1 - d1 * n0; m2 = n2 - d2 * n0; m3 = n3 - d3 * n0; m4 = -d4 * n0;
    } else {
        m1 = -(n1 - d1 * n0); m2 = -(n2 - d2 * n0); m3 = -(n3 - d3 * n0); m4 = d4 * n0;
    }
    double sn = n0 + n1 + n2 + n3;
    double sm = m1 + m2 + m3 + m4;
    out[0] = n0; out[1] = n1; out[2] = n2; out[3] = n3;
    out[4] = d1; out[5] = d2; out[6] = d3; out[7] = d4;
    out[8] = m1; out[9] = m2; out[10] = m3; out[11] = m4;
    out[12] = d1 * sn / sd; out[13] = d2 * sn / sd; out[14] = d3 * sn / sd; out[15] = d4 * sn / sd;
    out[16] = d1 * sm / sd; out[17] = d2 * sm / sd; out[18] = d3 * sm / sd; out[19] = d4 * sm / sd;
}

/* src/filters/recursive_gaussian.rs:472-528, operation for operation. */
void orc_filter_line(const double *data, size_t ln, const double c[20], double *outs,
                     double *scratch) {
    const double n0 = c[0], n1 = c[1], n2 = c[2], n3 = c[3];
    const double d1 = c[4], d2 = c[5], d3 = c[6], d4 = c[7];
    const double m1 = c[8], m2 = c[9], m3 = c[10], m4 = c[11];
    const double bn1 = c[12], bn2 = c[13], bn3 = c[14], bn4 = c[15];
    const double bm1 = c[16], bm2 = c[17], bm3 = c[18], bm4 = c[19];

    double v1 = data[0];
    outs[0] = v1 * n0 + v1 * n1 + v1 * n2 + v1 * n3;
    outs[1] = data[1] * n0 + v1 * n1 + v1 * n2 + v1 * n3;
    outs[2] = data[2] * n0 + data[1] * n1 + v1 * n2 + v1 * n3;
    outs[3] = data[3] * n0 + data[2] * n1 + data[1] * n2 + v1 * n3;
    outs[0] -= v1 * bn1 + v1 * bn2 + v1 * bn3 + v1 * bn4;
    outs[1] -= outs[0] * d1 + v1 * bn2 + v1 * bn3 + v1 * bn4;
    outs[2] -= outs[1] * d1 + outs[0] * d2 + v1 * bn3 + v1 * bn4;
    outs[3] -= outs[2] * d1 + outs[1] * d2 + outs[0] * d3 + v1 * bn4;
    for (size_t i = 4; i < ln; ++i) {
        outs[i] = data[i] * n0 + data[i - 1] * n1 + data[i - 2] * n2 + data[i - 3] * n3;
        outs[i] -= outs[i - 1] * d1 + outs[i - 2] * d2 + outs[i - 3] * d3 + outs[i - 4] * d4;
    }

    double v2 = data[ln - 1];
    scratch[ln - 1] = v2 * m1 + v2 * m2 + v2 * m3 + v2 * m4;
    scratch[ln - 2] = data[ln - 1] * m1 + v2 * m2 + v2 * m3 + v2 * m4;
    scratch[ln - 3] = data[ln - 2] * m1 + data[ln - 1] * m2 + v2 * m3 + v2 * m4;
    scratch[ln - 4] = data[ln - 3] * m1 + data[ln - 2] * m2 + data[ln - 1] * m3 + v2 * m4;
    scratch[ln - 1] -= v2 * bm1 + v2 * bm2 + v2 * bm3 + v2 * bm4;
    scratch[ln - 2] -= scratch[ln - 1] * d1 + v2 * bm2 + v2 * bm3 + v2 * bm4;
    scratch[ln - 3] -= scratch[ln - 2] * d1 + scratch[ln - 1] * d2 + v2 * bm3 + v2 * bm4;
    scratch[ln - 4] -=
        scratch[ln - 3] * d1 + scratch[ln - 2] * d2 + scratch[ln - 1] * d3 + v2 * bm4;
    for (size_t i = ln - 4; i > 0; --i) {
        scratch[i - 1] = data[i] * m1 + data[i + 1] * m2 + data[i + 2] * m3 + data[i + 3] * m4;
        scratch[i - 1] -= scratch[i] * d1 + scratch[i + 1] * d2 + scratch[i + 2] * d3 +
                          scratch[i + 3] * d4;
    }
    for (size_t i = 0; i < ln; ++i) outs[i] += scratch[i];
}

typedef struct {
    double *buf;
    const size_t *size;
    int axis;
    const double *coef;
    size_t line_begin, line_end;
} axis_job;

static void filter_lines(axis_job *j) {
    const size_t *size = j->size;
    int d = j->axis;
    size_t ln = size[d];
    size_t stride = d == 0 ? 1 : (d == 1 ? size[0] : size[0] * size[1]);
    double *line = (double *)malloc(ln * sizeof(double) * 3);
    double *outs = line + ln, *scratch = line + 2 * ln;
    for (size_t l = j->line_begin; l < j->line_end; ++l) {
        /* line l enumerates the other two axes, lower axis fastest */
        size_t start;
        if (d == 0) start = l * size[0];
        else if (d == 1) start = (l % size[0]) + (l / size[0]) * size[0] * size[1];
        else start = l;
        for (size_t k = 0; k < ln; ++k) line[k] = j->buf[start + k * stride];
        orc_filter_line(line, ln, j->coef, outs, scratch);
        for (size_t k = 0; k < ln; ++k) j->buf[start + k * stride] = outs[k];
    }
    free(line);
}

static void *filter_lines_thread(void *arg) { filter_lines((axis_job *)arg); return NULL; }

/* src/filters/recursive_gaussian.rs:171-204 (axis loop, in order 0,1,2). */
static int rg_f64(double *buf, const size_t size[3], const double spacing[3],
                  const double sigma[3], int threads) {
    for (int d = 0; d < 3; ++d)
        if (sigma[d] < 0.0) return -2;
    for (int d = 0; d < 3; ++d) {
        if (sigma[d] <= 0.0) continue;
        if (size[d] < 4) return -1;
        double coef[20];
        orc_deriche_coefficients(0, sigma[d] / spacing[d], sigma[d], 0, coef);
        size_t nlines = size[0] * size[1] * size[2] / size[d];
        if (threads <= 1) {
            axis_job j = {buf, size, d, coef, 0, nlines};
            filter_lines(&j);
        } else {
            pthread_t tid[256];
            axis_job jobs[256];
            int nt = threads > 256 ? 256 : threads;
            for (int t = 0; t < nt; ++t) {
                jobs[t].buf = buf; jobs[t].size = size; jobs[t].axis = d; jobs[t].coef = coef;
                jobs[t].line_begin = nlines * (size_t)t / (size_t)nt;
                jobs[t].line_end = nlines * (size_t)(t + 1) / (size_t)nt;
                pthread_create(&tid[t], NULL, filter_lines_thread, &jobs[t]);
            }
            for (int t = 0; t < nt; ++t) pthread_join(tid[t], NULL);
        }
    }
    return 0;
}

int orc_recursive_gaussian_f64(double *buf, const size_t size[3], const double spacing[3],
                               const double sigma[3]) {
    return rg_f64(buf, size, spacing, sigma, 1);
}

static int srg_f32(const float *in, float *out, const size_t size[3], const double spacing[3],
                   const double sigma[3], int threads) {
    size_t n = size[0] * size[1] * size[2];
    double *buf = (double *)malloc(n * sizeof(double));
    for (size_t i = 0; i < n; ++i) buf[i] = (double)in[i];
    int rc = rg_f64(buf, size, spacing, sigma, threads);
    if (rc == 0)
        for (size_t i = 0; i < n; ++i) out[i] = (float)buf[i]; /* src/core/pixel.rs:434 */
    free(buf);
    return rc;
}

int orc_smoothing_recursive_gaussian_f32(const float *in, float *out, const size_t size[3],
                                         const double spacing[3], const double sigma[3]) {
    return srg_f32(in, out, size, spacing, sigma, 1);
}

int orc_smoothing_recursive_gaussian_f32_mt(const float *in, float *out, const size_t size[3],
                                            const double spacing[3], const double sigma[3],
                                            int threads) {
    return srg_f32(in, out, size, spacing, sigma, threads);
}

/* ------------------------------------------------------------------------- */
/* src/core/coord.rs:50-74 */
void orc_index_to_physical_matrix(const double dir[9], const double spacing[3], double out[9]) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) out[r * 3 + c] = dir[r * 3 + c] * spacing[c];
}

/* src/core/matrix.rs:44-102 (Gauss-Jordan, partial pivoting, divide by pivot). */
int orc_invert3(const double m[9], double out[9]) {
    const int n = 3, w = 6;
    double a[18];
    memset(a, 0, sizeof(a));
    for (int r = 0; r < n; ++r) {
        for (int c = 0; c < n; ++c) a[r * w + c] = m[r * n + c];
        a[r * w + n + r] = 1.0;
    }
    for (int col = 0; col < n; ++col) {
        int pivot = col;
        double best = fabs(a[col * w + col]);
        for (int r = col + 1; r < n; ++r) {
            double v = fabs(a[r * w + col]);
            if (v > best) { best = v; pivot = r; }
        }
        if (best < 1e-12) return -1;
        if (pivot != col)
            for (int c = 0; c < w; ++c) {
                double t = a[col * w + c]; a[col * w + c] = a[pivot * w + c]; a[pivot * w + c] = t;
            }
        double diag = a[col * w + col];
        for (int c = 0; c < w; ++c) a[col * w + c] /= diag;
        for (int r = 0; r < n; ++r) {
            if (r == col) continue;
            double factor = a[r * w + col];
            if (factor == 0.0) continue;
            for (int c = 0; c < w; ++c) a[r * w + c] -= factor * a[col * w + c];
        }
    }
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) out[r * n + c] = a[r * w + n + c];
    return 0;
}

/* src/core/matrix.rs:9-21 */
static void mat_vec3(const double m[9], const double v[3], double out[3]) {
    for (int r = 0; r < 3; ++r) {
        double acc = 0.0;
        for (int c = 0; c < 3; ++c) acc += m[r * 3 + c] * v[c];
        out[r] = acc;
    }
}

/* src/core/matrix.rs:24-40 (skips zero a[r][k]) */
static void matmul3(const double a[9], const double b[9], double out[9]) {
    for (int i = 0; i < 9; ++i) out[i] = 0.0;
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) {
            double ark = a[r * 3 + k];
            if (ark == 0.0) continue;
            for (int c = 0; c < 3; ++c) out[r * 3 + c] += ark * b[k * 3 + c];
        }
}

/* ------------------------------------------------------------------------- */
/* Transforms. Translation parametric.rs:374-462; Affine :473-608 (offset
 * :513-519); Euler3D :1030-1179. */
void orc_transform_recompute(orc_transform *t) {
    if (t->kind == ORC_TRANSLATION) {
        t->nparams = 3;
        for (int i = 0; i < 9; ++i) t->matrix[i] = (i % 4 == 0) ? 1.0 : 0.0;
        for (int d = 0; d < 3; ++d) t->offset[d] = t->params[d];
    } else if (t->kind == ORC_EULER3D) {
        t->nparams = 6;
        double cx = cos(t->params[0]), sx = sin(t->params[0]);
        double cy = cos(t->params[1]), sy = sin(t->params[1]);
        double cz = cos(t->params[2]), sz = sin(t->params[2]);
        double rx[9] = {1.0, 0.0, 0.0, 0.0, cx, -sx, 0.0, sx, cx};
        double ry[9] = {cy, 0.0, sy, 0.0, 1.0, 0.0, -sy, 0.0, cy};
        double rz[9] = {cz, -sz, 0.0, sz, cz, 0.0, 0.0, 0.0, 1.0};
        double tmp[9];
        if (t->compute_zyx) { matmul3(rz, ry, tmp); matmul3(tmp, rx, t->matrix); }
        else { matmul3(rz, rx, tmp); matmul3(tmp, ry, t->matrix); }
        double mc[3];
        mat_vec3(t->matrix, t->center, mc);
        for (int i = 0; i < 3; ++i) t->offset[i] = t->params[3 + i] + t->center[i] - mc[i];
    } else if (t->kind == ORC_AFFINE) {
        t->nparams = 12;
        for (int i = 0; i < 9; ++i) t->matrix[i] = t->params[i];
        double ac[3];
        mat_vec3(t->matrix, t->center, ac);
        for (int d = 0; d < 3; ++d) t->offset[d] = t->center[d] + t->params[9 + d] - ac[d];
    }
}

/* ---- B-spline transform (src/transform/bspline.rs) ---- */
static int bspline_support(const orc_transform *t, const double p[3], long start[3],
                           double wts[3][4]);

void orc_transform_point(const orc_transform *t, const double p[3], double out[3]) {
    if (t->kind == ORC_TRANSLATION) {
        for (int d = 0; d < 3; ++d) out[d] = p[d] + t->params[d];
    } else if (t->kind == ORC_BSPLINE) {
        long start[3];
        double w[3][4];
        double disp[3] = {0.0, 0.0, 0.0};
        if (!bspline_support(t, p, start, w)) { /* outside: the point itself */
            out[0] = p[0]; out[1] = p[1]; out[2] = p[2];
            return;
        }
        {
            size_t gs0 = t->grid_size[0], gs1 = t->grid_size[1];
            size_t ncp = t->grid_size[0] * t->grid_size[1] * t->grid_size[2];
            for (int k = 0; k < 4; ++k)
                for (int j = 0; j < 4; ++j)
                    for (int i = 0; i < 4; ++i) {
                        double wt = w[0][i] * w[1][j] * w[2][k];
                        size_t cp = (size_t)(start[0] + i) + gs0 * ((size_t)(start[1] + j) + gs1 * (size_t)(start[2] + k));
                        for (int d = 0; d < 3; ++d) disp[d] += wt * t->coeffs[d * ncp + cp];
                    }
        }
        for (int d = 0; d < 3; ++d) out[d] = p[d] + disp[d];
    } else {
        double mx[3];
        mat_vec3(t->matrix, p, mx);
        for (int d = 0; d < 3; ++d) out[d] = mx[d] + t->offset[d];
    }
}

void orc_transform_jacobian(const orc_transform *t, const double point[3], double *j) {
    int np = t->nparams;
    for (int i = 0; i < 3 * np; ++i) j[i] = 0.0;
    if (t->kind == ORC_TRANSLATION) {
        for (int i = 0; i < 3; ++i) j[i * 3 + i] = 1.0;
    } else if (t->kind == ORC_AFFINE) {
        for (int i = 0; i < 3; ++i) {
            for (int c = 0; c < 3; ++c) j[i * np + (i * 3 + c)] = point[c] - t->center[c];
            j[i * np + (9 + i)] = 1.0;
        }
    } else if (t->kind == ORC_EULER3D) {
        double cx = cos(t->params[0]), sx = sin(t->params[0]);
        double cy = cos(t->params[1]), sy = sin(t->params[1]);
        double cz = cos(t->params[2]), sz = sin(t->params[2]);
        double px = point[0] - t->center[0], py = point[1] - t->center[1],
               pz = point[2] - t->center[2];
        if (t->compute_zyx) {
            j[0] = (cz * sy * cx + sz * sx) * py + (-cz * sy * sx + sz * cx) * pz;
            j[6] = (sz * sy * cx - cz * sx) * py + (-sz * sy * sx - cz * cx) * pz;
            j[12] = (cy * cx) * py + (-cy * sx) * pz;
            j[1] = (-cz * sy) * px + (cz * cy * sx) * py + (cz * cy * cx) * pz;
            j[7] = (-sz * sy) * px + (sz * cy * sx) * py + (sz * cy * cx) * pz;
            j[13] = (-cy) * px + (-sy * sx) * py + (-sy * cx) * pz;
            j[2] = (-sz * cy) * px + (-sz * sy * sx - cz * cx) * py + (-sz * sy * cx + cz * sx) * pz;
            j[8] = (cz * cy) * px + (cz * sy * sx - sz * cx) * py + (cz * sy * cx + sz * sx) * pz;
            j[14] = 0.0;
        } else {
            j[0] = (-sz * cx * sy) * px + (sz * sx) * py + (sz * cx * cy) * pz;
            j[6] = (cz * cx * sy) * px + (-cz * sx) * py + (-cz * cx * cy) * pz;
            j[12] = (sx * sy) * px + (cx) * py + (-sx * cy) * pz;
            j[1] = (-cz * sy - sz * sx * cy) * px + (cz * cy - sz * sx * sy) * pz;
            j[7] = (-sz * sy + cz * sx * cy) * px + (sz * cy + cz * sx * sy) * pz;
            j[13] = (-cx * cy) * px + (-cx * sy) * pz;
            j[2] = (-sz * cy - cz * sx * sy) * px + (-cz * cx) * py + (-sz * sy + cz * sx * cy) * pz;
            j[8] = (cz * cy - sz * sx * sy) * px + (-sz * cx) * py + (cz * sy + sz * sx * cy) * pz;
            j[14] = 0.0;
        }
        j[3] = 1.0;
        j[10] = 1.0;
        j[17] = 1.0;
    } else if (t->kind == ORC_BSPLINE) {
        /* bspline.rs:661-703: column d*ncp + flat holds the tap weight in row d;
         * a point outside the valid region has a zero Jacobian. */
        long start[3];
        double w[3][4];
        if (!bspline_support(t, point, start, w)) return;
        size_t gs0 = t->grid_size[0], gs1 = t->grid_size[1];
        size_t ncp = t->grid_size[0] * t->grid_size[1] * t->grid_size[2];
        for (int k = 0; k < 4; ++k)
            for (int jj = 0; jj < 4; ++jj)
                for (int i = 0; i < 4; ++i) {
                    double wt = w[0][i] * w[1][jj] * w[2][k];
                    size_t cp = (size_t)(start[0] + i) +
                                gs0 * ((size_t)(start[1] + jj) + gs1 * (size_t)(start[2] + k));
                    for (int d = 0; d < 3; ++d) j[(size_t)d * (size_t)np + d * ncp + cp] = wt;
                }
    }
}

void orc_transform_stage(const orc_transform *t, double stage[12]) {
    for (int i = 0; i < 9; ++i) stage[i] = t->matrix[i];
    for (int d = 0; d < 3; ++d) stage[9 + d] = t->offset[d];
}

/* src/transform/matrix_offset.rs:102-109 */
void orc_replay_stages(const double *stages, int nstage, const double p[3], double out[3]) {
    double q[3] = {p[0], p[1], p[2]};
    for (int s = 0; s < nstage; ++s) {
        double mq[3];
        mat_vec3(stages + 12 * s, q, mq);
        for (int d = 0; d < 3; ++d) q[d] = mq[d] + stages[12 * s + 9 + d];
    }
    out[0] = q[0]; out[1] = q[1]; out[2] = q[2];
}

/* ------------------------------------------------------------------------- */
/* src/registration/metric.rs:101-122 */
uint64_t orc_splitmix64_next(uint64_t *state) {
    *state += 0x9E3779B97F4A7C15ULL;
    uint64_t z = *state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* src/registration/metric.rs:419-437 */
size_t orc_draw_samples(size_t voxels, int strategy, double percentage, uint64_t seed,
                        int64_t *out) {
    if (strategy == 0) return (size_t)-1;
    if (strategy == 1) {
        size_t stride = (size_t)ceil(1.0 / percentage);
        if (stride < 1) stride = 1;
        size_t n = 0;
        for (size_t v = 0; v < voxels; v += stride) {
            if (out) out[n] = (int64_t)v;
            ++n;
        }
        return n;
    }
    size_t count = (size_t)((double)voxels * percentage);
    uint64_t state = seed;
    for (size_t i = 0; i < count; ++i) {
        uint64_t r = orc_splitmix64_next(&state) % (uint64_t)voxels;
        if (out) out[i] = (int64_t)r;
    }
    return count;
}

/* ------------------------------------------------------------------------- */
/* Sample enumeration: FixedSamples::from_image_with (metric.rs:494-580). The
 * mask filters the draw; points come from VirtualGrid::write_point. */
typedef struct {
    const orc_problem *pb;
    size_t n;          /* raw selection length */
} sample_iter;

/* Explicit points (FixedPoints::Explicit, cuda/ops/resident.rs:464-484):
 * sample s is the point pts[s] with the gathered value fvals[s]; there is no
 * grid index, so no fixed mask applies (error.rs MaskedExplicitPoints). */
static size_t selection_len(const orc_problem *pb) {
    if (pb->pts) return pb->nidx;
    return pb->idx ? pb->nidx : pb->fsize[0] * pb->fsize[1] * pb->fsize[2];
}
static size_t selection_flat(const orc_problem *pb, size_t s) {
    return pb->idx ? (size_t)pb->idx[s] : s;
}
static int selection_allowed(const orc_problem *pb, size_t flat) {
    return pb->fmask && !pb->pts ? pb->fmask[flat] != 0 : 1;
}

/* src/registration/scales.rs:259-271 (origin-first fold). */
static void write_point(const orc_problem *pb, size_t flat, double out[3]) {
    if (pb->pts) {
        for (int r = 0; r < 3; ++r) out[r] = pb->pts[3 * flat + r];
        return;
    }
    size_t id[3];
    for (int d = 0; d < 3; ++d) { id[d] = flat % pb->fsize[d]; flat /= pb->fsize[d]; }
    for (int r = 0; r < 3; ++r) {
        double acc = pb->forigin[r];
        for (int c = 0; c < 3; ++c) acc += pb->fi2p[r * 3 + c] * (double)id[c];
        out[r] = acc;
    }
}

int orc_fixed_value_range(const orc_problem *pb, double out[2]) {
    double lo = INFINITY, hi = -INFINITY;
    int any = 0;
    size_t n = selection_len(pb);
    for (size_t s = 0; s < n; ++s) {
        size_t flat = selection_flat(pb, s);
        if (!selection_allowed(pb, flat)) continue;
        double v = (double)pb->fvals[flat];
        lo = fmin(lo, v);
        hi = fmax(hi, v);
        any = 1;
    }
    out[0] = any ? lo : 0.0;
    out[1] = any ? hi : 0.0;
    return any;
}

/* src/registration/metric.rs:913-948 */
int orc_moving_value_range(const orc_problem *pb, double out[2]) {
    double lo = INFINITY, hi = -INFINITY;
    int any = 0;
    size_t n = pb->msize[0] * pb->msize[1] * pb->msize[2];
    for (size_t i = 0; i < n; ++i) {
        if (pb->mmask && !pb->mmask[i]) continue;
        double v = (double)pb->mvals[i];
        lo = fmin(lo, v);
        hi = fmax(hi, v);
        any = 1;
    }
    out[0] = any ? lo : 0.0;
    out[1] = any ? hi : 0.0;
    return any;
}

/* metric.rs:795-807; Rust's f64 sum seeds at -0.0, which adds exactly. */
static void continuous_index(const orc_problem *pb, const double p[3], double c[3]) {
    for (int r = 0; r < 3; ++r) {
        double acc = -0.0;
        for (int j = 0; j < 3; ++j) acc += pb->mp2i[r * 3 + j] * (p[j] - pb->morigin[j]);
        c[r] = acc;
    }
}

/* metric.rs:817-831 with coord.rs:189-191 */
static int mask_allows(const orc_problem *pb, const double c[3]) {
    if (!pb->mmask) return 1;
    size_t flat = 0, stride = 1;
    for (int d = 0; d < 3; ++d) {
        int64_t r = (int64_t)floor(c[d] + 0.5);
        if (r < 0 || (uint64_t)r >= pb->msize[d]) return 0;
        flat += (size_t)r * stride;
        stride *= pb->msize[d];
    }
    return pb->mmask[flat] != 0;
}

/* interpolator.rs:29-34 */
static int is_inside(const double c[3], const size_t size[3]) {
    for (int d = 0; d < 3; ++d)
        if (!(c[d] >= -0.5 && c[d] < (double)size[d] - 0.5)) return 0;
    return 1;
}

static long clampl(long v, long lo, long hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* interpolator.rs:58-91 */
static int linear_at(const orc_problem *pb, const double c[3], double *value) {
    if (!is_inside(c, pb->msize)) return 0;
    long base[3];
    double frac[3];
    for (int d = 0; d < 3; ++d) {
        double f = floor(c[d]);
        base[d] = (long)f;
        frac[d] = c[d] - f;
    }
    size_t strides[3] = {1, pb->msize[0], pb->msize[0] * pb->msize[1]};
    double acc = 0.0;
    for (int corner = 0; corner < 8; ++corner) {
        double weight = 1.0;
        size_t offset = 0;
        for (int d = 0; d < 3; ++d) {
            int bit = (corner >> d) & 1;
            weight *= bit ? frac[d] : 1.0 - frac[d];
            long idx = clampl(base[d] + bit, 0, (long)pb->msize[d] - 1);
            offset += (size_t)idx * strides[d];
        }
        if (weight != 0.0) acc += weight * (double)pb->mvals[offset];
    }
    *value = acc;
    return 1;
}

/* interpolator.rs:103-149 */
static int linear_value_and_gradient(const orc_problem *pb, const double c[3], double *value,
                                     double grad[3]) {
    if (!is_inside(c, pb->msize)) return 0;
    long base[3];
    double frac[3];
    for (int d = 0; d < 3; ++d) {
        double f = floor(c[d]);
        base[d] = (long)f;
        frac[d] = c[d] - f;
    }
    size_t strides[3] = {1, pb->msize[0], pb->msize[0] * pb->msize[1]};
    double v = 0.0;
    double g[3] = {0.0, 0.0, 0.0};
    for (int corner = 0; corner < 8; ++corner) {
        size_t offset = 0;
        double weight = 1.0;
        for (int d = 0; d < 3; ++d) {
            int bit = (corner >> d) & 1;
            weight *= bit ? frac[d] : 1.0 - frac[d];
            long idx = clampl(base[d] + bit, 0, (long)pb->msize[d] - 1);
            offset += (size_t)idx * strides[d];
        }
        double b = (double)pb->mvals[offset];
        v += weight * b;
        for (int jj = 0; jj < 3; ++jj) {
            double wj = 1.0;
            for (int d = 0; d < 3; ++d) {
                if (d == jj) continue;
                int bit = (corner >> d) & 1;
                wj *= bit ? frac[d] : 1.0 - frac[d];
            }
            double sign = ((corner >> jj) & 1) ? 1.0 : -1.0;
            g[jj] += sign * wj * b;
        }
    }
    *value = v;
    grad[0] = g[0]; grad[1] = g[1]; grad[2] = g[2];
    return 1;
}

/* ------------------------------------------------------------------------- */
/* Cubic B-spline interpolation (interpolator.rs:170-350). */
double orc_powi(double a, int b) {
    int recip = b < 0;
    unsigned pw = recip ? (unsigned)(-(long)b) : (unsigned)b;
    double r = 1.0;
    for (;;) {
        if (pw & 1u) r *= a;
        pw >>= 1;
        if (pw == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

/* bspline3_filter_1d (interpolator.rs:236-270) on one line of n >= 2 values. */
static void bspline3_filter_line(double *line, size_t n, double z) {
    double gain = (1.0 - z) * (1.0 - 1.0 / z);
    for (size_t k = 0; k < n; ++k) line[k] *= gain;
    double iz = 1.0 / z;
    double zn = z;
    double z2n = orc_powi(z, (int)n - 1);
    double sum = line[0] + z2n * line[n - 1];
    z2n = z2n * z2n * iz;
    for (size_t k = 1; k + 1 < n; ++k) {
        sum += (zn + z2n) * line[k];
        zn *= z;
        z2n *= iz;
    }
    line[0] = sum / (1.0 - zn * zn);
    for (size_t k = 1; k < n; ++k) line[k] += z * line[k - 1];
    line[n - 1] = (z / (z * z - 1.0)) * (z * line[n - 2] + line[n - 1]);
    for (size_t k = n - 1; k-- > 0;) line[k] = z * (line[k + 1] - line[k]);
}

void orc_bspline_coefficients(const float *in, const size_t size[3], double *out) {
    size_t total = size[0] * size[1] * size[2];
    for (size_t i = 0; i < total; ++i) out[i] = (double)in[i];
    double pole = sqrt(3.0) - 2.0;
    size_t strides[3] = {1, size[0], size[0] * size[1]};
    for (int axis = 0; axis < 3; ++axis) {
        size_t n = size[axis];
        if (n == 1) continue;
        double *line = (double *)malloc(n * sizeof(double));
        int o1 = axis == 0 ? 1 : 0, o2 = axis == 2 ? 1 : 2;
        for (size_t j2 = 0; j2 < size[o2]; ++j2)
            for (size_t j1 = 0; j1 < size[o1]; ++j1) {
                size_t base = j1 * strides[o1] + j2 * strides[o2];
                for (size_t k = 0; k < n; ++k) line[k] = out[base + k * strides[axis]];
                bspline3_filter_line(line, n, pole);
                for (size_t k = 0; k < n; ++k) out[base + k * strides[axis]] = line[k];
            }
        free(line);
    }
}

int orc_bspline_value_and_gradient(const double *coef, const size_t size[3], const double c[3],
                                   double *value, double grad[3]) {
    if (!is_inside(c, size)) return 0;
    size_t idx[3][4];
    double w[3][4], dw[3][4];
    size_t strides[3] = {1, size[0], size[0] * size[1]};
    for (int d = 0; d < 3; ++d) {
        long base = (long)floor(c[d]) - 1;
        long n = (long)size[d];
        for (int k = 0; k < 4; ++k) {
            long i = base + k;
            if (n == 1) {
                i = 0;
            } else {
                if (i < 0) i = -i;
                if (i > n - 1) i = 2 * (n - 1) - i;
            }
            idx[d][k] = (size_t)i;
        }
        double frac = c[d] - (double)(base + 1);
        double w3 = (1.0 / 6.0) * frac * frac * frac;
        double w0 = 1.0 / 6.0 + 0.5 * frac * (frac - 1.0) - w3;
        double w2 = frac + w0 - 2.0 * w3;
        double w1 = 1.0 - w0 - w2 - w3;
        w[d][0] = w0; w[d][1] = w1; w[d][2] = w2; w[d][3] = w3;
        double a = frac - 0.5;
        double dw2 = 0.75 - a * a;
        double dw3 = 0.5 * (a - dw2 + 1.0);
        double dw1 = 1.0 - dw2 - dw3;
        dw[d][0] = -dw1; dw[d][1] = dw1 - dw2; dw[d][2] = dw2 - dw3; dw[d][3] = dw3;
    }
    double v = 0.0, g[3] = {0.0, 0.0, 0.0};
    for (int corner = 0; corner < 64; ++corner) {
        int tap[3] = {corner % 4, (corner / 4) % 4, corner / 16};
        size_t offset = 0;
        double wprod = 1.0;
        for (int d = 0; d < 3; ++d) {
            offset += idx[d][tap[d]] * strides[d];
            wprod *= w[d][tap[d]];
        }
        double cv = coef[offset];
        v += wprod * cv;
        for (int q = 0; q < 3; ++q) {
            double gprod = 1.0;
            for (int d = 0; d < 3; ++d) gprod *= d == q ? dw[d][tap[d]] : w[d][tap[d]];
            g[q] += gprod * cv;
        }
    }
    *value = v;
    grad[0] = g[0]; grad[1] = g[1]; grad[2] = g[2];
    return 1;
}

static int moving_value_and_index_gradient(const orc_problem *pb, const double c[3],
                                           double *value, double gi[3]) {
    if (pb->mcoef) return orc_bspline_value_and_gradient(pb->mcoef, pb->msize, c, value, gi);
    return linear_value_and_gradient(pb, c, value, gi);
}

/* metric.rs:848-864 */
int orc_moving_value_and_physical_gradient(const orc_problem *pb, const double p[3],
                                           double *value, double grad[3]) {
    double c[3], gi[3];
    continuous_index(pb, p, c);
    if (!mask_allows(pb, c)) return 0;
    if (!moving_value_and_index_gradient(pb, c, value, gi)) return 0;
    for (int d = 0; d < 3; ++d) {
        double acc = -0.0;
        for (int j = 0; j < 3; ++j) acc += gi[j] * pb->mp2i[j * 3 + d];
        grad[d] = acc;
    }
    return 1;
}

/* metric.rs:881-890 */
int orc_moving_value_at(const orc_problem *pb, const double p[3], double *value) {
    double c[3];
    continuous_index(pb, p, c);
    if (!mask_allows(pb, c)) return 0;
    if (pb->mcoef) {
        double gi[3];
        return orc_bspline_value_and_gradient(pb->mcoef, pb->msize, c, value, gi);
    }
    return linear_at(pb, c, value);
}

/* ------------------------------------------------------------------------- */
/* mattes.rs:92-124 */
#define PADDING 2
double orc_cubic_bspline(double u) {
    double a = fabs(u);
    if (a < 1.0) {
        double sq = a * a;
        return (4.0 - 6.0 * sq + 3.0 * sq * a) / 6.0;
    } else if (a < 2.0) {
        double sq = a * a;
        return (8.0 - 12.0 * a + 6.0 * sq - sq * a) / 6.0;
    }
    return 0.0;
}

double orc_cubic_bspline_derivative(double u) {
    if (u >= 0.0 && u < 1.0) return -2.0 * u + 1.5 * u * u;
    if (u > -1.0 && u < 0.0) return -2.0 * u - 1.5 * u * u;
    if (u >= 1.0 && u < 2.0) return -2.0 + 2.0 * u - 0.5 * u * u;
    if (u > -2.0 && u <= -1.0) return 2.0 + 2.0 * u + 0.5 * u * u;
    return 0.0;
}

/* mattes.rs:152-186 */
int orc_mattes_geometry(const double fr[2], const double mr[2], size_t bins, orc_mattes_geom *g) {
    if (bins < 2 * PADDING + 1) return -1;
    if (fr[1] - fr[0] <= DBL_EPSILON) return -2;
    if (mr[1] - mr[0] <= DBL_EPSILON) return -3;
    double denom = (double)(bins - 2 * PADDING);
    g->num_bins = bins;
    g->moving_true_min = mr[0];
    g->moving_true_max = mr[1];
    g->fixed_bin_size = (fr[1] - fr[0]) / denom;
    g->moving_bin_size = (mr[1] - mr[0]) / denom;
    g->fixed_normalized_min = fr[0] / g->fixed_bin_size - (double)PADDING;
    g->moving_normalized_min = mr[0] / g->moving_bin_size - (double)PADDING;
    return 0;
}

/* mattes.rs:193-206 */
size_t orc_parzen_window_index(const orc_mattes_geom *g, double value, double bin_size,
                               double normalized_min) {
    double term = value / bin_size - normalized_min;
    long index = (long)term; /* `as isize`: truncation, saturating */
    long lo = PADDING, hi = (long)g->num_bins - PADDING - 1;
    if (index < lo) index = lo;
    else if (index > hi) index = hi;
    return (size_t)index;
}

/* The joint-PDF mass is Kahan-summed by the current reference
 * (mattes.rs:268-270); before its ledger row 2.161 it was a naive
 * `iter().sum()`. The reference-held value -1.063357897140305
 * (doc/upstream-findings.md:273) predates that fix, so the test that pins it
 * switches this to the naive walk. */
static int g_joint_sum_compensated = 1;
void orc_set_joint_sum_compensated(int on) { g_joint_sum_compensated = on; }

static double joint_pdf_sum(const double *v, size_t n) {
    if (g_joint_sum_compensated) return orc_compensated_sum(v, n);
    double s = -0.0;
    for (size_t i = 0; i < n; ++i) s += v[i];
    return s;
}

/* src/core/compensated.rs:115-128, 202-204 (residual not folded back). */
double orc_compensated_sum(const double *v, size_t n) {
    double sum = 0.0, comp = 0.0;
    for (size_t i = 0; i < n; ++i) {
        double ci = v[i] - comp;
        double ts = sum + ci;
        comp = (ts - sum) - ci;
        sum = ts;
    }
    return sum;
}

/* moving_point = moving_initial(optimized(x)) (registration/method.rs:198-240):
 * the optimized transform, then the moving-initial one, each rounding as its
 * own transform_point. */
static void composed_point(const orc_transform *t, const orc_transform *init, const double fp[3],
                           double mp[3]) {
    if (!init) {
        orc_transform_point(t, fp, mp);
        return;
    }
    double y[3];
    orc_transform_point(t, fp, y);
    orc_transform_point(init, y, mp);
}

static void mattes_histogram_composed(const orc_problem *pb, const orc_mattes_geom *g,
                                      const orc_transform *t, const orc_transform *init,
                                      double *joint, double *marginal, size_t *valid_out);

/* mattes.rs:442-486 */
void orc_mattes_histogram(const orc_problem *pb, const orc_mattes_geom *g,
                          const orc_transform *t, double *joint, double *marginal,
                          size_t *valid_out) {
    mattes_histogram_composed(pb, g, t, NULL, joint, marginal, valid_out);
}

static void mattes_histogram_composed(const orc_problem *pb, const orc_mattes_geom *g,
                                      const orc_transform *t, const orc_transform *init,
                                      double *joint, double *marginal, size_t *valid_out) {
    size_t bins = g->num_bins;
    memset(joint, 0, bins * bins * sizeof(double));
    memset(marginal, 0, bins * sizeof(double));
    size_t valid = 0;
    size_t n = selection_len(pb);
    for (size_t s = 0; s < n; ++s) {
        size_t flat = selection_flat(pb, s);
        if (!selection_allowed(pb, flat)) continue;
        double fp[3], mp[3], mv;
        write_point(pb, flat, fp);
        double fv = (double)pb->fvals[flat];
        composed_point(t, init, fp, mp);
        if (!orc_moving_value_at(pb, mp, &mv)) continue;
        if (mv < g->moving_true_min || mv > g->moving_true_max) continue;
        double mterm = mv / g->moving_bin_size - g->moving_normalized_min;
        size_t mi = orc_parzen_window_index(g, mv, g->moving_bin_size, g->moving_normalized_min);
        size_t fi = orc_parzen_window_index(g, fv, g->fixed_bin_size, g->fixed_normalized_min);
        marginal[fi] += 1.0;
        for (size_t m = mi - 1; m < mi + 3; ++m) {
            double arg = (double)m - mterm;
            joint[fi * bins + m] += orc_cubic_bspline(arg);
        }
        valid += 1;
    }
    *valid_out = valid;
}

/* mattes.rs:273-328 */
int orc_mattes_tail(double *joint, double *marginal, size_t valid, const orc_mattes_geom *g,
                    double *value, double *pratio) {
    size_t bins = g->num_bins;
    if (valid == 0) { *value = DBL_MAX; return 0; }
    double jsum = joint_pdf_sum(joint, bins * bins);
    if (jsum < DBL_EPSILON) { *value = DBL_MAX; return 0; }
    double n_factor = 1.0 / (g->moving_bin_size * (double)valid);
    double inv = 1.0 / jsum;
    for (size_t i = 0; i < bins * bins; ++i) joint[i] *= inv;
    for (size_t i = 0; i < bins; ++i) marginal[i] *= inv;
    double *mm = (double *)calloc(bins, sizeof(double));
    for (size_t f = 0; f < bins; ++f)
        for (size_t m = 0; m < bins; ++m) mm[m] += joint[f * bins + m];
    double sum = 0.0;
    if (pratio) memset(pratio, 0, bins * bins * sizeof(double));
    for (size_t f = 0; f < bins; ++f) {
        double fm = marginal[f];
        if (fm <= DBL_EPSILON) continue;
        double log_fm = log(fm);
        for (size_t m = 0; m < bins; ++m) {
            double mmv = mm[m];
            double jp = joint[f * bins + m];
            if (mmv > DBL_EPSILON && jp > DBL_EPSILON) {
                double pr = log(jp / mmv);
                sum += jp * (pr - log_fm);
                if (pratio) pratio[f * bins + m] = pr * n_factor;
            }
        }
    }
    free(mm);
    *value = -sum;
    return 1;
}

/* mattes.rs:495-501 */
double orc_mattes_value(const orc_problem *pb, const orc_mattes_geom *g, const orc_transform *t) {
    size_t bins = g->num_bins, valid;
    double *joint = (double *)malloc(bins * bins * sizeof(double));
    double *marg = (double *)malloc(bins * sizeof(double));
    orc_mattes_histogram(pb, g, t, joint, marg, &valid);
    double value;
    orc_mattes_tail(joint, marg, valid, g, &value, NULL);
    free(joint);
    free(marg);
    return value;
}

static double mattes_evaluate_sparse(const orc_problem *pb, const orc_mattes_geom *g,
                                     const orc_transform *t, const orc_transform *init,
                                     double *deriv, size_t *valid_out);

/* mattes.rs:522-676 */
static double mattes_evaluate_global(const orc_problem *pb, const orc_mattes_geom *g,
                                     const orc_transform *t, double *deriv, size_t *valid_out) {
    size_t bins = g->num_bins;
    int np = t->nparams;
    double *joint = (double *)calloc(bins * bins, sizeof(double));
    double *marg = (double *)calloc(bins, sizeof(double));
    double *jpd = (double *)calloc(bins * bins * (size_t)np, sizeof(double));
    double jac[36];
    size_t valid = 0;
    size_t n = selection_len(pb);
    for (size_t s = 0; s < n; ++s) {
        size_t flat = selection_flat(pb, s);
        if (!selection_allowed(pb, flat)) continue;
        double fp[3], mp[3], mv, grad[3];
        write_point(pb, flat, fp);
        double fv = (double)pb->fvals[flat];
        orc_transform_point(t, fp, mp);
        if (!orc_moving_value_and_physical_gradient(pb, mp, &mv, grad)) continue;
        if (mv < g->moving_true_min || mv > g->moving_true_max) continue;
        double mterm = mv / g->moving_bin_size - g->moving_normalized_min;
        size_t mi = orc_parzen_window_index(g, mv, g->moving_bin_size, g->moving_normalized_min);
        size_t fi = orc_parzen_window_index(g, fv, g->fixed_bin_size, g->fixed_normalized_min);
        marg[fi] += 1.0;
        orc_transform_jacobian(t, fp, jac);
        for (size_t m = mi - 1; m < mi + 3; ++m) {
            double arg = (double)m - mterm;
            joint[fi * bins + m] += orc_cubic_bspline(arg);
            double dw = orc_cubic_bspline_derivative(arg);
            size_t base = (fi * bins + m) * (size_t)np;
            for (int k = 0; k < np; ++k) {
                double inner = 0.0;
                for (int d = 0; d < 3; ++d) inner += jac[d * np + k] * grad[d];
                jpd[base + k] += inner * dw;
            }
        }
        valid += 1;
    }
    *valid_out = valid;
    double value;
    for (int k = 0; k < np; ++k) deriv[k] = 0.0;
    if (valid == 0) { value = DBL_MAX; goto done; }
    {
        double jsum = joint_pdf_sum(joint, bins * bins);
        if (jsum < DBL_EPSILON) { value = DBL_MAX; goto done; }
        double n_factor = 1.0 / (g->moving_bin_size * (double)valid);
        for (size_t i = 0; i < bins * bins * (size_t)np; ++i) jpd[i] *= n_factor;
        double inv = 1.0 / jsum;
        for (size_t i = 0; i < bins * bins; ++i) joint[i] *= inv;
        for (size_t i = 0; i < bins; ++i) marg[i] *= inv;
        double *mm = (double *)calloc(bins, sizeof(double));
        for (size_t f = 0; f < bins; ++f)
            for (size_t m = 0; m < bins; ++m) mm[m] += joint[f * bins + m];
        double sum = 0.0;
        for (size_t f = 0; f < bins; ++f) {
            double fm = marg[f];
            if (fm <= DBL_EPSILON) continue;
            double log_fm = log(fm);
            for (size_t m = 0; m < bins; ++m) {
                double mmv = mm[m];
                double jp = joint[f * bins + m];
                if (mmv > DBL_EPSILON && jp > DBL_EPSILON) {
                    double pr = log(jp / mmv);
                    sum += jp * (pr - log_fm);
                    size_t base = (f * bins + m) * (size_t)np;
                    for (int k = 0; k < np; ++k) deriv[k] += jpd[base + k] * pr;
                }
            }
        }
        free(mm);
        value = -sum;
    }
done:
    free(joint);
    free(marg);
    free(jpd);
    return value;
}

double orc_mattes_evaluate(const orc_problem *pb, const orc_mattes_geom *g,
                           const orc_transform *t, double *deriv, size_t *valid) {
    if (t->kind == ORC_BSPLINE) return mattes_evaluate_sparse(pb, g, t, NULL, deriv, valid);
    return mattes_evaluate_global(pb, g, t, deriv, valid);
}

double orc_mattes_evaluate_bspline_composed(const orc_problem *pb, const orc_mattes_geom *g,
                                            const orc_transform *t, const orc_transform *init,
                                            double *deriv, size_t *valid) {
    if (t->kind != ORC_BSPLINE) return -1.0;
    if (init && init->kind == ORC_BSPLINE) return -1.0;
    return mattes_evaluate_sparse(pb, g, t, init, deriv, valid);
}

/* The device derivative's moment form (src/cuda/ops/mattes.rs:75-91):
 * w = sum_taps pRatio[f, m] * B3'(arg); A[d] += w*gM[d]; B[d][e] += w*gM[d]*x[e]. */
void orc_mattes_moments(const orc_problem *pb, const orc_mattes_geom *g, const orc_transform *t,
                        const double *pratio, double mom[12]) {
    size_t bins = g->num_bins;
    for (int i = 0; i < 12; ++i) mom[i] = 0.0;
    size_t n = selection_len(pb);
    for (size_t s = 0; s < n; ++s) {
        size_t flat = selection_flat(pb, s);
        if (!selection_allowed(pb, flat)) continue;
        double fp[3], mp[3], mv, grad[3];
        write_point(pb, flat, fp);
        double fv = (double)pb->fvals[flat];
        orc_transform_point(t, fp, mp);
        if (!orc_moving_value_and_physical_gradient(pb, mp, &mv, grad)) continue;
        if (mv < g->moving_true_min || mv > g->moving_true_max) continue;
        double mterm = mv / g->moving_bin_size - g->moving_normalized_min;
        size_t mi = orc_parzen_window_index(g, mv, g->moving_bin_size, g->moving_normalized_min);
        size_t fi = orc_parzen_window_index(g, fv, g->fixed_bin_size, g->fixed_normalized_min);
        double w = 0.0;
        for (size_t m = mi - 1; m < mi + 3; ++m) {
            double arg = (double)m - mterm;
            w += pratio[fi * bins + m] * orc_cubic_bspline_derivative(arg);
        }
        for (int d = 0; d < 3; ++d) {
            mom[d] += w * grad[d];
            for (int e = 0; e < 3; ++e) mom[3 + d * 3 + e] += w * grad[d] * fp[e];
        }
    }
}

/* ------------------------------------------------------------------------- */
/* metric.rs:1065-1179 (dense branch: per-sample row, in-order fold). */
double orc_ms_evaluate(const orc_problem *pb, const orc_transform *t, double *deriv,
                       size_t *valid_out) {
    int np = t->nparams;
    double value_sum = 0.0;
    for (int k = 0; k < np; ++k) deriv[k] = 0.0;
    size_t valid = 0;
    double jac[36];
    size_t n = selection_len(pb);
    int sparse = t->kind == ORC_BSPLINE;
    for (size_t s = 0; s < n; ++s) {
        size_t flat = selection_flat(pb, s);
        if (!selection_allowed(pb, flat)) continue;
        double fp[3], mp[3], mv, grad[3];
        write_point(pb, flat, fp);
        orc_transform_point(t, fp, mp);
        if (!orc_moving_value_and_physical_gradient(pb, mp, &mv, grad)) continue;
        double diff = mv - (double)pb->fvals[flat];
        if (sparse) {
            value_sum += diff * diff;
            long start[3];
            double w[3][4];
            if (bspline_support(t, fp, start, w)) {
                size_t gs0 = t->grid_size[0], gs1 = t->grid_size[1];
                size_t ncp = gs0 * gs1 * t->grid_size[2];
                for (int kk = 0; kk < 4; ++kk)
                    for (int jj = 0; jj < 4; ++jj)
                        for (int ii = 0; ii < 4; ++ii) {
                            double wt = w[0][ii] * w[1][jj] * w[2][kk];
                            size_t cp = (size_t)(start[0] + ii) + gs0 * ((size_t)(start[1] + jj) + gs1 * (size_t)(start[2] + kk));
                            for (int d = 0; d < 3; ++d) {
                                double gg = -0.0;
                                gg += wt * grad[d];
                                deriv[d * ncp + cp] += 2.0 * diff * gg;
                            }
                        }
            }
        } else {
            value_sum += diff * diff;
            orc_transform_jacobian(t, fp, jac);
            for (int k = 0; k < np; ++k) {
                double gsum = 0.0;
                for (int d = 0; d < 3; ++d) gsum += grad[d] * jac[d * np + k];
                deriv[k] += 2.0 * diff * gsum;
            }
        }
        valid += 1;
    }
    *valid_out = valid;
    if (valid == 0) {
        for (int k = 0; k < np; ++k) deriv[k] = 0.0;
        return DBL_MAX;
    }
    double inv = 1.0 / (double)valid;
    for (int k = 0; k < np; ++k) deriv[k] = deriv[k] * inv;
    return value_sum * inv;
}

/* The dense branch of metric.rs:1065-1179 as the reference runs it:
 * map_rows_fold_in_order (core/parallel.rs:720-818) -- every sample's row
 * [diff^2, 2 diff g_k] computed by `threads` threads, chunk by chunk of 2^16
 * rows, then added on one thread in sample order. The sum sequence is the
 * serial loop's, so the result is orc_ms_evaluate's bit for bit; this only
 * times the reference's threading. Sparse transforms run the serial loop, as
 * in the reference. */
typedef struct {
    const orc_problem *pb;
    const orc_transform *t;
    size_t start, r0, r1;
    int width;
    double *rows;
    unsigned char *ok;
} ms_rows_job;

static void *ms_rows_thread(void *arg) {
    ms_rows_job *j = (ms_rows_job *)arg;
    const orc_problem *pb = j->pb;
    const orc_transform *t = j->t;
    int np = t->nparams;
    double jac[36];
    for (size_t r = j->r0; r < j->r1; ++r) {
        double *row = j->rows + r * (size_t)j->width;
        size_t flat = selection_flat(pb, j->start + r);
        j->ok[r] = 0;
        if (!selection_allowed(pb, flat)) continue;
        double fp[3], mp[3], mv, grad[3];
        write_point(pb, flat, fp);
        orc_transform_point(t, fp, mp);
        if (!orc_moving_value_and_physical_gradient(pb, mp, &mv, grad)) continue;
        double diff = mv - (double)pb->fvals[flat];
        row[0] = diff * diff;
        orc_transform_jacobian(t, fp, jac);
        for (int k = 0; k < np; ++k) {
            double g = 0.0;
            for (int d = 0; d < 3; ++d) g += grad[d] * jac[d * np + k];
            row[1 + k] = 2.0 * diff * g;
        }
        j->ok[r] = 1;
    }
    return NULL;
}

double orc_ms_evaluate_threads(const orc_problem *pb, const orc_transform *t, double *deriv,
                               size_t *valid_out, int threads) {
    if (threads <= 1 || t->kind == ORC_BSPLINE) return orc_ms_evaluate(pb, t, deriv, valid_out);
    int nt = threads > 256 ? 256 : threads;
    int np = t->nparams, width = 1 + np;
    size_t n = selection_len(pb);
    size_t chunk = (size_t)1 << 16;
    if (chunk > n) chunk = n;
    double *rows = (double *)malloc((chunk ? chunk : 1) * (size_t)width * sizeof(double));
    unsigned char *ok = (unsigned char *)malloc(chunk ? chunk : 1);
    double value_sum = 0.0;
    size_t valid = 0;
    for (int k = 0; k < np; ++k) deriv[k] = 0.0;
    pthread_t tid[256];
    ms_rows_job jobs[256];
    for (size_t start = 0; start < n; start += chunk) {
        size_t count = n - start < chunk ? n - start : chunk;
        for (int q = 0; q < nt; ++q) {
            jobs[q].pb = pb; jobs[q].t = t; jobs[q].start = start; jobs[q].width = width;
            jobs[q].rows = rows; jobs[q].ok = ok;
            jobs[q].r0 = count * (size_t)q / (size_t)nt;
            jobs[q].r1 = count * (size_t)(q + 1) / (size_t)nt;
            pthread_create(&tid[q], NULL, ms_rows_thread, &jobs[q]);
        }
        for (int q = 0; q < nt; ++q) pthread_join(tid[q], NULL);
        for (size_t r = 0; r < count; ++r) {
            if (!ok[r]) continue;
            const double *row = rows + r * (size_t)width;
            value_sum += row[0];
            for (int k = 0; k < np; ++k) deriv[k] += row[1 + k];
            valid += 1;
        }
    }
    free(rows);
    free(ok);
    *valid_out = valid;
    if (valid == 0) {
        for (int k = 0; k < np; ++k) deriv[k] = 0.0;
        return DBL_MAX;
    }
    double inv = 1.0 / (double)valid;
    for (int k = 0; k < np; ++k) deriv[k] = deriv[k] * inv;
    return value_sum * inv;
}

/* metric.rs:1181-1223 */
double orc_ms_value(const orc_problem *pb, const orc_transform *t, size_t *valid_out) {
    double value_sum = 0.0;
    size_t valid = 0;
    size_t n = selection_len(pb);
    for (size_t s = 0; s < n; ++s) {
        size_t flat = selection_flat(pb, s);
        if (!selection_allowed(pb, flat)) continue;
        double fp[3], mp[3], mv;
        write_point(pb, flat, fp);
        orc_transform_point(t, fp, mp);
        if (!orc_moving_value_at(pb, mp, &mv)) continue;
        double diff = mv - (double)pb->fvals[flat];
        value_sum += diff * diff;
        valid += 1;
    }
    *valid_out = valid;
    if (valid == 0) return DBL_MAX;
    return value_sum / (double)valid;
}

/* ------------------------------------------------------------------------- */
/* B-spline transform (src/transform/bspline.rs:488-556). */
static double bspline_kernel(double u) { return orc_cubic_bspline(u); } /* bspline.rs:94-105 */

static void bspline_continuous_index(const orc_transform *t, const double p[3], double c[3]) {
    for (int r = 0; r < 3; ++r) {
        double acc = -0.0;
        for (int cc = 0; cc < 3; ++cc) acc += t->grid_p2i[r * 3 + cc] * (p[cc] - t->grid_origin[cc]);
        c[r] = acc;
    }
}

/* bspline.rs:504-521 (InsideValidRegion with the far-boundary snap) */
static int bspline_inside_valid_region(const orc_transform *t, double idx[3]) {
    double min_limit = 0.5 * (3.0 - 1.0);
    for (int d = 0; d < 3; ++d) {
        double max_limit = (double)t->grid_size[d] - 0.5 * (3.0 - 1.0) - 1.0;
        double ulp = 4.0 * DBL_EPSILON * fmax(fabs(max_limit), 1.0);
        if (fabs(idx[d] - max_limit) <= ulp) {
            idx[d] = max_limit - 6.0 * DBL_EPSILON * fmax(fabs(max_limit), 1.0);
        } else if (idx[d] >= max_limit || idx[d] < min_limit) {
            return 0;
        }
    }
    return 1;
}

/* bspline.rs:525-556: per-axis support start and 1-D weights; the 64 product
 * weights are formed by the caller in support order (axis 0 fastest). */
static int bspline_support(const orc_transform *t, const double p[3], long start[3],
                           double wts[3][4]) {
    double idx[3];
    bspline_continuous_index(t, p, idx);
    if (!bspline_inside_valid_region(t, idx)) return 0;
    for (int d = 0; d < 3; ++d) {
        start[d] = (long)floor(idx[d] + 0.5 - 3.0 / 2.0);
        double x = idx[d] - (double)start[d];
        for (int k = 0; k < 4; ++k) {
            wts[d][k] = bspline_kernel(x);
            x -= 1.0;
        }
    }
    return 1;
}

/* bspline.rs:229-322 (new + from_grid_geometry) for a domain given by an
 * image's origin / size*spacing / direction (from_image_domain, :324-339). */
int orc_bspline_from_image_domain(orc_transform *t, const size_t img_size[3],
                                  const double img_spacing[3], const double img_origin[3],
                                  const double img_dir[9], const size_t mesh[3],
                                  const double *coeffs) {
    memset(t, 0, sizeof(*t));
    t->kind = ORC_BSPLINE;
    double phys[3], shift[3], rot[3];
    for (int i = 0; i < 3; ++i) {
        if (mesh[i] == 0) return -1;
        phys[i] = (double)img_size[i] * img_spacing[i];
        t->grid_size[i] = mesh[i] + 3;
        t->grid_spacing[i] = phys[i] / (double)mesh[i];
        shift[i] = -0.5 * t->grid_spacing[i] * (3.0 - 1.0);
    }
    mat_vec3(img_dir, shift, rot);
    for (int i = 0; i < 3; ++i) t->grid_origin[i] = img_origin[i] + rot[i];
    for (int i = 0; i < 9; ++i) t->grid_direction[i] = img_dir[i];
    double i2p[9];
    orc_index_to_physical_matrix(img_dir, t->grid_spacing, i2p);
    if (orc_invert3(i2p, t->grid_p2i) != 0) return -2;
    size_t ncp = t->grid_size[0] * t->grid_size[1] * t->grid_size[2];
    t->nparams = (int)(3 * ncp);
    t->coeffs = coeffs;
    return 0;
}

/* mattes.rs:722-796. With a moving-initial transform g (linear), the sparse
 * Jacobian's columns are left-multiplied by g's matrix (method.rs:313-340), so
 * the moving gradient enters as J_g' grad. */
static double mattes_evaluate_sparse(const orc_problem *pb, const orc_mattes_geom *g,
                                     const orc_transform *t, const orc_transform *init,
                                     double *deriv, size_t *valid_out) {
    size_t bins = g->num_bins;
    int np = t->nparams;
    double *joint = (double *)malloc(bins * bins * sizeof(double));
    double *marg = (double *)malloc(bins * sizeof(double));
    double *pratio = (double *)malloc(bins * bins * sizeof(double));
    size_t valid;
    mattes_histogram_composed(pb, g, t, init, joint, marg, &valid);
    *valid_out = valid;
    double value;
    for (int k = 0; k < np; ++k) deriv[k] = 0.0;
    if (!orc_mattes_tail(joint, marg, valid, g, &value, pratio)) goto done;
    {
        size_t n = selection_len(pb);
        size_t ncp = t->grid_size[0] * t->grid_size[1] * t->grid_size[2];
        for (size_t s = 0; s < n; ++s) {
            size_t flat = selection_flat(pb, s);
            if (!selection_allowed(pb, flat)) continue;
            double fp[3], mp[3], mv, grad[3];
            write_point(pb, flat, fp);
            double fv = (double)pb->fvals[flat];
            composed_point(t, init, fp, mp);
            if (!orc_moving_value_and_physical_gradient(pb, mp, &mv, grad)) continue;
            if (mv < g->moving_true_min || mv > g->moving_true_max) continue;
            long start[3];
            double w[3][4];
            if (!bspline_support(t, fp, start, w)) continue;
            if (init) {
                double gq[3];
                for (int d = 0; d < 3; ++d) {
                    gq[d] = 0.0;
                    for (int r = 0; r < 3; ++r) gq[d] += init->matrix[r * 3 + d] * grad[r];
                }
                grad[0] = gq[0]; grad[1] = gq[1]; grad[2] = gq[2];
            }
            double mterm = mv / g->moving_bin_size - g->moving_normalized_min;
            size_t mi = orc_parzen_window_index(g, mv, g->moving_bin_size, g->moving_normalized_min);
            size_t fi = orc_parzen_window_index(g, fv, g->fixed_bin_size, g->fixed_normalized_min);
            double acc = 0.0;
            for (size_t m = mi - 1; m < mi + 3; ++m) {
                double arg = (double)m - mterm;
                acc += orc_cubic_bspline_derivative(arg) * pratio[fi * bins + m];
            }
            size_t gs0 = t->grid_size[0], gs1 = t->grid_size[1];
            for (int kk = 0; kk < 4; ++kk)
                for (int jj = 0; jj < 4; ++jj)
                    for (int ii = 0; ii < 4; ++ii) {
                        double wt = w[0][ii] * w[1][jj] * w[2][kk];
                        size_t cp = (size_t)(start[0] + ii) + gs0 * ((size_t)(start[1] + jj) + gs1 * (size_t)(start[2] + kk));
                        for (int d = 0; d < 3; ++d) {
                            double inner = -0.0;
                            inner += wt * grad[d];
                            deriv[d * ncp + cp] += inner * acc;
                        }
                    }
        }
    }
done:
    free(joint);
    free(marg);
    free(pratio);
    return value;
}

/* ------------------------------------------------------------------------- */
/* tests/mattes_device.rs:42-100 */
void orc_fixture_volume(size_t n, const double shift[3], float *out) {
    double c = (double)n / 2.0;
    double s5 = (double)n / 5.0;
    double denom = 2.0 * (s5 * s5); /* powi(2) is an exact square */
    size_t o = 0;
    for (size_t k = 0; k < n; ++k)
        for (size_t j = 0; j < n; ++j)
            for (size_t i = 0; i < n; ++i) {
                double x = (double)i - c + shift[0];
                double y = (double)j - c + shift[1];
                double z = (double)k - c + shift[2];
                double d2 = x * x + y * y + z * z;
                double s = 120.0 * exp(-d2 / denom) + 10.0 * sin(x / 7.0) * cos(y / 9.0) * sin(z / 11.0);
                out[o++] = (float)s;
            }
}

void orc_fixture_inverted(size_t n, const double shift[3], float *out) {
    orc_fixture_volume(n, shift, out);
    size_t total = n * n * n;
    for (size_t i = 0; i < total; ++i) out[i] = (float)(120.0 - (double)out[i]);
}

void orc_fixture_commensurate(size_t n, float *out) {
    size_t o = 0;
    for (size_t k = 0; k < n; ++k)
        for (size_t j = 0; j < n; ++j)
            for (size_t i = 0; i < n; ++i) out[o++] = (float)((i * 7 + j * 13 + k * 3) % 64);
}

/* ------------------------------------------------------------------------- */
/* ResampleImageFilter::execute (src/transform/resample.rs:226-290) with the
 * linear interpolator, Float32 in and out: out index -> physical (origin
 * first) -> transform_point -> input continuous index (mat_vec seeded 0.0)
 * -> linear_at, default outside. Geometry arrays: size[3], spacing[3],
 * origin[3], direction[9]. */
/* interpolator.rs:37-54: nearest_at -- inside test first, then round half
 * up (coord.rs:189-191, floor(c + 0.5)) clamped into the buffer. */
static int nearest_at(const orc_problem *pb, const double c[3], double *v) {
    if (!is_inside(c, pb->msize)) return 0;
    size_t off = 0, stride = 1;
    for (int d = 0; d < 3; ++d) {
        int64_t i = (int64_t)floor(c[d] + 0.5);
        if (i < 0) i = 0;
        if (i > (int64_t)pb->msize[d] - 1) i = (int64_t)pb->msize[d] - 1;
        off += (size_t)i * stride;
        stride *= pb->msize[d];
    }
    *v = (double)pb->mvals[off];
    return 1;
}

static void resample_with(const float *in, const size_t in_size[3], const double in_spacing[3],
                          const double in_origin[3], const double in_dir[9],
                          const size_t out_size[3], const double out_spacing[3],
                          const double out_origin[3], const double out_dir[9],
                          const orc_transform *t, double dflt, int nearest, float *out);

/* ResampleImageFilter::execute (transform/resample.rs:226-290) with the
 * nearest-neighbour interpolator. */
void orc_resample_nearest(const float *in, const size_t in_size[3], const double in_spacing[3],
                          const double in_origin[3], const double in_dir[9],
                          const size_t out_size[3], const double out_spacing[3],
                          const double out_origin[3], const double out_dir[9],
                          const orc_transform *t, double dflt, float *out) {
    resample_with(in, in_size, in_spacing, in_origin, in_dir, out_size, out_spacing, out_origin,
                  out_dir, t, dflt, 1, out);
}

void orc_resample_linear(const float *in, const size_t in_size[3], const double in_spacing[3],
                         const double in_origin[3], const double in_dir[9],
                         const size_t out_size[3], const double out_spacing[3],
                         const double out_origin[3], const double out_dir[9],
                         const orc_transform *t, double dflt, float *out) {
    resample_with(in, in_size, in_spacing, in_origin, in_dir, out_size, out_spacing, out_origin,
                  out_dir, t, dflt, 0, out);
}

static void resample_with(const float *in, const size_t in_size[3], const double in_spacing[3],
                          const double in_origin[3], const double in_dir[9],
                          const size_t out_size[3], const double out_spacing[3],
                          const double out_origin[3], const double out_dir[9],
                          const orc_transform *t, double dflt, int nearest, float *out) {
    double o2p[9], i2p[9], p2i[9];
    orc_index_to_physical_matrix(out_dir, out_spacing, o2p);
    orc_index_to_physical_matrix(in_dir, in_spacing, i2p);
    orc_invert3(i2p, p2i);
    orc_problem pb;
    memset(&pb, 0, sizeof(pb));
    for (int d = 0; d < 3; ++d) { pb.msize[d] = in_size[d]; pb.morigin[d] = in_origin[d]; }
    pb.mvals = in;
    size_t n = out_size[0] * out_size[1] * out_size[2];
    size_t idx[3] = {0, 0, 0};
    for (size_t o = 0; o < n; ++o) {
        double phys[3], mapped[3], diff[3], c[3];
        for (int r = 0; r < 3; ++r) {
            double acc = out_origin[r];
            for (int k = 0; k < 3; ++k) acc += o2p[r * 3 + k] * (double)idx[k];
            phys[r] = acc;
        }
        orc_transform_point(t, phys, mapped);
        for (int d = 0; d < 3; ++d) diff[d] = mapped[d] - in_origin[d];
        mat_vec3(p2i, diff, c);
        double v;
        const int ok = nearest ? nearest_at(&pb, c, &v) : linear_at(&pb, c, &v);
        out[o] = (float)(ok ? v : dflt);
        for (int d = 0; d < 3; ++d) {
            if (++idx[d] < out_size[d]) break;
            idx[d] = 0;
        }
    }
}

/* shrink geometry (src/filters/shrink.rs:51-91) */
void orc_shrunk_geometry(const size_t size[3], const double spacing[3], const double origin[3],
                         const double dir[9], const size_t f[3], size_t out_size[3],
                         double out_spacing[3], double out_origin[3]) {
    double delta[3];
    for (int d = 0; d < 3; ++d) {
        size_t s = size[d] / f[d];
        out_size[d] = s < 1 ? 1 : s;
        out_spacing[d] = spacing[d] * (double)f[d];
        delta[d] = ((double)size[d] - 1.0) / 2.0 - (double)f[d] * ((double)out_size[d] - 1.0) / 2.0;
    }
    for (int i = 0; i < 3; ++i) {
        double acc = 0.0;
        for (int j = 0; j < 3; ++j) acc += dir[i * 3 + j] * spacing[j] * delta[j];
        out_origin[i] = origin[i] + acc;
    }
}
