// The registration driver: a multi-level ImageRegistrationMethod run with every
// volume resident on the device (the reference's execute_on_device,
// crates/sitk/src/registration/method.rs:2698-2879, with its level preparation
// :2416-2569 / :2919-2963 and optimizer dispatch `drive` :3140-3529).
//
// The optimizers and the scales / learning-rate estimator are host logic, as in
// the reference: each iteration hands the device one point-map stage, the device
// evaluates the metric (value + 12 or 14 moments) and the host contracts the
// moments with the transform's Jacobian decomposition.
//   RegularStepGradientDescent      src/registration/optimizer.rs:271-450
//   GradientDescentLineSearch       src/registration/optimizer.rs:462-817
//   WindowConvergenceMonitor        src/registration/convergence.rs
//   physical-shift scales / rates   src/registration/scales.rs:225-871
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cfloat>
#include <chrono>
#include <exception>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <vector>

#include "batch.hpp"
#include "host_pool.hpp"
#include "common.hpp"
#include "image.hpp"
#include "metric.hpp"
#include "transform.hpp"

namespace sitkhip {

// ITK's CompensatedSummation: the running sum without the final residual
// (src/core/compensated.rs:115-128).
static double compensated(const std::vector<double> &v) {
    double sum = 0.0, comp = 0.0;
    for (double x : v) {
        const double ci = x - comp;
        const double t = sum + ci;
        comp = (t - sum) - ci;
        sum = t;
    }
    return sum;
}

// ------------------------------------------------------------ scales estimator
// RegistrationParameterScalesFromPhysicalShift for a global linear transform:
// both sample sets are the 8 corners of the virtual grid (scales.rs:489-515),
// with the Jacobians cached once (they do not depend on the parameters).
struct ScalesEstimator {
    int np = 0;
    double max_step = 1.0;           // minimum virtual spacing
    double delta = 0.01;             // small parameter variation
    std::vector<double> jac;         // 8 x (3 x np)

    void build(const sitkhip_geometry &grid, const Transform &t) {
        np = t.nparams;
        max_step = std::min(grid.spacing[0], std::min(grid.spacing[1], grid.spacing[2]));
        jac.assign(static_cast<size_t>(8) * 3 * np, 0.0);
        double i2p[9];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) i2p[r * 3 + c] = grid.direction[r * 3 + c] * grid.spacing[c];
        for (int i = 0; i < 8; ++i) {
            double idx[3], x[3];
            for (int d = 0; d < 3; ++d)
                idx[d] = ((i >> d) & 1) ? static_cast<double>(grid.size[d] - 1) : 0.0;
            for (int r = 0; r < 3; ++r) {
                double v = grid.origin[r];
                for (int c = 0; c < 3; ++c) v += i2p[r * 3 + c] * idx[c];
                x[r] = v;
            }
            t.jacobian(x, jac.data() + static_cast<size_t>(i) * 3 * np);
        }
    }

    // shift_scales (scales.rs:659-718)
    std::vector<double> scales() const {
        std::vector<double> shifts(np, 0.0);
        for (int s = 0; s < 8; ++s) {
            const double *j = jac.data() + static_cast<size_t>(s) * 3 * np;
            for (int p = 0; p < np; ++p) {
                double sq = 0.0;
                for (int r = 0; r < 3; ++r) sq += j[r * np + p] * j[r * np + p];
                shifts[p] = std::max(shifts[p], delta * std::sqrt(sq));
            }
        }
        double min_nz = INFINITY;
        for (double s : shifts)
            if (s > DBL_EPSILON) min_nz = std::min(min_nz, s);
        if (!std::isfinite(min_nz)) return std::vector<double>(np, 1.0);
        const double inv_d2 = 1.0 / (delta * delta);
        std::vector<double> out(np);
        for (int p = 0; p < np; ++p) {
            const double base = shifts[p] <= DBL_EPSILON ? min_nz * min_nz : shifts[p] * shifts[p];
            out[p] = base * inv_d2;
        }
        return out;
    }

    double max_shift(const std::vector<double> &d) const {
        double m = 0.0;
        for (int s = 0; s < 8; ++s) {
            const double *j = jac.data() + static_cast<size_t>(s) * 3 * np;
            double sq = 0.0;
            for (int r = 0; r < 3; ++r) {
                double v = 0.0;
                for (int p = 0; p < np; ++p) v += j[r * np + p] * d[p];
                sq += v * v;
            }
            m = std::max(m, sq);
        }
        return std::sqrt(m);
    }

    // shift_step_scale (scales.rs:777-792) and estimate_learning_rate (:818-826)
    double learning_rate(const std::vector<double> &step) const {
        double mx = 0.0;
        for (double v : step) mx = std::max(mx, std::fabs(v));
        double scale = 0.0;
        if (mx > DBL_EPSILON) {
            const double factor = delta / mx;
            std::vector<double> small(step.size());
            for (size_t k = 0; k < step.size(); ++k) small[k] = step[k] * factor;
            scale = max_shift(small) / factor;
        }
        return scale <= DBL_EPSILON ? 1.0 : max_step / scale;
    }
};

// ------------------------------------------------------------ the objective
// The parameter gradient from the metric's moments and the transform's
// Jacobian decomposition (registration/cuda.rs:381-412, 505-524). Mattes:
// g_k = sum_d J0[d][k] A_d + sum_de C_e[d][k] B_de. Mean squares: the same
// contraction of S0, S1 times 2/count; returns the value sq/count (DBL_MAX and a
// zero gradient with no valid sample).
static double contract_objective(int kind, double mattes_value, const double *mom,
                                 const Transform &t, std::vector<double> &g, size_t &valid) {
    double j0[36], c[108];
    t.affine_form(j0, c);
    const int np = t.nparams;
    g.assign(np, 0.0);
    if (kind == SITKHIP_METRIC_MATTES) {
        for (int k = 0; k < np; ++k) {
            double acc = 0.0;
            for (int d = 0; d < 3; ++d) {
                acc += j0[d * np + k] * mom[d];
                for (int e = 0; e < 3; ++e) acc += c[(e * 3 + d) * np + k] * mom[3 + d * 3 + e];
            }
            g[k] = acc;
        }
        return mattes_value;
    }
    valid = static_cast<size_t>(mom[13]);
    if (mom[13] == 0.0) return DBL_MAX;
    const double inv = 1.0 / mom[13];
    for (int k = 0; k < np; ++k) {
        double acc = 0.0;
        for (int d = 0; d < 3; ++d) {
            acc += j0[d * np + k] * mom[1 + d];
            for (int e = 0; e < 3; ++e) acc += c[(e * 3 + d) * np + k] * mom[4 + d * 3 + e];
        }
        g[k] = 2.0 * acc * inv;
    }
    return mom[0] * inv;
}

struct LevelObjective {
    Metric *metric = nullptr;
    int kind = SITKHIP_METRIC_MATTES;
    Transform t;
    size_t evaluations = 0, value_evaluations = 0;
    size_t last_valid = 0;

    // Value and parameter derivative at p.
    double value_and_gradient(const std::vector<double> &p, std::vector<double> &g) {
        t.set_parameters(p.data());
        double st[12];
        t.stage(st);
        ++evaluations;
        if (kind == SITKHIP_METRIC_MATTES) {
            double value, mom[12];
            size_t valid;
            mattes_value_and_moments(*metric, st, 1, &value, mom, &valid);
            last_valid = valid;
            return contract_objective(kind, value, mom, t, g, last_valid);
        }
        double mom[14];
        ms_moments(*metric, st, 1, mom);
        return contract_objective(kind, 0.0, mom, t, g, last_valid);
    }

    double value(const std::vector<double> &p) {
        t.set_parameters(p.data());
        double st[12];
        t.stage(st);
        ++value_evaluations;
        if (kind == SITKHIP_METRIC_MATTES) {
            double v;
            size_t valid;
            mattes_value(*metric, st, 1, &v, &valid);
            return v;
        }
        double sq;
        size_t valid;
        ms_value(*metric, st, 1, &sq, &valid);
        return valid == 0 ? DBL_MAX : sq / static_cast<double>(valid);
    }
};

// --------------------------------------------------------------- optimizers
struct OptResult {
    std::vector<double> params;
    double value;
    int iterations;
    int stop;
};

// RegularStepGradientDescentOptimizer::optimize_with_lr (optimizer.rs:353-450).
static OptResult regular_step(const sitkhip_registration_config &cfg, const std::vector<double> &scales,
                              std::vector<double> p, LevelObjective &obj,
                              const std::function<double(const std::vector<double> &)> &lr_of) {
    const size_t n = p.size();
    std::vector<double> grad;
    double value = obj.value_and_gradient(p, grad);
    std::vector<double> prev(n, 0.0), scaled(n);
    double relaxation = 1.0;
    int stop = SITKHIP_STOP_MAX_ITERATIONS;
    int taken = 0;
    for (;;) {
        for (size_t k = 0; k < n; ++k) scaled[k] = grad[k] / scales[k];
        std::vector<double> sq(n), dot(n);
        for (size_t k = 0; k < n; ++k) { sq[k] = scaled[k] * scaled[k]; dot[k] = scaled[k] * prev[k]; }
        const double gm = std::sqrt(compensated(sq));
        if (gm < cfg.gradient_tolerance) { stop = SITKHIP_STOP_GRADIENT_CONVERGED; break; }
        if (compensated(dot) < 0.0) relaxation *= cfg.relaxation;
        const double step = relaxation * lr_of(scaled);
        if (step < cfg.minimum_step) { stop = SITKHIP_STOP_STEP_TOO_SMALL; break; }
        if (taken >= cfg.iterations) break;
        const double factor = step / gm;
        for (size_t k = 0; k < n; ++k) p[k] -= factor * scaled[k];
        prev = scaled;
        ++taken;
        value = obj.value_and_gradient(p, grad);
    }
    return {p, value, taken, stop};
}

// WindowConvergenceMonitor (convergence.rs): the negated slope of a line fit to
// the last `window` values normalised by the running total |energy|.
struct ConvergenceMonitor {
    size_t window;
    std::deque<double> values;
    double total = 0.0;
    explicit ConvergenceMonitor(size_t w) : window(w) {}
    void add(double v) {
        values.push_back(v);
        if (values.size() > window) values.pop_front();
        total += std::fabs(v);
    }
    bool value(double &cv) const {
        if (values.size() < window) return false;
        if (total == 0.0) { cv = 0.0; return true; }
        const double w = static_cast<double>(window);
        double o0 = 0, o1 = 0, d0 = 0, d1 = 0;
        size_t n = 0;
        for (double e : values) {
            const double x = static_cast<double>(n) / (w - 1.0);
            const double a = 1.0 - x, b = x, y = e / total, dd = a * a + b * b;
            o0 += a * a;
            o1 += b * b;
            d0 += a * a * a * y / dd;
            d1 += b * b * b * y / dd;
            ++n;
        }
        const double c0 = o0 != 0.0 ? d0 / o0 : 0.0;
        const double c1 = o1 != 0.0 ? d1 / o1 : 0.0;
        cv = c0 - c1;
        return true;
    }
};

// golden_section_search (optimizer.rs:475-545), iterative form of the same
// recursion.
static double golden_section(double eps, unsigned max_it, double a, double b, double c,
                             const std::function<double(double)> &f) {
    const double phi = 1.618034, resphi = 2.0 - phi;
    bool have_b = false;
    double fb = 0.0;
    unsigned it = 0;
    for (;;) {
        if (it > max_it) return (c + a) / 2.0;
        ++it;
        const double x = (c - b > b - a) ? b + resphi * (c - b) : b - resphi * (b - a);
        if (std::fabs(c - a) < eps * (std::fabs(b) + std::fabs(x))) return (c + a) / 2.0;
        const double fx = f(x);
        if (!have_b) { fb = f(b); have_b = true; }
        if (fx < fb) {
            if (c - b > b - a) { a = b; b = x; } else { c = b; b = x; }
            fb = fx;
        } else {
            if (c - b > b - a) { c = x; } else { a = x; }
        }
    }
}

// GradientDescentLineSearchOptimizer::run (optimizer.rs:724-817).
static OptResult line_search(const sitkhip_registration_config &cfg, const std::vector<double> &scales,
                             std::vector<double> p, LevelObjective &obj, double base_lr) {
    const size_t n = p.size();
    ConvergenceMonitor mon(10);
    std::vector<double> grad;
    double value = obj.value_and_gradient(p, grad);
    double best = value;
    std::vector<double> best_p = p;
    int stop = SITKHIP_STOP_MAX_ITERATIONS;
    int taken = 0;
    for (;;) {
        if (value < best) { best = value; best_p = p; }
        mon.add(value);
        double cv;
        if (mon.value(cv) && cv <= 1e-6) { stop = SITKHIP_STOP_CONVERGED; break; }
        if (taken >= cfg.iterations) break;
        std::vector<double> d(n);
        for (size_t k = 0; k < n; ++k) d[k] = grad[k] / scales[k];
        auto line = [&](double x) {
            std::vector<double> trial(n);
            for (size_t k = 0; k < n; ++k) trial[k] = p[k] - x * d[k];
            return obj.value(trial);
        };
        const double lr = golden_section(0.01, 20, base_lr * 0.0, base_lr, base_lr * 5.0, line);
        base_lr = lr;
        for (size_t k = 0; k < n; ++k) p[k] -= lr * d[k];
        ++taken;
        value = obj.value_and_gradient(p, grad);
    }
    return {best_p, best, taken, stop};
}

// ------------------------------------------------------------------- driver
// The masks' sizes against their images (method.rs:2070-2083).
static void check_masks(const sitkhip_registration_config &cfg, const Image &fixed,
                        const Image &moving) {
    auto same_size = [](const sitkhip_geometry &a, const sitkhip_geometry &b) {
        return a.size[0] == b.size[0] && a.size[1] == b.size[1] && a.size[2] == b.size[2];
    };
    if (cfg.fixed_mask && !same_size(reinterpret_cast<const Mask *>(cfg.fixed_mask)->geom, fixed.geom))
        fail(SITKHIP_ERR_MASK_SIZE_MISMATCH, "the fixed mask's size differs from the fixed image's");
    if (cfg.moving_mask &&
        !same_size(reinterpret_cast<const Mask *>(cfg.moving_mask)->geom, moving.geom))
        fail(SITKHIP_ERR_MASK_SIZE_MISMATCH, "the moving mask's size differs from the moving image's");
}

static void check_config(const sitkhip_registration_config &cfg, int kind) {
    if (cfg.levels < 1 || cfg.levels > SITKHIP_MAX_LEVELS)
        fail(SITKHIP_ERR_INVALID_ARGUMENT, "levels must be 1..8");
    if (kind != SITKHIP_TRANSFORM_TRANSLATION && kind != SITKHIP_TRANSFORM_EULER3D &&
        kind != SITKHIP_TRANSFORM_AFFINE)
        fail(SITKHIP_ERR_INVALID_ARGUMENT, "unsupported transform kind");
    if (cfg.optimizer != SITKHIP_OPTIMIZER_REGULAR_STEP && cfg.optimizer != SITKHIP_OPTIMIZER_LINE_SEARCH)
        fail(SITKHIP_ERR_INVALID_ARGUMENT, "unsupported optimizer");
    if (cfg.sampling < 0 || cfg.sampling > 2) fail(SITKHIP_ERR_INVALID_ARGUMENT, "unknown sampling strategy");
    if (cfg.interpolator != SITKHIP_INTERP_LINEAR && cfg.interpolator != SITKHIP_INTERP_BSPLINE)
        fail(SITKHIP_ERR_INVALID_ARGUMENT, "unknown interpolator");
    check_sampling_fraction(cfg.sampling, cfg.sampling_percentage, 1);
}

static Transform initial_transform(const sitkhip_transform &tr) {
    Transform T;
    T.kind = tr.kind;
    T.compute_zyx = tr.compute_zyx != 0;
    for (int d = 0; d < 3; ++d) T.center[d] = tr.center[d];
    T.recompute();
    T.set_parameters(tr.params);
    return T;
}

// One level's metric, built the way prepare_level does it (method.rs:
// 2416-2569): both images smoothed at full resolution, the fixed placed onto
// the shrunk grid, the moving kept at full size; then the level's sample draw
// (metric.rs:419-437) and, for Mattes, the histogram geometry from the value
// ranges.
struct LevelInputs {
    std::unique_ptr<Metric, void (*)(Metric *)> metric{nullptr, metric_destroy};
    sitkhip_geometry grid{};
};

struct ImageHolder {
    Image *img = nullptr;
    bool owned = false;
    ~ImageHolder() { if (owned) delete img; }
};

// Everything a level needs, enqueued without waiting on the stream: the
// metric's value ranges are left pending (prepare_level_end reads them).
// smoothed_fixed / smoothed_moving: the level's smoothed images when the
// caller made them (taken over and freed here), else this smooths.
// prep: the placement and the metric's set-up are recorded there (not for a
// level with a fixed mask, which is set up at once).
static LevelInputs prepare_level_begin(const Image &fixed, const Image &moving,
                                       const sitkhip_registration_config &cfg, int level,
                                       Image *smoothed_fixed = nullptr,
                                       Image *smoothed_moving = nullptr,
                                       PrepBatch *prep = nullptr) {
    if (cfg.fixed_mask) prep = nullptr;
    const double s = cfg.sigma[level];
    const double sig[3] = {s, s, s};
    const size_t *f = cfg.shrink[level];
    const bool smooth = s > 0.0;
    const bool native = f[0] == 1 && f[1] == 1 && f[2] == 1;
    ImageHolder sf, sm, fl;
    if (smoothed_fixed) sf.img = smoothed_fixed;
    else sf.img = smooth ? recursive_gaussian(fixed, sig, false) : const_cast<Image *>(&fixed);
    sf.owned = smooth || smoothed_fixed;
    if (smoothed_moving) sm.img = smoothed_moving;
    else sm.img = smooth ? recursive_gaussian(moving, sig, false) : const_cast<Image *>(&moving);
    sm.owned = smooth || smoothed_moving;
    if (native) {
        fl.img = sf.img;
    } else {
        sitkhip_geometry grid;
        long long off[3];
        shrunk_geometry(fixed.geom, f, grid, off);
        fl.img = prep ? resample_linear_into(*sf.img, grid, *prep)
                      : resample_linear(*sf.img, grid, nullptr, 0, 0.0, false);
        fl.owned = true;
    }
    const Image &fixed_level = *fl.img;
    const size_t nvox = fixed_level.len();
    LevelInputs out;
    out.grid = fixed_level.geom;
    // the level's own images go to the metric instead of being copied; the
    // value ranges only feed the Mattes histogram
    CreateOptions opt;
    opt.defer_ranges = true;
    opt.ranges = cfg.metric == SITKHIP_METRIC_MATTES;
    opt.adopt_moving = sm.owned ? sm.img : nullptr;
    opt.adopt_fixed = fl.owned || (fl.img == sf.img && sf.owned) ? fl.img : nullptr;
    check_sampling_fraction(cfg.sampling, cfg.sampling_percentage, nvox);
    // Masks (method.rs:2544-2559, 2760-2785, 3068-3110): the fixed mask placed
    // on the level grid by the nearest-neighbour resample (default 0), the
    // moving mask as given (the moving volume keeps its grid).
    std::unique_ptr<Mask> level_mask;
    if (cfg.fixed_mask) {
        level_mask.reset(resample_nearest_mask(*reinterpret_cast<const Mask *>(cfg.fixed_mask),
                                               out.grid, nullptr, 0, false));
        opt.fixed_mask = level_mask.get();
    }
    opt.moving_mask = reinterpret_cast<const Mask *>(cfg.moving_mask);
    opt.prep = prep;
    if (cfg.sampling == 2 && level_mask) {
        // the same draw on the host, filtered by the level mask at creation
        // (FixedSamples::from_image_with, metric.rs:494-580)
        const std::vector<int64_t> draws = host_draws(nvox, 2, cfg.sampling_percentage, cfg.seed);
        out.metric.reset(metric_create_with(cfg.metric, fixed_level, draws.data(), draws.size(),
                                            *sm.img, cfg.precision, opt));
    } else if (cfg.sampling == 2) {
        // random draws (metric.rs:419-437), made and sorted on the device
        const size_t cnt = static_cast<size_t>(static_cast<double>(nvox) * cfg.sampling_percentage);
        out.metric.reset(metric_create_drawn(cfg.metric, fixed_level, cnt, cfg.seed, *sm.img,
                                             cfg.precision, opt));
    } else {
        std::vector<int64_t> draws;
        const int64_t *idx = nullptr;
        size_t nidx = 0;
        if (cfg.sampling == 1) {
            size_t stride = static_cast<size_t>(std::ceil(1.0 / cfg.sampling_percentage));
            if (stride < 1) stride = 1;
            for (size_t v = 0; v < nvox; v += stride) draws.push_back(static_cast<int64_t>(v));
            idx = draws.data();
            nidx = draws.size();
        }
        out.metric.reset(metric_create_with(cfg.metric, fixed_level, idx, nidx, *sm.img, cfg.precision, opt));
    }
    return out;
}

// After the stream has drained: the value ranges, and for Mattes the
// histogram geometry from them (its fixed bins recorded in prep when given).
static void prepare_level_end(LevelInputs &in, const sitkhip_registration_config &cfg,
                              PrepBatch *prep = nullptr) {
    metric_finish_ranges(*in.metric);
    if (cfg.interpolator != SITKHIP_INTERP_LINEAR) metric_set_interpolator(*in.metric, cfg.interpolator);
    if (cfg.metric == SITKHIP_METRIC_MATTES) {
        double fr[2], mr[2];
        metric_value_ranges(*in.metric, fr, mr);
        sitkhip_mattes_bins b;
        mattes_bins_from_ranges(fr, mr, static_cast<size_t>(cfg.bins), &b);
        mattes_set_bins(*in.metric, b, prep);
    }
}

static LevelInputs prepare_level(const Image &fixed, const Image &moving,
                                 const sitkhip_registration_config &cfg, int level) {
    LevelInputs in = prepare_level_begin(fixed, moving, cfg, level);
    sync();
    prepare_level_end(in, cfg);
    return in;
}

void register_images(const Image &fixed, const Image &moving, const sitkhip_registration_config &cfg,
                     sitkhip_transform &tr, sitkhip_registration_result &res) {
    const auto t0 = std::chrono::steady_clock::now();
    check_config(cfg, tr.kind);
    check_masks(cfg, fixed, moving);
    std::memset(&res, 0, sizeof(res));
    Transform T = initial_transform(tr);
    std::vector<double> p(tr.params, tr.params + T.nparams);

    for (int level = 0; level < cfg.levels; ++level) {
        LevelInputs in = prepare_level(fixed, moving, cfg, level);
        LevelObjective obj;
        obj.metric = in.metric.get();
        obj.kind = cfg.metric;
        obj.t = T;

        ScalesEstimator est;
        est.build(in.grid, T);
        const int np = T.nparams;
        std::vector<double> scales = cfg.scales_mode == SITKHIP_SCALES_PHYSICAL_SHIFT
                                         ? est.scales()
                                         : std::vector<double>(np, 1.0);
        auto scaled_of = [&](const std::vector<double> &g) {
            std::vector<double> out(np);
            for (int k = 0; k < np; ++k) out[k] = g[k] / scales[k];
            return out;
        };
        OptResult r;
        if (cfg.optimizer == SITKHIP_OPTIMIZER_REGULAR_STEP) {
            if (cfg.learning_rate_mode == SITKHIP_LR_ONCE) {
                std::vector<double> g0;
                obj.value_and_gradient(p, g0);
                const std::vector<double> s0 = scaled_of(g0);
                double gm0 = 0.0;
                for (double v : s0) gm0 += v * v;
                gm0 = std::sqrt(gm0);
                const double lr = est.learning_rate(s0) * gm0;
                r = regular_step(cfg, scales, p, obj, [lr](const std::vector<double> &) { return lr; });
            } else if (cfg.learning_rate_mode == SITKHIP_LR_EACH_ITERATION) {
                r = regular_step(cfg, scales, p, obj, [&](const std::vector<double> &sg) {
                    double gm = 0.0;
                    for (double v : sg) gm += v * v;
                    return est.learning_rate(sg) * std::sqrt(gm);
                });
            } else {
                const double lr = cfg.learning_rate;
                r = regular_step(cfg, scales, p, obj, [lr](const std::vector<double> &) { return lr; });
            }
        } else {
            double lr = cfg.learning_rate;
            if (cfg.learning_rate_mode != SITKHIP_LR_MANUAL) {
                std::vector<double> g0;
                obj.value_and_gradient(p, g0);
                lr = est.learning_rate(scaled_of(g0));
            }
            r = line_search(cfg, scales, p, obj, lr);
        }
        p = r.params;
        // the level's result is evaluated at the returned parameters (method.rs:3504-3512)
        std::vector<double> gf;
        const double vf = obj.value_and_gradient(p, gf);
        if (obj.last_valid == 0) fail(SITKHIP_ERR_DEGENERATE_INPUT, "no valid samples at the result");
        res.level_value[level] = vf;
        res.level_iterations[level] = r.iterations;
        res.level_stop[level] = r.stop;
        res.iterations += r.iterations;
        res.evaluations += obj.evaluations;
        res.value_evaluations += obj.value_evaluations;
        res.value = vf;
        res.valid_points = obj.last_valid;
        T.set_parameters(p.data());
    }
    for (int k = 0; k < T.nparams; ++k) tr.params[k] = p[k];
    res.seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// ------------------------------------------------------------ batched driver
// RegularStepGradientDescent (optimizer.rs:353-450) as a state machine that
// consumes one evaluation at a time, so that the evaluations of many
// registrations can be made together. The sequence of parameters is the one
// regular_step() above produces (with LR Once the first evaluation also serves
// the rate estimate: it is made at the same parameters).
struct RsgdRun {
    const sitkhip_registration_config *cfg = nullptr;
    Transform T;
    ScalesEstimator est;
    std::vector<double> scales, p, prev, grad;
    double lr = 0.0, relaxation = 1.0, value = DBL_MAX;
    size_t valid = 0, evaluations = 0;
    int taken = 0, stop = SITKHIP_STOP_MAX_ITERATIONS;
    bool started = false;

    // Takes the evaluation (value, grad) at p; returns true when p has moved
    // and must be evaluated again, false when the level is finished.
    bool advance() {
        const size_t n = p.size();
        std::vector<double> scaled(n), sq(n), dot(n);
        for (size_t k = 0; k < n; ++k) scaled[k] = grad[k] / scales[k];
        if (!started) {
            started = true;
            if (cfg->learning_rate_mode == SITKHIP_LR_ONCE) {
                double gm0 = 0.0;
                for (double v : scaled) gm0 += v * v;
                lr = est.learning_rate(scaled) * std::sqrt(gm0);
            } else if (cfg->learning_rate_mode == SITKHIP_LR_MANUAL) {
                lr = cfg->learning_rate;
            }
        }
        for (size_t k = 0; k < n; ++k) { sq[k] = scaled[k] * scaled[k]; dot[k] = scaled[k] * prev[k]; }
        const double gm = std::sqrt(compensated(sq));
        if (gm < cfg->gradient_tolerance) { stop = SITKHIP_STOP_GRADIENT_CONVERGED; return false; }
        if (compensated(dot) < 0.0) relaxation *= cfg->relaxation;
        double rate = lr;
        if (cfg->learning_rate_mode == SITKHIP_LR_EACH_ITERATION) {
            double g2 = 0.0;
            for (double v : scaled) g2 += v * v;
            rate = est.learning_rate(scaled) * std::sqrt(g2);
        }
        const double step = relaxation * rate;
        if (step < cfg->minimum_step) { stop = SITKHIP_STOP_STEP_TOO_SMALL; return false; }
        if (taken >= cfg->iterations) { stop = SITKHIP_STOP_MAX_ITERATIONS; return false; }
        const double factor = step / gm;
        for (size_t k = 0; k < n; ++k) p[k] -= factor * scaled[k];
        prev = scaled;
        ++taken;
        return true;
    }
};

void register_batch(size_t n, const Image *const *fixed, const Image *const *moving,
                    const sitkhip_registration_config *cfg, sitkhip_transform *tr,
                    sitkhip_registration_result *res, bool autotune, const double *probe,
                    double *probe_values) {
    const auto t0 = std::chrono::steady_clock::now();
    // SITKHIP_TRACE=1: per-level phase times on stderr
    const bool trace = std::getenv("SITKHIP_TRACE") != nullptr;
    auto now = [] { return std::chrono::steady_clock::now(); };
    auto ms = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
        return std::chrono::duration<double, std::milli>(b - a).count();
    };
    if (n == 0) fail(SITKHIP_ERR_INVALID_ARGUMENT, "an empty batch");
    int max_levels = 0;
    for (size_t i = 0; i < n; ++i) {
        if (!fixed[i] || !moving[i]) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null image in batch");
        check_config(cfg[i], tr[i].kind);
        check_masks(cfg[i], *fixed[i], *moving[i]);
        if (cfg[i].optimizer != SITKHIP_OPTIMIZER_REGULAR_STEP)
            fail(SITKHIP_ERR_INVALID_ARGUMENT,
                 "batched registration runs RegularStepGradientDescent (pair " + std::to_string(i) + ")");
        if (cfg[i].interpolator != SITKHIP_INTERP_LINEAR)
            fail(SITKHIP_ERR_INVALID_ARGUMENT,
                 "batched registration interpolates linearly (pair " + std::to_string(i) + ")");
        if (cfg[i].precision != cfg[0].precision)
            fail(SITKHIP_ERR_INVALID_ARGUMENT, "every pair of a batch has the same precision");
        max_levels = std::max(max_levels, cfg[i].levels);
        std::memset(&res[i], 0, sizeof(res[i]));
    }
    std::vector<Transform> Ts(n);
    std::vector<std::vector<double>> P(n);
    for (size_t i = 0; i < n; ++i) {
        Ts[i] = initial_transform(tr[i]);
        P[i].assign(tr[i].params, tr[i].params + Ts[i].nparams);
    }
    for (int level = 0; level < max_levels; ++level) {
        std::vector<size_t> who;
        for (size_t i = 0; i < n; ++i)
            if (level < cfg[i].levels) who.push_back(i);
        const size_t m = who.size();
        const auto tp0 = now();
        std::vector<LevelInputs> inputs(m);
        std::vector<Metric *> metrics(m);
        // every pair's level enqueued (the smoothing of all fixed and moving
        // images in a few multi-volume launches), one stream wait, then the
        // histogram geometries from the ranges
        std::vector<const Image *> vols;
        std::vector<std::array<double, 3>> sig;
        std::vector<size_t> vol_of(m, SIZE_MAX);
        for (size_t k = 0; k < m; ++k) {
            const size_t i = who[k];
            const double s = cfg[i].sigma[level];
            if (!(s > 0.0)) continue;
            vol_of[k] = vols.size();
            vols.push_back(fixed[i]);
            vols.push_back(moving[i]);
            sig.push_back({s, s, s});
            sig.push_back({s, s, s});
        }
        std::vector<Image *> smoothed = recursive_gaussian_batch(vols, sig);
        // Each pair's set-up (placement, sample draws and their counting sort,
        // fixed values, copies, value ranges, then fixed bins) is recorded and
        // enqueued as one launch per kind of work for every pair.
        // Blocks freed while the work is only recorded (a pair's smoothed fixed
        // image once placed, say) are held back until it is enqueued, so no
        // allocation made meanwhile reuses them before the work reads them.
        PrepBatch prep;
        auto tpe = now();
        device_defer_frees(true);
        try {
            for (size_t k = 0; k < m; ++k) {
                const size_t i = who[k];
                Image *sf = nullptr, *sm = nullptr;
                if (vol_of[k] != SIZE_MAX) {
                    sf = smoothed[vol_of[k]];
                    sm = smoothed[vol_of[k] + 1];
                    smoothed[vol_of[k]] = smoothed[vol_of[k] + 1] = nullptr;
                }
                inputs[k] = prepare_level_begin(*fixed[i], *moving[i], cfg[i], level, sf, sm, &prep);
                metrics[k] = inputs[k].metric.get();
            }
            prep.flush();
            tpe = now();
            sync();
            for (size_t k = 0; k < m; ++k) prepare_level_end(inputs[k], cfg[who[k]], &prep);
            prep.flush();
        } catch (...) {
            (void)hipStreamSynchronize(backend().stream);
            device_defer_frees(false);
            for (Image *img : smoothed) delete img;
            throw;
        }
        device_defer_frees(false);
        const auto tp1 = now();
        // Two engines over alternate pairs once the batch is large: the host
        // updates one half's optimizers while the device evaluates the other
        // (each pair's evaluation is the same in either engine).
        const size_t H = m >= 64 ? 2 : 1;
        struct Half {
            std::vector<size_t> ks;  // level-local pair indices
            std::unique_ptr<BatchEngine, void (*)(BatchEngine *)> eng{nullptr, batch_destroy};
            std::vector<double> stages;
            std::vector<uint8_t> active;
            std::vector<BatchOutput> out;
            bool pending = false;
            int rounds = 0;
        };
        std::vector<Half> half(H);
        for (size_t k = 0; k < m; ++k) half[k % H].ks.push_back(k);
        for (Half &h : half) {
            std::vector<Metric *> mh;
            for (size_t k : h.ks) mh.push_back(metrics[k]);
            h.eng.reset(batch_create(mh.data(), mh.size()));
            h.stages.assign(12 * h.ks.size(), 0.0);
            h.active.assign(h.ks.size(), 1);
        }
        const auto tp2 = now();
        std::vector<RsgdRun> runs(m);
        for (Half &h : half)
            for (size_t j = 0; j < h.ks.size(); ++j) {
                const size_t k = h.ks[j];
                const size_t i = who[k];
                RsgdRun &r = runs[k];
                r.cfg = &cfg[i];
                r.T = Ts[i];
                r.p = P[i];
                r.prev.assign(r.p.size(), 0.0);
                r.est.build(inputs[k].grid, r.T);
                r.scales = cfg[i].scales_mode == SITKHIP_SCALES_PHYSICAL_SHIFT
                               ? r.est.scales()
                               : std::vector<double>(r.T.nparams, 1.0);
                r.T.set_parameters(r.p.data());
                r.T.stage(&h.stages[12 * j]);
            }
        // the metric value of each pair at its probe parameters (first level,
        // before optimizing), from the same prepared level
        if (probe && level == 0) {
            for (Half &h : half) {
                std::vector<double> pst(12 * h.ks.size());
                for (size_t j = 0; j < h.ks.size(); ++j) {
                    Transform T = Ts[who[h.ks[j]]];
                    T.set_parameters(probe + 12 * who[h.ks[j]]);
                    T.stage(&pst[12 * j]);
                }
                std::vector<BatchOutput> po(h.ks.size());
                batch_evaluate(*h.eng, pst.data(), nullptr, po);
                for (size_t j = 0; j < h.ks.size(); ++j) probe_values[who[h.ks[j]]] = po[j].value;
            }
        }
        // tuning pays only when one evaluation is long enough to time
        for (Half &h : half) {
            size_t total_samples = 0;
            for (size_t k : h.ks) total_samples += metric_sample_count(*metrics[k]);
            if (autotune && total_samples >= (static_cast<size_t>(1) << 22) / H)
                batch_autotune(*h.eng, h.stages.data(), nullptr);
            batch_prepare(*h.eng);
        }
        const auto tp3 = now();
        // on an error, let the evaluations in flight finish before the
        // engines and metrics go
        struct DrainOnError {
            ~DrainOnError() {
                if (std::uncaught_exceptions() > 0) (void)hipStreamSynchronize(backend().stream);
            }
        } drain;
        for (Half &h : half) {
            batch_launch(*h.eng, h.stages.data(), h.active.data());
            h.pending = true;
        }
        double t_wait = 0.0, t_update = 0.0, t_launch = 0.0;  // host time per phase (trace)
        for (bool busy = true; busy;) {
            busy = false;
            for (Half &h : half) {
                if (!h.pending) continue;
                const auto tw = trace ? now() : tp3;
                batch_collect(*h.eng, h.out);
                const auto tu = trace ? now() : tp3;
                if (trace) t_wait += ms(tw, tu);
                ++h.rounds;
                bool any = false;
                for (size_t j = 0; j < h.ks.size(); ++j) {
                    if (!h.active[j]) continue;
                    const size_t k = h.ks[j];
                    RsgdRun &r = runs[k];
                    ++r.evaluations;
                    size_t valid = h.out[j].valid;
                    r.value = contract_objective(cfg[who[k]].metric, h.out[j].value, h.out[j].moments,
                                                 r.T, r.grad, valid);
                    r.valid = h.out[j].valid;
                    if (r.advance()) {
                        r.T.set_parameters(r.p.data());
                        r.T.stage(&h.stages[12 * j]);
                        any = true;
                    } else {
                        h.active[j] = 0;
                    }
                }
                const auto tl = trace ? now() : tp3;
                if (any) batch_launch(*h.eng, h.stages.data(), h.active.data());
                if (trace) {
                    t_update += ms(tu, tl);
                    t_launch += ms(tl, now());
                }
                h.pending = any;
                busy = busy || any;
            }
        }
        int rounds = 0;
        for (const Half &h : half) rounds = std::max(rounds, h.rounds);
        if (trace) {
            const auto tp4 = now();
            std::fprintf(stderr,
                         "register_batch level %d: %zu pairs, prepare %.1f ms (enqueued at %.1f ms), "
                         "batch %.1f ms, tune %.1f ms, %d evaluations %.1f ms (%.3f ms each; host "
                         "waits %.1f, updates %.1f, launches %.1f ms)\n",
                         level, m, ms(tp0, tp1), ms(tp0, tpe), ms(tp1, tp2), ms(tp2, tp3), rounds,
                         ms(tp3, tp4), ms(tp3, tp4) / rounds, t_wait, t_update, t_launch);
        }
        for (size_t k = 0; k < m; ++k) {
            const size_t i = who[k];
            RsgdRun &r = runs[k];
            if (r.valid == 0)
                fail(SITKHIP_ERR_DEGENERATE_INPUT,
                     "pair " + std::to_string(i) + ": no valid samples at the result");
            res[i].level_value[level] = r.value;
            res[i].level_iterations[level] = r.taken;
            res[i].level_stop[level] = r.stop;
            res[i].iterations += r.taken;
            res[i].evaluations += r.evaluations;
            res[i].value = r.value;
            res[i].valid_points = r.valid;
            P[i] = r.p;
            Ts[i].set_parameters(P[i].data());
        }
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (size_t i = 0; i < n; ++i) {
        for (int k = 0; k < Ts[i].nparams; ++k) tr[i].params[k] = P[i][k];
        res[i].seconds = secs;
    }
}

}  // namespace sitkhip

extern "C" int sitkhip_register(sitkhip_image fixed, sitkhip_image moving,
                                const sitkhip_registration_config *config,
                                sitkhip_transform *transform, sitkhip_registration_result *result) {
    return sitkhip::run_guarded([&] {
        auto f = reinterpret_cast<sitkhip::Image *>(fixed);
        auto m = reinterpret_cast<sitkhip::Image *>(moving);
        if (!f || !m || !config || !transform || !result)
            sitkhip::fail(SITKHIP_ERR_INVALID_ARGUMENT, "null argument");
        sitkhip::register_images(*f, *m, *config, *transform, *result);
    });
}

extern "C" int sitkhip_register_batch(size_t n, const sitkhip_image *fixed,
                                      const sitkhip_image *moving,
                                      const sitkhip_registration_config *configs,
                                      sitkhip_transform *transforms,
                                      sitkhip_registration_result *results, int autotune) {
    return sitkhip_register_batch_probe(n, fixed, moving, configs, transforms, results, autotune,
                                        nullptr, nullptr);
}

extern "C" int sitkhip_register_batch_probe(size_t n, const sitkhip_image *fixed,
                                            const sitkhip_image *moving,
                                            const sitkhip_registration_config *configs,
                                            sitkhip_transform *transforms,
                                            sitkhip_registration_result *results, int autotune,
                                            const double *probe_params, double *probe_values) {
    return sitkhip::run_guarded([&] {
        if (!fixed || !moving || !configs || !transforms || !results)
            sitkhip::fail(SITKHIP_ERR_INVALID_ARGUMENT, "null argument");
        if ((probe_params == nullptr) != (probe_values == nullptr))
            sitkhip::fail(SITKHIP_ERR_INVALID_ARGUMENT, "probe parameters and values go together");
        sitkhip::register_batch(n, reinterpret_cast<sitkhip::Image *const *>(fixed),
                                reinterpret_cast<sitkhip::Image *const *>(moving), configs,
                                transforms, results, autotune != 0, probe_params, probe_values);
    });
}
