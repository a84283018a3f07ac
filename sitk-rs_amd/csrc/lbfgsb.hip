// Bound-constrained limited-memory BFGS on the host, as a reverse-
// communication state machine (the class comment in lbfgsb.hpp). The steps
// follow the reference's optimizer (crates/sitk/src/registration/lbfgsb.rs:
// minimize 1237-1588, cauchy 1776-2016, subsm 695-860, the Moré-Thuente line
// search 861-1200, projgr 1747-1774) -- the published algorithm of Byrd, Lu,
// Nocedal & Zhu with the 2011 projected subspace step -- with the small dense
// algebra written out here:
//
//   B = theta I - W M W',  W = [Y, theta S],  M = [[-D, L'], [L, theta S'S]]^-1
//
// with S, Y the stored corrections (oldest first), D = diag(s_i'y_i) and L the
// strictly lower triangle of S'Y. Products with M use the Cholesky factor of
// T = theta S'S + L D^-1 L' (block elimination); the subspace step solves
//
//   K = [[-(D + Y'ZZ'Y / theta), L_a' - R_z'], [L_a - R_z, theta S'AA'S]]
//
// (Z the free variables at the generalized Cauchy point, A the active ones)
// by the same elimination with two Cholesky factors. A factorization that
// fails discards the corrections and restarts the iteration from the
// steepest-descent model, as the reference does.
#include "lbfgsb.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>

#include "common.hpp"
#include "image.hpp"

namespace sitkhip {

namespace {

constexpr double kFtol = 1e-3;    // sufficient decrease
constexpr double kGtol = 0.9;     // curvature
constexpr double kXtol = 0.1;     // relative interval width
constexpr double kStpMin = 0.0;
constexpr double kBigStep = 1e10;
constexpr int kMaxBacktracks = 20;

double dot(const double *a, const double *b, size_t n) {
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

// In-place upper Cholesky factor U (A = U'U) of the k x k symmetric matrix in
// the upper triangle of a (row-major, leading dimension ld), column by column.
// False when a pivot is not positive.
bool cholesky_upper(double *a, int k, int ld) {
    for (int j = 0; j < k; ++j) {
        double s = 0.0;
        for (int i = 0; i < j; ++i) {
            double t = a[i * ld + j];
            for (int q = 0; q < i; ++q) t -= a[q * ld + i] * a[q * ld + j];
            t /= a[i * ld + i];
            a[i * ld + j] = t;
            s += t * t;
        }
        s = a[j * ld + j] - s;
        if (s <= 0.0) return false;
        a[j * ld + j] = std::sqrt(s);
    }
    return true;
}

// b := U'^-1 b
void solve_upper_t(const double *u, int k, int ld, double *b) {
    for (int i = 0; i < k; ++i) {
        double t = b[i];
        for (int q = 0; q < i; ++q) t -= u[q * ld + i] * b[q];
        b[i] = t / u[i * ld + i];
    }
}

// b := U^-1 b
void solve_upper(const double *u, int k, int ld, double *b) {
    for (int i = k - 1; i >= 0; --i) {
        double t = b[i];
        for (int q = i + 1; q < k; ++q) t -= u[i * ld + q] * b[q];
        b[i] = t / u[i * ld + i];
    }
}

// One safeguarded step of the Moré-Thuente search (MINPACK-2 dcstep): updates
// the interval [stx, sty] that brackets a step satisfying the conditions and
// computes the next trial step from cubic / quadratic interpolants.
void interval_step(double &stx, double &fx, double &dx, double &sty, double &fy, double &dy,
                   double &stp, double fp, double dp, bool &brackt, double stpmin, double stpmax) {
    const double sgnd = dp * (dx / std::fabs(dx));
    double stpf;
    if (fp > fx) {
        // higher value: the minimum is bracketed
        const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
        const double s = std::max({std::fabs(theta), std::fabs(dx), std::fabs(dp)});
        double gamma = s * std::sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
        if (stp < stx) gamma = -gamma;
        const double p = (gamma - dx) + theta;
        const double q = ((gamma - dx) + gamma) + dp;
        const double r = p / q;
        const double stpc = stx + r * (stp - stx);
        const double stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2.0) * (stp - stx);
        if (std::fabs(stpc - stx) < std::fabs(stpq - stx)) stpf = stpc;
        else stpf = stpc + (stpq - stpc) / 2.0;
        brackt = true;
    } else if (sgnd < 0.0) {
        // lower value, derivatives of opposite sign: bracketed
        const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
        const double s = std::max({std::fabs(theta), std::fabs(dx), std::fabs(dp)});
        double gamma = s * std::sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
        if (stp > stx) gamma = -gamma;
        const double p = (gamma - dp) + theta;
        const double q = ((gamma - dp) + gamma) + dx;
        const double r = p / q;
        const double stpc = stp + r * (stx - stp);
        const double stpq = stp + (dp / (dp - dx)) * (stx - stp);
        stpf = std::fabs(stpc - stp) > std::fabs(stpq - stp) ? stpc : stpq;
        brackt = true;
    } else if (std::fabs(dp) < std::fabs(dx)) {
        // lower value, same sign, the derivative's magnitude decreases
        const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
        const double s = std::max({std::fabs(theta), std::fabs(dx), std::fabs(dp)});
        double gamma = s * std::sqrt(std::max(0.0, (theta / s) * (theta / s) - (dx / s) * (dp / s)));
        if (stp > stx) gamma = -gamma;
        const double p = (gamma - dp) + theta;
        const double q = (gamma + (dx - dp)) + gamma;
        const double r = p / q;
        double stpc;
        if (r < 0.0 && gamma != 0.0) stpc = stp + r * (stx - stp);
        else if (stp > stx) stpc = stpmax;
        else stpc = stpmin;
        const double stpq = stp + (dp / (dp - dx)) * (stx - stp);
        if (brackt) {
            stpf = std::fabs(stpc - stp) < std::fabs(stpq - stp) ? stpc : stpq;
            if (stp > stx) stpf = std::min(stp + 0.66 * (sty - stp), stpf);
            else stpf = std::max(stp + 0.66 * (sty - stp), stpf);
        } else {
            stpf = std::fabs(stpc - stp) > std::fabs(stpq - stp) ? stpc : stpq;
            stpf = std::min(stpmax, stpf);
            stpf = std::max(stpmin, stpf);
        }
    } else {
        // lower value, same sign, the derivative's magnitude does not decrease
        if (brackt) {
            const double theta = 3.0 * (fp - fy) / (sty - stp) + dy + dp;
            const double s = std::max({std::fabs(theta), std::fabs(dy), std::fabs(dp)});
            double gamma = s * std::sqrt((theta / s) * (theta / s) - (dy / s) * (dp / s));
            if (stp > sty) gamma = -gamma;
            const double p = (gamma - dp) + theta;
            const double q = ((gamma - dp) + gamma) + dy;
            const double r = p / q;
            stpf = stp + r * (sty - stp);
        } else {
            stpf = stp > stx ? stpmax : stpmin;
        }
    }
    if (fp > fx) {
        sty = stp;
        fy = fp;
        dy = dp;
    } else {
        if (sgnd < 0.0) {
            sty = stx;
            fy = fx;
            dy = dx;
        }
        stx = stp;
        fx = fp;
        dx = dp;
    }
    stp = stpf;
}

}  // namespace

// MINPACK-2 dcsrch: the first call (started == false) sets up the interval and
// asks for the value at stp; later calls take the value f and directional
// derivative g at stp.
int Lbfgsb::Search::step(double f, double g, double &stp, double stpmax) {
    if (!started) {
        started = true;
        brackt = false;
        stage = 1;
        finit = f;
        ginit = g;
        gtest = kFtol * g;
        width = stpmax - kStpMin;
        width1 = width / 0.5;
        stx = 0.0;
        fx = finit;
        gx = ginit;
        sty = 0.0;
        fy = finit;
        gy = ginit;
        stmin = 0.0;
        stmax = stp + 4.0 * stp;
        return 0;
    }
    const double ftest = finit + stp * gtest;
    if (stage == 1 && f <= ftest && g >= 0.0) stage = 2;
    int task = 0;
    if (brackt && (stp <= stmin || stp >= stmax)) task = 2;
    if (brackt && stmax - stmin <= kXtol * stmax) task = 2;
    if (stp == stpmax && f <= ftest && g <= gtest) task = 2;
    if (stp == kStpMin && (f > ftest || g >= gtest)) task = 2;
    if (f <= ftest && std::fabs(g) <= kGtol * (-ginit)) task = 1;
    if (task) return task;

    if (stage == 1 && f <= fx && f > ftest) {
        // the modified function psi(stp) = f - stp * gtest while the decrease
        // is not yet sufficient
        const double fm = f - stp * gtest;
        double fxm = fx - stx * gtest, fym = fy - sty * gtest;
        const double gm = g - gtest;
        double gxm = gx - gtest, gym = gy - gtest;
        interval_step(stx, fxm, gxm, sty, fym, gym, stp, fm, gm, brackt, stmin, stmax);
        fx = fxm + stx * gtest;
        fy = fym + sty * gtest;
        gx = gxm + gtest;
        gy = gym + gtest;
    } else {
        interval_step(stx, fx, gx, sty, fy, gy, stp, f, g, brackt, stmin, stmax);
    }
    if (brackt) {
        if (std::fabs(sty - stx) >= 0.66 * width1) stp = stx + 0.5 * (sty - stx);
        width1 = width;
        width = std::fabs(sty - stx);
    }
    if (brackt) {
        stmin = std::min(stx, sty);
        stmax = std::max(stx, sty);
    } else {
        stmin = stp + 1.1 * (stp - stx);
        stmax = stp + 4.0 * (stp - stx);
    }
    stp = std::max(stp, kStpMin);
    stp = std::min(stp, stpmax);
    if (brackt && (stp <= stmin || stp >= stmax || stmax - stmin <= kXtol * stmax)) stp = stx;
    return 0;
}

Lbfgsb::Lbfgsb(size_t n, const LbfgsbOptions &opt, const double *lower, const double *upper,
               const int *kind)
    : n_(n), m_(opt.corrections), opt_(opt) {
    if (n == 0) fail(SITKHIP_ERR_INVALID_ARGUMENT, "L-BFGS-B needs at least one parameter");
    if (opt.corrections <= 0) fail(SITKHIP_ERR_INVALID_ARGUMENT, "max_corrections must be positive");
    if (!(opt.factr >= 0.0))
        fail(SITKHIP_ERR_INVALID_ARGUMENT, "cost function convergence factor must be non-negative");
    l_.assign(n, 0.0);
    u_.assign(n, 0.0);
    nbd_.assign(n, 0);
    if (kind) {
        for (size_t i = 0; i < n; ++i) {
            if (kind[i] < 0 || kind[i] > 3)
                fail(SITKHIP_ERR_INVALID_ARGUMENT, "bound selection values must be 0..3");
            nbd_[i] = kind[i];
            if (kind[i] == 0) continue;
            if ((kind[i] <= 2 && !lower) || (kind[i] >= 2 && !upper))
                fail(SITKHIP_ERR_INVALID_ARGUMENT, "a bounded variable needs its bound arrays");
            if (kind[i] <= 2) l_[i] = lower[i];
            if (kind[i] >= 2) u_[i] = upper[i];
            if (kind[i] == 2 && !(l_[i] <= u_[i]))
                fail(SITKHIP_ERR_INVALID_ARGUMENT,
                     "infeasible bound at index " + std::to_string(i) + ": lower > upper");
        }
    }
    x_.assign(n, 0.0);
    g_.assign(n, 0.0);
    z_.assign(n, 0.0);
    d_.assign(n, 0.0);
    t_.assign(n, 0.0);
    r_.assign(n, 0.0);
    where_.assign(n, 0);
    S_.assign(n * static_cast<size_t>(m_), 0.0);
    Y_.assign(n * static_cast<size_t>(m_), 0.0);
    sy_.assign(static_cast<size_t>(m_) * m_, 0.0);
    ss_.assign(static_cast<size_t>(m_) * m_, 0.0);
    syf_.assign(static_cast<size_t>(m_) * m_, 0.0);
    yy_.assign(static_cast<size_t>(m_) * m_, 0.0);
    tchol_.assign(static_cast<size_t>(m_) * m_, 0.0);
    c_.assign(2 * static_cast<size_t>(m_), 0.0);
    xbest_.assign(n, 0.0);
    stop_ = kLbMaxIterations;  // not running until start()
}

bool Lbfgsb::start(const double *x0) {
    if (!x0) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null initial parameters");
    constrained_ = false;
    boxed_ = true;
    for (size_t i = 0; i < n_; ++i) {
        double v = x0[i];
        if (nbd_[i] > 0) {
            if (nbd_[i] <= 2 && v <= l_[i]) v = std::max(v, l_[i]);
            else if (nbd_[i] >= 2 && v >= u_[i]) v = std::min(v, u_[i]);
        }
        x_[i] = v;
        if (nbd_[i] != 2) boxed_ = false;
        if (nbd_[i] == 0) {
            where_[i] = -1;
        } else {
            constrained_ = true;
            where_[i] = (nbd_[i] == 2 && u_[i] - l_[i] <= 0.0) ? 3 : 0;
        }
    }
    free_.clear();
    active_.clear();
    for (size_t i = 0; i < n_; ++i) free_.push_back(static_cast<int>(i));
    reset_memory();
    iter_ = 0;
    nevals_ = 0;
    step_ = Step::Initial;
    stop_ = kLbRunning;
    return true;
}

void Lbfgsb::reset_memory() {
    col_ = 0;
    head_ = 0;
    theta_ = 1.0;
    updated_ = false;
}

bool Lbfgsb::finish(int reason) {
    stop_ = reason;
    return false;
}

double Lbfgsb::projected_gradient_norm() const {
    double norm = 0.0;
    for (size_t i = 0; i < n_; ++i) {
        double gi = g_[i];
        if (nbd_[i] != 0) {
            if (gi < 0.0) {
                if (nbd_[i] >= 2) gi = std::max(x_[i] - u_[i], gi);
            } else {
                if (nbd_[i] <= 2) gi = std::min(x_[i] - l_[i], gi);
            }
        }
        norm = std::max(norm, std::fabs(gi));
    }
    return norm;
}

bool Lbfgsb::tell(double f, const double *g) {
    if (stop_ != kLbRunning) fail(SITKHIP_ERR_INVALID_ARGUMENT, "the optimizer is not running");
    if (!g) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null gradient");
    f_ = f;
    std::memcpy(g_.data(), g, n_ * sizeof(double));
    if (nevals_ == 0 || f < fbest_) {
        fbest_ = f;
        xbest_ = x_;
    }
    ++nevals_;
    if (nevals_ > opt_.max_evaluations) return finish(kLbEvaluations);
    if (step_ == Step::Initial) {
        sbgnrm_ = projected_gradient_norm();
        if (sbgnrm_ <= opt_.pgtol) return finish(kLbGradient);
        step_ = Step::Direction;
    }
    return run();
}

bool Lbfgsb::run() {
    for (;;) {
        if (step_ == Step::Direction) {
            bool ok = true;
            if (!constrained_ && col_ > 0) {
                z_ = x_;  // every variable free, no Cauchy step needed
            } else {
                ok = generalized_cauchy_point();
                if (ok) {
                    free_.clear();
                    active_.clear();
                    for (size_t i = 0; i < n_; ++i)
                        (where_[i] <= 0 ? free_ : active_).push_back(static_cast<int>(i));
                }
            }
            if (ok && !free_.empty() && col_ != 0) ok = subspace_minimization();
            if (!ok) {
                reset_memory();
                continue;
            }
            begin_line_search();
            step_ = Step::Search;
        }
        const double gd = dot(g_.data(), d_.data(), n_);
        if (ifun_ == 0) {
            gdold_ = gd;
            if (gd >= 0.0) {
                // not a descent direction
                if (!line_search_failed()) return false;
                continue;
            }
        }
        const int task = ls_.step(f_, gd, stp_, stpmx_);
        if (task != 0) {
            if (!accept_iterate(gd)) return false;
            step_ = Step::Direction;
            continue;
        }
        ++ifun_;
        if (ifun_ - 1 >= kMaxBacktracks) {
            if (!line_search_failed()) return false;
            continue;
        }
        if (stp_ == 1.0) {
            x_ = z_;
        } else {
            for (size_t i = 0; i < n_; ++i) {
                double v = stp_ * d_[i] + t_[i];
                if (nbd_[i] == 1 || nbd_[i] == 2) v = std::max(v, l_[i]);
                if (nbd_[i] == 2 || nbd_[i] == 3) v = std::min(v, u_[i]);
                x_[i] = v;
            }
        }
        return true;
    }
}

bool Lbfgsb::line_search_failed() {
    x_ = t_;
    g_ = r_;
    f_ = fold_;
    if (col_ == 0) return finish(kLbLineSearch);
    reset_memory();
    step_ = Step::Direction;
    return true;
}

bool Lbfgsb::accept_iterate(double gd) {
    ++iter_;
    sbgnrm_ = projected_gradient_norm();
    if (iter_ >= opt_.max_iterations) return finish(kLbMaxIterations);
    if (sbgnrm_ <= opt_.pgtol) return finish(kLbGradient);
    const double scale = std::max({std::fabs(fold_), std::fabs(f_), 1.0});
    if (fold_ - f_ <= opt_.factr * DBL_EPSILON * scale) return finish(kLbReduction);
    update_memory(gd);
    return true;
}

void Lbfgsb::begin_line_search() {
    for (size_t i = 0; i < n_; ++i) d_[i] = z_[i] - x_[i];
    dtd_ = dot(d_.data(), d_.data(), n_);
    const double dnorm = std::sqrt(dtd_);
    stpmx_ = kBigStep;
    if (constrained_) {
        if (iter_ == 0) {
            stpmx_ = 1.0;
        } else {
            for (size_t i = 0; i < n_; ++i) {
                const double a1 = d_[i];
                if (nbd_[i] == 0) continue;
                if (a1 < 0.0 && nbd_[i] <= 2) {
                    const double a2 = l_[i] - x_[i];
                    if (a2 >= 0.0) stpmx_ = 0.0;
                    else if (a1 * stpmx_ < a2) stpmx_ = a2 / a1;
                } else if (a1 > 0.0 && nbd_[i] >= 2) {
                    const double a2 = u_[i] - x_[i];
                    if (a2 <= 0.0) stpmx_ = 0.0;
                    else if (a1 * stpmx_ > a2) stpmx_ = a2 / a1;
                }
            }
        }
    }
    stp_ = (iter_ == 0 && !boxed_) ? std::min(1.0 / dnorm, stpmx_) : 1.0;
    // the search needs stp within [0, stpmax]; z is feasible, so only rounding
    // can put the unit step past the box
    stp_ = std::min(stp_, stpmx_);
    t_ = x_;
    r_ = g_;
    fold_ = f_;
    ifun_ = 0;
    ls_ = Search{};
}

// L(i,k) = s_i'y_k for i > k; D(k) = s_k'y_k.
double Lbfgsb::lower_factor(int i, int k) const { return sy_[i * m_ + k]; }

bool Lbfgsb::form_t() {
    const int c = col_;
    for (int i = 0; i < c; ++i)
        for (int j = i; j < c; ++j) {
            double acc = 0.0;
            for (int k = 0; k < i; ++k) acc += lower_factor(i, k) * lower_factor(j, k) / sy_[k * m_ + k];
            tchol_[i * m_ + j] = acc + theta_ * ss_[i * m_ + j];
        }
    return cholesky_upper(tchol_.data(), c, m_);
}

// out = M v for a 2*col vector v = [v1; v2]:
//   b = T^-1 (v2 + L D^-1 v1),  a = D^-1 (L' b - v1),  out = [a; b].
bool Lbfgsb::middle_product(const double *v, double *out) const {
    const int c = col_;
    if (c == 0) return true;
    double *a = out, *b = out + c;
    b[0] = v[c];
    for (int i = 1; i < c; ++i) {
        double acc = 0.0;
        for (int k = 0; k < i; ++k) acc += lower_factor(i, k) * v[k] / sy_[k * m_ + k];
        b[i] = v[c + i] + acc;
    }
    solve_upper_t(tchol_.data(), c, m_, b);
    for (int i = 0; i < c; ++i) a[i] = v[i] / std::sqrt(sy_[i * m_ + i]);
    solve_upper(tchol_.data(), c, m_, b);
    for (int i = 0; i < c; ++i) a[i] = -a[i] / std::sqrt(sy_[i * m_ + i]);
    for (int i = 0; i < c; ++i) {
        double acc = 0.0;
        for (int k = i + 1; k < c; ++k) acc += lower_factor(k, i) * b[k] / sy_[i * m_ + i];
        a[i] += acc;
    }
    return true;
}

// The generalized Cauchy point: the first local minimizer of the quadratic
// model along the projected steepest-descent path x(t) = P(x - t g), walking
// the breakpoints where variables reach their bounds in increasing order.
// Leaves the point in z_, the variables' states in where_ and W'(z - x) in c_.
bool Lbfgsb::generalized_cauchy_point() {
    const int c = col_;
    const int c2 = 2 * c;
    std::fill(c_.begin(), c_.end(), 0.0);
    if (sbgnrm_ <= 0.0) {
        z_ = x_;
        return true;
    }
    std::vector<double> p(c2, 0.0), v(c2, 0.0), wbp(c2, 0.0);
    struct Break {
        double t;
        int i;
    };
    std::vector<Break> bp;
    bool bounded = true;
    int unbounded_moving = 0;
    double f1 = 0.0;
    for (size_t i = 0; i < n_; ++i) {
        const double neggi = -g_[i];
        double tl = 0.0, tu = 0.0;
        if (where_[i] != 3 && where_[i] != -1) {
            if (nbd_[i] <= 2) tl = x_[i] - l_[i];
            if (nbd_[i] >= 2) tu = u_[i] - x_[i];
            const bool at_lower = nbd_[i] <= 2 && tl <= 0.0;
            const bool at_upper = nbd_[i] >= 2 && tu <= 0.0;
            where_[i] = 0;
            if (at_lower) {
                if (neggi <= 0.0) where_[i] = 1;
            } else if (at_upper) {
                if (neggi >= 0.0) where_[i] = 2;
            } else if (std::fabs(neggi) <= 0.0) {
                where_[i] = -3;
            }
        }
        if (where_[i] != 0 && where_[i] != -1) {
            d_[i] = 0.0;
            continue;
        }
        d_[i] = neggi;
        f1 -= neggi * neggi;
        for (int j = 0; j < c; ++j) {
            p[j] += y(j)[i] * neggi;
            p[c + j] += s(j)[i] * neggi;
        }
        if (nbd_[i] <= 2 && nbd_[i] != 0 && neggi < 0.0) {
            bp.push_back({tl / (-neggi), static_cast<int>(i)});
        } else if (nbd_[i] >= 2 && neggi > 0.0) {
            bp.push_back({tu / neggi, static_cast<int>(i)});
        } else {
            ++unbounded_moving;
            if (std::fabs(neggi) > 0.0) bounded = false;
        }
    }
    if (theta_ != 1.0)
        for (int j = 0; j < c; ++j) p[c + j] *= theta_;
    z_ = x_;
    if (bp.empty() && unbounded_moving == 0) return true;

    double f2 = -theta_ * f1;
    const double f2_org = f2;
    if (c > 0) {
        if (!middle_product(p.data(), v.data())) return false;
        f2 -= dot(v.data(), p.data(), c2);
    }
    double dtm = -f1 / f2;
    double tsum = 0.0;
    bool all_fixed = false;
    if (!bp.empty()) {
        std::stable_sort(bp.begin(), bp.end(), [](const Break &a, const Break &b) { return a.t < b.t; });
        const size_t nbreak = bp.size();
        size_t nleft = nbreak;
        double tj = 0.0;
        for (size_t k = 0; k < nbreak; ++k) {
            const double tj0 = tj;
            tj = bp[k].t;
            const int ib = bp[k].i;
            const double dt = tj - tj0;
            if (dtm < dt) break;  // the minimizer lies in this segment
            tsum += dt;
            --nleft;
            const double dib = d_[ib];
            d_[ib] = 0.0;
            double zib;
            if (dib > 0.0) {
                zib = u_[ib] - x_[ib];
                z_[ib] = u_[ib];
                where_[ib] = 2;
            } else {
                zib = l_[ib] - x_[ib];
                z_[ib] = l_[ib];
                where_[ib] = 1;
            }
            if (nleft == 0 && nbreak == n_) {
                dtm = dt;
                all_fixed = true;
                break;
            }
            const double dib2 = dib * dib;
            f1 = f1 + dt * f2 + dib2 - theta_ * dib * zib;
            f2 -= theta_ * dib2;
            if (c > 0) {
                for (int j = 0; j < c2; ++j) c_[j] += dt * p[j];
                for (int j = 0; j < c; ++j) {
                    wbp[j] = y(j)[ib];
                    wbp[c + j] = theta_ * s(j)[ib];
                }
                if (!middle_product(wbp.data(), v.data())) return false;
                const double wmc = dot(c_.data(), v.data(), c2);
                const double wmp = dot(p.data(), v.data(), c2);
                const double wmw = dot(wbp.data(), v.data(), c2);
                for (int j = 0; j < c2; ++j) p[j] -= dib * wbp[j];
                f1 += dib * wmc;
                f2 = f2 + dib * 2.0 * wmp - dib2 * wmw;
            }
            f2 = std::max(DBL_EPSILON * f2_org, f2);
            if (nleft > 0) {
                dtm = -f1 / f2;
                continue;
            }
            dtm = bounded ? 0.0 : -f1 / f2;
            break;
        }
    }
    if (!all_fixed) {
        if (dtm <= 0.0) dtm = 0.0;
        tsum += dtm;
        for (size_t i = 0; i < n_; ++i) z_[i] += tsum * d_[i];
    }
    if (c > 0)
        for (int j = 0; j < c2; ++j) c_[j] += dtm * p[j];
    return true;
}

// Minimizes the quadratic model over the free variables from the Cauchy
// point (the direct primal method), then projects the step onto the box and,
// if the projected step is not a descent direction, backtracks it to the
// box. Updates z_.
bool Lbfgsb::subspace_minimization() {
    const int c = col_;
    const int c2 = 2 * c;
    const size_t nsub = free_.size();
    // reduced gradient at the Cauchy point, negated: r = -Z'(g + theta (z - x) - W M c)
    std::vector<double> rs(nsub);
    if (!constrained_ && c > 0) {
        for (size_t q = 0; q < nsub; ++q) rs[q] = -g_[free_[q]];
    } else {
        std::vector<double> mc(c2, 0.0);
        if (!middle_product(c_.data(), mc.data())) return false;
        for (size_t q = 0; q < nsub; ++q) {
            const int k = free_[q];
            rs[q] = -theta_ * (z_[k] - x_[k]) - g_[k];
        }
        for (int j = 0; j < c; ++j) {
            const double a1 = mc[j], a2 = theta_ * mc[c + j];
            const double *yj = y(j), *sj = s(j);
            for (size_t q = 0; q < nsub; ++q) {
                const int k = free_[q];
                rs[q] += yj[k] * a1 + sj[k] * a2;
            }
        }
    }

    // K's blocks: P = D + Y'ZZ'Y / theta, K12 = L_a' - R_z', K22 = theta S'AA'S.
    std::vector<double> P(c * c), Q(c * c), S2(c * c);
    if (active_.empty()) {
        // every variable free: the sums over Z are the stored full products
        for (int i = 0; i < c; ++i)
            for (int j = i; j < c; ++j) {
                P[i * c + j] = yy_[i * m_ + j] / theta_ + (i == j ? sy_[i * m_ + i] : 0.0);
                S2[i * c + j] = theta_ * 0.0;
            }
        for (int i = 0; i < c; ++i)
            for (int j = 0; j < c; ++j) Q[i * c + j] = j > i ? 0.0 : -syf_[j * m_ + i];
    } else {
        for (int i = 0; i < c; ++i)
            for (int j = i; j < c; ++j) {
                const double *yi = y(i), *yj = y(j);
                double acc = 0.0;
                for (int k : free_) acc += yi[k] * yj[k];
                P[i * c + j] = acc / theta_ + (i == j ? sy_[i * m_ + i] : 0.0);
            }
        for (int i = 0; i < c; ++i)        // row: y index
            for (int j = 0; j < c; ++j) {  // column: s index
                const double *yi = y(i), *sj = s(j);
                double acc = 0.0;
                if (j > i) {
                    for (int k : active_) acc += sj[k] * yi[k];
                    Q[i * c + j] = acc;
                } else {
                    for (int k : free_) acc += sj[k] * yi[k];
                    Q[i * c + j] = -acc;
                }
            }
        for (int i = 0; i < c; ++i)
            for (int j = i; j < c; ++j) {
                const double *si = s(i), *sj = s(j);
                double acc = 0.0;
                for (int k : active_) acc += si[k] * sj[k];
                S2[i * c + j] = theta_ * acc;
            }
    }
    if (!cholesky_upper(P.data(), c, c)) return false;
    // Q := U1'^-1 K12, column by column; S2 := K22 + Q'Q
    for (int j = 0; j < c; ++j) {
        std::vector<double> colj(c);
        for (int i = 0; i < c; ++i) colj[i] = Q[i * c + j];
        solve_upper_t(P.data(), c, c, colj.data());
        for (int i = 0; i < c; ++i) Q[i * c + j] = colj[i];
    }
    for (int i = 0; i < c; ++i)
        for (int j = i; j < c; ++j) {
            double acc = 0.0;
            for (int k = 0; k < c; ++k) acc += Q[k * c + i] * Q[k * c + j];
            S2[i * c + j] += acc;
        }
    if (!cholesky_upper(S2.data(), c, c)) return false;

    // wv = W'Z r, then wv := K^-1 wv
    std::vector<double> w1(c), w2(c);
    for (int j = 0; j < c; ++j) {
        const double *yj = y(j), *sj = s(j);
        double a = 0.0, b = 0.0;
        for (size_t q = 0; q < nsub; ++q) {
            const int k = free_[q];
            a += yj[k] * rs[q];
            b += sj[k] * rs[q];
        }
        w1[j] = a;
        w2[j] = theta_ * b;
    }
    // -P a + K12 b = w1, K12' a + K22 b = w2:
    //   b = S2^-1 (w2 + Q' U1'^-1 w1),  a = U1^-1 (Q b - U1'^-1 w1)
    solve_upper_t(P.data(), c, c, w1.data());
    for (int j = 0; j < c; ++j) {
        double acc = 0.0;
        for (int k = 0; k < c; ++k) acc += Q[k * c + j] * w1[k];
        w2[j] += acc;
    }
    solve_upper_t(S2.data(), c, c, w2.data());
    solve_upper(S2.data(), c, c, w2.data());
    std::vector<double> a(c);
    for (int i = 0; i < c; ++i) {
        double acc = 0.0;
        for (int k = 0; k < c; ++k) acc += Q[i * c + k] * w2[k];
        a[i] = acc - w1[i];
    }
    solve_upper(P.data(), c, c, a.data());

    // Newton step on the free variables: d = (r + Z'Y a / theta + Z'S b) / theta
    std::vector<double> dsub(rs);
    for (int j = 0; j < c; ++j) {
        const double *yj = y(j), *sj = s(j);
        for (size_t q = 0; q < nsub; ++q) {
            const int k = free_[q];
            dsub[q] = dsub[q] + yj[k] * a[j] / theta_ + sj[k] * w2[j];
        }
    }
    const double inv_theta = 1.0 / theta_;
    for (double &v : dsub) v *= inv_theta;

    // project z + d onto the box
    const std::vector<double> zc = z_;
    bool clipped = false;
    for (size_t q = 0; q < nsub; ++q) {
        const int k = free_[q];
        const double v = zc[k] + dsub[q];
        if (nbd_[k] == 1) {
            z_[k] = std::max(l_[k], v);
            if (z_[k] == l_[k]) clipped = true;
        } else if (nbd_[k] == 2) {
            z_[k] = std::min(u_[k], std::max(l_[k], v));
            if (z_[k] == l_[k] || z_[k] == u_[k]) clipped = true;
        } else if (nbd_[k] == 3) {
            z_[k] = std::min(u_[k], v);
            if (z_[k] == u_[k]) clipped = true;
        } else {
            z_[k] = v;
        }
    }
    if (!clipped) return true;
    double ddp = 0.0;
    for (size_t i = 0; i < n_; ++i) ddp += (z_[i] - x_[i]) * g_[i];
    if (ddp <= 0.0) return true;
    // the projected point ascends: take the longest feasible step along d instead
    z_ = zc;
    double alpha = 1.0, t1 = 1.0;
    size_t ibd = 0;
    bool have_ibd = false;
    for (size_t q = 0; q < nsub; ++q) {
        const int k = free_[q];
        const double dk = dsub[q];
        if (nbd_[k] == 0) continue;
        if (dk < 0.0 && nbd_[k] <= 2) {
            const double t2 = l_[k] - z_[k];
            if (t2 >= 0.0) t1 = 0.0;
            else if (dk * alpha < t2) t1 = t2 / dk;
        } else if (dk > 0.0 && nbd_[k] >= 2) {
            const double t2 = u_[k] - z_[k];
            if (t2 <= 0.0) t1 = 0.0;
            else if (dk * alpha > t2) t1 = t2 / dk;
        }
        if (t1 < alpha) {
            alpha = t1;
            ibd = q;
            have_ibd = true;
        }
    }
    if (alpha < 1.0 && have_ibd) {
        const double dk = dsub[ibd];
        const int k = free_[ibd];
        if (dk > 0.0) {
            z_[k] = u_[k];
            dsub[ibd] = 0.0;
        } else if (dk < 0.0) {
            z_[k] = l_[k];
            dsub[ibd] = 0.0;
        }
    }
    for (size_t q = 0; q < nsub; ++q) z_[free_[q]] += alpha * dsub[q];
    return true;
}

// Stores the step s = stp d and the gradient change y = g - g_old as the newest
// correction (dropping the oldest when m are held) unless s'y is not safely
// positive, and refactors T.
void Lbfgsb::update_memory(double gd) {
    for (size_t i = 0; i < n_; ++i) r_[i] = g_[i] - r_[i];
    const double rr = dot(r_.data(), r_.data(), n_);
    double dr, ddum;
    if (stp_ == 1.0) {
        dr = gd - gdold_;
        ddum = -gdold_;
    } else {
        dr = (gd - gdold_) * stp_;
        for (double &v : d_) v *= stp_;
        ddum = -gdold_ * stp_;
    }
    if (dr <= DBL_EPSILON * ddum) {
        updated_ = false;
        return;
    }
    updated_ = true;
    int slot_new;
    if (col_ < m_) {
        slot_new = (head_ + col_) % m_;
        ++col_;
    } else {
        slot_new = head_;
        head_ = (head_ + 1) % m_;
        for (int i = 0; i + 1 < col_; ++i)
            for (int j = 0; j + 1 < col_; ++j) {
                sy_[i * m_ + j] = sy_[(i + 1) * m_ + j + 1];
                ss_[i * m_ + j] = ss_[(i + 1) * m_ + j + 1];
                syf_[i * m_ + j] = syf_[(i + 1) * m_ + j + 1];
                yy_[i * m_ + j] = yy_[(i + 1) * m_ + j + 1];
            }
    }
    std::memcpy(S_.data() + static_cast<size_t>(slot_new) * n_, d_.data(), n_ * sizeof(double));
    std::memcpy(Y_.data() + static_cast<size_t>(slot_new) * n_, r_.data(), n_ * sizeof(double));
    theta_ = rr / dr;
    const int cn = col_ - 1;
    for (int j = 0; j < cn; ++j) {
        sy_[cn * m_ + j] = dot(d_.data(), y(j), n_);
        const double v = dot(s(j), d_.data(), n_);
        ss_[j * m_ + cn] = v;
        ss_[cn * m_ + j] = v;
    }
    ss_[cn * m_ + cn] = stp_ == 1.0 ? dtd_ : stp_ * stp_ * dtd_;
    sy_[cn * m_ + cn] = dr;
    // the products the all-free subspace step reads, each summed over the
    // variables in index order with the general branch's operand order
    for (int i = 0; i < cn; ++i) {
        const double *si = s(i), *yi = y(i);
        double a = 0.0, b = 0.0;
        for (size_t k = 0; k < n_; ++k) {
            a += si[k] * r_[k];
            b += yi[k] * r_[k];
        }
        syf_[i * m_ + cn] = a;
        yy_[i * m_ + cn] = b;
        yy_[cn * m_ + i] = b;
    }
    {
        double a = 0.0, b = 0.0;
        for (size_t k = 0; k < n_; ++k) {
            a += d_[k] * r_[k];
            b += r_[k] * r_[k];
        }
        syf_[cn * m_ + cn] = a;
        yy_[cn * m_ + cn] = b;
    }
    if (!form_t()) reset_memory();
}

}  // namespace sitkhip

using namespace sitkhip;

extern "C" {

int sitkhip_lbfgsb_create(size_t n, const sitkhip_lbfgsb_options *opt, const double *lower,
                          const double *upper, const int *bound_selection, sitkhip_lbfgsb *out) {
    return run_guarded([&] {
        if (!opt || !out) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null argument");
        LbfgsbOptions o;
        o.corrections = opt->corrections;
        o.factr = opt->factr;
        o.pgtol = opt->pgtol;
        o.max_iterations = opt->max_iterations;
        o.max_evaluations = opt->max_evaluations;
        *out = reinterpret_cast<sitkhip_lbfgsb>(new Lbfgsb(n, o, lower, upper, bound_selection));
    });
}

void sitkhip_lbfgsb_free(sitkhip_lbfgsb h) { delete reinterpret_cast<Lbfgsb *>(h); }

int sitkhip_lbfgsb_start(sitkhip_lbfgsb h, const double *x0, double *x, int *want) {
    return run_guarded([&] {
        auto o = reinterpret_cast<Lbfgsb *>(h);
        if (!o || !x || !want) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null argument");
        *want = o->start(x0) ? 1 : 0;
        std::memcpy(x, o->point(), o->size() * sizeof(double));
    });
}

int sitkhip_lbfgsb_tell(sitkhip_lbfgsb h, double f, const double *g, double *x, int *want) {
    return run_guarded([&] {
        auto o = reinterpret_cast<Lbfgsb *>(h);
        if (!o || !x || !want) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null argument");
        *want = o->tell(f, g) ? 1 : 0;
        std::memcpy(x, o->point(), o->size() * sizeof(double));
    });
}

int sitkhip_lbfgsb_result(sitkhip_lbfgsb h, double *best, double *value, int *iterations,
                          int *evaluations, int *stop) {
    return run_guarded([&] {
        auto o = reinterpret_cast<Lbfgsb *>(h);
        if (!o) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null handle");
        if (best) std::memcpy(best, o->best(), o->size() * sizeof(double));
        if (value) *value = o->best_value();
        if (iterations) *iterations = o->iterations();
        if (evaluations) *evaluations = o->evaluations();
        if (stop) *stop = o->stop();
    });
}

}  // extern "C"
