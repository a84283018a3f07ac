// Limited-memory BFGS with simple bounds (Byrd, Lu, Nocedal & Zhu, SIAM J.
// Sci. Comput. 16(5), 1995; the subspace step's projection of Morales &
// Nocedal, ACM TOMS 38(1), 2011), the optimizer of the reference's deformable
// registration (crates/sitk/src/registration/lbfgsb.rs:1590-1746, itself the
// netlib 3.0 algorithm ITK wraps).
//
// Reverse communication: the optimizer never calls the objective. start()
// and tell() return true while they want the value and gradient at point();
// the caller evaluates there and calls tell(). This lets one device launch
// evaluate the trial points of many independent optimizers at once.
//
// Host code only; no device calls.
#pragma once

#include <cstddef>
#include <vector>

namespace sitkhip {

// Stop reasons, numbered as sitkhip_bspline_result.stop.
enum LbfgsbStop {
    kLbRunning = -1,
    kLbMaxIterations = 0,
    kLbGradient = 1,       // projected-gradient infinity norm <= pgtol
    kLbReduction = 2,      // relative reduction <= factr * eps
    kLbEvaluations = 3,    // evaluation count above the maximum
    kLbLineSearch = 4,     // no acceptable step from a steepest-descent start
};

struct LbfgsbOptions {
    int corrections = 5;          // m
    double factr = 1e7;
    double pgtol = 1e-5;
    int max_iterations = 100;
    int max_evaluations = 100;
};

class Lbfgsb {
public:
    // kind[i] (ITK BoundSelection): 0 unbounded, 1 lower only, 2 both, 3 upper
    // only. lower / upper / kind may be null for an unbounded problem. Throws
    // SITKHIP_ERR_INVALID_ARGUMENT for a bound pair with lower > upper.
    Lbfgsb(size_t n, const LbfgsbOptions &opt, const double *lower, const double *upper,
           const int *kind);

    // Starts from x0, projected onto the box.
    bool start(const double *x0);
    // The value and gradient at point().
    bool tell(double f, const double *g);

    const double *point() const { return x_.data(); }
    size_t size() const { return n_; }
    bool running() const { return stop_ == kLbRunning; }
    int stop() const { return stop_; }
    int iterations() const { return iter_; }
    int evaluations() const { return nevals_; }
    // The lowest value evaluated and where (the reference returns that point,
    // not the last iterate).
    double best_value() const { return fbest_; }
    const double *best() const { return xbest_.data(); }

private:
    enum class Step { Initial, Direction, Search };

    bool run();
    bool finish(int reason);
    void reset_memory();
    double projected_gradient_norm() const;
    bool generalized_cauchy_point();
    bool subspace_minimization();
    void begin_line_search();
    bool line_search_failed();
    bool accept_iterate(double gd);
    void update_memory(double gd);
    bool form_t();
    bool middle_product(const double *v, double *out) const;
    double lower_factor(int i, int k) const;  // L(i,k) = s_i'y_k, i > k
    const double *s(int j) const { return S_.data() + static_cast<size_t>(slot(j)) * n_; }
    const double *y(int j) const { return Y_.data() + static_cast<size_t>(slot(j)) * n_; }
    int slot(int j) const { return (head_ + j) % m_; }

    // Moré-Thuente line search (MINPACK-2 dcsrch / dcstep).
    struct Search {
        bool started = false, brackt = false;
        int stage = 1;
        double ginit = 0, gtest = 0, gx = 0, gy = 0, finit = 0, fx = 0, fy = 0;
        double stx = 0, sty = 0, stmin = 0, stmax = 0, width = 0, width1 = 0;
        // 0 evaluate at stp, 1 converged, 2 warning (no further progress)
        int step(double f, double g, double &stp, double stpmax);
    };

    size_t n_;
    int m_;
    LbfgsbOptions opt_;
    std::vector<double> l_, u_;
    std::vector<int> nbd_;
    bool constrained_ = false, boxed_ = false;

    std::vector<double> x_, g_, z_, d_, t_, r_;
    double f_ = 0.0;
    std::vector<int> where_;     // -1 always free, 0 free, 1 at lower, 2 at upper, 3 fixed, -3 free, zero gradient
    std::vector<int> free_, active_;

    // corrections: column j (0 oldest) of S and Y lives in slot (head_ + j) % m
    std::vector<double> S_, Y_;
    std::vector<double> sy_, ss_;   // m x m, logical order, row-major: sy_[i*m+j] = s_i'y_j (i >= j), ss_ symmetric
    // For steps with every variable free: syf_[i*m+j] = s_i'y_j for i <= j
    // (the diagonal as the dot product, not the step's dr) and yy_ = Y'Y,
    // kept as the corrections arrive so the subspace step needs no O(m^2 n)
    // sums (the same sums, in the same order, as the general branch).
    std::vector<double> syf_, yy_;
    std::vector<double> tchol_;     // Cholesky factor (upper, row-major) of T = theta SS + L D^-1 L'
    int col_ = 0, head_ = 0;
    double theta_ = 1.0;
    bool updated_ = false;
    std::vector<double> c_;         // W'(z - x) from the Cauchy step

    Step step_ = Step::Initial;
    int iter_ = 0;
    int nevals_ = 0;
    int stop_ = kLbRunning;
    double sbgnrm_ = 0.0;

    // line search state
    Search ls_;
    double stp_ = 0.0, stpmx_ = 0.0, fold_ = 0.0, gdold_ = 0.0, dtd_ = 0.0;
    int ifun_ = 0;

    double fbest_ = 0.0;
    std::vector<double> xbest_;
};

}  // namespace sitkhip
