"""The bound-constrained L-BFGS optimizer (host code in the HIP library).

The first group restates the reference's own unit tests
(crates/sitk/src/registration/lbfgsb.rs:2018-2174). The second pins the
iterates: scipy.optimize's L-BFGS-B implements the same published algorithm
(Byrd-Lu-Nocedal-Zhu 3.0 with the Moré-Thuente search) that the reference
restates (lbfgsb.rs:1-30), so with the same m, factr, pgtol and the 20-step
search limit both must ask for the same sequence of evaluation points -- the
Cauchy point, the subspace step and its projection, the line search and the
correction updates all show in that sequence. Only rounding may differ, which
an ill-conditioned problem amplifies after tens of iterations.

No device is used: these run in the CPU suite.
"""

import numpy as np
import pytest

import sitk_hip as S

scipy_opt = pytest.importorskip("scipy.optimize")


def quadratic(p):
    return (p[0] - 3.0) ** 2 + (p[1] + 2.0) ** 2, np.array([2.0 * (p[0] - 3.0), 2.0 * (p[1] + 2.0)])


def rosenbrock(p):
    v = (1.0 - p[0]) ** 2 + 100.0 * (p[1] - p[0] ** 2) ** 2
    return v, np.array([-2.0 * (1.0 - p[0]) - 400.0 * p[0] * (p[1] - p[0] ** 2),
                        200.0 * (p[1] - p[0] ** 2)])


# ---------------------------------------------------------- reference tests
def test_unconstrained_quadratic_reaches_the_minimum():
    r = S.LBFGSBOptimizer(200).optimize([0.0, 0.0], quadratic)
    assert abs(r.parameters[0] - 3.0) < 1e-6 and abs(r.parameters[1] + 2.0) < 1e-6
    assert r.value < 1e-10
    assert r.stop_reason == "GradientConverged"


def test_unconstrained_rosenbrock_converges():
    r = S.LBFGSBOptimizer(500).optimize([-1.2, 1.0], rosenbrock)
    assert np.all(np.abs(r.parameters - 1.0) < 1e-4)
    assert r.value < 1e-8


def test_active_lower_bound_pins_the_minimizer():
    o = S.LBFGSBOptimizer(200).set_bound_selection([1, 0]).set_lower_bound([5.0, 0.0]) \
        .set_upper_bound([0.0, 0.0])
    r = o.optimize([10.0, 10.0], quadratic)
    assert abs(r.parameters[0] - 5.0) < 1e-6 and abs(r.parameters[1] + 2.0) < 1e-6
    assert abs(r.value - 4.0) < 1e-8


def test_both_sided_bounds_are_respected_at_the_optimum():
    o = S.LBFGSBOptimizer(200).set_bound_selection([2, 2]).set_lower_bound([-1.0, -1.0]) \
        .set_upper_bound([1.0, 1.0])
    r = o.optimize([0.0, 0.0], quadratic)
    assert abs(r.parameters[0] - 1.0) < 1e-6 and abs(r.parameters[1] + 1.0) < 1e-6
    assert np.all((r.parameters >= -1.0) & (r.parameters <= 1.0))


def test_bounded_rosenbrock_converges_to_the_constrained_corner():
    o = S.LBFGSBOptimizer(500).set_bound_selection([2, 2]).set_lower_bound([-2.0, -2.0]) \
        .set_upper_bound([0.5, 0.5])
    r = o.optimize([-1.0, -1.0], rosenbrock)
    assert abs(r.parameters[0] - 0.5) < 1e-4 and abs(r.parameters[1] - 0.25) < 1e-3
    assert np.all((r.parameters >= -2.0) & (r.parameters <= 0.5))


def test_max_function_evaluations_stops_and_returns_best():
    o = S.LBFGSBOptimizer(10_000).set_max_function_evaluations(3) \
        .set_gradient_convergence_tolerance(0.0).set_cost_function_convergence_factor(0.0)
    seen = []

    def ev(p):
        v, g = rosenbrock(p)
        seen.append(v)
        return v, g

    r = o.optimize([-1.2, 1.0], ev)
    assert r.stop_reason == "MaxFunctionEvaluations"
    assert r.value <= rosenbrock([-1.2, 1.0])[0]
    # the best point ever evaluated, not the last one (the 4th evaluation stops)
    assert len(seen) == 4 and r.value == min(seen)


def test_max_iterations_stops_with_that_reason():
    r = S.LBFGSBOptimizer(2).optimize([-1.2, 1.0], rosenbrock)
    assert r.stop_reason == "MaxIterations" and r.iterations == 2


def test_already_at_the_minimum_converges_without_iterating():
    r = S.LBFGSBOptimizer(100).optimize([3.0, -2.0], quadratic)
    assert r.stop_reason == "GradientConverged" and r.iterations == 0
    assert list(r.parameters) == [3.0, -2.0]


def test_high_dimensional_quadratic_with_limited_memory():
    n = 20
    target = np.arange(n, dtype=np.float64)
    r = S.LBFGSBOptimizer(500).optimize(np.zeros(n), lambda p: (float(np.sum((p - target) ** 2)),
                                                                 2.0 * (p - target)))
    assert np.all(np.abs(r.parameters - target) < 1e-5)
    assert r.value < 1e-8


def test_infeasible_bound_is_refused():
    o = S.LBFGSBOptimizer(10).set_bound_selection([2]).set_lower_bound([5.0]).set_upper_bound([1.0])
    with pytest.raises(S.CudaError, match="infeasible bound"):
        o.optimize([0.0], lambda p: (p[0] ** 2, np.array([2.0 * p[0]])))


def test_invalid_settings_are_refused():
    with pytest.raises(S.CudaError):
        S.LBFGSBOptimizer(10).set_max_corrections(0).optimize([1.0], quadratic)
    with pytest.raises(S.CudaError):
        S.LBFGSBOptimizer(10).set_bound_selection([4]).set_lower_bound([0.0]) \
            .set_upper_bound([1.0]).optimize([1.0], lambda p: (p[0] ** 2, 2 * p))
    with pytest.raises(S.CudaError):
        S.LBFGSBOptimizer(10).set_cost_function_convergence_factor(-1.0).optimize([1.0], quadratic)


def test_start_point_is_projected_onto_the_box():
    o = S.LBFGSBOptimizer(50).set_bound_selection([2, 1]).set_lower_bound([-1.0, 2.0]) \
        .set_upper_bound([1.0, 0.0])
    first = []

    def ev(p):
        if not first:
            first.append(p.copy())
        return quadratic(p)

    o.optimize([5.0, -3.0], ev)
    assert list(first[0]) == [1.0, 2.0]


# ------------------------------------------- iterates against scipy's L-BFGS-B
def _selection(bounds, n):
    sel, lo, hi = [], [], []
    for a, b in bounds:
        sel.append(0 if a is None and b is None else 1 if b is None else 3 if a is None else 2)
        lo.append(0.0 if a is None else a)
        hi.append(0.0 if b is None else b)
    return sel, lo, hi


def _ours(fun, x0, bounds, m):
    pts = []

    def ev(x):
        pts.append(x.copy())
        return fun(x)

    o = S.LBFGSBOptimizer(500).set_max_corrections(m)
    if bounds is not None:
        sel, lo, hi = _selection(bounds, len(x0))
        o.set_bound_selection(sel).set_lower_bound(lo).set_upper_bound(hi)
    r = o.optimize(x0, ev)
    return np.array(pts), r


def _scipy(fun, x0, bounds, m):
    pts = []

    def ev(x):
        pts.append(np.array(x, dtype=np.float64))
        return fun(np.array(x, dtype=np.float64))

    r = scipy_opt.minimize(ev, np.asarray(x0, dtype=np.float64), jac=True, method="L-BFGS-B",
                           bounds=bounds,
                           options=dict(maxcor=m, ftol=1e7 * np.finfo(float).eps, gtol=1e-5,
                                        maxiter=500, maxfun=2000, maxls=20))
    return np.array(pts), r


def _spd_problem(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    h = a @ a.T / n + 0.1 * np.eye(n)
    b = rng.standard_normal(n) * 3.0
    return lambda x: (0.5 * x @ h @ x - b @ x, h @ x - b)


def _styblinski_tang(x):
    return 0.5 * np.sum(x ** 4 - 16.0 * x ** 2 + 5.0 * x), 0.5 * (4.0 * x ** 3 - 32.0 * x + 5.0)


def _extended_rosenbrock(x):
    f = np.sum(100.0 * (x[1:] - x[:-1] ** 2) ** 2 + (1.0 - x[:-1]) ** 2)
    g = np.zeros_like(x)
    g[:-1] += -400.0 * x[:-1] * (x[1:] - x[:-1] ** 2) - 2.0 * (1.0 - x[:-1])
    g[1:] += 200.0 * (x[1:] - x[:-1] ** 2)
    return f, g


_MIXED = [(-0.5, 0.5) if i % 3 == 0 else (-1.0, None) if i % 3 == 1 else (None, 0.7)
          for i in range(60)]

CASES = {
    # name: (objective, start, bounds, m)
    "rosenbrock": (rosenbrock, [-1.2, 1.0], None, 5),
    "rosenbrock_box": (rosenbrock, [-1.0, -1.0], [(-2.0, 0.5), (-2.0, 0.5)], 5),
    "quadratic_lower": (quadratic, [10.0, 10.0], [(5.0, None), (None, None)], 5),
    "spd60_mixed_bounds_m5": (_spd_problem(60, 3), np.zeros(60), _MIXED, 5),
    "spd60_mixed_bounds_m10": (_spd_problem(60, 4), np.random.default_rng(5).standard_normal(60),
                               _MIXED, 10),
    "styblinski_tang_box": (_styblinski_tang, np.random.default_rng(6).uniform(-4, 4, 30),
                            [(-3.0, 3.0)] * 30, 5),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_evaluation_points_are_scipys(name):
    fun, x0, bounds, m = CASES[name]
    a, ra = _ours(fun, x0, bounds, m)
    b, rb = _scipy(fun, x0, bounds, m)
    assert a.shape == b.shape, (len(a), len(b))
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-8)
    assert ra.iterations == rb.nit
    assert abs(ra.value - rb.fun) <= 1e-12 * max(1.0, abs(rb.fun))


def test_ill_conditioned_problem_follows_scipy_until_rounding_takes_over():
    x0 = -1.2 * np.ones(20)
    bounds = [(-1.5, 0.8)] * 20
    a, ra = _ours(_extended_rosenbrock, x0, bounds, 5)
    b, rb = _scipy(_extended_rosenbrock, x0, bounds, 5)
    np.testing.assert_allclose(a[:40], b[:40], rtol=0, atol=1e-8)
    assert abs(ra.value - rb.fun) <= 1e-9 * max(1.0, abs(rb.fun))
