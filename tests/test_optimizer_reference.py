"""The host restatements of the reference's optimizers (tests/
oracle/registration.py), which the device registration driver is checked
against iterate for iterate, pinned here by the reference's own optimizer
fixtures (crates/sitk/src/registration/optimizer.rs:1063-1296) and by the
window convergence monitor's published stop count.
"""

import math

import pytest

from oracle.registration import _kahan, gd_line_search, rsgd


def _bowl(p):
    # f(p) = (p0 - 3)^2 + (p1 + 2)^2, minimum at (3, -2)
    return ((p[0] - 3.0) ** 2 + (p[1] + 2.0) ** 2, [2.0 * (p[0] - 3.0), 2.0 * (p[1] + 2.0)])


def _aniso(p):
    return ((p[0] - 1.0) ** 2 + 1e6 * (p[1] - 1.0) ** 2,
            [2.0 * (p[0] - 1.0), 2e6 * (p[1] - 1.0)])


def test_the_reversal_test_is_compensated_and_a_naive_walk_flips_the_branch():
    """optimizer.rs:1063-1123: the products [1e16, 1, 1, -1e16, -1] sum to -1
    naively and +1 with compensation; the compensated walk does not relax, so
    parameter 1 moves by g1[1]/|g1| + g2[1]/|g2| (a naive walk would halve the
    second step)."""
    g1 = [1.0e16, 1.0, 1.0, -1.0e16, -1.0]
    g2 = [1.0] * 5
    naive = 0.0
    for a, b in zip(g1, g2):
        naive += a * b
    assert naive < 0.0 and _kahan([a * b for a, b in zip(g1, g2)]) > 0.0
    calls = [0]

    def ev(_p):
        calls[0] += 1
        return 0.0, list(g1 if calls[0] == 1 else g2)

    p, _, taken, _ = rsgd(ev, [0.0] * 5, lambda _s: 1.0, 1e-12, 2, 1e-4, relaxation=0.5)
    m1 = math.sqrt(_kahan([g * g for g in g1]))
    m2 = math.sqrt(_kahan([g * g for g in g2]))
    unrelaxed = -(g1[1] / m1) - (g2[1] / m2)
    relaxed = -(g1[1] / m1) - 0.5 * (g2[1] / m2)
    assert abs(unrelaxed - relaxed) > 0.2
    assert taken == 2
    assert abs(p[1] - unrelaxed) < 1e-12 * max(abs(unrelaxed), 1.0), (p[1], unrelaxed, relaxed)


def test_regular_step_minimizes_a_quadratic_bowl():
    # optimizer.rs:1178-1195 (learning rate 1, minimum step 1e-8, tolerance 1e-4)
    p, _, taken, stop = rsgd(_bowl, [0.0, 0.0], lambda _s: 1.0, 1e-8, 1000, 1e-4)
    assert abs(p[0] - 3.0) < 5e-5 and abs(p[1] + 2.0) < 5e-5, p
    assert stop == "GradientConverged" and taken < 1000


def test_regular_step_relaxes_on_overshoot_from_a_large_step():
    # optimizer.rs:1197-1213 (step 50, minimum step 1e-6, tolerance 1e-12)
    p, _, _, stop = rsgd(lambda q: ((q[0] - 3.0) ** 2, [2.0 * (q[0] - 3.0)]), [0.0],
                         lambda _s: 50.0, 1e-6, 1000, 1e-12)
    assert abs(p[0] - 3.0) < 1e-4, p
    assert stop == "StepTooSmall"


def test_regular_step_stops_immediately_at_a_stationary_start():
    # optimizer.rs:1215-1229
    p, _, taken, stop = rsgd(lambda q: (0.0, _bowl(q)[1]), [3.0, -2.0], lambda _s: 1.0, 1e-6,
                             1000, 1e-4)
    assert stop == "GradientConverged" and taken == 0 and p == [3.0, -2.0]


def test_regular_step_scales_balance_anisotropic_conditioning():
    # optimizer.rs:1231-1245 (scales [1, 1e6], minimum step 1e-7, 2000 iterations)
    p, _, _, _ = rsgd(_aniso, [0.0, 0.0], lambda _s: 1.0, 1e-7, 2000, 1e-4, scales=[1.0, 1e6])
    assert abs(p[0] - 1.0) < 1e-3 and abs(p[1] - 1.0) < 1e-3, p


def _ls(f, p0, lr, iterations, **kw):
    return gd_line_search(f, lambda q: f(q)[0], p0, lr, iterations, **kw)


def test_line_search_minimizes_a_quadratic_bowl():
    # optimizer.rs:1247-1262
    p, v, _, _ = _ls(_bowl, [0.0, 0.0], 1.0, 100)
    assert abs(p[0] - 3.0) < 1e-4 and abs(p[1] + 2.0) < 1e-4, p
    assert v < 1e-8


def test_line_search_tames_an_overshooting_base_rate():
    # optimizer.rs:1264-1278: a fixed rate of 1 oscillates 0 -> 6 -> 0; the
    # search finds the quadratic's optimal 0.5
    p, v, _, _ = _ls(lambda q: ((q[0] - 3.0) ** 2, [2.0 * (q[0] - 3.0)]), [0.0], 1.0, 50)
    assert abs(p[0] - 3.0) < 1e-4 and v < 1e-8, p


def test_line_search_scales_balance_anisotropic_conditioning():
    # optimizer.rs:1280-1294
    p, _, _, _ = _ls(_aniso, [0.0, 0.0], 1.0, 500, scales=[1.0, 1e6])
    assert abs(p[0] - 1.0) < 1e-3 and abs(p[1] - 1.0) < 1e-3, p


def test_line_search_stops_on_convergence_monitor():
    # optimizer.rs:1296-1314 (window 5, minimum convergence value 1e-8)
    p, _, taken, stop = _ls(_bowl, [0.0, 0.0], 1.0, 100_000, window=5, min_convergence=1e-8)
    assert stop == "Converged" and taken < 100_000
    assert abs(p[0] - 3.0) < 1e-3 and abs(p[1] + 2.0) < 1e-3, p


@pytest.mark.parametrize("window,minimum", [(10, 1e-6), (5, 1e-8)])
def test_line_search_monitor_window_fills_before_it_can_stop(window, minimum):
    """The window monitor (registration/convergence.rs:35-95) reports nothing
    until `window` values are in, so no run stops before that many iterations."""
    _, _, taken, stop = _ls(_bowl, [0.0, 0.0], 1.0, 100_000, window=window,
                            min_convergence=minimum)
    assert stop == "Converged" and taken >= window - 1
