"""The oracle's cubic B-spline interpolator of the moving image
(crates/sitk/src/transform/interpolator.rs:170-350: the mirror-boundary
prefilter and the 64-tap value and gradient), pinned by the reference's own
tests (interpolator.rs:832-1093, restated for volumes: a 2-D image is a volume
of one slice, a 1-D line one of one row and one slice), by an independent
dense solve of the interpolation condition, and by finite differences.
CPU only."""

import math

import numpy as np
import pytest

import oracle as O


def _vol(buf2d):
    """(h, w) image -> (1, h, w) float32 volume."""
    return np.ascontiguousarray(np.asarray(buf2d, dtype=np.float32)[None])


def _at(coef, x, y=0.0, z=0.0):
    r = O.bspline_value_and_gradient(coef, [x, y, z])
    assert r is not None, (x, y, z)
    return r


def _ramp(w, h):
    y, x = np.mgrid[0:h, 0:w]
    return 3.0 * x + 5.0 * y


def _noisy(w, h):
    """The reference's xorshift image (interpolator.rs:871-883), narrowed to
    float32 (the oracle takes float volumes, as the device does)."""
    state = 0x2545F4914F6CDD1D
    m = (1 << 64) - 1
    out = np.empty(w * h)
    for i in range(w * h):
        state ^= (state << 13) & m
        state ^= state >> 7
        state ^= (state << 17) & m
        out[i] = (state >> 11) / float(1 << 53) * 200.0 - 100.0
    return out.reshape(h, w).astype(np.float32)


def test_ramp_is_reproduced_away_from_the_boundary():
    # interpolator.rs:832-851
    c = O.bspline_coefficients(_vol(_ramp(40, 40)))
    v, _ = _at(c, 20.4, 19.6)
    assert abs(v - (3 * 20.4 + 5 * 19.6)) < 1e-9


def test_samples_at_integer_indices():
    # interpolator.rs:853-869
    buf = _ramp(10, 10)
    c = O.bspline_coefficients(_vol(buf))
    for y in range(2, 8):
        for x in range(2, 8):
            assert abs(_at(c, x, y)[0] - buf[y, x]) < 1e-9


def test_constant_image_has_constant_coefficients():
    # interpolator.rs:886-911
    c = O.bspline_coefficients(_vol(np.full((5, 7), 7.25)))
    assert np.all(np.abs(c - 7.25) < 1e-12)
    for x, y in ((-0.49, -0.49), (0.0, 0.0), (3.7, 2.2), (6.49, 4.49)):
        v, g = _at(c, x, y)
        assert abs(v - 7.25) < 1e-12 and np.all(np.abs(g) < 1e-12), (x, y, v, g)


def test_every_sample_of_a_noisy_image_including_the_edges():
    # interpolator.rs:913-934
    buf = _noisy(9, 7)
    c = O.bspline_coefficients(_vol(buf))
    for y in range(7):
        for x in range(9):
            assert abs(_at(c, x, y)[0] - float(buf[y, x])) < 1e-11, (x, y)


def test_ramp_bends_off_lattice_at_the_boundary():
    # interpolator.rs:936-975
    n = 41
    buf = (3.0 * np.arange(n) + 1.0).reshape(1, n)
    c = O.bspline_coefficients(_vol(buf))
    for i in range(n):
        assert abs(_at(c, i)[0] - (3.0 * i + 1.0)) < 1e-11
    for x in (19.5, 20.25, 20.5):
        assert abs(_at(c, x)[0] - (3.0 * x + 1.0)) < 1e-8
    for x in (0.5, n - 1 - 0.5):
        assert abs(_at(c, x)[0] - (3.0 * x + 1.0)) > 0.4


def test_index_folding_mirrors_about_the_edge_samples():
    # interpolator.rs:977-1010
    c = O.bspline_coefficients(_vol(_noisy(11, 1)))
    last = 10.0
    for t in (0.1, 0.25, 0.49):
        assert abs(_at(c, -t)[0] - _at(c, t)[0]) < 1e-12
        assert abs(_at(c, last + t)[0] - _at(c, last - t)[0]) < 1e-12


def test_a_size_one_axis_folds_onto_index_zero():
    # interpolator.rs:1012-1036: a line held as a volume of one row and one
    # slice interpolates like the row alone (checked against the 1-D spline
    # evaluated here from its dense-solved coefficients)
    row = _noisy(6, 1)
    c = O.bspline_coefficients(_vol(row))
    coef1 = _dense_coefficients_1d(row[0].astype(np.float64))
    for x in (0.0, 1.3, 2.5, 5.0):
        assert abs(_at(c, x)[0] - _eval_1d(coef1, x)) < 1e-12, x


def test_weights_at_a_half_index_are_the_cubic_taps():
    # interpolator.rs:1038-1050: coefficients fed directly
    coeffs = np.array([2.0, -1.0, 5.0, 3.0, -4.0, 0.5]).reshape(1, 1, 6)
    v, _ = _at(coeffs, 2.5)
    want = (coeffs[0, 0, 1] + 23 * coeffs[0, 0, 2] + 23 * coeffs[0, 0, 3] + coeffs[0, 0, 4]) / 48
    assert abs(v - want) < 1e-14


def test_gradient_matches_finite_difference_on_the_ramp():
    # interpolator.rs:1052-1093
    c = O.bspline_coefficients(_vol(_ramp(12, 12)))
    c0 = [5.4, 6.7]
    _, g = _at(c, *c0)
    h = 1e-4
    for k in range(2):
        cp, cm = list(c0), list(c0)
        cp[k] += h
        cm[k] -= h
        fd = (_at(c, *cp)[0] - _at(c, *cm)[0]) / (2 * h)
        assert abs(fd - g[k]) < 1e-6, (k, fd, g[k])


# ------------------------------------------- independent pins of the prefilter
def _dense_coefficients_1d(f):
    """Solve (c[k-1] + 4 c[k] + c[k+1]) / 6 = f[k] with the mirror extension
    c[-1] = c[1], c[n] = c[n-2] directly (numpy), not by the recursion."""
    n = f.size
    if n == 1:
        return f.copy()
    A = np.zeros((n, n))
    for k in range(n):
        for j, w in ((k - 1, 1.0), (k, 4.0), (k + 1, 1.0)):
            if j < 0:
                j = -j
            if j > n - 1:
                j = 2 * (n - 1) - j
            A[k, j] += w / 6.0
    return np.linalg.solve(A, f)


def _eval_1d(coef, x):
    """The cubic B-spline sum at x with mirror-folded taps, from the textbook
    basis (not the reference's weight formulas)."""
    n = coef.size
    base = math.floor(x) - 1
    v = 0.0
    for k in range(4):
        i = base + k
        j = 0 if n == 1 else (-i if i < 0 else i)
        if n > 1 and j > n - 1:
            j = 2 * (n - 1) - j
        v += O.cubic_bspline(x - i) * coef[j]
    return v


@pytest.mark.parametrize("shape", [(1, 1, 9), (1, 1, 2), (1, 1, 3), (5, 4, 7), (3, 1, 6)])
def test_prefilter_is_the_dense_solve(shape):
    rng = np.random.default_rng(sum(shape))
    v = rng.uniform(-50, 150, shape).astype(np.float32)
    got = O.bspline_coefficients(v)
    want = v.astype(np.float64)
    for axis in (2, 1, 0):  # x, y, z
        if want.shape[axis] > 1:
            want = np.apply_along_axis(_dense_coefficients_1d, axis, want)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-10), np.max(np.abs(got - want))


def test_off_lattice_values_equal_the_textbook_sum():
    rng = np.random.default_rng(4)
    v = rng.uniform(0, 100, (1, 1, 17)).astype(np.float32)
    c = O.bspline_coefficients(v)
    coef1 = _dense_coefficients_1d(v[0, 0].astype(np.float64))
    for x in rng.uniform(-0.5, 16.49, 25):
        assert abs(_at(c, x)[0] - _eval_1d(coef1, x)) < 1e-10, x


def test_powi_is_binary_exponentiation():
    # exact cases first
    assert O.powi(0.5, 10) == 2.0 ** -10
    assert O.powi(-2.0, 5) == -32.0
    assert O.powi(3.7, 0) == 1.0
    assert O.powi(2.0, -3) == 0.125
    # against the exact rational power: binary exponentiation rounds once per
    # multiply, and each squaring doubles the error it carries, so the bound
    # grows with k (k = 1, 2 are exact: one rounding of an exact product)
    from fractions import Fraction

    z = math.sqrt(3.0) - 2.0
    assert O.powi(z, 1) == z and O.powi(z, 2) == z * z
    for k in (3, 7, 31, 63, 127, 255):
        exact = Fraction(z) ** k
        err = abs(Fraction(O.powi(z, k)) - exact) / abs(exact)
        assert err <= Fraction(k, 2 ** 52), (k, float(err))


def test_mattes_derivative_with_bspline_interpolation_matches_finite_difference():
    """The Mattes derivative through the B-spline interpolant's gradient
    (metric.rs:848-864 with interpolator.rs:284-350) against central
    differences of the value."""
    sp = [0.8, 0.9, 1.1]
    org = np.array([-3.0, 2.0, 1.0])
    m = O.fixture_inverted(32, [1.5, -0.5, 0.75])
    f = np.ascontiguousarray(O.fixture_volume(32, [0, 0, 0])[8:24, 8:24, 8:24])
    forg = org + 8 * np.array(sp)
    pb = O.MetricProblem(f, sp, forg, m, sp, org, interpolator="bspline")
    g = pb.mattes_geometry(50)
    ctr = forg + 7.5 * np.array(sp)
    p0 = [0.02, -0.01, 0.03, 0.6, -0.4, 0.2]
    _, an, _ = pb.mattes_evaluate(g, O.make_transform(O.EULER3D, p0, center=ctr))
    for k in range(6):
        h = 1e-5
        pp, pm = list(p0), list(p0)
        pp[k] += h
        pm[k] -= h
        fp = pb.mattes_value(g, O.make_transform(O.EULER3D, pp, center=ctr))
        fm = pb.mattes_value(g, O.make_transform(O.EULER3D, pm, center=ctr))
        fd = (fp - fm) / (2 * h)
        assert abs(fd - an[k]) < 5e-3 * max(1.0, abs(an[k])), (k, fd, an[k])
