"""Pins the CPU oracle against the golden vectors the reference itself holds,
and against what the mathematics fixes where no vector exists.

Golden vectors (paths relative to physwkim/sitk-rs):
  - synth seed 42, first five voxels: crates/sitk/tests/bench_correctness_gate.rs:162-168
  - FNV-1a 'a' test vector: same file :170-175
  - input checksums 64^3 / 256^3 / 512^3: same file :192-211
  - smoothing_recursive_gaussian([2,2,2]) output checksums:
      64^3  0x8f3da2e917366232  bench/results/rust-cuda.ndjson:5
      256^3 0x9aaea53a51fda207  bench/results/rust-cuda.ndjson:53,
                                crates/sitk/tests/bench_correctness_gate.rs:60-63
      512^3 0xb60665c90074c378  bench/results/rust-cuda.ndjson:101
      64^3 seed 0x2545F4914F6CDD1D -> 0x9e5067c73b94ee1e  crates/sitk/tests/bit_parity.rs:151-160
  - Mattes value -1.063357897140305 (doc/upstream-findings.md:273) on the fixture
    of crates/sitk/tests/mattes_device.rs:42-81 (volume(32,[0;3]) vs
    inverted(32,[1.5,-0.5,0.75]), spacing [0.8,0.9,1.1], origin [-3,2,1], 50 bins,
    translation [1,0,0]).
"""

import math

import numpy as np
import pytest

import oracle as O

SEED = 42


def test_synth_first_five_values():
    assert list(O.synth(SEED, 5)) == [59.0, 641.0, 384.0, 121.0, 923.0]


def test_fnv1a64_test_vector():
    assert O.fnv1a64(np.frombuffer(b"a", dtype=np.uint8)) == 0xAF63DC4C8601EC8C


@pytest.mark.parametrize("dim,want", [(64, 0xA60A081F21AF857E), (256, 0xB04930CDA0BBCE53)])
def test_input_checksums(dim, want):
    assert O.fnv1a64(O.synth(SEED, dim ** 3)) == want


@pytest.mark.parametrize(
    "dim,seed,want",
    [
        (64, SEED, 0x8F3DA2E917366232),
        (64, 0x2545F4914F6CDD1D, 0x9E5067C73B94EE1E),
        (256, SEED, 0x9AAEA53A51FDA207),
    ],
)
def test_recursive_gaussian_golden_checksums(dim, seed, want):
    v = O.synth(seed, dim ** 3)
    out = O.smoothing_recursive_gaussian(v, (dim,) * 3, (1, 1, 1), (2, 2, 2), threads=8)
    assert O.fnv1a64(out) == want


@pytest.mark.slow
def test_recursive_gaussian_golden_checksum_512():
    v = O.synth(SEED, 512 ** 3)
    assert O.fnv1a64(v) == 0xFBF1951B8B4B69AA
    out = O.smoothing_recursive_gaussian(v, (512,) * 3, (1, 1, 1), (2, 2, 2), threads=8)
    assert O.fnv1a64(out) == 0xB60665C90074C378


def test_recursive_gaussian_thread_count_invariance():
    v = O.synth(7, 40 * 33 * 21)
    outs = [O.fnv1a64(O.smoothing_recursive_gaussian(v, (40, 33, 21), (0.9, 1.1, 1.5),
                                                    (1.3, 2.0, 0.7), threads=t))
            for t in (1, 3, 8)]
    assert len(set(outs)) == 1


def test_deriche_zero_order_dc_gain_is_one():
    # crates/sitk/src/core/deriche.rs:363-373: the zero-order filter has unity DC gain.
    for sigmad in (0.5, 1.0, 2.0, 3.7, 8.0):
        c = O.deriche_coefficients(sigmad)
        n, d, m = c[0:4], c[4:8], c[8:12]
        gain = (n.sum() + m.sum()) / (1.0 + d.sum())
        assert abs(gain - 1.0) < 1e-12


def test_constant_is_preserved_and_zero_sigma_is_identity():
    # recursive_gaussian.rs:545-572
    shape = (12, 10, 9)
    const = np.full(int(np.prod(shape)), 5.0, dtype=np.float32)
    out = O.smoothing_recursive_gaussian(const, shape[::-1], (1, 1, 1), (2, 2, 2))
    assert np.max(np.abs(out - 5.0)) < 1e-5
    v = O.synth(3, int(np.prod(shape)))
    out = O.smoothing_recursive_gaussian(v, shape[::-1], (1, 1, 1), (0, 0, 0))
    assert np.array_equal(out, v)


def test_axis_pass_is_filter_line_on_every_line():
    """Smoothing only axis 1 equals filter_line applied to each y-line: pins the
    line decomposition independently of the golden checksums."""
    nx, ny, nz = 5, 11, 4
    v = O.synth(11, nx * ny * nz)
    out = O.smoothing_recursive_gaussian(v, (nx, ny, nz), (1, 1, 1), (0, 1.7, 0))
    coef = O.deriche_coefficients(1.7)
    vol = v.reshape(nz, ny, nx).astype(np.float64)
    want = np.empty_like(vol)
    for z in range(nz):
        for x in range(nx):
            want[z, :, x] = O.filter_line(vol[z, :, x], coef)
    assert np.array_equal(out.reshape(nz, ny, nx), want.astype(np.float32))


def test_filter_line_is_a_linear_shift_invariant_filter_in_the_interior():
    """Far from the borders the IIR response to a unit impulse is a symmetric,
    positive, unit-mass Gaussian-like kernel."""
    coef = O.deriche_coefficients(3.0)
    x = np.zeros(201)
    x[100] = 1.0
    y = O.filter_line(x, coef)
    assert abs(y.sum() - 1.0) < 1e-9
    assert np.allclose(y[100 - 30:100], y[101:131][::-1], atol=1e-9)
    var = np.sum(y * (np.arange(201) - 100.0) ** 2)
    assert abs(math.sqrt(var) - 3.0) < 0.15


def test_axis_shorter_than_four_is_refused():
    v = O.synth(1, 3 * 8 * 8)
    with pytest.raises(ValueError):
        O.smoothing_recursive_gaussian(v, (3, 8, 8), (1, 1, 1), (1, 1, 1))
    # but not when that axis is not smoothed
    O.smoothing_recursive_gaussian(v, (3, 8, 8), (1, 1, 1), (0, 1, 1))


# ----------------------------------------------------------------- Mattes
def _ledger_problem():
    f = O.fixture_volume(32, [0, 0, 0])
    m = O.fixture_inverted(32, [1.5, -0.5, 0.75])
    return O.MetricProblem(f, [0.8, 0.9, 1.1], [-3, 2, 1], m, [0.8, 0.9, 1.1], [-3, 2, 1])


def test_mattes_ledger_value_is_reproduced_bit_for_bit():
    """The reference-held host value, recorded before the reference made its
    joint-PDF mass a Kahan sum (ledger row 2.161), so it is reproduced with the
    naive sum of that time; the current Kahan tail is 3 ulp away."""
    pb = _ledger_problem()
    g = pb.mattes_geometry(50)
    t = O.make_transform(O.TRANSLATION, [1.0, 0.0, 0.0])
    try:
        O.set_joint_sum_compensated(False)
        v = pb.mattes_value(g, t)
    finally:
        O.set_joint_sum_compensated(True)
    assert v == -1.063357897140305
    vk = pb.mattes_value(g, t)
    assert vk == -1.0633578971403057
    assert abs(vk - v) <= 4 * np.spacing(abs(v))


def test_mattes_matches_an_independent_numpy_restatement():
    """Brute force on a small pair with numpy (no shared code with the oracle)."""
    pb = _ledger_problem()
    g = pb.mattes_geometry(50)
    t = O.make_transform(O.TRANSLATION, [0.0, 0.0, 0.0])
    f = pb.fixed.astype(np.float64).ravel()
    m = pb.moving.astype(np.float64).ravel()  # identity on identical grids: mv == voxel
    B = 50
    fi = np.clip((f / g.fixed_bin_size - g.fixed_normalized_min).astype(np.int64), 2, B - 3)
    mt = m / g.moving_bin_size - g.moving_normalized_min
    mi = np.clip(mt.astype(np.int64), 2, B - 3)

    def b3(u):
        a = np.abs(u)
        return np.where(a < 1, (4 - 6 * a * a + 3 * a ** 3) / 6,
                        np.where(a < 2, (2 - a) ** 3 / 6, 0.0))

    J = np.zeros((B, B))
    for o in range(-1, 3):
        np.add.at(J, (fi, mi + o), b3(mi + o - mt))
    P = J / J.sum()
    pf, pm = P.sum(1), P.sum(0)
    nz = P > 0
    mi_val = (P[nz] * np.log(P[nz] / np.outer(pf, pm)[nz])).sum()
    assert abs(pb.mattes_value(g, t) - (-mi_val)) < 1e-12
    joint, marg, valid = pb.mattes_histogram(g, t)
    assert valid == 32 ** 3
    assert np.allclose(joint.reshape(B, B), J, rtol=1e-12, atol=1e-12)
    assert np.array_equal(marg, np.bincount(fi, minlength=B).astype(np.float64))


def test_cubic_bspline_partition_of_unity_and_derivative():
    # mattes.rs:821-846
    for i in range(20):
        frac = i / 20.0
        s = sum(O.cubic_bspline(k - frac) for k in (-1, 0, 1, 2))
        assert abs(s - 1.0) < 1e-12
    h = 1e-6
    for i in range(-40, 40):
        u = i / 10.0
        if abs(u - round(u)) < 1e-9:
            continue
        fd = (O.cubic_bspline(u + h) - O.cubic_bspline(u - h)) / (2 * h)
        assert abs(fd - O.cubic_bspline_derivative(u)) < 1e-4


def test_kahan_sum_is_itk_without_the_final_fold():
    # mattes.rs:855-903: the compensated sum differs from the naive one and is closer.
    h = np.array([(613.7 if i % 97 == 0 else 0.1) * (1.0 + (i % 13) / 7.0) for i in range(2500)])
    naive = -0.0
    for t in h:
        naive += t
    s, c = 0.0, 0.0
    for t in h:
        x = s + t
        c += (s - x) + t if abs(s) >= abs(t) else (t - x) + s
        s = x
    ref = s + c
    k = O.compensated_sum(h)
    assert naive != ref and k != naive and abs(k - ref) < abs(naive - ref)


def _interior_problem():
    """A 16^3 fixed block sitting in the middle of the 32^3 moving fixture, so
    small rotations never move a sample across the moving buffer's edge: the
    analytic derivative has no term for samples entering or leaving the domain,
    so a finite difference only agrees with it when none does."""
    sp = [0.8, 0.9, 1.1]
    org = np.array([-3.0, 2.0, 1.0])
    m = O.fixture_inverted(32, [1.5, -0.5, 0.75])
    f = np.ascontiguousarray(O.fixture_volume(32, [0, 0, 0])[8:24, 8:24, 8:24])
    forg = org + 8 * np.array(sp)
    return O.MetricProblem(f, sp, forg, m, sp, org), forg + 7.5 * np.array(sp)


def test_mattes_derivative_matches_finite_difference():
    # mattes.rs:1015-1041 (3-D version, Euler3D and translation)
    pb, ctr = _interior_problem()
    g = pb.mattes_geometry(50)
    for kind, p0, steps in [
        (O.TRANSLATION, [1.3, -0.7, 0.4], [1e-3] * 3),
        (O.EULER3D, [0.03, -0.02, 0.05, 1.1, -0.6, 0.3], [1e-3] * 6),
    ]:
        t = O.make_transform(kind, p0, center=ctr)
        _, an, _ = pb.mattes_evaluate(g, t)
        for k in range(len(p0)):
            pp = list(p0)
            pm = list(p0)
            pp[k] += steps[k]
            pm[k] -= steps[k]
            fd = (pb.mattes_value(g, O.make_transform(kind, pp, center=ctr)) -
                  pb.mattes_value(g, O.make_transform(kind, pm, center=ctr))) / (2 * steps[k])
            assert abs(fd - an[k]) < 5e-3 * max(1.0, abs(an[k])), (kind, k, fd, an[k])


def test_mattes_moment_form_equals_dense_form():
    """The 12-moment derivative (src/cuda/ops/mattes.rs:75-91) contracted with the
    affine decomposition equals the dense global-support derivative; the
    reference pins the two at 1e-9 (tests/mattes_device.rs:312-350)."""
    from helpers import affine_form, rel_vec

    pb = _ledger_problem()
    g = pb.mattes_geometry(50)
    for kind, p in [(O.EULER3D, [0.05, -0.03, 0.08, 1.2, -0.7, 0.4]),
                    (O.AFFINE, [1.02, 0.03, -0.01, -0.02, 0.99, 0.04, 0.01, -0.03, 1.01,
                                0.6, -1.1, 0.4])]:
        t = O.make_transform(kind, p, center=[1.0, -2.0, 0.5])
        v, dense, valid = pb.mattes_evaluate(g, t)
        joint, marg, valid2 = pb.mattes_histogram(g, t)
        v2, pratio = pb.mattes_tail(joint, marg, valid2, g)
        assert v == v2 and valid == valid2
        mom = pb.mattes_moments(g, t, pratio)
        j0, c = affine_form(t)
        deriv = np.zeros(t.nparams)
        for k in range(t.nparams):
            acc = 0.0
            for d in range(3):
                acc += j0[d, k] * mom[d]
                for e in range(3):
                    acc += c[e, d, k] * mom[3 + 3 * d + e]
            deriv[k] = acc
        assert rel_vec(deriv, dense) < 1e-9


def test_mattes_refusals():
    pb = _ledger_problem()
    with pytest.raises(ValueError, match="too few"):
        pb.mattes_geometry(4)
    pb.mattes_geometry(5)
    flat = np.full((4, 4, 4), 3.0, dtype=np.float32)
    pc = O.MetricProblem(flat, [1, 1, 1], [0, 0, 0], pb.moving[:4, :4, :4], [1, 1, 1], [0, 0, 0])
    with pytest.raises(ValueError, match="constant fixed"):
        pc.mattes_geometry(50)


# ----------------------------------------------------------------- mean squares
def test_mean_squares_identity_is_zero_with_zero_gradient():
    # metric.rs:1499-1509
    f = O.fixture_volume(16, [0, 0, 0])
    pb = O.MetricProblem(f, [1, 1, 1], [0, 0, 0], f, [1, 1, 1], [0, 0, 0])
    v, d, valid = pb.ms_evaluate(O.make_transform(O.TRANSLATION, [0, 0, 0]))
    assert v == 0.0 and np.all(d == 0.0) and valid == 16 ** 3


def test_mean_squares_thread_count_does_not_change_a_bit():
    """metric.rs:1370: rows computed in parallel and added in order give the
    serial loop's bits at any thread count (chunks of 2^16 rows: 40^3 samples
    span two chunks)."""
    sp = [0.8, 0.9, 1.1]
    f = O.fixture_volume(40, [0, 0, 0])
    m = O.fixture_inverted(40, [1.5, -0.5, 0.75])
    pb = O.MetricProblem(f, sp, [-3, 2, 1], m, sp, [-3, 2, 1])
    ctr = [13.0, 20.0, 22.0]
    t = O.make_transform(O.AFFINE, [1.01, 0.02, -0.01, -0.015, 0.99, 0.03, 0.01, -0.02, 1.0,
                                    0.7, -0.3, 0.4], center=ctr)
    v1, d1, n1 = pb.ms_evaluate(t)
    for threads in (2, 3, 8):
        vt, dt, nt = pb.ms_evaluate(t, threads=threads)
        assert nt == n1 and vt == v1 and np.array_equal(dt, d1), threads


def test_mean_squares_one_voxel_shift_is_hand_computable():
    """Moving = fixed shifted by one voxel along x; a translation of exactly one
    voxel re-aligns every interior sample, and the value is the mean of the
    squared differences of the clamped border column only."""
    n = 8
    f = (np.arange(n ** 3, dtype=np.float32) % 17).reshape(n, n, n)
    m = np.empty_like(f)
    m[:, :, 1:] = f[:, :, :-1]
    m[:, :, 0] = f[:, :, 0]
    pb = O.MetricProblem(f, [1, 1, 1], [0, 0, 0], m, [1, 1, 1], [0, 0, 0])
    v, valid = pb.ms_value(O.make_transform(O.TRANSLATION, [1.0, 0, 0]))
    # samples with x = n-1 map to x = n (outside); the others read m[x+1] = f[x]
    assert valid == n * n * (n - 1)
    assert v == 0.0


def test_mean_squares_derivative_matches_finite_difference():
    pb, ctr = _interior_problem()
    p0 = [0.02, -0.01, 0.03, 0.6, -0.4, 0.2]
    t = O.make_transform(O.EULER3D, p0, center=ctr)
    _, an, _ = pb.ms_evaluate(t)
    for k in range(6):
        # the interpolant is piecewise trilinear: steps small enough that no
        # sample crosses a cell wall make the difference exact to rounding
        h = 1e-6
        pp = list(p0)
        pm = list(p0)
        pp[k] += h
        pm[k] -= h
        fp = pb.ms_value(O.make_transform(O.EULER3D, pp, center=ctr))[0]
        fm = pb.ms_value(O.make_transform(O.EULER3D, pm, center=ctr))[0]
        fd = (fp - fm) / (2 * h)
        assert abs(fd - an[k]) < 1e-4 * max(1.0, abs(an[k])), (k, fd, an[k])


# ----------------------------------------------------------------- transforms
def test_stage_replay_is_transform_point_bitwise():
    # src/transform/matrix_offset.rs:102-109; tests/cuda_point_map.rs:84
    rng = np.random.default_rng(0)
    for kind, p in [(O.TRANSLATION, [0.7, -1.3, 2.1]),
                    (O.EULER3D, [0.05, -0.03, 0.08, 1.2, -0.7, 0.4]),
                    (O.AFFINE, [1.02, 0.03, -0.01, -0.02, 0.99, 0.04, 0.01, -0.03, 1.01,
                                0.6, -1.1, 0.4])]:
        t = O.make_transform(kind, p, center=[1.0, -2.0, 0.5])
        st = O.transform_stage(t)
        for _ in range(200):
            x = rng.uniform(-100, 100, 3)
            assert np.array_equal(O.replay_stages(st, x), O.transform_point(t, x))


def test_euler3d_jacobian_matches_finite_difference():
    for zyx in (False, True):
        p0 = np.array([0.3, -0.2, 0.5, 1.0, 2.0, 3.0])
        c = [1.0, -2.0, 0.5]
        x = np.array([4.0, -3.0, 7.0])
        J = O.transform_jacobian(O.make_transform(O.EULER3D, p0, c, zyx), x)
        for k in range(6):
            h = 1e-6
            pp, pm = p0.copy(), p0.copy()
            pp[k] += h
            pm[k] -= h
            fd = (O.transform_point(O.make_transform(O.EULER3D, pp, c, zyx), x) -
                  O.transform_point(O.make_transform(O.EULER3D, pm, c, zyx), x)) / (2 * h)
            assert np.allclose(fd, J[:, k], atol=1e-6)


def test_draw_samples():
    # metric.rs:419-437
    assert O.draw_samples(100, 0, 1.0, 0) is None
    assert list(O.draw_samples(10, 1, 0.3, 0)) == [0, 4, 8]
    a = O.draw_samples(1000, 2, 0.2, 5)
    b = O.draw_samples(1000, 2, 0.2, 5)
    assert a.size == 200 and np.array_equal(a, b) and a.min() >= 0 and a.max() < 1000
    # SplitMix64 first output for seed 0 (Vigna's reference implementation)
    assert O.draw_samples(1 << 63, 2, 1.0 / (1 << 63), 0)[0] == 0xE220A8397B1DCDAF % (1 << 63)


# ------------------------------------------------- pyramid placement (oracle)
def test_shrunk_geometry_preserves_the_physical_centre():
    # crates/sitk/src/filters/shrink.rs:51-91: out = max(1, size / f), spacing * f,
    # origin shifted by delta so the physical centre of the grid is kept.
    size, sp, org = (37, 29, 23), (0.9, 1.1, 1.5), (-3.0, 2.0, 1.0)
    for f in [(1, 1, 1), (2, 2, 2), (4, 4, 4), (3, 5, 7), (100, 1, 2)]:
        osz, osp, oor = O.shrunk_geometry(size, sp, org, f)
        for d in range(3):
            assert osz[d] == max(1, size[d] // f[d])
            assert osp[d] == sp[d] * f[d]
            c_in = org[d] + sp[d] * (size[d] - 1) / 2
            c_out = oor[d] + osp[d] * (osz[d] - 1) / 2
            assert abs(c_in - c_out) < 1e-12
        if f == (1, 1, 1):
            assert osz == list(size) and oor == list(org)


def test_linear_resample_identity_shift_and_linear_fields():
    # crates/sitk/src/transform/resample.rs:226-290 with the linear interpolator
    # (src/transform/interpolator.rs:58-91).
    v = O.synth(3, 12 * 10 * 8).reshape(8, 10, 12)
    sp, org = (1.0, 0.8, 1.2), (1.0, -2.0, 3.0)
    ident = O.make_transform(O.AFFINE, [1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0])
    same = O.resample_linear(v, sp, org, (12, 10, 8), sp, org, ident)
    assert np.array_equal(same, v)
    # a whole-voxel translation shifts the data; voxels mapped outside get the default
    t = O.make_transform(O.TRANSLATION, [2 * sp[0], 0.0, -1 * sp[2]])
    sh = O.resample_linear(v, sp, org, (12, 10, 8), sp, org, t, default=-5.0)
    assert np.array_equal(sh[1:, :, :10], v[:-1, :, 2:])
    assert np.all(sh[:, :, 10:] == -5.0) and np.all(sh[0] == -5.0)
    # trilinear interpolation reproduces an affine field exactly (to rounding)
    z, y, x = np.meshgrid(np.arange(8), np.arange(10), np.arange(12), indexing="ij")
    ramp = (3.0 * x - 2.0 * y + 0.5 * z + 7.0).astype(np.float32)
    e = O.make_transform(O.EULER3D, [0.05, -0.03, 0.08, 0.3, -0.2, 0.1], center=[6, 2, 7])
    out = O.resample_linear(ramp, sp, org, (6, 5, 4), (1.3, 1.1, 1.4), (3.0, 0.0, 5.0), e,
                            default=np.nan)
    for k, j, i in [(0, 0, 0), (1, 2, 3), (3, 4, 5), (2, 1, 4)]:
        p = np.array([3.0 + 1.3 * i, 0.0 + 1.1 * j, 5.0 + 1.4 * k])
        q = O.transform_point(e, p)
        c = (q - np.array(org)) / np.array(sp)
        if np.all(c >= 0) and np.all(c <= np.array([11, 9, 7])):
            assert abs(out[k, j, i] - (3.0 * c[0] - 2.0 * c[1] + 0.5 * c[2] + 7.0)) < 1e-4


# ---------------------------------------------- B-spline transform (oracle)
def _bspline(coeff_fn, mesh=(3, 3, 3)):
    size, sp, org = (16, 16, 16), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    ncp = (mesh[0] + 3) * (mesh[1] + 3) * (mesh[2] + 3)
    c = np.zeros(3 * ncp)
    coeff_fn(c, ncp)
    return O.make_bspline(size, sp, org, np.eye(3), mesh, c), c, ncp


def test_bspline_zero_is_identity_and_constant_is_a_translation():
    # crates/sitk/src/transform/bspline.rs:561 (transform_point): the cubic
    # B-spline weights are a partition of unity over the 4^3 support.
    t0, _, _ = _bspline(lambda c, n: None)
    tc, _, _ = _bspline(lambda c, n: (c.__setitem__(slice(0, n), 1.5),
                                      c.__setitem__(slice(n, 2 * n), -0.25),
                                      c.__setitem__(slice(2 * n, 3 * n), 2.0)))
    rng = np.random.default_rng(1)
    for _ in range(50):
        p = rng.uniform(0.5, 14.5, 3)
        assert np.array_equal(O.transform_point(t0, p), p)
        assert np.allclose(O.transform_point(tc, p), p + [1.5, -0.25, 2.0], atol=1e-12)


def test_bspline_jacobian_is_the_derivative_of_the_point_map():
    # bspline.rs:675 (sparse_jacobian_wrt_parameters): the map is linear in the
    # coefficients, so each column is the displacement of a unit coefficient.
    rng = np.random.default_rng(2)
    t, c, ncp = _bspline(lambda c, n: c.__setitem__(slice(None), rng.normal(0, 0.5, c.size)))
    p = np.array([5.3, 7.9, 4.4])
    J = O.transform_jacobian(t, p)
    base = O.transform_point(t, p)
    for k in rng.choice(3 * ncp, 40, replace=False):
        c[k] += 1.0
        moved = O.transform_point(t, p)
        c[k] -= 1.0
        assert np.allclose(moved - base, J[:, k], atol=1e-12)
    # the weights of each row sum to one
    for d in range(3):
        assert abs(J[d, d * ncp:(d + 1) * ncp].sum() - 1.0) < 1e-12


def test_bspline_sparse_mattes_derivative_matches_finite_difference():
    # crates/sitk/src/registration/mattes.rs:722-796, pinned by FD as in
    # mattes.rs:952-1012.
    n = 16
    f = O.fixture_volume(n, [0, 0, 0])
    m = O.fixture_volume(n, [0.6, -0.4, 0.3])
    sp, org = [1.0, 1.0, 1.0], [0.0, 0.0, 0.0]
    pb = O.MetricProblem(f, sp, org, m, sp, org)
    g = pb.mattes_geometry(20)
    rng = np.random.default_rng(3)
    t, c, ncp = _bspline(lambda c, n_: c.__setitem__(slice(None), rng.normal(0, 0.2, c.size)),
                         mesh=(2, 2, 2))
    v0, d0, _ = pb.mattes_evaluate(g, t)
    assert len(d0) == 3 * ncp
    big = np.argsort(-np.abs(d0))[:6]
    for k in big:
        h = 1e-5
        c[k] += h
        vp = pb.mattes_value(g, t)
        c[k] -= 2 * h
        vm = pb.mattes_value(g, t)
        c[k] += h
        fd = (vp - vm) / (2 * h)
        assert abs(fd - d0[k]) < 5e-3 * max(1.0, abs(d0[k])), (k, fd, d0[k])


def test_bspline_grid_geometry_is_the_references():
    # bspline.rs:229-262 and its own pin (bspline.rs tests, "fixed parameters are
    # the control-point grid geometry"): gridSize = mesh + 3, gridSpacing =
    # physDim / mesh, gridOrigin = origin - gridSpacing. Restated in 3-D with a
    # domain of physical size (4, 8, 6) at origin (1, 2, 3) and mesh (2, 4, 3).
    t = O.make_bspline((4, 8, 6), (1.0, 1.0, 1.0), (1.0, 2.0, 3.0), np.eye(3), (2, 4, 3),
                       np.zeros(3 * 5 * 7 * 6))
    assert list(t.grid_size) == [5, 7, 6]
    assert list(t.grid_spacing) == [2.0, 2.0, 2.0]
    assert list(t.grid_origin) == [-1.0, 0.0, 1.0]
    assert t.nparams == 3 * 5 * 7 * 6


# ------------------------- B-spline with a moving-initial transform (oracle)
def _composed_problem():
    n = 16
    f = O.fixture_volume(n, [0, 0, 0])
    m = O.fixture_volume(n, [0.6, -0.4, 0.3])
    sp, org = [1.0, 1.0, 1.0], [0.0, 0.0, 0.0]
    pb = O.MetricProblem(f, sp, org, m, sp, org)
    return pb, pb.mattes_geometry(20)


def test_composition_without_initial_is_the_plain_sparse_evaluation():
    pb, g = _composed_problem()
    rng = np.random.default_rng(4)
    t, c, ncp = _bspline(lambda c, n_: c.__setitem__(slice(None), rng.normal(0, 0.2, c.size)),
                         mesh=(2, 2, 2))
    v0, d0, n0 = pb.mattes_evaluate(g, t)
    v1, d1, n1 = pb.mattes_evaluate_composed(g, t, None)
    assert v1 == v0 and n1 == n0 and np.array_equal(d1, d0)


def test_zero_bspline_then_rigid_is_the_rigid_transform():
    # moving_initial(optimized(x)) (registration/method.rs:198-240): with every
    # coefficient zero the B-spline maps x to x exactly, so the composition is
    # the moving-initial transform alone, bit for bit.
    pb, g = _composed_problem()
    t, c, ncp = _bspline(lambda c, n_: None, mesh=(2, 2, 2))
    rig = O.make_transform(O.EULER3D, [0.04, -0.03, 0.05, 0.4, -0.3, 0.2], center=[7.5, 7.5, 7.5])
    v0, _, n0 = pb.mattes_evaluate(g, rig)
    v1, d1, n1 = pb.mattes_evaluate_composed(g, t, rig)
    assert v1 == v0 and n1 == n0
    assert np.any(d1 != 0.0)


def test_composed_sparse_derivative_matches_finite_difference():
    # The sparse Jacobian's columns are left-multiplied by the moving-initial
    # transform's matrix (method.rs:313-340); a rotation is not symmetric, so
    # applying its matrix untransposed, or not at all, fails this check.
    pb, g = _composed_problem()
    rng = np.random.default_rng(5)
    t, c, ncp = _bspline(lambda c, n_: c.__setitem__(slice(None), rng.normal(0, 0.2, c.size)),
                         mesh=(2, 2, 2))
    rig = O.make_transform(O.EULER3D, [0.25, -0.2, 0.3, 0.4, -0.3, 0.2], center=[7.5, 7.5, 7.5])
    v0, d0, _ = pb.mattes_evaluate_composed(g, t, rig)
    _, d_plain, _ = pb.mattes_evaluate(g, t)
    assert not np.allclose(d0, d_plain, rtol=1e-3)
    big = np.argsort(-np.abs(d0))[:8]
    for k in big:
        h = 1e-5
        c[k] += h
        vp, _, _ = pb.mattes_evaluate_composed(g, t, rig)
        c[k] -= 2 * h
        vm, _, _ = pb.mattes_evaluate_composed(g, t, rig)
        c[k] += h
        fd = (vp - vm) / (2 * h)
        assert abs(fd - d0[k]) < 5e-3 * max(1.0, abs(d0[k])), (k, fd, d0[k])


# ---------------------------------------------------------- explicit points
def _grid_points(shape, spacing, origin, direction):
    """Every voxel's physical point in dim-0-fastest order, by the origin-first
    fold of VirtualGrid::write_point (scales.rs:259-271): ((o + a i) + b j) + c k,
    one rounding per step, as the oracle evaluates it."""
    i2p = O.index_to_physical(np.asarray(direction, float).reshape(9), spacing).reshape(3, 3)
    nz, ny, nx = shape
    k, j, i = np.meshgrid(np.arange(nz, dtype=float), np.arange(ny, dtype=float),
                          np.arange(nx, dtype=float), indexing="ij")
    i, j, k = i.reshape(-1), j.reshape(-1), k.reshape(-1)
    pts = np.empty((i.size, 3))
    for r in range(3):
        pts[:, r] = ((origin[r] + i2p[r, 0] * i) + i2p[r, 1] * j) + i2p[r, 2] * k
    return pts


def test_explicit_points_on_the_grid_are_the_grid():
    """FixedPoints::Explicit (cuda/ops/resident.rs:464-484) carrying exactly the
    grid's own points and values evaluates the same samples as the grid: mean
    squares and Mattes value, derivative and valid count bit for bit, under an
    oblique direction and a multi-stage-free Euler map."""
    rng = np.random.default_rng(4)
    f = (rng.random((9, 10, 11)) * 100).astype(np.float32)
    m = (f * 0.7 + rng.random(f.shape).astype(np.float32) * 20).astype(np.float32)
    sp, org = [0.9, 1.1, 1.3], [-2.0, 1.5, 0.5]
    c, s = math.cos(0.2), math.sin(0.2)
    fdir = [c, -s, 0.0, s, c, 0.0, 0.0, 0.0, 1.0]
    grid = O.MetricProblem(f, sp, org, m, sp, org, fdir=fdir, mdir=fdir)
    pts = _grid_points(f.shape, sp, org, fdir)
    expl = O.MetricProblem(f.reshape(-1), sp, org, m, sp, org, mdir=fdir, points=pts)
    t = O.make_transform(O.EULER3D, [0.03, -0.02, 0.05, 0.4, -0.3, 0.2], center=[3.0, 5.0, 5.0])
    assert grid.ms_evaluate(t)[0] == expl.ms_evaluate(t)[0]
    assert np.array_equal(grid.ms_evaluate(t)[1], expl.ms_evaluate(t)[1])
    g1, g2 = grid.mattes_geometry(20), expl.mattes_geometry(20)
    v1, d1, n1 = grid.mattes_evaluate(g1, t)
    v2, d2, n2 = expl.mattes_evaluate(g2, t)
    assert v1 == v2 and n1 == n2 and np.array_equal(d1, d2)


def test_explicit_point_derivative_matches_finite_difference():
    """Off-grid explicit points: the mean-squares derivative is the derivative
    of its value."""
    rng = np.random.default_rng(6)
    m = np.zeros((14, 15, 16), np.float32)
    z, y, x = np.meshgrid(np.arange(14), np.arange(15), np.arange(16), indexing="ij")
    m[:] = (50 * np.exp(-((x - 8) ** 2 + (y - 7) ** 2 + (z - 6) ** 2) / 18.0)
            + 0.5 * x).astype(np.float32)
    pts = rng.uniform(3.0, 10.0, (400, 3))
    vals = rng.uniform(0.0, 40.0, 400).astype(np.float32)
    pb = O.MetricProblem(vals, [1, 1, 1], [0, 0, 0], m, [1, 1, 1], [0, 0, 0], points=pts)
    p0 = [0.02, -0.01, 0.03, 0.3, -0.2, 0.1]
    _, an, n = pb.ms_evaluate(O.make_transform(O.EULER3D, p0, center=[7, 7, 7]))
    assert n == 400
    for k in range(6):
        h = 1e-6
        pp, pm = list(p0), list(p0)
        pp[k] += h
        pm[k] -= h
        fd = (pb.ms_value(O.make_transform(O.EULER3D, pp, center=[7, 7, 7]))[0]
              - pb.ms_value(O.make_transform(O.EULER3D, pm, center=[7, 7, 7]))[0]) / (2 * h)
        assert abs(fd - an[k]) < 1e-4 * max(1.0, abs(an[k])), (k, fd, an[k])


# ------------------------------------------------- nearest-neighbour resample
def _nn(v, t, default=0.0, sp=(1.0, 1.0, 1.0), org=(0.0, 0.0, 0.0), out=None):
    nz, ny, nx = v.shape
    size, osp, oorg = out if out is not None else ((nx, ny, nz), sp, org)
    return O.resample_nearest(v, sp, org, size, osp, oorg, t, default=default)


def test_nearest_resample_identity_shift_and_half_ties():
    """ResampleImageFilter with the nearest-neighbour interpolator
    (transform/resample.rs:226-290, interpolator.rs:37-54): the identity
    returns the input; a whole-voxel shift moves it and fills the default; a
    shift of exactly half a voxel rounds every tie up (RoundHalfIntegerUp,
    core/coord.rs:189-191: 2.5 -> 3, not to even); a 0.3-voxel shift picks the
    nearer voxel's own value (no blending)."""
    rng = np.random.default_rng(3)
    v = rng.normal(size=(5, 6, 7)).astype(np.float32)
    ident = O.make_transform(O.TRANSLATION, [0.0, 0.0, 0.0])
    assert np.array_equal(_nn(v, ident), v)
    # out(x) = in(x + 2) along x: the last two columns fall outside
    got = _nn(v, O.make_transform(O.TRANSLATION, [2.0, 0.0, 0.0]), default=-9.0)
    assert np.array_equal(got[:, :, :5], v[:, :, 2:]) and np.all(got[:, :, 5:] == -9.0)
    # c = i + 0.5 exactly: voxel i + 1, for even and odd i alike; the last
    # column maps to size - 0.5, outside the half-open interval
    got = _nn(v, O.make_transform(O.TRANSLATION, [0.5, 0.0, 0.0]), default=-9.0)
    assert np.array_equal(got[:, :, :6], v[:, :, 1:]) and np.all(got[:, :, 6] == -9.0)
    # c = i - 0.5: voxel i (the tie goes up), and c = -0.5 is inside
    got = _nn(v, O.make_transform(O.TRANSLATION, [-0.5, 0.0, 0.0]), default=-9.0)
    assert np.array_equal(got, v)
    got = _nn(v, O.make_transform(O.TRANSLATION, [0.3, 0.0, -0.3]))
    assert np.array_equal(got[:, :, :6], v[:, :, :6])
    assert np.array_equal(got[:, :, 6], v[:, :, 6])


def test_nearest_resample_onto_a_shrunk_grid_picks_the_containing_voxel():
    """Onto a coarser grid (shrink 2 about the centre, as a pyramid level
    places its mask): every output voxel takes the input voxel whose cell
    contains its physical point, checked against a direct computation (odd
    sizes put the coarse points on input voxels, away from the ties the first
    test covers)."""
    rng = np.random.default_rng(8)
    v = rng.integers(0, 2, size=(13, 9, 11)).astype(np.float32)
    sp, org = (0.9, 1.1, 1.3), (-2.0, 1.0, 0.5)
    size, osp, oorg = O.shrunk_geometry((11, 9, 13), sp, org, (2, 2, 2))
    got = O.resample_nearest(v, sp, org, size, osp, oorg,
                             O.make_transform(O.TRANSLATION, [0.0, 0.0, 0.0]))
    for k in range(size[2]):
        for j in range(size[1]):
            for i in range(size[0]):
                p = np.asarray(oorg) + np.asarray(osp) * np.array([i, j, k])
                c = (p - np.asarray(org)) / np.asarray(sp)
                idx = np.clip(np.floor(c + 0.5).astype(int), 0, np.array([11, 9, 13]) - 1)
                assert got[k, j, i] == v[idx[2], idx[1], idx[0]]


# ------------------------------------------------------------------ Euler3D
def test_euler3d_fixtures_of_the_reference():
    """transform/parametric.rs:4105-4183: identity, single-axis quarter turns,
    the centre is fixed, both composition orders orthonormal and different,
    and the Jacobian at the identity is the so(3) generators applied to p."""
    assert list(O.transform_point(O.make_transform(O.EULER3D, [0.0] * 6), [3.0, -4.0, 5.0])) == \
        [3.0, -4.0, 5.0]
    h = math.pi / 2
    p = O.transform_point(O.make_transform(O.EULER3D, [0.0, 0.0, h, 0, 0, 0]), [1.0, 0.0, 7.0])
    assert np.allclose(p, [0.0, 1.0, 7.0], atol=1e-12, rtol=0)
    q = O.transform_point(O.make_transform(O.EULER3D, [h, 0.0, 0.0, 0, 0, 0]), [9.0, 1.0, 0.0])
    assert np.allclose(q, [9.0, 0.0, 1.0], atol=1e-12, rtol=0)
    r = O.transform_point(O.make_transform(O.EULER3D, [0.0, h, 0.0, 0, 0, 0]), [0.0, 4.0, 1.0])
    assert np.allclose(r, [1.0, 4.0, 0.0], atol=1e-12, rtol=0)
    c = [5.0, -2.0, 3.0]
    assert np.allclose(O.transform_point(O.make_transform(O.EULER3D, [0.3, -0.5, 0.7, 0, 0, 0],
                                                          center=c), c), c, atol=1e-12, rtol=0)
    mats = []
    for zyx in (False, True):
        t = O.make_transform(O.EULER3D, [0.4, -0.6, 0.8, 0, 0, 0], compute_zyx=zyx)
        m = O.transform_stage(t)[:9].reshape(3, 3)
        assert np.allclose(m @ m.T, np.eye(3), atol=1e-12, rtol=0)
        mats.append(m)
    assert np.max(np.abs(mats[0] - mats[1])) > 1e-6
    j = O.transform_jacobian(O.make_transform(O.EULER3D, [0.0] * 6), [2.0, 3.0, 5.0])
    assert np.array_equal(j, np.array([[0.0, 5.0, -3.0, 1.0, 0.0, 0.0],
                                       [-5.0, 0.0, 2.0, 0.0, 1.0, 0.0],
                                       [3.0, -2.0, 0.0, 0.0, 0.0, 1.0]]))


def test_euler3d_composition_orders_are_the_published_rotations():
    """Default order Rz . Rx . Ry (parametric.rs:1030-1053) is scipy's
    intrinsic 'ZXY' rotation; compute_zyx Rz . Ry . Rx is intrinsic 'ZYX'."""
    from scipy.spatial.transform import Rotation

    ax, ay, az = 0.37, -0.52, 0.81
    t = O.make_transform(O.EULER3D, [ax, ay, az, 0, 0, 0])
    assert np.allclose(O.transform_stage(t)[:9].reshape(3, 3),
                       Rotation.from_euler("ZXY", [az, ax, ay]).as_matrix(), atol=1e-14, rtol=0)
    t = O.make_transform(O.EULER3D, [ax, ay, az, 0, 0, 0], compute_zyx=True)
    assert np.allclose(O.transform_stage(t)[:9].reshape(3, 3),
                       Rotation.from_euler("ZYX", [az, ay, ax]).as_matrix(), atol=1e-14, rtol=0)


# ------------------------------------------------------- masked moving range
def _mask_fixture_volume(n, shift, outermost):
    c = n / 2.0
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    x, y, z = i - c - shift, j - c, k - c
    v = np.exp(-(x * x + y * y + z * z) / 32.0)
    edge = (i == 0) | (j == 0) | (k == 0) | (i == n - 1) | (j == n - 1) | (k == n - 1)
    return np.where(edge, outermost, v).astype(np.float32)


def test_a_masked_out_voxel_does_not_size_the_moving_axis():
    """tests/mattes_mask.rs:95-150: moving volumes that differ only in their
    outermost layer (1000 vs 0.5) give different Mattes values without a mask
    (the fixture is real), and bit-identical value, derivative and valid count
    with a mask that excludes a two-voxel shell, at three poses."""
    n, bins = 16, 32
    fixed = _mask_fixture_volume(n, 0.0, 0.5)
    outl = _mask_fixture_volume(n, 1.0, 1000.0)
    tame = _mask_fixture_volume(n, 1.0, 0.5)
    mask = np.zeros((n, n, n), np.uint8)
    mask[2:n - 2, 2:n - 2, 2:n - 2] = 1
    poses = [[0.0, 0.0, 0.0], [0.7, -0.4, 0.3], [-1.3, 0.9, -0.6]]
    one = [1.0, 1.0, 1.0]
    zero = [0.0, 0.0, 0.0]
    for mm in (None, mask):
        a = O.MetricProblem(fixed, one, zero, outl, one, zero, moving_mask=mm)
        b = O.MetricProblem(fixed, one, zero, tame, one, zero, moving_mask=mm)
        ga, gb = a.mattes_geometry(bins), b.mattes_geometry(bins)
        for p in poses:
            t = O.make_transform(O.TRANSLATION, p)
            va, da, na = a.mattes_evaluate(ga, t)
            vb, db, nb = b.mattes_evaluate(gb, t)
            if mm is None:
                assert va != vb, p
            else:
                assert va == vb and na == nb and np.array_equal(da, db), p
