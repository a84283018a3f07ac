"""Mattes MI through a cubic B-spline transform on the device against the
oracle's sparse-support evaluation (crates/sitk/src/registration/mattes.rs:
722-796, FD-pinned in tests/test_oracle_golden.py). Tolerances as for the
affine path: F64 value 1e-9, derivative 1e-7 (relative, vector norm);
F32 value 1e-6, derivative 1e-4; the valid count exact."""

import numpy as np
import pytest

import oracle as O
from helpers import blob_volume, rel_vec

pytestmark = pytest.mark.gpu

SP = [1.0, 1.1, 0.9]
ORG = [-2.0, 1.0, 0.5]


def _case(hip, precision, domain_size, mesh, sampled, seed):
    f = blob_volume((18, 20, 22), 30 + seed, spacing=SP)
    m = blob_volume((18, 20, 22), 60 + seed, spacing=SP) * 0.4 + f * 0.6
    idx = None
    if sampled:
        idx = hip.draw_samples(f.size, hip.SAMPLING_RANDOM, 0.5, 3 + seed)
    df = hip.DeviceImage.upload(f, spacing=SP, origin=ORG)
    dm = hip.DeviceImage.upload(m, spacing=SP, origin=ORG)
    met = hip.ResidentMattes(df, dm, bins=24, indices=idx, precision=precision)
    rng = np.random.default_rng(seed)
    t = O.make_bspline(domain_size, SP, ORG, np.eye(3), mesh, np.zeros(1))
    ncp = int(np.prod(list(t.grid_size)))
    coeffs = rng.normal(0.0, 0.6, 3 * ncp)
    t = O.make_bspline(domain_size, SP, ORG, np.eye(3), mesh, coeffs)
    bs = hip.BSplineMattes(met, list(t.grid_size), list(t.grid_origin), list(t.grid_spacing),
                           np.array(list(t.grid_direction)).reshape(3, 3))
    pb = O.MetricProblem(f, SP, ORG, m, SP, ORG, indices=idx)
    g = pb.mattes_geometry(24)
    return bs, pb, g, t, coeffs, met


@pytest.mark.parametrize("precision", [1, 0])
@pytest.mark.parametrize("domain,mesh,sampled", [((22, 20, 18), (3, 3, 3), False),
                                                 ((22, 20, 18), (2, 4, 3), True),
                                                 ((12, 10, 14), (2, 2, 2), False)])
def test_bspline_value_and_derivative(hip, precision, domain, mesh, sampled):
    bs, pb, g, t, coeffs, met = _case(hip, precision, domain, mesh, sampled, seed=len(domain) + mesh[1])
    want_v, want_d, want_n = pb.mattes_evaluate(g, t)
    v, d, n = bs.value_and_derivative(coeffs)
    assert n == want_n
    tv, td = (1e-9, 1e-7) if precision == 1 else (1e-6, 1e-4)
    assert abs(v - want_v) <= tv * abs(want_v), (v, want_v)
    assert rel_vec(d, want_d) < td, rel_vec(d, want_d)
    v2, n2 = bs.value(coeffs)
    assert v2 == v and n2 == n
    # deterministic
    v3, d3, _ = bs.value_and_derivative(coeffs)
    assert v3 == v and np.array_equal(d3, d)


def test_zero_coefficients_are_the_identity(hip):
    bs, pb, g, t, coeffs, met = _case(hip, 1, (22, 20, 18), (3, 3, 3), False, seed=9)
    zero = np.zeros_like(coeffs)
    v, d, n = bs.value_and_derivative(zero)
    ident = O.make_transform(O.TRANSLATION, [0.0, 0.0, 0.0])
    want_v, _, want_n = pb.mattes_evaluate(g, ident)
    assert n == want_n and abs(v - want_v) <= 1e-9 * abs(want_v)


def test_bspline_refusals(hip):
    f = blob_volume((12, 12, 12), 1)
    df = hip.DeviceImage.upload(f)
    ms = hip.ResidentMetric(df, df)
    with pytest.raises(hip.CudaError, match="HistogramShape"):
        hip.BSplineMattes(ms, (6, 6, 6), (0, 0, 0), (3, 3, 3))
    mt = hip.ResidentMattes(df, df, bins=20)
    with pytest.raises(hip.CudaError, match="InvalidArgument"):
        hip.BSplineMattes(mt, (3, 6, 6), (0, 0, 0), (3, 3, 3))


def test_bspline_registration_descends_and_reports_the_oracles_value(hip):
    """A deformable registration (L-BFGS over the device B-spline metric) on a
    smoothly warped pair: the Mattes value it reports is the oracle's value at
    the coefficients it returns (F64 sampler, 1e-9), and it is below the value
    at the start. (The metric is not regularised -- neither is the
    reference's -- so image-space error is not what it minimises.)"""
    n = 32
    sp, org = (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    f = blob_volume((n, n, n), 123, nblobs=16)
    rng = np.random.default_rng(4)
    mesh = (2, 2, 2)
    ncp = 5 ** 3
    true = rng.normal(0.0, 1.2, 3 * ncp)
    t_true = O.make_bspline((n, n, n), sp, org, np.eye(3), mesh, true)
    m = O.resample_linear(f, sp, org, (n, n, n), sp, org, t_true)
    df = hip.DeviceImage.upload(f, spacing=sp, origin=org)
    dm = hip.DeviceImage.upload(m, spacing=sp, origin=org)
    coeffs, res = hip.register_bspline(df, dm, mesh, bins=24, iterations=40,
                                       precision=hip.PRECISION_F64)
    pb = O.MetricProblem(f, sp, org, m, sp, org)
    g = pb.mattes_geometry(24)
    start = pb.mattes_value(g, O.make_bspline((n, n, n), sp, org, np.eye(3), mesh, np.zeros(3 * ncp)))
    got = pb.mattes_value(g, O.make_bspline((n, n, n), sp, org, np.eye(3), mesh, coeffs))
    assert res.iterations > 0 and res.evaluations >= res.iterations
    assert abs(res.value - got) <= 1e-9 * abs(got), (res.value, got)
    assert got < start - 1e-3, (start, got)


# ------------------------------------- moving-initial transform and many starts
def _stage(params, center):
    return O.transform_stage(O.make_transform(O.EULER3D, params, center=center))


@pytest.mark.parametrize("precision", [1, 0])
def test_moving_initial_stage_against_the_oracle(hip, precision):
    """moving point = moving_initial(B(x)) (registration/method.rs:198-340):
    value, derivative (the chain rule through the stage's matrix) and valid
    count against the oracle's composed sparse evaluation."""
    bs, pb, g, t, coeffs, met = _case(hip, precision, (22, 20, 18), (3, 3, 3), True, seed=11)
    params, center = [0.12, -0.08, 0.15, 0.7, -0.4, 0.3], [9.0, 11.0, 10.0]
    rig = O.make_transform(O.EULER3D, params, center=center)
    bs.set_moving_initial(O.transform_stage(rig))
    want_v, want_d, want_n = pb.mattes_evaluate_composed(g, t, rig)
    v, d, n = bs.value_and_derivative(coeffs)
    assert n == want_n
    tv, td = (1e-9, 1e-7) if precision == 1 else (1e-6, 1e-4)
    assert abs(v - want_v) <= tv * abs(want_v), (v, want_v)
    assert rel_vec(d, want_d) < td, rel_vec(d, want_d)
    bs.set_moving_initial(None)
    v0, d0, _ = bs.value_and_derivative(coeffs)
    want0, _, _ = pb.mattes_evaluate(g, t)
    assert abs(v0 - want0) <= tv * abs(want0)


@pytest.mark.parametrize("precision", [1, 0])
def test_many_starts_equal_one_start_at_a_time(hip, precision):
    """Each start of the batched evaluation is bit-equal to the single-start
    evaluation with the same stage, whichever starts are active."""
    bs, pb, g, t, coeffs, met = _case(hip, precision, (22, 20, 18), (3, 3, 3), True, seed=12)
    rng = np.random.default_rng(7)
    S = 5
    stages = [None if s == 0 else _stage(rng.uniform(-0.15, 0.15, 3).tolist()
                                         + rng.uniform(-1, 1, 3).tolist(), [9.0, 11.0, 10.0])
              for s in range(S)]
    st = np.stack([np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0.0]) if x is None else x for x in stages])
    cs = np.stack([rng.normal(0.0, 0.5, bs.nparams) for _ in range(S)])
    multi = hip.BSplineMulti(bs, S, st)
    for active in (None, [3, 1], [4]):
        vals, ders, valid = multi.value_and_derivative(cs, active)
        for s in (range(S) if active is None else active):
            bs.set_moving_initial(stages[s])
            v, d, n = bs.value_and_derivative(cs[s])
            assert vals[s] == v and valid[s] == n and np.array_equal(ders[s], d), s
    bs.set_moving_initial(None)
    for bad in ([1, 1], [5], [-1]):
        with pytest.raises(hip.CudaError, match="InvalidArgument"):
            multi.value_and_derivative(cs, bad)


def test_many_starts_grouped_four_cells_per_workgroup(hip):
    """With 16 or more active starts the value pass groups four cells per
    workgroup: still bit-equal per start, the outside-the-grid bucket
    included (the grid covers only part of the image here)."""
    bs, pb, g, t, coeffs, met = _case(hip, 0, (12, 10, 14), (2, 2, 2), True, seed=13)
    rng = np.random.default_rng(8)
    S = 20
    st = np.stack([_stage(rng.uniform(-0.1, 0.1, 3).tolist() + rng.uniform(-1, 1, 3).tolist(),
                          [9.0, 11.0, 10.0]) for _ in range(S)])
    cs = np.stack([rng.normal(0.0, 0.5, bs.nparams) for _ in range(S)])
    multi = hip.BSplineMulti(bs, S, st)
    vals, ders, valid = multi.value_and_derivative(cs)
    for s in (0, 7, 19):
        bs.set_moving_initial(st[s])
        v, d, n = bs.value_and_derivative(cs[s])
        assert vals[s] == v and valid[s] == n and np.array_equal(ders[s], d), s
    bs.set_moving_initial(None)


def _warped_pair(hip, n=32, seed=4):
    sp, org = (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    f = blob_volume((n, n, n), 123, nblobs=16)
    rng = np.random.default_rng(seed)
    mesh = (2, 2, 2)
    true = rng.normal(0.0, 1.2, 3 * 5 ** 3)
    t_true = O.make_bspline((n, n, n), sp, org, np.eye(3), mesh, true)
    m = O.resample_linear(f, sp, org, (n, n, n), sp, org, t_true)
    return (hip.DeviceImage.upload(f, spacing=sp, origin=org),
            hip.DeviceImage.upload(m, spacing=sp, origin=org), mesh)


def test_multistart_registration_equals_one_start_at_a_time(hip):
    """Each start of a batched multi-start registration ends exactly where the
    same start registered alone ends (same evaluations, so the same L-BFGS-B
    iterates), and the best start is the lowest value."""
    df, dm, mesh = _warped_pair(hip)
    S = 8  # two lanes of four starts
    rng = np.random.default_rng(9)
    stages = np.stack([_stage(rng.uniform(-0.05, 0.05, 3).tolist() + rng.uniform(-0.8, 0.8, 3).tolist(),
                              [16.0, 16.0, 16.0]) for _ in range(S)])
    kw = dict(bins=24, iterations=15, sampling=hip.SAMPLING_RANDOM, sampling_percentage=0.5, seed=5)
    xs, rs, best = hip.register_bspline_multistart(df, dm, mesh, S, stages=stages, **kw)
    for s in (0, 3, 4, 7):
        x1, r1, _ = hip.register_bspline_multistart(df, dm, mesh, 1, stages=stages[s:s + 1], **kw)
        assert np.array_equal(x1[0], xs[s]), s
        assert (r1[0].value, r1[0].iterations, r1[0].evaluations, r1[0].stop) == \
            (rs[s].value, rs[s].iterations, rs[s].evaluations, rs[s].stop), s
    assert rs[best].value == min(r.value for r in rs)
    assert all(r.iterations > 0 for r in rs)


def test_bounded_multistart_stays_in_the_box(hip):
    df, dm, mesh = _warped_pair(hip, seed=6)
    npar = 3 * 5 ** 3
    lo, hi = -0.4 * np.ones(npar), 0.3 * np.ones(npar)
    sel = np.full(npar, 2, dtype=np.int32)
    sel[::7] = 0  # some unbounded
    xs, rs, best = hip.register_bspline_multistart(
        df, dm, mesh, 2, bins=24, iterations=20, precision=hip.PRECISION_F64,
        coeffs=np.stack([np.zeros(npar), np.full(npar, 0.1)]),
        bound_selection=sel, lower=lo, upper=hi)
    for s in range(2):
        b = xs[s][sel == 2]
        assert np.all((b >= -0.4) & (b <= 0.3)), s
        assert rs[s].iterations > 0
    # some coefficients end on a bound: the box is active
    assert np.any(np.isclose(xs[0][sel == 2], -0.4) | np.isclose(xs[0][sel == 2], 0.3))


@pytest.mark.parametrize("precision", [1, 0])
def test_deformable_pair_at_scale_against_the_oracle(hip, precision):
    """The deformable workload's generator at 96^3 (RBF bumps on an Euler3D
    offset), 10% random draws, a 10^3-point control grid over the image and a
    rigid moving-initial stage: a batch of two starts against the oracle's
    composed sparse evaluation, value, derivative and valid count."""
    import workload as W

    f, m, truth = W.config_deformable(edge=96, nblobs=17, seed=21)
    idx = hip.draw_samples(f.data.size, hip.SAMPLING_RANDOM, 0.10, 9)
    df = hip.DeviceImage.upload(f.data, spacing=f.spacing, origin=f.origin)
    dm = hip.DeviceImage.upload(m.data, spacing=m.spacing, origin=m.origin)
    met = hip.ResidentMattes(df, dm, bins=32, indices=idx, precision=precision)
    mesh = (7, 7, 7)
    gs, go, gsp = W.bspline_grid(f, mesh)
    bs = hip.BSplineMattes(met, list(gs), list(go), list(gsp))
    rng = np.random.default_rng(3)
    coeffs = np.stack([rng.normal(0.0, 0.8, bs.nparams) for _ in range(2)])
    rig = [O.make_transform(O.EULER3D, list(rng.uniform(-0.05, 0.05, 3)) + list(rng.uniform(-2, 2, 3)),
                            center=list(f.center())) for _ in range(2)]
    multi = hip.BSplineMulti(bs, 2, np.stack([O.transform_stage(r) for r in rig]))
    vals, ders, valid = multi.value_and_derivative(coeffs)
    pb = O.MetricProblem(f.data, f.spacing, f.origin, m.data, m.spacing, m.origin, indices=idx)
    g = pb.mattes_geometry(32)
    tv, td = (1e-9, 1e-7) if precision == 1 else (1e-6, 1e-4)
    for s in range(2):
        t = O.make_bspline(f.size, f.spacing, f.origin, np.eye(3), mesh, coeffs[s])
        want_v, want_d, want_n = pb.mattes_evaluate_composed(g, t, rig[s])
        assert valid[s] == want_n, s
        assert abs(vals[s] - want_v) <= tv * abs(want_v), (s, vals[s], want_v)
        assert rel_vec(ders[s], want_d) < td, (s, rel_vec(ders[s], want_d))
