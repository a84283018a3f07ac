"""The cubic B-spline interpolator of the moving image on the device
(MovingImage::from_image_with_interpolator, crates/sitk/src/registration/
metric.rs:720-750, with crates/sitk/src/transform/interpolator.rs:170-350)
against the oracle.

The coefficients are FP64 on the device in the reference's operation order:
bit for bit the oracle's (narrowed to float on the FP32 path). Metric
tolerances are those of the linear interpolator: F64 value 1e-9 relative and
derivative 1e-7 relative (vector norm), F32 value 1e-6 and derivative 1e-4;
valid counts exact.
"""

import numpy as np
import pytest

import oracle as O
from helpers import affine_form, blob_volume, record_error, rel_vec, stages_of
from oracle.registration import register_rsgd
from test_gpu_metric import ORG, SP, TRANSFORMS, contract, ledger_pair

pytestmark = pytest.mark.gpu

TOL = {0: (1e-6, 1e-4), 1: (1e-9, 1e-7)}


def _mattes(hip, f, m, precision, bins=50, indices=None, mmask=None):
    df = hip.DeviceImage.upload(f, spacing=SP, origin=ORG)
    dm = hip.DeviceImage.upload(m, spacing=SP, origin=ORG)
    kw = dict(precision=precision, indices=indices)
    if mmask is not None:
        kw["moving_mask"] = hip.DeviceMask.upload(mmask, spacing=SP, origin=ORG)
    d = hip.ResidentMattes(df, dm, bins=bins, **kw)
    d.set_interpolator(hip.INTERP_BSPLINE)
    pb = O.MetricProblem(f, SP, ORG, m, SP, ORG, indices=indices, moving_mask=mmask,
                         interpolator="bspline")
    return d, pb


@pytest.mark.parametrize("shape", [(24, 20, 17), (9, 1, 13), (3, 7, 1), (1, 5, 6)])
@pytest.mark.parametrize("precision", [0, 1])
def test_coefficients_are_the_oracles_bit_for_bit(hip, shape, precision):
    v = blob_volume(shape, 4) if min(shape) > 1 else \
        np.random.default_rng(7).uniform(-40, 200, shape).astype(np.float32)
    want = O.bspline_coefficients(v).reshape(-1)
    df = hip.DeviceImage.upload(v)
    d = hip.ResidentMetric(df, df, precision=precision)
    d.set_interpolator(hip.INTERP_BSPLINE)
    got = d.bspline_coefficients(v.size)
    if precision == 1:
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    else:
        assert np.array_equal(got.astype(np.float32).view(np.uint32),
                              want.astype(np.float32).view(np.uint32))


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("case", range(len(TRANSFORMS)))
def test_mattes_value_and_derivative(hip, precision, case):
    if precision == 0 and case == 0:
        pytest.skip("F32 on the lattice (the identity between equal grids puts every sample on a "
                    "voxel, the moving extremes included): whether the oracle's FP64 64-tap sum "
                    "at the extreme voxels lands inside the moving range [min, max] or a rounding "
                    "outside it is decided at 1e-13, below what float coefficients resolve, so "
                    "those samples' validity is not comparable; F64 covers this pose")
    kind, p, ctr = TRANSFORMS[case]
    f, m = ledger_pair()
    d, pb = _mattes(hip, f, m, precision)
    g = pb.mattes_geometry(50)
    t = O.make_transform(kind, p, center=ctr)
    st = stages_of(t)
    wv, wd, wn = pb.mattes_evaluate(g, t)
    v, mom, valid = d.value_and_moments(st)
    assert valid == wn
    tv, td = TOL[precision]
    ev, ed = abs(v - wv) / abs(wv), rel_vec(contract(mom, t), wd)
    record_error(f"cubic B-spline interpolation, ledger pair, transform {case}",
                 precision="f64" if precision else "f32", value_rel=ev, derivative_rel=ed)
    assert ev <= tv, (v, wv)
    assert ed < td
    # value only, and the derivative at the oracle's pRatio (the gathering pass)
    v2, n2 = d.value(st)
    assert n2 == wn and abs(v2 - wv) <= tv * abs(wv)
    if precision == 1:
        joint, marg, n3 = d.joint_histogram(st)
        wj, wm, _ = pb.mattes_histogram(g, t)
        assert n3 == wn and np.max(np.abs(joint - wj)) < 1e-6


def test_derivative_at_the_callers_pratio(hip):
    f, m = ledger_pair()
    d, pb = _mattes(hip, f, m, 1)
    g = pb.mattes_geometry(50)
    kind, p, ctr = TRANSFORMS[4]
    t = O.make_transform(kind, p, center=ctr)
    joint, marg, valid = pb.mattes_histogram(g, t)
    _, pratio = pb.mattes_tail(joint, marg, valid, g)
    want = pb.mattes_moments(g, t, pratio)
    got = d.derivative_moments(stages_of(t), pratio)
    assert rel_vec(got, want) < 1e-9


def test_sampled_and_moving_masked(hip):
    f, m = ledger_pair()
    idx = O.draw_samples(f.size, 2, 0.4, 11)
    mm = np.ones(m.shape, np.uint8)
    mm[:, :, :4] = 0
    mm[20:, :, :] = 0
    for precision in (0, 1):
        d, pb = _mattes(hip, f, m, precision, bins=32, indices=idx, mmask=mm)
        g = pb.mattes_geometry(32)
        t = O.make_transform(*TRANSFORMS[5][:2], center=TRANSFORMS[5][2])
        wv, wd, wn = pb.mattes_evaluate(g, t)
        v, mom, valid = d.value_and_moments(stages_of(t))
        tv, td = TOL[precision]
        assert valid == wn
        assert abs(v - wv) <= tv * abs(wv)
        assert rel_vec(contract(mom, t), wd) < td


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("case", [1, 4, 5])
def test_mean_squares(hip, precision, case):
    kind, p, ctr = TRANSFORMS[case]
    f = blob_volume((20, 24, 28), 3)
    m = (blob_volume((20, 24, 28), 3) * 0.9 + 4.0).astype(np.float32)
    df = hip.DeviceImage.upload(f, spacing=SP, origin=ORG)
    dm = hip.DeviceImage.upload(m, spacing=SP, origin=ORG)
    d = hip.ResidentMetric(df, dm, precision=precision)
    d.set_interpolator(hip.INTERP_BSPLINE)
    pb = O.MetricProblem(f, SP, ORG, m, SP, ORG, interpolator="bspline")
    t = O.make_transform(kind, p, center=ctr)
    mom = d.evaluate(stages_of(t))
    j0, c = affine_form(t)
    v, deriv, valid = hip.contract_ms(mom, j0, c, t.nparams)
    wv, wd, wn = pb.ms_evaluate(t)
    assert valid == wn
    tv, td = (1e-12, 1e-9) if precision == 1 else (1e-6, 1e-4)
    assert abs(v - wv) <= tv * abs(wv)
    assert rel_vec(deriv, wd) < td


def test_back_to_linear_is_the_linear_metric(hip):
    f, m = ledger_pair()
    df = hip.DeviceImage.upload(f, spacing=SP, origin=ORG)
    dm = hip.DeviceImage.upload(m, spacing=SP, origin=ORG)
    lin = hip.ResidentMattes(df, dm, bins=50)
    d = hip.ResidentMattes(df, dm, bins=50)
    st = stages_of(O.make_transform(*TRANSFORMS[4][:2], center=TRANSFORMS[4][2]))
    want = lin.value_and_moments(st)
    d.set_interpolator(hip.INTERP_BSPLINE)
    cub = d.value_and_moments(st)
    assert cub[0] != want[0]
    d.set_interpolator(hip.INTERP_LINEAR)
    got = d.value_and_moments(st)
    assert got[0] == want[0] and np.array_equal(got[1], want[1]) and got[2] == want[2]


def test_refusals(hip):
    f, m = ledger_pair()
    df = hip.DeviceImage.upload(f, spacing=SP, origin=ORG)
    dm = hip.DeviceImage.upload(m, spacing=SP, origin=ORG)
    d = hip.ResidentMattes(df, dm, bins=50)
    with pytest.raises(hip.CudaError) as e:
        d.set_interpolator(7)
    assert e.value.status == "InvalidArgument"
    with pytest.raises(hip.CudaError):
        d.bspline_coefficients(f.size)  # linear: no coefficients
    # a moving axis of two voxels: the mirror fold reads index -1 there
    two = hip.DeviceImage.upload(np.ones((5, 2, 6), np.float32) * np.arange(6, dtype=np.float32))
    ms = hip.ResidentMetric(two, two)
    with pytest.raises(hip.CudaError) as e:
        ms.set_interpolator(hip.INTERP_BSPLINE)
    assert e.value.status == "UnsupportedGeometry"
    # batched evaluation takes linear metrics only
    d.set_interpolator(hip.INTERP_BSPLINE)
    with pytest.raises(hip.CudaError, match="linear"):
        hip.Batch([d])


def test_registration_with_the_cubic_interpolator_matches_the_host_loop(hip):
    """set_interpolator(BSpline) on the registration (method.rs:1507-1520):
    a single-level Mattes RSGD registration through the device driver takes
    the host loop's iterates on the oracle with the cubic interpolator (F64:
    the same iteration count and stop, parameters within 1e-9)."""
    n = 32
    f = O.fixture_volume(n, [0, 0, 0])
    m = O.fixture_inverted(n, [-2.0, 1.0, 0.0])
    pb = O.MetricProblem(f, SP, ORG, m, SP, ORG, interpolator="bspline")
    g = pb.mattes_geometry(50)
    want, want_it, want_stop = register_rsgd(pb, g, O.TRANSLATION, [0.0, 0.0, 0.0], [0, 0, 0],
                                             (n, n, n), SP, ORG, minimum_step=0.01,
                                             iterations=200, grad_tol=1e-6)
    df = hip.DeviceImage.upload(f, spacing=SP, origin=ORG)
    dm = hip.DeviceImage.upload(m, spacing=SP, origin=ORG)
    cfg = hip.registration_config(metric=hip.MATTES, bins=50, precision=1,
                                  optimizer=hip.REGULAR_STEP, minimum_step=0.01, iterations=200,
                                  gradient_tolerance=1e-6, learning_rate_mode=hip.LR_ONCE,
                                  interpolator=hip.INTERP_BSPLINE)
    got, res = hip.register(df, dm, cfg, hip.TRANSLATION, [0.0, 0.0, 0.0])
    assert res.iterations == want_it, (res.iterations, want_it, want_stop)
    assert hip.STOP_NAMES[res.level_stop[0]] == want_stop
    assert np.max(np.abs(got - want)) < 1e-9, (got, want)
    # batched registration refuses the cubic interpolator by name
    with pytest.raises(hip.CudaError, match="linearly"):
        hip.register_batch([df], [dm], [cfg], [hip.TRANSLATION], [[0.0, 0.0, 0.0]], [[0, 0, 0]])
