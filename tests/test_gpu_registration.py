"""Device pyramid placement and the device registration driver.

- shrink geometry / shrink / linear resample: bit-exact against the oracle's
  restatement of ShrinkImageFilter and ResampleImageFilter
  (crates/sitk/src/filters/shrink.rs:51-91, src/transform/resample.rs:226-290).
- a single-level Mattes registration lands where the host loop lands, with the
  same number of iterations (the reference's pin, crates/sitk/tests/
  mattes_device.rs:459-511: parameters within 1e-3, equal iteration counts).
- a three-level registration of a synthetic blob pair recovers the inverse of
  the warp that made it (TRE below 0.5 mm).
"""

import numpy as np
import pytest

import oracle as O
from oracle.registration import register_gd_line_search, register_rsgd

pytestmark = pytest.mark.gpu


def test_shrunk_geometry_is_the_filters(hip):
    g = hip.Geometry.make((256, 256, 160), (0.9, 0.9, 1.5), (-3.0, 2.0, 1.0))
    for f in [(4, 4, 4), (2, 2, 2), (1, 1, 1), (3, 5, 7), (300, 1, 2)]:
        out = hip.shrunk_geometry(g, f)
        size, sp, org = O.shrunk_geometry((256, 256, 160), (0.9, 0.9, 1.5), (-3.0, 2.0, 1.0), f)
        assert list(out.size) == size and list(out.spacing) == sp and list(out.origin) == org


@pytest.mark.parametrize("factors", [(2, 2, 2), (4, 4, 4), (3, 2, 5)])
def test_pyramid_placement_is_bit_exact(hip, factors):
    v = O.synth(77, 37 * 29 * 23).reshape(23, 29, 37)
    sp, org = (0.9, 1.1, 1.5), (-3.0, 2.0, 1.0)
    img = hip.DeviceImage.upload(v, spacing=sp, origin=org)
    grid = hip.shrunk_geometry(img.geometry(), factors)
    got = hip.resample_linear(img, grid).to_host()
    ident = O.make_transform(O.AFFINE, [1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0])
    want = O.resample_linear(v, sp, org, list(grid.size), list(grid.spacing), list(grid.origin),
                             ident)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_resample_through_a_transform_is_bit_exact(hip):
    v = O.synth(5, 40 * 30 * 20).reshape(20, 30, 40)
    sp, org = (1.0, 0.8, 1.2), (1.0, -2.0, 3.0)
    img = hip.DeviceImage.upload(v, spacing=sp, origin=org)
    t = O.make_transform(O.EULER3D, [0.1, -0.05, 0.2, 1.5, -2.0, 0.7], center=[20, 10, 15])
    st = O.transform_stage(t).reshape(1, 12)
    out_geom = hip.Geometry.make((33, 25, 17), (1.1, 1.0, 1.3), (0.5, -1.0, 2.5))
    got = hip.resample_linear(img, out_geom, st, default_value=-7.0).to_host()
    want = O.resample_linear(v, sp, org, (33, 25, 17), (1.1, 1.0, 1.3), (0.5, -1.0, 2.5), t,
                             default=-7.0)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_shrink_gathers_every_fth_voxel(hip):
    v = O.synth(9, 20 * 17 * 13).reshape(13, 17, 20)
    img = hip.DeviceImage.upload(v)
    out = hip.shrink(img, (2, 3, 4)).to_host()
    # offset floor(delta + 0.5), delta = (size-1)/2 - f(out-1)/2: x 0.5 -> 1,
    # y 2 -> 2, z 2 -> 2
    assert np.array_equal(out, v[2::4, 2::3, 1::2][:3, :5, :10])


@pytest.mark.parametrize("precision", [1, 0])
def test_single_level_mattes_registration_matches_the_host_loop(hip, precision):
    n = 32
    sp, org = [0.8, 0.9, 1.1], [-3.0, 2.0, 1.0]
    f = O.fixture_volume(n, [0, 0, 0])
    m = O.fixture_inverted(n, [-2.0, 1.0, 0.0])
    pb = O.MetricProblem(f, sp, org, m, sp, org)
    g = pb.mattes_geometry(50)
    want, want_it, want_stop = register_rsgd(pb, g, O.TRANSLATION, [0.0, 0.0, 0.0], [0, 0, 0],
                                             (n, n, n), sp, org, minimum_step=0.01,
                                             iterations=200, grad_tol=1e-6)
    df = hip.DeviceImage.upload(f, spacing=sp, origin=org)
    dm = hip.DeviceImage.upload(m, spacing=sp, origin=org)
    cfg = hip.registration_config(metric=hip.MATTES, bins=50, precision=precision,
                                  optimizer=hip.REGULAR_STEP, minimum_step=0.01, iterations=200,
                                  gradient_tolerance=1e-6, learning_rate_mode=hip.LR_ONCE)
    got, res = hip.register(df, dm, cfg, hip.TRANSLATION, [0.0, 0.0, 0.0])
    assert np.max(np.abs(got)) > 0.5  # it moved
    tol = 1e-3 if precision == 1 else 1e-2
    assert np.max(np.abs(got - want)) < tol, (got, want)
    if precision == 1:
        assert res.iterations == want_it, (res.iterations, want_it, want_stop)
    assert hip.STOP_NAMES[res.level_stop[0]] == want_stop or precision == 0


@pytest.mark.parametrize("precision", [1, 0])
def test_single_level_line_search_registration_matches_the_host_loop(hip, precision):
    """Gradient descent with the golden-section line search and the window
    convergence stop (optimizer.rs:587-785) against its host restatement on
    the oracle's metric: the same iterates, so the same result, iteration
    count and stop reason (F64)."""
    n = 32
    sp, org = [0.8, 0.9, 1.1], [-3.0, 2.0, 1.0]
    f = O.fixture_volume(n, [0, 0, 0])
    m = O.fixture_inverted(n, [-2.0, 1.0, 0.0])
    pb = O.MetricProblem(f, sp, org, m, sp, org)
    g = pb.mattes_geometry(50)
    want, want_it, want_stop = register_gd_line_search(pb, g, O.TRANSLATION, [0.0, 0.0, 0.0],
                                                       [0, 0, 0], (n, n, n), sp, org,
                                                       iterations=40)
    df = hip.DeviceImage.upload(f, spacing=sp, origin=org)
    dm = hip.DeviceImage.upload(m, spacing=sp, origin=org)
    cfg = hip.registration_config(metric=hip.MATTES, bins=50, precision=precision,
                                  optimizer=hip.LINE_SEARCH, iterations=40,
                                  learning_rate_mode=hip.LR_ONCE)
    got, res = hip.register(df, dm, cfg, hip.TRANSLATION, [0.0, 0.0, 0.0])
    assert np.max(np.abs(got)) > 0.5  # it moved
    tol = 1e-6 if precision == 1 else 1e-2
    assert np.max(np.abs(got - want)) < tol, (got, want)
    if precision == 1:
        assert res.iterations == want_it, (res.iterations, want_it, want_stop)
        assert hip.STOP_NAMES[res.level_stop[0]] == want_stop


def test_three_level_registration_recovers_the_warp(hip):
    import workload as W

    fixed = W.blob_volume((64, 64, 48), (1.0, 1.0, 1.5), seed=11, nblobs=30, bias=0.1)
    rng = np.random.default_rng(4)
    (ang, t), T_gt = W.random_euler(rng, 5.0, 3.0, fixed.center())
    moving = W.resample_linear(fixed, T_gt)
    df = hip.DeviceImage.upload(fixed.data, spacing=fixed.spacing, origin=fixed.origin)
    dm = hip.DeviceImage.upload(moving.data, spacing=moving.spacing, origin=moving.origin)
    cfg = hip.registration_config(metric=hip.MATTES, bins=50, sampling=2, sampling_percentage=0.2,
                                  seed=3, optimizer=hip.REGULAR_STEP, minimum_step=1e-4,
                                  iterations=200, learning_rate_mode=hip.LR_ONCE,
                                  shrink_factors=((4, 4, 4), (2, 2, 2), (1, 1, 1)),
                                  sigmas=(2.0, 1.0, 0.0))
    params, res = hip.register(df, dm, cfg, hip.EULER3D, [0.0] * 6, center=fixed.center())
    est = W.AffineMap(W.euler_matrix(*params[:3]), params[3:], fixed.center())
    err = W.tre(est, T_gt.inverse(), fixed)
    assert err < 0.5, (err, params, res.iterations)


@pytest.mark.parametrize("factors", [(2, 2, 2), (4, 4, 4), (3, 2, 5)])
def test_nearest_placement_is_bit_exact(hip, factors):
    """resample_nearest onto a shrunk grid (crates/sitk/src/cuda/ops/
    pyramid.rs:739-790) equals the oracle's nearest-neighbour resample
    (transform/resample.rs:226-290, round half up) bit for bit, ties of the
    even shrink factors included."""
    v = O.synth(31, 37 * 29 * 24).reshape(24, 29, 37)
    sp, org = (0.9, 1.1, 1.5), (-3.0, 2.0, 1.0)
    img = hip.DeviceImage.upload(v, spacing=sp, origin=org)
    grid = hip.shrunk_geometry(img.geometry(), factors)
    got = hip.resample_nearest(img, grid).to_host()
    ident = O.make_transform(O.AFFINE, [1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0])
    want = O.resample_nearest(v, sp, org, list(grid.size), list(grid.spacing), list(grid.origin),
                              ident)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_nearest_resample_through_a_transform_is_bit_exact(hip):
    v = O.synth(6, 40 * 30 * 20).reshape(20, 30, 40)
    sp, org = (1.0, 0.8, 1.2), (1.0, -2.0, 3.0)
    img = hip.DeviceImage.upload(v, spacing=sp, origin=org)
    t = O.make_transform(O.EULER3D, [0.1, -0.05, 0.2, 1.5, -2.0, 0.7], center=[20, 10, 15])
    st = O.transform_stage(t).reshape(1, 12)
    out_geom = hip.Geometry.make((33, 25, 17), (1.1, 1.0, 1.3), (0.5, -1.0, 2.5))
    got = hip.resample_nearest(img, out_geom, st, default_value=-7.0).to_host()
    want = O.resample_nearest(v, sp, org, (33, 25, 17), (1.1, 1.0, 1.3), (0.5, -1.0, 2.5), t,
                              default=-7.0)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("factors", [(2, 2, 2), (1, 1, 1), (3, 2, 4)])
def test_mask_placement_is_the_thresholded_nearest_resample(hip, factors):
    """A mask placed on a level grid is the nearest-neighbour resample of the
    0/1 volume with default 0, thresholded (level_mask_on_device,
    crates/sitk/src/registration/method.rs:3068-3110)."""
    rng = np.random.default_rng(9)
    raw = rng.integers(0, 3, size=(21, 26, 31)).astype(np.uint8) * 3
    sp, org = (0.9, 1.1, 1.5), (-3.0, 2.0, 1.0)
    mask = hip.DeviceMask.upload(raw, spacing=sp, origin=org)
    img = hip.DeviceImage.upload(raw.astype(np.float32), spacing=sp, origin=org)
    grid = hip.shrunk_geometry(img.geometry(), factors)
    t = O.make_transform(O.EULER3D, [0.05, -0.02, 0.03, 0.7, -0.4, 0.2], center=[12, 14, 15])
    st = O.transform_stage(t).reshape(1, 12)
    got = hip.mask_resample_nearest(mask, grid, st).to_host(tuple(reversed(list(grid.size))))
    want = O.resample_nearest((raw != 0).astype(np.float32), sp, org, list(grid.size),
                              list(grid.spacing), list(grid.origin), t)
    assert np.array_equal(got, (want != 0).astype(np.uint8))


def _masks(n):
    fm = np.zeros((n, n, n), np.uint8)
    fm[3:n - 4, 2:n - 2, 4:n - 3] = 1
    fm[10:14, 10:14, 10:14] = 0
    mm = np.ones((n, n, n), np.uint8)
    mm[:, :, :5] = 0
    mm[n - 3:, :, :] = 0
    return fm, mm


def test_masked_registration_matches_the_host_loop(hip):
    """A fixed and a moving mask through the device driver
    (set_metric_fixed_mask / set_metric_moving_mask, method.rs:1555-1590,
    2760-2785): F64, one level, the same iterates as the host loop on the
    oracle's masked problem (iteration count, stop reason, parameters)."""
    n = 32
    sp, org = [0.8, 0.9, 1.1], [-3.0, 2.0, 1.0]
    f = O.fixture_volume(n, [0, 0, 0])
    m = O.fixture_inverted(n, [-2.0, 1.0, 0.0])
    fm, mm = _masks(n)
    pb = O.MetricProblem(f, sp, org, m, sp, org, fixed_mask=fm, moving_mask=mm)
    g = pb.mattes_geometry(50)
    want, want_it, want_stop = register_rsgd(pb, g, O.TRANSLATION, [0.0, 0.0, 0.0], [0, 0, 0],
                                             (n, n, n), sp, org, minimum_step=0.01,
                                             iterations=200, grad_tol=1e-6)
    df = hip.DeviceImage.upload(f, spacing=sp, origin=org)
    dm = hip.DeviceImage.upload(m, spacing=sp, origin=org)
    dfm = hip.DeviceMask.upload(fm, spacing=sp, origin=org)
    dmm = hip.DeviceMask.upload(mm, spacing=sp, origin=org)
    cfg = hip.registration_config(metric=hip.MATTES, bins=50, precision=1,
                                  optimizer=hip.REGULAR_STEP, minimum_step=0.01, iterations=200,
                                  gradient_tolerance=1e-6, learning_rate_mode=hip.LR_ONCE,
                                  fixed_mask=dfm, moving_mask=dmm)
    got, res = hip.register(df, dm, cfg, hip.TRANSLATION, [0.0, 0.0, 0.0])
    plain, _ = hip.register(df, dm, hip.registration_config(
        metric=hip.MATTES, bins=50, precision=1, optimizer=hip.REGULAR_STEP, minimum_step=0.01,
        iterations=200, gradient_tolerance=1e-6, learning_rate_mode=hip.LR_ONCE),
        hip.TRANSLATION, [0.0, 0.0, 0.0])
    assert np.max(np.abs(got - plain)) > 1e-6  # the masks change the problem
    assert res.iterations == want_it, (res.iterations, want_it, want_stop)
    assert hip.STOP_NAMES[res.level_stop[0]] == want_stop
    assert np.max(np.abs(got - want)) < 1e-9, (got, want)


@pytest.mark.parametrize("sampling", [0, 2])
def test_a_level_places_the_fixed_mask_on_its_grid(hip, sampling):
    """On a shrunk level the fixed mask is the nearest-neighbour resample of
    the user's mask onto the level grid (level_mask_on_device, method.rs:
    3068-3110), random draws are filtered by it (metric.rs:494-580), and the
    moving mask is used at full resolution: a zero-iteration registration
    reports the oracle's value and valid count on that level."""
    n = 40
    sp, org = [1.0, 0.9, 1.2], [-2.0, 1.0, 0.5]
    f = O.fixture_volume(n, [0, 0, 0])
    m = O.fixture_inverted(n, [-1.0, 0.5, 0.0])
    fm, mm = _masks(n)
    df = hip.DeviceImage.upload(f, spacing=sp, origin=org)
    dm = hip.DeviceImage.upload(m, spacing=sp, origin=org)
    cfg = hip.registration_config(metric=hip.MATTES, bins=30, sampling=sampling,
                                  sampling_percentage=0.5, seed=7, precision=1, iterations=0,
                                  shrink_factors=((2, 2, 2),), sigmas=(0.0,),
                                  fixed_mask=hip.DeviceMask.upload(fm, spacing=sp, origin=org),
                                  moving_mask=hip.DeviceMask.upload(mm, spacing=sp, origin=org))
    p0 = [0.3, -0.2, 0.1]
    _, res = hip.register(df, dm, cfg, hip.TRANSLATION, p0)
    gsize, gsp, gorg = O.shrunk_geometry((n, n, n), sp, org, (2, 2, 2))
    ident = O.make_transform(O.TRANSLATION, [0.0, 0.0, 0.0])
    fl = O.resample_linear(f, sp, org, gsize, gsp, gorg, ident)
    lm = O.resample_nearest(fm.astype(np.float32), sp, org, gsize, gsp, gorg, ident) != 0
    idx = O.draw_samples(fl.size, 2, 0.5, 7) if sampling == 2 else None
    pb = O.MetricProblem(fl, gsp, gorg, m, sp, org, indices=idx, fixed_mask=lm.astype(np.uint8),
                         moving_mask=mm)
    wv, _, wn = pb.mattes_evaluate(pb.mattes_geometry(30), O.make_transform(O.TRANSLATION, p0))
    assert res.valid_points == wn
    assert abs(res.value - wv) <= 1e-9 * abs(wv), (res.value, wv)


def test_mask_size_mismatch_is_refused(hip):
    f = O.fixture_volume(16, [0, 0, 0])
    df = hip.DeviceImage.upload(f)
    bad = hip.DeviceMask.upload(np.ones((16, 16, 15), np.uint8))
    for kw in (dict(fixed_mask=bad), dict(moving_mask=bad)):
        cfg = hip.registration_config(metric=hip.MATTES, bins=20, iterations=1, **kw)
        with pytest.raises(hip.CudaError) as e:
            hip.register(df, df, cfg, hip.TRANSLATION, [0.0, 0.0, 0.0])
        assert e.value.status == "MaskSizeMismatch"
