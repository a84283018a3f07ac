"""Batched evaluation of many resident metrics, and the batched registration
driver.

Oracle for the batch (it has no reference counterpart): each pair's result
equals the single-pair evaluation of the same metric -- the Mattes value and
valid count exactly (the histogram is an exact integer sum whatever the
tiling), the FP64 moments within 1e-12 relative to their largest magnitude
(only the order of their sum differs). One pair is also checked against the CPU
oracle at the single-pair bar (value 1e-6 relative, derivative 1e-4).
"""

import numpy as np
import pytest

import oracle as O
from helpers import blob_volume, rel_vec, stages_of
from test_gpu_metric import contract

pytestmark = pytest.mark.gpu


def _euler_stage(rng, centre, deg=8.0, mm=3.0):
    ang = np.deg2rad(rng.uniform(-deg, deg, 3))
    t = O.make_transform(O.EULER3D, list(ang) + list(rng.uniform(-mm, mm, 3)), center=centre)
    return t, stages_of(t)[0]


def _pairs(hip, precision=0):
    """Uneven pairs: (kind, shape, spacing, sampling fraction, bins)."""
    specs = [
        (hip.MATTES, (20, 24, 28), (1.0, 1.1, 0.9), 0.25, 32),
        (hip.MEAN_SQUARES, (16, 30, 22), (0.9, 1.0, 1.2), None, None),
        (hip.MATTES, (33, 17, 40), (1.2, 0.8, 1.0), 0.5, 50),
        (hip.MEAN_SQUARES, (24, 24, 24), (1.0, 1.0, 1.0), 0.3, None),
        (hip.MATTES, (41, 35, 19), (0.8, 0.9, 1.1), None, 20),
    ]
    out = []
    for i, (kind, shape, sp, frac, bins) in enumerate(specs):
        f = blob_volume(shape, 100 + i, spacing=sp)
        m = blob_volume(shape, 200 + i, spacing=sp) * 0.5 + f * 0.5
        org = (-2.0 + i, 1.0, 0.5 * i)
        df = hip.DeviceImage.upload(f, spacing=sp, origin=org)
        dm = hip.DeviceImage.upload(m, spacing=sp, origin=org)
        idx = None
        if frac is not None:
            idx = hip.draw_samples(f.size, hip.SAMPLING_RANDOM, frac, 7 + i)
        if kind == hip.MATTES:
            met = hip.ResidentMattes(df, dm, bins=bins, indices=idx, precision=precision)
        else:
            met = hip.ResidentMetric(df, dm, indices=idx, precision=precision)
        nx, ny, nz = shape[2], shape[1], shape[0]
        centre = np.asarray(org) + 0.5 * (np.array([nx, ny, nz]) - 1) * np.asarray(sp)
        out.append(dict(kind=kind, metric=met, f=f, m=m, sp=sp, org=org, idx=idx, bins=bins,
                        centre=centre))
    return out


def _single(p, stage):
    if p["kind"] == 1:  # Mattes
        v, mom, valid = p["metric"].value_and_moments(stage.reshape(1, 12))
        return v, np.concatenate([mom, [0.0, 0.0]]), valid
    mom = p["metric"].evaluate(stage.reshape(1, 12))
    count = mom[13]
    return (mom[0] / count if count else np.finfo(float).max), mom, int(count)


def _close(a, b, tol=1e-12):
    scale = max(np.max(np.abs(b)), 1e-300)
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) <= tol * scale


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_equals_single_pair_evaluations(hip, precision):
    pairs = _pairs(hip, precision)
    rng = np.random.default_rng(1)
    stages = np.stack([_euler_stage(rng, p["centre"])[1] for p in pairs])
    batch = hip.Batch([p["metric"] for p in pairs])
    for tile, grid in [(0, 0), (256, 8), (1024, 64), (4096, 2048)]:
        batch.set_tiling(tile, grid)
        vals, mom, valid = batch.evaluate(stages)
        for i, p in enumerate(pairs):
            v1, m1, n1 = _single(p, stages[i])
            assert int(valid[i]) == n1, (i, tile)
            if p["kind"] == hip.MATTES:
                assert vals[i] == v1, (i, tile, vals[i], v1)
            else:
                assert abs(vals[i] - v1) <= 1e-12 * abs(v1)
            assert _close(mom[i], m1), (i, tile, mom[i], m1)


def test_batch_pair_against_the_oracle(hip):
    pairs = _pairs(hip, 0)
    rng = np.random.default_rng(5)
    tr = [_euler_stage(rng, p["centre"]) for p in pairs]
    batch = hip.Batch([p["metric"] for p in pairs])
    vals, mom, valid = batch.evaluate(np.stack([s for _, s in tr]))
    for i in (0, 2):
        p = pairs[i]
        pb = O.MetricProblem(p["f"], p["sp"], p["org"], p["m"], p["sp"], p["org"],
                             indices=p["idx"])
        g = pb.mattes_geometry(p["bins"])
        want_v, want_d, want_n = pb.mattes_evaluate(g, tr[i][0])
        assert int(valid[i]) == want_n
        assert abs(vals[i] - want_v) <= 1e-6 * abs(want_v)
        assert rel_vec(contract(mom[i][:12], tr[i][0]), want_d) < 1e-4


def test_graph_replays_match_direct_launches(hip):
    pairs = _pairs(hip, 0)
    rng = np.random.default_rng(2)
    all_stages = [np.stack([_euler_stage(rng, p["centre"])[1] for p in pairs]) for _ in range(4)]
    batch = hip.Batch([p["metric"] for p in pairs])
    batch.set_tiling(2048, 256, use_graph=False)
    direct = [batch.evaluate(s) for s in all_stages]
    batch.set_tiling(2048, 256, use_graph=True)
    for rep in range(2):
        for s, (dv, dm, dn) in zip(all_stages, direct):
            v, m, n = batch.evaluate(s)
            assert np.array_equal(v, dv) and np.array_equal(m, dm) and np.array_equal(n, dn)
            # eager single-pair work between replays
            _single(pairs[1], s[1])


def test_inactive_pairs_are_skipped(hip):
    pairs = _pairs(hip, 0)
    rng = np.random.default_rng(3)
    stages = np.stack([_euler_stage(rng, p["centre"])[1] for p in pairs])
    batch = hip.Batch([p["metric"] for p in pairs])
    full_v, full_m, full_n = batch.evaluate(stages)
    active = np.array([1, 0, 1, 0, 1], dtype=np.uint8)
    v, m, n = batch.evaluate(stages, active)
    for i in range(len(pairs)):
        if active[i]:
            assert v[i] == full_v[i] and np.array_equal(m[i], full_m[i])
        else:
            assert np.isnan(v[i]) and not m[i].any() and n[i] == 0
    # the skipped pairs' histograms were left clean: a full evaluation agrees
    v2, m2, n2 = batch.evaluate(stages)
    assert np.array_equal(v2, full_v) and np.array_equal(m2, full_m)


def test_autotune_picks_a_valid_tiling(hip):
    pairs = _pairs(hip, 0)
    rng = np.random.default_rng(4)
    stages = np.stack([_euler_stage(rng, p["centre"])[1] for p in pairs])
    batch = hip.Batch([p["metric"] for p in pairs])
    tile, grid, ms = batch.autotune(stages)
    assert tile >= 256 and grid % 8 == 0 and ms > 0
    assert batch.tiling() == (tile, grid)
    vals, mom, valid = batch.evaluate(stages)
    for i, p in enumerate(pairs):
        v1, m1, n1 = _single(p, stages[i])
        assert int(valid[i]) == n1
        assert _close(mom[i], m1)


def test_batch_refusals(hip):
    f = blob_volume((12, 12, 12), 1)
    df = hip.DeviceImage.upload(f)
    a = hip.ResidentMattes(df, df, bins=None)
    with pytest.raises(hip.CudaError, match="HistogramShape"):
        hip.Batch([a])
    b = hip.ResidentMattes(df, df, bins=80)
    with pytest.raises(hip.CudaError, match="TooManyHistogramBins"):
        hip.Batch([b])
    c = hip.ResidentMetric(df, df, precision=hip.PRECISION_F64)
    d = hip.ResidentMetric(df, df, precision=hip.PRECISION_F32)
    with pytest.raises(hip.CudaError, match="InvalidArgument"):
        hip.Batch([c, d])
    cfg = hip.registration_config(optimizer=hip.LINE_SEARCH)
    with pytest.raises(hip.CudaError, match="RegularStep"):
        hip.register_batch([df], [df], [cfg], [hip.TRANSLATION], [[0, 0, 0]], [[0, 0, 0]])


def test_batched_registration_matches_one_at_a_time(hip):
    import workload as W

    rng = np.random.default_rng(11)
    fixed, moving, configs, kinds, p0, centres = [], [], [], [], [], []
    shapes = [(28, 32, 24), (24, 26, 30), (32, 28, 28), (26, 30, 22)]
    for i, shp in enumerate(shapes):
        vol = W.blob_volume(shp, (1.0, 1.0, 1.0), seed=40 + i, nblobs=20)
        (ang, t), T_gt = W.random_euler(rng, 4.0, 2.0, vol.center())
        mov = W.resample_linear(vol, T_gt)
        fixed.append(hip.DeviceImage.upload(vol.data, spacing=vol.spacing, origin=vol.origin))
        moving.append(hip.DeviceImage.upload(mov.data, spacing=mov.spacing, origin=mov.origin))
        metric = hip.MATTES if i % 2 == 0 else hip.MEAN_SQUARES
        configs.append(hip.registration_config(metric=metric, bins=32, sampling=2 if i % 2 == 0 else 0,
                                               sampling_percentage=0.25, seed=9 + i,
                                               minimum_step=1e-3, iterations=60,
                                               learning_rate_mode=hip.LR_ONCE))
        kinds.append(hip.EULER3D)
        p0.append([0.0] * 6)
        centres.append(vol.center())
    got, res = hip.register_batch(fixed, moving, configs, kinds, p0, centres)
    for i in range(len(shapes)):
        want, r1 = hip.register(fixed[i], moving[i], configs[i], kinds[i], p0[i], center=centres[i])
        assert np.max(np.abs(got[i] - want)) < 1e-6, (i, got[i], want)
        assert res[i].iterations == r1.iterations, (i, res[i].iterations, r1.iterations)
        assert res[i].level_stop[0] == r1.level_stop[0]
        assert res[i].valid_points == r1.valid_points


def test_batched_pyramid_registrations_match_one_at_a_time(hip):
    """Twelve pairs with two-level pyramids (shrink 2 and sigma 1, then full
    resolution): every level of every pair is set up by one launch per kind of
    work (placement, draws, gathers, ranges, fixed bins), so this checks that
    set-up against the one-at-a-time path for every sampling mode (device
    draws, a regular grid, every voxel)."""
    import workload as W

    rng = np.random.default_rng(17)
    fixed, moving, configs, kinds, p0, centres = [], [], [], [], [], []
    for i in range(12):
        shp = (22 + 2 * (i % 4), 26 - (i % 3), 20 + 3 * (i % 2))
        vol = W.blob_volume(shp, (1.0, 1.0, 1.0), seed=90 + i, nblobs=16)
        (ang, t), T_gt = W.random_euler(rng, 4.0, 2.0, vol.center())
        mov = W.resample_linear(vol, T_gt)
        fixed.append(hip.DeviceImage.upload(vol.data, spacing=vol.spacing, origin=vol.origin))
        moving.append(hip.DeviceImage.upload(mov.data, spacing=mov.spacing, origin=mov.origin))
        metric = hip.MATTES if i % 3 != 1 else hip.MEAN_SQUARES
        sampling = (2, 0, 1)[i % 3]
        configs.append(hip.registration_config(metric=metric, bins=24, sampling=sampling,
                                               sampling_percentage=0.3, seed=5 + i,
                                               minimum_step=1e-3, iterations=25,
                                               learning_rate_mode=hip.LR_ONCE,
                                               shrink_factors=((2, 2, 2), (1, 1, 1)),
                                               sigmas=(1.0, 0.0)))
        kinds.append(hip.EULER3D)
        p0.append([0.0] * 6)
        centres.append(vol.center())
    got, res = hip.register_batch(fixed, moving, configs, kinds, p0, centres)
    for i in range(12):
        want, r1 = hip.register(fixed[i], moving[i], configs[i], kinds[i], p0[i], center=centres[i])
        assert np.max(np.abs(got[i] - want)) < 1e-6, (i, got[i], want)
        assert res[i].iterations == r1.iterations, (i, res[i].iterations, r1.iterations)
        assert res[i].valid_points == r1.valid_points


def test_batched_masked_registrations_match_one_at_a_time(hip):
    """Masked and unmasked pairs in one batch with two-level pyramids: a pair
    with a fixed mask is set up at once (its level mask placed by the
    nearest-neighbour resample), one with only a moving mask through the
    batched set-up (its mask copied there), the rest batched as well; every
    pair registers as it does alone."""
    import workload as W

    rng = np.random.default_rng(23)
    fixed, moving, configs, kinds, p0, centres, keep = [], [], [], [], [], [], []
    for i in range(10):
        shp = (24 + 2 * (i % 3), 28 - (i % 4), 22 + (i % 2))
        vol = W.blob_volume(shp, (1.0, 1.0, 1.0), seed=130 + i, nblobs=16)
        (ang, t), T_gt = W.random_euler(rng, 4.0, 2.0, vol.center())
        mov = W.resample_linear(vol, T_gt)
        fixed.append(hip.DeviceImage.upload(vol.data, spacing=vol.spacing, origin=vol.origin))
        moving.append(hip.DeviceImage.upload(mov.data, spacing=mov.spacing, origin=mov.origin))
        fm = mm = None
        if i % 3 == 0:  # both masks
            fm = hip.DeviceMask.upload((vol.data > np.percentile(vol.data, 20)).astype(np.uint8),
                                       spacing=vol.spacing, origin=vol.origin)
        if i % 3 != 2:  # a moving mask
            mm = hip.DeviceMask.upload((mov.data > np.percentile(mov.data, 10)).astype(np.uint8),
                                       spacing=mov.spacing, origin=mov.origin)
        keep += [fm, mm]
        configs.append(hip.registration_config(metric=hip.MATTES if i % 2 == 0 else hip.MEAN_SQUARES,
                                               bins=24, sampling=2 if i % 4 != 3 else 0,
                                               sampling_percentage=0.3, seed=3 + i,
                                               minimum_step=1e-3, iterations=25,
                                               learning_rate_mode=hip.LR_ONCE,
                                               shrink_factors=((2, 2, 2), (1, 1, 1)),
                                               sigmas=(1.0, 0.0), fixed_mask=fm, moving_mask=mm))
        kinds.append(hip.EULER3D)
        p0.append([0.0] * 6)
        centres.append(vol.center())
    got, res = hip.register_batch(fixed, moving, configs, kinds, p0, centres)
    for i in range(len(fixed)):
        want, r1 = hip.register(fixed[i], moving[i], configs[i], kinds[i], p0[i], center=centres[i])
        assert np.max(np.abs(got[i] - want)) < 1e-6, (i, got[i], want)
        assert res[i].iterations == r1.iterations, (i, res[i].iterations, r1.iterations)
        assert res[i].valid_points == r1.valid_points, i


def test_probe_values_are_the_zero_iteration_values(hip):
    """register_batch's probe: each pair's metric value at the probe
    parameters on the first prepared level equals the value a zero-iteration
    batched registration started there reports (same level, same samples;
    Mattes bit for bit, mean squares to 1e-12), and the probe does not change
    the registration itself."""
    import workload as W

    rng = np.random.default_rng(21)
    fixed, moving, configs, zero, kinds, p0, probe, centres = [], [], [], [], [], [], [], []
    for i, shp in enumerate([(28, 30, 24), (26, 24, 30), (30, 28, 26)]):
        vol = W.blob_volume(shp, (1.0, 1.0, 1.0), seed=70 + i, nblobs=20)
        (ang, t), T_gt = W.random_euler(rng, 5.0, 3.0, vol.center())
        mov = W.resample_linear(vol, T_gt)
        fixed.append(hip.DeviceImage.upload(vol.data, spacing=vol.spacing, origin=vol.origin))
        moving.append(hip.DeviceImage.upload(mov.data, spacing=mov.spacing, origin=mov.origin))
        metric = hip.MATTES if i != 1 else hip.MEAN_SQUARES
        kw = dict(metric=metric, bins=32, sampling=2 if metric == hip.MATTES else 0,
                  sampling_percentage=0.3, seed=3 + i, minimum_step=1e-3,
                  learning_rate_mode=hip.LR_ONCE)
        configs.append(hip.registration_config(iterations=20, **kw))
        zero.append(hip.registration_config(iterations=0, **kw))
        kinds.append(hip.EULER3D)
        p0.append([0.02, -0.01, 0.03, 0.5, -0.4, 0.2])
        probe.append([0.0] * 6)
        centres.append(vol.center())
    got, res, pv = hip.register_batch(fixed, moving, configs, kinds, p0, centres, probe=probe)
    got2, res2 = hip.register_batch(fixed, moving, configs, kinds, p0, centres)
    _, res0 = hip.register_batch(fixed, moving, zero, kinds, probe, centres)
    for i in range(3):
        assert np.array_equal(got[i], got2[i]) and res[i].value == res2[i].value
        if configs[i].metric == hip.MATTES:
            assert pv[i] == res0[i].value, (i, pv[i], res0[i].value)
        else:
            assert abs(pv[i] - res0[i].value) <= 1e-12 * abs(res0[i].value)
