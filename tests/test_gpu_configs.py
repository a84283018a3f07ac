"""The benchmark workloads at their full sizes against the oracle.

Each configuration the benchmark runs is checked here at the size it runs at,
not only on small fixtures:

- the CPU-runnable rigid pair (64^3, every voxel, Euler3D, 50 bins, RSGD with
  physical-shift scales, 100 iterations): the device registration against the
  host restatement of the reference's RSGD loop driving the oracle's metric
  (crates/sitk/src/registration/optimizer.rs:353-450, method.rs:3140-3529);
- the single large pair (256x256x160, 0.9x0.9x1.5 mm, bias field, affine,
  50 bins, 20% draws): value and derivative at every pyramid level (the coarse
  levels through the device's smoothing and placement, each compared bit for
  bit with the oracle's), and the three-level line-search registration;
- the 512-pair rigid batch (edges 32..128, half mean squares on every voxel,
  half Mattes 32 bins at 25%): every pair's batched evaluation equals its
  single-pair evaluation, and four pairs against the oracle;
- the deformable workload at 192^3 (10% draws, 10^3 control points, two
  starts with rigid moving-initial stages) against the oracle's composed
  sparse evaluation.

Tolerances are the single-pair bar: F32 value 1e-6 relative and derivative
1e-4 relative (vector norm), F64 value 1e-9 and derivative 1e-7, valid counts
exact. Every measured error is appended to $SITKHIP_ERROR_LOG when it is set.
"""

import numpy as np
import pytest

import oracle as O
from helpers import affine_form, record_error, rel_vec, stages_of
from oracle.registration import register_rsgd
from test_gpu_metric import contract

pytestmark = pytest.mark.gpu

TOL = {0: (1e-6, 1e-4), 1: (1e-9, 1e-7)}
PREC = {0: "f32", 1: "f64"}


def _upload(hip, vol):
    return hip.DeviceImage.upload(vol.data, spacing=vol.spacing, origin=vol.origin)


def _affine(A, offset):
    return O.make_transform(O.AFFINE, list(np.asarray(A).reshape(9)) + list(offset))


# ------------------------------------------------------------- large pair
@pytest.fixture(scope="module")
def large_pair():
    import workload as W

    fixed, moving, T_gt = W.config_large_pair(seed=2)
    inv = T_gt.inverse()
    poses = {"identity": _affine(np.eye(3), np.zeros(3)),
             "solution": _affine(inv.A, inv.offset)}
    return fixed, moving, T_gt, poses


@pytest.fixture(scope="module")
def large_finest_oracle(large_pair):
    fixed, moving, _, poses = large_pair
    idx = O.draw_samples(fixed.data.size, 2, 0.2, 7)
    pb = O.MetricProblem(fixed.data, fixed.spacing, fixed.origin, moving.data, moving.spacing,
                         moving.origin, indices=idx)
    g = pb.mattes_geometry(50)
    return idx, {name: pb.mattes_evaluate(g, t) for name, t in poses.items()}


@pytest.mark.parametrize("precision", [0, 1])
def test_large_pair_finest_level_against_the_oracle(hip, large_pair, large_finest_oracle,
                                                    precision):
    fixed, moving, _, poses = large_pair
    idx, want = large_finest_oracle
    assert idx.size == 2_097_152
    d = hip.ResidentMattes(_upload(hip, fixed), _upload(hip, moving), bins=50, indices=idx,
                           precision=precision)
    tv, td = TOL[precision]
    for name, t in poses.items():
        wv, wd, wn = want[name]
        v, mom, n = d.value_and_moments(stages_of(t))
        ev, ed = abs(v - wv) / abs(wv), rel_vec(contract(mom, t), wd)
        record_error("large pair, finest level (256x256x160, 2097152 draws, 50 bins, affine)",
                     precision=PREC[precision], pose=name, value_rel=ev, derivative_rel=ed,
                     valid=n)
        assert n == wn, (name, n, wn)
        assert ev <= tv, (name, v, wv)
        assert ed < td, (name, ed)


@pytest.mark.parametrize("level", [0, 1])
def test_large_pair_coarse_levels_against_the_oracle(hip, large_pair, level):
    """Level 0 (shrink 4, sigma 2 mm) and level 1 (shrink 2, sigma 1 mm): both
    volumes smoothed at full resolution, the fixed placed on the shrunk grid,
    the moving kept at full size (method.rs:2416-2569); the device's images
    are the oracle's bit for bit, and the metric matches at the bar."""
    fixed, moving, _, poses = large_pair
    shrink, sigma = [(4, 2.0), (2, 1.0)][level]
    size, sp, org = fixed.size, fixed.spacing, fixed.origin
    shape = fixed.data.shape
    fs = O.smoothing_recursive_gaussian(fixed.data, size, sp, [sigma] * 3, threads=8).reshape(shape)
    ms = O.smoothing_recursive_gaussian(moving.data, size, sp, [sigma] * 3, threads=8).reshape(shape)
    gsize, gsp, gorg = O.shrunk_geometry(size, sp, org, (shrink,) * 3)
    ident = _affine(np.eye(3), np.zeros(3))
    fl = O.resample_linear(fs, sp, org, gsize, gsp, gorg, ident)

    dfs = hip.recursive_gaussian(_upload(hip, fixed), [sigma] * 3)
    dms = hip.recursive_gaussian(_upload(hip, moving), [sigma] * 3)
    grid = hip.shrunk_geometry(dfs.geometry(), (shrink,) * 3)
    assert list(grid.size) == gsize
    dfl = hip.resample_linear(dfs, grid)
    assert np.array_equal(dms.to_host().reshape(-1).view(np.uint32), ms.reshape(-1).view(np.uint32))
    assert np.array_equal(dfl.to_host().reshape(-1).view(np.uint32), fl.reshape(-1).view(np.uint32))

    idx = O.draw_samples(fl.size, 2, 0.2, 7 + level)
    d = hip.ResidentMattes(dfl, dms, bins=50, indices=idx)
    pb = O.MetricProblem(fl, gsp, gorg, ms, sp, org, indices=idx)
    g = pb.mattes_geometry(50)
    tv, td = TOL[0]
    for name, t in poses.items():
        wv, wd, wn = pb.mattes_evaluate(g, t)
        v, mom, n = d.value_and_moments(stages_of(t))
        ev, ed = abs(v - wv) / abs(wv), rel_vec(contract(mom, t), wd)
        record_error(f"large pair, level {level} (shrink {shrink}, sigma {sigma} mm, "
                     f"{idx.size} draws, 50 bins, affine)", precision="f32", pose=name,
                     value_rel=ev, derivative_rel=ed, valid=n)
        assert n == wn, (name, n, wn)
        assert ev <= tv, (name, v, wv)
        assert ed < td, (name, ed)


def test_large_pair_line_search_registration_recovers_the_warp(hip, large_pair):
    import workload as W

    fixed, moving, T_gt, _ = large_pair
    cfg = hip.registration_config(metric=hip.MATTES, bins=50, sampling=hip.SAMPLING_RANDOM,
                                  sampling_percentage=0.2, seed=1, optimizer=hip.LINE_SEARCH,
                                  iterations=100, learning_rate_mode=hip.LR_ONCE,
                                  shrink_factors=((4, 4, 4), (2, 2, 2), (1, 1, 1)),
                                  sigmas=(2.0, 1.0, 0.0))
    ident = list(np.eye(3).reshape(9)) + [0.0, 0.0, 0.0]
    p, res = hip.register(_upload(hip, fixed), _upload(hip, moving), cfg, hip.AFFINE, ident,
                          center=fixed.center())
    est = W.AffineMap(np.asarray(p[:9]).reshape(3, 3), np.asarray(p[9:]), fixed.center())
    err = W.tre(est, T_gt.inverse(), fixed)
    record_error("large pair registration (3 levels, line search)", precision="f32", tre_mm=err,
                 iterations=int(res.iterations))
    assert err <= 0.5, (err, res.iterations)


# ----------------------------------------------------------- CPU rigid pair
@pytest.fixture(scope="module")
def cpu_pair():
    import workload as W

    fixed, moving, T_gt = W.config_cpu_pair()
    pb = O.MetricProblem(fixed.data, fixed.spacing, fixed.origin, moving.data, moving.spacing,
                         moving.origin)
    g = pb.mattes_geometry(50)
    c = list(fixed.center())
    want = register_rsgd(pb, g, O.EULER3D, [0.0] * 6, c, fixed.size, fixed.spacing, fixed.origin,
                         minimum_step=1e-4, iterations=100, grad_tol=1e-6)
    return fixed, moving, T_gt, want


@pytest.mark.parametrize("precision", [1, 0])
def test_cpu_pair_registration_matches_the_oracle_driven_loop(hip, cpu_pair, precision):
    """64^3, every voxel (262,144 samples), 50 bins, Euler3D about the centre,
    RSGD with physical-shift scales and a once-estimated step, minimum step
    1e-4, gradient tolerance 1e-6, at most 100 iterations (it stops on the
    minimum step after about 20): F64 takes the same
    iterates as the host loop on the oracle (iteration count and stop reason
    equal, parameters within 1e-9); F32 lands within 1e-3 of it."""
    import workload as W

    fixed, moving, T_gt, (want, want_it, want_stop) = cpu_pair
    cfg = hip.registration_config(metric=hip.MATTES, bins=50, precision=precision,
                                  optimizer=hip.REGULAR_STEP, minimum_step=1e-4, iterations=100,
                                  gradient_tolerance=1e-6, learning_rate_mode=hip.LR_ONCE)
    got, res = hip.register(_upload(hip, fixed), _upload(hip, moving), cfg, hip.EULER3D,
                            [0.0] * 6, center=fixed.center())
    tre = {}
    for name, p in (("device", got), ("oracle", want)):
        est = W.AffineMap(W.euler_matrix(*p[:3]), np.asarray(p[3:]), fixed.center())
        tre[name] = W.tre(est, T_gt.inverse(), fixed)
    dev = float(np.max(np.abs(got - want)))
    record_error("CPU rigid pair registration (64^3, every voxel, 50 bins, RSGD)",
                 precision=PREC[precision], parameter_max_abs=dev, iterations=int(res.iterations),
                 oracle_iterations=int(want_it), tre_mm=tre["device"], oracle_tre_mm=tre["oracle"])
    assert tre["oracle"] <= 0.5 and tre["device"] <= 0.5, tre
    if precision == 1:
        assert res.iterations == want_it, (res.iterations, want_it, want_stop)
        assert hip.STOP_NAMES[res.level_stop[0]] == want_stop
        assert dev < 1e-9, (got, want)
    else:
        assert dev < 1e-3, (got, want)


# --------------------------------------------------------- 512-pair batch
def _batch_metrics(hip, pairs):
    out = []
    for i, p in enumerate(pairs):
        f = _upload(hip, p["fixed"])
        m = hip.resample_linear(f, f.geometry(), p["T_gt"].stage().reshape(1, 12))
        if p["metric"] == "mattes":
            idx = hip.draw_samples(p["fixed"].data.size, hip.SAMPLING_RANDOM, 0.25, 17 + i)
            met = hip.ResidentMattes(f, m, bins=32, indices=idx)
        else:
            idx = None
            met = hip.ResidentMetric(f, m)
        out.append(dict(metric=met, moving=m, idx=idx))
    return out


def test_batched_workload_at_full_size(hip):
    """All 512 pairs of the batched rigid workload at the finest level, each at
    a pose near its solution: the batch gives every pair its single-pair
    result (Mattes value and valid count exactly, moments within 1e-12 of
    their largest magnitude), and two mean-squares and two Mattes pairs match
    the oracle at the single-pair bar."""
    import workload as W

    pairs = W.config_batched_pairs(512, seed=3)
    mets = _batch_metrics(hip, pairs)
    rng = np.random.default_rng(29)
    trs = []
    for p in pairs:
        inv = p["T_gt"].inverse()
        _, P = W.random_euler(rng, 2.0, 2.0, p["fixed"].center())
        A, off = inv.A @ P.A, inv.A @ P.offset + inv.offset
        t = O.make_transform(O.AFFINE, list(A.reshape(9)) + list(off))
        trs.append(t)
    stages = np.stack([stages_of(t)[0] for t in trs])
    batch = hip.Batch([m["metric"] for m in mets])
    vals, mom, valid = batch.evaluate(stages)
    worst = 0.0
    for i, (p, m) in enumerate(zip(pairs, mets)):
        st = stages[i].reshape(1, 12)
        if p["metric"] == "mattes":
            v1, m1, n1 = m["metric"].value_and_moments(st)
            assert vals[i] == v1, (i, vals[i], v1)
            m1 = np.concatenate([m1, [0.0, 0.0]])
        else:
            m1 = m["metric"].evaluate(st)
            n1 = int(m1[13])
            assert abs(vals[i] - m1[0] / n1) <= 1e-12 * abs(vals[i]), i
        assert int(valid[i]) == n1, i
        scale = max(np.max(np.abs(m1)), 1e-300)
        err = float(np.max(np.abs(mom[i] - m1)) / scale)
        worst = max(worst, err)
        assert err <= 1e-12, (i, err)
    record_error("rigid batch (512 pairs): batched vs single-pair", moments_rel_max=worst)

    checked = {"ms": 0, "mattes": 0}
    for i, (p, m) in enumerate(zip(pairs, mets)):
        if checked[p["metric"]] >= 2:
            continue
        checked[p["metric"]] += 1
        v = p["fixed"]
        mov = m["moving"].to_host()
        pb = O.MetricProblem(v.data, v.spacing, v.origin, mov, v.spacing, v.origin,
                             indices=m["idx"])
        t = trs[i]
        if p["metric"] == "mattes":
            g = pb.mattes_geometry(32)
            wv, wd, wn = pb.mattes_evaluate(g, t)
            got_d = contract(mom[i][:12], t)
            got_v = vals[i]
        else:
            wv, wd, wn = pb.ms_evaluate(t)
            j0, c = affine_form(t)
            got_v, got_d, _ = hip.contract_ms(mom[i], j0, c, t.nparams)
        ev, ed = abs(got_v - wv) / abs(wv), rel_vec(got_d, wd)
        record_error(f"rigid batch pair {i} ({p['metric']}, {v.size[0]}x{v.size[1]}x{v.size[2]})",
                     precision="f32", value_rel=ev, derivative_rel=ed, valid=int(valid[i]))
        assert int(valid[i]) == wn, i
        assert ev <= 1e-6, (i, got_v, wv)
        assert ed < 1e-4, (i, ed)


def _tiny_pairs(hip, n, seed):
    import workload as W

    rng = np.random.default_rng(seed)
    fixed, moving, configs, kinds, p0, centres = [], [], [], [], [], []
    for i in range(n):
        shp = tuple(int(v) for v in rng.integers(14, 22, 3))
        vol = W.blob_volume(shp, (1.0, 1.0, 1.0), seed=300 + i, nblobs=10, sigma=(2.0, 4.0))
        _, T_gt = W.random_euler(rng, 3.0, 1.5, vol.center())
        mov = W.resample_linear(vol, T_gt)
        fixed.append(_upload(hip, vol))
        moving.append(_upload(hip, mov))
        mattes = i % 2 == 1
        configs.append(hip.registration_config(
            metric=hip.MATTES if mattes else hip.MEAN_SQUARES, bins=16,
            sampling=hip.SAMPLING_RANDOM if mattes else hip.SAMPLING_NONE,
            sampling_percentage=0.5, seed=11 + i, minimum_step=1e-3, iterations=30,
            learning_rate_mode=hip.LR_ONCE))
        kinds.append(hip.EULER3D)
        p0.append([0.0] * 6)
        centres.append(vol.center())
    return fixed, moving, configs, kinds, p0, centres


def test_two_engine_batched_registration_matches_one_at_a_time(hip):
    """From 64 pairs on, register_batch splits a level's pairs over two engines
    that take turns on the device; every pair must still register exactly as
    it does alone (parameters, iteration count, stop reason, valid points)."""
    fixed, moving, configs, kinds, p0, centres = _tiny_pairs(hip, 67, seed=41)
    got, res = hip.register_batch(fixed, moving, configs, kinds, p0, centres)
    for i in range(len(fixed)):
        want, r1 = hip.register(fixed[i], moving[i], configs[i], kinds[i], p0[i], center=centres[i])
        assert np.max(np.abs(got[i] - want)) < 1e-6, (i, got[i], want)
        assert res[i].iterations == r1.iterations, (i, res[i].iterations, r1.iterations)
        assert res[i].level_stop[0] == r1.level_stop[0], i
        assert res[i].valid_points == r1.valid_points, i


def test_autotuned_batched_registration_matches_one_at_a_time(hip):
    """A batch large enough for register_batch's autotuner to run (more than
    2^22 samples per engine): the tuned tiling changes only the order of the
    FP64 moment sums, so each pair lands where it does alone."""
    import workload as W

    rng = np.random.default_rng(8)
    fixed, moving, configs, kinds, p0, centres = [], [], [], [], [], []
    for i, shp in enumerate([(128, 128, 128), (128, 128, 128), (96, 80, 72)]):
        vol = W.blob_volume(shp, (1.0, 1.0, 1.0), seed=500 + i, nblobs=30)
        _, T_gt = W.random_euler(rng, 4.0, 2.0, vol.center())
        fixed.append(_upload(hip, vol))
        moving.append(hip.resample_linear(fixed[-1], fixed[-1].geometry(),
                                          T_gt.stage().reshape(1, 12)))
        mattes = i == 2
        configs.append(hip.registration_config(
            metric=hip.MATTES if mattes else hip.MEAN_SQUARES, bins=32,
            sampling=hip.SAMPLING_RANDOM if mattes else hip.SAMPLING_NONE,
            sampling_percentage=0.25, seed=3 + i, minimum_step=1e-3, iterations=15,
            learning_rate_mode=hip.LR_ONCE))
        kinds.append(hip.EULER3D)
        p0.append([0.0] * 6)
        centres.append(vol.center())
    hip.profile_reset()
    hip.profile_enable(True)
    try:
        got, res = hip.register_batch(fixed, moving, configs, kinds, p0, centres, autotune=True)
    finally:
        hip.profile_enable(False)
    tuned, _ = hip.profile_query("batch_autotune")
    assert tuned >= 1
    for i in range(len(fixed)):
        want, r1 = hip.register(fixed[i], moving[i], configs[i], kinds[i], p0[i], center=centres[i])
        assert np.max(np.abs(got[i] - want)) < 1e-6, (i, got[i], want)
        assert res[i].iterations == r1.iterations, (i, res[i].iterations, r1.iterations)
        assert res[i].valid_points == r1.valid_points, i


# ------------------------------------------------------------ deformable
@pytest.mark.parametrize("precision", [0, 1])
def test_deformable_workload_at_full_size_against_the_oracle(hip, precision):
    """The deformable workload at its benchmark size: 192^3, 10% draws
    (707,788 samples), 32 bins, a 10^3-point control grid (3000 parameters),
    two starts, each a B-spline followed by its own rigid moving-initial stage,
    in one batched launch, against the oracle's composed sparse evaluation."""
    import workload as W

    f, m, _ = W.config_deformable()
    idx = O.draw_samples(f.data.size, 2, 0.10, 1)
    assert idx.size == 707_788
    met = hip.ResidentMattes(_upload(hip, f), _upload(hip, m), bins=32, indices=idx,
                             precision=precision)
    mesh = (7, 7, 7)
    gs, go, gsp = W.bspline_grid(f, mesh)
    bs = hip.BSplineMattes(met, list(gs), list(go), list(gsp))
    assert bs.nparams == 3000
    rng = np.random.default_rng(13)
    coeffs = np.stack([rng.normal(0.0, 0.8, bs.nparams) for _ in range(2)])
    rig = [O.make_transform(O.EULER3D, list(rng.uniform(-0.05, 0.05, 3)) + list(rng.uniform(-3, 3, 3)),
                            center=list(f.center())) for _ in range(2)]
    multi = hip.BSplineMulti(bs, 2, np.stack([O.transform_stage(r) for r in rig]))
    vals, ders, valid = multi.value_and_derivative(coeffs)
    pb = O.MetricProblem(f.data, f.spacing, f.origin, m.data, m.spacing, m.origin, indices=idx)
    g = pb.mattes_geometry(32)
    tv, td = TOL[precision]
    for s in range(2):
        t = O.make_bspline(f.size, f.spacing, f.origin, np.eye(3), mesh, coeffs[s])
        wv, wd, wn = pb.mattes_evaluate_composed(g, t, rig[s])
        ev, ed = abs(vals[s] - wv) / abs(wv), rel_vec(ders[s], wd)
        record_error("deformable (192^3, 707788 draws, 32 bins, 3000 B-spline parameters + rigid "
                     "stage)", precision=PREC[precision], start=s, value_rel=ev, derivative_rel=ed,
                     valid=int(valid[s]))
        assert valid[s] == wn, s
        assert ev <= tv, (s, vals[s], wv)
        assert ed < td, (s, ed)
