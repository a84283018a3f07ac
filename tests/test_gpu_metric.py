"""Device Mattes and mean-squares metrics against the oracle.

Tolerances (stated per precision):
  F64 sampler : histogram bins within 1e-6 absolute (64-bit fixed point, 2^-32
                units), valid count and fixed marginal exact; value within 1e-9
                relative; derivative within 1e-7 relative (vector norm).
  F32 sampler : value within 1e-6 relative, derivative within 1e-4 relative
                (vector norm) -- the bar the north star sets for the FP32 path;
                valid count exact (every discrete decision runs in FP64).
Reference pins mirrored: crates/sitk/tests/mattes_device.rs:189-350 (values,
derivatives, valid counts), :280-305 (determinism), :357-448 (refusals);
tests/sampled_metric.rs:241-262 (a repeated index counts twice);
tests/mattes_mask.rs (masked moving range); tests/cuda_mean_squares.rs:121-214.
"""

import numpy as np
import pytest

import oracle as O
from helpers import affine_form, blob_volume, rel_vec, stages_of

pytestmark = pytest.mark.gpu

SP = [0.8, 0.9, 1.1]
ORG = [-3.0, 2.0, 1.0]

TRANSFORMS = [
    (O.TRANSLATION, [0.0, 0.0, 0.0], [0, 0, 0]),
    (O.TRANSLATION, [1.0, 0.0, 0.0], [0, 0, 0]),
    (O.TRANSLATION, [0.7, -1.3, 2.1], [0, 0, 0]),
    (O.TRANSLATION, [-3.4, 2.2, -0.9], [0, 0, 0]),
    (O.EULER3D, [0.05, -0.03, 0.08, 1.2, -0.7, 0.4], [9.0, 16.0, 18.0]),
    (O.AFFINE, [1.02, 0.03, -0.01, -0.02, 0.99, 0.04, 0.01, -0.03, 1.01, 0.6, -1.1, 0.4],
     [1.0, -2.0, 0.5]),
]


def ledger_pair():
    return O.fixture_volume(32, [0, 0, 0]), O.fixture_inverted(32, [1.5, -0.5, 0.75])


def mattes_pair(hip, f, m, precision, bins=50, indices=None, fmask=None, mmask=None,
                sp=SP, org=ORG):
    df = hip.DeviceImage.upload(f, spacing=sp, origin=org)
    dm = hip.DeviceImage.upload(m, spacing=sp, origin=org)
    kw = dict(precision=precision, indices=indices)
    if fmask is not None:
        kw["fixed_mask"] = hip.DeviceMask.upload(fmask, spacing=sp, origin=org)
    if mmask is not None:
        kw["moving_mask"] = hip.DeviceMask.upload(mmask, spacing=sp, origin=org)
    d = hip.ResidentMattes(df, dm, bins=bins, **kw)
    pb = O.MetricProblem(f, sp, org, m, sp, org, indices=indices, fixed_mask=fmask,
                         moving_mask=mmask)
    return d, pb


def contract(moments, t):
    j0, c = affine_form(t)
    out = np.zeros(t.nparams)
    for k in range(t.nparams):
        g = 0.0
        for d in range(3):
            g += j0[d, k] * moments[d]
            for e in range(3):
                g += c[e, d, k] * moments[3 + 3 * d + e]
        out[k] = g
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_value_ranges_are_the_hosts(hip, precision):
    f, m = ledger_pair()
    d, pb = mattes_pair(hip, f, m, precision)
    fr, mr = d.value_ranges()
    assert fr == pb.fixed_range() and mr == pb.moving_range()
    g = pb.mattes_geometry(50)
    assert d.bins.moving_bin_size == g.moving_bin_size
    assert d.bins.fixed_normalized_min == g.fixed_normalized_min


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("case", range(len(TRANSFORMS)))
def test_mattes_histogram_value_and_derivative(hip, precision, case):
    kind, p, ctr = TRANSFORMS[case]
    f, m = ledger_pair()
    d, pb = mattes_pair(hip, f, m, precision)
    g = pb.mattes_geometry(50)
    t = O.make_transform(kind, p, center=ctr)
    st = stages_of(t)

    joint, marg, valid = d.joint_histogram(st)
    wj, wm, wv = pb.mattes_histogram(g, t)
    assert valid == wv
    assert np.array_equal(marg, wm)
    if precision == 1:
        assert np.max(np.abs(joint - wj)) < 1e-6

    want_v, want_d, _ = pb.mattes_evaluate(g, t)
    v, mom, v_valid = d.value_and_moments(st)
    assert v_valid == wv
    deriv = contract(mom, t)
    if precision == 1:
        assert abs(v - want_v) <= 1e-9 * abs(want_v)
        assert rel_vec(deriv, want_d) < 1e-7
    else:
        assert abs(v - want_v) <= 1e-6 * abs(want_v)
        assert rel_vec(deriv, want_d) < 1e-4


@pytest.mark.parametrize("precision", [0, 1])
def test_the_ledger_value(hip, precision):
    """doc/upstream-findings.md:273: host -1.063357897140305 (naive mass) /
    -1.0633578971403057 (current Kahan mass), translation [1,0,0]."""
    f, m = ledger_pair()
    d, _ = mattes_pair(hip, f, m, precision)
    v, valid = d.value(stages_of(O.make_transform(O.TRANSLATION, [1.0, 0.0, 0.0])))
    tol = 1e-9 if precision == 1 else 1e-6
    # the x = 31 column maps past the moving buffer's edge
    assert abs(v - (-1.063357897140305)) <= tol * 1.07 and valid == 31 * 32 * 32


@pytest.mark.parametrize("precision", [0, 1])
def test_commensurate_straddles(hip, precision):
    """tests/mattes_device.rs:83-100,189-217: the identity on commensurate grids
    puts every sample exactly on a voxel and many on Parzen bin edges."""
    c = O.fixture_commensurate(32)
    d, pb = mattes_pair(hip, c, c, precision, sp=[1, 1, 1], org=[0, 0, 0])
    g = pb.mattes_geometry(50)
    for tr in ([0, 0, 0], [1.0, 0, 0], [0.7, -1.3, 2.1]):
        t = O.make_transform(O.TRANSLATION, tr)
        v, mom, valid = d.value_and_moments(stages_of(t))
        wv, wd, wn = pb.mattes_evaluate(g, t)
        assert valid == wn
        tol = 1e-9 if precision == 1 else 1e-6
        assert abs(v - wv) <= tol * abs(wv)
        assert rel_vec(contract(mom, t), wd) < (1e-7 if precision == 1 else 1e-4)


@pytest.mark.parametrize("precision", [0, 1])
def test_sampled_with_duplicates_and_masks(hip, precision):
    f, m = ledger_pair()
    idx = O.draw_samples(32 ** 3, 2, 0.25, 7)
    assert len(np.unique(idx)) < idx.size  # draws with replacement repeat voxels
    fmask = np.ones((32, 32, 32), np.uint8)
    fmask[:, :, :5] = 0
    mmask = np.ones((32, 32, 32), np.uint8)
    mmask[20:, :, :] = 0
    mmask[0, 0, 0] = 0
    for fm_, mm_ in [(None, None), (fmask, None), (None, mmask), (fmask, mmask)]:
        d, pb = mattes_pair(hip, f, m, precision, indices=idx, fmask=fm_, mmask=mm_)
        assert d.sample_count() == (idx.size if fm_ is None else int((fm_.reshape(-1)[idx] != 0).sum()))
        fr, mr = d.value_ranges()
        assert (fr, mr) == (pb.fixed_range(), pb.moving_range())
        g = pb.mattes_geometry(50)
        t = O.make_transform(O.EULER3D, [0.04, 0.02, -0.05, 0.9, 1.3, -0.2], [9.0, 16.0, 18.0])
        v, mom, valid = d.value_and_moments(stages_of(t))
        wv, wd, wn = pb.mattes_evaluate(g, t)
        assert valid == wn
        tol = 1e-9 if precision == 1 else 1e-6
        assert abs(v - wv) <= tol * abs(wv)
        assert rel_vec(contract(mom, t), wd) < (1e-7 if precision == 1 else 1e-4)


def test_a_repeated_index_counts_twice(hip):
    f, m = ledger_pair()
    base = np.arange(0, 32 ** 3, 7, dtype=np.int64)
    d1, _ = mattes_pair(hip, f, m, 1, indices=base)
    d2, _ = mattes_pair(hip, f, m, 1, indices=np.concatenate([base, base]))
    st = stages_of(O.make_transform(O.TRANSLATION, [0.3, 0.2, -0.1]))
    j1, m1, v1 = d1.joint_histogram(st)
    j2, m2, v2 = d2.joint_histogram(st)
    assert v2 == 2 * v1 and np.array_equal(m2, 2 * m1) and np.max(np.abs(j2 - 2 * j1)) < 1e-6


@pytest.mark.parametrize("precision", [0, 1])
def test_determinism(hip, precision):
    f, m = ledger_pair()
    d, _ = mattes_pair(hip, f, m, precision)
    st = stages_of(O.make_transform(O.EULER3D, [0.05, -0.03, 0.08, 1.2, -0.7, 0.4],
                                    [9.0, 16.0, 18.0]))
    first = d.value_and_moments(st)
    for _ in range(8):
        again = d.value_and_moments(st)
        assert again[0] == first[0] and np.array_equal(again[1], first[1])


def test_fused_path_equals_the_two_call_path(hip):
    """value_and_moments (device tail) equals joint_histogram -> host tail ->
    derivative_moments, the reference's own sequence (ops/mattes.rs:697-841)."""
    f, m = ledger_pair()
    d, pb = mattes_pair(hip, f, m, 1)
    g = pb.mattes_geometry(50)
    t = O.make_transform(O.AFFINE, TRANSFORMS[5][1], TRANSFORMS[5][2])
    st = stages_of(t)
    joint, marg, valid = d.joint_histogram(st)
    hv, pratio = pb.mattes_tail(joint, marg, valid, g)
    mom2 = d.derivative_moments(st, pratio)
    v, mom, _ = d.value_and_moments(st)
    assert abs(v - hv) <= 1e-12 * abs(hv)
    assert rel_vec(mom, mom2) < 1e-9


def test_refusals(hip):
    f, m = ledger_pair()
    df = hip.DeviceImage.upload(f)
    dm = hip.DeviceImage.upload(m)
    d = hip.ResidentMattes(df, dm, bins=50)
    one = stages_of(O.make_transform(O.TRANSLATION, [0, 0, 0]))
    for bad in (np.zeros((0, 12)), np.tile(one, (9, 1))):
        with pytest.raises(hip.CudaError) as e:
            d.value(bad)
        assert e.value.status == "PointMapStageCount"
    d.value(np.tile(one, (8, 1)))  # 8 stages are replayed
    with pytest.raises(hip.CudaError) as e:
        hip.ResidentMattes(df, dm, bins=50, indices=[0, 5, 32 ** 3])
    assert e.value.status == "SampleIndexOutOfGrid"
    with pytest.raises(hip.CudaError) as e:
        hip.ResidentMattes(df, dm, bins=4)
    assert e.value.status == "TooFewHistogramBins"
    nob = hip.ResidentMattes(df, dm)
    with pytest.raises(hip.CudaError) as e:
        nob.value(one)
    assert e.value.status == "HistogramShape"
    flat = hip.DeviceImage.upload(np.full((8, 8, 8), 2.0, np.float32))
    with pytest.raises(hip.CudaError) as e:
        hip.ResidentMattes(flat, dm, bins=50)
    assert e.value.status == "ConstantIntensity"


def test_multi_stage_point_map(hip):
    """Two stages (a transform composed with a moving-initial) are replayed in
    order (matrix_offset.rs:102-109)."""
    f, m = ledger_pair()
    d, pb = mattes_pair(hip, f, m, 1)
    g = pb.mattes_geometry(50)
    t1 = O.make_transform(O.EULER3D, [0.02, 0.01, -0.03, 0.5, 0.2, -0.1], [9, 16, 18])
    t2 = O.make_transform(O.TRANSLATION, [0.3, -0.4, 0.25])
    st = np.stack([O.transform_stage(t1), O.transform_stage(t2)])
    # compose on the host: an affine whose point map is t2(t1(x)) bit for bit is
    # not available, so compare against the histogram of replayed points
    v, valid = d.value(st)
    # oracle: same map expressed as a single affine (rounding may differ slightly)
    A = O.transform_stage(t1)[:9].reshape(3, 3)
    off = O.transform_stage(t1)[9:] + np.array([0.3, -0.4, 0.25])
    ta = O.make_transform(O.AFFINE, list(A.reshape(9)) + [0, 0, 0])
    ta.offset[:] = list(off)
    wv = pb.mattes_value(g, ta)
    assert abs(v - wv) <= 1e-7 * abs(wv)


# ------------------------------------------------------------- mean squares
def ms_pair(hip, f, m, precision, indices=None, fmask=None, mmask=None, sp=SP, org=ORG):
    df = hip.DeviceImage.upload(f, spacing=sp, origin=org)
    dm = hip.DeviceImage.upload(m, spacing=sp, origin=org)
    kw = dict(precision=precision, indices=indices)
    if fmask is not None:
        kw["fixed_mask"] = hip.DeviceMask.upload(fmask, spacing=sp, origin=org)
    if mmask is not None:
        kw["moving_mask"] = hip.DeviceMask.upload(mmask, spacing=sp, origin=org)
    d = hip.ResidentMetric(df, dm, **kw)
    pb = O.MetricProblem(f, sp, org, m, sp, org, indices=indices, fixed_mask=fmask,
                         moving_mask=mmask)
    return d, pb


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("case", [1, 2, 4, 5])
def test_mean_squares_moments(hip, precision, case):
    kind, p, ctr = TRANSFORMS[case]
    f = blob_volume((20, 24, 28), 3)
    m = blob_volume((20, 24, 28), 3) * 0.9 + 4.0
    d, pb = ms_pair(hip, f, m.astype(np.float32), precision)
    t = O.make_transform(kind, p, center=ctr)
    mom = d.evaluate(stages_of(t))
    j0, c = affine_form(t)
    v, deriv, valid = hip.contract_ms(mom, j0, c, t.nparams)
    wv, wd, wn = pb.ms_evaluate(t)
    assert valid == wn
    tol_v, tol_d = (1e-12, 1e-9) if precision == 1 else (1e-6, 1e-4)
    assert abs(v - wv) <= tol_v * abs(wv)
    assert rel_vec(deriv, wd) < tol_d
    sq, n = d.value(stages_of(t))
    assert n == wn and abs(sq / n - wv) <= tol_v * abs(wv)


def test_mean_squares_identity_is_zero(hip):
    f = blob_volume((16, 16, 16), 5)
    d, _ = ms_pair(hip, f, f, 0)
    mom = d.evaluate(stages_of(O.make_transform(O.TRANSLATION, [0, 0, 0])))
    assert mom[0] == 0.0 and not mom[1:13].any() and mom[13] == 16 ** 3


def test_mean_squares_sampled_and_masked(hip):
    f = blob_volume((20, 24, 28), 8)
    m = blob_volume((20, 24, 28), 9)
    idx = O.draw_samples(f.size, 2, 0.3, 3)
    fmask = (np.arange(f.size).reshape(f.shape) % 5 != 0).astype(np.uint8)
    mmask = np.ones(f.shape, np.uint8)
    mmask[:, :3, :] = 0
    d, pb = ms_pair(hip, f, m, 1, indices=idx, fmask=fmask, mmask=mmask)
    t = O.make_transform(O.EULER3D, [0.03, 0.01, -0.02, 0.4, -0.3, 0.8], [10, 12, 9])
    mom = d.evaluate(stages_of(t))
    j0, c = affine_form(t)
    v, deriv, valid = hip.contract_ms(mom, j0, c, 6)
    wv, wd, wn = pb.ms_evaluate(t)
    assert valid == wn and abs(v - wv) <= 1e-12 * abs(wv) and rel_vec(deriv, wd) < 1e-9


@pytest.mark.parametrize("kind", ["mattes", "ms"])
def test_device_draws_equal_host_draws(hip, kind):
    """Samples drawn on the device (SplitMix64, counting sort) are the host's
    draws (metric.rs:419-437) in the same ascending order: the metrics agree
    bit for bit."""
    f, m = ledger_pair()
    df = hip.DeviceImage.upload(f, spacing=SP, origin=ORG)
    dm = hip.DeviceImage.upload(m, spacing=SP, origin=ORG)
    idx = hip.draw_samples(f.size, hip.SAMPLING_RANDOM, 0.3, 99)
    t = O.make_transform(O.EULER3D, [0.05, -0.03, 0.08, 1.2, -0.7, 0.4], center=[9.0, 16.0, 18.0])
    st = stages_of(t)
    if kind == "mattes":
        a = hip.ResidentMattes(df, dm, bins=40, indices=idx)
        b = hip.ResidentMattes(df, dm, bins=40, device_draws=(idx.size, 99))
        assert a.sample_count() == b.sample_count() == idx.size
        va, ma, na = a.value_and_moments(st)
        vb, mb, nb = b.value_and_moments(st)
        assert va == vb and na == nb and np.array_equal(ma, mb)
    else:
        a = hip.ResidentMetric(df, dm, indices=idx)
        b = hip.ResidentMetric(df, dm, device_draws=(idx.size, 99))
        assert np.array_equal(a.evaluate(st), b.evaluate(st))


@pytest.mark.parametrize("precision", [1, 0])
@pytest.mark.parametrize("bins", [76, 77, 90, 91, 128, 255])
def test_bin_counts_across_the_lds_limits(hip, precision, bins):
    """Every histogram placement: LDS value pass up to 90 bins (global atomics
    above), staged tail up to 76 bins (global reads above), LDS pRatio in the
    derivative pass while it fits."""
    f, m = ledger_pair()
    d, pb = mattes_pair(hip, f, m, precision, bins=bins)
    g = pb.mattes_geometry(bins)
    t = O.make_transform(O.EULER3D, [0.05, -0.03, 0.08, 1.2, -0.7, 0.4], center=[9.0, 16.0, 18.0])
    want_v, want_d, want_n = pb.mattes_evaluate(g, t)
    v, mom, n = d.value_and_moments(stages_of(t))
    assert n == want_n
    tv, td = (1e-9, 1e-7) if precision == 1 else (1e-6, 1e-4)
    assert abs(v - want_v) <= tv * abs(want_v), (v, want_v)
    assert rel_vec(contract(mom, t), want_d) < td


@pytest.mark.parametrize("precision", [1, 0])
@pytest.mark.parametrize("n", [7, 63, 1030, 2049, 3075])
def test_sample_counts_that_leave_workgroups_empty(hip, precision, n):
    """Sample sets whose per-workgroup slices overrun n (ceil(n / G) * G well
    past n, G the fixed workgroup count), so the last workgroups have no
    samples: the record-driven derivative pass must load nothing for them.
    Value, derivative and valid count against the oracle."""
    f = blob_volume((16, 18, 20), 5, spacing=SP)
    m = blob_volume((16, 18, 20), 6, spacing=SP) * 0.5 + f * 0.5
    rng = np.random.default_rng(n)
    idx = np.sort(rng.choice(f.size, size=n, replace=False)).astype(np.uint32)
    d, pb = mattes_pair(hip, f, m, precision, bins=24, indices=idx)
    g = pb.mattes_geometry(24)
    t = O.make_transform(O.EULER3D, [0.04, -0.02, 0.03, 0.4, -0.3, 0.2], center=[5.0, 7.0, 9.0])
    want_v, want_d, want_n = pb.mattes_evaluate(g, t)
    v, mom, got_n = d.value_and_moments(stages_of(t))
    assert got_n == want_n
    if want_n == 0:
        return
    tv, td = (1e-9, 1e-7) if precision == 1 else (1e-6, 1e-4)
    assert abs(v - want_v) <= tv * abs(want_v) + 1e-15, (v, want_v)
    if np.linalg.norm(want_d) > 0:
        assert rel_vec(contract(mom, t), want_d) < td
