"""The host-data constructor of the resident metrics, ResidentMetric::new
(crates/sitk/src/cuda/ops/mean_squares.rs:201-256) with the three FixedPoints
forms (crates/sitk/src/cuda/ops/resident.rs:464-484): the one the reference's
CudaMetricBackend calls (crates/sitk/src/registration/cuda.rs:195-200).

- explicit points (the host's gathered sample points and values) against the
  oracle's explicit-point evaluation, mean squares and Mattes, F32 at the
  single-pair bar (value 1e-6, derivative 1e-4 relative) and F64 at 1e-12 /
  1e-9 (mean squares) and 1e-9 / 1e-7 (Mattes);
- explicit points that are exactly the grid's points give the grid metric's
  results (F64: the same per-sample arithmetic in the same sample order, so
  the same bits);
- a grid and an index list given as host values with an oblique direction and
  both masks against the oracle;
- an index outside the grid is refused by name.
"""

import math

import numpy as np
import pytest

import oracle as O
from helpers import affine_form, blob_volume, rel_vec, stages_of
from test_gpu_metric import contract

pytestmark = pytest.mark.gpu

SP = [0.8, 0.9, 1.1]
ORG = [-3.0, 2.0, 1.0]


def _moving():
    return blob_volume((20, 24, 28), 41, spacing=SP)


def _p2i(sp, direction=None):
    d = np.eye(3).reshape(9) if direction is None else np.asarray(direction, float).reshape(9)
    return O.invert3(O.index_to_physical(d, sp)).reshape(3, 3)


def _euler(centre):
    return O.make_transform(O.EULER3D, [0.04, -0.03, 0.06, 0.7, -0.5, 0.3], center=centre)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("kind", ["ms", "mattes"])
def test_explicit_points_against_the_oracle(hip, precision, kind):
    m = _moving()
    nz, ny, nx = m.shape
    rng = np.random.default_rng(12)
    lo = np.asarray(ORG) - 2.0
    hi = np.asarray(ORG) + np.array([nx, ny, nz]) * np.asarray(SP) + 2.0
    pts = rng.uniform(lo, hi, (3000, 3))  # some map outside the moving buffer
    vals = rng.uniform(0.0, 150.0, 3000)
    cls = hip.ResidentMattes if kind == "mattes" else hip.ResidentMetric
    kw = dict(bins=24) if kind == "mattes" else {}
    d = cls.from_host(vals, m, (nx, ny, nz), ORG, _p2i(SP), points=pts, precision=precision, **kw)
    assert d.sample_count() == 3000
    pb = O.MetricProblem(vals.astype(np.float32), SP, ORG, m, SP, ORG, points=pts)
    t = _euler([5.0, 12.0, 14.0])
    st = stages_of(t)
    if kind == "ms":
        j0, c = affine_form(t)
        v, deriv, n = hip.contract_ms(d.evaluate(st), j0, c, t.nparams)
        wv, wd, wn = pb.ms_evaluate(t)
        tv, td = (1e-12, 1e-9) if precision == 1 else (1e-6, 1e-4)
    else:
        v, mom, n = d.value_and_moments(st)
        deriv = contract(mom, t)
        wv, wd, wn = pb.mattes_evaluate(pb.mattes_geometry(24), t)
        tv, td = (1e-9, 1e-7) if precision == 1 else (1e-6, 1e-4)
    assert 0 < n == wn < 3000
    assert abs(v - wv) <= tv * abs(wv), (v, wv)
    assert rel_vec(deriv, wd) < td


def test_explicit_points_on_the_grid_equal_the_grid_metric(hip):
    """F64: the explicit points of every grid voxel, in grid order, with the
    grid's values, give the grid metric's value, moments and valid count bit
    for bit (mean squares and Mattes)."""
    f = blob_volume((12, 14, 16), 43, spacing=SP)
    m = blob_volume((12, 14, 16), 44, spacing=SP) * 0.6 + f * 0.4
    nz, ny, nx = f.shape
    # the origin-first fold of the grid point with an axis-aligned grid:
    # ((o + s0 i) + 0 j) + 0 k == o + s0 i, and so on per axis
    i2p = O.index_to_physical(np.eye(3).reshape(9), SP).reshape(3, 3)
    k, j, i = np.meshgrid(np.arange(nz, dtype=float), np.arange(ny, dtype=float),
                          np.arange(nx, dtype=float), indexing="ij")
    idx = np.stack([i.reshape(-1), j.reshape(-1), k.reshape(-1)])
    pts = np.empty((idx.shape[1], 3))
    for r in range(3):
        pts[:, r] = ((ORG[r] + i2p[r, 0] * idx[0]) + i2p[r, 1] * idx[1]) + i2p[r, 2] * idx[2]
    df = hip.DeviceImage.upload(f, spacing=SP, origin=ORG)
    dm = hip.DeviceImage.upload(m, spacing=SP, origin=ORG)
    t = _euler([5.0, 6.0, 7.0])
    st = stages_of(t)
    g_ms = hip.ResidentMetric(df, dm, precision=1)
    p_ms = hip.ResidentMetric.from_host(f.reshape(-1), m, (nx, ny, nz), ORG, _p2i(SP), points=pts,
                                        precision=1)
    assert np.array_equal(g_ms.evaluate(st), p_ms.evaluate(st))
    g_mi = hip.ResidentMattes(df, dm, bins=30, precision=1)
    p_mi = hip.ResidentMattes.from_host(f.reshape(-1), m, (nx, ny, nz), ORG, _p2i(SP), points=pts,
                                        precision=1, bins=30)
    v1, m1, n1 = g_mi.value_and_moments(st)
    v2, m2, n2 = p_mi.value_and_moments(st)
    assert v1 == v2 and n1 == n2 and np.array_equal(m1, m2)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("sampled", [False, True])
def test_host_grid_with_direction_and_masks_against_the_oracle(hip, precision, sampled):
    f = blob_volume((16, 18, 20), 45, spacing=SP)
    m = blob_volume((16, 18, 20), 46, spacing=SP) * 0.5 + f * 0.5
    nz, ny, nx = f.shape
    c, s = math.cos(0.15), math.sin(0.15)
    fdir = [c, 0.0, s, 0.0, 1.0, 0.0, -s, 0.0, c]
    mdir = [1.0, 0.0, 0.0, 0.0, c, -s, 0.0, s, c]
    i2p = O.index_to_physical(np.asarray(fdir), SP).reshape(3, 3)
    p2i = _p2i(SP, mdir)
    fmask = (np.arange(f.size) % 7 != 0).astype(np.uint8)
    mmask = np.ones(m.shape, np.uint8)
    mmask[:, :, :3] = 0
    idx = O.draw_samples(f.size, 2, 0.4, 5) if sampled else None
    d = hip.ResidentMattes.from_host(f.reshape(-1), m, (nx, ny, nz), ORG, p2i, indices=idx,
                                     grid=((nx, ny, nz), ORG, i2p), fixed_mask=fmask,
                                     moving_mask=mmask, precision=precision, bins=28)
    pb = O.MetricProblem(f, SP, ORG, m, SP, ORG, fdir=fdir, mdir=mdir, indices=idx,
                         fixed_mask=fmask.reshape(f.shape), moving_mask=mmask)
    t = _euler([4.0, 8.0, 9.0])
    v, mom, n = d.value_and_moments(stages_of(t))
    wv, wd, wn = pb.mattes_evaluate(pb.mattes_geometry(28), t)
    tv, td = (1e-9, 1e-7) if precision == 1 else (1e-6, 1e-4)
    assert n == wn
    assert abs(v - wv) <= tv * abs(wv), (v, wv)
    assert rel_vec(contract(mom, t), wd) < td


def test_host_indices_outside_the_grid_are_refused(hip):
    m = _moving()
    nz, ny, nx = m.shape
    with pytest.raises(hip.CudaError) as e:
        hip.ResidentMetric.from_host(np.zeros(8), m, (nx, ny, nz), ORG, _p2i(SP),
                                     indices=[0, 3, 8], grid=((2, 2, 2), ORG, np.eye(3)))
    assert e.value.status == "SampleIndexOutOfGrid"
    d = hip.ResidentMetric.from_host(np.zeros(4), m, (nx, ny, nz), ORG, _p2i(SP),
                                     points=np.zeros((4, 3)))
    with pytest.raises(hip.CudaError) as e:
        hip.Batch([d])
    assert e.value.status == "InvalidArgument"
