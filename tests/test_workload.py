"""The synthetic workloads' own arithmetic (CPU): the deformable truth and the
B-spline displacement used for TRE."""

import numpy as np

import oracle as O
import workload as W


def test_bspline_displacement_is_the_oracles_point_map():
    # crates/sitk/src/transform/bspline.rs:485-583 through the oracle
    vol = W.Volume(np.zeros((18, 20, 22), np.float32), (1.0, 1.1, 0.9), (-2.0, 1.0, 0.5))
    mesh = (3, 2, 4)
    gs, go, gsp = W.bspline_grid(vol, mesh)
    rng = np.random.default_rng(1)
    c = rng.normal(0, 1, 3 * int(np.prod(gs)))
    t = O.make_bspline(vol.size, vol.spacing, vol.origin, np.eye(3), mesh, c)
    assert list(t.grid_size) == list(gs)
    assert np.allclose(list(t.grid_origin), go) and np.allclose(list(t.grid_spacing), gsp)
    ext = np.asarray(vol.size) * np.asarray(vol.spacing)
    pts = np.asarray(vol.origin) + rng.uniform(0, 1, (400, 3)) * ext
    want = np.array([O.transform_point(t, p) for p in pts]) - pts
    got = W.bspline_displacement(pts, gs, go, gsp, c)
    assert np.abs(want - got).max() < 1e-12


def test_deformable_truth_inverts_the_generating_map():
    vol = W.Volume(np.zeros((40, 40, 40), np.float32), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0))
    rng = np.random.default_rng(2)
    _, R = W.random_euler(rng, 5.0, 5.0, vol.center())
    u = W.rbf_field(rng, vol, nbumps=6, peak=4.0, sigma=(8.0, 12.0))
    truth = W.DeformableTruth(R, u)
    x = rng.uniform(5, 35, (200, 3))
    y = truth(x)
    z = R.inverse().apply(y)
    assert np.abs((z - u(z)) - x).max() < 1e-9  # V(T(x)) = x
    g = np.stack(np.meshgrid(*[np.arange(0, 40, 4)] * 3, indexing="ij"), -1).reshape(-1, 3)
    assert abs(np.linalg.norm(u(g.astype(float)), axis=1).max() - 4.0) < 1e-9


def test_deformable_tre_is_zero_for_the_truths_own_rigid_part_without_bumps():
    vol = W.Volume(np.zeros((48, 48, 48), np.float32), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0))
    rng = np.random.default_rng(3)
    _, R = W.random_euler(rng, 5.0, 5.0, vol.center())
    u = W.RbfField(np.zeros((1, 3)), np.zeros((1, 3)), np.ones(1))
    truth = W.DeformableTruth(R, u)
    mesh = (3, 3, 3)
    assert W.deformable_tre(truth, vol, mesh, np.zeros(3 * 6 ** 3), R.stage(), margin=8.0) < 1e-12
