"""Seeded synthetic registration workloads (the benchmark configurations).

The reference has no generator for these; they are defined here. Every
random draw comes from numpy's PCG64 generator seeded with the stated seed, so
a workload is reproducible bit for bit on any host.

Volumes are float32 numpy arrays of shape (nz, ny, nx) (dimension 0 fastest in
memory), identity direction. Geometry is (spacing, origin) in (x, y, z).

Resample semantics follow ResampleImageFilter
(crates/sitk/src/registration/method.rs:2502-2515): the output voxel at physical
x takes the input's value at T(x), linearly interpolated, 0 outside. A moving
image made this way from the fixed image through T_gt is registered by T_gt^-1.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class Volume:
    data: np.ndarray           # float32 (nz, ny, nx)
    spacing: tuple
    origin: tuple

    @property
    def size(self):
        nz, ny, nx = self.data.shape
        return (nx, ny, nz)

    def center(self) -> np.ndarray:
        return np.asarray(self.origin) + 0.5 * (np.asarray(self.size) - 1) * np.asarray(self.spacing)


def blob_volume(size, spacing, seed, nblobs=40, amp=(50.0, 200.0), sigma=(2.0, 8.0),
                noise=5.0, bias=0.0, origin=(0.0, 0.0, 0.0)) -> Volume:
    """Sum of `nblobs` 3-D Gaussian blobs (amplitude U[amp], sigma U[sigma] mm,
    centres uniform over the volume) plus N(0, noise) voxel noise, optionally
    times a smooth second-order polynomial bias field of relative amplitude
    `bias` (values in [1 - bias, 1 + bias])."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = size
    sp = np.asarray(spacing, dtype=np.float64)
    ext = np.array([nx, ny, nz]) * sp
    vol = np.zeros((nz, ny, nx), dtype=np.float64)
    centres = rng.uniform(0.0, 1.0, (nblobs, 3)) * ext
    amps = rng.uniform(amp[0], amp[1], nblobs)
    sigs = rng.uniform(sigma[0], sigma[1], nblobs)
    xs = [np.arange(n) * s for n, s in zip((nx, ny, nz), sp)]
    for c, a, s in zip(centres, amps, sigs):
        # separable: exp(-(dx^2+dy^2+dz^2)/2s^2) = gx * gy * gz, within 4.5 sigma
        g = []
        lo_hi = []
        for d in range(3):
            lo = max(0, int(math.floor((c[d] - 4.5 * s) / sp[d])))
            hi = min(len(xs[d]), int(math.ceil((c[d] + 4.5 * s) / sp[d])) + 1)
            lo_hi.append((lo, hi))
            g.append(np.exp(-((xs[d][lo:hi] - c[d]) ** 2) / (2.0 * s * s)))
        (x0, x1), (y0, y1), (z0, z1) = lo_hi
        if x1 <= x0 or y1 <= y0 or z1 <= z0:
            continue
        vol[z0:z1, y0:y1, x0:x1] += a * g[2][:, None, None] * g[1][None, :, None] * g[0][None, None, :]
    if bias:
        coef = rng.uniform(-1.0, 1.0, 10)
        u = [np.linspace(-1.0, 1.0, n) for n in (nx, ny, nz)]
        X, Y, Z = u[0][None, None, :], u[1][None, :, None], u[2][:, None, None]
        terms = [X, Y, Z, X * X, Y * Y, Z * Z, X * Y, X * Z, Y * Z]
        field = np.zeros((nz, ny, nx))
        for k, t in enumerate(terms):
            field = field + coef[k] * t
        field /= max(np.abs(field).max(), 1e-12)
        vol *= 1.0 + bias * field
    if noise:
        vol += rng.normal(0.0, noise, vol.shape)
    return Volume(vol.astype(np.float32), tuple(map(float, spacing)), tuple(map(float, origin)))


def euler_matrix(ax, ay, az):
    """Rz . Rx . Ry (ITK's default Euler3D order, parametric.rs:1030-1047)."""
    cx, sx, cy, sy, cz, sz = math.cos(ax), math.sin(ax), math.cos(ay), math.sin(ay), math.cos(az), math.sin(az)
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return rz @ rx @ ry


@dataclass
class AffineMap:
    """y = A (x - c) + t + c, i.e. y = A x + offset."""
    A: np.ndarray
    t: np.ndarray
    c: np.ndarray

    @property
    def offset(self):
        return self.c + self.t - self.A @ self.c

    def apply(self, pts):
        return pts @ self.A.T + self.offset

    def inverse(self) -> "AffineMap":
        Ai = np.linalg.inv(self.A)
        # x = Ai (y - offset)  ->  about the same centre: t' = -Ai t' ...
        off_i = -Ai @ self.offset
        c = self.c
        t_i = off_i - c + Ai @ c
        return AffineMap(Ai, t_i, c)

    def stage(self) -> np.ndarray:
        return np.concatenate([self.A.reshape(9), self.offset])


def random_euler(rng, max_deg, max_mm, centre) -> tuple:
    ang = np.deg2rad(rng.uniform(-max_deg, max_deg, 3))
    t = rng.uniform(-max_mm, max_mm, 3)
    return (ang, t), AffineMap(euler_matrix(*ang), t, np.asarray(centre, dtype=np.float64))


def random_affine(rng, max_deg, max_scale, max_mm, centre) -> AffineMap:
    ang = np.deg2rad(rng.uniform(-max_deg, max_deg, 3))
    s = 1.0 + rng.uniform(-max_scale, max_scale, 3)
    t = rng.uniform(-max_mm, max_mm, 3)
    return AffineMap(euler_matrix(*ang) @ np.diag(s), t, np.asarray(centre, dtype=np.float64))


def resample_linear(src: Volume, T: AffineMap, out_size=None, out_spacing=None,
                    out_origin=None) -> Volume:
    """out(x) = src(T(x)), trilinear, 0 outside the source buffer."""
    from scipy import ndimage

    out_size = src.size if out_size is None else out_size
    out_spacing = src.spacing if out_spacing is None else out_spacing
    out_origin = src.origin if out_origin is None else out_origin
    nx, ny, nz = out_size
    osp = np.asarray(out_spacing)
    oorg = np.asarray(out_origin)
    ssp = np.asarray(src.spacing)
    sorg = np.asarray(src.origin)
    # continuous source index of T(x) is affine in the output index: M idx + b
    M = (T.A * osp[None, :]) / ssp[:, None]
    b = (T.A @ oorg + T.offset - sorg) / ssp
    # map_coordinates works in (z, y, x) array order
    Mz = M[::-1, ::-1]
    bz = b[::-1]
    out = ndimage.affine_transform(src.data.astype(np.float32), Mz, offset=bz,
                                   output_shape=(nz, ny, nx), order=1, mode="constant", cval=0.0,
                                   prefilter=False)
    return Volume(out.astype(np.float32), tuple(map(float, osp)), tuple(map(float, oorg)))


def tre(T_est: AffineMap, T_true_inv: AffineMap, vol: Volume, stride: int = 8) -> float:
    """Mean target registration error (mm) over a grid of points of `vol`:
    |T_est(x) - T_gt^-1(x)| averaged over every `stride`-th voxel."""
    nx, ny, nz = vol.size
    g = np.stack(np.meshgrid(np.arange(0, nx, stride), np.arange(0, ny, stride),
                             np.arange(0, nz, stride), indexing="ij"), -1).reshape(-1, 3)
    pts = np.asarray(vol.origin) + g * np.asarray(vol.spacing)
    return float(np.linalg.norm(T_est.apply(pts) - T_true_inv.apply(pts), axis=1).mean())


# --------------------------------------------------------------- configurations
def config_cpu_pair(seed: int = 1):
    """The CPU-runnable rigid pair: 64^3 float32, spacing 1 mm, identity
    direction, 40 blobs (amplitude U[50,200], sigma U[2,8] mm) plus N(0,5)
    noise; moving = fixed resampled through a seeded Euler3D (rotations
    +-5 deg, translations +-3 mm about the volume centre), linear
    interpolation. Returns (fixed, moving, T_gt)."""
    fixed = blob_volume((64, 64, 64), (1.0, 1.0, 1.0), seed, nblobs=40)
    rng = np.random.default_rng(seed + 500)
    _, T_gt = random_euler(rng, 5.0, 3.0, fixed.center())
    moving = resample_linear(fixed, T_gt)
    return fixed, moving, T_gt


def config_large_pair(seed: int = 2):
    """256x256x160 float32, spacing 0.9x0.9x1.5 mm, 40 blobs, +-20% bias field;
    moving = fixed resampled through a seeded affine (rotations +-10 deg, scale
    +-5%, translations +-10 mm about the volume centre)."""
    fixed = blob_volume((256, 256, 160), (0.9, 0.9, 1.5), seed, nblobs=40, bias=0.2)
    rng = np.random.default_rng(seed + 1000)
    T_gt = random_affine(rng, 10.0, 0.05, 10.0, fixed.center())
    moving = resample_linear(fixed, T_gt)
    return fixed, moving, T_gt


def config_batched_pairs(n: int = 512, seed: int = 3):
    """`n` independent rigid pairs: per-pair edge lengths drawn in [32, 128]
    per axis, spacing 1 mm, the blob generator (40 blobs), each with its own
    Euler3D perturbation (rotations +-15 deg, translations +-15 mm about the
    volume centre). Even pairs use mean squares, odd pairs Mattes MI.
    Returns a list of dicts {fixed: Volume, T_gt: AffineMap, metric: "ms"|"mattes"};
    the moving image of a pair is fixed resampled through T_gt."""
    rng = np.random.default_rng(seed)
    pairs = []
    for i in range(n):
        size = tuple(int(v) for v in rng.integers(32, 129, 3))
        vol = blob_volume(size, (1.0, 1.0, 1.0), seed=seed * 1_000_003 + i, nblobs=40)
        _, T_gt = random_euler(rng, 15.0, 15.0, vol.center())
        pairs.append({"fixed": vol, "T_gt": T_gt, "metric": "ms" if i % 2 == 0 else "mattes"})
    return pairs


# ------------------------------------------------------- deformable workload
@dataclass
class RbfField:
    """A smooth displacement u(x) = sum_k a_k exp(-|x - c_k|^2 / (2 s_k^2)) (mm),
    a_k a 3-vector per bump."""
    centres: np.ndarray   # (K, 3)
    vectors: np.ndarray   # (K, 3)
    sigmas: np.ndarray    # (K,)

    def __call__(self, pts) -> np.ndarray:
        pts = np.asarray(pts, dtype=np.float64)
        out = np.zeros_like(pts)
        for c, a, s in zip(self.centres, self.vectors, self.sigmas):
            d2 = np.sum((pts - c) ** 2, axis=-1)
            out += np.exp(-d2 / (2.0 * s * s))[..., None] * a
        return out


def rbf_field(rng, vol: Volume, nbumps=20, peak=4.0, sigma=(15.0, 30.0)) -> RbfField:
    """`nbumps` Gaussian bumps with centres uniform over the middle 80% of the
    volume, sigma U[sigma] mm and random directions, scaled so the largest
    displacement over a 4-voxel grid of the volume is `peak` mm."""
    size = np.asarray(vol.size, dtype=np.float64)
    sp = np.asarray(vol.spacing)
    org = np.asarray(vol.origin)
    ext = size * sp
    centres = org + rng.uniform(0.1, 0.9, (nbumps, 3)) * ext
    dirs = rng.normal(0.0, 1.0, (nbumps, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    vectors = dirs * rng.uniform(0.5, 1.0, (nbumps, 1))
    sigmas = rng.uniform(sigma[0], sigma[1], nbumps)
    f = RbfField(centres, vectors, sigmas)
    g = np.stack(np.meshgrid(*[np.arange(0, n, 4) for n in vol.size], indexing="ij"), -1).reshape(-1, 3)
    m = np.linalg.norm(f(org + g * sp), axis=1).max()
    return RbfField(centres, vectors * (peak / m), sigmas)


@dataclass
class DeformableTruth:
    """The registration map that aligns a deformable pair, T(x) = R(z) with
    z = x + u(z): the moving image was made as moving(y) = fixed(V(y)),
    V(y) = z - u(z) with z = R^-1(y), so V(T(x)) = x."""
    R: AffineMap
    u: RbfField

    def __call__(self, pts, iterations=30) -> np.ndarray:
        x = np.asarray(pts, dtype=np.float64)
        z = x.copy()
        for _ in range(iterations):  # |grad u| < 1: a contraction
            z = x + self.u(z)
        return self.R.apply(z)


def config_deformable(seed: int = 5, edge: int = 192, nblobs: int = 135, peak: float = 4.0,
                      max_deg: float = 5.0, max_mm: float = 5.0):
    """192^3 float32, 1 mm, the blob generator at the 128^3 pairs' density
    (135 blobs); moving(y) = fixed(V(y)) with V(y) = z - u(z), z = R^-1(y), R a
    seeded Euler3D offset (+-5 deg, +-5 mm about the centre) and u 20 Gaussian
    RBF bumps of peak 4 mm. Returns (fixed, moving, DeformableTruth)."""
    from scipy import ndimage

    fixed = blob_volume((edge,) * 3, (1.0, 1.0, 1.0), seed, nblobs=nblobs)
    rng = np.random.default_rng(seed + 2000)
    _, R = random_euler(rng, max_deg, max_mm, fixed.center())
    u = rbf_field(rng, fixed, peak=peak)
    Ri = R.inverse()
    out = np.empty_like(fixed.data)
    n = edge
    xs = np.arange(n, dtype=np.float64)
    for k in range(n):  # one z slice at a time keeps the temporaries small
        X, Y = np.meshgrid(xs, xs, indexing="xy")
        y = np.stack([X.ravel(), Y.ravel(), np.full(X.size, float(k))], -1)
        z = Ri.apply(y)
        v = z - u(z)
        coords = np.stack([v[:, 2], v[:, 1], v[:, 0]])  # (z, y, x) array order, spacing 1, origin 0
        out[k] = ndimage.map_coordinates(fixed.data, coords, order=1, mode="constant",
                                         cval=0.0, prefilter=False).reshape(n, n)
    moving = Volume(out, fixed.spacing, fixed.origin)
    return fixed, moving, DeformableTruth(R, u)


def bspline_grid(vol: Volume, mesh):
    """The control grid of a B-spline transform over the volume's extent
    (crates/sitk/src/transform/bspline.rs:324-339; identity direction):
    (grid size, grid origin, grid spacing)."""
    size = np.asarray(vol.size, dtype=np.float64)
    sp = np.asarray(vol.spacing)
    gsp = size * sp / np.asarray(mesh, dtype=np.float64)
    return np.asarray(mesh) + 3, np.asarray(vol.origin) - gsp, gsp


def _b3(u):
    a = np.abs(u)
    return np.where(a < 1.0, (4.0 - 6.0 * a * a + 3.0 * a ** 3) / 6.0,
                    np.where(a < 2.0, (2.0 - a) ** 3 / 6.0, 0.0))


def bspline_displacement(pts, grid_size, grid_origin, grid_spacing, coeffs) -> np.ndarray:
    """The cubic B-spline displacement at physical points (identity grid
    direction; points outside the valid region get zero), coefficients
    [3][ncp], cp = x + gs0 (y + gs1 z) (bspline.rs:485-583)."""
    pts = np.asarray(pts, dtype=np.float64)
    gs = np.asarray(grid_size, dtype=np.int64)
    c = (pts - np.asarray(grid_origin)) / np.asarray(grid_spacing)
    ok = np.all((c >= 1.0) & (c < gs - 2.0), axis=1)
    start = np.floor(c - 1.0).astype(np.int64)
    start = np.clip(start, 0, gs - 4)
    u = c - start
    ncp = int(np.prod(gs))
    coef = np.asarray(coeffs, dtype=np.float64).reshape(3, ncp)
    out = np.zeros_like(pts)
    for k in range(4):
        wk = _b3(u[:, 2] - k)
        for j in range(4):
            wj = _b3(u[:, 1] - j)
            for i in range(4):
                w = _b3(u[:, 0] - i) * wj * wk
                cp = (start[:, 0] + i) + gs[0] * ((start[:, 1] + j) + gs[1] * (start[:, 2] + k))
                out += w[:, None] * coef[:, cp].T
    out[~ok] = 0.0
    return out


def deformable_tre(truth: DeformableTruth, vol: Volume, mesh, coeffs, stage, stride=8,
                   margin=16.0) -> float:
    """Mean |T_est(x) - T(x)| (mm) over every `stride`-th voxel at least `margin`
    mm inside the volume, T_est(x) = stage(x + B(x))."""
    g = np.stack(np.meshgrid(*[np.arange(0, n, stride) for n in vol.size], indexing="ij"),
                 -1).reshape(-1, 3)
    pts = np.asarray(vol.origin) + g * np.asarray(vol.spacing)
    lo = np.asarray(vol.origin) + margin
    hi = np.asarray(vol.origin) + (np.asarray(vol.size) - 1) * np.asarray(vol.spacing) - margin
    pts = pts[np.all((pts >= lo) & (pts <= hi), axis=1)]
    gs, go, gsp = bspline_grid(vol, mesh)
    y = pts + bspline_displacement(pts, gs, go, gsp, coeffs)
    st = np.asarray(stage, dtype=np.float64).reshape(12)
    est = y @ st[:9].reshape(3, 3).T + st[9:]
    return float(np.linalg.norm(est - truth(pts), axis=1).mean())
