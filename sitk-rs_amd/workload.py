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
