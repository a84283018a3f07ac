"""CPU oracle for the sitk-rs registration metric path -- TEST INFRASTRUCTURE ONLY.

A ctypes binding of ``liboracle.so`` (built from ``oracle.c`` by the Makefile next
to it). Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker or the timed CPU
baseline; the product library never loads it.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

TRANSLATION, EULER3D, AFFINE, BSPLINE = 0, 1, 2, 3


def build() -> None:
    """Compile liboracle.so with the Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class Transform(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("nparams", C.c_int),
        ("params", C.c_double * 12),
        ("center", C.c_double * 3),
        ("compute_zyx", C.c_int),
        ("matrix", C.c_double * 9),
        ("offset", C.c_double * 3),
        ("grid_size", C.c_size_t * 3),
        ("grid_origin", C.c_double * 3),
        ("grid_spacing", C.c_double * 3),
        ("grid_direction", C.c_double * 9),
        ("grid_p2i", C.c_double * 9),
        ("coeffs", C.POINTER(C.c_double)),
    ]


class Problem(C.Structure):
    _fields_ = [
        ("fsize", C.c_size_t * 3),
        ("forigin", C.c_double * 3),
        ("fi2p", C.c_double * 9),
        ("fvals", C.POINTER(C.c_float)),
        ("idx", C.POINTER(C.c_int64)),
        ("nidx", C.c_size_t),
        ("pts", C.POINTER(C.c_double)),
        ("fmask", C.POINTER(C.c_uint8)),
        ("msize", C.c_size_t * 3),
        ("morigin", C.c_double * 3),
        ("mp2i", C.c_double * 9),
        ("mvals", C.POINTER(C.c_float)),
        ("mmask", C.POINTER(C.c_uint8)),
        ("mcoef", C.POINTER(C.c_double)),
    ]


class MattesGeom(C.Structure):
    _fields_ = [
        ("num_bins", C.c_size_t),
        ("moving_true_min", C.c_double),
        ("moving_true_max", C.c_double),
        ("fixed_bin_size", C.c_double),
        ("moving_bin_size", C.c_double),
        ("fixed_normalized_min", C.c_double),
        ("moving_normalized_min", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        fp = C.POINTER(C.c_float)
        szp = C.POINTER(C.c_size_t)
        L.orc_synth.argtypes = [C.c_uint64, C.c_size_t, fp]
        L.orc_fnv1a64.argtypes = [C.c_void_p, C.c_size_t]
        L.orc_fnv1a64.restype = C.c_uint64
        L.orc_deriche_coefficients.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, dp]
        L.orc_filter_line.argtypes = [dp, C.c_size_t, dp, dp, dp]
        L.orc_recursive_gaussian_f64.argtypes = [dp, szp, dp, dp]
        L.orc_smoothing_recursive_gaussian_f32.argtypes = [fp, fp, szp, dp, dp]
        L.orc_smoothing_recursive_gaussian_f32_mt.argtypes = [fp, fp, szp, dp, dp, C.c_int]
        L.orc_index_to_physical_matrix.argtypes = [dp, dp, dp]
        L.orc_invert3.argtypes = [dp, dp]
        tp = C.POINTER(Transform)
        pp = C.POINTER(Problem)
        gp = C.POINTER(MattesGeom)
        L.orc_transform_recompute.argtypes = [tp]
        L.orc_transform_point.argtypes = [tp, dp, dp]
        L.orc_transform_jacobian.argtypes = [tp, dp, dp]
        L.orc_transform_stage.argtypes = [tp, dp]
        L.orc_replay_stages.argtypes = [dp, C.c_int, dp, dp]
        L.orc_bspline_from_image_domain.argtypes = [tp, szp, dp, dp, dp, szp, dp]
        L.orc_draw_samples.argtypes = [C.c_size_t, C.c_int, C.c_double, C.c_uint64,
                                       C.POINTER(C.c_int64)]
        L.orc_draw_samples.restype = C.c_size_t
        L.orc_fixed_value_range.argtypes = [pp, dp]
        L.orc_moving_value_range.argtypes = [pp, dp]
        L.orc_moving_value_and_physical_gradient.argtypes = [pp, dp, dp, dp]
        L.orc_moving_value_at.argtypes = [pp, dp, dp]
        L.orc_mattes_geometry.argtypes = [dp, dp, C.c_size_t, gp]
        L.orc_cubic_bspline.argtypes = [C.c_double]
        L.orc_cubic_bspline.restype = C.c_double
        L.orc_cubic_bspline_derivative.argtypes = [C.c_double]
        L.orc_cubic_bspline_derivative.restype = C.c_double
        L.orc_compensated_sum.argtypes = [dp, C.c_size_t]
        L.orc_compensated_sum.restype = C.c_double
        L.orc_set_joint_sum_compensated.argtypes = [C.c_int]
        L.orc_mattes_histogram.argtypes = [pp, gp, tp, dp, dp, szp]
        L.orc_mattes_tail.argtypes = [dp, dp, C.c_size_t, gp, dp, dp]
        L.orc_mattes_value.argtypes = [pp, gp, tp]
        L.orc_mattes_value.restype = C.c_double
        L.orc_mattes_evaluate.argtypes = [pp, gp, tp, dp, szp]
        L.orc_mattes_evaluate.restype = C.c_double
        L.orc_mattes_evaluate_bspline_composed.argtypes = [pp, gp, tp, tp, dp, szp]
        L.orc_mattes_evaluate_bspline_composed.restype = C.c_double
        L.orc_mattes_moments.argtypes = [pp, gp, tp, dp, dp]
        L.orc_ms_evaluate.argtypes = [pp, tp, dp, szp]
        L.orc_ms_evaluate.restype = C.c_double
        L.orc_ms_value.argtypes = [pp, tp, szp]
        L.orc_ms_evaluate_threads.argtypes = [pp, tp, dp, szp, C.c_int]
        L.orc_ms_evaluate_threads.restype = C.c_double
        L.orc_ms_value.restype = C.c_double
        L.orc_resample_linear.argtypes = [fp, szp, dp, dp, dp, szp, dp, dp, dp, tp, C.c_double, fp]
        L.orc_resample_nearest.argtypes = L.orc_resample_linear.argtypes
        L.orc_shrunk_geometry.argtypes = [szp, dp, dp, dp, szp, szp, dp, dp]
        L.orc_fixture_volume.argtypes = [C.c_size_t, dp, fp]
        L.orc_fixture_inverted.argtypes = [C.c_size_t, dp, fp]
        L.orc_fixture_commensurate.argtypes = [C.c_size_t, fp]
        L.orc_bspline_coefficients.argtypes = [fp, szp, dp]
        L.orc_powi.argtypes = [C.c_double, C.c_int]
        L.orc_powi.restype = C.c_double
        L.orc_bspline_value_and_gradient.argtypes = [dp, szp, dp, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _szp(vals):
    return (C.c_size_t * 3)(*[int(v) for v in vals])


def _d3(vals):
    return (C.c_double * 3)(*[float(v) for v in vals])


# ---------------------------------------------------------------- generators
def synth(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.float32)
    lib().orc_synth(C.c_uint64(seed), n, _fp(out))
    return out


def fnv1a64(arr: np.ndarray) -> int:
    a = np.ascontiguousarray(arr)
    return int(lib().orc_fnv1a64(a.ctypes.data, a.nbytes))


# ---------------------------------------------------------------- Gaussian
def deriche_coefficients(sigmad: float, order: int = 0, sigma: float = 1.0,
                         normalize: bool = False) -> np.ndarray:
    out = np.empty(20, dtype=np.float64)
    lib().orc_deriche_coefficients(order, sigmad, sigma, int(normalize), _dp(out))
    return out


def filter_line(data: np.ndarray, coef: np.ndarray) -> np.ndarray:
    data = np.ascontiguousarray(data, dtype=np.float64)
    outs = np.empty_like(data)
    scratch = np.empty_like(data)
    lib().orc_filter_line(_dp(data), data.size, _dp(np.ascontiguousarray(coef)), _dp(outs),
                          _dp(scratch))
    return outs


def smoothing_recursive_gaussian(vol: np.ndarray, size, spacing, sigma, threads: int = 1):
    """vol: flat float32, dim-0 fastest, `size` = (nx, ny, nz)."""
    vol = np.ascontiguousarray(vol, dtype=np.float32).reshape(-1)
    out = np.empty_like(vol)
    rc = lib().orc_smoothing_recursive_gaussian_f32_mt(_fp(vol), _fp(out), _szp(size),
                                                       _d3(spacing), _d3(sigma), threads)
    if rc != 0:
        raise ValueError(f"recursive gaussian refused ({rc})")
    return out


# ---------------------------------------------------------------- geometry
def bspline_coefficients(vol: np.ndarray) -> np.ndarray:
    """Cubic B-spline coefficients of a float32 volume (nz, ny, nx), FP64
    (interpolator.rs:170-190)."""
    v = np.ascontiguousarray(vol, dtype=np.float32)
    nz, ny, nx = v.shape
    out = np.empty(v.shape, dtype=np.float64)
    lib().orc_bspline_coefficients(_fp(v), _szp([nx, ny, nz]), _dp(out))
    return out


def powi(a: float, b: int) -> float:
    return float(lib().orc_powi(float(a), int(b)))


def bspline_value_and_gradient(coef: np.ndarray, c):
    """(value, index gradient) of the cubic B-spline interpolant at continuous
    index c (x, y, z), or None outside (interpolator.rs:273-350)."""
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    nz, ny, nx = coef.shape
    v = C.c_double()
    g = np.empty(3)
    ok = lib().orc_bspline_value_and_gradient(_dp(coef), _szp([nx, ny, nz]), _d3(c), C.byref(v),
                                              _dp(g))
    return (v.value, g) if ok else None


def index_to_physical(direction, spacing) -> np.ndarray:
    d = np.ascontiguousarray(direction, dtype=np.float64).reshape(9)
    s = np.ascontiguousarray(spacing, dtype=np.float64)
    out = np.empty(9, dtype=np.float64)
    lib().orc_index_to_physical_matrix(_dp(d), _dp(s), _dp(out))
    return out


def invert3(m) -> np.ndarray:
    m = np.ascontiguousarray(m, dtype=np.float64).reshape(9)
    out = np.empty(9, dtype=np.float64)
    if lib().orc_invert3(_dp(m), _dp(out)) != 0:
        raise ValueError("singular")
    return out


def make_transform(kind: int, params, center=(0.0, 0.0, 0.0), compute_zyx=False) -> Transform:
    t = Transform()
    t.kind = kind
    for i, p in enumerate(params):
        t.params[i] = float(p)
    for i in range(3):
        t.center[i] = float(center[i])
    t.compute_zyx = int(compute_zyx)
    lib().orc_transform_recompute(C.byref(t))
    return t


def make_bspline(img_size, img_spacing, img_origin, img_dir, mesh, coeffs) -> Transform:
    t = Transform()
    c = np.ascontiguousarray(coeffs, dtype=np.float64)
    rc = lib().orc_bspline_from_image_domain(
        C.byref(t), _szp(img_size), _d3(img_spacing), _d3(img_origin),
        (C.c_double * 9)(*[float(v) for v in np.asarray(img_dir).reshape(9)]),
        _szp(mesh), _dp(c))
    if rc != 0:
        raise ValueError("bad b-spline domain")
    t._keep = c  # keep the coefficient buffer alive
    return t


def transform_point(t: Transform, p) -> np.ndarray:
    p = np.ascontiguousarray(p, dtype=np.float64)
    out = np.empty(3, dtype=np.float64)
    lib().orc_transform_point(C.byref(t), _dp(p), _dp(out))
    return out


def transform_jacobian(t: Transform, p) -> np.ndarray:
    p = np.ascontiguousarray(p, dtype=np.float64)
    out = np.empty(3 * t.nparams, dtype=np.float64)
    lib().orc_transform_jacobian(C.byref(t), _dp(p), _dp(out))
    return out.reshape(3, t.nparams)


def transform_stage(t: Transform) -> np.ndarray:
    out = np.empty(12, dtype=np.float64)
    lib().orc_transform_stage(C.byref(t), _dp(out))
    return out


def replay_stages(stages: np.ndarray, p) -> np.ndarray:
    s = np.ascontiguousarray(stages, dtype=np.float64).reshape(-1)
    p = np.ascontiguousarray(p, dtype=np.float64)
    out = np.empty(3, dtype=np.float64)
    lib().orc_replay_stages(_dp(s), s.size // 12, _dp(p), _dp(out))
    return out


def draw_samples(voxels: int, strategy: int, percentage: float, seed: int):
    L = lib()
    n = L.orc_draw_samples(voxels, strategy, percentage, C.c_uint64(seed), None)
    if n == C.c_size_t(-1).value:
        return None
    out = np.empty(n, dtype=np.int64)
    L.orc_draw_samples(voxels, strategy, percentage, C.c_uint64(seed),
                       out.ctypes.data_as(C.POINTER(C.c_int64)))
    return out


# ---------------------------------------------------------------- metric problem
class MetricProblem:
    """Fixed samples + moving image, as the host metric holds them.

    fixed / moving: float32 arrays of shape (nz, ny, nx) (dim 0 fastest in memory).
    Geometry: spacing, origin (x, y, z) and a row-major 3x3 direction.
    """

    def __init__(self, fixed, fspacing, forigin, moving, mspacing, morigin,
                 fdir=None, mdir=None, indices=None, fixed_mask=None, moving_mask=None,
                 points=None, interpolator="linear"):
        """points: explicit sample points (n x 3, physical), with `fixed` then
        the n gathered sample values (FixedPoints::Explicit); the fixed
        geometry is unused. interpolator: "linear" or "bspline" (cubic
        B-spline of the moving image, MovingImage::from_image_with_interpolator,
        registration/metric.rs:720-750)."""
        eye = np.eye(3).reshape(9)
        fdir = eye if fdir is None else np.asarray(fdir, dtype=np.float64).reshape(9)
        mdir = eye if mdir is None else np.asarray(mdir, dtype=np.float64).reshape(9)
        self.fixed = np.ascontiguousarray(fixed, dtype=np.float32)
        self.moving = np.ascontiguousarray(moving, dtype=np.float32)
        self._pts = None
        if points is not None:
            self._pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
            self.fixed = self.fixed.reshape(1, 1, -1)
            assert self.fixed.size == self._pts.shape[0]
        fz, fy, fx = self.fixed.shape
        mz, my, mx = self.moving.shape
        pb = Problem()
        pb.fsize[:] = [fx, fy, fz]
        pb.forigin[:] = [float(v) for v in forigin]
        fi2p = index_to_physical(fdir, fspacing)
        pb.fi2p[:] = list(fi2p)
        pb.fvals = _fp(self.fixed)
        self._idx = None
        if indices is not None:
            self._idx = np.ascontiguousarray(indices, dtype=np.int64)
            pb.idx = self._idx.ctypes.data_as(C.POINTER(C.c_int64))
            pb.nidx = self._idx.size
        if self._pts is not None:
            pb.pts = self._pts.ctypes.data_as(C.POINTER(C.c_double))
            pb.nidx = self._pts.shape[0]
        self._fmask = None
        if fixed_mask is not None:
            self._fmask = np.ascontiguousarray(fixed_mask, dtype=np.uint8)
            pb.fmask = self._fmask.ctypes.data_as(C.POINTER(C.c_uint8))
        pb.msize[:] = [mx, my, mz]
        pb.morigin[:] = [float(v) for v in morigin]
        mi2p = index_to_physical(mdir, mspacing)
        pb.mp2i[:] = list(invert3(mi2p))
        pb.mvals = _fp(self.moving)
        self._mmask = None
        if moving_mask is not None:
            self._mmask = np.ascontiguousarray(moving_mask, dtype=np.uint8)
            pb.mmask = self._mmask.ctypes.data_as(C.POINTER(C.c_uint8))
        self._mcoef = None
        if interpolator == "bspline":
            self._mcoef = bspline_coefficients(self.moving)
            pb.mcoef = self._mcoef.ctypes.data_as(C.POINTER(C.c_double))
        elif interpolator != "linear":
            raise ValueError(f"unknown interpolator {interpolator!r}")
        self.pb = pb

    def fixed_range(self):
        out = np.empty(2, dtype=np.float64)
        lib().orc_fixed_value_range(C.byref(self.pb), _dp(out))
        return float(out[0]), float(out[1])

    def moving_range(self):
        out = np.empty(2, dtype=np.float64)
        lib().orc_moving_value_range(C.byref(self.pb), _dp(out))
        return float(out[0]), float(out[1])

    def mattes_geometry(self, bins: int) -> MattesGeom:
        g = MattesGeom()
        fr = np.array(self.fixed_range())
        mr = np.array(self.moving_range())
        rc = lib().orc_mattes_geometry(_dp(fr), _dp(mr), bins, C.byref(g))
        if rc != 0:
            raise ValueError({-1: "too few bins", -2: "constant fixed",
                              -3: "constant moving"}[rc])
        return g

    def mattes_histogram(self, g: MattesGeom, t: Transform):
        b = g.num_bins
        joint = np.empty(b * b, dtype=np.float64)
        marg = np.empty(b, dtype=np.float64)
        valid = C.c_size_t(0)
        lib().orc_mattes_histogram(C.byref(self.pb), C.byref(g), C.byref(t), _dp(joint),
                                   _dp(marg), C.byref(valid))
        return joint, marg, valid.value

    def mattes_value(self, g: MattesGeom, t: Transform) -> float:
        return lib().orc_mattes_value(C.byref(self.pb), C.byref(g), C.byref(t))

    def mattes_evaluate(self, g: MattesGeom, t: Transform):
        deriv = np.empty(t.nparams, dtype=np.float64)
        valid = C.c_size_t(0)
        v = lib().orc_mattes_evaluate(C.byref(self.pb), C.byref(g), C.byref(t), _dp(deriv),
                                      C.byref(valid))
        return v, deriv, valid.value

    def mattes_evaluate_composed(self, g: MattesGeom, t: Transform, init):
        """B-spline t followed by the linear moving-initial transform init
        (None = no composition)."""
        deriv = np.empty(t.nparams, dtype=np.float64)
        valid = C.c_size_t(0)
        v = lib().orc_mattes_evaluate_bspline_composed(
            C.byref(self.pb), C.byref(g), C.byref(t), C.byref(init) if init is not None else None,
            _dp(deriv), C.byref(valid))
        return v, deriv, valid.value

    def mattes_tail(self, joint, marg, valid, g: MattesGeom):
        j = np.array(joint, dtype=np.float64)
        m = np.array(marg, dtype=np.float64)
        pr = np.empty(g.num_bins ** 2, dtype=np.float64)
        value = C.c_double(0)
        ok = lib().orc_mattes_tail(_dp(j), _dp(m), valid, C.byref(g), C.byref(value), _dp(pr))
        return value.value, (pr if ok else None)

    def mattes_moments(self, g: MattesGeom, t: Transform, pratio) -> np.ndarray:
        pr = np.ascontiguousarray(pratio, dtype=np.float64)
        out = np.empty(12, dtype=np.float64)
        lib().orc_mattes_moments(C.byref(self.pb), C.byref(g), C.byref(t), _dp(pr), _dp(out))
        return out

    def ms_evaluate(self, t: Transform, threads: int = 1):
        """Mean squares value and derivative (metric.rs:1065-1179); with
        threads > 1 the rows are computed in parallel and added in order, as
        the reference does (the same bits)."""
        deriv = np.empty(t.nparams, dtype=np.float64)
        valid = C.c_size_t(0)
        if threads > 1:
            v = lib().orc_ms_evaluate_threads(C.byref(self.pb), C.byref(t), _dp(deriv),
                                              C.byref(valid), int(threads))
        else:
            v = lib().orc_ms_evaluate(C.byref(self.pb), C.byref(t), _dp(deriv), C.byref(valid))
        return v, deriv, valid.value

    def ms_value(self, t: Transform):
        valid = C.c_size_t(0)
        v = lib().orc_ms_value(C.byref(self.pb), C.byref(t), C.byref(valid))
        return v, valid.value

    def value_and_gradient_at(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        v = C.c_double(0)
        g = np.empty(3, dtype=np.float64)
        ok = lib().orc_moving_value_and_physical_gradient(C.byref(self.pb), _dp(p), C.byref(v),
                                                          _dp(g))
        return (v.value, g) if ok else None


# ---------------------------------------------------------------- fixtures
def fixture_volume(n: int, shift) -> np.ndarray:
    out = np.empty(n ** 3, dtype=np.float32)
    lib().orc_fixture_volume(n, _d3(shift), _fp(out))
    return out.reshape(n, n, n)


def fixture_inverted(n: int, shift) -> np.ndarray:
    out = np.empty(n ** 3, dtype=np.float32)
    lib().orc_fixture_inverted(n, _d3(shift), _fp(out))
    return out.reshape(n, n, n)


def fixture_commensurate(n: int) -> np.ndarray:
    out = np.empty(n ** 3, dtype=np.float32)
    lib().orc_fixture_commensurate(n, _fp(out))
    return out.reshape(n, n, n)


def cubic_bspline(u: float) -> float:
    return lib().orc_cubic_bspline(u)


def cubic_bspline_derivative(u: float) -> float:
    return lib().orc_cubic_bspline_derivative(u)


def set_joint_sum_compensated(on: bool) -> None:
    lib().orc_set_joint_sum_compensated(int(on))


def compensated_sum(v) -> float:
    a = np.ascontiguousarray(v, dtype=np.float64)
    return lib().orc_compensated_sum(_dp(a), a.size)


def resample_nearest(vol, in_spacing, in_origin, out_size, out_spacing, out_origin, t,
                     default=0.0, in_dir=None, out_dir=None):
    """As resample_linear with the nearest-neighbour interpolator (round half
    up, transform/interpolator.rs:37-54)."""
    return resample_linear(vol, in_spacing, in_origin, out_size, out_spacing, out_origin, t,
                           default, in_dir, out_dir, _fn="orc_resample_nearest")


def resample_linear(vol, in_spacing, in_origin, out_size, out_spacing, out_origin, t,
                    default=0.0, in_dir=None, out_dir=None, _fn="orc_resample_linear"):
    """vol: (nz, ny, nx) float32; out_size (nx, ny, nz). Returns (nz', ny', nx')."""
    v = np.ascontiguousarray(vol, dtype=np.float32)
    nz, ny, nx = v.shape
    eye = (C.c_double * 9)(*np.eye(3).reshape(9))
    idir = eye if in_dir is None else (C.c_double * 9)(*np.asarray(in_dir, float).reshape(9))
    odir = eye if out_dir is None else (C.c_double * 9)(*np.asarray(out_dir, float).reshape(9))
    out = np.empty(int(np.prod(out_size)), dtype=np.float32)
    getattr(lib(), _fn)(_fp(v), _szp((nx, ny, nz)), _d3(in_spacing), _d3(in_origin), idir,
                              _szp(out_size), _d3(out_spacing), _d3(out_origin), odir,
                              C.byref(t), default, _fp(out))
    return out.reshape(out_size[2], out_size[1], out_size[0])


def shrunk_geometry(size, spacing, origin, factors, direction=None):
    d = (C.c_double * 9)(*(np.eye(3) if direction is None else np.asarray(direction, float)).reshape(9))
    osz = (C.c_size_t * 3)()
    osp = (C.c_double * 3)()
    oor = (C.c_double * 3)()
    lib().orc_shrunk_geometry(_szp(size), _d3(spacing), _d3(origin), d, _szp(factors), osz, osp, oor)
    return list(osz), list(osp), list(oor)
