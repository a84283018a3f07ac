"""ctypes binding of libsitk_hip (include/sitk_hip.h).

Mirrors the reference's device surface `crate::cuda` (crates/sitk/src/cuda/mod.rs:35-56)
with the same names: ``DeviceImage``, ``DeviceMask``, ``recursive_gaussian``,
``ResidentMetric`` (mean squares) and ``ResidentMattes``. Errors raise ``CudaError``
carrying the status name of the C ABI, mirroring the reference's ``CudaError``
variants (crates/sitk/src/cuda/error.rs:11-108).

This module only marshals arguments: every computation runs in the HIP library.
Loading fails loudly if ``libsitk_hip.so`` has not been built.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SITK_HIP_LIBRARY points the binding at another build of the same library.
LIB_PATH = os.environ.get("SITK_HIP_LIBRARY", os.path.join(_HERE, "libsitk_hip.so"))

OK = 0
PIXEL = {
    np.dtype(np.uint8): 0, np.dtype(np.int8): 1, np.dtype(np.uint16): 2, np.dtype(np.int16): 3,
    np.dtype(np.uint32): 4, np.dtype(np.int32): 5, np.dtype(np.uint64): 6, np.dtype(np.int64): 7,
    np.dtype(np.float32): 8, np.dtype(np.float64): 9,
}
MEAN_SQUARES, MATTES = 0, 1
INTERP_LINEAR, INTERP_BSPLINE = 0, 1  # SITKHIP_INTERP_*
PRECISION_F32, PRECISION_F64 = 0, 1
MAX_STAGES = 8


class CudaError(RuntimeError):
    """A refused or failed device operation: ``.status`` is the C status name."""

    def __init__(self, status: str, message: str):
        super().__init__(f"{status}: {message}")
        self.status = status
        self.message = message


POINTS_GRID, POINTS_EXPLICIT, POINTS_INDICES = 0, 1, 2


class FixedPointsSpec(C.Structure):
    """sitkhip_fixed_points (FixedPoints, crates/sitk/src/cuda/ops/resident.rs:464-484)."""
    _fields_ = [
        ("mode", C.c_int),
        ("points", C.POINTER(C.c_double)),
        ("indices", C.POINTER(C.c_int64)),
        ("size", C.c_size_t * 3),
        ("origin", C.c_double * 3),
        ("idx_to_phys", C.c_double * 9),
    ]


class MovingGeometrySpec(C.Structure):
    """sitkhip_moving_geometry (MovingGeometry, resident.rs:424-437)."""
    _fields_ = [
        ("len", C.c_size_t),
        ("size", C.c_size_t * 3),
        ("strides", C.c_size_t * 3),
        ("origin", C.c_double * 3),
        ("phys_to_index", C.c_double * 9),
        ("mask", C.POINTER(C.c_uint8)),
    ]


class Geometry(C.Structure):
    _fields_ = [
        ("size", C.c_size_t * 3),
        ("spacing", C.c_double * 3),
        ("origin", C.c_double * 3),
        ("direction", C.c_double * 9),
    ]

    @staticmethod
    def make(size, spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), direction=None):
        g = Geometry()
        g.size[:] = [int(v) for v in size]
        g.spacing[:] = [float(v) for v in spacing]
        g.origin[:] = [float(v) for v in origin]
        d = np.eye(3) if direction is None else np.asarray(direction, dtype=np.float64)
        g.direction[:] = [float(v) for v in d.reshape(9)]
        return g


class MattesBins(C.Structure):
    _fields_ = [
        ("bins", C.c_size_t),
        ("padding", C.c_size_t),
        ("fixed_bin_size", C.c_double),
        ("moving_bin_size", C.c_double),
        ("fixed_normalized_min", C.c_double),
        ("moving_normalized_min", C.c_double),
        ("moving_true_min", C.c_double),
        ("moving_true_max", C.c_double),
    ]


MAX_LEVELS = 8
TRANSLATION, EULER3D, AFFINE = 0, 1, 2
REGULAR_STEP, LINE_SEARCH = 0, 1
LR_MANUAL, LR_ONCE, LR_EACH_ITERATION = 0, 1, 2
SCALES_UNIT, SCALES_PHYSICAL_SHIFT = 0, 1
STOP_NAMES = {0: "MaxIterations", 1: "GradientConverged", 2: "StepTooSmall", 3: "Converged"}


class TransformSpec(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("compute_zyx", C.c_int),
        ("center", C.c_double * 3),
        ("params", C.c_double * 12),
    ]


class RegistrationConfig(C.Structure):
    _fields_ = [
        ("metric", C.c_int),
        ("bins", C.c_int),
        ("sampling", C.c_int),
        ("sampling_percentage", C.c_double),
        ("seed", C.c_uint64),
        ("precision", C.c_int),
        ("optimizer", C.c_int),
        ("learning_rate", C.c_double),
        ("minimum_step", C.c_double),
        ("iterations", C.c_int),
        ("relaxation", C.c_double),
        ("gradient_tolerance", C.c_double),
        ("learning_rate_mode", C.c_int),
        ("scales_mode", C.c_int),
        ("levels", C.c_int),
        ("shrink", (C.c_size_t * 3) * MAX_LEVELS),
        ("sigma", C.c_double * MAX_LEVELS),
        ("fixed_mask", C.c_void_p),
        ("moving_mask", C.c_void_p),
        ("interpolator", C.c_int),
    ]


class BSplineConfig(C.Structure):
    _fields_ = [("bins", C.c_int), ("sampling", C.c_int), ("sampling_percentage", C.c_double),
                ("seed", C.c_uint64), ("precision", C.c_int), ("iterations", C.c_int),
                ("max_evaluations", C.c_int), ("corrections", C.c_int), ("factr", C.c_double),
                ("pgtol", C.c_double), ("mesh", C.c_size_t * 3)]


class BSplineResult(C.Structure):
    _fields_ = [("value", C.c_double), ("valid_points", C.c_size_t), ("iterations", C.c_int),
                ("evaluations", C.c_int), ("stop", C.c_int), ("seconds", C.c_double)]


class LbfgsbOptions(C.Structure):
    _fields_ = [("corrections", C.c_int), ("factr", C.c_double), ("pgtol", C.c_double),
                ("max_iterations", C.c_int), ("max_evaluations", C.c_int)]


class RegistrationResult(C.Structure):
    _fields_ = [
        ("value", C.c_double),
        ("valid_points", C.c_size_t),
        ("iterations", C.c_int),
        ("level_iterations", C.c_int * MAX_LEVELS),
        ("level_stop", C.c_int * MAX_LEVELS),
        ("level_value", C.c_double * MAX_LEVELS),
        ("evaluations", C.c_size_t),
        ("value_evaluations", C.c_size_t),
        ("seconds", C.c_double),
    ]


_lib = None


def lib():
    """The loaded library (raises if it was never built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp, dp, szp = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_size_t)
        gp = C.POINTER(Geometry)
        L.sitkhip_metric_create_host.argtypes = [C.c_int, C.c_size_t, dp, C.POINTER(FixedPointsSpec),
                                                 C.POINTER(C.c_uint8), C.POINTER(MovingGeometrySpec),
                                                 dp, C.c_int, C.POINTER(vp)]
        L.sitkhip_status_name.restype = C.c_char_p
        L.sitkhip_status_name.argtypes = [C.c_int]
        L.sitkhip_last_error.restype = C.c_char_p
        L.sitkhip_backend_init.argtypes = [C.c_int]
        L.sitkhip_backend_stream.argtypes = [C.POINTER(vp)]
        L.sitkhip_device_name.argtypes = [C.c_char_p, C.c_size_t]
        L.sitkhip_image_upload.argtypes = [C.c_int, vp, C.c_size_t, gp, C.POINTER(vp)]
        L.sitkhip_image_filled.argtypes = [gp, C.c_float, C.POINTER(vp)]
        L.sitkhip_image_to_host.argtypes = [vp, C.POINTER(C.c_float), C.c_size_t]
        L.sitkhip_image_geometry.argtypes = [vp, gp]
        L.sitkhip_image_len.argtypes = [vp]
        L.sitkhip_image_len.restype = C.c_size_t
        L.sitkhip_image_id.argtypes = [vp]
        L.sitkhip_image_id.restype = C.c_uint64
        L.sitkhip_image_device_ptr.argtypes = [vp]
        L.sitkhip_image_device_ptr.restype = C.c_void_p
        L.sitkhip_image_free.argtypes = [vp]
        L.sitkhip_image_free.restype = None
        L.sitkhip_mask_upload.argtypes = [C.POINTER(C.c_uint8), C.c_size_t, gp, C.POINTER(vp)]
        L.sitkhip_mask_from_image.argtypes = [vp, C.POINTER(vp)]
        L.sitkhip_mask_intersect.argtypes = [vp, vp, C.POINTER(vp)]
        L.sitkhip_mask_to_host.argtypes = [vp, C.POINTER(C.c_uint8), C.c_size_t]
        L.sitkhip_mask_free.argtypes = [vp]
        L.sitkhip_mask_free.restype = None
        L.sitkhip_recursive_gaussian.argtypes = [vp, dp, C.POINTER(vp)]
        L.sitkhip_recursive_gaussian_batch.argtypes = [C.c_size_t, C.POINTER(vp), dp, C.POINTER(vp)]
        L.sitkhip_recursive_gaussian_into.argtypes = [vp, dp, vp]
        L.sitkhip_metric_set_interpolator.argtypes = [vp, C.c_int]
        L.sitkhip_mattes_value_and_moments_repeat.argtypes = [vp, dp, C.c_int, C.c_int, dp, dp, szp]
        L.sitkhip_metric_bspline_coefficients.argtypes = [vp, dp, C.c_size_t]
        L.sitkhip_metric_create.argtypes = [C.c_int, vp, C.POINTER(C.c_int64), C.c_size_t, vp, vp,
                                            vp, C.c_int, C.POINTER(vp)]
        L.sitkhip_metric_create_drawn.argtypes = [C.c_int, vp, C.c_size_t, C.c_uint64, vp, C.c_int,
                                                  C.POINTER(vp)]
        L.sitkhip_metric_free.argtypes = [vp]
        L.sitkhip_metric_free.restype = None
        L.sitkhip_metric_sample_count.argtypes = [vp, szp]
        L.sitkhip_metric_volume_bytes.argtypes = [vp, szp]
        L.sitkhip_metric_value_ranges.argtypes = [vp, dp, dp]
        L.sitkhip_mattes_bins_from_ranges.argtypes = [dp, dp, C.c_size_t, C.POINTER(MattesBins)]
        L.sitkhip_mattes_set_bins.argtypes = [vp, C.POINTER(MattesBins)]
        L.sitkhip_mattes_joint_histogram.argtypes = [vp, dp, C.c_int, dp, dp, szp]
        L.sitkhip_mattes_derivative_moments.argtypes = [vp, dp, C.c_int, dp, dp]
        L.sitkhip_mattes_value.argtypes = [vp, dp, C.c_int, dp, szp]
        L.sitkhip_mattes_value_and_moments.argtypes = [vp, dp, C.c_int, dp, dp, szp]
        L.sitkhip_ms_moments.argtypes = [vp, dp, C.c_int, dp]
        L.sitkhip_ms_value.argtypes = [vp, dp, C.c_int, dp, szp]
        L.sitkhip_contract_ms.argtypes = [dp, dp, dp, C.c_int, dp, dp, szp]
        L.sitkhip_contract_mattes.argtypes = [dp, dp, dp, C.c_int, dp]
        L.sitkhip_shrunk_geometry.argtypes = [gp, szp, gp]
        L.sitkhip_shrink.argtypes = [vp, szp, C.POINTER(vp)]
        L.sitkhip_resample_linear.argtypes = [vp, gp, dp, C.c_int, C.c_double, C.POINTER(vp)]
        L.sitkhip_resample_nearest.argtypes = [vp, gp, dp, C.c_int, C.c_double, C.POINTER(vp)]
        L.sitkhip_mask_resample_nearest.argtypes = [vp, gp, dp, C.c_int, C.POINTER(vp)]
        L.sitkhip_register.argtypes = [vp, vp, C.POINTER(RegistrationConfig),
                                       C.POINTER(TransformSpec), C.POINTER(RegistrationResult)]
        L.sitkhip_draw_samples.argtypes = [C.c_size_t, C.c_int, C.c_double, C.c_uint64,
                                           C.POINTER(C.c_int64), szp]
        L.sitkhip_batch_create.argtypes = [C.POINTER(vp), C.c_size_t, C.POINTER(vp)]
        L.sitkhip_batch_free.argtypes = [vp]
        L.sitkhip_batch_free.restype = None
        L.sitkhip_batch_set_tiling.argtypes = [vp, C.c_size_t, C.c_uint, C.c_int]
        L.sitkhip_batch_get_tiling.argtypes = [vp, szp, C.POINTER(C.c_uint)]
        L.sitkhip_batch_autotune.argtypes = [vp, dp, szp, C.POINTER(C.c_uint), dp]
        L.sitkhip_batch_evaluate.argtypes = [vp, dp, C.POINTER(C.c_uint8), dp, dp, szp]
        L.sitkhip_register_batch.argtypes = [C.c_size_t, C.POINTER(vp), C.POINTER(vp),
                                             C.POINTER(RegistrationConfig),
                                             C.POINTER(TransformSpec),
                                             C.POINTER(RegistrationResult), C.c_int]
        L.sitkhip_register_batch_probe.argtypes = [C.c_size_t, C.POINTER(vp), C.POINTER(vp),
                                                   C.POINTER(RegistrationConfig),
                                                   C.POINTER(TransformSpec),
                                                   C.POINTER(RegistrationResult), C.c_int, dp, dp]
        fp = C.POINTER(C.c_float)
        L.sitkhip_bspline_create.argtypes = [vp, szp, dp, dp, dp, C.POINTER(vp)]
        L.sitkhip_bspline_free.argtypes = [vp]
        L.sitkhip_bspline_free.restype = None
        L.sitkhip_bspline_parameter_count.argtypes = [vp, szp]
        L.sitkhip_bspline_mattes_value.argtypes = [vp, dp, dp, szp]
        L.sitkhip_bspline_mattes_value_and_derivative.argtypes = [vp, dp, dp, dp, szp]
        L.sitkhip_register_bspline.argtypes = [vp, vp, C.POINTER(BSplineConfig), dp, C.c_size_t,
                                               C.POINTER(BSplineResult)]
        L.sitkhip_bspline_set_moving_initial.argtypes = [vp, dp]
        L.sitkhip_bspline_multi_create.argtypes = [vp, C.c_size_t, dp, C.POINTER(vp)]
        L.sitkhip_bspline_multi_free.argtypes = [vp]
        L.sitkhip_bspline_multi_free.restype = None
        L.sitkhip_bspline_multi_value_and_derivative.argtypes = [vp, C.c_size_t, C.POINTER(C.c_int),
                                                                 dp, dp, dp, szp]
        L.sitkhip_register_bspline_multistart.argtypes = [
            vp, vp, C.POINTER(BSplineConfig), C.c_size_t, dp, dp, C.c_size_t, C.POINTER(C.c_int),
            dp, dp, C.POINTER(BSplineResult), szp]
        ip = C.POINTER(C.c_int)
        L.sitkhip_lbfgsb_create.argtypes = [C.c_size_t, C.POINTER(LbfgsbOptions), dp, dp, ip,
                                            C.POINTER(vp)]
        L.sitkhip_lbfgsb_free.argtypes = [vp]
        L.sitkhip_lbfgsb_free.restype = None
        L.sitkhip_lbfgsb_start.argtypes = [vp, dp, dp, ip]
        L.sitkhip_lbfgsb_tell.argtypes = [vp, C.c_double, dp, dp, ip]
        L.sitkhip_lbfgsb_result.argtypes = [vp, dp, dp, ip, ip, ip]
        L.sitkhip_synth_blobs.argtypes = [gp, C.c_int, dp, dp, dp, C.c_double, C.c_uint64,
                                          C.POINTER(vp)]
        L.sitkhip_cnn_create.argtypes = [C.c_int, C.c_uint64, C.POINTER(vp)]
        L.sitkhip_cnn_free.argtypes = [vp]
        L.sitkhip_cnn_free.restype = None
        L.sitkhip_cnn_param_count.argtypes = [vp, szp]
        L.sitkhip_cnn_get_params.argtypes = [vp, fp, C.c_size_t]
        L.sitkhip_cnn_set_params.argtypes = [vp, fp, C.c_size_t]
        L.sitkhip_cnn_get_grads.argtypes = [vp, fp, C.c_size_t]
        L.sitkhip_cnn_set_input.argtypes = [vp, C.c_int, vp, vp]
        L.sitkhip_cnn_set_input_raw.argtypes = [vp, C.c_int, fp, fp]
        L.sitkhip_cnn_set_inputs.argtypes = [vp, C.c_int, C.c_int, C.POINTER(vp), C.POINTER(vp)]
        L.sitkhip_cnn_forward.argtypes = [vp, C.c_int, fp]
        L.sitkhip_cnn_backward.argtypes = [vp, C.c_int, fp]
        L.sitkhip_cnn_adam.argtypes = [vp, C.c_float, C.c_float, C.c_float, C.c_float]
        L.sitkhip_cnn_train_step.argtypes = [vp, C.c_int, fp, C.c_float, fp]
        L.sitkhip_cnn_activation.argtypes = [vp, C.c_int, C.c_int, C.c_int, fp, C.c_size_t]
        L.sitkhip_empty_cache.argtypes = []
        L.sitkhip_profile_enable.argtypes = [C.c_int]
        L.sitkhip_profile_query.argtypes = [C.c_char_p, C.POINTER(C.c_uint64), dp]
        L.sitkhip_profile_dump.argtypes = [C.c_char_p, C.c_size_t]
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != OK:
        L = lib()
        raise CudaError(L.sitkhip_status_name(rc).decode(), L.sitkhip_last_error().decode())


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _stages(stages) -> tuple:
    s = np.ascontiguousarray(stages, dtype=np.float64).reshape(-1)
    if s.size % 12:
        raise ValueError("stages must be a multiple of 12 doubles")
    return s, s.size // 12


def backend_init(device: int = 0) -> None:
    _check(lib().sitkhip_backend_init(device))


def device_name() -> str:
    buf = C.create_string_buffer(256)
    _check(lib().sitkhip_device_name(buf, 256))
    return buf.value.decode()


def backend_stream() -> int:
    p = C.c_void_p()
    _check(lib().sitkhip_backend_stream(C.byref(p)))
    return p.value or 0


class DeviceImage:
    """DeviceImage (crates/sitk/src/cuda/image.rs:162-413): a float32 device volume.

    Arrays are (nz, ny, nx) numpy arrays (dimension 0 = x fastest in memory).
    """

    def __init__(self, handle):
        self._h = handle

    @staticmethod
    def upload(arr: np.ndarray, spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), direction=None):
        a = np.ascontiguousarray(arr)
        if a.dtype not in PIXEL:
            raise CudaError("UnsupportedPixelType", f"no upload cast for {a.dtype}")
        nz, ny, nx = a.shape
        g = Geometry.make((nx, ny, nz), spacing, origin, direction)
        h = C.c_void_p()
        _check(lib().sitkhip_image_upload(PIXEL[a.dtype], a.ctypes.data, a.size, C.byref(g),
                                          C.byref(h)))
        return DeviceImage(h)

    @staticmethod
    def filled(size, value: float, spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0),
               direction=None):
        g = Geometry.make(size, spacing, origin, direction)
        h = C.c_void_p()
        _check(lib().sitkhip_image_filled(C.byref(g), float(value), C.byref(h)))
        return DeviceImage(h)

    def geometry(self) -> Geometry:
        g = Geometry()
        _check(lib().sitkhip_image_geometry(self._h, C.byref(g)))
        return g

    def __len__(self) -> int:
        return int(lib().sitkhip_image_len(self._h))

    def id(self) -> int:
        return int(lib().sitkhip_image_id(self._h))

    def device_ptr(self) -> int:
        return int(lib().sitkhip_image_device_ptr(self._h) or 0)

    def to_host(self) -> np.ndarray:
        g = self.geometry()
        out = np.empty(len(self), dtype=np.float32)
        _check(lib().sitkhip_image_to_host(self._h, out.ctypes.data_as(C.POINTER(C.c_float)),
                                           out.size))
        return out.reshape(g.size[2], g.size[1], g.size[0])

    def free(self) -> None:
        if self._h:
            lib().sitkhip_image_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceMask:
    """DeviceMask (crates/sitk/src/cuda/mask.rs:82-216): 1 inside, 0 outside."""

    def __init__(self, handle):
        self._h = handle

    @staticmethod
    def upload(arr: np.ndarray, spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), direction=None):
        a = np.ascontiguousarray(arr, dtype=np.uint8)
        nz, ny, nx = a.shape
        g = Geometry.make((nx, ny, nz), spacing, origin, direction)
        h = C.c_void_p()
        _check(lib().sitkhip_mask_upload(a.ctypes.data_as(C.POINTER(C.c_uint8)), a.size,
                                         C.byref(g), C.byref(h)))
        return DeviceMask(h)

    @staticmethod
    def from_device_image(img: DeviceImage):
        h = C.c_void_p()
        _check(lib().sitkhip_mask_from_image(img._h, C.byref(h)))
        return DeviceMask(h)

    def intersect(self, other: "DeviceMask") -> "DeviceMask":
        h = C.c_void_p()
        _check(lib().sitkhip_mask_intersect(self._h, other._h, C.byref(h)))
        return DeviceMask(h)

    def to_host(self, shape) -> np.ndarray:
        n = int(np.prod(shape))
        out = np.empty(n, dtype=np.uint8)
        _check(lib().sitkhip_mask_to_host(self._h, out.ctypes.data_as(C.POINTER(C.c_uint8)), n))
        return out.reshape(shape)

    def free(self) -> None:
        if self._h:
            lib().sitkhip_mask_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def recursive_gaussian(src: DeviceImage, sigma) -> DeviceImage:
    """recursive_gaussian (crates/sitk/src/cuda/ops/pyramid.rs:217)."""
    s = np.ascontiguousarray(sigma, dtype=np.float64)
    if s.size != 3:
        raise CudaError("UnsupportedGeometry", "sigma must have 3 entries")
    h = C.c_void_p()
    _check(lib().sitkhip_recursive_gaussian(src._h, _dp(s), C.byref(h)))
    return DeviceImage(h)


# SITKHIP_GAUSSIAN_INTO_MAX_VOXELS: the largest image recursive_gaussian_into takes
GAUSSIAN_INTO_MAX_VOXELS = 8_388_608


def recursive_gaussian_into(src: DeviceImage, sigma, out: DeviceImage) -> DeviceImage:
    """recursive_gaussian into the caller's image `out` (the input's geometry),
    allocating no device memory (no counterpart in the reference): for callers
    that cannot allocate. At most GAUSSIAN_INTO_MAX_VOXELS voxels; calls that
    overlap on one device share its scratch and race."""
    s = np.ascontiguousarray(sigma, dtype=np.float64)
    if s.size != 3:
        raise CudaError("UnsupportedGeometry", "sigma must have 3 entries")
    _check(lib().sitkhip_recursive_gaussian_into(src._h, _dp(s), out._h))
    return out


def recursive_gaussian_batch(srcs, sigmas) -> list:
    """recursive_gaussian of several images in a few multi-image launches
    (no counterpart in the reference); the same bits as one call per image."""
    n = len(srcs)
    s = np.ascontiguousarray(sigmas, dtype=np.float64).reshape(n, 3)
    ins = (C.c_void_p * n)(*[x._h for x in srcs])
    outs = (C.c_void_p * n)()
    _check(lib().sitkhip_recursive_gaussian_batch(C.c_size_t(n), ins, _dp(s), outs))
    return [DeviceImage(outs[i]) for i in range(n)]


class _Resident:
    def __init__(self, kind, fixed: DeviceImage, moving: DeviceImage, indices=None,
                 fixed_mask: DeviceMask | None = None, moving_mask: DeviceMask | None = None,
                 precision: int = PRECISION_F32, device_draws=None):
        h = C.c_void_p()
        if device_draws is not None:
            n, seed = device_draws
            _check(lib().sitkhip_metric_create_drawn(kind, fixed._h, int(n), int(seed), moving._h,
                                                     precision, C.byref(h)))
            self._h = h
            return
        if indices is None:
            ip, n = None, 0
        else:
            self._idx = np.ascontiguousarray(indices, dtype=np.int64)
            ip, n = self._idx.ctypes.data_as(C.POINTER(C.c_int64)), self._idx.size
        _check(lib().sitkhip_metric_create(kind, fixed._h, ip, n,
                                           fixed_mask._h if fixed_mask else None, moving._h,
                                           moving_mask._h if moving_mask else None, precision,
                                           C.byref(h)))
        self._h = h

    @classmethod
    def from_host(cls, fixed_values, moving_values, moving_size, moving_origin, moving_p2i,
                  points=None, indices=None, grid=None, fixed_mask=None, moving_mask=None,
                  precision: int = PRECISION_F32, bins=None):
        """ResidentMetric::new (crates/sitk/src/cuda/ops/mean_squares.rs:201-256):
        both volumes from host values. The sample set is `points` (explicit,
        n x 3 physical; fixed_values then holds one value per point), or the
        fixed grid `grid` = (size, origin, idx_to_phys 3x3) with every voxel
        or the voxels `indices` (fixed_values then holds the whole grid)."""
        self = cls.__new__(cls)
        self._h = None
        fv = np.ascontiguousarray(fixed_values, dtype=np.float64).reshape(-1)
        mv = np.ascontiguousarray(moving_values, dtype=np.float64).reshape(-1)
        spec = FixedPointsSpec()
        keep = [fv, mv]
        if points is not None:
            pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
            keep.append(pts)
            spec.mode = POINTS_EXPLICIT
            spec.points = _dp(pts)
            n = pts.shape[0]
        else:
            size, origin, i2p = grid
            spec.size[:] = [int(v) for v in size]
            spec.origin[:] = [float(v) for v in origin]
            spec.idx_to_phys[:] = [float(v) for v in np.asarray(i2p, dtype=np.float64).reshape(9)]
            if indices is not None:
                idx = np.ascontiguousarray(indices, dtype=np.int64)
                keep.append(idx)
                spec.mode = POINTS_INDICES
                spec.indices = idx.ctypes.data_as(C.POINTER(C.c_int64))
                n = idx.size
            else:
                spec.mode = POINTS_GRID
                n = int(np.prod(size))
        mg = MovingGeometrySpec()
        ms = [int(v) for v in moving_size]
        mg.len = ms[0] * ms[1] * ms[2]
        mg.size[:] = ms
        mg.strides[:] = [1, ms[0], ms[0] * ms[1]]
        mg.origin[:] = [float(v) for v in moving_origin]
        mg.phys_to_index[:] = [float(v) for v in np.asarray(moving_p2i, dtype=np.float64).reshape(9)]
        if moving_mask is not None:
            mm = np.ascontiguousarray(moving_mask, dtype=np.uint8).reshape(-1)
            keep.append(mm)
            mg.mask = mm.ctypes.data_as(C.POINTER(C.c_uint8))
        fmp = None
        if fixed_mask is not None:
            fm = np.ascontiguousarray(fixed_mask, dtype=np.uint8).reshape(-1)
            keep.append(fm)
            fmp = fm.ctypes.data_as(C.POINTER(C.c_uint8))
        kind = MATTES if cls is ResidentMattes else MEAN_SQUARES
        h = C.c_void_p()
        _check(lib().sitkhip_metric_create_host(kind, n, _dp(fv), C.byref(spec), fmp, C.byref(mg),
                                                _dp(mv), precision, C.byref(h)))
        self._h = h
        if cls is ResidentMattes:
            self.bins = None
            if bins is not None:
                fr, mr = self.value_ranges()
                self.set_bins(mattes_bins(fr, mr, bins))
        return self

    def sample_count(self) -> int:
        v = C.c_size_t()
        _check(lib().sitkhip_metric_sample_count(self._h, C.byref(v)))
        return v.value

    def volume_bytes(self) -> int:
        v = C.c_size_t()
        _check(lib().sitkhip_metric_volume_bytes(self._h, C.byref(v)))
        return v.value

    def value_ranges(self):
        fr = np.empty(2)
        mr = np.empty(2)
        _check(lib().sitkhip_metric_value_ranges(self._h, _dp(fr), _dp(mr)))
        return (float(fr[0]), float(fr[1])), (float(mr[0]), float(mr[1]))

    def set_interpolator(self, interpolator: int) -> None:
        """The moving image's interpolator (MovingImage::from_image_with_interpolator,
        registration/metric.rs:720-750): INTERP_LINEAR or INTERP_BSPLINE (cubic;
        its coefficients are computed on the device here)."""
        _check(lib().sitkhip_metric_set_interpolator(self._h, int(interpolator)))

    def bspline_coefficients(self, n: int) -> np.ndarray:
        """The cubic B-spline coefficients of the moving volume (n voxels), FP64."""
        out = np.empty(n)
        _check(lib().sitkhip_metric_bspline_coefficients(self._h, _dp(out), n))
        return out

    def free(self) -> None:
        if self._h:
            lib().sitkhip_metric_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class ResidentMetric(_Resident):
    """Mean squares (crates/sitk/src/cuda/ops/mean_squares.rs:186-439)."""

    def __init__(self, fixed, moving, **kw):
        super().__init__(MEAN_SQUARES, fixed, moving, **kw)

    def evaluate(self, stages) -> np.ndarray:
        """The 14 moments [sq, S0[3], S1[3][3], count]."""
        s, n = _stages(stages)
        out = np.empty(14)
        _check(lib().sitkhip_ms_moments(self._h, _dp(s), n, _dp(out)))
        return out

    def value(self, stages):
        s, n = _stages(stages)
        sq = C.c_double()
        valid = C.c_size_t()
        _check(lib().sitkhip_ms_value(self._h, _dp(s), n, C.byref(sq), C.byref(valid)))
        return sq.value, valid.value


def mattes_bins(fixed_range, moving_range, bins: int) -> MattesBins:
    """MattesGeometry::new (crates/sitk/src/registration/mattes.rs:152-186)."""
    fr = np.asarray(fixed_range, dtype=np.float64)
    mr = np.asarray(moving_range, dtype=np.float64)
    b = MattesBins()
    _check(lib().sitkhip_mattes_bins_from_ranges(_dp(fr), _dp(mr), bins, C.byref(b)))
    return b


class ResidentMattes(_Resident):
    """Mattes MI (crates/sitk/src/cuda/ops/mattes.rs:500-934)."""

    def __init__(self, fixed, moving, bins: int | None = None, **kw):
        super().__init__(MATTES, fixed, moving, **kw)
        self.bins = None
        if bins is not None:
            fr, mr = self.value_ranges()
            self.set_bins(mattes_bins(fr, mr, bins))

    def set_bins(self, b: MattesBins) -> None:
        _check(lib().sitkhip_mattes_set_bins(self._h, C.byref(b)))
        self.bins = b

    def joint_histogram(self, stages):
        s, n = _stages(stages)
        B = int(self.bins.bins)
        joint = np.empty(B * B)
        marg = np.empty(B)
        valid = C.c_size_t()
        _check(lib().sitkhip_mattes_joint_histogram(self._h, _dp(s), n, _dp(joint), _dp(marg),
                                                    C.byref(valid)))
        return joint, marg, valid.value

    def derivative_moments(self, stages, pratio) -> np.ndarray:
        s, n = _stages(stages)
        pr = np.ascontiguousarray(pratio, dtype=np.float64)
        out = np.empty(12)
        _check(lib().sitkhip_mattes_derivative_moments(self._h, _dp(s), n, _dp(pr), _dp(out)))
        return out

    def value(self, stages):
        s, n = _stages(stages)
        v = C.c_double()
        valid = C.c_size_t()
        _check(lib().sitkhip_mattes_value(self._h, _dp(s), n, C.byref(v), C.byref(valid)))
        return v.value, valid.value

    def value_and_moments(self, stages):
        s, n = _stages(stages)
        v = C.c_double()
        valid = C.c_size_t()
        out = np.empty(12)
        _check(lib().sitkhip_mattes_value_and_moments(self._h, _dp(s), n, C.byref(v), _dp(out),
                                                      C.byref(valid)))
        return v.value, out, valid.value

    def value_and_moments_repeat(self, stages, repeats: int):
        """value_and_moments `repeats` times in one call (timing without a
        foreign call per evaluation); returns the last result."""
        s, n = _stages(stages)
        v = C.c_double()
        valid = C.c_size_t()
        out = np.empty(12)
        _check(lib().sitkhip_mattes_value_and_moments_repeat(self._h, _dp(s), n, int(repeats),
                                                             C.byref(v), _dp(out), C.byref(valid)))
        return v.value, out, valid.value


def contract_ms(moments, j0, c, nparams: int):
    """contract (crates/sitk/src/registration/cuda.rs:381-412)."""
    m = np.ascontiguousarray(moments, dtype=np.float64)
    j0 = np.ascontiguousarray(j0, dtype=np.float64)
    c = np.ascontiguousarray(c, dtype=np.float64)
    v = C.c_double()
    d = np.empty(nparams)
    valid = C.c_size_t()
    _check(lib().sitkhip_contract_ms(_dp(m), _dp(j0), _dp(c), nparams, C.byref(v), _dp(d),
                                     C.byref(valid)))
    return v.value, d, valid.value


def contract_mattes(moments, j0, c, nparams: int) -> np.ndarray:
    """contract_mattes (crates/sitk/src/registration/cuda.rs:505-524)."""
    m = np.ascontiguousarray(moments, dtype=np.float64)
    j0 = np.ascontiguousarray(j0, dtype=np.float64)
    c = np.ascontiguousarray(c, dtype=np.float64)
    d = np.empty(nparams)
    _check(lib().sitkhip_contract_mattes(_dp(m), _dp(j0), _dp(c), nparams, _dp(d)))
    return d


SAMPLING_NONE, SAMPLING_REGULAR, SAMPLING_RANDOM = 0, 1, 2


def draw_samples(voxels: int, strategy: int, percentage: float, seed: int):
    """draw_samples (crates/sitk/src/registration/metric.rs:419-437); None for
    the identity selection."""
    n = C.c_size_t()
    _check(lib().sitkhip_draw_samples(voxels, strategy, percentage, C.c_uint64(seed), None,
                                      C.byref(n)))
    if n.value == C.c_size_t(-1).value:
        return None
    out = np.empty(n.value, dtype=np.int64)
    _check(lib().sitkhip_draw_samples(voxels, strategy, percentage, C.c_uint64(seed),
                                      out.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(n)))
    return out


def profile_enable(on: bool) -> None:
    _check(lib().sitkhip_profile_enable(int(on)))


def profile_reset() -> None:
    _check(lib().sitkhip_profile_reset())


def profile_query(prefix: str):
    cnt = C.c_uint64()
    ms = C.c_double()
    n = lib().sitkhip_profile_query(prefix.encode(), C.byref(cnt), C.byref(ms))
    if n < 0:
        _check(-n)
    return cnt.value, ms.value


def profile_dump() -> str:
    buf = C.create_string_buffer(1 << 16)
    _check(lib().sitkhip_profile_dump(buf, 1 << 16))
    return buf.value.decode()


# ------------------------------------------------------------- pyramid placement
def shrunk_geometry(geometry: Geometry, factors) -> Geometry:
    """shrunk_geometry (crates/sitk/src/cuda/ops/pyramid.rs:355-420)."""
    out = Geometry()
    f = (C.c_size_t * 3)(*[int(v) for v in factors])
    _check(lib().sitkhip_shrunk_geometry(C.byref(geometry), f, C.byref(out)))
    return out


def shrink(img: DeviceImage, factors) -> DeviceImage:
    f = (C.c_size_t * 3)(*[int(v) for v in factors])
    h = C.c_void_p()
    _check(lib().sitkhip_shrink(img._h, f, C.byref(h)))
    return DeviceImage(h)


def _stage_arg(stages):
    if stages is None:
        return None, 0, None
    s = np.ascontiguousarray(stages, dtype=np.float64).reshape(-1)
    return _dp(s), s.size // 12, s


def resample_linear(img: DeviceImage, out_geometry: Geometry, stages=None,
                    default_value: float = 0.0) -> DeviceImage:
    """resample_linear[_through] (crates/sitk/src/cuda/ops/pyramid.rs:677-738)."""
    sp, n, keep = _stage_arg(stages)
    h = C.c_void_p()
    _check(lib().sitkhip_resample_linear(img._h, C.byref(out_geometry), sp, n, default_value,
                                         C.byref(h)))
    return DeviceImage(h)


def resample_nearest(img: DeviceImage, out_geometry: Geometry, stages=None,
                     default_value: float = 0.0) -> DeviceImage:
    """resample_nearest[_through] (crates/sitk/src/cuda/ops/pyramid.rs:739-790)."""
    sp, n, keep = _stage_arg(stages)
    h = C.c_void_p()
    _check(lib().sitkhip_resample_nearest(img._h, C.byref(out_geometry), sp, n, default_value,
                                          C.byref(h)))
    return DeviceImage(h)


def mask_resample_nearest(mask: DeviceMask, out_geometry: Geometry, stages=None) -> DeviceMask:
    """A mask placed on another grid by the nearest-neighbour resample with
    default 0, then thresholded (level_mask_on_device,
    crates/sitk/src/registration/method.rs:3068-3110)."""
    sp, n, keep = _stage_arg(stages)
    h = C.c_void_p()
    _check(lib().sitkhip_mask_resample_nearest(mask._h, C.byref(out_geometry), sp, n, C.byref(h)))
    return DeviceMask(h)


# -------------------------------------------------------------------- registration
def registration_config(metric=MATTES, bins=50, sampling=0, sampling_percentage=1.0, seed=0,
                        precision=PRECISION_F32, optimizer=REGULAR_STEP, learning_rate=1.0,
                        minimum_step=1e-4, iterations=100, relaxation=0.5,
                        gradient_tolerance=1e-4, learning_rate_mode=LR_ONCE,
                        scales_mode=SCALES_PHYSICAL_SHIFT, shrink_factors=((1, 1, 1),),
                        sigmas=(0.0,), fixed_mask=None, moving_mask=None,
                        interpolator=INTERP_LINEAR) -> RegistrationConfig:
    """fixed_mask / moving_mask: DeviceMask on the fixed / moving image's grid
    (set_metric_fixed_mask / set_metric_moving_mask). The configuration holds
    the C handles, and a reference to each mask object so that the masks
    outlive it (a mask made inline in the call would otherwise be freed when
    this returns, leaving the handle dangling)."""
    c = RegistrationConfig()
    c.fixed_mask = fixed_mask._h.value if fixed_mask is not None else None
    c.moving_mask = moving_mask._h.value if moving_mask is not None else None
    c._masks = (fixed_mask, moving_mask)
    c.interpolator = interpolator
    c.metric, c.bins, c.sampling = metric, bins, sampling
    c.sampling_percentage, c.seed, c.precision = sampling_percentage, seed, precision
    c.optimizer, c.learning_rate, c.minimum_step = optimizer, learning_rate, minimum_step
    c.iterations, c.relaxation, c.gradient_tolerance = iterations, relaxation, gradient_tolerance
    c.learning_rate_mode, c.scales_mode = learning_rate_mode, scales_mode
    if len(shrink_factors) != len(sigmas) or not 1 <= len(sigmas) <= MAX_LEVELS:
        raise ValueError("one shrink factor triple and one sigma per level (1..8 levels)")
    c.levels = len(sigmas)
    for i, (f, s) in enumerate(zip(shrink_factors, sigmas)):
        c.shrink[i][:] = [int(v) for v in f]
        c.sigma[i] = float(s)
    return c


def register(fixed: DeviceImage, moving: DeviceImage, config: RegistrationConfig, kind: int,
             params, center=(0.0, 0.0, 0.0), compute_zyx=False):
    """ImageRegistrationMethod::execute_on_device
    (crates/sitk/src/registration/method.rs:2698-2879). Returns (params, result)."""
    t = TransformSpec()
    t.kind = kind
    t.compute_zyx = int(compute_zyx)
    t.center[:] = [float(v) for v in center]
    for i, v in enumerate(params):
        t.params[i] = float(v)
    r = RegistrationResult()
    _check(lib().sitkhip_register(fixed._h, moving._h, C.byref(config), C.byref(t), C.byref(r)))
    n = {TRANSLATION: 3, EULER3D: 6, AFFINE: 12}[kind]
    return np.array(t.params[:n]), r


def _transform_spec(kind, params, center=(0.0, 0.0, 0.0), compute_zyx=False) -> TransformSpec:
    t = TransformSpec()
    t.kind = kind
    t.compute_zyx = int(compute_zyx)
    t.center[:] = [float(v) for v in center]
    for i, v in enumerate(params):
        t.params[i] = float(v)
    return t


# ------------------------------------------------------------------------ batches
class Batch:
    """Many resident metrics evaluated together (one captured graph per
    evaluation). The metrics must outlive the batch."""

    def __init__(self, metrics):
        self._metrics = list(metrics)  # keep them alive
        n = len(self._metrics)
        arr = (C.c_void_p * n)(*[m._h for m in self._metrics])
        h = C.c_void_p()
        _check(lib().sitkhip_batch_create(arr, n, C.byref(h)))
        self._h = h
        self.n = n

    def set_tiling(self, tile=0, grid=0, use_graph=True) -> None:
        _check(lib().sitkhip_batch_set_tiling(self._h, tile, grid, int(use_graph)))

    def tiling(self):
        t, g = C.c_size_t(), C.c_uint()
        _check(lib().sitkhip_batch_get_tiling(self._h, C.byref(t), C.byref(g)))
        return t.value, g.value

    def autotune(self, stages):
        st = np.ascontiguousarray(stages, dtype=np.float64).reshape(self.n, 12)
        t, g, ms = C.c_size_t(), C.c_uint(), C.c_double()
        _check(lib().sitkhip_batch_autotune(self._h, _dp(st), C.byref(t), C.byref(g), C.byref(ms)))
        return t.value, g.value, ms.value

    def evaluate(self, stages, active=None):
        """Returns (values[n], moments[n, 14], valid[n])."""
        st = np.ascontiguousarray(stages, dtype=np.float64).reshape(self.n, 12)
        vals = np.full(self.n, np.nan)
        mom = np.zeros((self.n, 14))
        valid = np.zeros(self.n, dtype=np.uint64)
        ap = None
        if active is not None:
            self._act = np.ascontiguousarray(active, dtype=np.uint8)
            ap = self._act.ctypes.data_as(C.POINTER(C.c_uint8))
        _check(lib().sitkhip_batch_evaluate(self._h, _dp(st), ap, _dp(vals), _dp(mom),
                                            valid.ctypes.data_as(C.POINTER(C.c_size_t))))
        return vals, mom, valid

    def free(self) -> None:
        if self._h:
            lib().sitkhip_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def register_batch(fixed, moving, configs, kinds, params, centers, autotune=False,
                   compute_zyx=False, probe=None):
    """n registrations run together, one batched evaluation per iteration.
    Returns (list of parameter arrays, list of RegistrationResult); with
    `probe` (n parameter lists), also each pair's metric value at its probe
    parameters on the first level (sitkhip_register_batch_probe)."""
    n = len(fixed)
    fa = (C.c_void_p * n)(*[f._h for f in fixed])
    ma = (C.c_void_p * n)(*[m._h for m in moving])
    ca = (RegistrationConfig * n)(*configs)
    ta = (TransformSpec * n)(*[_transform_spec(k, p, c, compute_zyx)
                               for k, p, c in zip(kinds, params, centers)])
    ra = (RegistrationResult * n)()
    npar = {TRANSLATION: 3, EULER3D: 6, AFFINE: 12}
    if probe is None:
        _check(lib().sitkhip_register_batch(n, fa, ma, ca, ta, ra, int(autotune)))
    else:
        pp = np.zeros((n, 12))
        for i, p in enumerate(probe):
            pp[i, :len(p)] = p
        pv = np.zeros(n)
        _check(lib().sitkhip_register_batch_probe(n, fa, ma, ca, ta, ra, int(autotune), _dp(pp),
                                                  _dp(pv)))
    out = ([np.array(ta[i].params[:npar[kinds[i]]]) for i in range(n)], [ra[i] for i in range(n)])
    return out if probe is None else out + (pv,)


# ------------------------------------------------------------- learned initializer
def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class CNN:
    """The learned initializer network (64^3 x 2 input, 12 affine outputs)."""

    EDGE = 64
    CHANNELS = (8, 16, 32, 64, 64, 128)

    def __init__(self, max_batch: int = 16, seed: int = 0):
        h = C.c_void_p()
        _check(lib().sitkhip_cnn_create(max_batch, seed, C.byref(h)))
        self._h = h
        self.max_batch = max_batch
        n = C.c_size_t()
        _check(lib().sitkhip_cnn_param_count(self._h, C.byref(n)))
        self.nparams = n.value

    def params(self) -> np.ndarray:
        out = np.empty(self.nparams, dtype=np.float32)
        _check(lib().sitkhip_cnn_get_params(self._h, _fp(out), out.size))
        return out

    def set_params(self, p) -> None:
        a = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().sitkhip_cnn_set_params(self._h, _fp(a), a.size))

    def grads(self) -> np.ndarray:
        out = np.empty(self.nparams, dtype=np.float32)
        _check(lib().sitkhip_cnn_get_grads(self._h, _fp(out), out.size))
        return out

    def set_input(self, slot: int, fixed: DeviceImage, moving: DeviceImage) -> None:
        _check(lib().sitkhip_cnn_set_input(self._h, slot, fixed._h, moving._h))

    def set_inputs(self, first: int, fixed, moving) -> None:
        """Slots first .. first + len(fixed) - 1 from image pairs, in four launches."""
        n = len(fixed)
        if len(moving) != n:
            raise ValueError("one moving image per fixed image")
        fa = (C.c_void_p * n)(*[f._h for f in fixed])
        ma = (C.c_void_p * n)(*[m._h for m in moving])
        _check(lib().sitkhip_cnn_set_inputs(self._h, first, n, fa, ma))

    def set_input_raw(self, slot: int, fixed64, moving64) -> None:
        f = np.ascontiguousarray(fixed64, dtype=np.float32)
        m = np.ascontiguousarray(moving64, dtype=np.float32)
        if f.size != self.EDGE ** 3 or m.size != self.EDGE ** 3:
            raise ValueError("64^3 volumes expected")
        _check(lib().sitkhip_cnn_set_input_raw(self._h, slot, _fp(f), _fp(m)))

    def forward(self, batch: int) -> np.ndarray:
        out = np.empty((batch, 12), dtype=np.float32)
        _check(lib().sitkhip_cnn_forward(self._h, batch, _fp(out)))
        return out

    def backward(self, batch: int, dout) -> None:
        d = np.ascontiguousarray(dout, dtype=np.float32).reshape(batch, 12)
        _check(lib().sitkhip_cnn_backward(self._h, batch, _fp(d)))

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8) -> None:
        _check(lib().sitkhip_cnn_adam(self._h, lr, beta1, beta2, eps))

    def train_step(self, batch: int, targets, lr=1e-3) -> float:
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(batch, 12)
        loss = C.c_float()
        _check(lib().sitkhip_cnn_train_step(self._h, batch, _fp(t), lr, C.byref(loss)))
        return loss.value

    def activation(self, index: int, batch: int, gradient=False) -> np.ndarray:
        e = self.EDGE
        for l in range(index):
            e //= (1 if l == 0 else 2)
        c = self.CHANNELS[index]
        out = np.empty((batch, e, e, e, c), dtype=np.float32)
        _check(lib().sitkhip_cnn_activation(self._h, index, int(gradient), batch, _fp(out), out.size))
        return out

    def free(self) -> None:
        if self._h:
            lib().sitkhip_cnn_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------- generators
def synth_blobs(size, spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), centres=(), amplitudes=(),
                sigmas=(), noise=0.0, seed=0) -> DeviceImage:
    """Sum of Gaussian blobs plus seeded voxel noise, generated on the device."""
    g = Geometry.make(size, spacing, origin)
    c = np.ascontiguousarray(centres, dtype=np.float64).reshape(-1)
    a = np.ascontiguousarray(amplitudes, dtype=np.float64)
    s = np.ascontiguousarray(sigmas, dtype=np.float64)
    nb = a.size
    if c.size != 3 * nb or s.size != nb:
        raise ValueError("one centre (3), amplitude and sigma per blob")
    h = C.c_void_p()
    _check(lib().sitkhip_synth_blobs(C.byref(g), nb, _dp(c) if nb else None, _dp(a) if nb else None,
                                     _dp(s) if nb else None, float(noise), int(seed), C.byref(h)))
    return DeviceImage(h)


# ------------------------------------------------------------------------ B-spline
class BSplineMattes:
    """Mattes MI through a cubic B-spline transform with a fixed control grid
    (grid_size = mesh + 3, origin, spacing, direction). Uses `metric` (a
    ResidentMattes with bins set), which must outlive it."""

    def __init__(self, metric, grid_size, grid_origin, grid_spacing, grid_direction=None):
        self._metric = metric
        gs = (C.c_size_t * 3)(*[int(v) for v in grid_size])
        o = np.ascontiguousarray(grid_origin, dtype=np.float64)
        sp = np.ascontiguousarray(grid_spacing, dtype=np.float64)
        d = np.ascontiguousarray(np.eye(3) if grid_direction is None else grid_direction,
                                 dtype=np.float64).reshape(9)
        h = C.c_void_p()
        _check(lib().sitkhip_bspline_create(metric._h, gs, _dp(o), _dp(sp), _dp(d), C.byref(h)))
        self._h = h
        n = C.c_size_t()
        _check(lib().sitkhip_bspline_parameter_count(self._h, C.byref(n)))
        self.nparams = n.value

    def value(self, coeffs):
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        v, n = C.c_double(), C.c_size_t()
        _check(lib().sitkhip_bspline_mattes_value(self._h, _dp(c), C.byref(v), C.byref(n)))
        return v.value, n.value

    def value_and_derivative(self, coeffs):
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        if c.size != self.nparams:
            raise ValueError("3 x control-point coefficients expected")
        d = np.empty(self.nparams)
        v, n = C.c_double(), C.c_size_t()
        _check(lib().sitkhip_bspline_mattes_value_and_derivative(self._h, _dp(c), C.byref(v), _dp(d),
                                                                 C.byref(n)))
        return v.value, d, n.value

    def set_moving_initial(self, stage=None):
        """The moving-initial transform applied after the B-spline, as one
        matrix-offset stage (12 doubles), or None for none."""
        if stage is None:
            _check(lib().sitkhip_bspline_set_moving_initial(self._h, None))
        else:
            st = np.ascontiguousarray(stage, dtype=np.float64).reshape(12)
            _check(lib().sitkhip_bspline_set_moving_initial(self._h, _dp(st)))

    def free(self) -> None:
        if self._h:
            lib().sitkhip_bspline_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class BSplineMulti:
    """Several starts of one BSplineMattes evaluated in one set of launches:
    start s has its own coefficients and moving-initial stage (stages:
    nstarts x 12 or None). Uses `bspline`, which must outlive it."""

    def __init__(self, bspline: BSplineMattes, nstarts: int, stages=None):
        self._bs = bspline
        self.nstarts = int(nstarts)
        self.nparams = bspline.nparams
        h = C.c_void_p()
        st = None if stages is None else np.ascontiguousarray(stages, dtype=np.float64).reshape(-1)
        if st is not None and st.size != 12 * self.nstarts:
            raise ValueError("12 doubles per start expected")
        _check(lib().sitkhip_bspline_multi_create(bspline._h, self.nstarts,
                                                  _dp(st) if st is not None else None, C.byref(h)))
        self._h = h

    def value_and_derivative(self, coeffs, active=None):
        """coeffs: nstarts x nparams; evaluates the starts in `active` (all by
        default); returns values, derivatives and valid counts per start (only
        the active rows are written)."""
        c = np.ascontiguousarray(coeffs, dtype=np.float64).reshape(self.nstarts, self.nparams)
        act = np.arange(self.nstarts, dtype=np.int32) if active is None else \
            np.ascontiguousarray(active, dtype=np.int32)
        vals = np.zeros(self.nstarts)
        ders = np.zeros((self.nstarts, self.nparams))
        valid = np.zeros(self.nstarts, dtype=np.uint64)
        _check(lib().sitkhip_bspline_multi_value_and_derivative(
            self._h, act.size, act.ctypes.data_as(C.POINTER(C.c_int)), _dp(c), _dp(vals), _dp(ders),
            valid.ctypes.data_as(C.POINTER(C.c_size_t))))
        return vals, ders, valid

    def free(self) -> None:
        if self._h:
            lib().sitkhip_bspline_multi_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def empty_cache() -> None:
    """Return the library's cached device blocks to the runtime."""
    _check(lib().sitkhip_empty_cache())


def register_bspline(fixed: DeviceImage, moving: DeviceImage, mesh, coeffs=None, bins=32,
                     sampling=SAMPLING_NONE, sampling_percentage=1.0, seed=0,
                     precision=PRECISION_F32, iterations=100, max_evaluations=2000, corrections=5,
                     factr=1e7, pgtol=1e-5):
    """Deformable Mattes registration through a cubic B-spline transform over
    the fixed image's extent. Returns (coefficients [3 x ncp], BSplineResult)."""
    c = BSplineConfig()
    c.bins, c.sampling, c.sampling_percentage, c.seed = bins, sampling, sampling_percentage, seed
    c.precision, c.iterations, c.max_evaluations = precision, iterations, max_evaluations
    c.corrections, c.factr, c.pgtol = corrections, factr, pgtol
    c.mesh[:] = [int(v) for v in mesh]
    ncp = int(np.prod([m + 3 for m in mesh]))
    x = np.zeros(3 * ncp) if coeffs is None else np.ascontiguousarray(coeffs, dtype=np.float64).copy()
    r = BSplineResult()
    _check(lib().sitkhip_register_bspline(fixed._h, moving._h, C.byref(c), _dp(x), x.size,
                                          C.byref(r)))
    return x, r


def register_bspline_multistart(fixed: DeviceImage, moving: DeviceImage, mesh, nstarts, stages=None,
                                coeffs=None, bins=32, sampling=SAMPLING_NONE,
                                sampling_percentage=1.0, seed=0, precision=PRECISION_F32,
                                iterations=100, max_evaluations=2000, corrections=5, factr=1e7,
                                pgtol=1e-5, bound_selection=None, lower=None, upper=None):
    """nstarts deformable registrations of one pair, evaluated together each
    round; start s from moving-initial stage stages[s] (12 doubles) and
    coefficients coeffs[s]. Returns (best coefficients per start [nstarts x
    3 ncp], [BSplineResult per start], index of the best start)."""
    c = BSplineConfig()
    c.bins, c.sampling, c.sampling_percentage, c.seed = bins, sampling, sampling_percentage, seed
    c.precision, c.iterations, c.max_evaluations = precision, iterations, max_evaluations
    c.corrections, c.factr, c.pgtol = corrections, factr, pgtol
    c.mesh[:] = [int(v) for v in mesh]
    ncp = int(np.prod([m + 3 for m in mesh]))
    npar = 3 * ncp
    x = np.zeros((nstarts, npar)) if coeffs is None else \
        np.ascontiguousarray(coeffs, dtype=np.float64).reshape(nstarts, npar).copy()
    st = None if stages is None else np.ascontiguousarray(stages, dtype=np.float64).reshape(-1)
    sel = None if bound_selection is None else np.ascontiguousarray(bound_selection, dtype=np.int32)
    lo = None if lower is None else np.ascontiguousarray(lower, dtype=np.float64)
    hi = None if upper is None else np.ascontiguousarray(upper, dtype=np.float64)
    for a in (sel, lo, hi):
        if a is not None and a.size != npar:
            raise ValueError("bounds are per coefficient")
    res = (BSplineResult * nstarts)()
    best = C.c_size_t()
    _check(lib().sitkhip_register_bspline_multistart(
        fixed._h, moving._h, C.byref(c), nstarts, _dp(st) if st is not None else None, _dp(x), npar,
        sel.ctypes.data_as(C.POINTER(C.c_int)) if sel is not None else None,
        _dp(lo) if lo is not None else None, _dp(hi) if hi is not None else None, res,
        C.byref(best)))
    return x, list(res), best.value


# ------------------------------------------------------------------ L-BFGS-B
LBFGSB_STOP_NAMES = {0: "MaxIterations", 1: "GradientConverged", 2: "Converged",
                     3: "MaxFunctionEvaluations", 4: "LineSearchFailed"}


class OptimizerResult:
    """The best point found, its value, the iteration count and the stop reason
    (crates/sitk/src/registration/optimizer.rs OptimizerResult; stop_reason is
    the StopReason variant's name)."""

    def __init__(self, parameters, value, iterations, evaluations, stop_reason):
        self.parameters = parameters
        self.value = value
        self.iterations = iterations
        self.evaluations = evaluations
        self.stop_reason = stop_reason

    def __repr__(self):
        return (f"OptimizerResult(value={self.value!r}, iterations={self.iterations}, "
                f"evaluations={self.evaluations}, stop_reason={self.stop_reason!r})")


class LBFGSBOptimizer:
    """Bound-constrained limited-memory BFGS (src/registration/lbfgsb.rs:1590-1746),
    with the reference's builder API and ITK's defaults (no bounds, factr 1e7,
    pgtol 1e-5, 2000 evaluations, 5 corrections). The iterations run in the
    library (reverse communication); ``optimize`` calls ``eval(p) -> (value,
    gradient)`` wherever the library asks. Invalid settings raise CudaError
    (INVALID_ARGUMENT) where the reference panics."""

    def __init__(self, max_iterations: int):
        self.max_iterations = int(max_iterations)
        self.lower_bound = None
        self.upper_bound = None
        self.bound_selection = None
        self.factr = 1e7
        self.pgtol = 1e-5
        self.max_function_evaluations = 2000
        self.max_corrections = 5

    def set_lower_bound(self, lower):
        self.lower_bound = np.ascontiguousarray(lower, dtype=np.float64)
        return self

    def set_upper_bound(self, upper):
        self.upper_bound = np.ascontiguousarray(upper, dtype=np.float64)
        return self

    def set_bound_selection(self, selection):
        self.bound_selection = np.ascontiguousarray(selection, dtype=np.int32)
        return self

    def set_cost_function_convergence_factor(self, factr):
        self.factr = float(factr)
        return self

    def set_gradient_convergence_tolerance(self, pgtol):
        self.pgtol = float(pgtol)
        return self

    def set_max_function_evaluations(self, n):
        self.max_function_evaluations = int(n)
        return self

    def set_max_corrections(self, m):
        self.max_corrections = int(m)
        return self

    def _create(self, n):
        o = LbfgsbOptions(self.max_corrections, self.factr, self.pgtol, self.max_iterations,
                          self.max_function_evaluations)
        sel = self.bound_selection
        if sel is not None and sel.size != n:
            raise CudaError("INVALID_ARGUMENT", "bound selection length must equal parameter count")
        needs = sel is not None and bool(np.any(sel != 0))
        for b in (self.lower_bound, self.upper_bound):
            if needs and b is not None and b.size != n:
                raise CudaError("INVALID_ARGUMENT", "bound length must equal parameter count")
        h = C.c_void_p()
        ip = C.POINTER(C.c_int)
        _check(lib().sitkhip_lbfgsb_create(
            n, C.byref(o),
            _dp(self.lower_bound) if needs and self.lower_bound is not None else None,
            _dp(self.upper_bound) if needs and self.upper_bound is not None else None,
            sel.ctypes.data_as(ip) if needs else None, C.byref(h)))
        return h

    def optimize(self, initial, evaluate) -> OptimizerResult:
        x0 = np.ascontiguousarray(initial, dtype=np.float64).copy()
        n = x0.size
        if n == 0:
            raise CudaError("INVALID_ARGUMENT", "initial parameters must be non-empty")
        L = lib()
        h = self._create(n)
        try:
            x = np.empty(n)
            want = C.c_int()
            _check(L.sitkhip_lbfgsb_start(h, _dp(x0), _dp(x), C.byref(want)))
            while want.value:
                f, g = evaluate(x.copy())
                g = np.ascontiguousarray(g, dtype=np.float64)
                _check(L.sitkhip_lbfgsb_tell(h, float(f), _dp(g), _dp(x), C.byref(want)))
            best = np.empty(n)
            v = C.c_double()
            it, ev, st = C.c_int(), C.c_int(), C.c_int()
            _check(L.sitkhip_lbfgsb_result(h, _dp(best), C.byref(v), C.byref(it), C.byref(ev),
                                           C.byref(st)))
        finally:
            L.sitkhip_lbfgsb_free(h)
        return OptimizerResult(best, v.value, it.value, ev.value, LBFGSB_STOP_NAMES[st.value])
