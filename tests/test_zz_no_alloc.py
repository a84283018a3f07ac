"""The allocation-free smoothing entry point, sitkhip_recursive_gaussian_into:
through the C ABI, its output equals sitkhip_recursive_gaussian's bit for bit
(and the oracle's, crates/sitk/src/filters/recursive_gaussian.rs:171-204,
472-528) on two calls in a row whose second image is smaller, so the second
call runs on the scratch the first one left behind. Also its refusals."""

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def _case(hip, shape, spacing, sigma, seed):
    nz, ny, nx = shape
    v = O.synth(seed, nx * ny * nz).reshape(shape)
    img = hip.DeviceImage.upload(v, spacing=spacing)
    want = hip.recursive_gaussian(img, sigma).to_host()
    out = hip.DeviceImage.filled((nx, ny, nz), -7.0, spacing=spacing)
    got = hip.recursive_gaussian_into(img, sigma, out).to_host()
    return v, want, got


@pytest.mark.parametrize(
    "first,second",
    [
        # all three axes smoothed, then a smaller image with two axes
        (((160, 192, 176), (0.9, 0.9, 1.5), (2.0, 2.0, 2.0)),
         ((37, 24, 13), (1.0, 1.0, 1.0), (1.0, 0.0, 3.5))),
        # the largest image it takes (2^23 voxels), then a small anisotropic one
        (((128, 256, 256), (1.0, 1.0, 1.0), (2.0, 1.0, 3.0)),
         ((39, 41, 40), (0.9, 0.9, 1.5), (0.31, 4.7, 0.9))),
        # one axis, then none (a copy)
        (((64, 64, 64), (1.0, 1.0, 1.0), (0.0, 2.5, 0.0)),
         ((20, 24, 32), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0))),
    ],
)
def test_into_equals_the_allocating_entry_point(hip, first, second):
    for k, (shape, spacing, sigma) in enumerate((first, second)):
        v, want, got = _case(hip, shape, spacing, sigma, 42 + k)
        assert np.array_equal(_bits(got), _bits(want)), (k, shape, sigma)
        if v.size <= 64 ** 3:
            nz, ny, nx = shape
            ref = O.smoothing_recursive_gaussian(v, (nx, ny, nz), spacing, sigma)
            assert np.array_equal(_bits(got).reshape(-1), _bits(ref).reshape(-1)), (k, shape)


def test_into_golden_checksum(hip):
    v = O.synth(42, 64 ** 3).reshape(64, 64, 64)
    img = hip.DeviceImage.upload(v)
    out = hip.DeviceImage.filled((64, 64, 64), 0.0)
    hip.recursive_gaussian_into(img, [2.0, 2.0, 2.0], out)
    assert O.fnv1a64(out.to_host()) == 0x8F3DA2E917366232


def test_into_refusals(hip):
    small = hip.DeviceImage.filled((16, 16, 16), 1.0)
    other = hip.DeviceImage.filled((16, 16, 17), 0.0)
    with pytest.raises(hip.CudaError, match="geometry"):
        hip.recursive_gaussian_into(small, [1.0, 1.0, 1.0], other)
    with pytest.raises(hip.CudaError, match="input image"):
        hip.recursive_gaussian_into(small, [1.0, 1.0, 1.0], small)
    with pytest.raises(hip.CudaError, match="sigma"):
        hip.recursive_gaussian_into(small, [1.0, -1.0, 1.0],
                                    hip.DeviceImage.filled((16, 16, 16), 0.0))
    # one voxel beyond the scratch: refused on the host, before any launch
    big = (hip.GAUSSIAN_INTO_MAX_VOXELS // (256 * 256) + 1, 256, 256)
    src = hip.DeviceImage.filled((big[2], big[1], big[0]), 1.0)
    dst = hip.DeviceImage.filled((big[2], big[1], big[0]), 0.0)
    with pytest.raises(hip.CudaError, match="exceed"):
        hip.recursive_gaussian_into(src, [1.0, 1.0, 1.0], dst)
