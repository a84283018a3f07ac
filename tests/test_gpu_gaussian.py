"""Device recursive Gaussian: bit-exact against the reference's golden checksums
and against the oracle (crates/sitk/src/filters/recursive_gaussian.rs), and the
device image residency contract (crates/sitk/src/cuda/image.rs)."""

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize(
    "dim,seed,want",
    [
        (64, 42, 0x8F3DA2E917366232),
        (64, 0x2545F4914F6CDD1D, 0x9E5067C73B94EE1E),
        (256, 42, 0x9AAEA53A51FDA207),
    ],
)
def test_golden_checksums(hip, dim, seed, want):
    v = O.synth(seed, dim ** 3).reshape(dim, dim, dim)
    img = hip.DeviceImage.upload(v)
    out = hip.recursive_gaussian(img, [2.0, 2.0, 2.0]).to_host()
    assert O.fnv1a64(out) == want


def test_golden_checksum_512(hip):
    v = O.synth(42, 512 ** 3).reshape(512, 512, 512)
    img = hip.DeviceImage.upload(v)
    out = hip.recursive_gaussian(img, [2.0, 2.0, 2.0]).to_host()
    assert O.fnv1a64(out) == 0xB60665C90074C378


@pytest.mark.parametrize(
    "shape,spacing,sigma",
    [
        ((13, 24, 37), (1.0, 1.0, 1.0), (1.0, 0.0, 3.5)),       # tests/pyramid_parity.rs:130-135
        ((13, 24, 37), (1.0, 1.0, 1.0), (0.31, 4.7, 0.9)),
        ((40, 41, 39), (0.9, 0.9, 1.5), (2.0, 2.0, 2.0)),      # anisotropic spacing
        ((4, 5, 6), (1.0, 1.0, 1.0), (1.5, 1.5, 1.5)),         # shortest legal lines
        ((9, 17, 260), (1.0, 1.0, 1.0), (2.5, 0.0, 0.0)),      # one axis, long x lines
        ((160, 12, 7), (1.0, 1.0, 1.0), (0.0, 0.0, 1.25)),     # one axis, z only
        ((33, 50, 3), (1.0, 1.0, 1.0), (0.0, 1.1, 0.7)),       # short unsmoothed x axis
    ],
)
def test_bit_exact_against_the_oracle(hip, shape, spacing, sigma):
    nz, ny, nx = shape
    v = O.synth(1234 + nx, nx * ny * nz).reshape(shape)
    want = O.smoothing_recursive_gaussian(v, (nx, ny, nz), spacing, sigma, threads=8)
    img = hip.DeviceImage.upload(v, spacing=spacing)
    got = hip.recursive_gaussian(img, sigma).to_host().reshape(-1)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_zero_sigma_is_identity_and_geometry_is_kept(hip):
    v = O.synth(5, 10 * 11 * 12).reshape(12, 11, 10)
    img = hip.DeviceImage.upload(v, spacing=(0.5, 0.7, 1.9), origin=(1, 2, 3))
    out = hip.recursive_gaussian(img, [0, 0, 0])
    assert np.array_equal(out.to_host(), v)
    g = out.geometry()
    assert list(g.size) == [10, 11, 12] and list(g.spacing) == [0.5, 0.7, 1.9]
    assert list(g.origin) == [1.0, 2.0, 3.0]


def test_refusals(hip):
    img = hip.DeviceImage.upload(np.ones((8, 8, 3), np.float32))
    with pytest.raises(hip.CudaError) as e:
        hip.recursive_gaussian(img, [1.0, 1.0, 1.0])
    assert e.value.status == "UnsupportedGeometry"
    hip.recursive_gaussian(img, [0.0, 1.0, 1.0])  # the short axis is not smoothed
    with pytest.raises(hip.CudaError):
        hip.recursive_gaussian(img, [-1.0, 1.0, 1.0])


def test_determinism_across_calls(hip):
    v = O.synth(9, 64 * 48 * 40).reshape(40, 48, 64)
    img = hip.DeviceImage.upload(v)
    outs = {O.fnv1a64(hip.recursive_gaussian(img, [1.2, 2.0, 0.8]).to_host()) for _ in range(5)}
    assert len(outs) == 1


@pytest.mark.parametrize("dtype", [np.uint8, np.int8, np.uint16, np.int16, np.uint32, np.int32,
                                   np.uint64, np.int64, np.float32, np.float64])
def test_upload_cast_is_float_of_double(hip, dtype):
    # crates/sitk/src/cuda/image.rs:93-106: (float)(double)x on the device
    rng = np.random.default_rng(3)
    if np.issubdtype(dtype, np.integer):
        info = np.iinfo(dtype)
        v = rng.integers(info.min, info.max, size=(5, 6, 7), dtype=dtype, endpoint=True)
        v.flat[0] = info.max
        v.flat[1] = info.min
    else:
        v = (rng.normal(size=(5, 6, 7)) * 1e3).astype(dtype)
    got = hip.DeviceImage.upload(v).to_host()
    want = v.astype(np.float64).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_filled_and_masks(hip):
    f = hip.DeviceImage.filled((4, 5, 6), 2.5)
    assert np.all(f.to_host() == 2.5) and len(f) == 120
    a = np.zeros((6, 5, 4), np.uint8)
    a[1:4] = 7
    b = np.zeros((6, 5, 4), np.uint8)
    b[:, 2:] = 1
    ma, mb = hip.DeviceMask.upload(a), hip.DeviceMask.upload(b)
    assert np.array_equal(ma.to_host(a.shape), (a != 0).astype(np.uint8))
    assert np.array_equal(ma.intersect(mb).to_host(a.shape), ((a != 0) & (b != 0)).astype(np.uint8))
    img = hip.DeviceImage.upload(np.where(a != 0, 3.0, 0.0).astype(np.float32))
    assert np.array_equal(hip.DeviceMask.from_device_image(img).to_host(a.shape),
                          (a != 0).astype(np.uint8))
    mc = hip.DeviceMask.upload(np.ones((6, 5, 5), np.uint8))
    with pytest.raises(hip.CudaError):
        ma.intersect(mc)


def test_batch_equals_one_image_at_a_time(hip):
    """The multi-image launches give every image the bits of its own call,
    across uneven shapes and spacings, per-image sigmas, and images the batch
    hands to the one-image path (an axis shorter than 8 voxels)."""
    rng = np.random.default_rng(17)
    shapes = [(9, 12, 40), (33, 17, 8), (64, 20, 12), (7, 30, 30), (16, 16, 16), (128, 9, 11)]
    imgs, sig = [], []
    for i, shp in enumerate(shapes):
        v = rng.standard_normal(shp).astype(np.float32)
        sp = tuple(float(x) for x in rng.uniform(0.6, 1.8, 3))
        imgs.append(hip.DeviceImage.upload(v, spacing=sp))
        sig.append([1.0 + 0.5 * i, 2.0, 0.8 + 0.1 * i])
    got = hip.recursive_gaussian_batch(imgs, sig)
    for img, s, g in zip(imgs, sig, got):
        want = hip.recursive_gaussian(img, s).to_host()
        assert np.array_equal(g.to_host(), want)
