"""The C ABI without a GPU: the library loads, exports every function
include/sitk_hip.h declares, and its host-only entry points (status names,
MattesGeometry refusals, the Jacobian contractions) behave as the reference's
(crates/sitk/src/registration/mattes.rs:152-186, registration/cuda.rs:381-524)."""

import ctypes as C
import os
import re

import numpy as np
import pytest

import sitk_hip as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "sitk_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sitkhip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = S.lib()
    names = declared_functions()
    assert len(names) >= 38
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_status_names_cover_every_code():
    L = S.lib()
    names = [L.sitkhip_status_name(i).decode() for i in range(16)]
    assert names[0] == "Ok" and "Unknown" not in names
    assert L.sitkhip_status_name(99).decode() == "Unknown"


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_no_device_is_reported_by_name():
    with pytest.raises(S.CudaError) as e:
        S.backend_init(0)
    assert e.value.status == "NoDevice"


def test_mattes_bins_from_ranges_matches_mattes_geometry():
    import oracle as O

    fr, mr = (3.25, 211.5), (-7.0, 140.125)
    b = S.mattes_bins(fr, mr, 50)
    g = O.MattesGeom()
    O.lib().orc_mattes_geometry(np.array(fr).ctypes.data_as(C.POINTER(C.c_double)),
                                np.array(mr).ctypes.data_as(C.POINTER(C.c_double)), 50,
                                C.byref(g))
    assert b.bins == 50 and b.padding == 2
    assert b.fixed_bin_size == g.fixed_bin_size
    assert b.moving_bin_size == g.moving_bin_size
    assert b.fixed_normalized_min == g.fixed_normalized_min
    assert b.moving_normalized_min == g.moving_normalized_min
    assert (b.moving_true_min, b.moving_true_max) == mr


def test_mattes_bins_refusals_are_named():
    with pytest.raises(S.CudaError) as e:
        S.mattes_bins((0, 1), (0, 1), 4)
    assert e.value.status == "TooFewHistogramBins"
    S.mattes_bins((0, 1), (0, 1), 5)
    with pytest.raises(S.CudaError) as e:
        S.mattes_bins((2, 2), (0, 1), 50)
    assert e.value.status == "ConstantIntensity" and "fixed" in e.value.message
    with pytest.raises(S.CudaError) as e:
        S.mattes_bins((0, 1), (3, 3), 50)
    assert e.value.status == "ConstantIntensity" and "moving" in e.value.message


def _contract_reference(mom, j0, c, np_):
    out = np.zeros(np_)
    for k in range(np_):
        g = 0.0
        for d in range(3):
            g += j0[d, k] * mom[d]
            for e in range(3):
                g += c[e, d, k] * mom[3 + 3 * d + e]
        out[k] = g
    return out


def test_contractions_match_the_reference_formula():
    rng = np.random.default_rng(1)
    for nparams in (3, 6, 12):
        j0 = rng.normal(size=(3, nparams))
        c = rng.normal(size=(3, 3, nparams))
        mom12 = rng.normal(size=12)
        d = S.contract_mattes(mom12, j0, c, nparams)
        assert np.array_equal(d, _contract_reference(mom12, j0, c, nparams))
        mom14 = np.concatenate([[5.0], rng.normal(size=12), [7.0]])
        v, dm, valid = S.contract_ms(mom14, j0, c, nparams)
        assert valid == 7 and v == 5.0 * (1.0 / 7.0)
        want = np.array([2.0 * g * (1.0 / 7.0) for g in
                         _contract_reference(mom14[1:13], j0, c, nparams)])
        assert np.array_equal(dm, want)
    # no valid sample: f64::MAX and a zero derivative (registration/cuda.rs:385-392)
    v, dm, valid = S.contract_ms(np.zeros(14), np.zeros((3, 6)), np.zeros((3, 3, 6)), 6)
    assert v == np.finfo(np.float64).max and valid == 0 and not dm.any()


def test_draw_samples_is_the_references():
    import oracle as O

    assert S.draw_samples(100, 0, 1.0, 0) is None
    for strategy, pct, seed in [(1, 0.3, 0), (1, 0.05, 0), (2, 0.2, 5), (2, 0.25, 12345)]:
        got = S.draw_samples(10007, strategy, pct, seed)
        want = O.draw_samples(10007, strategy, pct, seed)
        assert np.array_equal(got, want)


def test_draw_samples_refuses_fractions_it_cannot_count():
    """ceil(1/p) and voxels*p are cast to sizes: NaN, infinities, zero and
    negative fractions are refused by name instead of being undefined casts.
    A fraction above 1 stays allowed, as in the reference (stride 1 / more draws
    than voxels)."""
    for strategy in (1, 2):
        for bad in (0.0, -0.2, float("nan"), float("inf"), 1e-320):
            with pytest.raises(S.CudaError) as e:
                S.draw_samples(1000, strategy, bad, 1)
            assert e.value.status == "InvalidArgument"
    assert S.draw_samples(10, 1, 1.5, 0).tolist() == list(range(10))
    assert S.draw_samples(10, 2, 1.5, 3).size == 15


def test_host_constructor_refusals_before_any_device_work():
    """sitkhip_metric_create_host validates on the host, in the order of
    Resident::build (crates/sitk/src/cuda/ops/resident.rs:530-660), before it
    touches a device: an empty sample set and a moving length that does not
    match its size are DegenerateInput, a non-dense moving layout is refused,
    and a fixed mask with explicit points is MaskedExplicitPoints
    (src/cuda/error.rs)."""
    m = np.ones((4, 4, 4))
    p2i = np.eye(3)
    pts = np.zeros((5, 3))
    with pytest.raises(S.CudaError) as e:
        S.ResidentMetric.from_host(np.zeros(0), m, (4, 4, 4), (0, 0, 0), p2i, points=np.zeros((0, 3)))
    assert e.value.status == "DegenerateInput"
    with pytest.raises(S.CudaError) as e:
        S.ResidentMetric.from_host(np.zeros(5), m, (4, 4, 4), (0, 0, 0), p2i, points=pts,
                                   fixed_mask=np.ones(5, np.uint8))
    assert e.value.status == "MaskedExplicitPoints"
    spec = S.FixedPointsSpec()
    spec.mode = S.POINTS_EXPLICIT
    p = np.zeros(15)
    spec.points = p.ctypes.data_as(C.POINTER(C.c_double))
    mg = S.MovingGeometrySpec()
    mg.len = 63
    mg.size[:] = [4, 4, 4]
    mg.strides[:] = [1, 4, 16]
    fv, mv = np.zeros(5), np.zeros(64)
    h = C.c_void_p()
    dp = C.POINTER(C.c_double)
    rc = S.lib().sitkhip_metric_create_host(0, 5, fv.ctypes.data_as(dp), C.byref(spec), None,
                                            C.byref(mg), mv.ctypes.data_as(dp), 0, C.byref(h))
    assert S.lib().sitkhip_status_name(rc).decode() == "DegenerateInput"
    mg.len = 64
    mg.strides[:] = [1, 4, 17]
    rc = S.lib().sitkhip_metric_create_host(0, 5, fv.ctypes.data_as(dp), C.byref(spec), None,
                                            C.byref(mg), mv.ctypes.data_as(dp), 0, C.byref(h))
    assert S.lib().sitkhip_status_name(rc).decode() == "UnsupportedGeometry"
