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
