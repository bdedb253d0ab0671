"""C ABI of libapsm_amd.so: exports, defaults and host-side argument validation (no GPU)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "apsm_capi.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*([a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    from libapsm_amd._lib import SIGNATURES, lib
    L = lib()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), f"{n} declared in apsm_capi.h but not exported"
    assert set(names) == set(SIGNATURES), "ctypes signature table out of sync with the header"


def test_params_default_matches_reference_factory():
    from libapsm_amd._lib import ApsmParamsC, lib
    p = ApsmParamsC()
    lib().apsm_params_default(ctypes.byref(p))
    # cpp/lib/apsm/apsm_noma_detector.cu:361-372
    assert p.gaussian_kernel_weight == pytest.approx(0.5)
    assert p.gaussian_kernel_variance == pytest.approx(0.05)
    assert p.eB == pytest.approx(0.001)
    assert (p.window_size, p.sample_step, p.train_passes, p.dictionary_size) == (20, 2, 1, 685)
    assert p.detect_version == 0x10007
    assert p.linear_kernel_weight == pytest.approx(0.5)
    assert p.detect_precision == 0  # APSM_PRECISION_FP32


@pytest.mark.parametrize("field,value,code", [
    ("sample_step", 3, -1),
    ("sample_step", 0, -1),
    ("window_size", 700, -1),
    ("detect_version", 12345, -1),
    ("detect_precision", 5, -1),
    ("detect_precision", -1, -1),
    ("norm_constraint", -1.0, -1),
    ("gaussian_kernel_weight", 1.5, -1),
])
def test_create_rejects_bad_parameters(field, value, code):
    from libapsm_amd._lib import ApsmParamsC, lib
    L = lib()
    p = ApsmParamsC()
    L.apsm_params_default(ctypes.byref(p))
    setattr(p, field, value)
    h = ctypes.c_void_p()
    rc = L.apsm_create(ctypes.byref(p), 8, 1, ctypes.byref(h))
    assert rc == code
    assert not h.value
    assert L.apsm_last_error().decode() != ""


def test_create_rejects_bad_sizes():
    from libapsm_amd._lib import ApsmParamsC, lib
    L = lib()
    p = ApsmParamsC()
    L.apsm_params_default(ctypes.byref(p))
    h = ctypes.c_void_p()
    assert L.apsm_create(ctypes.byref(p), 0, 1, ctypes.byref(h)) == -1
    assert L.apsm_create(ctypes.byref(p), 65, 1, ctypes.byref(h)) == -1
    assert L.apsm_create(ctypes.byref(p), 8, 0, ctypes.byref(h)) == -1
    assert L.lls_create(0, ctypes.byref(h)) == -1


def test_null_handles_are_errors_not_crashes():
    from libapsm_amd._lib import lib
    L = lib()
    assert L.apsm_train(None, None, None, 10, None) == -1
    assert L.apsm_detect(None, None, 10, None, None) == -1
    assert L.apsm_reset(None) == -1
    assert L.lls_train(None, None, None, 10, None) == -1
    L.apsm_destroy(None)
    L.lls_destroy(None)
    assert L.apsm_device_count() >= 0


def test_factory_validation_matches_reference():
    from libapsm_amd import ApsmNomaDetectorFactory, NomaDetectorFactory
    with pytest.raises(ValueError, match="Invalid key: windowsize"):
        NomaDetectorFactory.build({"otype": "APSM", "windowsize": 3}, 8)
    with pytest.raises(ValueError, match="Unknown detect version"):
        NomaDetectorFactory.build({"otype": "APSM", "detectVersion": "fast"}, 8)
    with pytest.raises(ValueError, match="Unknown detect precision"):
        NomaDetectorFactory.build({"otype": "APSM", "detectPrecision": "fp8"}, 8)
    with pytest.raises(ValueError, match="Unknown algorithm"):
        NomaDetectorFactory.build({"otype": "KNN"}, 8)
    with pytest.raises(ValueError, match="Invalid key"):
        NomaDetectorFactory.build({"otype": "LLS", "windowSize": 3}, 8)
    par = ApsmNomaDetectorFactory.params_from_config({"otype": "APSM", "gaussianKernelWeight": 0.25})
    assert par.linearKernelWeight == pytest.approx(0.75)
    assert par.dictionarySize == 685 and par.windowSize == 20 and par.sampleStep == 2
    assert par.detectPrecision == 0
    assert ApsmNomaDetectorFactory.params_from_config({"otype": "APSM", "detectPrecision": "fp16"}).detectPrecision == 4


def test_noalloc_least_squares_checks_sizes_before_any_device_call():
    """The allocation-free least-squares entry point rejects a problem beyond its fixed scratch
    (and bad arguments) on the host, before touching the device (runs without a GPU)."""
    import ctypes

    from libapsm_amd._lib import APSM_ERR_ARG, lib
    L = lib()
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: the call fails before any launch
    assert L.lls_train_lowlevel_noalloc(fake, fake, 128, 1, 1000000, fake, fake, None) == APSM_ERR_ARG
    assert b"scratch" in L.apsm_last_error()
    assert L.lls_train_lowlevel_noalloc(fake, fake, 129, 1, 100, fake, fake, None) == APSM_ERR_ARG
    assert L.lls_train_lowlevel_noalloc(fake, fake, 32, 2, 100, fake, fake, None) == APSM_ERR_ARG
    assert L.lls_train_lowlevel_noalloc(None, fake, 32, 1, 100, fake, fake, None) == APSM_ERR_ARG
