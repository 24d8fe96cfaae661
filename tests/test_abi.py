"""CPU: the C-ABI library loads, exports every symbol include/nrfft.h
declares, and validates arguments (error paths return before any HIP call)."""
import ctypes
import subprocess

import numpy as np
import pytest

import numrs_amd
from numrs_amd import _lib


def test_library_exports_header_symbols():
    L = _lib.lib()
    declared = _lib.header_symbols()
    assert len(declared) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    for s in declared:
        assert hasattr(L, s)


def test_status_strings():
    assert _lib.status_string(0) == "ok"
    assert _lib.status_string(_lib.E_DIMS) == "Invalid dimensions"
    assert _lib.status_string(_lib.E_RESP_TOO_LONG) == "Response function longer than data"


def _d(n):
    return np.zeros(max(n, 1)), np.zeros(max(n, 1), np.float32)


def test_four1_validation():
    L = _lib.lib()
    d, f = _d(8)
    assert L.nrfft_four1_f64(d.ctypes.data_as(_lib._dp), 3, 1) == _lib.E_NOT_POW2
    assert L.nrfft_four1_f64(d.ctypes.data_as(_lib._dp), 4, 0) == _lib.E_ISIGN
    assert L.nrfft_four1_f64(d.ctypes.data_as(_lib._dp), 0, 1) == _lib.E_EMPTY
    assert L.nrfft_four1_f32(f.ctypes.data_as(_lib._fp), 6, 1) == _lib.E_NOT_POW2
    assert L.nrfft_four1_f64(None, 4, 1) == _lib.E_INVALID


def test_fourn_validation_matches_reference_order():
    # Fourn.rs:367-378: ndim, then isign, then sizes.
    L = _lib.lib()
    d, _ = _d(64)

    def call(nn, ndim, isign):
        arr = (ctypes.c_size_t * max(len(nn), 1))(*nn)
        return L.nrfft_fourn_f64(d.ctypes.data_as(_lib._dp), arr, ndim, isign)

    assert call([1], 1, 1) == _lib.E_DIMS          # Fourn.rs:472
    assert call([8], 1, 0) == _lib.E_ISIGN         # Fourn.rs:475
    assert call([8, 8], 0, 1) == _lib.E_DIMS
    assert call([8, 1], 2, 0) == _lib.E_ISIGN
    assert call([8, 6], 2, 1) == _lib.E_NOT_POW2
    with pytest.raises(numrs_amd.FournError):
        numrs_amd.fourn(np.zeros(16), [1], 1, 1)
    with pytest.raises(numrs_amd.FournError):
        numrs_amd.fourn(np.zeros(16), [8], 1, 0)


def test_convlv_validation():
    # Convolve.rs:405-424
    with pytest.raises(numrs_amd.ConvlvError) as e:
        numrs_amd.convlv([], [1.0], 1)
    assert e.value.kind == "EmptyInput"
    with pytest.raises(numrs_amd.ConvlvError) as e:
        numrs_amd.convlv([1.0, 2.0], [1.0, 2.0, 3.0], 1)
    assert e.value.kind == "ResponseTooLong"
    with pytest.raises(numrs_amd.ConvlvError) as e:
        numrs_amd.convlv([1.0, 2.0], [1.0], 0)
    assert e.value.kind == "InvalidIsign"
    with pytest.raises(numrs_amd.ConvlvError) as e:
        numrs_amd.convlv([1.0, 2.0, 3.0], [1.0], 1)
    assert e.value.code == _lib.E_NOT_POW2


def test_correl_validation():
    # Correlation.rs:453-462
    with pytest.raises(numrs_amd.CorrelError) as e:
        numrs_amd.correl([], [1.0])
    assert e.value.kind == "EmptyInput"
    with pytest.raises(numrs_amd.CorrelError) as e:
        numrs_amd.correl([1.0, 2.0], [1.0])
    assert e.value.kind == "LengthMismatch"
    with pytest.raises(numrs_amd.CorrelError) as e:
        numrs_amd.correl_normalized([1.0, 1.0], [1.0, 2.0])
    assert e.value.kind == "ZeroStdDev"


def test_realft_validation():
    L = _lib.lib()
    d, _ = _d(16)
    assert L.nrfft_realft_f64(d.ctypes.data_as(_lib._dp), 12, 1) == _lib.E_NOT_POW2
    assert L.nrfft_realft_f64(d.ctypes.data_as(_lib._dp), 16, 2) == _lib.E_ISIGN
    with pytest.raises(AssertionError):
        numrs_amd.realft(np.zeros(8), 7, 1)      # Real_FT.rs:5
    with pytest.raises(AssertionError):
        numrs_amd.realft(np.zeros(8), 16, 1)     # Real_FT.rs:6


def test_plan_validation():
    L = _lib.lib()
    h = ctypes.c_void_p()
    nn = (ctypes.c_size_t * 1)(12)
    assert L.nrfft_plan_c2c(ctypes.byref(h), 0, nn, 1, 1) == _lib.E_NOT_POW2
    assert L.nrfft_plan_convlv(ctypes.byref(h), 0, 16, 32, 1) == _lib.E_RESP_TOO_LONG
    assert L.nrfft_execute_c2c(None, None, None, 1, None) == _lib.E_INVALID
    assert L.nrfft_plan_destroy(None) == 0


def test_host_helpers():
    c = numrs_amd.real_to_complex([1.0, 2.0])
    assert np.array_equal(c, [1.0, 0.0, 2.0, 0.0])
    assert np.array_equal(numrs_amd.complex_to_real(c), [1.0, 2.0])
    z = np.array([3.0, 4.0, 0.0, 1.0])
    assert np.allclose(numrs_amd.magnitude_spectrum(z), [5.0, 1.0])
    assert np.allclose(numrs_amd.power_spectrum(z), [25.0, 1.0])


def test_spectral_validation():
    # twofft / cosft1 / cosft2 / rlft3: argument checks that run before any
    # device work (FFT_2.rs:5-7, Cos_FT2.rs:8-12, Real_FT3.rs:16-19).
    L = _lib.lib()
    d, _ = _d(8)
    p = d.ctypes.data_as(_lib._dp)
    assert L.nrfft_twofft_f64(p, p, 3, p, p) == _lib.E_NOT_POW2
    assert L.nrfft_twofft_f64(p, p, 0, p, p) == _lib.E_EMPTY
    assert L.nrfft_twofft_f64(None, p, 4, p, p) == _lib.E_INVALID
    assert L.nrfft_cosft1_f64(p, 6) == _lib.E_NOT_POW2
    assert L.nrfft_cosft1_f64(p, 1) == _lib.E_NOT_POW2
    assert L.nrfft_cosft2_f64(p, 4, 0) == _lib.E_ISIGN
    assert L.nrfft_cosft2_f64(p, 6, 1) == _lib.E_NOT_POW2
    assert L.nrfft_rlft3_f64(p, p, 2, 2, 2, 2) == _lib.E_ISIGN
    assert L.nrfft_rlft3_f64(p, p, 2, 3, 2, 1) == _lib.E_NOT_POW2
    assert L.nrfft_rlft3_f64(p, p, 2, 2, 1, 1) == _lib.E_NOT_POW2
    h = ctypes.c_void_p()
    assert L.nrfft_plan_cosft2(ctypes.byref(h), 1, 12, 1) == _lib.E_NOT_POW2
    assert L.nrfft_execute_rlft3(None, None, None, 1, None) == _lib.E_INVALID
    with pytest.raises(AssertionError):
        numrs_amd.twofft(np.zeros(8), np.zeros(4), np.zeros(18), np.zeros(18))   # FFT_2.rs:5
    with pytest.raises(AssertionError):
        numrs_amd.twofft(np.zeros(8), np.zeros(8), np.zeros(16), np.zeros(18))   # FFT_2.rs:6
    with pytest.raises(ValueError, match="Invalid isign value"):
        numrs_amd.cosft2(np.zeros(8), 8, 2)                                        # Cos_FT2.rs:11
    with pytest.raises(AssertionError):
        numrs_amd.rlft3(np.zeros((2, 2, 4)), np.zeros((2, 4)), 2, 2, 4, 0)         # Real_FT3.rs:16
    with pytest.raises(AssertionError):
        numrs_amd.rlft3(np.zeros((2, 2, 4)), np.zeros((2, 2)), 2, 2, 4, 1)         # Real_FT3.rs:18
    assert numrs_amd.extract_real_imag(numrs_amd.combine_real_imag([1.0, 2.0, 3.0], [4.0, 5.0, 6.0]))[1].tolist() \
        == [4.0, 5.0, 6.0]                                                         # FFT_2.rs:470-479


def test_fourfs_validation(tmp_path):
    # Four_FS.rs:328-339 order (checked before any device work) and the
    # reference's test_invalid_inputs (Four_FS.rs:427-437).
    L = _lib.lib()
    d, _ = _d(64)
    p = d.ctypes.data_as(_lib._dp)
    nn = (ctypes.c_size_t * 3)(8, 8, 1)
    assert L.nrfft_fourfs_f64(p, None, nn, 0, 1, 0) == _lib.E_DIMS
    assert L.nrfft_fourfs_f64(p, None, nn, 2, 0, 0) == _lib.E_ISIGN
    assert L.nrfft_fourfs_f64(p, None, nn, 3, 1, 0) == _lib.E_DIMS
    bad = (ctypes.c_size_t * 1)(12)
    assert L.nrfft_fourfs_f64(p, None, bad, 1, 1, 0) == _lib.E_NOT_POW2
    assert L.nrfft_fourfs_f64(None, None, nn, 2, 1, 0) == _lib.E_INVALID
    fs = numrs_amd.FourFS.new(tmp_path)
    assert len(fs.paths) == 4 and fs.with_memory_mapping().memory_mapped
    with pytest.raises(numrs_amd.FournError, match="Invalid dimension size"):
        fs.fourfs([1], 1, 1)
    with pytest.raises(numrs_amd.FournError, match="isign must be 1 or -1"):
        fs.fourfs([8], 1, 0)
    with pytest.raises(numrs_amd.FournError, match="Invalid dimensions"):
        fs.fourfs([8], 2, 1)


def test_missing_library_fails_loudly():
    # No CPU fallback: with the HIP library absent, a call raises instead of
    # computing anything.
    import os
    import sys

    code = ("import numrs_amd._lib as L; L.LIB_PATH = '/nonexistent/libnrfft.so'; "
            "import numpy as np, numrs_amd; numrs_amd.four1(np.zeros(16), 8, 1)")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300,
                         cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    assert out.returncode != 0
    assert "is missing" in out.stderr
