"""ctypes view of the CPU FP64 oracle (oracle/nr_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the checker or the timed CPU baseline,
never as the thing measured or shipped.  The product package numrs_amd never
imports this module.

Parity status of each function is documented in nr_oracle.c and pinned by
tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liborc.so")
_SO_NATIVE = os.path.join(_HERE, "liborc_native.so")
_lib = None
_lib_native = None

_dp = ctypes.POINTER(ctypes.c_double)
_szp = ctypes.POINTER(ctypes.c_size_t)


def build() -> None:
    """Compile liborc.so with oracle/Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def native_lib():
    """The oracle built with -march=native on this host (make -C oracle
    native), for the CPU-baseline timings; None if it cannot be built here."""
    global _lib_native
    if _lib_native is None:
        try:
            subprocess.run(["make", "-s", "-C", _HERE, "native"], check=True, timeout=120,
                           capture_output=True)
            _lib_native = _load(_SO_NATIVE)
        except (OSError, subprocess.SubprocessError):
            return None
    return _lib_native


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            raise RuntimeError(f"oracle library missing: {_SO} (run `make -C oracle`)")
        _lib = _load(_SO)
    return _lib


def _load(path: str) -> ctypes.CDLL:
    L = ctypes.CDLL(path)
    sz, i32, u64 = ctypes.c_size_t, ctypes.c_int, ctypes.c_uint64
    sig = {
        "orc_gen_uniform": (None, [u64, sz, _dp]),
        "orc_gen_normal": (None, [u64, sz, _dp]),
        "orc_gen_skip": (u64, [u64, sz]),
        "orc_four1_literal": (None, [_dp, sz, i32]),
        "orc_four1_nr": (None, [_dp, sz, i32]),
        "orc_dft_brute": (None, [_dp, sz, i32, _dp]),
        "orc_realft": (None, [_dp, sz, i32]),
        "orc_realft_literal": (None, [_dp, sz, i32]),
        "orc_fourn": (i32, [_dp, _szp, sz, i32]),
        "orc_prepare_response": (None, [_dp, _dp, sz, sz]),
        "orc_convlv": (i32, [_dp, sz, _dp, sz, i32, _dp]),
        "orc_convlv_direct": (i32, [_dp, sz, _dp, sz, _dp]),
        "orc_correl": (i32, [_dp, _dp, sz, _dp]),
        "orc_correl_direct_circular": (None, [_dp, _dp, sz, _dp]),
        "orc_four1_batch": (i32, [_dp, sz, sz, i32, i32]),
        "orc_twofft": (None, [_dp, _dp, sz, _dp, _dp]),
        "orc_cosft1": (None, [_dp, sz]),
        "orc_cosft2": (None, [_dp, sz, i32]),
        "orc_rlft3": (i32, [_dp, _dp, sz, sz, sz, i32]),
        "orc_convlv_batch": (i32, [_dp, sz, _dp, sz, sz, i32, _dp, i32]),
        "orc_uneven_batch": (i32, [_dp, ctypes.POINTER(i32), _szp, _szp, sz, i32, i32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dp)


class OracleError(RuntimeError):
    def __init__(self, code: int):
        super().__init__(f"oracle status {code}")
        self.code = code


def _chk(rc: int) -> None:
    if rc != 0:
        raise OracleError(rc)


# ---- inputs ----------------------------------------------------------------
def gen_uniform(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    lib().orc_gen_uniform(seed, n, _p(out))
    return out


def gen_normal(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    lib().orc_gen_normal(seed, n, _p(out))
    return out


def gen_skip(seed: int, draws: int) -> int:
    return int(lib().orc_gen_skip(seed, draws))


# ---- transforms (all return new arrays; inputs are not modified) ----------
def four1_literal(data, isign: int) -> np.ndarray:
    d = _f64(data).copy()
    lib().orc_four1_literal(_p(d), d.size // 2, isign)
    return d


def four1_nr(data, isign: int) -> np.ndarray:
    d = _f64(data).copy()
    lib().orc_four1_nr(_p(d), d.size // 2, isign)
    return d


def dft_brute(data, isign: int) -> np.ndarray:
    d = _f64(data)
    out = np.empty_like(d)
    lib().orc_dft_brute(_p(d), d.size // 2, isign, _p(out))
    return out


def realft(data, isign: int) -> np.ndarray:
    d = _f64(data).copy()
    lib().orc_realft(_p(d), d.size, isign)
    return d


def realft_literal(buf, n: int, isign: int) -> np.ndarray:
    """The reference's defective realft on a buffer of at least n+1 values (the
    reference callers pass 2n, Convolve.rs:81); returns the whole buffer."""
    d = _f64(buf).copy()
    assert d.size > n
    lib().orc_realft_literal(_p(d), n, isign)
    return d


def fourn(data, nn, isign: int) -> np.ndarray:
    d = _f64(data).copy()
    nn_a = (ctypes.c_size_t * len(nn))(*nn)
    _chk(lib().orc_fourn(_p(d), nn_a, len(nn), isign))
    return d


def fourn_status(nn, ndim: int, isign: int) -> int:
    """Validation only (no data touched when the status is an error)."""
    total = int(np.prod(nn[:ndim])) if ndim and ndim <= len(nn) and all(v > 0 for v in nn[:ndim]) else 1
    d = np.zeros(2 * max(total, 1), np.float64)
    nn_a = (ctypes.c_size_t * max(len(nn), 1))(*nn)
    return int(lib().orc_fourn(_p(d), nn_a, ndim, isign))


def prepare_response(respns, n: int) -> np.ndarray:
    r = _f64(respns)
    out = np.empty(n, np.float64)
    lib().orc_prepare_response(_p(out), _p(r), n, r.size)
    return out


def convlv(data, respns, isign: int) -> np.ndarray:
    d, r = _f64(data), _f64(respns)
    out = np.empty(d.size, np.float64)
    _chk(lib().orc_convlv(_p(d), d.size, _p(r), r.size, isign, _p(out)))
    return out


def convlv_status(n: int, m: int, isign: int) -> int:
    d = np.zeros(max(n, 1))
    r = np.zeros(max(m, 1))
    out = np.zeros(max(n, 1))
    return int(lib().orc_convlv(_p(d), n, _p(r), m, isign, _p(out)))


def convlv_direct(data, respns) -> np.ndarray:
    d, r = _f64(data), _f64(respns)
    out = np.empty(d.size, np.float64)
    _chk(lib().orc_convlv_direct(_p(d), d.size, _p(r), r.size, _p(out)))
    return out


def correl(d1, d2) -> np.ndarray:
    a, b = _f64(d1), _f64(d2)
    out = np.empty(a.size, np.float64)
    _chk(lib().orc_correl(_p(a), _p(b), a.size, _p(out)))
    return out


def correl_direct_circular(d1, d2) -> np.ndarray:
    a, b = _f64(d1), _f64(d2)
    out = np.empty(a.size, np.float64)
    lib().orc_correl_direct_circular(_p(a), _p(b), a.size, _p(out))
    return out


# ---- CPU baseline drivers ---------------------------------------------------
# `L` selects the library (default: the portable build; native_lib() for
# timings).  Each returns the OpenMP threads used.
def four1_batch_inplace(data: np.ndarray, nn: int, batch: int, isign: int, nthreads: int = 0, L=None) -> int:
    assert data.dtype == np.float64 and data.flags.c_contiguous and data.size == 2 * nn * batch
    return int((L or lib()).orc_four1_batch(_p(data), nn, batch, isign, nthreads))


def convlv_batch(data: np.ndarray, n: int, respns: np.ndarray, m: int, batch: int, isign: int,
                 out: np.ndarray, nthreads: int = 0, L=None) -> int:
    return int((L or lib()).orc_convlv_batch(_p(data), n, _p(respns), m, batch, isign, _p(out), nthreads))


def uneven_batch_inplace(data: np.ndarray, items, offsets, isign: int, nthreads: int = 0, L=None) -> int:
    """Items (N,) through the literal four1, (n0, n1) through fourn, in place
    at complex offsets `offsets` of data; parallel over items."""
    assert data.dtype == np.float64 and data.flags.c_contiguous
    n = len(items)
    nd = (ctypes.c_int * n)(*[len(it) for it in items])
    dims = (ctypes.c_size_t * (2 * n))(*[v for it in items for v in (it[0], it[1] if len(it) > 1 else 1)])
    offs = (ctypes.c_size_t * n)(*[int(o) for o in offsets])
    return int((L or lib()).orc_uneven_batch(_p(data), nd, dims, offs, n, isign, nthreads))


def four1_literal_with(L, data: np.ndarray, isign: int) -> None:
    """In-place literal four1 through library L (timing driver)."""
    (L or lib()).orc_four1_literal(_p(data), data.size // 2, isign)


def fourn_with(L, data: np.ndarray, nn, isign: int) -> None:
    """In-place NR fourn through library L (timing driver)."""
    arr = (ctypes.c_size_t * len(nn))(*[int(v) for v in nn])
    _chk((L or lib()).orc_fourn(_p(data), arr, len(nn), isign))


def twofft(d1, d2):
    """FFT_2.rs:3-132 (NR twofft): the isign=+1 spectra (bins 0..n-1,
    interleaved) of two real sequences from one complex FFT."""
    a, b = _f64(d1), _f64(d2)
    n = a.size
    f1, f2 = np.zeros(2 * n), np.zeros(2 * n)
    lib().orc_twofft(_p(a), _p(b), n, _p(f1), _p(f2))
    return f1, f2


def cosft1(y):
    """Cos_FT.rs:7-68 (NR cosft1, DCT-I): y has n + 1 values."""
    d = _f64(y).copy()
    lib().orc_cosft1(_p(d), d.size - 1)
    return d


def cosft2(y, isign: int):
    """Cos_FT2.rs:7-173 (NR cosft2): DCT-II (isign=1) / DCT-III (isign=-1)."""
    d = _f64(y).copy()
    lib().orc_cosft2(_p(d), d.size, isign)
    return d


def rlft3(data, speq, nn1: int, nn2: int, nn3: int, isign: int):
    """Real_FT3.rs:8-141 (NR rlft3): returns (data, speq) after the call."""
    d = _f64(data).copy()
    s = _f64(speq).copy()
    _chk(lib().orc_rlft3(_p(d), _p(s), nn1, nn2, nn3, isign))
    return d, s
