"""Host-side mirror of the numrs FFT-path API on top of the C ABI.

Names, argument meaning and error behaviour follow the Rust routines each
function replaces (cited per function).  Every transform runs in the HIP
kernels of libnrfft.so; nothing here computes an FFT on the CPU.  Rust panics
(slice out of bounds, ``assert!``) surface as ``ValueError`` /
``AssertionError``; Rust ``Result`` errors as the exception classes below.
"""
from __future__ import annotations

import ctypes
from typing import Iterable, Sequence

import numpy as np

from . import _lib

# ---------------------------------------------------------------------------
# Errors
# ---------------------------------------------------------------------------


class NrfftError(RuntimeError):
    """A non-zero nrfft_status from the C ABI."""

    def __init__(self, code: int, message: str = ""):
        self.code = code
        text = message or _lib.status_string(code)
        detail = _lib.last_error() if code in (_lib.E_HIP, _lib.E_NOMEM, _lib.E_UNSUPPORTED) else ""
        super().__init__(f"{text}" + (f" ({detail})" if detail and detail != text else ""))


class ConvlvError(NrfftError):
    """Convolve.rs:226-238 ``ConvlvError``; ``kind`` names the variant."""

    _KIND = {_lib.E_EMPTY: "EmptyInput", _lib.E_RESP_TOO_LONG: "ResponseTooLong",
             _lib.E_ISIGN: "InvalidIsign"}
    _MSG = {"EmptyInput": "Input arrays cannot be empty",
            "ResponseTooLong": "Response function longer than data",
            "InvalidIsign": "isign must be 1 (convolution) or -1 (deconvolution)",
            "DivisionByZero": "Division by zero in deconvolution"}

    def __init__(self, code: int):
        self.kind = self._KIND.get(code, "FftError")
        msg = self._MSG.get(self.kind) or f"FFT computation error: {_lib.status_string(code)}"
        super().__init__(code, msg)


class CorrelError(NrfftError):
    """Correlation.rs:389-399 ``CorrelError``; ``kind`` names the variant."""

    _MSG = {"EmptyInput": "Input arrays cannot be empty",
            "LengthMismatch": "Input arrays must have the same length",
            "ZeroStdDev": "Normalization error: standard deviation is zero"}

    def __init__(self, kind: str, code: int = 0):
        self.kind = kind
        msg = self._MSG.get(kind) or f"FFT computation error: {_lib.status_string(code)}"
        super().__init__(code, msg)


class FournError(ValueError):
    """``io::Error`` of kind InvalidInput with the messages of Fourn.rs:367-378."""

    def __init__(self, code: int, message: str | None = None):
        self.code = code
        if message is None:
            message = _lib.status_string(code) if code != _lib.E_DIMS else (_lib.last_error() or "Invalid dimensions")
        super().__init__(message)


def _check(code: int) -> None:
    if code != _lib.OK:
        raise NrfftError(code)


# ---------------------------------------------------------------------------
# Array plumbing
# ---------------------------------------------------------------------------

def _buf(a: np.ndarray):
    """(ctypes pointer, suffix) for a contiguous float32/float64 array."""
    if not isinstance(a, np.ndarray):
        raise TypeError("expected a numpy array")
    if a.dtype == np.float64:
        return a.ctypes.data_as(_lib._dp), "f64"
    if a.dtype == np.float32:
        return a.ctypes.data_as(_lib._fp), "f32"
    raise TypeError(f"unsupported dtype {a.dtype} (float32 or float64)")


def _inplace_view(data: np.ndarray, count: int) -> np.ndarray:
    if data.ndim != 1 or not data.flags.c_contiguous or not data.flags.writeable:
        raise ValueError("data must be a writable contiguous 1-D array")
    if data.size < count:
        raise ValueError(f"index out of bounds: the len is {data.size} but the transform needs {count}")
    return data[:count]


def _as_input(x, dtype=None) -> np.ndarray:
    a = np.asarray(x)
    if dtype is None:
        dtype = a.dtype if a.dtype in (np.float32, np.float64) else np.float64
    return np.ascontiguousarray(a, dtype=dtype).reshape(-1)


# ---------------------------------------------------------------------------
# four1 (FFT_1.rs)
# ---------------------------------------------------------------------------

def four1(data: np.ndarray, nn: int, isign: int) -> None:
    """FFT_1.rs:5 ``four1(data, nn, isign)``: in-place complex FFT of 2*nn
    interleaved values, kernel exp(+isign 2 pi i jk/nn), unnormalised."""
    v = _inplace_view(data, 2 * nn)
    p, sfx = _buf(v)
    _check(getattr(_lib.lib(), f"nrfft_four1_{sfx}")(p, nn, isign))


# FFT_1.rs:110 -- same transform, all stages with the recurrence on the CPU.
four1_optimized = four1


class FFTProcessor:
    """FFT_1.rs:142-190.  ``fft`` transforms one buffer (nn = len/2);
    ``fft_batch`` runs a list of buffers, equal-length groups batched into one
    GPU execution each."""

    def __init__(self) -> None:
        self.max_threads = 1
        self.use_optimized = True

    @classmethod
    def new(cls) -> "FFTProcessor":
        return cls()

    def with_threads(self, threads: int) -> "FFTProcessor":
        self.max_threads = threads
        return self

    def with_optimized(self, use_optimized: bool) -> "FFTProcessor":
        self.use_optimized = use_optimized
        return self

    def fft(self, data: np.ndarray, isign: int) -> None:
        four1(data, data.size // 2, isign)

    def fft_batch(self, batches: Sequence[np.ndarray], isign: int) -> None:
        """All buffers of one dtype go to the GPU in one uneven-batch launch
        (nrfft_fft_batch_*), whatever their lengths."""
        for dt, ctype, sfx in ((np.float64, ctypes.c_double, "f64"), (np.float32, ctypes.c_float, "f32")):
            idx = [i for i, b in enumerate(batches) if b.dtype == dt]
            if not idx:
                continue
            for i in idx:
                _inplace_view(batches[i], 2 * (batches[i].size // 2))
            ptrs = (ctypes.POINTER(ctype) * len(idx))(*[batches[i].ctypes.data_as(ctypes.POINTER(ctype)) for i in idx])
            nns = (ctypes.c_size_t * len(idx))(*[batches[i].size // 2 for i in idx])
            _check(getattr(_lib.lib(), f"nrfft_fft_batch_{sfx}")(ptrs, nns, len(idx), isign))
        other = [b for b in batches if b.dtype not in (np.float32, np.float64)]
        if other:
            raise TypeError("fft_batch buffers must be float32 or float64")


def real_to_complex(real_data) -> np.ndarray:
    """FFT_1.rs:193-200."""
    r = np.asarray(real_data, dtype=np.float64)
    out = np.zeros(2 * r.size, np.float64)
    out[0::2] = r
    return out


def complex_to_real(complex_data) -> np.ndarray:
    """FFT_1.rs:202-204."""
    return np.asarray(complex_data)[0::2].copy()


def magnitude_spectrum(complex_data) -> np.ndarray:
    """FFT_1.rs:206-216."""
    c = np.asarray(complex_data, dtype=np.float64)
    m = c.size // 2
    return np.sqrt(c[0:2 * m:2] ** 2 + c[1:2 * m:2] ** 2)


def power_spectrum(complex_data) -> np.ndarray:
    """FFT_1.rs:218-228."""
    c = np.asarray(complex_data, dtype=np.float64)
    m = c.size // 2
    return c[0:2 * m:2] ** 2 + c[1:2 * m:2] ** 2


# ---------------------------------------------------------------------------
# realft (Real_FT.rs)
# ---------------------------------------------------------------------------

def realft(data: np.ndarray, n: int, isign: int) -> None:
    """Real_FT.rs:4 ``realft(data, n, isign)`` with NR semantics: forward
    (isign=1) packs [F_0, F_{n/2}, Re F_1, Im F_1, ...]; inverse returns (n/2) x."""
    assert n % 2 == 0, "n must be even"                      # Real_FT.rs:5
    assert data.size >= n, "data length must be at least n"  # Real_FT.rs:6
    v = _inplace_view(data, n)
    p, sfx = _buf(v)
    _check(getattr(_lib.lib(), f"nrfft_realft_{sfx}")(p, n, 1 if isign == 1 else -1))


realft_optimized = realft  # Real_FT.rs:216
# Real_FT.rs:373 -- the same transform (sequential recurrence on the CPU).
realft_sequential = realft


class RealFTProcessor:
    """Real_FT.rs:331-370.  ``process`` is one realft (the three Rust variants
    it chooses between compute the same transform); ``process_batch`` takes
    (data, n, isign) items like the Rust ``&mut [(&mut [f64], usize, i32)]``
    and runs every group of equal (n, isign, dtype) as one batched GPU
    execution (nrfft_realft_batch_*), where the Rust spreads items over rayon
    threads (Real_FT.rs:365-369)."""

    def __init__(self) -> None:
        self.use_optimized = True
        self.parallel_threshold = 1024

    @classmethod
    def new(cls) -> "RealFTProcessor":
        return cls()

    def with_optimized(self, use_optimized: bool) -> "RealFTProcessor":
        self.use_optimized = use_optimized
        return self

    def with_threshold(self, threshold: int) -> "RealFTProcessor":
        self.parallel_threshold = threshold
        return self

    def process(self, data: np.ndarray, n: int, isign: int) -> None:
        realft(data, n, isign)

    def process_batch(self, batches: Iterable[tuple]) -> None:
        items = list(batches)
        groups: dict = {}
        for data, n, isign in items:
            n = int(n)
            assert n % 2 == 0, "n must be even"                      # Real_FT.rs:5
            assert data.size >= n, "data length must be at least n"  # Real_FT.rs:6
            v = _inplace_view(data, n)
            if v.dtype not in (np.float32, np.float64):
                raise TypeError(f"unsupported dtype {v.dtype} (float32 or float64)")
            groups.setdefault((n, 1 if isign == 1 else -1, v.dtype.str), []).append(v)
        for (n, isign, _), views in groups.items():
            f32 = views[0].dtype == np.float32
            ct = ctypes.c_float if f32 else ctypes.c_double
            ptrs = (ctypes.POINTER(ct) * len(views))(*[v.ctypes.data_as(ctypes.POINTER(ct)) for v in views])
            fn = _lib.lib().nrfft_realft_batch_f32 if f32 else _lib.lib().nrfft_realft_batch_f64
            _check(fn(ptrs, n, len(views), isign))


# ---------------------------------------------------------------------------
# fourn (Fourn.rs / Real_FT3.rs:35)
# ---------------------------------------------------------------------------

def fourn(data: np.ndarray, nn: Sequence[int], ndim: int, isign: int) -> None:
    """In-memory NR fourn as Real_FT3.rs:35 calls it: in place, C order (nn[0]
    slowest).  Errors: Fourn.rs:367-378 (``FournError``)."""
    nn = [int(v) for v in nn]
    arr = (ctypes.c_size_t * max(len(nn), 1))(*nn)
    if ndim == 0 or ndim > len(nn):
        code = _lib.E_DIMS
        _lib.lib()  # message from the library for consistency
        raise FournError(code)
    total = int(np.prod(nn[:ndim]))
    v = _inplace_view(data, 2 * total) if all(x > 1 for x in nn[:ndim]) else data
    p, sfx = _buf(v)
    code = getattr(_lib.lib(), f"nrfft_fourn_{sfx}")(p, arr, ndim, isign)
    if code != _lib.OK:
        if code in (_lib.E_DIMS, _lib.E_ISIGN, _lib.E_NOT_POW2):
            raise FournError(code)
        raise NrfftError(code)


# ---------------------------------------------------------------------------
# convlv (Convolve.rs)
# ---------------------------------------------------------------------------

def _convlv_call(data: np.ndarray, respns: np.ndarray, isign: int, batch: int) -> np.ndarray:
    n = data.size // batch if batch else 0
    m = respns.size // batch if batch else 0
    ans = np.empty(max(n * batch, 1), data.dtype)
    pd, sfx = _buf(data if data.size else np.zeros(1, data.dtype))
    pr, _ = _buf(respns if respns.size else np.zeros(1, data.dtype))
    pa, _ = _buf(ans)
    L = _lib.lib()
    if batch == 1:
        code = getattr(L, f"nrfft_convlv_{sfx}")(pd, n, pr, m, isign, pa)
    else:
        code = getattr(L, f"nrfft_convlv_batch_{sfx}")(pd, n, pr, m, batch, isign, pa)
    if code != _lib.OK:
        raise ConvlvError(code)
    return ans[:n * batch]


def convlv(data, respns, isign: int) -> np.ndarray:
    """Convolve.rs:8 ``convlv(data, respns, isign) -> Result<Array1, ConvlvError>``:
    circular convolution (isign=1) or deconvolution (isign=-1) of data with the
    response placed as Convolve.rs:41-63 places it."""
    d = _as_input(data)
    r = _as_input(respns, d.dtype)
    if d.size == 0 or r.size == 0:
        raise ConvlvError(_lib.E_EMPTY)
    return _convlv_call(d, r, isign, 1)


def convlv_batch(data_batch: Sequence, respns, isign: int) -> list:
    """Convolve.rs:241-250: every signal convolved with the same response."""
    if len(data_batch) == 0:
        return []
    out = []
    groups: dict = {}
    for i, d in enumerate(data_batch):
        a = _as_input(d)
        groups.setdefault((a.dtype.str, a.size), []).append(i)
    res: list = [None] * len(data_batch)
    for (_, n), idx in groups.items():
        stack = np.ascontiguousarray(np.stack([_as_input(data_batch[i]) for i in idx]))
        r = _as_input(respns, stack.dtype)
        if n == 0 or r.size == 0:
            raise ConvlvError(_lib.E_EMPTY)
        rr = np.ascontiguousarray(np.tile(r, len(idx)))
        ans = _convlv_call(stack.reshape(-1), rr, isign, len(idx)).reshape(len(idx), n)
        for k, i in enumerate(idx):
            res[i] = ans[k].copy()
    out.extend(res)
    return out


def convlv_batch_responses(data: np.ndarray, respns: np.ndarray, isign: int) -> np.ndarray:
    """convlv over a (batch, n) array with one response per signal, (batch, m):
    the batched entry point nrfft_convlv_batch_*."""
    d = np.ascontiguousarray(data)
    r = np.ascontiguousarray(respns, dtype=d.dtype)
    if d.ndim != 2 or r.ndim != 2 or d.shape[0] != r.shape[0]:
        raise ValueError("data (batch, n) and respns (batch, m) expected")
    if d.size == 0 or r.size == 0:
        raise ConvlvError(_lib.E_EMPTY)
    return _convlv_call(d.reshape(-1), r.reshape(-1), isign, d.shape[0]).reshape(d.shape)


class ConvlvProcessor:
    """Convolve.rs:253-339."""

    def __init__(self) -> None:
        self.use_optimized = True
        self.parallel_threshold = 1024

    @classmethod
    def new(cls) -> "ConvlvProcessor":
        return cls()

    def with_optimized(self, use_optimized: bool) -> "ConvlvProcessor":
        self.use_optimized = use_optimized
        return self

    def with_threshold(self, threshold: int) -> "ConvlvProcessor":
        self.parallel_threshold = threshold
        return self

    def process(self, data, respns, isign: int) -> np.ndarray:
        return convlv(data, respns, isign)

    def process_batch(self, data_batch, respns, isign: int) -> list:
        return convlv_batch(data_batch, respns, isign)


# ---------------------------------------------------------------------------
# correl (Correlation.rs)
# ---------------------------------------------------------------------------

def correl(data1, data2) -> np.ndarray:
    """Correlation.rs:8 ``correl(data1, data2) -> Result<Array1, CorrelError>``.
    n <= 32: linear correlation at lags 0..n-1; n > 32: circular correlation
    (negative lags wrapped to the top)."""
    a = _as_input(data1)
    b = _as_input(data2, a.dtype)
    if a.size == 0:
        raise CorrelError("EmptyInput", _lib.E_EMPTY)
    if b.size != a.size:
        raise CorrelError("LengthMismatch", _lib.E_LEN)
    ans = np.empty(a.size, a.dtype)
    pa, sfx = _buf(a)
    pb, _ = _buf(b)
    po, _ = _buf(ans)
    code = getattr(_lib.lib(), f"nrfft_correl_{sfx}")(pa, pb, a.size, po)
    if code != _lib.OK:
        kind = {_lib.E_EMPTY: "EmptyInput", _lib.E_LEN: "LengthMismatch"}.get(code, "FftError")
        raise CorrelError(kind, code)
    return ans


def _moments(x: np.ndarray):
    mean = float(np.sum(x) / x.size)
    std = float(np.sqrt(np.sum((x - mean) ** 2) / x.size))
    return mean, std


def correl_normalized(data1, data2) -> np.ndarray:
    """Correlation.rs:189-223: z-score both inputs, then correl."""
    a = _as_input(data1, np.float64)
    b = _as_input(data2, np.float64)
    if a.size == 0:
        raise CorrelError("EmptyInput")
    if b.size != a.size:
        raise CorrelError("LengthMismatch")
    m1, s1 = _moments(a)
    m2, s2 = _moments(b)
    if s1 == 0.0 or s2 == 0.0:
        raise CorrelError("ZeroStdDev")
    return correl((a - m1) / s1, (b - m2) / s2)


def correl_normalized_fast(data1, data2) -> np.ndarray:
    """Correlation.rs:226-270 (single-pass moments).  For n <= 32 the
    reference divides the direct sum by (std1 std2 n); for n > 32 it returns
    correl of the z-scored inputs."""
    a = _as_input(data1, np.float64)
    b = _as_input(data2, np.float64)
    n = a.size
    if n == 0:
        raise CorrelError("EmptyInput")
    if b.size != n:
        raise CorrelError("LengthMismatch")
    m1, m2 = a.sum() / n, b.sum() / n
    s1 = float(np.sqrt((a * a).sum() / n - m1 * m1))
    s2 = float(np.sqrt((b * b).sum() / n - m2 * m2))
    if s1 == 0.0 or s2 == 0.0:
        raise CorrelError("ZeroStdDev")
    if n <= 32:
        return correl(a - m1, b - m2) * (1.0 / (s1 * s2 * n))
    return correl((a - m1) / s1, (b - m2) / s2)


def correl_batch(data_pairs: Sequence[tuple]) -> list:
    """Correlation.rs:273-278."""
    return [correl(a, b) for a, b in data_pairs]


def autocorrel(data) -> np.ndarray:
    """Correlation.rs:281-283."""
    return correl(data, data)


def autocorrel_fast(data) -> np.ndarray:
    """Correlation.rs:286-323 (|F|^2 then inverse; equal to autocorrel)."""
    a = _as_input(data)
    if a.size == 0:
        raise CorrelError("EmptyInput")
    return correl(a, a)


# ---------------------------------------------------------------------------
# twofft (FFT_2.rs), cosft1 (Cos_FT.rs), cosft2 (Cos_FT2.rs), rlft3 (Real_FT3.rs)
# NR semantics, 0-based (the Rust sources are 1-based NR translations).
# ---------------------------------------------------------------------------

def twofft(data1, data2, fft1: np.ndarray, fft2: np.ndarray) -> None:
    """FFT_2.rs:3 ``twofft(data1, data2, fft1, fft2)``: the isign = +1 spectra
    of two real sequences of n points from one complex FFT.  fft1 / fft2 are
    outputs of 2n + 2 values (asserted as FFT_2.rs:5-7): bins 0..n-1
    interleaved, then 0, 0."""
    d1 = np.asarray(data1)
    n = d1.size
    assert np.asarray(data2).size == n, "data2 length must equal data1 length"   # FFT_2.rs:5
    assert fft1.size == 2 * n + 2, "fft1 must have length 2*n + 2"               # FFT_2.rs:6
    assert fft2.size == 2 * n + 2, "fft2 must have length 2*n + 2"               # FFT_2.rs:7
    dt = fft1.dtype if fft1.dtype in (np.float32, np.float64) else np.float64
    a, b = _as_input(d1, dt), _as_input(data2, dt)
    for f in (fft1, fft2):
        _inplace_view(f, 2 * n + 2)
        if f.dtype != dt:
            raise TypeError("fft1 and fft2 must have the same float dtype")
    p1, sfx = _buf(a)
    p2, _ = _buf(b)
    q1, _ = _buf(fft1)
    q2, _ = _buf(fft2)
    _check(getattr(_lib.lib(), f"nrfft_twofft_{sfx}")(p1, p2, n, q1, q2))


twofft_optimized = twofft  # FFT_2.rs:134


def cosft1(y: np.ndarray, n: int) -> None:
    """Cos_FT.rs:7 ``cosft1(y, n)``: DCT-I in place on y[0..n] (n + 1 values),
    Y_k = (y_0 + (-1)^k y_n)/2 + sum_{j=1}^{n-1} y_j cos(pi j k / n)."""
    v = _inplace_view(y, n + 1)
    p, sfx = _buf(v)
    _check(getattr(_lib.lib(), f"nrfft_cosft1_{sfx}")(p, n))


cosft1_optimized = cosft1  # Cos_FT.rs:78


def cosft2(y: np.ndarray, n: int, isign: int) -> None:
    """Cos_FT2.rs:7 ``cosft2(y, n, isign)``: in place on y[0..n-1]; isign = 1
    the DCT-II Y_k = sum_j y_j cos(pi k (j + 1/2) / n), isign = -1 its
    inverse scaled by n/2.  Any other isign panics in the reference
    (Cos_FT2.rs:11); here it raises ``ValueError``."""
    if isign not in (1, -1):
        raise ValueError(f"Invalid isign value: {isign}. Must be 1 or -1")
    v = _inplace_view(y, n)
    p, sfx = _buf(v)
    _check(getattr(_lib.lib(), f"nrfft_cosft2_{sfx}")(p, n, isign))


def rlft3(data: np.ndarray, speq: np.ndarray, nn1: int, nn2: int, nn3: int, isign: int) -> None:
    """Real_FT3.rs:8 ``rlft3(data, speq, nn1, nn2, nn3, isign)`` with the
    asserts of Real_FT3.rs:16-19.  data: (nn1, nn2, nn3) reals, in place --
    forward leaves the half spectrum F[k1, k2, k3 < nn3/2] as interleaved
    complex values; speq: (nn1, 2 nn2) reals, the k3 = nn3/2 plane."""
    assert isign in (1, -1), "isign must be 1 or -1"
    assert tuple(data.shape) == (nn1, nn2, nn3), "data dimensions mismatch"
    assert tuple(speq.shape) == (nn1, 2 * nn2), "speq dimensions mismatch"
    for a in (data, speq):
        if not a.flags.c_contiguous or not a.flags.writeable:
            raise ValueError("data and speq must be writable C-contiguous arrays")
    if data.dtype != speq.dtype:
        raise TypeError("data and speq must have the same dtype")
    p, sfx = _buf(data.reshape(-1))
    q, _ = _buf(speq.reshape(-1))
    _check(getattr(_lib.lib(), f"nrfft_rlft3_{sfx}")(p, q, nn1, nn2, nn3, isign))


def twofft_batch(pairs: Sequence[tuple]) -> None:
    """Each (data1, data2, fft1, fft2) through ``twofft``."""
    for d1, d2, f1, f2 in pairs:
        twofft(d1, d2, f1, f2)


class TwoFFTProcessor:
    """FFT_2.rs:221-264: ``process`` / ``process_batch`` over (data1, data2,
    fft1, fft2) tuples; the size threshold only picks between CPU variants in
    the reference, every size runs on the GPU here."""

    def __init__(self) -> None:
        self.use_optimized = True
        self.parallel_threshold = 1024

    @classmethod
    def new(cls) -> "TwoFFTProcessor":
        return cls()

    def with_optimized(self, use_optimized: bool) -> "TwoFFTProcessor":
        self.use_optimized = use_optimized
        return self

    def with_threshold(self, threshold: int) -> "TwoFFTProcessor":
        self.parallel_threshold = threshold
        return self

    def process(self, data1, data2, fft1: np.ndarray, fft2: np.ndarray) -> None:
        twofft(data1, data2, fft1, fft2)

    def process_batch(self, batches: Sequence[tuple]) -> None:
        twofft_batch(batches)


def extract_real_imag(fft) -> tuple:
    """FFT_2.rs:361-373."""
    f = np.asarray(fft, dtype=np.float64)
    n = f.size // 2
    return f[0:2 * n:2].copy(), f[1:2 * n:2].copy()


def combine_real_imag(real, imag) -> np.ndarray:
    """FFT_2.rs:375-384."""
    r, i = np.asarray(real, dtype=np.float64), np.asarray(imag, dtype=np.float64)
    assert r.size == i.size, "Real and imaginary parts must have same length"
    out = np.empty(2 * r.size)
    out[0::2], out[1::2] = r, i
    return out


cosft2_simd = cosft2  # Cos_FT2.rs:202


def rlft3_optimized(data: np.ndarray, speq: np.ndarray, nn1: int, nn2: int, nn3: int, isign: int) -> None:
    """Real_FT3.rs:145: rlft3 on flat slices (asserts of Real_FT3.rs:153-154)."""
    assert data.size >= nn1 * nn2 * nn3, "data array too small"
    assert speq.size >= nn1 * 2 * nn2, "speq array too small"
    d = _inplace_view(data, nn1 * nn2 * nn3).reshape(nn1, nn2, nn3)
    s = _inplace_view(speq, nn1 * 2 * nn2).reshape(nn1, 2 * nn2)
    rlft3(d, s, nn1, nn2, nn3, isign)


# ---------------------------------------------------------------------------
# Out-of-core fourn (Four_FS.rs)
# ---------------------------------------------------------------------------

def _fourfs_call(data: np.ndarray, scratch, nn: Sequence[int], ndim: int, isign: int, device_bytes: int) -> None:
    nn = [int(v) for v in nn]
    if ndim == 0 or ndim > len(nn):                       # Four_FS.rs:329-331
        raise FournError(_lib.E_DIMS, "Invalid dimensions")
    arr = (ctypes.c_size_t * len(nn))(*nn)
    p, sfx = _buf(data)
    q = None
    if scratch is not None:
        if scratch.dtype != data.dtype or scratch.size < data.size:
            raise ValueError("scratch must match data's dtype and size")
        q, _ = _buf(scratch)
    code = getattr(_lib.lib(), f"nrfft_fourfs_{sfx}")(p, q, arr, ndim, isign, int(device_bytes))
    if code != _lib.OK:
        if code in (_lib.E_DIMS, _lib.E_ISIGN, _lib.E_NOT_POW2):
            raise FournError(code)
        raise NrfftError(code)


def fourfs(data: np.ndarray, nn: Sequence[int], ndim: int, isign: int, device_bytes: int = 0,
           scratch: np.ndarray | None = None) -> None:
    """The out-of-core n-D FFT of Four_FS.rs:55 on a host array (a numpy
    memmap of a file works): in place, C order, fourn's results, streamed
    through HBM with at most ``device_bytes`` of device memory (0: 1 GiB).
    A 1-D transform larger than half the budget needs ``scratch`` (same size;
    allocated here when not given)."""
    total = int(np.prod([int(v) for v in nn][:max(ndim, 0)])) if ndim > 0 else 0
    v = _inplace_view(data, 2 * total) if total > 0 else data
    if scratch is None and ndim == 1 and total > 0 and 2 * total * v.itemsize * 2 > (device_bytes or (1 << 30)):
        scratch = np.empty_like(v)
    _fourfs_call(v, scratch, nn, ndim, isign, device_bytes)


class FourFS:
    """Four_FS.rs:13-75.  ``new(temp_dir)`` creates four temporary files
    (fft_1 .. fft_4, Four_FS.rs:20-40).  The data to transform live in fft_1
    as interleaved complex float64 values in C order (``write_data`` /
    ``read_data``); ``fourfs(nn, ndim, isign)`` transforms them in place,
    memory-mapped and streamed through HBM; fft_2 is the scratch of 1-D
    transforms larger than the device budget."""

    def __init__(self, temp_dir, device_bytes: int = 0) -> None:
        import os
        import tempfile

        self.paths = []
        for k in range(1, 5):
            fd, path = tempfile.mkstemp(prefix=f"fft_{k}", dir=str(temp_dir))
            os.close(fd)
            self.paths.append(path)
        self.device_bytes = device_bytes
        self.memory_mapped = False

    @classmethod
    def new(cls, temp_dir) -> "FourFS":
        return cls(temp_dir)

    def with_memory_mapping(self) -> "FourFS":
        """Four_FS.rs:42-53 (the transform always memory-maps the files)."""
        self.memory_mapped = True
        return self

    def write_data(self, data) -> None:
        np.ascontiguousarray(data, dtype=np.float64).tofile(self.paths[0])

    def read_data(self) -> np.ndarray:
        return np.fromfile(self.paths[0], dtype=np.float64)

    def fourfs(self, nn: Sequence[int], ndim: int, isign: int) -> None:
        nn = [int(v) for v in nn]
        if ndim == 0 or ndim > len(nn):                          # Four_FS.rs:329-331
            raise FournError(_lib.E_DIMS, "Invalid dimensions")
        if isign not in (1, -1):                                   # Four_FS.rs:332-334
            raise FournError(_lib.E_ISIGN, "isign must be 1 or -1")
        if any(v <= 1 for v in nn[:ndim]):                         # Four_FS.rs:335-337
            raise FournError(_lib.E_DIMS, "Invalid dimension size")
        total = int(np.prod(nn[:ndim]))
        data = np.memmap(self.paths[0], dtype=np.float64, mode="r+", shape=(2 * total,))
        scratch = None
        if ndim == 1:
            scratch = np.memmap(self.paths[1], dtype=np.float64, mode="w+", shape=(2 * total,))
        try:
            _fourfs_call(data, scratch, nn, ndim, isign, self.device_bytes)
            data.flush()
        finally:
            del data, scratch
