"""Device-pointer plans (data resident in HBM), for torch tensors or raw
device addresses.  Execution is enqueued on the caller's HIP stream.

Every execute checks each torch tensor it is given before anything is
launched: contiguous, on the plan's GPU, of the plan's dtype and holding at
least the values the plan reads or writes (a short or wrongly typed buffer
would make the kernels access memory past its end).  Raw integer addresses
are passed through unchecked: the caller vouches for them."""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from . import _lib
from .api import NrfftError


def _dtype_code(dtype) -> int:
    name = str(dtype)
    if name in ("float32", "torch.float32", "f32", "<f4"):
        return _lib.F32
    if name in ("float64", "torch.float64", "f64", "<f8"):
        return _lib.F64
    raise TypeError(f"unsupported dtype {dtype}")


_DT_NAME = {_lib.F32: "float32", _lib.F64: "float64"}


def _check(code: int) -> None:
    if code != _lib.OK:
        raise NrfftError(code)


def _current_device():
    try:
        import torch

        if torch.cuda.is_available():
            return torch.cuda.current_device()
    except Exception:  # noqa: BLE001 - no torch or no GPU: device unknown
        pass
    return None


def _ptr(x) -> int:
    """Device address of a torch tensor (or an int address)."""
    if isinstance(x, int):
        return x
    if not x.is_contiguous():
        raise ValueError("tensor must be contiguous")
    if x.device.type != "cuda":
        raise ValueError("tensor must live in GPU memory")
    return x.data_ptr()


_RAW_STREAM = None  # torch's current-stream lookup (the cheapest one this torch has)


def _stream(stream) -> int:
    global _RAW_STREAM
    if stream is None:
        if _RAW_STREAM is None:
            try:
                import torch

                raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
                if raw is not None:
                    _RAW_STREAM = lambda: raw(torch.cuda.current_device())  # noqa: E731
                else:
                    _RAW_STREAM = lambda: torch.cuda.current_stream().cuda_stream  # noqa: E731
            except Exception:  # noqa: BLE001 - no torch: default stream
                _RAW_STREAM = lambda: 0  # noqa: E731
        return _RAW_STREAM()
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


class _Plan:
    def __init__(self):
        self._h = ctypes.c_void_p()
        self._dtype = None
        self._tdtype = None
        self._device = _current_device()

    def __del__(self):
        try:
            self.destroy()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    def destroy(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _lib.lib().nrfft_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def _buf(self, x, values: int, name: str) -> int:
        """Address of buffer x after checking that it can hold `values`
        scalars of the plan's dtype on the plan's device."""
        if not self._h.value:
            raise ValueError("plan has been destroyed")
        if isinstance(x, int):
            return x
        tdt = self._tdtype
        if tdt is None:
            import torch

            tdt = self._tdtype = torch.float32 if self._dtype == _lib.F32 else torch.float64
        if not x.is_cuda:
            raise ValueError(f"{name}: tensor must live in GPU memory")
        if not x.is_contiguous():
            raise ValueError(f"{name}: tensor must be contiguous")
        if x.dtype is not tdt:
            raise ValueError(f"{name}: dtype {x.dtype} does not match the plan's {_DT_NAME[self._dtype]}")
        if self._device is not None and x.get_device() != self._device:
            raise ValueError(f"{name}: tensor on cuda:{x.get_device()}, plan on cuda:{self._device}")
        if x.numel() < values:
            raise ValueError(f"{name}: {x.numel()} values, the plan needs {values}")
        return x.data_ptr()

    @property
    def workspace_bytes(self) -> int:
        return int(_lib.lib().nrfft_plan_workspace_bytes(self._h))

    @property
    def num_launches(self) -> int:
        return int(_lib.lib().nrfft_plan_num_launches(self._h))


class C2CPlan(_Plan):
    """Batched C2C of shape nn (C order); ndim 1 = four1, ndim >= 2 = fourn.
    ``split`` = "f16" / "bf16" selects the split-precision MFMA form (float32).
    ``tune`` = (split_log1, tilog_a, tilog_b, chunk, wg_per_cu, xcd_remap)
    (a prefix is enough; -1 = default) sets the tunables of this plan only."""

    def __init__(self, dtype, nn: Sequence[int], batch: int = 1, tune: Sequence[int] | None = None,
                 split: str | None = None):
        super().__init__()
        nn = [int(v) for v in nn]
        arr = (ctypes.c_size_t * len(nn))(*nn)
        self._dtype = _dtype_code(dtype)
        prec = 0
        if split is not None:
            if self._dtype != _lib.F32:
                raise TypeError("split precision applies to float32 plans")
            prec = {"f16": _lib.SPLIT_F16, "bf16": _lib.SPLIT_BF16}[split]
        if tune is not None:
            tv = [int(v) for v in tune][:6]
            tv += [-1] * (6 - len(tv))
            tv.append(prec)
            t = (ctypes.c_int * 7)(*tv)
            _check(_lib.lib().nrfft_plan_c2c_tuned(ctypes.byref(self._h), self._dtype, arr, len(nn), batch, t, 7))
        elif prec:
            _check(_lib.lib().nrfft_plan_c2c_split(ctypes.byref(self._h), arr, len(nn), batch, prec))
        else:
            _check(_lib.lib().nrfft_plan_c2c(ctypes.byref(self._h), self._dtype, arr, len(nn), batch))
        self.nn, self.batch = nn, batch
        self._values = 2 * batch * int(np.prod(nn))

    def execute(self, x_in, x_out, isign: int, stream=None) -> None:
        a = self._buf(x_in, self._values, "x_in")
        b = self._buf(x_out, self._values, "x_out")
        _check(_lib.lib().nrfft_execute_c2c(self._h, a, b, isign, _stream(stream)))

    @property
    def tune(self) -> tuple:
        """The tunables this plan was built with: (split_log1, tilog_a,
        tilog_b, chunk, wg_per_cu, xcd_remap), -1 = planner default."""
        t = (ctypes.c_int * 6)()
        _check(_lib.lib().nrfft_plan_tune(self._h, t, 6))
        return tuple(t)

    def execute_launches(self, x_in, x_out, isign: int, first: int, count: int, stream=None) -> None:
        """Enqueue only launches [first, first+count) (e.g. one pass)."""
        a = self._buf(x_in, self._values, "x_in")
        b = self._buf(x_out, self._values, "x_out")
        _check(_lib.lib().nrfft_execute_c2c_launches(self._h, a, b, isign, _stream(stream), first, count))


def autotune_c2c(dtype, nn: Sequence[int], batch: int = 1, iters: int = 5):
    """Time every candidate tuning of a C2C plan on the current GPU and cache
    the fastest.  Returns ((split_log1, tilog_a, tilog_b, chunk, wg_per_cu,
    xcd_remap), ms)."""
    nn = [int(v) for v in nn]
    arr = (ctypes.c_size_t * len(nn))(*nn)
    best = (ctypes.c_int * 6)()
    ms = ctypes.c_double()
    _check(_lib.lib().nrfft_autotune_c2c(_dtype_code(dtype), arr, len(nn), batch, iters, best, ctypes.byref(ms)))
    return tuple(best), ms.value


def tune_cache_save(path: str) -> None:
    _check(_lib.lib().nrfft_tune_cache_save(path.encode()))


def tune_cache_load(path: str) -> None:
    _check(_lib.lib().nrfft_tune_cache_load(path.encode()))


class RealFTPlan(_Plan):
    def __init__(self, dtype, n: int, batch: int = 1):
        super().__init__()
        self._dtype = _dtype_code(dtype)
        _check(_lib.lib().nrfft_plan_realft(ctypes.byref(self._h), self._dtype, n, batch))
        self.n, self.batch = n, batch

    def execute(self, x, isign: int, stream=None) -> None:
        _check(_lib.lib().nrfft_execute_realft(self._h, self._buf(x, self.n * self.batch, "x"), isign,
                                               _stream(stream)))


class ConvlvPlan(_Plan):
    def __init__(self, dtype, n: int, m: int, batch: int = 1):
        super().__init__()
        self._dtype = _dtype_code(dtype)
        _check(_lib.lib().nrfft_plan_convlv(ctypes.byref(self._h), self._dtype, n, m, batch))
        self.n, self.m, self.batch = n, m, batch

    def execute(self, data, respns, ans, isign: int = 1, stream=None) -> None:
        nb, mb = self.n * self.batch, self.m * self.batch
        _check(_lib.lib().nrfft_execute_convlv(self._h, self._buf(data, nb, "data"), self._buf(respns, mb, "respns"),
                                               self._buf(ans, nb, "ans"), isign, _stream(stream)))


class CorrelPlan(_Plan):
    def __init__(self, dtype, n: int, batch: int = 1):
        super().__init__()
        self._dtype = _dtype_code(dtype)
        _check(_lib.lib().nrfft_plan_correl(ctypes.byref(self._h), self._dtype, n, batch))
        self.n, self.batch = n, batch

    def execute(self, d1, d2, ans, stream=None) -> None:
        nb = self.n * self.batch
        _check(_lib.lib().nrfft_execute_correl(self._h, self._buf(d1, nb, "d1"), self._buf(d2, nb, "d2"),
                                               self._buf(ans, nb, "ans"), _stream(stream)))


class TwofftPlan(_Plan):
    """Batched twofft: batch pairs of real n-point sequences -> two spectra of
    n complex values each (2n reals per transform)."""

    def __init__(self, dtype, n: int, batch: int = 1):
        super().__init__()
        self._dtype = _dtype_code(dtype)
        _check(_lib.lib().nrfft_plan_twofft(ctypes.byref(self._h), self._dtype, n, batch))
        self.n, self.batch = n, batch

    def execute(self, d1, d2, fft1, fft2, stream=None) -> None:
        nb = self.n * self.batch
        _check(_lib.lib().nrfft_execute_twofft(self._h, self._buf(d1, nb, "d1"), self._buf(d2, nb, "d2"),
                                               self._buf(fft1, 2 * nb, "fft1"), self._buf(fft2, 2 * nb, "fft2"),
                                               _stream(stream)))


class Cosft1Plan(_Plan):
    """Batched cosft1 (DCT-I): n + 1 reals per transform."""

    def __init__(self, dtype, n: int, batch: int = 1):
        super().__init__()
        self._dtype = _dtype_code(dtype)
        _check(_lib.lib().nrfft_plan_cosft1(ctypes.byref(self._h), self._dtype, n, batch))
        self.n, self.batch = n, batch

    def execute(self, x, y, stream=None) -> None:
        nb = (self.n + 1) * self.batch
        _check(_lib.lib().nrfft_execute_cosft1(self._h, self._buf(x, nb, "x"), self._buf(y, nb, "y"),
                                               _stream(stream)))


class Cosft2Plan(_Plan):
    """Batched cosft2 (DCT-II forward, scaled DCT-III inverse): n reals per transform."""

    def __init__(self, dtype, n: int, batch: int = 1):
        super().__init__()
        self._dtype = _dtype_code(dtype)
        _check(_lib.lib().nrfft_plan_cosft2(ctypes.byref(self._h), self._dtype, n, batch))
        self.n, self.batch = n, batch

    def execute(self, x, y, isign: int, stream=None) -> None:
        nb = self.n * self.batch
        _check(_lib.lib().nrfft_execute_cosft2(self._h, self._buf(x, nb, "x"), self._buf(y, nb, "y"), isign,
                                               _stream(stream)))


class Rlft3Plan(_Plan):
    """Batched rlft3: data nn1 x nn2 x nn3 reals (in place), speq 2 nn1 nn2 reals per transform."""

    def __init__(self, dtype, nn1: int, nn2: int, nn3: int, batch: int = 1):
        super().__init__()
        self._dtype = _dtype_code(dtype)
        _check(_lib.lib().nrfft_plan_rlft3(ctypes.byref(self._h), self._dtype, nn1, nn2, nn3, batch))
        self.shape, self.batch = (nn1, nn2, nn3), batch

    def execute(self, data, speq, isign: int, stream=None) -> None:
        nn1, nn2, nn3 = self.shape
        _check(_lib.lib().nrfft_execute_rlft3(self._h, self._buf(data, nn1 * nn2 * nn3 * self.batch, "data"),
                                              self._buf(speq, 2 * nn1 * nn2 * self.batch, "speq"), isign,
                                              _stream(stream)))


class UnevenPlan(_Plan):
    """Uneven batch: items of different lengths / dimensionality in one
    persistent work-queue launch (replayed as a captured hipGraph).
    ``items`` is a list of shapes, (N,) or (n0, n1); ``offsets`` are the items'
    complex-element offsets in the buffers (default: packed back to back)."""

    def __init__(self, dtype, items, offsets=None):
        super().__init__()
        self._dtype = _dtype_code(dtype)
        items = [tuple(int(v) for v in it) for it in items]
        n = len(items)
        nd = (ctypes.c_int * n)(*[len(it) for it in items])
        dims = (ctypes.c_size_t * (2 * n))(*[v for it in items for v in (it[0], it[1] if len(it) > 1 else 1)])
        if offsets is None:
            offsets, acc = [], 0
            for it in items:
                offsets.append(acc)
                acc += int(np.prod(it))
        self.offsets = [int(o) for o in offsets]
        self.items = items
        offs = (ctypes.c_size_t * n)(*self.offsets)
        _check(_lib.lib().nrfft_plan_uneven(ctypes.byref(self._h), self._dtype, n, nd, dims, offs))
        self.total = max(o + int(np.prod(it)) for o, it in zip(self.offsets, items))

    def destroy(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _lib.lib().nrfft_uneven_destroy(self._h)
            self._h = ctypes.c_void_p()

    @property
    def workspace_bytes(self) -> int:
        return self.info()["workspace_bytes"]

    @property
    def num_launches(self) -> int:
        return 3  # counter reset, phase 1, phase 2

    def execute(self, x_in, x_out, isign: int, stream=None, graph: bool = True) -> None:
        fn = _lib.lib().nrfft_execute_uneven if graph else _lib.lib().nrfft_execute_uneven_nograph
        a = self._buf(x_in, 2 * self.total, "x_in")
        b = self._buf(x_out, 2 * self.total, "x_out")
        _check(fn(self._h, a, b, isign, _stream(stream)))

    def check(self, stream=None) -> None:
        """Wait for the stream; raise if the last execution left tasks untaken."""
        _check(_lib.lib().nrfft_uneven_check(self._h, _stream(stream)))

    def info(self) -> dict:
        a, b, w = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
        g = ctypes.c_uint()
        _check(_lib.lib().nrfft_uneven_info(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(w),
                                            ctypes.byref(g)))
        return {"tasks_phase1": a.value, "tasks_phase2": b.value, "workspace_bytes": w.value, "groups": g.value}


def autotune_uneven(dtype, log2n_min: int = 14, log2n_max: int = 20, iters: int = 5) -> dict:
    """Autotune the four-step split of the uneven-batch 1-D size classes
    2^log2n_min .. 2^log2n_max; returns {log2 N: (log2 N1, ms)} for the
    classes that take the four-step path.  Cached like the C2C tunings."""
    k = log2n_max - log2n_min + 1
    l1 = (ctypes.c_int * k)()
    ms = (ctypes.c_float * k)()
    _check(_lib.lib().nrfft_autotune_uneven(_dtype_code(dtype), log2n_min, log2n_max, iters, l1, ms))
    return {log2n_min + i: (l1[i], ms[i]) for i in range(k) if l1[i] > 0}


def gen_uniform(x, seed: int, start: int = 0, stream=None) -> None:
    """Fill tensor x (float32/float64, contiguous, on the GPU) with the seeded
    uniform[-1,1) stream of numrs_amd.synth, generated in HBM."""
    if str(x.dtype) not in ("torch.float32", "torch.float64"):
        raise ValueError(f"gen_uniform: dtype {x.dtype} is not float32 / float64")
    n = x.numel()
    fn = _lib.lib().nrfft_gen_uniform_f32 if str(x.dtype) == "torch.float32" else _lib.lib().nrfft_gen_uniform_f64
    _check(fn(_ptr(x), n, seed, start, _stream(stream)))
