"""GPU: the allocation-free batched four1 (nrfft_execute_four1_batch_noalloc)
through the C ABI against the plan-based entry point (nrfft_plan_c2c +
nrfft_execute_c2c) on the same inputs: bit-identical results (same kernels,
only the workspace and the batch chunking differ), and the oracle on the
first and last transform.  Two calls in a row, the second with a smaller
input; sizes beyond the scratch are refused before anything runs."""
import ctypes

import numpy as np
import pytest

import numrs_amd
import oracle
from conftest import rel_l2
from numrs_amd import _lib

pytestmark = pytest.mark.gpu


def _noalloc(dtype_code, x, y, n, batch, isign):
    import torch

    return _lib.lib().nrfft_execute_four1_batch_noalloc(dtype_code, ctypes.c_void_p(x.data_ptr()),
                                                        ctypes.c_void_p(y.data_ptr()), n, batch, isign,
                                                        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))


def _plan_result(dtype, x, n, batch, isign):
    import torch

    # the plan-based entry point through the C ABI
    h = ctypes.c_void_p()
    arr = (ctypes.c_size_t * 1)(n)
    code = _lib.F32 if dtype == "float32" else _lib.F64
    assert _lib.lib().nrfft_plan_c2c(ctypes.byref(h), code, arr, 1, batch) == 0, _lib.last_error()
    y = torch.empty_like(x)
    rc = _lib.lib().nrfft_execute_c2c(h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), isign,
                                      ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    _lib.lib().nrfft_plan_destroy(h)
    assert rc == 0, _lib.last_error()
    return y


@pytest.mark.parametrize("dtype,first,second", [
    ("float32", (1 << 20, 20, 1), (1 << 16, 10, -1)),   # scratch chunks 8 + 8 + 4, then 64 + ... one chunk
    ("float64", (1 << 20, 5, -1), (1 << 14, 3, 1)),     # chunks 4 + 1, then one chunk
    ("float32", (1 << 12, 7, 1), (1 << 5, 2, -1)),      # single-pass lengths (no scratch)
])
def test_noalloc_matches_plan(dtype, first, second):
    import torch

    assert _lib.lib().nrfft_init(0) == 0, _lib.last_error()
    code = _lib.F32 if dtype == "float32" else _lib.F64
    tdt = torch.float32 if dtype == "float32" else torch.float64
    for seed, (n, batch, isign) in enumerate((first, second)):
        print(f"noalloc case {dtype} n={n} batch={batch} isign={isign}", flush=True)
        x = torch.empty(2 * n * batch, dtype=tdt, device="cuda")
        numrs_amd.gen_uniform(x, 50 + seed)
        y = torch.full_like(x, float("nan"))
        assert _noalloc(code, x, y, n, batch, isign) == 0, _lib.last_error()
        torch.cuda.synchronize()
        ref = _plan_result(dtype, x, n, batch, isign)
        assert torch.equal(y, ref), (n, batch)
        xh, yh = x.cpu().numpy().astype(np.float64), y.cpu().numpy().astype(np.float64)
        tol = 2e-6 if dtype == "float32" else 1e-12
        for i in (0, batch - 1):
            seg = slice(2 * n * i, 2 * n * (i + 1))
            assert rel_l2(yh[seg], oracle.four1_nr(xh[seg], isign)) < tol, (n, i)


def test_noalloc_in_place_and_limits():
    import torch

    assert _lib.lib().nrfft_init(0) == 0, _lib.last_error()
    n, batch = 1 << 18, 9
    x = torch.empty(2 * n * batch, dtype=torch.float32, device="cuda")
    numrs_amd.gen_uniform(x, 60)
    ref = _plan_result("float32", x, n, batch, 1)
    assert _noalloc(_lib.F32, x, x, n, batch, 1) == 0, _lib.last_error()
    torch.cuda.synchronize()
    assert torch.equal(x, ref)
    # one FP32 transform of 2^24 points needs 128 MiB of scratch: refused
    big = torch.zeros(2 * (1 << 24), dtype=torch.float32, device="cuda")
    out = torch.empty_like(big)
    assert _noalloc(_lib.F32, big, out, 1 << 24, 1, 1) == _lib.E_UNSUPPORTED
    assert "scratch" in _lib.last_error()
    assert _noalloc(_lib.F64, big, out, 1 << 23, 1, 1) == _lib.E_UNSUPPORTED
    assert _noalloc(_lib.F32, big, out, 3, 1, 1) == _lib.E_NOT_POW2
    assert _noalloc(_lib.F32, big, out, 8, 1, 0) == _lib.E_ISIGN
    assert _noalloc(_lib.F32, big, out, 8, 0, 1) == _lib.E_EMPTY
    torch.cuda.synchronize()
