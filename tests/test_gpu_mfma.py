"""GPU: split-precision MFMA C2C (FP16 hi/lo and BF16 hi/mid/lo DFT-matrix
products, FP32 accumulation) against the FP64 oracle on the same FP32 inputs.

Tolerance: relative L2 <= 1e-6 versus FP64 for every length (the bar the
split-precision form is held to); the plain FP32 path on the same data is
checked to the same bar for comparison.
"""
import numpy as np
import pytest

import numrs_amd
import oracle
from conftest import rel_l2
from numrs_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-6


def _run(x32: np.ndarray, n: int, batch: int, isign: int, split):
    import torch

    xt = torch.from_numpy(x32).cuda()
    yt = torch.empty_like(xt)
    plan = numrs_amd.C2CPlan("float32", [n], batch, split=split)
    plan.execute(xt, yt, isign)
    torch.cuda.synchronize()
    return yt.cpu().numpy()


@pytest.mark.parametrize("split", ["f16", "bf16", None])
@pytest.mark.parametrize("logn,batch", [(8, 3), (9, 2), (10, 17), (14, 4), (16, 2), (18, 2), (20, 2)])
def test_split_c2c_accuracy(split, logn, batch):
    n = 1 << logn
    x32 = synth.uniform(3 + logn, 2 * n * batch).astype(np.float32)
    for isign in (1, -1):
        y = _run(x32, n, batch, isign, split)
        errs = []
        for b in range(batch if n <= 65536 else 1):
            seg = slice(2 * n * b, 2 * n * (b + 1))
            ref = oracle.four1_nr(x32[seg].astype(np.float64), isign)
            errs.append(rel_l2(y[seg], ref))
        assert max(errs) < TOL, (split, n, isign, max(errs))


@pytest.mark.parametrize("split", ["f16", "bf16"])
def test_split_c2c_large_and_small_magnitudes(split):
    # The FP16 form rescales each tile by a power of two; values far from 1
    # must keep the same relative accuracy.
    n, batch = 1 << 10, 16
    base = synth.uniform(9, 2 * n * batch)
    for scale in (1e-6, 3.7e4, 1e9):
        x32 = (base * scale).astype(np.float32)
        y = _run(x32, n, batch, 1, split)
        ref = oracle.four1_nr(x32[:2 * n].astype(np.float64), 1)
        assert rel_l2(y[:2 * n], ref) < TOL, (split, scale)


@pytest.mark.parametrize("split", ["f16", "bf16"])
def test_split_c2c_roundtrip(split):
    n, batch = 1 << 20, 2
    x32 = synth.uniform(30, 2 * n * batch).astype(np.float32)
    y = _run(x32, n, batch, 1, split)
    z = _run(y, n, batch, -1, split) / n
    assert rel_l2(z, x32.astype(np.float64)) < 2 * TOL
