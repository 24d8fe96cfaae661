"""GPU parity at the benchmark configurations' full sizes, against the oracle.

* fourn 256 x 256 x 256, FP64 and FP32, N(0,1) from seed 1, both signs: the
  whole transform against the oracle's NR fourn (Real_FT3.rs:33-35
  convention); FP64 <= 1e-12, FP32 <= 2e-6 relative L2.
* batch 64 x 2^20 FP32 C2C (uniform[-1,1) from seed 3) in the plain and both
  split-precision forms: the first and last transform of every 16-transform
  chunk against the oracle's four1, <= 2e-6 (plain) and <= 1e-6 (split).
* the 10,000-item uneven batch (seed 4): 240 items drawn at random (seeded)
  against the oracle's four1 / fourn, FP32 <= 2e-6.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import numrs_amd
import oracle
from conftest import rel_l2
from numrs_amd import synth

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 2e-6


def _pool():
    return ThreadPoolExecutor(max_workers=8)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("isign", [1, -1])
def test_fourn_256cubed_full(dtype, isign):
    # The fourn configuration at full size: 2^24 complex points, every
    # value compared with the oracle.
    import torch

    shape = [256, 256, 256]
    x = synth.normal(1, 2 * 256 ** 3)
    if dtype == "float32":
        x = x.astype(np.float32).astype(np.float64)
    t = torch.from_numpy(x.astype(np.float32 if dtype == "float32" else np.float64)).cuda()
    plan = numrs_amd.C2CPlan(dtype, shape, 1)
    plan.execute(t, t, isign)
    got = t.cpu().numpy().astype(np.float64)
    del t
    ref = oracle.fourn(x, shape, isign)
    tol = TOL32 if dtype == "float32" else TOL64
    err = rel_l2(got, ref)
    assert err < tol, (dtype, isign, err)


@pytest.mark.parametrize("split", [None, "f16", "bf16"])
def test_c2c_batch64_2p20_chunks(split):
    # The headline batch (64 x 2^20 FP32, seed 3) through the plain and the
    # split-precision plans; each plan runs 16-transform chunks, and the
    # transforms on both sides of every chunk boundary are checked.
    import torch

    batch, n = 64, 1 << 20
    x = torch.empty(2 * n * batch, dtype=torch.float32, device="cuda")
    numrs_amd.gen_uniform(x, 3)
    y = torch.empty_like(x)
    plan = numrs_amd.C2CPlan("float32", [n], batch, split=split)
    assert plan.num_launches == 8  # 4 chunks x (pass A, pass B)
    plan.execute(x, y, 1)
    torch.cuda.synchronize()
    idx = sorted({c + d for c in range(0, batch, 16) for d in (0, 15)})
    xs = {i: x[2 * n * i:2 * n * (i + 1)].cpu().numpy().astype(np.float64) for i in idx}
    ys = {i: y[2 * n * i:2 * n * (i + 1)].cpu().numpy().astype(np.float64) for i in idx}
    del x, y
    with _pool() as ex:
        refs = dict(zip(idx, ex.map(lambda i: oracle.four1_nr(xs[i], 1), idx)))
    tol = TOL32 if split is None else 1e-6
    for i in idx:
        err = rel_l2(ys[i], refs[i])
        assert err < tol, (split, i, err)


def test_uneven_10000_sampled():
    # The uneven configuration at full size (10,000 items, 80% 1-D with
    # log2 N on 8..20, 20% 2-D with sides 2^6..2^11, inputs from seed 4
    # generated in HBM), one graph-replayed queue execution; 240 items drawn
    # at random against the oracle.
    import torch

    items = synth.uneven_shapes(4, 10000)
    plan = numrs_amd.UnevenPlan("float32", items)
    x = torch.empty(2 * plan.total, dtype=torch.float32, device="cuda")
    numrs_amd.gen_uniform(x, 4)
    y = torch.empty_like(x)
    plan.execute(x, y, 1)
    plan.check()
    rng = np.random.default_rng(2024)
    pick = sorted(rng.choice(len(items), size=240, replace=False).tolist())
    # every 1-D length class and both kinds are in the sample
    kinds = {(len(items[i]), int(np.prod(items[i])) if len(items[i]) == 1 else 0) for i in pick}
    assert sum(1 for k in kinds if k[0] == 2) == 1 and len(kinds) >= 10, kinds
    ins, outs = {}, {}
    for i in pick:
        o, c = plan.offsets[i], int(np.prod(items[i]))
        ins[i] = x[2 * o:2 * (o + c)].cpu().numpy().astype(np.float64)
        outs[i] = y[2 * o:2 * (o + c)].cpu().numpy().astype(np.float64)
    del x, y

    def ref(i):
        it = items[i]
        return oracle.four1_nr(ins[i], 1) if len(it) == 1 else oracle.fourn(ins[i], list(it), 1)

    with _pool() as ex:
        refs = dict(zip(pick, ex.map(ref, pick)))
    worst = max((rel_l2(outs[i], refs[i]), i) for i in pick)
    assert worst[0] < TOL32, (worst, items[worst[1]])
