"""GPU: uneven batches in one persistent work-queue launch (captured graph),
against the oracle item by item.

Tolerances: FP64 1e-12 relative L2 per item; FP32 2e-6 versus the FP64 oracle
on the FP32-rounded inputs.
"""
import numpy as np
import pytest

import numrs_amd
import oracle
from conftest import rel_l2
from numrs_amd import synth

pytestmark = pytest.mark.gpu


def _items(seed: int, count: int, max_log1d: int, max_side: int):
    rng = np.random.default_rng(seed)
    items = []
    for _ in range(count):
        if rng.random() < 0.75:
            items.append((1 << int(rng.integers(6, max_log1d + 1)),))
        else:
            items.append((1 << int(rng.integers(6, max_side + 1)), 1 << int(rng.integers(6, max_side + 1))))
    return items


def _ref(x64: np.ndarray, it, isign: int) -> np.ndarray:
    if len(it) == 1:
        return oracle.four1_nr(x64, isign)
    return oracle.fourn(x64, list(it), isign)


@pytest.mark.parametrize("dtype,tol", [("float64", 1e-12), ("float32", 2e-6)])
def test_uneven_vs_oracle(dtype, tol):
    import torch

    items = _items(11, 40, 18, 9)
    plan = numrs_amd.UnevenPlan(dtype, items)
    tdt = getattr(torch, dtype)
    npdt = np.float64 if dtype == "float64" else np.float32
    x = synth.uniform(4, 2 * plan.total).astype(npdt)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    for isign in (1, -1):
        plan.execute(xt, yt, isign)
        plan.check()
        y = yt.cpu().numpy()
        for it, off in zip(items, plan.offsets):
            n = int(np.prod(it))
            seg = slice(2 * off, 2 * (off + n))
            assert rel_l2(y[seg], _ref(x[seg].astype(np.float64), it, isign)) < tol, (dtype, it, isign)


def test_uneven_graph_replays():
    # Replays of the captured graph, with eager work between them, each checked.
    import torch

    items = _items(12, 25, 16, 8)
    plan = numrs_amd.UnevenPlan("float64", items)
    x = torch.from_numpy(synth.uniform(5, 2 * plan.total)).cuda()
    y = torch.empty_like(x)
    plan.execute(x, y, 1)
    plan.check()
    first = y.cpu().numpy()
    for rep in range(4):
        x.mul_(1.0 + rep)           # eager work between replays
        y.fill_(0.0)
        plan.execute(x, y, 1)
        plan.check()
        got = y.cpu().numpy()
        assert rel_l2(got, first * np.prod([1.0 + r for r in range(rep + 1)])) < 1e-12, rep
    # the same work without the graph
    plan.execute(x, y, 1, graph=False)
    plan.check()
    assert rel_l2(y.cpu().numpy(), got) < 1e-15


def test_uneven_in_place_and_info():
    import torch

    items = [(1 << 17,), (256, 128), (64,), (1 << 13,), (1 << 14,)]
    plan = numrs_amd.UnevenPlan("float32", items)
    info = plan.info()
    assert info["tasks_phase1"] > 0 and info["tasks_phase2"] > 0 and info["groups"] > 0
    x = synth.uniform(6, 2 * plan.total).astype(np.float32)
    xt = torch.from_numpy(x).cuda()
    plan.execute(xt, xt, -1)
    plan.check()
    y = xt.cpu().numpy()
    for it, off in zip(items, plan.offsets):
        n = int(np.prod(it))
        seg = slice(2 * off, 2 * (off + n))
        assert rel_l2(y[seg], _ref(x[seg].astype(np.float64), it, -1)) < 2e-6


def test_fft_batch_mixed_lengths_host():
    # FFTProcessor::fft_batch (FFT_1.rs:185-189) with mixed lengths, one launch.
    proc = numrs_amd.FFTProcessor.new()
    bufs = [synth.uniform(40 + i, 2 * (1 << (3 + (i * 5) % 15))) for i in range(20)]
    refs = [oracle.four1_literal(b, -1) for b in bufs]
    proc.fft_batch(bufs, -1)
    for b, r in zip(bufs, refs):
        assert rel_l2(b, r) < 1e-12


def test_gen_uniform_matches_synth():
    import torch

    for dt in (torch.float32, torch.float64):
        x = torch.empty(100003, dtype=dt, device="cuda")
        numrs_amd.gen_uniform(x, 9, start=17)
        torch.cuda.synchronize()
        ref = synth.uniform(9, 100003, start=17)
        if dt == torch.float32:
            ref = ref.astype(np.float32)
        assert np.array_equal(x.cpu().numpy(), ref)


def test_uneven_full_size_range():
    # The size classes of the 10,000-item benchmark batch (1-D 2^8..2^20,
    # 2-D sides 2^6..2^11), 40 items, FP32 against the FP64 oracle.
    import torch

    items = synth.uneven_shapes(4, 40)
    assert max(int(np.prod(it)) for it in items if len(it) == 1) >= 1 << 17
    plan = numrs_amd.UnevenPlan("float32", items)
    x = torch.empty(2 * plan.total, dtype=torch.float32, device="cuda")
    numrs_amd.gen_uniform(x, 4)
    y = torch.empty_like(x)
    plan.execute(x, y, 1)
    plan.check()
    xh, yh = x.cpu().numpy(), y.cpu().numpy()
    for it, off in zip(items, plan.offsets):
        n = int(np.prod(it))
        seg = slice(2 * off, 2 * (off + n))
        assert rel_l2(yh[seg], _ref(xh[seg].astype(np.float64), it, 1)) < 2e-6, it


@pytest.mark.parametrize("dtype,items,tol", [
    ("float32", [(1 << 24,), (8192, 64), (64, 8192), (1 << 13,), (1 << 14,), (1 << 15,)], 3e-6),
    ("float64", [(1 << 22,), (4096, 64), (64, 8192), (1 << 13,), (1 << 14,)], 1e-12),
])
def test_uneven_size_limits(dtype, items, tol):
    # Largest items the queue takes (1-D four-step at the limit, long rows and
    # long columns, the longest single-pass 1-D length and the shortest
    # four-step one).
    import torch

    plan = numrs_amd.UnevenPlan(dtype, items)
    x = torch.empty(2 * plan.total, dtype=getattr(torch, dtype), device="cuda")
    numrs_amd.gen_uniform(x, 8)
    y = torch.empty_like(x)
    plan.execute(x, y, -1)
    plan.check()
    xh, yh = x.cpu().numpy(), y.cpu().numpy()
    for it, off in zip(items, plan.offsets):
        n = int(np.prod(it))
        seg = slice(2 * off, 2 * (off + n))
        assert rel_l2(yh[seg], _ref(xh[seg].astype(np.float64), it, -1)) < tol, it


@pytest.mark.parametrize("dtype,item", [("float32", (1 << 25,)), ("float32", (16384, 64)), ("float32", (64, 16384)),
                                        ("float64", (1 << 23,)), ("float64", (8192, 64)), ("float32", (32,))])
def test_uneven_rejects_out_of_range(dtype, item):
    with pytest.raises(numrs_amd.NrfftError):
        numrs_amd.UnevenPlan(dtype, [item])


def test_uneven_autotune_keeps_results():
    # The autotuner's per-class four-step split is used by later plans, whose
    # results stay those of the oracle.
    import torch

    choice = numrs_amd.autotune_uneven("float64", 14, 17, iters=2)
    assert set(choice) == {14, 15, 16, 17}
    for logn, (l1, ms) in choice.items():
        assert 6 <= l1 <= logn - 6 and ms > 0
    items = [(1 << 14,), (1 << 15,), (1 << 16,), (1 << 17,), (128, 64)]
    plan = numrs_amd.UnevenPlan("float64", items)
    x = torch.empty(2 * plan.total, dtype=torch.float64, device="cuda")
    numrs_amd.gen_uniform(x, 12)
    y = torch.empty_like(x)
    plan.execute(x, y, 1)
    plan.check()
    xh, yh = x.cpu().numpy(), y.cpu().numpy()
    for it, off in zip(items, plan.offsets):
        n = int(np.prod(it))
        seg = slice(2 * off, 2 * (off + n))
        assert rel_l2(yh[seg], _ref(xh[seg], it, 1)) < 1e-12, it
