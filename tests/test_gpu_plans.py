"""GPU tests of the plan machinery: the batched RealFTProcessor, argument
checks of the device plans, the C2C autotuner and its cache, explicit
per-plan tunings, four-step defaults whose last batch chunk is partial, and
the chunked four-step along long strided fourn axes.  Results against the
oracle (FP64 <= 1e-12, FP32 <= 2e-6 relative L2)."""
import numpy as np
import pytest

import numrs_amd
import oracle
from conftest import rel_l2
from numrs_amd import synth

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 2e-6


def f32round(x):
    return np.asarray(x, np.float64).astype(np.float32).astype(np.float64)


# ---- RealFTProcessor (Real_FT.rs:331-370) -------------------------------------
def test_realft_processor_batch():
    # process_batch over (data, n, isign) items: groups of equal n and isign
    # (FP64 forward 2^12 x 5, FP64 inverse 2^12 x 2, FP64 forward 2^15 x 3,
    # FP32 forward 2^10 x 4, one single 2^6 item) each run as one batched
    # execution; items longer than n keep their tail.
    proc = numrs_amd.RealFTProcessor.new().with_threshold(512)
    specs = ([(1 << 12, 1, np.float64)] * 5 + [(1 << 12, -1, np.float64)] * 2 + [(1 << 15, 1, np.float64)] * 3 +
             [(1 << 10, 1, np.float32)] * 4 + [(1 << 6, 1, np.float64)])
    items, refs = [], []
    for k, (n, isign, dt) in enumerate(specs):
        x = synth.normal(900 + k, n + 3)
        if dt == np.float32:
            x = f32round(x)
        buf = x.astype(dt)
        items.append((buf, n, isign))
        refs.append((oracle.realft(x[:n], isign), x[n:]))
    proc.process_batch(items)
    for (buf, n, isign), (ref, tail) in zip(items, refs):
        tol = TOL32 * 5 if buf.dtype == np.float32 else TOL64
        assert rel_l2(buf[:n].astype(np.float64), ref) < tol, (n, isign, buf.dtype)
        assert np.array_equal(buf[n:].astype(np.float64), tail.astype(buf.dtype).astype(np.float64))
    # process() and the variants are the same transform
    x = synth.normal(950, 2048)
    for fn in (proc.process, lambda d, n, s: numrs_amd.realft_sequential(d, n, s)):
        y = x.copy()
        fn(y, 2048, 1)
        assert rel_l2(y, oracle.realft(x, 1)) < TOL64


def test_realft_processor_batch_errors():
    proc = numrs_amd.RealFTProcessor.new()
    with pytest.raises(AssertionError):
        proc.process_batch([(np.zeros(8), 7, 1)])
    with pytest.raises(AssertionError):
        proc.process_batch([(np.zeros(8), 16, 1)])


# ---- device-plan argument checks ----------------------------------------------
def test_device_plan_buffer_checks():
    # A buffer of the wrong dtype, too short, or not contiguous is refused
    # before anything is launched; the right buffers then run normally.
    import torch

    n, b = 1 << 12, 3
    p = numrs_amd.C2CPlan("float64", [n], b)
    good = torch.zeros(2 * n * b, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="dtype"):
        p.execute(good.float(), good, 1)
    with pytest.raises(ValueError, match="values"):
        p.execute(good[:-2], good, 1)
    with pytest.raises(ValueError, match="contiguous"):
        p.execute(good.view(-1, 2).t(), good, 1)
    p.execute(good, good, 1)
    r = numrs_amd.RealFTPlan("float32", 256, 4)
    with pytest.raises(ValueError, match="values"):
        r.execute(torch.zeros(256 * 4 - 1, dtype=torch.float32, device="cuda"), 1)
    c = numrs_amd.ConvlvPlan("float32", 1024, 33, 2)
    d = torch.zeros(2048, dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError, match="respns"):
        c.execute(d, torch.zeros(33, dtype=torch.float32, device="cuda"), d, 1)
    t = numrs_amd.TwofftPlan("float32", 64, 2)
    with pytest.raises(ValueError, match="fft2"):
        t.execute(d[:128], d[:128], d[:256], d[:255])
    torch.cuda.synchronize()


# ---- autotuner and tune cache ----------------------------------------------------
def test_autotune_c2c_cache_roundtrip(tmp_path):
    # nrfft_autotune_c2c on a small shape (2^14 x 6, FP32): the tuning it
    # returns is what later plans of that key are built with, their results
    # match the oracle, and nrfft_tune_cache_save / load round-trip it.
    import torch

    n, b = 1 << 14, 6
    best, ms = numrs_amd.autotune_c2c("float32", [n], b, iters=2)
    assert ms > 0 and 1 <= best[0] <= 13
    plan = numrs_amd.C2CPlan("float32", [n], b)
    assert plan.tune == best
    x = f32round(synth.uniform(71, 2 * n * b))
    xt = torch.from_numpy(x.astype(np.float32)).cuda()
    yt = torch.empty_like(xt)
    plan.execute(xt, yt, -1)
    y = yt.cpu().numpy().astype(np.float64)
    for i in range(b):
        seg = slice(2 * n * i, 2 * n * (i + 1))
        assert rel_l2(y[seg], oracle.four1_nr(x[seg], -1)) < TOL32, i
    path = str(tmp_path / "tune.txt")
    numrs_amd.tune_cache_save(path)
    lines = [ln.split() for ln in open(path) if not ln.startswith("#")]
    key = ["0", "0", str(b), "1", str(n), ":"]
    mine = [ln for ln in lines if ln[:6] == key]
    assert mine and tuple(int(v) for v in mine[0][6:]) == best
    # a different tuning for the same key, loaded from a file, is used next
    other = (best[0] + 1 if best[0] < 9 else best[0] - 1, 2, 2, 2, -1, 0)
    with open(path, "w") as f:
        f.write("# edited\n" + " ".join(key) + " " + " ".join(str(v) for v in other) + "\n")
    numrs_amd.tune_cache_load(path)
    p2 = numrs_amd.C2CPlan("float32", [n], b)
    assert p2.tune == other and p2.num_launches == 2 * 3
    p2.execute(xt, yt, -1)
    y2 = yt.cpu().numpy().astype(np.float64)
    for i in (0, b - 1):
        seg = slice(2 * n * i, 2 * n * (i + 1))
        assert rel_l2(y2[seg], oracle.four1_nr(x[seg], -1)) < TOL32, i
    # restore the autotuned entry
    with open(path, "w") as f:
        f.write(" ".join(key) + " " + " ".join(str(v) for v in best) + "\n")
    numrs_amd.tune_cache_load(path)


def test_explicit_tune_is_per_plan():
    # C2CPlan(tune=...) applies to that plan only: a later plan of the same
    # key takes the default (or cached) tuning, split or not.
    import torch

    n, b = 1 << 16, 5
    base = numrs_amd.C2CPlan("float32", [n], b).tune
    tuned = numrs_amd.C2CPlan("float32", [n], b, tune=(6, 2, 3, 2))
    assert tuned.tune[:4] == (6, 2, 3, 2) and tuned.num_launches == 6
    assert numrs_amd.C2CPlan("float32", [n], b).tune == base
    split = numrs_amd.C2CPlan("float32", [n], b, tune=(8, 4, 4, 0), split="bf16")
    assert split.tune[:4] == (8, 4, 4, 0)
    x = f32round(synth.uniform(72, 2 * n * b))
    xt = torch.from_numpy(x.astype(np.float32)).cuda()
    for p, tol in ((tuned, TOL32), (split, 1e-6)):
        yt = torch.empty_like(xt)
        p.execute(xt, yt, 1)
        y = yt.cpu().numpy().astype(np.float64)
        for i in (0, 2, b - 1):
            seg = slice(2 * n * i, 2 * n * (i + 1))
            assert rel_l2(y[seg], oracle.four1_nr(x[seg], 1)) < tol, i


# ---- four-step defaults with a partial last chunk -------------------------------
@pytest.mark.parametrize("dtype,logn,batch", [("float32", 14, 300), ("float32", 15, 300), ("float64", 14, 600),
                                              ("float64", 16, 150), ("float64", 18, 40)])
def test_c2c_default_tuning_partial_chunk(dtype, logn, batch):
    # Default tunings whose batch chunk (32 / 64 / 128 MiB) leaves a short
    # last chunk: the transforms at the first chunk boundary and the last
    # two against the oracle.
    import torch

    n = 1 << logn
    x = synth.uniform(logn + 77, 2 * n * batch)
    if dtype == "float32":
        x = f32round(x)
    xt = torch.from_numpy(x.astype(np.float32 if dtype == "float32" else np.float64)).cuda()
    yt = torch.empty_like(xt)
    plan = numrs_amd.C2CPlan(dtype, [n], batch)
    chunk = plan.tune[3]
    assert 0 < chunk < batch and batch % chunk != 0, plan.tune
    plan.execute(xt, yt, 1)
    y = yt.cpu().numpy().astype(np.float64)
    tol = TOL32 if dtype == "float32" else TOL64
    for i in sorted({0, chunk - 1, chunk, batch - 2, batch - 1}):
        seg = slice(2 * n * i, 2 * n * (i + 1))
        assert rel_l2(y[seg], oracle.four1_nr(x[seg], 1)) < tol, (i, chunk)


# ---- chunked four-step along long strided axes ----------------------------------
@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_fourn_strided_four_step_chunks(dtype):
    # Batched 2-D plans whose first axis (4096 / 8192 points, strided) runs
    # the four-step in batch chunks of at most 128 MiB of workspace: several
    # chunks, a partial last one; first, boundary and last transforms
    # against the oracle, both signs.
    import torch

    shape = [4096, 512] if dtype == "float32" else [8192, 256]
    es = 8 if dtype == "float32" else 16
    per = shape[0] * shape[1] * es
    chunk = max(1, (128 << 20) // per)
    batch = 2 * chunk + 3
    plan = numrs_amd.C2CPlan(dtype, shape, batch)
    assert plan.workspace_bytes == chunk * per
    vol = shape[0] * shape[1]
    x = synth.normal(91, 2 * vol * batch)
    if dtype == "float32":
        x = f32round(x)
    xt = torch.from_numpy(x.astype(np.float32 if dtype == "float32" else np.float64)).cuda()
    tol = TOL32 if dtype == "float32" else TOL64
    for isign in (1, -1):
        yt = torch.empty_like(xt)
        plan.execute(xt, yt, isign)
        y = yt.cpu().numpy().astype(np.float64)
        for i in sorted({0, chunk - 1, chunk, batch - 1}):
            seg = slice(2 * vol * i, 2 * vol * (i + 1))
            assert rel_l2(y[seg], oracle.fourn(x[seg], shape, isign)) < tol, (i, isign)


@pytest.mark.parametrize("shape", [(4096, 8), (8192, 4)])
def test_fourn_narrow_strided_axis_one_pass(shape):
    # Long strided axes with fewer than 16 columns side by side run as one
    # column pass (no workspace).
    plan = numrs_amd.C2CPlan("float32", list(shape), 2)
    assert plan.workspace_bytes == 0
    x = f32round(synth.normal(93, 2 * int(np.prod(shape))))
    for isign in (1, -1):
        y = x.astype(np.float32)
        numrs_amd.fourn(y, list(shape), 2, isign)
        assert rel_l2(y.astype(np.float64), oracle.fourn(x, list(shape), isign)) < TOL32, isign

