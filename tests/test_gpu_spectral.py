"""GPU: twofft, cosft1, cosft2 and rlft3 against the oracle.

Tolerances: FP64 1e-12 relative L2; FP32 2e-6 relative L2 against the FP64
oracle run on the FP32-rounded inputs.  Sizes cover the fused single-group
kernels (transform up to 8192 complex points) and the multi-pass path.
"""
import numpy as np
import pytest

import numrs_amd
import oracle
from conftest import rel_l2
from numrs_amd import synth

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-12, np.float32: 2e-6}


def _oracle_bound(dt, oracle_out, exact):
    """NR's cosft1/cosft2 build their twiddles by recurrence, whose drift
    reaches ~1e-12 of the output at n = 2^17 (measured against scipy's DCT).
    The bound against the oracle is the tolerance or three times the
    oracle's own error, whichever is larger."""
    return max(TOL[dt], 3.0 * rel_l2(oracle_out, exact))


def _in(seed, n, dt):
    return synth.normal(seed, n).astype(dt)


# ---- twofft -------------------------------------------------------------------
@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("n", [2, 4, 64, 1024, 8192, 16384, 1 << 17])
def test_twofft_vs_oracle(dt, n):
    d1, d2 = _in(70, n, dt), _in(71, n, dt)
    f1, f2 = np.full(2 * n + 2, 7, dt), np.full(2 * n + 2, 7, dt)
    numrs_amd.twofft(d1, d2, f1, f2)
    r1, r2 = oracle.twofft(d1.astype(np.float64), d2.astype(np.float64))
    assert rel_l2(f1[:2 * n], r1) < TOL[dt] and rel_l2(f2[:2 * n], r2) < TOL[dt], n
    assert f1[2 * n] == 0 and f1[2 * n + 1] == 0 and f2[2 * n] == 0 and f2[2 * n + 1] == 0


def test_twofft_reference_signals():
    # FFT_2.rs:391-401 test signals: sin(2 pi 5 t), cos(2 pi 10 t), n = 256.
    n = 256
    t = np.arange(n) / n
    d1, d2 = np.sin(2 * np.pi * 5 * t), np.cos(2 * np.pi * 10 * t)
    f1, f2 = np.zeros(2 * n + 2), np.zeros(2 * n + 2)
    numrs_amd.twofft(d1, d2, f1, f2)
    assert abs(f1[1]) < 1e-10 and abs(f2[1]) < 1e-10             # FFT_2.rs:414-415
    F1, F2 = f1[0:2 * n:2] + 1j * f1[1:2 * n:2], f2[0:2 * n:2] + 1j * f2[1:2 * n:2]
    e1 = np.zeros(n, complex)
    e1[5], e1[n - 5] = 0.5j * n, -0.5j * n                         # exp(+2 pi i jk/n) convention
    e2 = np.zeros(n, complex)
    e2[10] = e2[n - 10] = 0.5 * n
    assert np.abs(F1 - e1).max() < 1e-10 and np.abs(F2 - e2).max() < 1e-10
    # Hermitian symmetry, bin k against bin n - k (FFT_2.rs:418-425 pairs
    # index 2n + 2 - 2k, NR's 1-based mirror; 0-based it is 2n - 2k)
    k = np.arange(1, n // 2)
    assert np.abs(F1[k] - np.conj(F1[n - k])).max() < 1e-10


def test_twofft_processor_batch():
    proc = numrs_amd.TwoFFTProcessor.new()
    items = []
    for i, n in enumerate((16, 512, 4096)):
        items.append((_in(80 + i, n, np.float64), _in(90 + i, n, np.float64), np.zeros(2 * n + 2), np.zeros(2 * n + 2)))
    proc.process_batch(items)
    for d1, d2, f1, f2 in items:
        r1, r2 = oracle.twofft(d1, d2)
        assert rel_l2(f1[:-2], r1) < 1e-12 and rel_l2(f2[:-2], r2) < 1e-12


# ---- cosft1 -------------------------------------------------------------------
@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("n", [2, 4, 64, 4096, 8192, 16384, 1 << 17])
def test_cosft1_vs_oracle(dt, n):
    import scipy.fft as sfft

    y = _in(72, n + 1, dt)
    ref = oracle.cosft1(y.astype(np.float64))
    exact = 0.5 * sfft.dct(y.astype(np.float64), type=1)
    g = y.copy()
    numrs_amd.cosft1(g, n)
    assert rel_l2(g, ref) < _oracle_bound(dt, ref, exact), n
    if dt == np.float64:
        assert rel_l2(g, exact) < 2e-13, n
    numrs_amd.cosft1(g, n)                     # its own inverse up to n/2
    assert rel_l2(g, 0.5 * n * y.astype(np.float64)) < 10 * TOL[dt], n


@pytest.mark.parametrize("dt,n,b", [(np.float32, 16384, 5), (np.float32, 4096, 3), (np.float64, 8192, 4),
                                    (np.float32, 16384, 600), (np.float64, 8192, 520)])
@pytest.mark.parametrize("inplace", [True, False])
def test_cosft1_fused_batched(dt, n, b, inplace):
    # Batches of the one-group cosft1 whose signals of n + 1 values start at
    # any offset: every signal against the oracle and scipy's DCT-I, in place
    # and out of place.  FP32 16384 and FP64 8192 run persistent (one group
    # per CU, each taking every gridDim-th signal): 600 / 520 signals give
    # groups two or three each.
    import scipy.fft as sfft
    import torch

    y = _in(86, b * (n + 1), dt)
    t = torch.from_numpy(y).cuda()
    o = t if inplace else torch.empty_like(t)
    numrs_amd.Cosft1Plan(np.dtype(dt).name, n, b).execute(t, o)
    got = o.cpu().numpy()
    for i in range(b):
        seg = slice(i * (n + 1), (i + 1) * (n + 1))
        y64 = y[seg].astype(np.float64)
        ref, exact = oracle.cosft1(y64), 0.5 * sfft.dct(y64, type=1)
        assert rel_l2(got[seg], ref) < _oracle_bound(dt, ref, exact), i
        if dt == np.float64:
            assert rel_l2(got[seg], exact) < 2e-13, i


# ---- cosft2 -------------------------------------------------------------------
@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("n", [2, 4, 8, 64, 16384, 32768, 1 << 17])
@pytest.mark.parametrize("isign", [1, -1])
def test_cosft2_vs_oracle(dt, n, isign):
    import scipy.fft as sfft

    y = _in(73, n, dt)
    ref = oracle.cosft2(y.astype(np.float64), isign)
    exact = 0.5 * sfft.dct(y.astype(np.float64), type=2 if isign == 1 else 3)
    g = y.copy()
    numrs_amd.cosft2(g, n, isign)
    assert rel_l2(g, ref) < _oracle_bound(dt, ref, exact), (n, isign)
    if dt == np.float64:
        assert rel_l2(g, exact) < 2e-13, (n, isign)


@pytest.mark.parametrize("n", [16, 1 << 15])
def test_cosft2_round_trip(n):
    # Cos_FT2.rs:247-262 data (sin(i), i = 1..n in its 1-based slots).  The
    # inverse is NR's, scaled by n/2 (the code the reference test calls
    # returns that scale; the test's unscaled expectation is a divergence).
    y = np.sin(np.arange(1, n + 1, dtype=np.float64))
    g = y.copy()
    numrs_amd.cosft2(g, n, 1)
    numrs_amd.cosft2(g, n, -1)
    assert rel_l2(g, 0.5 * n * y) < 1e-12


@pytest.mark.parametrize("dt,n,b", [("float32", 1 << 16, 150), ("float64", 1 << 16, 80), ("float32", 1 << 15, 300)])
def test_cosft1_plan_chunked(dt, n, b):
    # A batched device plan beyond one group runs its four-step C2C passes
    # chunk by chunk; these batches span two chunks, the second partial.
    # Signals of both chunks against the oracle, and the plan's result is its
    # own inverse up to n/2.
    import torch

    npdt = np.float32 if dt == "float32" else np.float64
    tdt = torch.float32 if dt == "float32" else torch.float64
    x = _in(80, (n + 1) * b, npdt)
    plan = numrs_amd.Cosft1Plan(dt, n, b)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    plan.execute(xt, yt)
    y = yt.cpu().numpy().astype(np.float64)
    for i in (0, b // 2, b - 1):
        seg = slice(i * (n + 1), (i + 1) * (n + 1))
        ref = oracle.cosft1(x[seg].astype(np.float64))
        assert rel_l2(y[seg], ref) < 5 * TOL[npdt], i
    zt = torch.empty_like(xt)
    plan.execute(yt, zt)
    assert rel_l2(zt.cpu().numpy().astype(np.float64), 0.5 * n * x.astype(np.float64)) < 20 * TOL[npdt]
    assert zt.dtype == tdt


@pytest.mark.parametrize("dt,n,b", [("float32", 1 << 16, 300), ("float64", 1 << 16, 150), ("float32", 1 << 17, 80)])
@pytest.mark.parametrize("isign", [1, -1])
def test_cosft2_plan_chunked(dt, n, b, isign):
    # As above for cosft2 in both directions.  The forward direction also runs
    # its permute-pack and its untangle-and-rotate step per chunk, through a
    # chunk-sized scratch; the inverse runs those two over the whole batch.
    import torch

    npdt = np.float32 if dt == "float32" else np.float64
    x = _in(81, n * b, npdt)
    plan = numrs_amd.Cosft2Plan(dt, n, b)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    plan.execute(xt, yt, isign)
    y = yt.cpu().numpy().astype(np.float64)
    for i in (0, b // 2, b - 1):
        seg = slice(i * n, (i + 1) * n)
        ref = oracle.cosft2(x[seg].astype(np.float64), isign)
        assert rel_l2(y[seg], ref) < 5 * TOL[npdt], (i, isign)


@pytest.mark.parametrize("dt,n,b", [("float64", 1 << 20, 3), ("float32", 1 << 20, 5), ("float32", 1 << 22, 2),
                                    ("float32", 1 << 23, 2)])
def test_cosft1_plan_long(dt, n, b):
    # Long DCT-Is whose first C2C pass reads the even extension of each
    # signal straight from its n + 1 reals (signals at odd real offsets):
    # 512- and 2048-point first passes, 8-wide tiles; at 2^23 the first of
    # three passes.
    import scipy.fft as sfft
    import torch

    npdt = np.float32 if dt == "float32" else np.float64
    x = _in(82, (n + 1) * b, npdt)
    plan = numrs_amd.Cosft1Plan(dt, n, b)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    plan.execute(xt, yt)
    y = yt.cpu().numpy().astype(np.float64)
    for i in range(b):
        seg = slice(i * (n + 1), (i + 1) * (n + 1))
        ref = oracle.cosft1(x[seg].astype(np.float64))
        exact = 0.5 * sfft.dct(x[seg].astype(np.float64), type=1)
        assert rel_l2(y[seg], ref) < _oracle_bound(npdt, ref, exact), i
        if npdt == np.float64:
            assert rel_l2(y[seg], exact) < 2e-13, i


@pytest.mark.parametrize("dt,n,b", [("float64", 1 << 21, 3), ("float64", 1 << 19, 3), ("float32", 1 << 21, 5),
                                    ("float32", 1 << 24, 2)])
@pytest.mark.parametrize("isign", [1, -1])
def test_cosft2_plan_long(dt, n, b, isign):
    # As above for cosft2: the forward first pass reads the even samples
    # ascending and the odd ones descending straight from the signal (FP64
    # 512-point first passes 8 wide at 2^21, 256-point 16 wide at 2^19).
    import scipy.fft as sfft
    import torch

    npdt = np.float32 if dt == "float32" else np.float64
    x = _in(83, n * b, npdt)
    plan = numrs_amd.Cosft2Plan(dt, n, b)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    plan.execute(xt, yt, isign)
    y = yt.cpu().numpy().astype(np.float64)
    for i in range(b):
        seg = slice(i * n, (i + 1) * n)
        ref = oracle.cosft2(x[seg].astype(np.float64), isign)
        exact = 0.5 * sfft.dct(x[seg].astype(np.float64), type=2 if isign == 1 else 3)
        assert rel_l2(y[seg], ref) < _oracle_bound(npdt, ref, exact), (i, isign)
        if npdt == np.float64:
            assert rel_l2(y[seg], exact) < 2e-13, (i, isign)


def test_cosft2_invalid_isign():
    with pytest.raises(ValueError, match="Invalid isign value"):   # Cos_FT2.rs:265-271
        numrs_amd.cosft2(np.zeros(8), 8, 0)


# ---- rlft3 --------------------------------------------------------------------
# From nn2 >= 2 and 32 <= nn3 <= 4096 the forward fuses the untangle into
# the last axis (rows of nn3/2 points with their mirror rows; nn1 = 1 is the
# 2-D case).
SHAPES = [(4, 8, 16), (1, 8, 16), (8, 8, 8), (2, 2, 4), (1, 1, 8), (4, 4, 2), (16, 32, 64), (8, 64, 256),
          (1, 1, 1 << 15), (2, 2, 32), (32, 8, 128), (4, 16, 4096), (2, 256, 64), (1, 16, 64), (1, 256, 512)]


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("shape", SHAPES)
def test_rlft3_vs_oracle(dt, shape):
    nn1, nn2, nn3 = shape
    x = _in(74, nn1 * nn2 * nn3, dt)
    d, s = x.reshape(shape).copy(), np.zeros((nn1, 2 * nn2), dt)
    numrs_amd.rlft3(d, s, nn1, nn2, nn3, 1)
    rd, rs = oracle.rlft3(x.astype(np.float64), np.zeros(2 * nn1 * nn2), nn1, nn2, nn3, 1)
    assert rel_l2(d.reshape(-1), rd) < TOL[dt] and rel_l2(s.reshape(-1), rs) < TOL[dt], shape
    numrs_amd.rlft3(d, s, nn1, nn2, nn3, -1)
    assert rel_l2(d.reshape(-1), 0.5 * nn1 * nn2 * nn3 * x.astype(np.float64)) < 10 * TOL[dt], shape


def test_rlft3_reference_pattern():
    # Real_FT3.rs:267-311 data (i + j + k), 8 x 8 x 8; the inverse returns
    # (nn1 nn2 nn3 / 2) x (NR), the reference test divides by nn1 nn2 nn3.
    shape = (8, 8, 8)
    i, j, k = np.meshgrid(*[np.arange(8.0)] * 3, indexing="ij")
    x = i + j + k
    d, s = x.copy(), np.zeros((8, 16))
    numrs_amd.rlft3(d, s, 8, 8, 8, 1)
    assert abs(d[0, 0, 0] - x.sum()) < 1e-9 and abs(d[0, 0, 1]) < 1e-9      # F[0,0,0] = sum, real
    numrs_amd.rlft3(d, s, 8, 8, 8, -1)
    assert np.abs(d / 256.0 - x).max() < 1e-10


def test_rlft3_asserts():
    with pytest.raises(AssertionError):
        numrs_amd.rlft3(np.zeros((4, 4, 8)), np.zeros((4, 4)), 4, 4, 8, 1)


# ---- device plans, batched ----------------------------------------------------
def test_device_plans_batched():
    import torch

    b = 3
    n = 2048
    d1, d2 = _in(75, b * n, np.float64), _in(76, b * n, np.float64)
    p = numrs_amd.TwofftPlan("float64", n, b)
    f1, f2 = torch.empty(2 * b * n, dtype=torch.float64, device="cuda"), torch.empty(2 * b * n, dtype=torch.float64,
                                                                                      device="cuda")
    p.execute(torch.from_numpy(d1).cuda(), torch.from_numpy(d2).cuda(), f1, f2)
    g1, g2 = f1.cpu().numpy(), f2.cpu().numpy()
    for i in range(b):
        r1, r2 = oracle.twofft(d1[i * n:(i + 1) * n], d2[i * n:(i + 1) * n])
        assert rel_l2(g1[2 * i * n:2 * (i + 1) * n], r1) < 1e-12 and rel_l2(g2[2 * i * n:2 * (i + 1) * n], r2) < 1e-12

    import scipy.fft as sfft

    for n, b in ((1024, 5), (1 << 15, 2)):
        y = _in(77, b * (n + 1), np.float64)
        t = torch.from_numpy(y).cuda()
        numrs_amd.Cosft1Plan("float64", n, b).execute(t, t)   # in place
        got = t.cpu().numpy()
        for i in range(b):
            seg = slice(i * (n + 1), (i + 1) * (n + 1))
            ref, exact = oracle.cosft1(y[seg]), 0.5 * sfft.dct(y[seg], type=1)
            assert rel_l2(got[seg], ref) < _oracle_bound(np.float64, ref, exact)
            assert rel_l2(got[seg], exact) < 2e-13

    for n, b in ((256, 4), (1 << 16, 2)):
        y = _in(78, b * n, np.float64)
        plan = numrs_amd.Cosft2Plan("float64", n, b)
        for isign in (1, -1):
            t = torch.from_numpy(y).cuda()
            o = torch.empty_like(t)
            plan.execute(t, o, isign)
            got = o.cpu().numpy()
            for i in range(b):
                seg = slice(i * n, (i + 1) * n)
                ref = oracle.cosft2(y[seg], isign)
                exact = 0.5 * sfft.dct(y[seg], type=2 if isign == 1 else 3)
                assert rel_l2(got[seg], ref) < _oracle_bound(np.float64, ref, exact)
                assert rel_l2(got[seg], exact) < 2e-13

    shape, b = (4, 8, 32), 2
    vol = int(np.prod(shape))
    x = _in(79, b * vol, np.float64)
    t = torch.from_numpy(x).cuda()
    sp = torch.zeros(b * 2 * shape[0] * shape[1], dtype=torch.float64, device="cuda")
    plan = numrs_amd.Rlft3Plan("float64", *shape, b)
    plan.execute(t, sp, 1)
    gd, gs = t.cpu().numpy(), sp.cpu().numpy()
    ns = 2 * shape[0] * shape[1]
    for i in range(b):
        rd, rs = oracle.rlft3(x[i * vol:(i + 1) * vol], np.zeros(ns), *shape, 1)
        assert rel_l2(gd[i * vol:(i + 1) * vol], rd) < 1e-12 and rel_l2(gs[i * ns:(i + 1) * ns], rs) < 1e-12
    plan.execute(t, sp, -1)
    assert rel_l2(t.cpu().numpy(), 0.5 * vol * x) < 1e-12


def test_device_plans_multipass_chunked_f32():
    # FP32 at n = 2^20, 36 transforms: the inner C2C runs the tuned four-step
    # in chunks of 16 transforms (16, 16, 4).  Transforms of the first, a
    # middle and the tail chunk against FP64 numpy / scipy.
    import scipy.fft as sfft
    import torch

    n, b, pick = 1 << 20, 36, (0, 17, 35)
    tol = 2e-6
    d1, d2 = _in(80, b * n, np.float32), _in(81, b * n, np.float32)
    f1 = torch.empty(2 * b * n, dtype=torch.float32, device="cuda")
    f2 = torch.empty_like(f1)
    numrs_amd.TwofftPlan("float32", n, b).execute(torch.from_numpy(d1).cuda(), torch.from_numpy(d2).cuda(), f1, f2)
    g1, g2 = f1.cpu().numpy(), f2.cpu().numpy()
    for i in pick:
        for g, d in ((g1, d1), (g2, d2)):
            z = n * np.fft.ifft(d[i * n:(i + 1) * n].astype(np.float64))   # sum x_j exp(+2 pi i jk / n)
            ref = np.stack([z.real, z.imag], axis=1).ravel()
            assert rel_l2(g[2 * i * n:2 * (i + 1) * n], ref) < tol, ("twofft", i)
    del f1, f2

    y = _in(82, b * (n + 1), np.float32)
    t = torch.from_numpy(y).cuda()
    numrs_amd.Cosft1Plan("float32", n, b).execute(t, t)
    got = t.cpu().numpy()
    for i in pick:
        seg = slice(i * (n + 1), (i + 1) * (n + 1))
        assert rel_l2(got[seg], 0.5 * sfft.dct(y[seg].astype(np.float64), type=1)) < tol, ("cosft1", i)

    y = _in(83, b * n, np.float32)
    plan = numrs_amd.Cosft2Plan("float32", n, b)
    for isign in (1, -1):
        t = torch.from_numpy(y).cuda()
        o = torch.empty_like(t)
        plan.execute(t, o, isign)
        got = o.cpu().numpy()
        for i in pick:
            seg = slice(i * n, (i + 1) * n)
            exact = 0.5 * sfft.dct(y[seg].astype(np.float64), type=2 if isign == 1 else 3)
            assert rel_l2(got[seg], exact) < tol, ("cosft2", isign, i)


def test_twofft_chunked_f64():
    # FP64 n = 2^16, 130 signals: each array's 2^15-point C2C runs in chunks
    # of 128 transforms (128, 2) before the whole-batch untangle.  Signals of
    # both chunks against the oracle.
    import torch

    n, b = 1 << 16, 130
    d1, d2 = _in(84, b * n, np.float64), _in(85, b * n, np.float64)
    f1 = torch.empty(2 * b * n, dtype=torch.float64, device="cuda")
    f2 = torch.empty_like(f1)
    numrs_amd.TwofftPlan("float64", n, b).execute(torch.from_numpy(d1).cuda(), torch.from_numpy(d2).cuda(), f1, f2)
    g1, g2 = f1.cpu().numpy(), f2.cpu().numpy()
    for i in (0, 63, 64, 100, 128, 129):
        r1, r2 = oracle.twofft(d1[i * n:(i + 1) * n], d2[i * n:(i + 1) * n])
        seg = slice(2 * i * n, 2 * (i + 1) * n)
        assert rel_l2(g1[seg], r1) < 1e-12 and rel_l2(g2[seg], r2) < 1e-12, i


@pytest.mark.parametrize("dt,shape,b", [(np.float32, (8, 4, 64), 3), (np.float64, (2, 16, 256), 2),
                                        (np.float32, (16, 2, 4096), 2), (np.float64, (1, 8, 128), 3)])
def test_rlft3_plan_fused_batched(dt, shape, b):
    # Batched device plans on the fused forward (row pairs enumerated per
    # volume, nn1 > nn2 and nn1 < nn2, the longest fused rows, 2-D): every
    # volume's data and speq against the oracle, then the inverse round trip.
    import torch

    nn1, nn2, nn3 = shape
    vol, ns = nn1 * nn2 * nn3, 2 * nn1 * nn2
    x = _in(87, b * vol, dt)
    t = torch.from_numpy(x).cuda()
    sp = torch.zeros(b * ns, dtype=t.dtype, device="cuda")
    plan = numrs_amd.Rlft3Plan(np.dtype(dt).name, nn1, nn2, nn3, b)
    plan.execute(t, sp, 1)
    gd, gs = t.cpu().numpy(), sp.cpu().numpy()
    for i in range(b):
        rd, rs = oracle.rlft3(x[i * vol:(i + 1) * vol].astype(np.float64), np.zeros(ns), nn1, nn2, nn3, 1)
        assert rel_l2(gd[i * vol:(i + 1) * vol], rd) < TOL[dt], ("data", i)
        assert rel_l2(gs[i * ns:(i + 1) * ns], rs) < TOL[dt], ("speq", i)
    plan.execute(t, sp, -1)
    assert rel_l2(t.cpu().numpy(), 0.5 * vol * x.astype(np.float64)) < 10 * TOL[dt]
