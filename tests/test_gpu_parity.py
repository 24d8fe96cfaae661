"""GPU parity: the HIP kernels (through the C ABI) against the oracle.

Tolerances (relative L2, stated per test):
  FP64 path vs the oracle: 1e-12 (the required bar for four1).
  FP32 path vs the FP64 oracle on the FP32-rounded inputs: 2e-6 for the
  C2C transforms (radix-16 Stockham, ~1e-7 per stage), 1e-5 for realft
  packed spectra and convolutions of N(0,1) data.
"""
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


# ---- four1 -------------------------------------------------------------------
@pytest.mark.parametrize("logn", list(range(0, 17)))
def test_four1_f64_all_lengths(logn):
    nn = 1 << logn
    x = synth.uniform(logn, 2 * nn)
    for isign in (1, -1):
        y = x.copy()
        numrs_amd.four1(y, nn, isign)
        assert rel_l2(y, oracle.four1_literal(x, isign)) < TOL64, (nn, isign)


def test_four1_2p16_roundtrip():
    # N = 2^16 FP64, U[-1,1) seed 0: forward (+1) then inverse (-1) * 1/N.
    nn = 1 << 16
    x = synth.uniform(0, 2 * nn)
    fwd = x.copy()
    numrs_amd.four1(fwd, nn, 1)
    ref = oracle.four1_literal(x, 1)
    assert rel_l2(fwd, ref) < TOL64
    back = fwd.copy()
    numrs_amd.four1(back, nn, -1)
    back /= nn
    assert rel_l2(back, x) < TOL64
    ref_back = oracle.four1_literal(ref, -1) / nn
    assert rel_l2(back, ref_back) < TOL64


@pytest.mark.parametrize("logn", [17, 18, 20, 22])
def test_four1_f64_large(logn):
    nn = 1 << logn
    x = synth.uniform(100 + logn, 2 * nn)
    y = x.copy()
    numrs_amd.four1(y, nn, -1)
    assert rel_l2(y, oracle.four1_nr(x, -1)) < TOL64


@pytest.mark.parametrize("logn", [1, 4, 7, 10, 13, 14, 16, 20])
def test_four1_f32(logn):
    nn = 1 << logn
    x = f32round(synth.uniform(200 + logn, 2 * nn))
    for isign in (1, -1):
        y = x.astype(np.float32)
        numrs_amd.four1(y, nn, isign)
        assert rel_l2(y, oracle.four1_nr(x, isign)) < TOL32, (nn, isign)


def test_four1_reference_test_signal():
    # FFT_1.rs:245-267 on the GPU path.
    t = np.arange(1024) / 1024.0
    sig = np.sin(2 * np.pi * 5 * t) + 0.5 * np.cos(2 * np.pi * 20 * t)
    z = numrs_amd.real_to_complex(sig)
    numrs_amd.four1(z, 1024, 1)
    mag = numrs_amd.magnitude_spectrum(z)
    assert np.allclose(mag * mag, numrs_amd.power_spectrum(z), atol=1e-10)  # FFT_1.rs:304-321
    numrs_amd.four1(z, 1024, -1)
    z /= 1024.0
    assert np.max(np.abs(numrs_amd.complex_to_real(z) - sig)) < 1e-10


def test_fft_processor_batch_uneven():
    # FFTProcessor::fft_batch (FFT_1.rs:185-189) with mixed lengths.
    proc = numrs_amd.FFTProcessor.new()
    bufs = [synth.uniform(300 + i, 2 * (1 << (4 + (i % 5)))) for i in range(12)]
    refs = [oracle.four1_literal(b, 1) for b in bufs]
    proc.fft_batch(bufs, 1)
    for b, r in zip(bufs, refs):
        assert rel_l2(b, r) < TOL64


def test_four1_batch_f32_2pow20():
    # Batched C2C at N = 2^20, FP32 (two-pass four-step), 3 transforms.
    nn = 1 << 20
    x = f32round(synth.uniform(3, 2 * nn * 3)).astype(np.float32)
    y = x.copy()
    proc = numrs_amd.FFTProcessor.new()
    bufs = [y[i * 2 * nn:(i + 1) * 2 * nn].copy() for i in range(3)]
    proc.fft_batch(bufs, -1)
    for i in range(3):
        ref = oracle.four1_nr(x[i * 2 * nn:(i + 1) * 2 * nn].astype(np.float64), -1)
        assert rel_l2(bufs[i], ref) < TOL32


@pytest.mark.parametrize("dtype,logn,batch", [(np.float32, 16, 256), (np.float64, 15, 200)])
def test_four1_batch_host_pipelined(dtype, logn, batch):
    # nrfft_four1_batch_* on host arrays large enough to be streamed through
    # HBM in 32 MB chunks (4 chunks; the FP64 case ends with a short one):
    # transforms on both sides of every chunk boundary against the oracle.
    import ctypes

    from numrs_amd import _lib

    n = 1 << logn
    x = synth.uniform(500 + logn, 2 * n * batch)
    if dtype == np.float32:
        x = f32round(x)
    y = x.astype(dtype)
    ptr = y.ctypes.data_as(ctypes.POINTER(ctypes.c_float if dtype == np.float32 else ctypes.c_double))
    fn = _lib.lib().nrfft_four1_batch_f32 if dtype == np.float32 else _lib.lib().nrfft_four1_batch_f64
    assert fn(ptr, n, batch, -1) == 0, _lib.last_error()
    chunk = (32 << 20) // (n * 2 * np.dtype(dtype).itemsize)
    for i in sorted({0, chunk - 1, chunk, 2 * chunk + 1, batch - 2, batch - 1}):
        ref = oracle.four1_nr(x[i * 2 * n:(i + 1) * 2 * n], -1)
        got = y[i * 2 * n:(i + 1) * 2 * n].astype(np.float64)
        assert rel_l2(got, ref) < (TOL32 if dtype == np.float32 else TOL64), (i, chunk)


def test_four1_batch_host_pipelined_threads():
    # Host-pointer batch calls from two threads at once (ctypes releases the
    # GIL): both stream through the library's shared pinned staging and copy
    # pool, which serialise the pipelined calls; each result against the oracle.
    import ctypes
    import threading

    from numrs_amd import _lib

    n, batch = 1 << 16, 160
    xs = [f32round(synth.uniform(600 + k, 2 * n * batch)) for k in range(2)]
    ys = [x.astype(np.float32) for x in xs]
    rcs = [None, None]

    def run(k):
        ptr = ys[k].ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        rcs[k] = _lib.lib().nrfft_four1_batch_f32(ptr, n, batch, 1 if k == 0 else -1)

    ts = [threading.Thread(target=run, args=(k,)) for k in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert rcs == [0, 0], _lib.last_error()
    for k in range(2):
        for i in (0, 63, 64, batch - 1):
            ref = oracle.four1_nr(xs[k][i * 2 * n:(i + 1) * 2 * n], 1 if k == 0 else -1)
            assert rel_l2(ys[k][i * 2 * n:(i + 1) * 2 * n].astype(np.float64), ref) < TOL32, (k, i)


# ---- realft ------------------------------------------------------------------
@pytest.mark.parametrize("logn", list(range(1, 16)) + [16, 17, 18, 19, 20, 21, 22, 23])
def test_realft_f64(logn):
    n = 1 << logn
    x = synth.normal(logn, n)
    y = x.copy()
    numrs_amd.realft(y, n, 1)
    assert rel_l2(y, oracle.realft(x, 1)) < TOL64, n
    numrs_amd.realft(y, n, -1)
    assert rel_l2(y, (n / 2) * x) < TOL64, n


def test_realft_golden(golden):
    for n in (16, 256, 2048):
        y = golden[f"realft_{n}_in"].copy()
        numrs_amd.realft(y, n, 1)
        assert rel_l2(y, golden[f"realft_{n}_out"]) < TOL64


@pytest.mark.parametrize("logn", [3, 10, 14, 15, 16, 17, 18, 20, 22, 24])
def test_realft_f32(logn):
    n = 1 << logn
    x = f32round(synth.normal(logn, n))
    y = x.astype(np.float32)
    numrs_amd.realft(y, n, 1)
    assert rel_l2(y, oracle.realft(x, 1)) < 1e-5
    numrs_amd.realft(y, n, -1)
    assert rel_l2(y, (n / 2) * x) < 1e-5


@pytest.mark.parametrize("logn,batch", [(14, 300), (15, 300), (16, 300), (17, 150), (18, 70), (19, 40), (21, 10),
                                        (22, 6)])
def test_c2c_f32_default_tuning(logn, batch):
    # The FP32 defaults: the per-length split table (2^14 .. 2^18), 2048-point
    # pass-B rows from 2^19, three passes from 2^22; 32 / 64 / 128 MiB chunks
    # (256 transforms at 2^14 and 2^15), each batch here ending in a partial
    # chunk.  First and last transform against numpy (isign = +1 is N ifft).
    import torch

    n = 1 << logn
    x = synth.uniform(logn, 2 * n * batch).astype(np.float32)
    t = torch.from_numpy(x).cuda()
    o = torch.empty_like(t)
    numrs_amd.C2CPlan("float32", [n], batch).execute(t, o, 1)
    got = o.cpu().numpy()
    for i in (0, batch - 1):
        xc = x[2 * i * n:2 * (i + 1) * n].astype(np.float64).view(np.complex128)
        z = n * np.fft.ifft(xc)
        assert rel_l2(got[2 * i * n:2 * (i + 1) * n], z.view(np.float64)) < TOL32, (logn, i)


@pytest.mark.parametrize("logn,batch", [(14, 600), (16, 150), (17, 40), (18, 40), (19, 20), (21, 6)])
def test_c2c_f64_default_tuning(logn, batch):
    # FP64 four-step defaults (8-wide pass-A tiles from 512-point columns;
    # 32 MiB chunks at 2^14, 64 MiB to 2^17, 128 MiB above) with a partial
    # last chunk; first and last transform against numpy.
    import torch

    n = 1 << logn
    x = synth.uniform(logn + 40, 2 * n * batch)
    t = torch.from_numpy(x).cuda()
    o = torch.empty_like(t)
    numrs_amd.C2CPlan("float64", [n], batch).execute(t, o, 1)
    got = o.cpu().numpy()
    for i in (0, batch - 1):
        z = n * np.fft.ifft(x[2 * i * n:2 * (i + 1) * n].view(np.complex128))
        assert rel_l2(got[2 * i * n:2 * (i + 1) * n], z.view(np.float64)) < TOL64, (logn, i)


@pytest.mark.parametrize("n,b,isign", [(1 << 16, 6, 1), (1 << 16, 6, -1), (1 << 21, 20, 1), (1 << 24, 2, 1)])
def test_convlv_batch_f32_multipass(n, b, isign):
    # FP32 convlv beyond one group: response placement, forward C2Cs (in
    # 16-transform chunks at n = 2^21) + untangles, the spectrum operation
    # fused with the inverse untangle, inverse C2C; first and last signal
    # against the oracle.
    m = 257
    d = f32round(synth.normal(n + 5, n * b)).reshape(b, n)
    if isign == 1:
        x = np.arange(m) - (m - 1) / 2
        r = np.stack([f32round(np.exp(-0.5 * (x / 40.0) ** 2) * (1 + 0.1 * synth.normal(60 + i, m)))
                      for i in range(b)])
    else:
        r = np.stack([f32round(0.3 / m * synth.normal(60 + i, m)) for i in range(b)])
        r[:, (m - 1) // 2] = 1.0
    y = numrs_amd.convlv_batch_responses(d.astype(np.float32), r.astype(np.float32), isign)
    for i in (0, b - 1):
        ref = oracle.convlv(d[i], r[i], isign)
        assert rel_l2(y[i], ref) < 1e-5, (n, isign, i)


@pytest.mark.parametrize("m,isign", [(257, 1), (257, -1), (1001, 1)])
def test_convlv_batch_f32_fused_2p15(m, isign):
    # FP32 signals of 2^15 in one group each (both 16384-point spectra and
    # the inverse in LDS, the full-FFT form): every signal against the oracle.
    n, b = 1 << 15, 5
    d = f32round(synth.normal(n + 7, n * b)).reshape(b, n)
    if isign == 1:
        x = np.arange(m) - (m - 1) / 2
        r = np.stack([f32round(np.exp(-0.5 * (x / 40.0) ** 2) * (1 + 0.1 * synth.normal(70 + i, m)))
                      for i in range(b)])
    else:
        r = np.stack([f32round(0.3 / m * synth.normal(70 + i, m)) for i in range(b)])
        r[:, (m - 1) // 2] = 1.0
    y = numrs_amd.convlv_batch_responses(d.astype(np.float32), r.astype(np.float32), isign)
    for i in range(b):
        assert rel_l2(y[i], oracle.convlv(d[i], r[i], isign)) < 1e-5, (m, isign, i)
    a1, a2 = f32round(synth.normal(91, n)), f32round(synth.normal(92, n))
    assert rel_l2(numrs_amd.correl(a1.astype(np.float32), a2.astype(np.float32)), oracle.correl(a1, a2)) < 1e-5


@pytest.mark.parametrize("isign", [1, -1])
def test_convlv_batch_f64_multipass(isign):
    # FP64 convlv beyond one group with 8 signals of 2^16: the response's
    # C2C reads its taps through the placement in its first pass (16-wide
    # tiles), both forward untangles inside the spectrum step; every signal
    # against the oracle (deconvolution to 1e-9, as for single signals).
    n, b, m = 1 << 16, 8, 301
    d = synth.normal(n + 7, n * b).reshape(b, n)
    if isign == 1:
        r = np.stack([synth.normal(70 + i, m) for i in range(b)])
    else:
        r = np.stack([0.3 / m * synth.normal(70 + i, m) for i in range(b)])
        r[:, (m - 1) // 2] = 1.0
    y = numrs_amd.convlv_batch_responses(d, r, isign)
    for i in range(b):
        ref = oracle.convlv(d[i], r[i], isign)
        assert rel_l2(y[i], ref) < (TOL64 if isign == 1 else 1e-9), (isign, i)


def test_correl_plan_f64_multipass():
    # FP64 correl device plan beyond one group, 8 pairs of 2^16: every pair
    # against the oracle.
    import torch

    n, b = 1 << 16, 8
    d1 = synth.normal(90, n * b)
    d2 = synth.normal(91, n * b)
    plan = numrs_amd.CorrelPlan("float64", n, b)
    t1, t2 = torch.from_numpy(d1).cuda(), torch.from_numpy(d2).cuda()
    out = torch.empty_like(t1)
    plan.execute(t1, t2, out)
    got = out.cpu().numpy()
    for i in range(b):
        s = slice(i * n, (i + 1) * n)
        assert rel_l2(got[s], oracle.correl(d1[s], d2[s])) < TOL64, i


@pytest.mark.parametrize("dtype,logn,batch", [("float32", 22, 6), ("float64", 22, 4), ("float32", 23, 5),
                                              ("float64", 23, 3), ("float32", 24, 3), ("float64", 24, 2),
                                              ("float32", 25, 2)])
def test_c2c_three_pass(dtype, logn, batch):
    # Transforms of 2^22 points and more run as three passes N1 x Na x Mb
    # (2^8 x 2^7 x 2^7 at 2^22 ... 2^9 x 2^8 x 2^8 at 2^25) in chunks (2^23
    # FP32: chunks of 2, 2, 1): the first and last transform against numpy's
    # FP64 FFT, both signs.
    import torch

    n = 1 << logn
    x = synth.uniform(logn + 40, 2 * n * batch)
    if dtype == "float32":
        x = f32round(x)
    t = torch.from_numpy(x.astype(np.float32 if dtype == "float32" else np.float64)).cuda()
    o = torch.empty_like(t)
    plan = numrs_amd.C2CPlan(dtype, [n], batch)
    tol = TOL32 if dtype == "float32" else TOL64
    for isign in (1, -1):
        plan.execute(t, o, isign)
        got = o.cpu().numpy().astype(np.float64)
        for i in (0, batch - 1):
            z = x[2 * i * n:2 * (i + 1) * n].view(np.complex128)
            ref = n * np.fft.ifft(z) if isign == 1 else np.fft.fft(z)
            assert rel_l2(got[2 * i * n:2 * (i + 1) * n], ref.view(np.float64)) < tol, (isign, i)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_c2c_max_length(dtype):
    # The longest 1-D transform the planner takes, N = 2^26 (two passes of
    # 8192 points): against numpy's FP64 FFT, then the round trip (x N).
    import torch

    n = 1 << 26
    x = synth.uniform(26, 2 * n)
    if dtype == "float32":
        x = f32round(x)
    t = torch.from_numpy(x.astype(np.float32 if dtype == "float32" else np.float64)).cuda()
    o = torch.empty_like(t)
    plan = numrs_amd.C2CPlan(dtype, [n], 1)
    plan.execute(t, o, -1)
    tol = TOL32 if dtype == "float32" else TOL64
    ref = np.fft.fft(x.view(np.complex128))
    assert rel_l2(o.cpu().numpy().astype(np.float64), ref.view(np.float64)) < tol
    del ref
    plan.execute(o, t, 1)
    assert rel_l2(t.cpu().numpy().astype(np.float64), n * x) < 2 * tol


def test_realft_plan_multipass_chunked_f32():
    # 36 FP32 signals of 2^21 reals: the 2^20-point C2C runs in chunks of 16
    # transforms (16, 16, 4), each followed by its untangle.  Signals of the
    # first, a middle and the tail chunk against numpy's FP64 rfft (NR's
    # sign: F = conj(rfft), packed [F_0, F_{n/2}, Re F_1, Im F_1, ...]);
    # then the inverse (untangle in its first pass) returns (n/2) x.
    import torch

    n, b = 1 << 21, 36
    x = synth.normal(30, b * n).astype(np.float32)
    t = torch.from_numpy(x).cuda()
    plan = numrs_amd.RealFTPlan("float32", n, b)
    plan.execute(t, 1)
    got = t.cpu().numpy()
    for i in (0, 17, 35):
        f = np.conj(np.fft.rfft(x[i * n:(i + 1) * n].astype(np.float64)))
        ref = np.empty(n)
        ref[0], ref[1] = f[0].real, f[n // 2].real
        ref[2::2], ref[3::2] = f[1:n // 2].real, f[1:n // 2].imag
        assert rel_l2(got[i * n:(i + 1) * n], ref) < TOL32, i
    plan.execute(t, -1)
    back = t.cpu().numpy()
    for i in (0, 17, 35):
        assert rel_l2(back[i * n:(i + 1) * n], (n / 2) * x[i * n:(i + 1) * n].astype(np.float64)) < TOL32, i


@pytest.mark.parametrize("b", [1, 37])
def test_realft_plan_fused_2p15_f32(b):
    # FP32 signals of 2^15 reals run in one group each (16384-point complex
    # FFT and the untangle in LDS): every signal against numpy's FP64 rfft
    # in NR's packing, then the inverse returns (n/2) x, in place.
    import torch

    n = 1 << 15
    x = synth.normal(33, b * n).astype(np.float32)
    t = torch.from_numpy(x).cuda()
    plan = numrs_amd.RealFTPlan("float32", n, b)
    plan.execute(t, 1)
    got = t.cpu().numpy()
    for i in range(b):
        f = np.conj(np.fft.rfft(x[i * n:(i + 1) * n].astype(np.float64)))
        ref = np.empty(n)
        ref[0], ref[1] = f[0].real, f[n // 2].real
        ref[2::2], ref[3::2] = f[1:n // 2].real, f[1:n // 2].imag
        assert rel_l2(got[i * n:(i + 1) * n], ref) < TOL32, i
    plan.execute(t, -1)
    back = t.cpu().numpy()
    for i in range(b):
        assert rel_l2(back[i * n:(i + 1) * n], (n / 2) * x[i * n:(i + 1) * n].astype(np.float64)) < TOL32, i


def test_realft_plan_three_pass_f32():
    # 3 FP32 signals of 2^24 reals: the 2^23-point C2C runs as three passes
    # (chunks of 2 and 1) with one untangle pass after it; the inverse reads
    # the untangle through its first pass, in place.  First and last signal
    # against numpy's FP64 rfft, then the inverse returns (n/2) x.
    import torch

    n, b = 1 << 24, 3
    x = synth.normal(32, b * n).astype(np.float32)
    t = torch.from_numpy(x).cuda()
    plan = numrs_amd.RealFTPlan("float32", n, b)
    plan.execute(t, 1)
    got = t.cpu().numpy()
    for i in (0, b - 1):
        f = np.conj(np.fft.rfft(x[i * n:(i + 1) * n].astype(np.float64)))
        ref = np.empty(n)
        ref[0], ref[1] = f[0].real, f[n // 2].real
        ref[2::2], ref[3::2] = f[1:n // 2].real, f[1:n // 2].imag
        assert rel_l2(got[i * n:(i + 1) * n], ref) < TOL32, i
    plan.execute(t, -1)
    back = t.cpu().numpy()
    for i in (0, b - 1):
        assert rel_l2(back[i * n:(i + 1) * n], (n / 2) * x[i * n:(i + 1) * n].astype(np.float64)) < TOL32, i


def test_realft_plan_multipass_chunked_f64():
    # 40 FP64 signals of 2^18 reals: the 2^17-point four-step over several
    # chunks and a partial one, each chunk's untangle after its second pass;
    # every signal against the oracle, then the inverse (its untangle read
    # by the first pass of each chunk, in place) returns (n/2) x.
    import torch

    n, b = 1 << 18, 40
    x = synth.normal(31, b * n)
    t = torch.from_numpy(x.copy()).cuda()
    plan = numrs_amd.RealFTPlan("float64", n, b)
    plan.execute(t, 1)
    got = t.cpu().numpy()
    for i in range(b):
        assert rel_l2(got[i * n:(i + 1) * n], oracle.realft(x[i * n:(i + 1) * n], 1)) < TOL64, i
    plan.execute(t, -1)
    back = t.cpu().numpy()
    for i in (0, 21, 39):
        assert rel_l2(back[i * n:(i + 1) * n], (n / 2) * x[i * n:(i + 1) * n]) < TOL64, i


# ---- fourn -------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(2, 2), (4, 8, 16), (16, 2), (2, 64, 4), (32, 32, 32), (8, 4096), (64, 256, 2),
                                   (2, 2, 2, 2, 2)])
def test_fourn_f64(shape):
    x = synth.normal(sum(shape), 2 * int(np.prod(shape)))
    for isign in (1, -1):
        y = x.copy()
        numrs_amd.fourn(y, list(shape), len(shape), isign)
        assert rel_l2(y, oracle.fourn(x, list(shape), isign)) < TOL64, (shape, isign)


def test_fourn_golden(golden):
    for isign in (1, -1):
        y = golden[f"fourn_4x8x16_{isign}_in"].copy()
        numrs_amd.fourn(y, [4, 8, 16], 3, isign)
        assert rel_l2(y, golden[f"fourn_4x8x16_{isign}_out"]) < TOL64


@pytest.mark.parametrize("shape", [(64, 64, 64), (128, 2048)])
def test_fourn_f32(shape):
    x = f32round(synth.normal(1, 2 * int(np.prod(shape))))
    y = x.astype(np.float32)
    numrs_amd.fourn(y, list(shape), len(shape), 1)
    assert rel_l2(y, oracle.fourn(x, list(shape), 1)) < TOL32


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(4096, 64), (8192, 32), (2, 4096, 16), (4096, 4), (8192, 2, 8)])
def test_fourn_long_strided_axis(dtype, shape):
    # Strided axes of 4096 / 8192 points run as two column passes (64 x 64,
    # 64 x 128) with the W_L^{k1 l2} twiddle between them; wide, narrow and
    # batched (leading axis) layouts, both signs.
    x = synth.normal(len(shape) + shape[0], 2 * int(np.prod(shape)))
    if dtype == np.float32:
        x = f32round(x)
    for isign in (1, -1):
        y = x.astype(dtype)
        numrs_amd.fourn(y, list(shape), len(shape), isign)
        tol = TOL32 if dtype == np.float32 else TOL64
        assert rel_l2(y.astype(np.float64), oracle.fourn(x, list(shape), isign)) < tol, (shape, isign)


@pytest.mark.parametrize("shape", [(512, 64), (256, 8), (2, 256, 32)])
def test_fourn_f32_column_pass_tiles(shape):
    # Strided passes of 256 / 512 points: 16-wide tiles take the
    # LDS-DMA-staged column kernel (512 x 64, and 2 x 256 x 32 with a batch
    # stride), an 8-wide tile (256 x 8) the plain one.
    x = f32round(synth.normal(7, 2 * int(np.prod(shape))))
    for isign in (1, -1):
        y = x.astype(np.float32)
        numrs_amd.fourn(y, list(shape), len(shape), isign)
        assert rel_l2(y, oracle.fourn(x, list(shape), isign)) < TOL32, (shape, isign)


def test_four1_f32_device_plan_alignment():
    # Four-step at N = 2^18 (first pass over 512-point columns) on device
    # buffers: input 16-byte aligned (LDS-DMA-staged pass) and offset by one
    # complex value, 8 bytes (plain pass); both against the oracle.
    import torch

    n, batch = 1 << 18, 2
    x = f32round(synth.uniform(11, 2 * n * batch))
    ref = np.concatenate([oracle.four1_nr(x[i * 2 * n:(i + 1) * 2 * n], 1) for i in range(batch)])
    plan = numrs_amd.C2CPlan("float32", [n], batch)
    for off in (0, 2):
        buf = torch.zeros(2 * n * batch + 4, dtype=torch.float32, device="cuda")
        xin = buf[off:off + 2 * n * batch]
        xin.copy_(torch.from_numpy(x.astype(np.float32)))
        out = torch.empty(2 * n * batch, dtype=torch.float32, device="cuda")
        plan.execute(xin, out, 1)
        torch.cuda.synchronize()
        assert rel_l2(out.cpu().numpy().astype(np.float64), ref) < TOL32, off


# ---- convlv ------------------------------------------------------------------
def test_convlv_reference_case():
    # Convolve.rs:346-360 (circular: index 0 is 1 + 4 = 5).
    y = numrs_amd.convlv([1.0, 2.0, 3.0, 4.0], [1.0, 1.0], 1)
    assert np.allclose(y, [5.0, 3.0, 5.0, 7.0], atol=1e-12)
    proc = numrs_amd.ConvlvProcessor.new()
    assert abs(proc.process([1.0, 2.0, 3.0, 4.0], [1.0, 1.0], 1)[1] - 3.0) < 1e-10


def test_convlv_golden(golden):
    d, r = golden["convlv_1024_257_data"], golden["convlv_1024_257_respns"]
    y = numrs_amd.convlv(d, r, 1)
    assert rel_l2(y, golden["convlv_1024_257_out"]) < TOL64
    dec = numrs_amd.convlv(golden["convlv_1024_257_out"], r, -1)
    assert rel_l2(dec, oracle.convlv(golden["convlv_1024_257_out"], r, -1)) < 1e-10


@pytest.mark.parametrize("n,m", [(2, 1), (2, 2), (4, 3), (8, 8), (64, 7), (1024, 257), (4096, 257), (8192, 1),
                                 (8192, 100), (8192, 383), (16384, 257), (16384, 511), (32768, 101), (65536, 1000)])
@pytest.mark.parametrize("isign", [1, -1])
def test_convlv_f64(n, m, isign):
    d = synth.normal(n + m, n)
    r = np.exp(-0.5 * ((np.arange(m) - (m - 1) / 2) / max(m / 6, 0.5)) ** 2) * (1 + 0.1 * synth.normal(3, m))
    y = numrs_amd.convlv(d, r, isign)
    ref = oracle.convlv(d, r, isign)
    assert rel_l2(y, ref) < (TOL64 if isign == 1 else 1e-9), (n, m)


def test_convlv_batch_per_signal_f32():
    # Config-3 shape on a subset: n = 2^14, m = 257 Gaussian-windowed responses.
    n, m, b = 1 << 14, 257, 16
    d = f32round(synth.normal(2, n * b)).reshape(b, n)
    r = f32round(synth.gaussian_windowed_responses(2, b, m, 40.0, start=n * b))
    y = numrs_amd.convlv_batch_responses(d.astype(np.float32), r.astype(np.float32), 1)
    for i in range(b):
        ref = oracle.convlv(d[i], r[i], 1)
        assert rel_l2(y[i], ref) < 1e-5


@pytest.mark.parametrize("n,m", [(64, 7), (512, 17), (1024, 257), (2048, 101), (4096, 33), (8192, 257), (16384, 257),
                                 (16384, 1001), (32768, 257)])
@pytest.mark.parametrize("isign", [1, -1])
def test_convlv_batch_f32_lengths(n, m, isign):
    # FP32 one-launch convlv / deconvolution across tile shapes (TI > 1 up
    # to one signal per group, where the response's spectrum is formed
    # first; L = 512 and 8192 are the 4-waves-per-SIMD builds; m = 257 at
    # n = 8192 / 16384 takes the windowed response, m = 1001 the full FFT;
    # n = 32768 is the multi-pass form), per-signal responses, against the
    # FP64 oracle on the FP32-rounded inputs.
    b = 6
    d = f32round(synth.normal(n + 3, n * b)).reshape(b, n)
    if isign == 1:
        x = np.arange(m) - (m - 1) / 2
        r = np.stack([f32round(np.exp(-0.5 * (x / max(m / 6, 0.5)) ** 2) * (1 + 0.1 * synth.normal(40 + i, m)))
                      for i in range(b)])
    else:
        # deconvolution needs a response whose spectrum stays away from 0:
        # a unit centre tap plus taps summing to |.| <= 0.3, so |R| >= 0.7
        r = np.stack([f32round(0.3 / m * synth.normal(40 + i, m)) for i in range(b)])
        r[:, (m - 1) // 2] = 1.0
    y = numrs_amd.convlv_batch_responses(d.astype(np.float32), r.astype(np.float32), isign)
    for i in range(b):
        ref = oracle.convlv(d[i], r[i], isign)
        assert rel_l2(y[i], ref) < 1e-5, (n, m, isign, i)


@pytest.mark.parametrize("n", [64, 512, 1024, 4096, 8192, 16384, 32768])
def test_correl_f32(n):
    d1 = f32round(synth.normal(n + 11, n))
    d2 = f32round(synth.normal(n + 12, n))
    y = numrs_amd.correl(d1.astype(np.float32), d2.astype(np.float32))
    assert rel_l2(y, oracle.correl(d1, d2)) < 1e-5, n


def test_convlv_batch_shared_response():
    data = [synth.normal(5 + i, 256) for i in range(4)]
    resp = [0.5, 1.0, 0.5]
    out = numrs_amd.convlv_batch(data, resp, 1)
    for d, y in zip(data, out):
        assert rel_l2(y, oracle.convlv(d, resp, 1)) < TOL64


# ---- correl ------------------------------------------------------------------
def test_correl_reference_cases():
    y = numrs_amd.correl([1.0, 2.0, 3.0, 4.0], [1.0, 2.0, 3.0, 4.0])
    assert abs(y[0] - 30.0) < 1e-10 and y[0] > y[1]
    res = numrs_amd.correl_batch([([1.0, 2.0], [1.0, 2.0]), ([3.0, 4.0], [3.0, 4.0])])
    assert abs(res[0][0] - 5.0) < 1e-10 and abs(res[1][0] - 25.0) < 1e-10
    assert np.allclose(numrs_amd.correl([1.0, 2.0], [1.0, 2.0]), [5.0, 2.0])
    # correl_normalized correlates z-scores without dividing by n
    # (Correlation.rs:189-223), so lag 0 is n * rho = 4 here; the reference's
    # own test_normalized_correlation (Correlation.rs:435-450) expects 1.0,
    # which its code cannot produce.
    assert abs(numrs_amd.correl_normalized([1.0, 2.0, 3.0, 4.0], [1.0, 2.0, 3.0, 4.0])[0] - 4.0) < 1e-10
    assert abs(numrs_amd.correl_normalized_fast([1.0, 2.0, 3.0, 4.0], [1.0, 2.0, 3.0, 4.0])[0] - 1.0) < 1e-10


@pytest.mark.parametrize("n", [1, 3, 16, 32, 64, 1024, 16384, 32768, 131072])
def test_correl_f64(n):
    a, b = synth.normal(n, n), synth.normal(n + 1, n)
    y = numrs_amd.correl(a, b)
    assert rel_l2(y, oracle.correl(a, b)) < TOL64, n


def test_correl_golden(golden):
    assert rel_l2(numrs_amd.correl(golden["correl_64_d1"], golden["correl_64_d2"]), golden["correl_64_out"]) < TOL64
    assert rel_l2(numrs_amd.correl(golden["correl_16_d1"], golden["correl_16_d2"]), golden["correl_16_out"]) < TOL64


def test_autocorrel():
    x = synth.normal(77, 512)
    assert rel_l2(numrs_amd.autocorrel(x), oracle.correl(x, x)) < TOL64
    assert rel_l2(numrs_amd.autocorrel_fast(x), oracle.correl(x, x)) < TOL64


# ---- errors of GPU calls equal the CPU-side ones -------------------------------
def test_errors_on_gpu():
    with pytest.raises(numrs_amd.ConvlvError):
        numrs_amd.convlv([1.0, 2.0], [1.0, 2.0, 3.0], 1)
    with pytest.raises(numrs_amd.FournError):
        numrs_amd.fourn(np.zeros(16), [8], 1, 0)
