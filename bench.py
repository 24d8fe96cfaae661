"""Benchmark of the numrs FFT path on one MI355X.

Headline workload (one JSON line on stdout): batched FP32 C2C four1,
64 transforms x N = 2^20 complex, re/im i.i.d. uniform[-1,1) from seed 3,
isign = +1, inputs resident in HBM, out of place.  One step = one plan
execution over the whole batch.  GFLOP/s uses 5 N log2 N per transform.

The same line carries
  roofline:      the dominant kernel (slower four-step pass), algorithmic bytes
                 per launch (one read + one write of its chunk, 16 B per
                 point) / its average duration between HIP events recorded
                 after every launch of the plan, in sequence, on its stream;
  cpu_baseline:  the oracle's restatement of FFT_1::four1 (FP64, OpenMP over
                 transforms like FFTProcessor::fft_batch) timed on the host
                 cores on a bounded sample of the same workload;
  extra:         the other configurations (four1 2^16 FP64, fourn 256^3,
                 batched convlv 4096 x 2^14, the uneven batch of 10,000
                 mixed 1-D/2-D transforms) and twofft / cosft1 / cosft2 /
                 rlft3, each with its own GB/s figure.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "batched FFT GFLOP/s + %HBM peak; split-MFMA vs plain FP32 at rel-L2≤1e-6"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def fill_uniform(t, seed: int, start: int, count: int, chunk: int = 1 << 22) -> None:
    """Write count uniform[-1,1) draws (stream position start..) into tensor t."""
    import torch

    from numrs_amd import synth

    for off in range(0, count, chunk):
        c = min(chunk, count - off)
        v = synth.uniform(seed, c, start=start + off).astype(np.float32)
        t[off:off + c].copy_(torch.from_numpy(v))


def fill_normal(t, seed: int, count: int, dtype=np.float32, chunk: int = 1 << 22) -> None:
    import torch

    from numrs_amd import synth

    for off in range(0, count, chunk):
        c = min(chunk, count - off)
        v = synth.normal(seed, c, start=off).astype(dtype)
        t[off:off + c].copy_(torch.from_numpy(v))


def time_events(fn, reps: int) -> float:
    """Average ms of fn() over reps, HIP events on torch's current stream (the
    stream every launch here is enqueued on)."""
    import torch

    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def measured_traffic_ratio():
    """Bytes past L2 / algorithmic bytes of the FP32 four-step passes, from the
    committed rocprofv3 summary of raw TCC request counts (last column)."""
    path = os.path.join(ROOT, "profiles", "c2c_fp32_batch64_n2p20_traffic.txt")
    try:
        ratios = []
        for line in open(path):
            parts = line.split()
            if line.startswith(" nrfft::k_c2c_dma<1, 512>") or line.startswith(" nrfft::k_c2c<float, 1, 2048"):
                ratios.append(float(parts[-1]))
        return (max(ratios), os.path.relpath(path, ROOT)) if ratios else (None, None)
    except (OSError, ValueError):
        return None, None


def _rel_l2_gpu(a, ref) -> float:
    import torch

    d = a.double() - ref
    return float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(ref))


def _settle(fn, seconds: float) -> None:
    """Run fn back to back for `seconds` before a timed region, so the GPU
    clocks have settled whatever the warm-up step count (a 20-step run right
    after start-up measured ~7-10% slower per step than a settled one)."""
    import torch

    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(10):
            fn()
        torch.cuda.synchronize()


def _time_plan(plan, x, y, steps: int, warmup: int, settle: float = 1.0) -> float:
    import torch

    _settle(lambda: plan.execute(x, y, 1), settle)
    for _ in range(warmup):
        plan.execute(x, y, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.execute(x, y, 1)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def _time_launches_in_sequence(plan, x, y, steps: int) -> list:
    """Average ms of each launch of the plan, measured between consecutive
    HIP events in `steps` back-to-back executions."""
    import torch

    nl = plan.num_launches
    s = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps * nl + 1)]
    torch.cuda.synchronize()
    ev[0].record(s)
    for st in range(steps):
        for i in range(nl):
            plan.execute_launches(x, y, 1, i, 1)
            ev[st * nl + i + 1].record(s)
    torch.cuda.synchronize()
    return [sum(ev[st * nl + i].elapsed_time(ev[st * nl + i + 1]) for st in range(steps)) / steps for i in range(nl)]


def headline(args) -> dict:
    import torch

    import numrs_amd

    batch, logn = 64, 20
    n = 1 << logn
    pts = batch * n
    log(f"[bench] c2c fp32 batch {batch} x 2^{logn}: generating inputs")
    x = torch.empty(2 * pts, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    fill_uniform(x, 3, 0, 2 * pts)
    plan = numrs_amd.C2CPlan("float32", [n], batch)
    nl = plan.num_launches
    log(f"[bench] timing {args.steps} steps ({nl} launches per step)")
    ms = _time_plan(plan, x, y, args.steps, args.warmup)
    flops = 5.0 * n * logn * batch
    gflops = flops / (ms * 1e-3) / 1e9
    compulsory = 2.0 * pts * 8  # one read + one write, 8 B per complex
    # Per-launch durations: the plan's launches issued one by one in their
    # normal order for `steps` more steps, with a HIP event recorded after
    # each on the stream the library launches on (torch's current stream), so
    # every launch is timed inside the real sequence (the chunk written by
    # pass A is what pass B reads).  Four-step plans alternate pass A / pass B
    # over batch chunks; each launch reads and writes every point of its
    # chunk once.
    per = _time_launches_in_sequence(plan, x, y, args.steps)
    kinds = ["A", "B"] if nl % 2 == 0 and nl > 1 else ["pass"]
    groups = {k: [per[i] for i in range(nl) if kinds[i % len(kinds)] == k] for k in kinds}
    dom = max(groups, key=lambda k: sum(groups[k]))
    launch_bytes = compulsory / (nl // len(kinds))
    dom_ms = sum(groups[dom]) / len(groups[dom])
    achieved = launch_bytes / (dom_ms * 1e-3) / 1e9
    ratio, tsrc = measured_traffic_ratio()
    out = {
        "metric": METRIC,
        "value": round(gflops, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: re/im i.i.d. uniform[-1,1) from SplitMix64 seed 3",
        "config": {"workload": "four1 C2C FP32, batch 64 x N=2^20, isign=+1, out of place, HBM-resident",
                   "batch": batch, "n": n, "launches": nl, "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": int(ratio * launch_bytes) if ratio else None,
                     "traffic_source": tsrc,
                     # FP32 pass A over 512-point columns runs as k_c2c_dma (16-byte aligned input)
                     "kernel": f"four-step pass {dom} ({'k_c2c_dma' if dom == 'A' else 'k_c2c'}), "
                               f"{len(groups[dom])} launches per step",
                     "kernel_ms": round(dom_ms, 4), "algorithmic_bytes_per_launch": int(launch_bytes)},
        "hbm_effective": {"GBps": round(compulsory / (ms * 1e-3) / 1e9, 1),
                          "frac": round(compulsory / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                          "note": "compulsory bytes (one read + one write of the batch) / whole step"},
        "traffic_rate": ({"GBps": round(ratio * launch_bytes * nl / (ms * 1e-3) / 1e9, 1),
                          "frac": round(ratio * launch_bytes * nl / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                          "note": "counter-measured bytes past L2 per step (rocprofv3 TCC read / write request "
                                  "counts of every launch, from traffic_source) / whole step; each of the two "
                                  "passes reads and writes every point"} if ratio else None),
        "launch_ms": [round(p, 4) for p in per],
    }
    # FP64 reference of the same FP32 inputs from the FP64 path (parity-tested
    # against the oracle at 1e-12), for the relative L2 of every FP32 form.
    log("[bench] fp64 reference and split-precision MFMA forms")
    x64 = x.double()
    y64 = torch.empty_like(x64)
    p64 = numrs_amd.C2CPlan("float64", [n], batch)
    p64.execute(x64, y64, 1)
    torch.cuda.synchronize()
    out["rel_l2_vs_fp64"] = _rel_l2_gpu(y, y64)
    # The FP64 plan on the same batch (the reference computes in f64 only).
    y64b = torch.empty_like(x64)
    ms64 = _time_plan(p64, x64, y64b, args.steps, args.warmup)
    comp64 = 2.0 * pts * 16
    out["fp64"] = {"ms": round(ms64, 4), "GFLOPs": round(flops / (ms64 * 1e-3) / 1e9, 1),
                   "GBps_compulsory": round(comp64 / (ms64 * 1e-3) / 1e9, 1),
                   "frac_hbm": round(comp64 / (ms64 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                   "launches": p64.num_launches}
    del x64, y64b, p64
    split = {}
    for name in ("f16", "bf16"):
        sp = numrs_amd.C2CPlan("float32", [n], batch, split=name)
        sms = _time_plan(sp, x, y, args.steps, args.warmup)
        split[name] = {"ms": round(sms, 4), "GFLOPs": round(flops / (sms * 1e-3) / 1e9, 1),
                       "rel_l2_vs_fp64": _rel_l2_gpu(y, y64),
                       "speedup_vs_plain_fp32": round(ms / sms, 3), "launches": sp.num_launches}
        del sp
    out["split_mfma"] = split
    del x, y, y64
    torch.cuda.empty_cache()
    return out


def _host_threads() -> dict:
    """Host cores this process may use.  OpenMP pools are capped at 16
    threads: the GPU host's share of CPU for one GPU is 16 workers, whatever
    nproc reports there."""
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = nproc
    return {"nproc": nproc, "affinity": aff, "threads": max(1, min(aff, 16))}


def _median_time(run, reps: int):
    """(median seconds, value returned by the last run) over reps timed runs
    after one untimed warm-up run."""
    out = run()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        out = run()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), out


def cpu_baseline(args) -> dict:
    """The reference's CPU path, restated by the oracle (FP64, the reference's
    only precision), timed on the host cores beside the GPU run, on bounded
    samples of each configuration:
      headline (64 x 2^20): the literal four1 of FFT_1.rs:5-107 over
        transforms on all usable cores (OpenMP, as FFTProcessor::fft_batch
        spreads them with rayon, FFT_1.rs:185-189), and on one core;
      four1 2^16: one transform (serial, as the reference runs one call);
      fourn 256^3: NR fourn (serial: the reference has no parallel fourn);
      convlv 4096 x 2^14, m = 257: over signals (Convolve.rs:246-249) on all
        cores and one core;
      the uneven batch: fft_batch over its items (1-D literal four1, 2-D
        fourn) on all cores and one core.
    The oracle is built with -march=native on this host when gcc is here
    (oracle/liborc_native.so), else the portable build is used."""
    import oracle
    from numrs_amd import synth

    host = _host_threads()
    L = oracle.native_lib()
    build = "gcc -O3 -march=native -ffp-contract=off -fopenmp (built on this host)"
    if L is None:
        L = oracle.lib()
        build = "gcc -O3 -ffp-contract=off -fopenmp (portable build; -march=native build failed)"
    threads = host["threads"]
    res = {}

    # headline shape: 64 transforms of 2^20 on all cores, 8 on one core
    n = 1 << 20
    sample = args.cpu_sample
    src = np.empty(2 * n * sample, np.float64)
    for i in range(sample):
        src[2 * n * i:2 * n * (i + 1)] = synth.uniform(3, 2 * n, start=2 * n * i).astype(np.float32)

    def four1_batch(count, nthreads):
        data = src[:2 * n * count].copy()
        return oracle.four1_batch_inplace(data, n, count, 1, nthreads, L)

    med, used = _median_time(lambda: four1_batch(sample, threads), args.cpu_reps)
    one = max(1, sample // 8)
    med1, _ = _median_time(lambda: four1_batch(one, 1), 3)
    del src
    fl = 5.0 * n * 20
    out = {"value": round(fl * sample / med / 1e9, 3), "unit": "GFLOP/s", "cores": used, "kind": "port",
           "sample": f"{sample} of the 64 transforms of N=2^20 (FP64 literal four1 of FFT_1.rs), median of "
                     f"{args.cpu_reps} runs after one warm-up: {med:.2f} s wall on {used} threads",
           "single_core": {"value": round(fl * one / med1 / 1e9, 3), "unit": "GFLOP/s", "cores": 1,
                           "sample": f"{one} transforms of N=2^20, median of 3 runs: {med1:.2f} s"},
           "host": dict(host, build=build)}

    # four1 2^16 FP64, uniform seed 0, one transform
    n1 = 1 << 16
    x1 = synth.uniform(0, 2 * n1)

    def c1():
        d = x1.copy()
        oracle.four1_literal_with(L, d, 1)
        return d

    t1, _ = _median_time(c1, 21)
    res["four1_f64_2p16"] = {"us": round(t1 * 1e6, 1), "GFLOPs": round(5.0 * n1 * 16 / t1 / 1e9, 3),
                                     "cores": 1, "sample": "one transform, median of 21 runs"}

    # convlv 4096 x 2^14, m = 257 (FP64 on the FP32-rounded inputs)
    b3, n3, m3 = 4096, 1 << 14, 257
    d3 = synth.normal(2, b3 * n3).astype(np.float32).astype(np.float64)
    r3 = synth.gaussian_windowed_responses(2, b3, m3, 40.0, start=b3 * n3).astype(np.float32).astype(np.float64)
    a3 = np.empty_like(d3)
    t3, used3 = _median_time(lambda: oracle.convlv_batch(d3, n3, r3.reshape(-1), m3, b3, 1, a3, threads, L), 3)
    one3 = 256
    t31, _ = _median_time(lambda: oracle.convlv_batch(d3, n3, r3.reshape(-1), m3, one3, 1, a3, 1, L), 3)
    fl3 = 3 * 2.5 * n3 * 14
    res["convlv_4096x2p14_m257"] = {
        "ms": round(t3 * 1e3, 2), "signals_per_s": round(b3 / t3, 1), "GFLOPs": round(fl3 * b3 / t3 / 1e9, 3),
        "GBps": round((2.0 * n3 + m3) * 8 * b3 / t3 / 1e9, 2), "cores": used3,
        "sample": f"all 4096 signals, median of 3 runs on {used3} threads",
        "single_core": {"signals_per_s": round(one3 / t31, 1), "GFLOPs": round(fl3 * one3 / t31 / 1e9, 3),
                        "sample": f"{one3} signals, median of 3 runs"}}
    del d3, r3, a3

    # the uneven batch's first items (the list is in random order)
    items = synth.uneven_shapes(4, 10000)
    offs, acc = [], 0
    for it in items:
        offs.append(acc)
        acc += int(np.prod(it))

    def uneven(count, nthreads):
        tot = offs[count - 1] + int(np.prod(items[count - 1]))
        data = synth.uniform(4, 2 * tot).astype(np.float32).astype(np.float64)
        t0 = time.perf_counter()
        used_ = oracle.uneven_batch_inplace(data, items[:count], offs[:count], 1, nthreads, L)
        return time.perf_counter() - t0, used_

    def uneven_flops(count):
        return sum(5.0 * int(np.prod(it)) * math.log2(int(np.prod(it))) for it in items[:count])

    k5, k51 = args.cpu_uneven_items, max(1, args.cpu_uneven_items // 8)
    uneven(k5, threads)  # warm-up
    r5 = sorted(uneven(k5, threads) for _ in range(3))[1]
    r51 = sorted(uneven(k51, 1) for _ in range(3))[1]
    res["uneven_first_items"] = {
        "GFLOPs": round(uneven_flops(k5) / r5[0] / 1e9, 3), "items_per_s": round(k5 / r5[0], 1),
        "cores": r5[1], "sample": f"items 0..{k5 - 1} of the 10,000 (shapes drawn in random order), "
                                  f"median of 3 runs: {r5[0]:.2f} s",
        "single_core": {"GFLOPs": round(uneven_flops(k51) / r51[0] / 1e9, 3),
                        "sample": f"items 0..{k51 - 1}, median of 3 runs: {r51[0]:.2f} s"}}

    # NR fourn 256^3 FP64 (serial), one run after none: ~10 s
    x2 = synth.normal(1, 2 * 256 ** 3).astype(np.float32).astype(np.float64)
    t0 = time.perf_counter()
    oracle.fourn_with(L, x2, [256, 256, 256], 1)
    t2 = time.perf_counter() - t0
    res["fourn_256cubed_f64"] = {"s": round(t2, 3), "GFLOPs": round(5.0 * 256 ** 3 * 24 / t2 / 1e9, 3),
                                 "cores": 1, "sample": "one forward transform (FP32-rounded N(0,1) "
                                                       "seed 1, the GPU bench's input)"}
    out["workloads"] = res
    out["_fourn_oracle_output"] = x2
    return out


def _fourn_gpu_vs_oracle(ref: np.ndarray) -> float:
    """Relative L2 of the FP64 GPU fourn 256^3 (the bench's input) against the
    oracle's output for the same input, computed by the CPU-baseline leg."""
    import torch

    import numrs_amd

    pts = 256 ** 3
    x = torch.empty(2 * pts, dtype=torch.float64, device="cuda")
    fill_normal(x, 1, 2 * pts, np.float32)
    numrs_amd.C2CPlan("float64", [256, 256, 256], 1).execute(x, x, 1)
    got = x.cpu().numpy()
    del x
    torch.cuda.empty_cache()
    return float(np.linalg.norm(got - ref) / np.linalg.norm(ref))


def copy_calibration() -> dict:
    """A plain 512 MiB device-to-device copy: the reference kernel whose
    FETCH_SIZE / WRITE_SIZE counters scale the FFT kernels' counters to bytes."""
    import torch

    nbytes = 512 << 20
    a = torch.zeros(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    ms = time_events(lambda: b.copy_(a), 10)
    del a, b
    torch.cuda.empty_cache()
    return {"bytes_read": nbytes, "bytes_written": nbytes, "ms": round(ms, 4),
            "GBps_rw": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}


def extra_configs(args, only: str = "all") -> dict:
    import torch

    import numrs_amd

    res = {}
    if only in ("all", "four1"):
        res.update(_extra_four1())
        res.update(_extra_host_call())
    if only in ("all", "fourn"):
        res.update(_extra_fourn())
    if only in ("all", "convlv"):
        res.update(_extra_convlv())
    if only in ("all", "uneven"):
        res.update(_extra_uneven())
    if only in ("all", "spectral"):
        res.update(_extra_spectral())
    if only in ("all", "long"):
        res.update(_extra_long())
    if only in ("all", "fourfs"):
        res.update(_extra_fourfs())
    return res


def _extra_long() -> dict:
    """Long 1-D transforms (2^23, 2^26 points: three passes of at most 2^9
    points), 512 MiB batches, device resident, forward.  Bytes are
    compulsory: the batch read once and written once."""
    import torch

    import numrs_amd

    res = {}
    for dt, logn, b in (("float32", 23, 8), ("float32", 26, 1), ("float64", 23, 4)):
        n = 1 << logn
        tdt = getattr(torch, dt)
        x = torch.empty(2 * n * b, dtype=tdt, device="cuda")
        numrs_amd.gen_uniform(x, 50 + logn)
        y = torch.empty_like(x)
        plan = numrs_amd.C2CPlan(dt, [n], b)
        _settle(lambda: plan.execute(x, y, 1), 0.3)
        ms = time_events(lambda: plan.execute(x, y, 1), 20)
        byt = 2.0 * x.numel() * x.element_size()
        res[f"c2c_{'f32' if dt == 'float32' else 'f64'}_{b}x2p{logn}"] = {
            "ms": round(ms, 4), "GBps": round(byt / (ms * 1e-3) / 1e9, 1),
            "frac_hbm": round(byt / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "GFLOPs": round(5.0 * n * logn * b / (ms * 1e-3) / 1e9, 1), "launches": plan.num_launches}
        del plan, x, y
    return res


def _extra_four1() -> dict:
    import torch

    import numrs_amd

    res = {}
    # four1 N=2^16 FP64, single transform, device resident.
    n = 1 << 16
    x = torch.empty(2 * n, dtype=torch.float64, device="cuda")
    fill_uniform(x, 0, 0, 2 * n)
    x = x.double()
    y = torch.empty_like(x)
    plan = numrs_amd.C2CPlan("float64", [n], 1)
    _settle(lambda: plan.execute(x, y, 1), 0.3)
    ms = time_events(lambda: plan.execute(x, y, 1), 200)
    res["four1_f64_2p16"] = {"us": round(ms * 1e3, 2), "GFLOPs": round(5 * n * 16 / (ms * 1e-3) / 1e9, 2),
                             "GBps": round(2 * n * 16 / (ms * 1e-3) / 1e9, 1)}
    del plan, x, y
    return res


def _extra_host_call() -> dict:
    """The headline batch through the host-pointer drop-in call
    (nrfft_four1_batch_f32: host -> HBM, transform, HBM -> host inside the
    call), i.e. the PCIe-inclusive rate.  Never the headline value."""
    from numrs_amd import _lib

    batch, n = 64, 1 << 20
    src = np.tile(np.linspace(-1.0, 1.0, 2 * n, dtype=np.float32), batch)
    h = src.copy()
    p = h.ctypes.data_as(_lib._fp)
    rc = _lib.lib().nrfft_four1_batch_f32(p, n, batch, 1)
    if rc:
        raise RuntimeError(_lib.last_error())
    times = []
    for _ in range(3):
        h[:] = src
        t0 = time.perf_counter()
        _lib.lib().nrfft_four1_batch_f32(p, n, batch, 1)
        times.append(time.perf_counter() - t0)
    ms = float(np.median(times)) * 1e3
    return {"four1_batch_host_f32_64x2p20": {
        "ms": round(ms, 2), "GFLOPs": round(5.0 * n * 20 * batch / (ms * 1e-3) / 1e9, 1),
        "note": "host-pointer call on pageable numpy memory: 512 MiB copied in and out over PCIe per call"}}


def _extra_fourn() -> dict:
    import torch

    import numrs_amd

    res = {}
    # fourn 256^3 FP32 and FP64: total, and each axis pass
    # timed between HIP events inside the plan's own sequence (a pass timed
    # repeated on its own finds its 128-256 MiB array warm in the MALL).
    # The FP32 result's relative L2 is against the FP64 path on the same
    # (FP32-rounded) input; the FP64 path is checked against the oracle in
    # the tests and in the CPU-baseline leg.
    shape = [256, 256, 256]
    pts = 256 ** 3
    ref = None
    for dt, tdt, es in (("float64", torch.float64, 16), ("float32", torch.float32, 8)):
        # N(0,1) from seed 1 rounded to FP32 (the FP64 run holds the same values)
        x = torch.empty(2 * pts, dtype=tdt, device="cuda")
        fill_normal(x, 1, 2 * pts, np.float32)
        y = torch.empty_like(x)
        plan = numrs_amd.C2CPlan(dt, shape, 1)
        plan.execute(x, y, 1)
        torch.cuda.synchronize()
        if ref is None:
            ref = y.clone()
            err = None
        else:
            err = _rel_l2_gpu(y, ref)
        _settle(lambda: plan.execute(x, y, 1), 0.5)
        total = time_events(lambda: plan.execute(x, y, 1), 20)
        per = _time_launches_in_sequence(plan, x, y, 20)
        byt = 2.0 * pts * es
        res[f"fourn_256cubed_{dt}"] = {
            "ms": round(total, 4), "GFLOPs": round(5 * pts * 24 / (total * 1e-3) / 1e9, 1),
            "GBps_total_compulsory": round(byt / (total * 1e-3) / 1e9, 1),
            "frac_hbm": round(byt / (total * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "pass_ms_in_sequence": [round(p, 4) for p in per],
            "pass_GBps_in_sequence": [round(byt / (p * 1e-3) / 1e9, 1) for p in per]}
        if err is not None:
            res[f"fourn_256cubed_{dt}"]["rel_l2_vs_fp64"] = err
        del plan, x, y
    del ref
    torch.cuda.empty_cache()
    return res


def _extra_convlv() -> dict:
    import torch

    import numrs_amd
    from numrs_amd import synth

    res = {}
    # batched convlv 4096 x 2^14 FP32, m = 257 per-signal responses.

    b, n, m = 4096, 1 << 14, 257
    d = torch.empty(b * n, dtype=torch.float32, device="cuda")
    fill_normal(d, 2, b * n)
    r = torch.from_numpy(synth.gaussian_windowed_responses(2, b, m, 40.0, start=b * n).astype(np.float32)).cuda()
    a = torch.empty_like(d)
    plan = numrs_amd.ConvlvPlan("float32", n, m, b)
    _settle(lambda: plan.execute(d, r, a, 1), 0.5)
    ms = time_events(lambda: plan.execute(d, r, a, 1), 20)
    byt = (2.0 * n + m) * 4 * b
    res["convlv_f32_4096x2p14_m257"] = {
        "ms": round(ms, 4), "GBps": round(byt / (ms * 1e-3) / 1e9, 1),
        "frac_hbm": round(byt / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "GFLOPs": round(3 * 2.5 * n * 14 * b / (ms * 1e-3) / 1e9, 1), "launches": plan.num_launches}
    del plan, d, r, a
    torch.cuda.empty_cache()

    # batched realft beyond one group: 64 x 2^21 FP32 reals in place, forward
    # (C2C of 2^20 points in 16-transform chunks, each chunk's untangle after
    # its second pass) and inverse (the untangle in each chunk's first pass).
    # The inverse scales by n/2, so repeating it on one buffer would overflow:
    # it is timed over 20 copies of a forward spectrum, one per call, after
    # a settle phase of inverses on a scratch copy (the first launch of a
    # plan's inverse kernels costs ~9 ms once).
    b, n = 64, 1 << 21
    x = torch.empty(b * n, dtype=torch.float32, device="cuda")
    fill_normal(x, 2, b * n)
    plan = numrs_amd.RealFTPlan("float32", n, b)
    _settle(lambda: plan.execute(x, 1), 0.5)
    ms = time_events(lambda: plan.execute(x, 1), 20)
    byt = 2.0 * n * 4 * b
    res["realft_f32_64x2p21_fwd"] = {
        "ms": round(ms, 4), "GBps": round(byt / (ms * 1e-3) / 1e9, 1),
        "frac_hbm": round(byt / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "GFLOPs": round(2.5 * n * 21 * b / (ms * 1e-3) / 1e9, 1), "launches": plan.num_launches}
    fill_normal(x, 2, b * n)
    plan.execute(x, 1)
    copies = [x.clone() for _ in range(20)]
    it = iter(copies)
    _settle(lambda: plan.execute(x, -1), 0.3)
    ms = time_events(lambda: plan.execute(next(it), -1), 20)
    res["realft_f32_64x2p21_inv"] = {
        "ms": round(ms, 4), "GBps": round(byt / (ms * 1e-3) / 1e9, 1),
        "frac_hbm": round(byt / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "GFLOPs": round(2.5 * n * 21 * b / (ms * 1e-3) / 1e9, 1)}
    del plan, x, copies
    torch.cuda.empty_cache()

    # FP32 realft and convlv at n = 2^15, one group per signal (the
    # 16384-point FFTs in LDS): realft 4096 x 2^15 forward, convlv 2048 x
    # 2^15 with m = 257 per-signal responses.
    b, n = 4096, 1 << 15
    x = torch.empty(b * n, dtype=torch.float32, device="cuda")
    fill_normal(x, 3, b * n)
    plan = numrs_amd.RealFTPlan("float32", n, b)
    _settle(lambda: plan.execute(x, 1), 0.3)
    ms = time_events(lambda: plan.execute(x, 1), 20)
    byt = 2.0 * n * 4 * b
    res["realft_f32_4096x2p15_fwd"] = {
        "ms": round(ms, 4), "GBps": round(byt / (ms * 1e-3) / 1e9, 1),
        "frac_hbm": round(byt / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "GFLOPs": round(2.5 * n * 15 * b / (ms * 1e-3) / 1e9, 1), "launches": plan.num_launches}
    del plan, x
    b, n, m = 2048, 1 << 15, 257
    d = torch.empty(b * n, dtype=torch.float32, device="cuda")
    fill_normal(d, 4, b * n)
    r = torch.from_numpy(synth.gaussian_windowed_responses(4, b, m, 40.0, start=b * n).astype(np.float32)).cuda()
    a = torch.empty_like(d)
    plan = numrs_amd.ConvlvPlan("float32", n, m, b)
    _settle(lambda: plan.execute(d, r, a, 1), 0.3)
    ms = time_events(lambda: plan.execute(d, r, a, 1), 20)
    byt = (2.0 * n + m) * 4 * b
    res["convlv_f32_2048x2p15_m257"] = {
        "ms": round(ms, 4), "GBps": round(byt / (ms * 1e-3) / 1e9, 1),
        "frac_hbm": round(byt / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "launches": plan.num_launches}
    del plan, d, r, a
    torch.cuda.empty_cache()
    return res


def _extra_spectral() -> dict:
    """twofft, cosft1, cosft2 and rlft3 (FP32, batched, device resident).
    Bytes are compulsory: every input read once, every output written once."""
    import torch

    import numrs_amd

    res = {}

    def entry(ms, byt, flops=None):
        e = {"ms": round(ms, 4), "GBps": round(byt / (ms * 1e-3) / 1e9, 1),
             "frac_hbm": round(byt / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        if flops:
            e["GFLOPs"] = round(flops / (ms * 1e-3) / 1e9, 1)
        return e

    f32 = torch.float32
    for b, n in ((4096, 1 << 14), (64, 1 << 20)):
        d1 = torch.empty(b * n, dtype=f32, device="cuda")
        d2 = torch.empty_like(d1)
        numrs_amd.gen_uniform(d1, 40)
        numrs_amd.gen_uniform(d2, 41)
        f1 = torch.empty(2 * b * n, dtype=f32, device="cuda")
        f2 = torch.empty_like(f1)
        p = numrs_amd.TwofftPlan("float32", n, b)
        _settle(lambda: p.execute(d1, d2, f1, f2), 0.3)
        ms = time_events(lambda: p.execute(d1, d2, f1, f2), 10)
        res[f"twofft_f32_{b}x2p{n.bit_length() - 1}"] = entry(
            ms, (2 * n + 4 * n) * 4.0 * b, 5.0 * n * math.log2(n) * b)
        del p, d1, d2, f1, f2
    for b, n in ((4096, 1 << 14), (64, 1 << 20)):
        x = torch.empty(b * (n + 1), dtype=f32, device="cuda")
        numrs_amd.gen_uniform(x, 42)
        y = torch.empty_like(x)
        p = numrs_amd.Cosft1Plan("float32", n, b)
        _settle(lambda: p.execute(x, y), 0.3)
        ms = time_events(lambda: p.execute(x, y), 10)
        res[f"cosft1_f32_{b}x2p{n.bit_length() - 1}"] = entry(ms, 2.0 * (n + 1) * 4 * b)
        del p, x, y
    for b, n in ((4096, 1 << 14), (64, 1 << 20)):
        x = torch.empty(b * n, dtype=f32, device="cuda")
        numrs_amd.gen_uniform(x, 43)
        y = torch.empty_like(x)
        p = numrs_amd.Cosft2Plan("float32", n, b)
        for isign, tag in ((1, "fwd"), (-1, "inv")):
            _settle(lambda: p.execute(x, y, isign), 0.3)
            ms = time_events(lambda: p.execute(x, y, isign), 10)
            res[f"cosft2_{tag}_f32_{b}x2p{n.bit_length() - 1}"] = entry(ms, 2.0 * n * 4 * b)
        del p, x, y
    shape = (256, 256, 256)
    vol = 256 ** 3
    x = torch.empty(vol, dtype=f32, device="cuda")
    numrs_amd.gen_uniform(x, 44)
    sp = torch.empty(2 * 256 * 256, dtype=f32, device="cuda")
    p = numrs_amd.Rlft3Plan("float32", *shape, 1)
    _settle(lambda: p.execute(x, sp, 1), 0.3)
    ms = time_events(lambda: p.execute(x, sp, 1), 10)
    res["rlft3_f32_256cubed_fwd"] = entry(ms, 2.0 * vol * 4 + sp.numel() * 4, 2.5 * vol * 24)
    del p, x, sp
    torch.cuda.empty_cache()
    return res


def _extra_fourfs() -> dict:
    """Out-of-core fourn (Four_FS.rs): a host array larger than the device
    budget streamed through HBM (PCIe-bound by construction)."""
    import numrs_amd

    res = {}
    # (name, shape, budget, passes over PCIe: the 1-D four-step makes two)
    for name, shape, budget, passes in (("fourfs_f32_1d_2p25_budget64MiB", [1 << 25], 64 << 20, 2),
                                        ("fourfs_f32_512x512x256_budget64MiB", [512, 512, 256], 64 << 20, 3)):
        total = int(np.prod(shape))
        x = np.tile(np.linspace(-1.0, 1.0, 2 * 4096, dtype=np.float32), total // 4096)
        g = x.copy()
        scratch = np.empty_like(g) if len(shape) == 1 else None
        numrs_amd.fourfs(g, shape, len(shape), 1, device_bytes=budget, scratch=scratch)  # warm-up
        ts = []
        for _ in range(3):
            g[:] = x
            t0 = time.perf_counter()
            numrs_amd.fourfs(g, shape, len(shape), 1, device_bytes=budget, scratch=scratch)
            ts.append(time.perf_counter() - t0)
        s = float(np.median(ts))
        res[name] = {"s": round(s, 4), "data_MiB": total * 8 >> 20, "device_budget_MiB": budget >> 20,
                     "GBps_data": round(total * 8 / s / 1e9, 2),
                     "GBps_pcie_each_way": round(passes * total * 8 / s / 1e9, 2),
                     "note": f"host array in pageable memory; {passes} passes, each moving the data over "
                             "PCIe once in each direction; median of 3"}
    return res


def _segment_sums(v, lengths):
    """Sums of consecutive segments of v (lengths in elements), on the GPU,
    from differences of an FP64 prefix sum."""
    import torch

    c = torch.cumsum(v, 0)
    ends = torch.as_tensor(np.cumsum(lengths) - 1, device=v.device)
    tot = c[ends]
    return torch.cat([tot[:1], tot[1:] - tot[:-1]])


def _extra_uneven() -> dict:
    import torch

    import numrs_amd
    from numrs_amd import synth

    res = {}
    # 10,000 transforms, 80% 1-D (log2 N on 8..20), 20% 2-D (sides 2^6..2^11),
    # values uniform[-1,1) from seed 4, FP32, one graph-replayed queue launch.
    items = synth.uneven_shapes(4, 10000)
    # the autotuner picks the four-step split of every 1-D size class that
    # takes two passes (2^14 .. 2^20) before the plan is built
    tuned = numrs_amd.autotune_uneven("float32", 14, 20, iters=3)
    plan = numrs_amd.UnevenPlan("float32", items)
    total = plan.total
    x = torch.empty(2 * total, dtype=torch.float32, device="cuda")
    numrs_amd.gen_uniform(x, 4)
    y = torch.empty_like(x)
    _settle(lambda: plan.execute(x, y, 1), 0.5)
    plan.check()
    ms = time_events(lambda: plan.execute(x, y, 1), 10)
    plan.check()
    flops = sum(5.0 * int(np.prod(it)) * math.log2(int(np.prod(it))) for it in items)
    byt = 2.0 * total * 8
    info = plan.info()
    del plan
    # accuracy: the FP64 queue path on the same (FP32-rounded) inputs
    y64 = x.double()
    del x
    p64 = numrs_amd.UnevenPlan("float64", items)
    p64.execute(y64, y64, 1)
    p64.check()
    del p64
    worst, acc_e, acc_r = 0.0, 0.0, 0.0
    lens = [2 * int(np.prod(it)) for it in items]
    pos, i0 = 0, 0
    while i0 < len(items):
        i1, span = i0, 0
        while i1 < len(items) and (span == 0 or span + lens[i1] <= (1 << 28)):
            span += lens[i1]
            i1 += 1
        d = y[pos:pos + span].double() - y64[pos:pos + span]
        e = _segment_sums(d * d, lens[i0:i1])
        r = _segment_sums(y64[pos:pos + span] ** 2, lens[i0:i1])
        worst = max(worst, float(torch.sqrt(e / r).max()))
        acc_e += float(e.sum())
        acc_r += float(r.sum())
        pos += span
        i0 = i1
    del y, y64
    torch.cuda.empty_cache()
    n1d = sum(1 for it in items if len(it) == 1)
    res["uneven_f32_10000"] = {
        "items": len(items), "items_1d": n1d, "items_2d": len(items) - n1d, "points": total,
        "ms": round(ms, 4), "GFLOPs": round(flops / (ms * 1e-3) / 1e9, 1),
        "GBps": round(byt / (ms * 1e-3) / 1e9, 1), "frac_hbm": round(byt / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "max_rel_l2_vs_fp64": worst, "rel_l2_all_vs_fp64": math.sqrt(acc_e / acc_r),
        "tasks": [info["tasks_phase1"], info["tasks_phase2"]], "groups": info["groups"],
        "autotuned_split_log2_n1": {str(k): v[0] for k, v in tuned.items()},
        "launch": "graph replay: counter-reset fill kernel + the persistent queue kernel, one dispatch per pass phase"}
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 200 timed steps (~0.1 s) after 50 warm-up steps: shorter runs start
    # before the clocks settle (20 steps after 3 warm-up: ~7% slower per step).
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-sample", type=int, default=64, help="transforms in the CPU baseline sample")
    ap.add_argument("--cpu-reps", type=int, default=5, help="timed runs of the CPU sample (median kept)")
    ap.add_argument("--cpu-uneven-items", type=int, default=400, help="uneven items in the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--only", choices=["headline", "four1", "fourn", "convlv", "uneven", "spectral", "long", "fourfs"],
                    default=None,
                    help="run one workload only (profiling runs); the JSON line then holds just that workload")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a 512 MiB copy kernel (counter calibration for the traffic figure)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the FFT path shards by transform with no exchange, and this build "
              "runs and reports the single-GPU figure only (use --gpus 1).", file=sys.stderr)
        return 2
    import torch

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        return 1
    from numrs_amd import _lib

    rc = _lib.lib().nrfft_init(0)
    if rc:
        raise RuntimeError(_lib.last_error())
    calib = copy_calibration() if args.calib_copy else None
    if args.only and args.only != "headline":
        out = {"metric": METRIC, "only": args.only, "extra": extra_configs(args, args.only)}
        if calib:
            out["copy_calibration"] = calib
        print(json.dumps(out), flush=True)
        return 0
    out = headline(args)
    if calib:
        out["copy_calibration"] = calib
    if not args.no_extra and args.only is None:
        log("[bench] secondary configurations")
        out["extra"] = extra_configs(args)
    if not args.no_cpu and args.only is None:
        log("[bench] cpu baseline")
        cb = cpu_baseline(args)
        ref = cb.pop("_fourn_oracle_output")
        cb["workloads"]["fourn_256cubed_f64"]["gpu_fp64_rel_l2_vs_oracle"] = _fourn_gpu_vs_oracle(ref)
        del ref
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
