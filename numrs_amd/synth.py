"""Seeded synthetic inputs shared by the tests and the benchmark.

The reference has no random generator on its FFT path, so every seeded
workload draws from this one definition:

* SplitMix64 stream from ``seed``: state_i = seed + (i+1) * 0x9E3779B97F4A7C15,
  x_i = mix(state_i); u_i = (x_i >> 11) * 2^-53.
* uniform[-1, 1): 2u - 1.
* N(0, 1): Box-Muller on consecutive pairs (u1, u2):
  r = sqrt(-2 ln(1 - u1)), (r cos 2 pi u2, r sin 2 pi u2).
* Values are drawn in storage order (re0, im0, re1, ...).  FP32 inputs are the
  FP64 draws rounded to FP32.

SplitMix64 is counter based, so the stream is generated vectorised.
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _u01(seed: int, n: int, start: int = 0) -> np.ndarray:
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def uniform(seed: int, n: int, start: int = 0) -> np.ndarray:
    """n draws of uniform[-1,1), skipping the first ``start`` draws of the stream."""
    return 2.0 * _u01(seed, n, start) - 1.0


def normal(seed: int, n: int, start: int = 0) -> np.ndarray:
    """n draws of N(0,1) (Box-Muller), starting at uniform draw ``start``."""
    m = (n + 1) // 2
    u = _u01(seed, 2 * m, start)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log(1.0 - u1))
    a = 2.0 * np.pi * u2
    out = np.empty(2 * m, np.float64)
    out[0::2] = r * np.cos(a)
    out[1::2] = r * np.sin(a)
    return out[:n]


def gaussian_windowed_responses(seed: int, count: int, m: int, sigma: float, start: int) -> np.ndarray:
    """``count`` responses of length m: r_t = exp(-(t-c)^2 / (2 sigma^2)) * g_t with
    g ~ N(0,1) drawn from the stream at ``start``, c = (m-1)/2, each normalised to
    unit L1 norm.  Shape (count, m), float64."""
    g = normal(seed, count * m, start).reshape(count, m)
    t = np.arange(m, dtype=np.float64)
    c = (m - 1) / 2.0
    w = np.exp(-((t - c) ** 2) / (2.0 * sigma * sigma))
    r = g * w
    r /= np.abs(r).sum(axis=1, keepdims=True)
    return r


def uneven_shapes(seed: int, count: int, frac_1d: float = 0.8, log_1d=(8, 20), log_2d=(6, 11)) -> list:
    """Shapes of an uneven batch: each item is 1-D with probability frac_1d,
    log2 N uniform on log_1d (inclusive), else 2-D with each side's log2
    uniform on log_2d.  Drawn with numpy's PCG64 generator from ``seed``
    (shape draws only; values come from the SplitMix64 stream)."""
    rng = np.random.default_rng(seed)
    is1d = rng.random(count) < frac_1d
    l1 = rng.integers(log_1d[0], log_1d[1] + 1, count)
    a = rng.integers(log_2d[0], log_2d[1] + 1, count)
    b = rng.integers(log_2d[0], log_2d[1] + 1, count)
    return [(1 << int(l1[i]),) if is1d[i] else (1 << int(a[i]), 1 << int(b[i])) for i in range(count)]
