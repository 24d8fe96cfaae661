"""Writes tests/golden/golden.npz: inputs and expected outputs for the FFT path,
computed with numpy.fft (pocketfft) -- an implementation independent of the
oracle and of the HIP kernels.

Conventions pinned (numrs, FFT_1.rs:5-44 / Real_FT.rs:4-21):
  four1(+1) = N * ifft(x),  four1(-1) = fft(x)
  realft(+1) = packed [F_0, F_{n/2}, Re F_1, Im F_1, ...], F = n * ifft(x)
  fourn(+1) = prod(nn) * ifftn(x) (C order), fourn(-1) = fftn(x)
  convlv   = circular convolution with the Convolve.rs:41-63 placement
  correl   = circular correlation r[l] = sum_j d1[(j+l) mod n] d2[j] (n > 32)
             linear correlation at lags 0..n-1 (n <= 32, Correlation.rs:37-50)
  twofft   = (n * ifft(d1), n * ifft(d2)), bins 0..n-1 (FFT_2.rs:3-17)
  cosft1   = dct(y, type=1) / 2 on n + 1 points (Cos_FT.rs:7)
  cosft2   = dct(y, type=2) / 2 (isign 1), dct(y, type=3) / 2 (isign -1) (Cos_FT2.rs:7)
  rlft3    = prod(nn) * ifftn(x): data holds k3 < nn3/2, speq holds k3 = nn3/2
             (Real_FT3.rs:8)

Run: python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy.fft as sfft

HERE = os.path.dirname(os.path.abspath(__file__))


def rng(seed):
    return np.random.default_rng(seed)


def place_response(r, n):
    m = r.size
    mid = (m - 1) // 2
    out = np.zeros(n)
    for i in range(n):
        if i < mid:
            out[i] = r[m - mid + i]
        elif i < m:
            out[i] = r[i]
        elif i < n - mid:
            out[i] = 0.0
        else:
            out[i] = r[i - (n - mid)]
    return out


def pack_real_spectrum(x):
    n = x.size
    F = n * np.fft.ifft(x)
    out = np.empty(n)
    out[0] = F[0].real
    out[1] = F[n // 2].real
    out[2::2] = F[1:n // 2].real
    out[3::2] = F[1:n // 2].imag
    return out


def main():
    g = {}
    for nn in (8, 64, 1024):
        x = rng(nn).uniform(-1, 1, 2 * nn)
        z = x[0::2] + 1j * x[1::2]
        for isign, y in ((1, nn * np.fft.ifft(z)), (-1, np.fft.fft(z))):
            out = np.empty(2 * nn)
            out[0::2], out[1::2] = y.real, y.imag
            g[f"four1_{nn}_{isign}_in"] = x
            g[f"four1_{nn}_{isign}_out"] = out
    for n in (16, 256, 2048):
        x = rng(100 + n).normal(size=n)
        g[f"realft_{n}_in"] = x
        g[f"realft_{n}_out"] = pack_real_spectrum(x)
    shape = (4, 8, 16)
    x = rng(7).normal(size=2 * int(np.prod(shape)))
    z = (x[0::2] + 1j * x[1::2]).reshape(shape)
    for isign, y in ((1, np.prod(shape) * np.fft.ifftn(z)), (-1, np.fft.fftn(z))):
        out = np.empty(x.size)
        out[0::2], out[1::2] = y.reshape(-1).real, y.reshape(-1).imag
        g[f"fourn_4x8x16_{isign}_in"] = x
        g[f"fourn_4x8x16_{isign}_out"] = out
    n, m = 1024, 257
    d = rng(11).normal(size=n)
    t = np.arange(m)
    r = np.exp(-((t - 128.0) ** 2) / (2 * 40.0 ** 2)) * rng(12).normal(size=m)
    r /= np.abs(r).sum()
    p = place_response(r, n)
    conv = np.real(np.fft.ifft(np.fft.fft(d) * np.fft.fft(p)))
    g["convlv_1024_257_data"] = d
    g["convlv_1024_257_respns"] = r
    g["convlv_1024_257_placed"] = p
    g["convlv_1024_257_out"] = conv
    # deconvolution of the convolution recovers the data where |R|^2 >= 1e-12
    R = np.fft.fft(p)
    mask = np.abs(R) ** 2 >= 1e-12
    D = np.where(mask, np.fft.fft(conv) / np.where(mask, R, 1.0), 0.0)
    g["deconvlv_1024_257_out"] = np.real(np.fft.ifft(D))
    n = 64
    d1, d2 = rng(21).normal(size=n), rng(22).normal(size=n)
    g["correl_64_d1"], g["correl_64_d2"] = d1, d2
    g["correl_64_out"] = np.real(np.fft.ifft(np.fft.fft(d1) * np.conj(np.fft.fft(d2))))
    n = 16
    d1, d2 = rng(23).normal(size=n), rng(24).normal(size=n)
    g["correl_16_d1"], g["correl_16_d2"] = d1, d2
    g["correl_16_out"] = np.array([np.dot(d1[lag:], d2[:n - lag]) for lag in range(n)])
    n = 64
    d1, d2 = rng(31).normal(size=n), rng(32).normal(size=n)
    g["twofft_64_d1"], g["twofft_64_d2"] = d1, d2
    for name, d in (("fft1", d1), ("fft2", d2)):
        F = n * np.fft.ifft(d)
        out = np.empty(2 * n)
        out[0::2], out[1::2] = F.real, F.imag
        g[f"twofft_64_{name}"] = out
    y = rng(33).normal(size=n + 1)
    g["cosft1_64_in"], g["cosft1_64_out"] = y, 0.5 * sfft.dct(y, type=1)
    y = rng(34).normal(size=n)
    g["cosft2_64_in"] = y
    g["cosft2_64_1_out"] = 0.5 * sfft.dct(y, type=2)
    g["cosft2_64_-1_out"] = 0.5 * sfft.dct(y, type=3)
    shape = (4, 8, 16)
    x = rng(35).normal(size=shape)
    F = np.prod(shape) * np.fft.ifftn(x)
    data = np.empty(x.size)
    data[0::2] = F[:, :, :shape[2] // 2].reshape(-1).real
    data[1::2] = F[:, :, :shape[2] // 2].reshape(-1).imag
    speq = np.empty(2 * shape[0] * shape[1])
    speq[0::2] = F[:, :, shape[2] // 2].reshape(-1).real
    speq[1::2] = F[:, :, shape[2] // 2].reshape(-1).imag
    g["rlft3_4x8x16_in"], g["rlft3_4x8x16_data"], g["rlft3_4x8x16_speq"] = x.reshape(-1), data, speq
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **g)
    print(f"wrote {len(g)} arrays")


if __name__ == "__main__":
    main()
