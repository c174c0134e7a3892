"""Seeded synthetic workloads for the phase-retrieval benchmarks and tests.

No data ships with the reference (its test_data/ is git-ignored), so every
workload is generated here, on the host, from fixed seeds, and handed to the
GPU and to the CPU oracle as identical bytes.

2D patterns (the shrinkwrap-HIO configurations):
  object   real, uniform[0, 1) inside a centred disc of radius r, 0 outside;
  I        |Forward2D(object)|^2 (centred DFT) scaled to `photons` total,
           Poisson-sampled, scaled back to the object's Fourier units;
  A        sqrt(I_noisy) (float32);
  x0       Backward2D(A * exp(i phi)), phi ~ U[0, 2 pi) per start (complex64);
  support  autocorrelation support: |Backward2D(I_noisy)| >= 0.04 * max (u8).
Seeds: numpy default_rng(1000 * config + pattern) for the object and noise,
default_rng(1000 * config + 500_000 + start) for the start phases.
"""
from __future__ import annotations

import numpy as np


def centred_fft2(x):
    return np.fft.fftshift(np.fft.fft2(np.fft.fftshift(x, axes=(-2, -1))), axes=(-2, -1))


def centred_ifft2(x):
    return np.fft.fftshift(np.fft.ifft2(np.fft.fftshift(x, axes=(-2, -1))), axes=(-2, -1))


def disc_object(n: int, radius: float, rng) -> np.ndarray:
    yy, xx = np.mgrid[0:n, 0:n]
    r2 = (yy - n / 2) ** 2 + (xx - n / 2) ** 2
    obj = rng.uniform(0.0, 1.0, size=(n, n))
    obj[r2 > radius * radius] = 0.0
    return obj


def pattern_from_object(obj: np.ndarray, photons: float, rng):
    I = np.abs(centred_fft2(obj)) ** 2
    s = photons / I.sum()
    counts = rng.poisson(I * s)
    In = counts / s
    return In


def autocorr_support(In: np.ndarray, frac: float = 0.04) -> np.ndarray:
    ac = np.abs(centred_ifft2(In))
    return (ac >= frac * ac.max(axis=(-2, -1), keepdims=True)).astype(np.uint8)


def start_from_amplitude(A: np.ndarray, rng) -> np.ndarray:
    phi = rng.uniform(0.0, 2 * np.pi, size=A.shape)
    return centred_ifft2(A * np.exp(1j * phi))


def make_2d_batch(n_patterns: int, n: int = 256, radius_lo: float = 35.0,
                  radius_hi: float = 45.0, photons: float = 1e6, starts_per_pattern: int = 1,
                  config: int = 1):
    """Returns dict(x0 complex64 [B,n,n], amp float32 [P,n,n], support u8 [B,n,n],
    obj float64 [P,n,n], amp_batch P) with B = P * starts_per_pattern; pattern b
    uses amp[b % P] (the reference's i % (num/batch) broadcast)."""
    P = n_patterns
    objs = np.zeros((P, n, n))
    amps = np.zeros((P, n, n))
    sups = np.zeros((P, n, n), dtype=np.uint8)
    for p in range(P):
        rng = np.random.default_rng(1000 * config + p)
        r = rng.uniform(radius_lo, radius_hi)
        obj = disc_object(n, r, rng)
        In = pattern_from_object(obj, photons, rng)
        objs[p] = obj
        amps[p] = np.sqrt(In)
        sups[p] = autocorr_support(In)
    B = P * starts_per_pattern
    x0 = np.zeros((B, n, n), dtype=np.complex128)
    support = np.zeros((B, n, n), dtype=np.uint8)
    for b in range(B):
        rng = np.random.default_rng(1000 * config + 500_000 + b)
        x0[b] = start_from_amplitude(amps[b % P], rng)
        support[b] = sups[b % P]
    return {
        "x0": x0.astype(np.complex64),
        "amp": amps.astype(np.float32),
        "support": support,
        "obj": objs,
        "amp_batch": P,
    }


def config0(n: int = 256):
    """8 random starts of one 256x256 pattern, disc radius 40 +- 5 px."""
    return make_2d_batch(1, n, 35.0, 45.0, 1e6, starts_per_pattern=8, config=0)


def config1(n_patterns: int = 512, n: int = 256):
    """512 patterns of 256x256, each its own object, radii U[30, 50] px."""
    return make_2d_batch(n_patterns, n, 30.0, 50.0, 1e6, starts_per_pattern=1, config=1)


HIO_SCHEDULE = dict(iters=200, beta=0.9, real_positive=1, dc_skip_zero=1, sw_every=20,
                    sw_sigma0=3.0, sw_sigma1=1.5, sw_frac=0.1)
