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


def multistart_256(n: int = 256):
    """8 random starts of one 256x256 pattern, disc radius 40 +- 5 px."""
    return make_2d_batch(1, n, 35.0, 45.0, 1e6, starts_per_pattern=8, config=0)


def batch_512x256(n_patterns: int = 512, n: int = 256):
    """512 patterns of 256x256, each its own object, radii U[30, 50] px."""
    return make_2d_batch(n_patterns, n, 30.0, 50.0, 1e6, starts_per_pattern=1, config=1)


HIO_SCHEDULE = dict(iters=200, beta=0.9, real_positive=1, dc_skip_zero=1, sw_every=20,
                    sw_sigma0=3.0, sw_sigma1=1.5, sw_frac=0.1)


# ---- 3D: random ellipsoids, carousel projections -----------------------------
#
# Object: `count` ellipsoids with semi-axes U[5, 25] voxels, random orientation,
# densities U[0.2, 1.0], centres uniform in the ball of radius
# (sphere_radius - largest semi-axis) so each ellipsoid stays inside the
# sphere of radius `sphere_radius` about the volume centre c = n/2.  Volume
# layout [z][y][x] (index i at coordinate i - c).
# Projection p (angle th_p = p*pi/P about z) is an [n][n] image, rows u along
# z and columns v: the line integral of the object along
# d = (-sin th, cos th, 0) through (x, y, z) - c = ((v-c) cos th,
# (v-c) sin th, u-c), computed exactly per ellipsoid (chord length x density);
# by the Fourier-slice theorem its centred 2D DFT is the central slice of the
# object's centred 3D DFT sampled by the carousel merge.
# Seeds: default_rng(1000 * config) for the object, + 1 + p for the noise and
# start phases of projection p.

def random_ellipsoids(count: int = 12, sphere_radius: float = 64.0, axes=(5.0, 25.0),
                      dens=(0.2, 1.0), config: int = 2):
    rng = np.random.default_rng(1000 * config)
    out = []
    for _ in range(count):
        a = rng.uniform(axes[0], axes[1], 3)
        q, r = np.linalg.qr(rng.standard_normal((3, 3)))
        R = q * np.sign(np.diag(r))
        Q = R @ np.diag(1.0 / a ** 2) @ R.T
        lim = sphere_radius - a.max()
        while True:
            m = rng.uniform(-lim, lim, 3)
            if np.dot(m, m) <= lim * lim:
                break
        out.append((m, Q, rng.uniform(dens[0], dens[1])))
    return out


def ellipsoid_volume(n: int, ells) -> np.ndarray:
    c = n / 2
    vol = np.zeros((n, n, n))
    coord = np.arange(n) - c
    for m, Q, rho in ells:
        # bounding box of the ellipsoid: half-extent sqrt(Q^-1_ii) per axis
        ext = np.sqrt(np.diag(np.linalg.inv(Q))) + 1
        sl = []
        for ax in (2, 1, 0):  # z, y, x ranges (vol is [z][y][x], m is (x, y, z))
            lo = int(max(0, np.floor(m[ax] - ext[ax] + c)))
            hi = int(min(n, np.ceil(m[ax] + ext[ax] + c) + 1))
            sl.append((lo, hi))
        z = coord[sl[0][0]:sl[0][1]][:, None, None] - m[2]
        y = coord[sl[1][0]:sl[1][1]][None, :, None] - m[1]
        x = coord[sl[2][0]:sl[2][1]][None, None, :] - m[0]
        qf = (Q[0, 0] * x * x + Q[1, 1] * y * y + Q[2, 2] * z * z
              + 2 * (Q[0, 1] * x * y + Q[0, 2] * x * z + Q[1, 2] * y * z))
        vol[sl[0][0]:sl[0][1], sl[1][0]:sl[1][1], sl[2][0]:sl[2][1]] += rho * (qf <= 1.0)
    return vol


def ellipsoid_projections(n: int, angles, ells) -> np.ndarray:
    c = n / 2
    th = np.asarray(angles, dtype=np.float64)[:, None, None]
    u = (np.arange(n) - c)[None, :, None]
    v = (np.arange(n) - c)[None, None, :]
    cs, sn = np.cos(th), np.sin(th)
    d = (-sn, cs, 0.0)
    out = np.zeros((len(angles), n, n))
    for m, Q, rho in ells:
        w = ((v * cs - m[0]) + 0 * u, (v * sn - m[1]) + 0 * u, (u - m[2]) + 0 * v)
        Qd = [Q[i, 0] * d[0] + Q[i, 1] * d[1] + Q[i, 2] * d[2] for i in range(3)]
        a = d[0] * Qd[0] + d[1] * Qd[1] + d[2] * Qd[2]
        b = w[0] * Qd[0] + w[1] * Qd[1] + w[2] * Qd[2]
        Qw = [Q[i, 0] * w[0] + Q[i, 1] * w[1] + Q[i, 2] * w[2] for i in range(3)]
        cc = w[0] * Qw[0] + w[1] * Qw[1] + w[2] * Qw[2] - 1.0
        disc = b * b - a * cc
        out += rho * 2.0 * np.sqrt(np.maximum(disc, 0.0)) / a
    return out


def make_carousel(n: int = 128, P: int = 180, count: int = 12, photons: float = 1e7,
                  config: int = 2, sphere_radius: float = None):
    """The 3D carousel workload: dict(obj [n][n][n] float64, angles [P] float32,
    proj [P][n][n] float64, amp [P][n][n] float32, x0 [P][n][n] complex64,
    support2d u8 [P][n][n] (disc of radius n/2: the shadow of the ball),
    support3d u8 [n][n][n] (ball of radius n/2))."""
    R = sphere_radius if sphere_radius is not None else n / 2
    ells = random_ellipsoids(count, R, config=config)
    angles = (np.arange(P) * np.pi / P).astype(np.float32)
    proj = ellipsoid_projections(n, angles.astype(np.float64), ells)
    amp = np.zeros((P, n, n))
    x0 = np.zeros((P, n, n), dtype=np.complex128)
    for p in range(P):
        rng = np.random.default_rng(1000 * config + 1 + p)
        In = pattern_from_object(proj[p], photons, rng)
        amp[p] = np.sqrt(In)
        x0[p] = start_from_amplitude(amp[p], rng)
    c = n / 2
    k = np.arange(n) - c
    disc = ((k[:, None] ** 2 + k[None, :] ** 2) <= c * c).astype(np.uint8)
    ball = ((k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2) <= c * c)
    return {
        "obj": ellipsoid_volume(n, ells),
        "angles": angles,
        "proj": proj,
        "amp": amp.astype(np.float32),
        "x0": x0.astype(np.complex64),
        "support2d": np.ascontiguousarray(np.broadcast_to(disc, (P, n, n))),
        "support3d": ball.astype(np.uint8),
    }


def ellipsoid_amplitude_volume(n: int = 256, count: int = 12, config: int = 4,
                               sphere_radius: float = 64.0):
    """A 3D object of random ellipsoids (as the carousel's) in an n^3 volume and
    its exact centred 3D DFT magnitude (float32): the 3D amplitude-loss case."""
    ells = random_ellipsoids(count, sphere_radius, config=config)
    obj = ellipsoid_volume(n, ells)
    A = np.abs(np.fft.fftshift(np.fft.fftn(np.fft.fftshift(obj)))).astype(np.float32)
    return obj, A
