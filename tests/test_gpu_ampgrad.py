"""GPU parity of the amplitude loss L_b = sum_k (|F x_b|_k - A_k)^2, its adjoint
(cpra_hip_amp_grad / _3d) and the fused gradient descent (cpra_hip_amp_gd_run
/ _3d) against the FP64 oracle (orc_amp_loss / orc_amp_grad, which
tests/test_oracle.py pins by finite differences and by the zero gradient at an
exact solution), and against central finite differences (h = 1e-6) of the
library's own FP64 loss.  The reference has no loss or gradient; this is the
differentiable-retrieval workload (256 x 256^2 and a 256^3 volume under
gradient descent with step 0.5/max(I)) at test sizes."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def centred_fft2(x):
    return np.fft.fftshift(np.fft.fft2(np.fft.fftshift(x, axes=(-2, -1))), axes=(-2, -1))


def disc_problem(B, n, seed, noise=0.05, P=None):
    """Objects (discs of U[0,1) density), exact amplitudes of the first P of
    them, and starts = object + complex Gaussian noise."""
    P = P or B
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[:n, :n]
    obj = np.zeros((P, n, n))
    for p in range(P):
        r = rng.uniform(0.12, 0.2) * n
        obj[p] = ((yy - n / 2) ** 2 + (xx - n / 2) ** 2 <= r * r) * rng.random((n, n))
    A = np.abs(centred_fft2(obj))
    x = obj[np.arange(B) % P] + noise * (rng.standard_normal((B, n, n)) +
                                        1j * rng.standard_normal((B, n, n)))
    return x, A


def ellipsoid_volume(n, seed):
    rng = np.random.default_rng(seed)
    z, y, x = np.mgrid[:n, :n, :n] - n / 2
    vol = np.zeros((n, n, n))
    for _ in range(4):
        c = rng.uniform(-n / 8, n / 8, 3)
        a = rng.uniform(n / 16, n / 6, 3)
        vol += ((((z - c[0]) / a[0]) ** 2 + ((y - c[1]) / a[1]) ** 2 + ((x - c[2]) / a[2]) ** 2) <= 1) \
            * rng.uniform(0.2, 1.0)
    return vol


def centred_fftn(v):
    return np.fft.fftshift(np.fft.fftn(np.fft.fftshift(v)))


# ------------------------------------------------------------ loss, adjoint --

def test_loss_and_grad_f64_match_oracle(hipmod, oracle, gpu):
    B, n = 16, 64
    x, A = disc_problem(B, n, 1, P=4)
    xd, ad = hipmod.to_device(x.astype(np.complex128)), hipmod.to_device(A)
    g = hipmod.zeros((B, n, n), np.complex128)
    with hipmod.HipImpl(n, n, 1, B, dtype="float64") as h:
        L = h.amp_loss(xd, ad)
        L2 = h.amp_grad(xd, ad, g)
    Lr = oracle.amp_loss(x, A)
    gr = oracle.amp_grad(x, A)
    assert np.abs(L - Lr).max() <= 1e-12 * np.abs(Lr).max()
    assert np.abs(L2 - Lr).max() <= 1e-12 * np.abs(Lr).max()
    assert relerr(g.numpy(), gr) < 1e-11
    assert np.array_equal(xd.numpy(), x)                     # x untouched


def test_grad_f64_central_differences(hipmod, gpu):
    """16 x 64^2 FP64, h = 1e-6: every sampled partial derivative within 1e-6
    of the pattern's largest gradient component, and 1e-6 relative over the
    sampled set.  A coordinate is perturbed in all 16 patterns at once (the
    patterns are independent), so one pair of loss calls yields 16 FDs."""
    B, n, hstep = 16, 64, 1e-6
    x, A = disc_problem(B, n, 2)
    ad = hipmod.to_device(A)
    g = hipmod.zeros((B, n, n), np.complex128)
    rng = np.random.default_rng(3)
    fds, ans = [], []
    with hipmod.HipImpl(n, n, 1, B, dtype="float64") as h:
        h.amp_grad(hipmod.to_device(x), ad, g)
        gn = g.numpy()
        scale = np.abs(gn).reshape(B, -1).max(axis=1)
        for _ in range(16):
            r, c = rng.integers(n), rng.integers(n)
            for unit in (1.0, 1j):
                xp, xm = x.copy(), x.copy()
                xp[:, r, c] += hstep * unit
                xm[:, r, c] -= hstep * unit
                fd = (h.amp_loss(hipmod.to_device(xp), ad) - h.amp_loss(hipmod.to_device(xm), ad)) \
                    / (2 * hstep)
                an = gn[:, r, c].real if unit == 1.0 else gn[:, r, c].imag
                assert (np.abs(fd - an) <= 1e-6 * scale).all(), (r, c, unit, fd, an)
                fds.append(fd)
                ans.append(an)
    assert relerr(np.array(fds), np.array(ans)) <= 1e-6


def test_grad_zero_at_exact_solution(hipmod, gpu):
    B, n = 2, 128
    x, A = disc_problem(B, n, 4, noise=0.0)
    g = hipmod.zeros((B, n, n), np.complex128)
    with hipmod.HipImpl(n, n, 1, B, dtype="float64") as h:
        L = h.amp_grad(hipmod.to_device(x.astype(np.complex128)), hipmod.to_device(A), g)
    assert np.abs(g.numpy()).max() < 1e-9 * np.abs(x).max() * n * n
    assert (L < 1e-16 * (A ** 2).sum(axis=(1, 2))).all()


@pytest.mark.parametrize("n,B,P", [(64, 4, 2), (192, 2, 1), (256, 4, 4), (96, 2, 2)])
def test_loss_and_grad_f32_match_oracle(hipmod, oracle, gpu, n, B, P):
    rng = np.random.default_rng(n)
    A = np.abs(rng.standard_normal((P, n, n))) * 3.0
    x = rng.standard_normal((B, n, n)) + 1j * rng.standard_normal((B, n, n))
    xd = hipmod.to_device(x.astype(np.complex64))
    ad = hipmod.to_device(A.astype(np.float32))
    g = hipmod.zeros((B, n, n), np.complex64)
    with hipmod.HipImpl(n, n, 1, B) as h:
        L = h.amp_grad(xd, ad, g)
        L1 = h.amp_loss(xd, ad)
    x32 = x.astype(np.complex64).astype(complex)
    A32 = A.astype(np.float32).astype(float)
    Lr = oracle.amp_loss(x32, A32)
    assert np.abs(L - Lr).max() <= 1e-5 * np.abs(Lr).max()
    assert np.abs(L1 - L).max() <= 1e-6 * np.abs(Lr).max()
    assert relerr(g.numpy(), oracle.amp_grad(x32, A32)) < 1e-5


@pytest.mark.parametrize("n,dtype,tol", [(64, "float32", 1e-5), (32, "float64", 1e-11),
                                         (128, "float32", 1e-5), (256, "float32", 1e-5)])
def test_loss_and_grad_3d_match_oracle(hipmod, oracle, gpu, n, dtype, tol):
    vol = ellipsoid_volume(n, 5)
    A = np.abs(centred_fftn(vol))
    rng = np.random.default_rng(6)
    x = vol + 0.05 * (rng.standard_normal(vol.shape) + 1j * rng.standard_normal(vol.shape))
    cdt, rdt = (np.complex64, np.float32) if dtype == "float32" else (np.complex128, np.float64)
    xd, ad = hipmod.to_device(x[None].astype(cdt)), hipmod.to_device(A[None].astype(rdt))
    g = hipmod.zeros((1, n, n, n), cdt)
    with hipmod.HipImpl(n, n, n, 1, dtype=dtype) as h:
        L = h.amp_grad(xd, ad, g, volume=True)
        L1 = h.amp_loss(xd, ad, volume=True)
    xr = x.astype(cdt).astype(complex)
    Ar = A.astype(rdt).astype(float)
    Lr = oracle.amp_loss_3d(xr, Ar)
    assert abs(L[0] - Lr) <= tol * abs(Lr)
    assert abs(L1[0] - Lr) <= tol * abs(Lr)
    assert relerr(g.numpy()[0], oracle.amp_grad_3d(xr, Ar)) < tol * 10


def test_grad_argument_errors(hipmod, gpu):
    n, B = 64, 2
    x = hipmod.zeros((B, n, n), np.complex64)
    a = hipmod.zeros((1, n, n), np.float32)
    with hipmod.HipImpl(n, n, 1, B) as h:
        with pytest.raises(hipmod.CpraHipError):
            h.amp_grad(x, a, x)                      # grad aliases x
        with pytest.raises(hipmod.CpraHipError):
            h.amp_loss(x, a, amp_batch=3)            # does not divide the batch
    with hipmod.HipImpl(n, n, 1, B, dtype="float64") as h:
        xd = hipmod.zeros((B, n, n), np.complex128)
        with pytest.raises(hipmod.CpraHipError):
            h.amp_gd_run(xd, a, 1, 1e-3)             # fused descent is FP32
    with hipmod.HipImpl(66, 64, 1, 1) as h:
        with pytest.raises(hipmod.CpraHipError):
            h.amp_gd_run(hipmod.zeros((1, 66, 64), np.complex64),
                         hipmod.zeros((1, 66, 64), np.float32), 1, 1e-3)   # outer 66


# ------------------------------------------------------- gradient descent ---

def oracle_gd(oracle, x, A, mu, steps, volume=False):
    x = x.astype(complex)
    for _ in range(steps):
        x = x - mu * (oracle.amp_grad_3d(x, A) if volume else oracle.amp_grad(x, A))
    return x


@pytest.mark.parametrize("n,B,P,steps", [(64, 8, 4, 25), (256, 2, 2, 10), (192, 2, 1, 6),
                                         (128, 3, 3, 0)])
def test_fused_gd_matches_oracle_loop(hipmod, oracle, gpu, n, B, P, steps):
    x, A = disc_problem(B, n, 10 + n, noise=0.3, P=P)
    A32 = A.astype(np.float32)
    mu = 0.5 / float((A32.astype(float) ** 2).max())     # step 0.5 / max(I)
    x32 = x.astype(np.complex64)
    xd = hipmod.to_device(x32)
    with hipmod.HipImpl(n, n, 1, B) as h:
        L = h.amp_gd_run(xd, hipmod.to_device(A32), steps, mu)
    ref = oracle_gd(oracle, x32, A32.astype(float), np.float32(mu), steps)
    assert relerr(xd.numpy(), ref) < 2e-6 * max(steps, 1) ** 0.5 + 1e-6
    Lr = oracle.amp_loss(ref, A32.astype(float))
    assert np.abs(L - Lr).max() <= 1e-4 * np.abs(Lr).max()


def test_fused_gd_decreases_loss(hipmod, gpu):
    n, B = 128, 4
    x, A = disc_problem(B, n, 21, noise=0.5)
    A32 = A.astype(np.float32)
    mu = 0.5 / float((A32.astype(float) ** 2).max())
    xd, ad = hipmod.to_device(x.astype(np.complex64)), hipmod.to_device(A32)
    with hipmod.HipImpl(n, n, 1, B) as h:
        hist = [h.amp_gd_run(xd, ad, 0, mu)]
        for _ in range(5):
            hist.append(h.amp_gd_run(xd, ad, 20, mu))
    hist = np.array(hist)
    assert (np.diff(hist, axis=0) < 0).all(), hist


@pytest.mark.parametrize("n,steps", [(64, 5), (128, 2), (256, 3)])
def test_fused_gd_3d_matches_oracle_loop(hipmod, oracle, gpu, n, steps):
    vol = ellipsoid_volume(n, 30 + n)
    A32 = np.abs(centred_fftn(vol)).astype(np.float32)
    rng = np.random.default_rng(31)
    x32 = (vol + 0.3 * rng.standard_normal(vol.shape)).astype(np.complex64)
    mu = 0.5 / float((A32.astype(float) ** 2).max())
    xd = hipmod.to_device(x32[None])
    with hipmod.HipImpl(n, n, n, 1) as h:
        L = h.amp_gd_run(xd, hipmod.to_device(A32[None]), steps, mu, volume=True)
    ref = oracle_gd(oracle, x32, A32.astype(float), np.float32(mu), steps, volume=True)
    assert relerr(xd.numpy()[0], ref) < 3e-6 * steps ** 0.5 + 1e-6
    Lr = oracle.amp_loss_3d(ref, A32.astype(float))
    assert abs(L[0] - Lr) <= 1e-4 * abs(Lr)


def test_torch_autograd_harness(gpu):
    """torch.autograd.Function over the C ABI (FP64 gradcheck, 16 x 64^2 and a
    32^3 volume), in a child process that imports torch."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "amp_autograd_harness.py"), ROOT],
                       capture_output=True, text=True, timeout=600)
    print(r.stdout, r.stderr[-3000:])
    assert r.returncode == 0, r.stdout + r.stderr[-3000:]


@pytest.mark.parametrize("n", [64, 256, 384, 512])
def test_fused_gd_tiled_state_matches_row_major(hipmod, gpu, n):
    """The 2D descent keeps its state tile-major by default (square patterns);
    the tuning key tiled=0 keeps it row-major in x.  Same arithmetic per
    column, so the two agree to FP32 rounding (the entry and exit row passes
    are separate kernels), the loss included."""
    B = 2
    x, A = disc_problem(B, n, 40 + n, noise=0.3, P=B)
    A32 = A.astype(np.float32)
    mu = 0.5 / float((A32.astype(float) ** 2).max())
    outs = []
    for tl in (0, 1):
        xd = hipmod.to_device(x.astype(np.complex64))
        with hipmod.HipImpl(n, n, 1, B) as h:
            assert h.set_tuning(f"tiled={tl};")
            L = h.amp_gd_run(xd, hipmod.to_device(A32), 8, mu)
        outs.append((xd.numpy(), L))
    assert relerr(outs[1][0], outs[0][0]) < 2e-6
    assert np.abs(outs[1][1] - outs[0][1]).max() <= 1e-5 * np.abs(outs[0][1]).max()
