"""The CPU baseline: the reference's MKL backend restated in FP32 over
libmkl_rt (oracle/mkl_path.c), checked before it is timed beside the GPU.

Pinned against numpy's pocketfft, the reference's own known answers
(test/unitest_mkl.cpp:15-71) and the FP64 oracle.  Skipped where libmkl_rt
cannot be loaded (bench.py then falls back to torch's CPU FFT, labelled as
such)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mkl_path as mp  # noqa: E402

pytestmark = pytest.mark.skipif(not mp.available(), reason="libmkl_rt not loadable here")


def cfft2(x):
    return np.fft.fftshift(np.fft.fft2(np.fft.fftshift(x, axes=(-2, -1))), axes=(-2, -1))


def rnd(shape, seed):
    r = np.random.default_rng(seed)
    return (r.standard_normal(shape) + 1j * r.standard_normal(shape)).astype(np.complex64)


@pytest.mark.parametrize("M,N,B", [(16, 16, 3), (64, 128, 2), (256, 256, 2), (10, 10, 1)])
def test_forward_backward_vs_pocketfft(M, N, B):
    x = rnd((B, M, N), M + N + B)
    p = mp.Plan(M, N, B)
    y = p.forward2d(x)
    ref = cfft2(x.astype(np.complex128))
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-6
    xb = p.backward2d(y)
    assert np.linalg.norm(xb - x) / np.linalg.norm(x) < 1e-6


def test_reference_known_answer_2d_dft():
    # test/unitest_mkl.cpp:42-71 (in double there): 10x10 ones -> 100 at (5, 5)
    p = mp.Plan(10, 10, 1)
    y = p.forward2d(np.ones((1, 10, 10), np.complex64))
    expect = np.zeros((10, 10))
    expect[5, 5] = 100.0
    assert np.abs(y[0] - expect).max() < 1e-4
    assert np.abs(p.backward2d(y)[0] - 1.0).max() < 1e-6


def test_fftshift_is_numpy_fftshift():
    x = rnd((3, 8, 12), 5)
    got = mp.fftshift2d(x, 8, 12, 3)
    assert np.array_equal(got, np.fft.fftshift(x, axes=(-2, -1)))


def test_data_constraint_semantics(oracle):
    """MKL_Impl.inl:86-106: A == 0 skipped, x == 0 kept, else polar(A, arg x);
    the same as the oracle's FP32 restatement, with A broadcast over the
    batch."""
    x = rnd((2, 6, 6), 6)
    x[0, 0, 0] = 0
    A = np.abs(rnd((1, 6, 6), 7).real) + 0.5
    A[0, 1, 1] = 0.0
    got = mp.data_constraint(x, A, 2)
    ref, _ = oracle.data_constraint_f32(x, A, 2)
    assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max()
    assert got[0, 0, 0] == 0 and got[1, 1, 1] == x[1, 1, 1] and got[0, 1, 1] == x[0, 1, 1]
    np.testing.assert_allclose(np.abs(got[:, 2:, 2:]), np.broadcast_to(A[:, 2:, 2:], (2, 4, 4)), rtol=1e-6)


def test_merge_add_reference_unit_test():
    # test/unitest_mkl.cpp:15-40: y = 2 x + 3 y on ones -> 5; alpha 0 is a no-op
    x = np.ones(100, np.complex64)
    y = np.ones(100, np.complex64)
    out, ok = mp.merge_add(x, y, 2.0, 3.0)
    assert ok and np.all(out == 5.0)
    out, ok = mp.merge_add(x, y, 0.0, 3.0)
    assert not ok and np.all(out == 1.0)


def test_hio_steps_match_fp64_oracle(oracle):
    """Each HIO step (complex and real-positive) within FP32 rounding of the
    FP64 oracle's step from the same state."""
    n, B = 32, 2
    rng = np.random.default_rng(8)
    A = np.abs(rnd((B, n, n), 9)).astype(np.float32) + 0.1
    sup = np.zeros((B, n, n), np.uint8)
    sup[:, 8:24, 8:24] = 1
    x = rnd((B, n, n), 10)
    p = mp.Plan(n, n, B)
    for rp in (False, True):
        got, s = p.hio(x, A, sup, 1, beta=0.9, real_positive=rp)
        ref = oracle.hio_step(x, A, sup, 0.9, rp, True)
        pm = oracle.hio_step(x, A, np.ones_like(sup), 0.0, False, True)
        near = rp & (sup > 0) & (np.abs(pm.real) < 1e-5 * np.abs(pm).max())
        err = np.linalg.norm((got - ref)[~near]) / np.linalg.norm(ref[~near])
        assert err < 1e-5, (rp, err)
        assert np.array_equal(s, sup)
    del rng


def test_hio_shrinkwrap_schedule_matches_fp64_oracle(oracle):
    """A short schedule with shrinkwrap updates after iterations 2 and 4
    (sigma 3.0 then 1.5): the support masks and the iterate follow the FP64
    oracle's run (complex mode, where HIO is continuous)."""
    n, B = 64, 2
    yy, xx = np.mgrid[0:n, 0:n]
    disc = ((yy - n / 2) ** 2 + (xx - n / 2) ** 2 <= 100).astype(float)
    rng = np.random.default_rng(11)
    A = np.abs(cfft2(rng.random((B, n, n)) * disc)).astype(np.float32)
    sup = np.zeros((B, n, n), np.uint8)
    sup[:, 16:48, 16:48] = 1
    x = (np.fft.ifft2(A * np.exp(2j * np.pi * rng.random((B, n, n))))).astype(np.complex64)
    p = mp.Plan(n, n, B)
    got, s = p.hio(x, A, sup, 4, beta=0.9, real_positive=False, sw_every=2)
    ref, sr = oracle.hio(x, A, sup, 4, 0.9, False, True, 2, 3.0, 1.5, 0.1)
    assert (s != sr).mean() < 1e-3
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-4
