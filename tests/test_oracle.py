"""Pins of the CPU oracle (runs without a GPU).

Each oracle function is checked against something it does not share code
with: numpy's pocketfft, the reference's own known-answer tests
(test/unitest_mkl.cpp), brute-force sums on tiny inputs, closed forms and
invariants (fixed points, adjoint identities, finite differences) chosen so
that a dropped term, a wrong sign or index, or a transposed operand fails.
"""
import numpy as np
import pytest


def cfft2(x):
    return np.fft.fftshift(np.fft.fft2(np.fft.fftshift(x, axes=(-2, -1))), axes=(-2, -1))


def cifft2(x):
    return np.fft.fftshift(np.fft.ifft2(np.fft.fftshift(x, axes=(-2, -1))), axes=(-2, -1))


def rnd(shape, seed=0, cplx=True):
    r = np.random.default_rng(seed)
    a = r.standard_normal(shape)
    return a + 1j * r.standard_normal(shape) if cplx else a


# ---------------------------------------------------------------- FFT ---------

@pytest.mark.parametrize("M,N,B", [(8, 8, 1), (10, 10, 2), (12, 12, 1), (6, 10, 3), (16, 16, 2),
                                   (64, 128, 1), (192, 384, 1), (14, 22, 1)])
def test_forward_backward_2d_vs_pocketfft(oracle, M, N, B):
    x = rnd((B, M, N), seed=M * N + B)
    y = oracle.forward2d(x)
    ref = cfft2(x)
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()
    xb = oracle.backward2d(ref)
    assert np.abs(xb - x).max() <= 1e-12 * np.abs(x).max()


def test_reference_known_answer_2d_dft(oracle):
    # test/unitest_mkl.cpp:42-71: double 10x10 ones -> 100 at (5,5), 0 elsewhere
    x = np.ones((10, 10), dtype=np.complex128)
    y = oracle.forward2d(x)
    expect = np.zeros((10, 10))
    expect[5, 5] = 100.0
    assert np.abs(y.real - expect).max() < 1e-14 * 100  # the reference bound is absolute 1e-14
    assert np.abs(y.imag).max() < 1e-12
    assert abs(y[5, 5].real - 100.0) < 1e-13
    xb = oracle.backward2d(y)
    assert np.abs(xb.real - 1.0).max() < 1e-14
    assert np.abs(xb.imag).max() < 1e-14


def test_fft_brute_force_tiny(oracle):
    M, N = 4, 6
    x = rnd((M, N), seed=3)
    # shift -> direct double sum -> shift, with Python loops
    s = np.roll(np.roll(x, -M // 2, 0), -N // 2, 1)
    d = np.zeros((M, N), dtype=complex)
    for kr in range(M):
        for kc in range(N):
            acc = 0j
            for r in range(M):
                for c in range(N):
                    acc += s[r, c] * np.exp(-2j * np.pi * (r * kr / M + c * kc / N))
            d[kr, kc] = acc
    d = np.roll(np.roll(d, -M // 2, 0), -N // 2, 1)
    assert np.allclose(oracle.forward2d(x), d, atol=1e-12)


@pytest.mark.parametrize("L,M,N", [(6, 6, 6), (4, 6, 8), (8, 8, 8)])
def test_forward_backward_3d_vs_pocketfft(oracle, L, M, N):
    x = rnd((2, L, M, N), seed=L + M + N)
    y = oracle.forward3d(x)
    ref = np.fft.fftshift(np.fft.fftn(np.fft.fftshift(x, axes=(1, 2, 3)), axes=(1, 2, 3)),
                          axes=(1, 2, 3))
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.abs(oracle.backward3d(y) - x).max() <= 1e-12 * np.abs(x).max()


@pytest.mark.parametrize("n", [2, 3, 5, 7, 11, 13, 49, 60, 97, 120])
def test_fft_lines_any_length(oracle, n):
    x = rnd((3, n), seed=n)
    y = oracle.fft_lines(x, n, 1, n, 3, -1, 1.0)
    assert np.allclose(y, np.fft.fft(x, axis=1), atol=1e-11 * n)
    # strided: transform along axis 0 of a [n][3] array
    xt = np.ascontiguousarray(x.T)
    yt = oracle.fft_lines(xt, n, 3, 1, 3, +1, 1.0 / n)
    assert np.allclose(yt, np.fft.ifft(xt, axis=0), atol=1e-12 * n)


def test_parseval_and_linearity(oracle):
    a, b = rnd((32, 32), 1), rnd((32, 32), 2)
    fa, fb = oracle.forward2d(a), oracle.forward2d(b)
    assert np.isclose(np.vdot(fa, fa).real, 32 * 32 * np.vdot(a, a).real, rtol=1e-12)
    assert np.allclose(oracle.forward2d(2 * a - 3j * b), 2 * fa - 3j * fb, atol=1e-10)


# ---------------------------------------------------------------- shifts -------

@pytest.mark.parametrize("M,N,B", [(4, 6, 2), (10, 10, 1), (8, 12, 3)])
def test_fftshift2d_is_numpy_fftshift(oracle, M, N, B):
    x = np.arange(B * M * N, dtype=np.float64).reshape(B, M, N).astype(np.complex64)
    y = oracle.fftshift2d(x, M, N, B)
    assert np.array_equal(y, np.fft.fftshift(x, axes=(1, 2)))
    assert np.array_equal(oracle.fftshift2d(y, M, N, B), x)


def test_fftshift3d_is_numpy_fftshift(oracle):
    L, M, N, B = 4, 6, 8, 2
    x = np.arange(B * L * M * N, dtype=np.float64).reshape(B, L, M, N)
    y = oracle.fftshift3d(x, M, N, L, B)
    assert np.array_equal(y, np.fft.fftshift(x, axes=(1, 2, 3)))


# ---------------------------------------------------------------- pointwise ----

def test_space_constraint_mkl_known_answers(oracle):
    x = np.array([1 + 2j, -3 + 4j, 5 - 6j, 7 + 8j, -1 - 1j, 2 + 2j], dtype=np.complex64)
    nxt = np.nextafter(np.float32(0.9), np.float32(1.0))
    S = np.array([0.0, np.float32(0.9), nxt, np.nan, 1.0, 0.5], dtype=np.float32)
    y, ok = oracle.space_constraint_f32(x, S, 1)
    assert ok
    # (double)0.9f < 0.9 -> zeroed; the next float above is kept; NaN is kept;
    # negative real parts are kept (no positivity in the MKL path)
    expect = np.array([0, 0, 5, 7, -1, 0], dtype=np.complex64)
    assert np.array_equal(y, expect)
    # broadcast: S has num/batch entries reused per batch member
    x2 = np.tile(x, 3)
    y2, _ = oracle.space_constraint_f32(x2, S, 3)
    assert np.array_equal(y2, np.tile(expect, 3))
    assert oracle.space_constraint_f32(x[:0], S, 1)[1] is False


def test_space_constraint_positive_known_answers(oracle):
    x = np.array([1 + 2j, -3 + 4j, 5 - 6j], dtype=np.complex64)
    S = np.array([0.5, 1.0, 0.0], dtype=np.float32)
    y, ok = oracle.space_constraint_pos_f32(x, S, 1)
    assert np.array_equal(y, np.array([0.5, 0, 0], dtype=np.complex64))


def test_data_constraint_real(oracle):
    x = np.array([3 + 4j, 0j, -1 + 0j, 1j, 2 + 2j], dtype=np.complex64)
    A = np.array([10.0, 5.0, 2.0, 0.0, 1.0], dtype=np.float32)
    y, ok = oracle.data_constraint_f32(x, A, 1, skip_zero=True)
    assert ok
    assert np.allclose(y, [6 + 8j, 0, -2, 1j, (1 + 1j) / np.sqrt(2)], atol=1e-6)
    y2, _ = oracle.data_constraint_f32(x, A, 1, skip_zero=False)
    assert y2[3] == 0  # the CUDA kernel sets A == 0 pixels to 0
    # magnitude = A and phase preserved
    xr = rnd(50, 4).astype(np.complex64)
    Ar = np.abs(rnd(50, 5, cplx=False)).astype(np.float32) + 0.1
    yr, _ = oracle.data_constraint_f32(xr, Ar, 1)
    assert np.allclose(np.abs(yr), Ar, rtol=1e-6)
    assert np.allclose(np.angle(yr), np.angle(xr), atol=1e-5)


def test_data_constraint_complex(oracle):
    x = np.array([1 + 1j, 2 + 2j, 3 + 3j], dtype=np.complex64)
    Cc = np.array([0j, 5 - 1j], dtype=np.complex64)
    y, ok = oracle.data_constraint_cplx_f32(np.tile(x[:2], 2), Cc, 2)
    assert np.array_equal(y, np.array([1 + 1j, 5 - 1j, 1 + 1j, 5 - 1j], dtype=np.complex64))


def test_merge_norm_reference_unit_test(oracle):
    # test/unitest_mkl.cpp:15-40
    x = np.full(1000, 10 + 10j, dtype=np.complex64)
    y, ok = oracle.merge_add_f32(x, x, 0.5, 0.5)
    assert ok and np.array_equal(y, x)
    z, ok = oracle.normalize_f32(y, 10.0)
    assert ok and np.array_equal(z, np.full(1000, 1 + 1j, dtype=np.complex64))
    w, ok = oracle.merge_add_f32(x, y, 0.0, 3.0)
    assert not ok and np.array_equal(w, y)  # alpha == 0 is a no-op (MKL_Impl.inl:34)


# ---------------------------------------------------------------- compositions --

def exact_problem(M=16, N=16, seed=0, r=4.0):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:M, 0:N]
    sup = ((yy - M / 2) ** 2 + (xx - N / 2) ** 2 <= r * r).astype(np.uint8)
    obj = rng.uniform(0.1, 1.0, (M, N)) * sup
    A = np.abs(cfft2(obj))
    return obj, A, sup


def pm(z, A, skip=True):
    f = cfft2(z)
    m = np.abs(f)
    out = np.where(m > 0, A * f / np.where(m > 0, m, 1), f)
    if skip:
        out = np.where(A == 0, f, out)
    return cifft2(out)


def test_dm_ref_fixed_point(oracle):
    obj, A, sup = exact_problem()
    S = sup.astype(np.float64)  # 1 inside (>= 0.9 keeps), 0 outside
    psi = obj[None].astype(complex)
    p, t1, t2 = oracle.dm_ref(psi, psi, psi, A[None], S[None], 0.9, 5)
    assert np.abs(p - psi).max() < 1e-10
    assert np.abs(t1 - psi).max() < 1e-10 and np.abs(t2 - psi).max() < 1e-10


def test_dm_ref_first_step_closed_form_beta1(oracle):
    obj, A, sup = exact_problem(seed=1)
    S = sup.astype(np.float64)
    psi = rnd((16, 16), 9)
    p, t1, t2 = oracle.dm_ref(psi[None], psi[None], psi[None], A[None], S[None], 1.0, 1)
    PS = lambda z: np.where(sup > 0, z.real, 0.0) + 0j  # noqa: E731
    e_t1 = PS(2 * pm(psi, A) - psi)
    e_t2 = pm(psi, A)
    assert np.allclose(t1[0], e_t1, atol=1e-10)
    assert np.allclose(t2[0], e_t2, atol=1e-10)
    assert np.allclose(p[0], psi + e_t1 - e_t2, atol=1e-10)


def test_dm_ref_carries_temporaries(oracle):
    # t1, t2 carried across iterations (src/CPRA_MKL.cpp:79-80 outside the loop):
    # after 2 iterations the reference variant differs from the textbook map
    obj, A, sup = exact_problem(seed=2)
    S = sup.astype(np.float64)
    psi = rnd((1, 16, 16), 10)
    p_ref, _, _ = oracle.dm_ref(psi, psi, psi, A[None], S[None], 0.9, 2)
    p_tb = oracle.dm_textbook(psi, A[None], 0.9, 2, S=S[None], space_mode=0)
    p_ref1, _, _ = oracle.dm_ref(psi, psi, psi, A[None], S[None], 0.9, 1)
    p_tb1 = oracle.dm_textbook(psi, A[None], 0.9, 1, S=S[None], space_mode=0)
    assert np.allclose(p_ref1, p_tb1, atol=1e-12)  # identical first step
    assert np.abs(p_ref - p_tb).max() > 1e-6


def test_final_projection_fixed_point(oracle):
    obj, A, sup = exact_problem(seed=3)
    p = oracle.final_projection(obj[None].astype(complex), A[None], sup[None].astype(float), 0.9)
    assert np.abs(p[0] - obj).max() < 1e-10


def cfft3(x):
    ax = (-3, -2, -1)
    return np.fft.fftshift(np.fft.fftn(np.fft.fftshift(x, axes=ax), axes=ax), axes=ax)


def cifft3(x):
    ax = (-3, -2, -1)
    return np.fft.fftshift(np.fft.ifftn(np.fft.fftshift(x, axes=ax), axes=ax), axes=ax)


def exact_problem_3d(L=8, M=10, N=12, seed=0, r=2.5):
    rng = np.random.default_rng(seed)
    zz, yy, xx = np.mgrid[0:L, 0:M, 0:N]
    sup = ((zz - L / 2) ** 2 + (yy - M / 2) ** 2 + (xx - N / 2) ** 2 <= r * r).astype(np.uint8)
    obj = rng.uniform(0.1, 1.0, (L, M, N)) * sup
    return obj, np.abs(cfft3(obj)), sup


def pm3(z, A):
    f = cfft3(z)
    m = np.abs(f)
    out = np.where((m > 0) & (A != 0), A * f / np.where(m > 0, m, 1), f)
    return cifft3(out)


def test_dm_ref3d_fixed_point(oracle):
    obj, A, sup = exact_problem_3d()
    psi = obj[None].astype(complex)
    p, t1, t2 = oracle.dm_ref3d(psi, psi, psi, A[None], sup[None].astype(float), 0.9, 3)
    assert np.abs(p - psi).max() < 1e-10
    assert np.abs(t1 - psi).max() < 1e-10 and np.abs(t2 - psi).max() < 1e-10


def test_dm_ref3d_first_step_closed_form_beta1(oracle):
    """beta = 1: t1 = P_S(2 P_M psi - psi), t2 = P_M psi, psi + t1 - t2, with
    P_M through numpy's n-dimensional FFT (src/CONV_MKL.cpp:47-63)."""
    obj, A, sup = exact_problem_3d(seed=1)
    psi = rnd((8, 10, 12), 21)
    p, t1, t2 = oracle.dm_ref3d(psi[None], psi[None], psi[None], A[None],
                                sup[None].astype(float), 1.0, 1)
    PS = lambda z: np.where(sup > 0, z.real, 0.0) + 0j  # noqa: E731
    e_t1 = PS(2 * pm3(psi, A) - psi)
    e_t2 = pm3(psi, A)
    assert np.allclose(t1[0], e_t1, atol=1e-10)
    assert np.allclose(t2[0], e_t2, atol=1e-10)
    assert np.allclose(p[0], psi + e_t1 - e_t2, atol=1e-10)


def test_dm_ref3d_carries_temporaries_and_batches(oracle):
    """Two iterations differ from two separate one-iteration runs restarted at
    t1 = t2 = psi (the temporaries carry over), and a batch of two volumes
    with one shared constraint equals the members run alone."""
    obj, A, sup = exact_problem_3d(seed=2)
    S = sup[None].astype(float)
    psi = rnd((2, 8, 10, 12), 22)
    p2, _, _ = oracle.dm_ref3d(psi, psi, psi, A[None], S, 0.9, 2)
    q, _, _ = oracle.dm_ref3d(psi, psi, psi, A[None], S, 0.9, 1)
    q2, _, _ = oracle.dm_ref3d(q, q, q, A[None], S, 0.9, 1)
    assert np.abs(p2 - q2).max() > 1e-6
    for b in range(2):
        pb, _, _ = oracle.dm_ref3d(psi[b:b + 1], psi[b:b + 1], psi[b:b + 1], A[None], S, 0.9, 2)
        assert np.allclose(pb[0], p2[b], atol=1e-12)


def test_conv_driver_fixed_point(oracle):
    """At an exact solution the conventional driver (src/CONV_MKL.cpp:11-101)
    returns it with zero convergence error."""
    obj, A, sup = exact_problem_3d(seed=3)
    p, err = oracle.conv_driver(A, sup.astype(float), obj[None].astype(complex), 4, 0.9)
    assert np.abs(p[0] - obj).max() < 1e-10
    assert err[0] < 1e-10


def test_dm_textbook_fixed_point_real_positive(oracle):
    obj, A, sup = exact_problem(seed=4)
    p = oracle.dm_textbook(obj[None].astype(complex), A[None], 0.8, 4, mask=sup[None],
                           space_mode=1)
    assert np.abs(p[0] - obj).max() < 1e-10


def test_hio_fixed_point_and_special_cases(oracle):
    obj, A, sup = exact_problem(seed=5)
    x = oracle.hio_step(obj[None].astype(complex), A[None], sup[None], 0.9, True, True)
    assert np.abs(x[0] - obj).max() < 1e-10
    z = rnd((1, 16, 16), 11)
    # support everywhere, complex mode: HIO step = P_M
    ones = np.ones((1, 16, 16), np.uint8)
    x1 = oracle.hio_step(z, A[None], ones, 0.7, False, True)
    assert np.allclose(x1[0], pm(z[0], A), atol=1e-10)
    # empty support: x - beta P_M x
    zeros = np.zeros((1, 16, 16), np.uint8)
    x0 = oracle.hio_step(z, A[None], zeros, 0.7, False, True)
    assert np.allclose(x0[0], z[0] - 0.7 * pm(z[0], A), atol=1e-10)
    # real-positive: inside support with Re x' >= 0 -> Re x'; Re x' < 0 -> x - beta x'
    xr = oracle.hio_step(z, A[None], ones, 0.7, True, True)
    p = pm(z[0], A)
    expect = np.where(p.real >= 0, p.real + 0j, z[0] - 0.7 * p)
    assert np.allclose(xr[0], expect, atol=1e-10)


def test_shrinkwrap_blur_closed_form(oracle):
    n = 64
    x = np.zeros((1, n, n), complex)
    x[0, n // 2, n // 2] = 1.0
    for sigma, tol in [(3.0, 1e-12), (1.5, 1e-4)]:
        s, blur = oracle.shrinkwrap(x, sigma, 0.1, want_blur=True)
        d = np.arange(n) - n // 2
        g1 = sum(np.exp(-((d + k * n) ** 2) / (2 * sigma ** 2)) for k in (-1, 0, 1))
        g = np.outer(g1, g1) / (2 * np.pi * sigma ** 2)
        assert np.abs(blur[0] - g).max() < tol * g.max()
        assert np.array_equal(s[0], (blur[0] >= 0.1 * blur[0].max()).astype(np.uint8))
    const = np.full((1, 16, 16), 2.0 + 0j)
    s, blur = oracle.shrinkwrap(const, 2.0, 0.5, want_blur=True)
    assert np.allclose(blur, 2.0) and s.all()


def test_rfactor_closed_forms(oracle):
    obj, A, sup = exact_problem(seed=6)
    assert oracle.rfactor(obj[None], A[None])[0] < 1e-12
    assert abs(oracle.rfactor(np.zeros((1, 16, 16)), A[None])[0] - 1.0) < 1e-12
    assert abs(oracle.rfactor(2 * obj[None], A[None])[0] - 1.0) < 1e-12


def test_converge_error(oracle):
    old, new = rnd((3, 4, 5), 1), rnd((3, 4, 5), 2)
    err, d = oracle.converge_error(old, new, 4, 5, 1, 3)
    assert np.allclose(d, new - old)
    for b in range(3):
        assert np.isclose(err[b], np.linalg.norm(new[b] - old[b]) / np.linalg.norm(new[b]))


# ---------------------------------------------------------------- slices -------

def test_slice_extract_on_grid_angles(oracle):
    n = 8
    vol = rnd((n, n, n), 7)
    s = oracle.slice_extract(vol, np.array([0.0, np.pi / 2]))
    c = n // 2
    assert np.allclose(s[0], vol[:, c, :], atol=1e-12)  # theta 0: x = v, y = c
    assert np.allclose(s[1], vol[:, :, c], atol=1e-12)  # theta pi/2: y = v, x = c


def test_slice_pair_adjoint_and_merge(oracle):
    n = 8
    ang = np.linspace(0, np.pi, 7, endpoint=False)
    v = rnd((n, n, n), 8)
    s = rnd((7, n, n), 9)
    Rv = oracle.slice_extract(v, ang)
    RTs, w = oracle.slice_scatter(s, ang)
    assert np.isclose(np.vdot(s, Rv), np.vdot(RTs, v), rtol=1e-12)
    merged = oracle.slice_merge(np.full((7, n, n), 3.0 + 1j), ang)
    assert np.allclose(merged[w > 0], 3.0 + 1j)
    assert np.all(merged[w == 0] == 0)


# ---------------------------------------------------------------- gradient -----

def test_amp_grad_finite_differences(oracle):
    M = N = 8
    x = rnd((2, M, N), 12)
    A = np.abs(rnd((1, M, N), 13)) + 0.2
    g = oracle.amp_grad(x, A)
    h = 1e-6
    rng = np.random.default_rng(0)
    for _ in range(12):
        b, r, c = rng.integers(2), rng.integers(M), rng.integers(N)
        for unit in (1.0, 1j):
            xp, xm = x.copy(), x.copy()
            xp[b, r, c] += h * unit
            xm[b, r, c] -= h * unit
            fd = (oracle.amp_loss(xp, A)[b] - oracle.amp_loss(xm, A)[b]) / (2 * h)
            an = g[b, r, c].real if unit == 1.0 else g[b, r, c].imag
            assert abs(fd - an) <= 1e-6 * max(1.0, np.abs(g).max())
    obj, Aex, _ = exact_problem(M, N, seed=14, r=2.5)
    assert np.abs(oracle.amp_grad(obj[None].astype(complex), Aex[None])).max() < 1e-9


def test_amp_grad_3d_finite_differences(oracle):
    n = 4
    x = rnd((n, n, n), 15)
    A = np.abs(rnd((n, n, n), 16)) + 0.2
    g = oracle.amp_grad_3d(x, A)
    h = 1e-6
    for idx in [(0, 1, 2), (3, 3, 0), (2, 0, 1)]:
        xp, xm = x.copy(), x.copy()
        xp[idx] += h
        xm[idx] -= h
        fd = (oracle.amp_loss_3d(xp, A) - oracle.amp_loss_3d(xm, A)) / (2 * h)
        assert abs(fd - g[idx].real) <= 1e-6 * max(1.0, np.abs(g).max())


# ------------------------------------------------- schedules and drivers -------

def hio_problem(n=32, B=2, seed=30):
    """Disc objects with their exact amplitudes and a loose square support,
    random starts: shrinkwrap then moves the support at every update."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:n, 0:n]
    A = np.zeros((B, n, n))
    for b in range(B):
        r = 5.0 + 2.0 * b
        obj = rng.uniform(0.2, 1.0, (n, n)) * (((yy - n / 2) ** 2 + (xx - n / 2) ** 2) <= r * r)
        A[b] = np.abs(cfft2(obj))
    sup = np.zeros((B, n, n), np.uint8)
    sup[:, n // 4:3 * n // 4, n // 4:3 * n // 4] = 1
    x0 = cifft2(A * np.exp(2j * np.pi * rng.random((B, n, n))))
    return x0, A, sup


@pytest.mark.parametrize("iters,sw_every,sigmas", [
    # the 200-iteration schedule's shape at a small scale: updates at
    # iterations 2, 4, 6, 8, sigma linear from 3.0 (first update) to 1.5 (last)
    (8, 2, [3.0, 2.5, 2.0, 1.5]),
    # iters not a multiple of sw_every: the trailing iteration has no update
    (9, 2, [3.0, 2.5, 2.0, 1.5]),
    # one update: sigma0
    (3, 3, [3.0]),
])
def test_hio_shrinkwrap_schedule_composition(oracle, iters, sw_every, sigmas):
    """orc_hio's schedule equals hio_step iterations with a shrinkwrap after
    every sw_every-th one, sigma annealed linearly 3.0 -> 1.5 over the updates
    (one per 20 iterations in the 200-iteration workload: 3.0 at iteration 20,
    1.5 at iteration 200).  The sigmas and update iterations are written out
    here, so a wrong update index, cadence or interpolation fails."""
    x0, A, sup0 = hio_problem()
    x, s = x0.copy(), sup0.copy()
    upd = [sw_every * (k + 1) for k in range(len(sigmas))]
    for it in range(1, iters + 1):
        x = oracle.hio_step(x, A, s, 0.9, True, True)
        if it in upd:
            s = oracle.shrinkwrap(x, sigmas[upd.index(it)], 0.1)
    xo, so = oracle.hio(x0, A, sup0.copy(), iters, 0.9, True, True, sw_every, 3.0, 1.5, 0.1)
    assert np.array_equal(so, s)
    assert np.abs(xo - x).max() == 0.0
    # the check is sensitive: the same loop with the first update one index
    # off in sigma (or one iteration late) ends elsewhere
    if len(sigmas) > 1:
        x2, s2 = x0.copy(), sup0.copy()
        for it in range(1, iters + 1):
            x2 = oracle.hio_step(x2, A, s2, 0.9, True, True)
            if it in upd:
                k = upd.index(it)
                s2 = oracle.shrinkwrap(x2, sigmas[min(k + 1, len(sigmas) - 1)], 0.1)
        assert not np.array_equal(s2, so) or np.abs(x2 - xo).max() > 1e-9
    x3, s3 = x0.copy(), sup0.copy()
    for it in range(1, iters + 1):
        x3 = oracle.hio_step(x3, A, s3, 0.9, True, True)
        if it - 1 in upd:
            s3 = oracle.shrinkwrap(x3, sigmas[upd.index(it - 1)], 0.1)
    assert not np.array_equal(s3, so) or np.abs(x3 - xo).max() > 1e-9


def test_hio_schedule_sigmas_of_the_200_iteration_workload(oracle):
    """The 200-iteration workload (shrinkwrap every 20, sigma 3.0 -> 1.5):
    ten updates at sigma = linspace(3.0, 1.5, 10).  Checked on 20 iterations
    with sw_every = 2 (ten updates as well), against the composition with
    numpy's linspace."""
    x0, A, sup0 = hio_problem(n=24, B=1, seed=31)
    sig = np.linspace(3.0, 1.5, 10)
    x, s = x0.copy(), sup0.copy()
    for it in range(1, 21):
        x = oracle.hio_step(x, A, s, 0.9, True, True)
        if it % 2 == 0:
            s = oracle.shrinkwrap(x, float(sig[it // 2 - 1]), 0.1)
    xo, so = oracle.hio(x0, A, sup0.copy(), 20, 0.9, True, True, 2, 3.0, 1.5, 0.1)
    assert np.array_equal(so, s)
    assert np.abs(xo - x).max() == 0.0


def exact_problem_complex(M=16, N=16, seed=0, r=4.0):
    """A complex object inside a disc support, with its exact amplitudes."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:M, 0:N]
    sup = ((yy - M / 2) ** 2 + (xx - N / 2) ** 2 <= r * r).astype(np.uint8)
    obj = (rng.uniform(-1.0, 1.0, (M, N)) + 1j * rng.uniform(-1.0, 1.0, (M, N))) * sup
    return obj, np.abs(cfft2(obj)), sup


@pytest.mark.parametrize("beta", [0.8, 0.7, -0.6])
def test_dm_textbook_support_only_closed_forms(oracle, beta):
    """space_mode 2 (P_S: zero outside the support, complex values kept).
    With the support everywhere P_S is the identity and one DM step is P_M psi
    for any beta; with an empty support P_S = 0 and the step is
    psi - beta P_M psi (P_M(psi / beta) = sign(beta) P_M psi when A > 0)."""
    _, A, _ = exact_problem(seed=40)
    A = A + 0.1  # no zero amplitudes: P_M is scale invariant for positive scales
    psi = rnd((1, 16, 16), 41)
    ones = np.ones((1, 16, 16), np.uint8)
    p1 = oracle.dm_textbook(psi, A[None], beta, 1, mask=ones, space_mode=2)
    assert np.abs(p1[0] - pm(psi[0], A)).max() < 1e-10
    zeros = np.zeros((1, 16, 16), np.uint8)
    p0 = oracle.dm_textbook(psi, A[None], beta, 1, mask=zeros, space_mode=2)
    expect = psi[0] - beta * np.sign(beta) * pm(np.sign(beta) * psi[0], A) * np.sign(beta)
    assert np.abs(p0[0] - expect).max() < 1e-10


def test_dm_textbook_fixed_points_tell_the_modes_apart(oracle):
    """An exact complex solution inside the support is a fixed point of
    space_mode 2 only; an exact real solution with negative values is a fixed
    point of modes 0 (MKL SpaceConstraint with S = support) and 2, not of 1
    (positivity)."""
    objc, Ac, supc = exact_problem_complex(seed=42)
    p2 = oracle.dm_textbook(objc[None], Ac[None], 0.8, 3, mask=supc[None], space_mode=2)
    assert np.abs(p2[0] - objc).max() < 1e-10
    p1 = oracle.dm_textbook(objc[None], Ac[None], 0.8, 3, mask=supc[None], space_mode=1)
    assert np.abs(p1[0] - objc).max() > 1e-3
    p0 = oracle.dm_textbook(objc[None], Ac[None], 0.8, 3, S=supc[None].astype(float), space_mode=0)
    assert np.abs(p0[0] - objc).max() > 1e-3
    objr = objc.real
    Ar = np.abs(cfft2(objr))
    for mode, kw in ((0, dict(S=supc[None].astype(float))), (2, dict(mask=supc[None]))):
        p = oracle.dm_textbook(objr[None].astype(complex), Ar[None], 0.8, 3, space_mode=mode, **kw)
        assert np.abs(p[0] - objr).max() < 1e-10, mode
    p = oracle.dm_textbook(objr[None].astype(complex), Ar[None], 0.8, 3, mask=supc[None],
                           space_mode=1)
    assert np.abs(p[0] - objr).max() > 1e-3


def test_dm_textbook_mode0_empty_constraint(oracle):
    """space_mode 0 with S < 0.9 everywhere: SpaceConstraint zeroes every
    pixel, so one step is psi - beta P_M psi."""
    _, A, _ = exact_problem(seed=43)
    A = A + 0.1
    psi = rnd((1, 16, 16), 44)
    S = np.full((1, 16, 16), 0.5)
    p = oracle.dm_textbook(psi, A[None], 0.8, 1, S=S, space_mode=0)
    assert np.abs(p[0] - (psi[0] - 0.8 * pm(psi[0], A))).max() < 1e-10


def test_carousel_driver_fixed_point(oracle):
    """Every projection at the same exact solution: the warm start averages
    equal neighbours, the DM iterations and the final projection keep it, and
    the convergence error is zero (src/CPRA_MKL.cpp:58-140)."""
    obj, A, sup = exact_problem(seed=50)
    P, nb = 4, 2
    AP = np.broadcast_to(A, (P, 16, 16)).copy()
    SP = np.broadcast_to(sup.astype(float), (P, 16, 16)).copy()
    init = np.broadcast_to(obj.astype(complex), (P, nb, 16, 16)).copy()
    psi, err = oracle.carousel_driver(AP, SP, init, 2, 3, 5, 0.9)
    assert np.abs(psi - obj).max() < 1e-10
    assert np.abs(err).max() < 1e-10


@pytest.mark.parametrize("preit", [0, 3])
def test_carousel_driver_sweep_by_hand(oracle, preit):
    """One sweep over P = 3 projections started at c_p * obj (obj an exact
    real solution, c_p > 0), worked by hand from src/CPRA_MKL.cpp:35-140:
      - Step A (`preit` DM iterations on projection 0): one iteration maps
        c * obj to obj and later ones keep it (P_M(c obj) = obj for c > 0),
        so c_0 becomes 1;
      - projections are finished one at a time, p = 0, 1, 2: warm start
        (:63-76), then the DM iterations, so the warm start of p reads p - 1
        as it left its own iterations in this sweep:
          p = 0:     (c_1 + c_0) / 2
          0 < p < 2: (c_{p-1}'' + c_{p+1}) / 2   (the own value has weight 0)
          p = 2:     (c_1'' + c_2) / 2
        with c'' = 1 after at least one DM iteration, c'' = c' with none;
      - with iters = 1 the convergence error of the last iteration is
        ||F(c' obj) - F obj|| / ||F obj|| = |c' - 1|, and psi ends at obj;
      - with iters = 0 nothing runs but the warm starts, and the final
        projection P_S((1 + 1/beta) P_M psi - psi/beta) gives
        (1 + 1/beta - c'/beta) obj.
    Each projection has two batch members with their own c."""
    obj, A, sup = exact_problem(seed=51)
    beta = 0.9
    c = np.array([[0.5, 0.6], [1.0, 1.2], [1.5, 1.3]])   # [P][member]
    P, nb = c.shape
    AP = np.broadcast_to(A, (P, 16, 16)).copy()
    SP = np.broadcast_to(sup.astype(float), (P, 16, 16)).copy()
    init = c[:, :, None, None] * obj[None, None].astype(complex)
    c0 = np.ones(nb) if preit > 0 else c[0]
    # iters = 1
    one = np.ones(nb)
    cp = np.stack([(c[1] + c0) / 2, (one + c[2]) / 2, (one + c[2]) / 2])
    psi, err = oracle.carousel_driver(AP, SP, init, 1, 1, preit, beta)
    assert np.abs(psi - obj).max() < 1e-10
    assert np.allclose(err, np.abs(cp - 1.0), atol=1e-10)
    # a Jacobi-style sweep (every warm start from the values before the
    # sweep) gives other errors for p = 1 and, for member 1, p = 2
    jac = np.stack([(c[1] + c0) / 2, (c0 + c[2]) / 2, (c[1] + c[2]) / 2])
    assert np.abs(np.abs(jac - 1.0) - np.abs(cp - 1.0)).max() > 0.05
    # iters = 0: warm starts only
    c0p = (c[1] + c0) / 2
    c1p = (c0p + c[2]) / 2
    c2p = (c1p + c[2]) / 2
    psi0, err0 = oracle.carousel_driver(AP, SP, init, 1, 0, preit, beta)
    coef = 1 + 1 / beta - np.stack([c0p, c1p, c2p]) / beta
    assert np.allclose(psi0, coef[:, :, None, None] * obj, atol=1e-10)
    assert np.all(err0 == 0)


def test_carousel3d_composition(oracle):
    """oracle.carousel3d against one outer iteration written out with numpy's
    n-dimensional FFT: x <- HIO steps; V = merge of Forward2D(x);
    w = Backward3D(V); v = max(Re w, 0) inside support3d, 0 outside;
    x = Backward2D(slices of Forward3D(v)).
    Angles (0, 0): both slices are the grid plane y = n/2 (sin 0 = 0,
    cos 0 = 1 exactly), so the merge is their mean on that plane and 0
    elsewhere, all by hand.  Angles (0, pi/2): the merge and extraction come
    from orc_slice_merge / orc_slice_extract, pinned above.  A random complex
    spectrum makes Re w take both signs and w complex, so using |w|, keeping
    Im w, dropping positivity or the support each fail."""
    n = 8
    c = n // 2
    rng = np.random.default_rng(60)
    Vt = rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))
    F = np.stack([Vt[:, c, :], Vt[:, c + 1, :]])
    x = cifft2(F)
    s3 = (rng.random((n, n, n)) < 0.6).astype(np.uint8)
    A = np.abs(F) + 0.1
    s2 = np.ones((2, n, n), np.uint8)
    for ang in (np.array([0.0, 0.0]), np.array([0.0, np.pi / 2])):
        for inner in (0, 1):
            xin = x if inner == 0 else oracle.hio_step(x, A, s2, 0.9, True, True)
            Fin = cfft2(xin)
            if ang[1] == 0.0:
                V = np.zeros((n, n, n), complex)
                V[:, c, :] = 0.5 * (Fin[0] + Fin[1])
            else:
                V = oracle.slice_merge(Fin, ang)
            w = cifft3(V)
            v = np.where((s3 > 0) & (w.real > 0), w.real, 0.0)
            Fv = cfft3(v)
            if ang[1] == 0.0:
                xe = cifft2(np.stack([Fv[:, c, :], Fv[:, c, :]]))
            else:
                xe = cifft2(oracle.slice_extract(Fv, ang))
            xo, vo = oracle.carousel3d(x, A, s2, ang, s3, 1, inner, 0.9, True)
            assert np.abs(vo - v).max() < 1e-12 * np.abs(v).max()
            assert np.abs(xo - xe).max() < 1e-12 * np.abs(xe).max()
            assert (w.real < 0).any() and (np.abs(w.imag) > 1e-3 * np.abs(w).max()).any()
            assert ((s3 == 0) & (w.real > 0)).any()


def _cfft(x, axis, inverse=False):
    """Centred DFT along one axis (index n/2 is frequency 0), unnormalised
    forward, 1/n inverse."""
    f = np.fft.ifft if inverse else np.fft.fft
    return np.fft.fftshift(f(np.fft.ifftshift(x, axes=axis), axis=axis), axes=axis)


def test_merge_and_extraction_commute_with_the_rotation_axis_transform(oracle):
    """The identity the GPU carousel's outer iteration rests on: slice row u
    lies in the plane z = u, and the merge and the re-projection combine
    whole lines along the rotation axis with weights that do not depend on
    the position along it, so they commute with the centred transform along
    that axis:
        IDFT_z(merge(F)) = merge(IDFT_u(F)),  extract(DFT_z(V)) = DFT_u(extract(V)).
    First the 1D centred transforms here are pinned to the oracle's 2D and 3D
    ones (their compositions give orc_forward2d/backward2d/backward3d).  A
    transform along the other slice axis (v), or along the volume's y, breaks
    both identities: the check cannot pass by construction."""
    rng = np.random.default_rng(11)
    n, P = 16, 7
    ang = rng.uniform(0, np.pi, P)
    F = rng.standard_normal((P, n, n)) + 1j * rng.standard_normal((P, n, n))
    V = rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))
    # the 1D transforms against the oracle's
    assert np.allclose(oracle.forward2d(F), _cfft(_cfft(F, -1), -2), rtol=1e-12, atol=1e-9)
    assert np.allclose(oracle.backward2d(F), _cfft(_cfft(F, -1, True), -2, True), rtol=1e-12, atol=1e-12)
    assert np.allclose(oracle.backward3d(V), _cfft(_cfft(_cfft(V, -1, True), -2, True), 0, True),
                       rtol=1e-12, atol=1e-12)
    # merge: slices [P][u][v] (u along z), volume [z][y][x]
    lhs = _cfft(oracle.slice_merge(F, ang, True), 0, True)
    rhs = oracle.slice_merge(_cfft(F, 1, True), ang, True)
    assert np.abs(lhs).max() > 1e-3
    assert np.allclose(lhs, rhs, rtol=1e-12, atol=1e-12)
    assert not np.allclose(lhs, oracle.slice_merge(_cfft(F, 2, True), ang, True), atol=1e-6)
    # re-projection
    lhs = oracle.slice_extract(_cfft(V, 0), ang, True)
    rhs = _cfft(oracle.slice_extract(V, ang, True), 1)
    assert np.allclose(lhs, rhs, rtol=1e-12, atol=1e-9)
    assert not np.allclose(oracle.slice_extract(_cfft(V, 1), ang, True), rhs, atol=1e-6)
