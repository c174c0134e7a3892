"""GPU parity of the CpraImpl primitives (through the C ABI) against the
oracle: permutations and masks bit-exact, FFTs within 1e-5 relative L2 (FP32)
or 1e-12 (FP64), and the reference's own unit tests (test/unitest_mkl.cpp,
test/unitest_cuda.cu) restated on the HIP backend."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


# ------------------------------------------------------------------ reference tests

def test_reference_initialize_range(hipmod, gpu):
    with hipmod.HipImpl(1, 1, 1, 1) as h:
        x = hipmod.zeros(1000, np.float32)
        assert h.Initialize(x, 1000, seed=5)
        h.Sync()
        v = x.numpy()
        assert ((v >= 0) & (v <= 1)).all() and v.std() > 0.2
        assert h.Initialize(x, 0) is False


def test_reference_norm_merge(hipmod, gpu):
    # test/unitest_mkl.cpp:15-40 and test/unitest_cuda.cu:14-41
    with hipmod.HipImpl(1, 1, 1, 1) as h:
        x = hipmod.to_device(np.full(1000, 10 + 10j, dtype=np.complex64))
        assert h.MergeAddData(x, x, 0.5, 0.5, 1000)
        h.Sync()
        assert (x.numpy() == 10 + 10j).all()
        assert h.Normalization(x, 10.0, 1000)
        h.Sync()
        assert (x.numpy() == 1 + 1j).all()
        y = hipmod.to_device(np.full(1000, 10 + 10j, dtype=np.complex64))
        assert h.MergeAddData(y, y, 1.0, 1.0, 1000)
        assert h.Normalization(y, 20.0, 1000)
        h.Sync()
        assert (y.numpy() == 1 + 1j).all()
        assert h.MergeAddData(y, x, 0.0, 2.0, 1000) is False  # alpha == 0: no-op
        h.Sync()
        assert (x.numpy() == 1 + 1j).all()


def test_reference_2d_dft_known_answer_double(hipmod, gpu):
    # test/unitest_mkl.cpp:42-71 (double 10x10; generic-length FFT path)
    with hipmod.HipImpl(10, 10, 1, 1, dtype="float64") as h:
        x = hipmod.to_device(np.ones((10, 10), dtype=np.complex128))
        assert h.Forward2D(x)
        h.Sync()
        y = x.numpy()
        expect = np.zeros((10, 10))
        expect[5, 5] = 100
        assert np.abs(y.real - expect).max() < 1e-13
        assert np.abs(y.imag).max() < 1e-13
        assert h.Backward2D(x)
        h.Sync()
        z = x.numpy()
        assert np.abs(z.real - 1).max() < 1e-14 and np.abs(z.imag).max() < 1e-14


# ------------------------------------------------------------------ shifts

@pytest.mark.parametrize("M,N,B", [(256, 256, 3), (10, 14, 2), (128, 64, 1)])
@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_fftshift2d_bit_exact(hipmod, oracle, gpu, M, N, B, dtype):
    cd = np.complex64 if dtype == "float32" else np.complex128
    x = (np.random.default_rng(M + N).standard_normal((B, M, N)) * (1 + 1j)).astype(cd)
    d = hipmod.to_device(x)
    with hipmod.HipImpl(M, N, 1, B, dtype=dtype) as h:
        assert h.FFTShift2D(d, M, N, B)
        h.Sync()
    assert np.array_equal(d.numpy(), oracle.fftshift2d(x, M, N, B))


@pytest.mark.parametrize("L,M,N,B", [(8, 6, 4, 2), (16, 16, 16, 1)])
def test_fftshift3d_bit_exact(hipmod, oracle, gpu, L, M, N, B):
    x = np.random.default_rng(1).standard_normal((B, L, M, N)).astype(np.complex64)
    d = hipmod.to_device(x)
    with hipmod.HipImpl(M, N, L, B) as h:
        assert h.FFTShift3D(d, M, N, L, B)
        h.Sync()
    assert np.array_equal(d.numpy(), oracle.fftshift3d(x, M, N, L, B))


def test_odd_shift_rejected(hipmod, gpu):
    d = hipmod.zeros(9 * 8, np.complex64)
    with hipmod.HipImpl(9, 8, 1, 1) as h:
        with pytest.raises(hipmod.CpraHipError) as e:
            h.FFTShift2D(d, 9, 8, 1)
        assert e.value.code == hipmod.ERR_UNSUPPORTED


# ------------------------------------------------------------------ pointwise

def test_space_constraint_bit_exact(hipmod, oracle, gpu):
    rng = np.random.default_rng(2)
    per, B = 4096, 3
    x = (rng.standard_normal(per * B) + 1j * rng.standard_normal(per * B)).astype(np.complex64)
    S = rng.uniform(0.85, 0.95, per).astype(np.float32)
    S[:4] = [0.9, np.nextafter(np.float32(0.9), np.float32(1)), np.nan, 0.0]
    for mode, ref in [(hipmod.SPACE_MKL, oracle.space_constraint_f32),
                      (hipmod.SPACE_POSITIVE, oracle.space_constraint_pos_f32)]:
        d = hipmod.to_device(x)
        with hipmod.HipImpl(1, 1, 1, 1) as h:
            assert h.SpaceConstraint(d, hipmod.to_device(S), per * B, B, mode)
            h.Sync()
        e, _ = ref(x, S, B)
        assert np.array_equal(d.numpy().view(np.uint32), e.view(np.uint32)), mode


def test_data_constraints(hipmod, oracle, gpu):
    rng = np.random.default_rng(3)
    per, B = 5000, 2
    x = (rng.standard_normal(per * B) + 1j * rng.standard_normal(per * B)).astype(np.complex64)
    x[:3] = 0
    A = np.abs(rng.standard_normal(per)).astype(np.float32)
    A[5:50] = 0
    for skip in (True, False):
        d = hipmod.to_device(x)
        with hipmod.HipImpl(1, 1, 1, 1) as h:
            assert h.DataConstraint(d, hipmod.to_device(A), per * B, B, skip_zero=skip)
            h.Sync()
        e, _ = oracle.data_constraint_f32(x, A, B, skip_zero=skip)
        assert relerr(d.numpy(), e) < 1e-6
    Cc = (rng.standard_normal(per) + 1j * rng.standard_normal(per)).astype(np.complex64)
    Cc[::7] = 0
    d = hipmod.to_device(x)
    with hipmod.HipImpl(1, 1, 1, 1) as h:
        assert h.DataConstraint(d, hipmod.to_device(Cc), per * B, B)
        h.Sync()
    e, _ = oracle.data_constraint_cplx_f32(x, Cc, B)
    assert np.array_equal(d.numpy(), e)


def test_converge_error(hipmod, oracle, gpu):
    rng = np.random.default_rng(4)
    m, n, l, B = 32, 16, 2, 3
    old = (rng.standard_normal((B, l, m, n)) + 1j * rng.standard_normal((B, l, m, n))).astype(np.complex64)
    new = (rng.standard_normal((B, l, m, n)) + 1j * rng.standard_normal((B, l, m, n))).astype(np.complex64)
    do, dn = hipmod.to_device(old), hipmod.to_device(new)
    with hipmod.HipImpl(m, n, l, B) as h:
        ok, err = h.ConvergeError(do, dn, m, n, l, B)
    e, diff = oracle.converge_error(old, new, m, n, l, B)
    assert ok and np.allclose(err, e, rtol=1e-5)
    assert relerr(do.numpy(), diff) < 1e-6


# ------------------------------------------------------------------ FFTs

FFT2_CASES = [(64, 64, 4), (128, 128, 2), (192, 192, 2), (256, 256, 3), (384, 384, 1),
              (512, 512, 1), (256, 128, 2), (128, 512, 1), (192, 256, 1), (10, 10, 2),
              (12, 20, 2), (96, 160, 1)]


@pytest.mark.parametrize("M,N,B", FFT2_CASES)
def test_forward_backward_2d_f32(hipmod, oracle, gpu, M, N, B):
    rng = np.random.default_rng(M * 7 + N)
    x = (rng.standard_normal((B, M, N)) + 1j * rng.standard_normal((B, M, N))).astype(np.complex64)
    d = hipmod.to_device(x)
    with hipmod.HipImpl(M, N, 1, B) as h:
        assert h.Forward2D(d)
        h.Sync()
        y = d.numpy()
        assert relerr(y, oracle.forward2d(x)) < 1e-5, "forward"
        assert h.Backward2D(d)
        h.Sync()
    assert relerr(d.numpy(), x.astype(np.complex128)) < 1e-5, "round trip"
    d2 = hipmod.to_device(x)
    with hipmod.HipImpl(M, N, 1, B) as h:
        h.Backward2D(d2)
        h.Sync()
    assert relerr(d2.numpy(), oracle.backward2d(x)) < 1e-5, "backward"


@pytest.mark.parametrize("M,N,B", [(64, 64, 2), (256, 256, 1), (10, 10, 1), (128, 192, 1)])
def test_forward_backward_2d_f64(hipmod, oracle, gpu, M, N, B):
    rng = np.random.default_rng(M + 3 * N)
    x = rng.standard_normal((B, M, N)) + 1j * rng.standard_normal((B, M, N))
    d = hipmod.to_device(x)
    with hipmod.HipImpl(M, N, 1, B, dtype="float64") as h:
        assert h.Forward2D(d)
        h.Sync()
        assert relerr(d.numpy(), oracle.forward2d(x)) < 1e-13
        assert h.Backward2D(d)
        h.Sync()
    assert relerr(d.numpy(), x) < 1e-13


@pytest.mark.parametrize("L,M,N,B", [(64, 64, 64, 1), (128, 128, 128, 1), (16, 16, 16, 2),
                                     (32, 64, 128, 1), (256, 256, 256, 1)])
def test_forward_backward_3d_f32(hipmod, oracle, gpu, L, M, N, B):
    rng = np.random.default_rng(L + M + N)
    x = (rng.standard_normal((B, L, M, N)) + 1j * rng.standard_normal((B, L, M, N))).astype(np.complex64)
    d = hipmod.to_device(x)
    with hipmod.HipImpl(M, N, L, B) as h:
        assert h.Forward3D(d)
        h.Sync()
        assert relerr(d.numpy(), oracle.forward3d(x)) < 1e-5
        assert h.Backward3D(d)
        h.Sync()
    assert relerr(d.numpy(), x.astype(np.complex128)) < 1e-5


def test_forward_backward_3d_512_separable(hipmod, gpu):
    """Forward3D at 512^3 (the reference benchmark's Step D volume) on a sum of
    two separable volumes a(z) b(y) c(x): its centred 3D DFT is the sum of the
    outer products of the three centred 1D DFTs, computed with numpy in FP64
    line by line, so the whole 1 GiB result is checked without an FP64 3D
    transform.  Backward3D then returns the input."""
    n = 512
    rng = np.random.default_rng(512)
    f = [[rng.standard_normal(n) + 1j * rng.standard_normal(n) for _ in range(3)] for _ in range(2)]
    c1 = lambda v: np.fft.fftshift(np.fft.fft(np.fft.fftshift(v)))  # noqa: E731

    def outer_sum(parts):
        out = np.zeros((n, n, n), np.complex64)
        for a, b, c in parts:
            out += (a.astype(np.complex64)[:, None, None] * b.astype(np.complex64)[None, :, None]
                    * c.astype(np.complex64)[None, None, :])
        return out

    x = outer_sum(f)
    d = hipmod.to_device(x[None])
    with hipmod.HipImpl(n, n, n, 1) as h:
        assert h.Forward3D(d)
        h.Sync()
        y = d.numpy()[0]
        expect = outer_sum([[c1(a), c1(b), c1(c)] for a, b, c in f])
        assert relerr(y, expect) < 1e-5
        del y, expect
        assert h.Backward3D(d)
        h.Sync()
    assert relerr(d.numpy()[0], x) < 1e-5


def test_fft_known_answers_f32(hipmod, gpu):
    M = N = 256
    with hipmod.HipImpl(M, N, 1, 1) as h:
        # constant -> impulse of M*N at the centre (DC at (M/2, N/2))
        d = hipmod.to_device(np.ones((M, N), dtype=np.complex64))
        h.Forward2D(d)
        h.Sync()
        y = d.numpy()
        assert abs(y[M // 2, N // 2] - M * N) < 1e-2
        y[M // 2, N // 2] = 0
        assert np.abs(y).max() < 1e-2
        # centred impulse -> constant 1
        imp = np.zeros((M, N), dtype=np.complex64)
        imp[M // 2, N // 2] = 1
        d = hipmod.to_device(imp)
        h.Forward2D(d)
        h.Sync()
        assert np.abs(d.numpy() - 1).max() < 1e-6
        # single frequency: cos along columns -> two impulses at N/2 +- k
        k = 5
        c = np.arange(N)
        row = np.cos(2 * np.pi * k * (c - N // 2) / N).astype(np.complex64)
        d = hipmod.to_device(np.tile(row, (M, 1)))
        h.Forward2D(d)
        h.Sync()
        y = d.numpy()
        assert abs(y[M // 2, N // 2 + k] - M * N / 2) < 0.05
        assert abs(y[M // 2, N // 2 - k] - M * N / 2) < 0.05


def test_graph_capture_replays_primitive_sequence(hipmod, gpu):
    """cpra_hip_graph_begin/_end record the calls made on a handle (here the
    reference's DM iteration as primitives, src/CPRA_MKL.cpp:94-110, plus one
    fused engine call, which runs eagerly inside the capture); three replays
    give the bits of three eager runs.  Calls that need a host
    synchronisation are refused while the capture is open."""
    n, B = 64, 2
    rng = np.random.default_rng(5)
    shp = (B, n, n)
    x0 = (rng.standard_normal(shp) + 1j * rng.standard_normal(shp)).astype(np.complex64)
    A = rng.random(shp).astype(np.float32)
    S = rng.random(shp).astype(np.float32)
    px = B * n * n
    outs = []
    for mode in ("eager", "graph"):
        with hipmod.HipImpl(n, n, 1, B) as h:
            psi, t1, t2 = (hipmod.to_device(x0) for _ in range(3))
            dA, dS = hipmod.to_device(A), hipmod.to_device(S)
            mask = hipmod.zeros(shp, np.uint8)
            h.support_from_constraint(dS, mask, px)

            def body():
                h.Forward2D(t1)
                h.DataConstraint(t1, dA, px, B)
                h.Backward2D(t1)
                h.MergeAddData(psi, t1, -1 / 0.9, 1 + 1 / 0.9, px)
                h.SpaceConstraint(t1, dS, px, B)
                h.dm_run(psi, dA, mask, t1=t1, t2=t2, iters=3, beta=0.9, textbook=0, use_graph=1)

            h.dm_run(psi, dA, mask, t1=t1, t2=t2, iters=1, use_graph=0)  # size the workspaces
            h.Sync()
            if mode == "eager":
                for _ in range(3):
                    body()
            else:
                assert h.graph_begin()
                body()
                with pytest.raises(hipmod.CpraHipError) as e:
                    h.Sync()
                assert e.value.code == hipmod.ERR_ARG
                gid = h.graph_end()
                for _ in range(3):
                    assert h.graph_launch(gid)
                assert h.graph_destroy(gid)
            h.Sync()
            outs.append([a.numpy() for a in (psi, t1, t2)])
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_graph_capture_of_adaptive_hio(hipmod, gpu):
    """cpra_hip_hio_run inside a caller's capture: the adaptive row mode
    decides on the device (no synchronisation), so a captured run with
    shrinkwrap updates replays to the bits of the same run made eagerly."""
    n, B = 64, 3
    rng = np.random.default_rng(9)
    shp = (B, n, n)
    x0 = (rng.standard_normal(shp) + 1j * rng.standard_normal(shp)).astype(np.complex64)
    A = rng.random(shp).astype(np.float32)
    yy, xx = np.mgrid[0:n, 0:n]
    sup0 = np.broadcast_to((((yy - n / 2) ** 2 + (xx - n / 2) ** 2) < 15 ** 2), shp).astype(np.uint8)
    kw = dict(iters=6, beta=0.9, real_positive=1, sw_every=3, use_graph=1)
    outs = []
    for mode in ("eager", "graph"):
        with hipmod.HipImpl(n, n, 1, B) as h:
            x, sup, dA = hipmod.to_device(x0), hipmod.to_device(sup0), hipmod.to_device(A)
            h.hio_run(hipmod.to_device(x0), dA, hipmod.to_device(sup0), iters=1, sw_every=0)  # workspaces
            h.Sync()
            if mode == "eager":
                h.hio_run(x, dA, sup, **kw)
            else:
                assert h.graph_begin()
                h.hio_run(x, dA, sup, **kw)
                gid = h.graph_end()
                assert h.graph_launch(gid)
                h.Sync()
                assert h.graph_destroy(gid)
            h.Sync()
            outs.append((x.numpy(), sup.numpy()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
