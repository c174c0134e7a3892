"""GPU parity of the carousel merge (gather-trilinear R^T(s)/R^T(1)) and
re-projection R against the FP64 oracle (scatter formulation)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("n,P", [(16, 7), (32, 64), (64, 180), (128, 180)])
@pytest.mark.parametrize("cplx", [True, False])
def test_merge_matches_oracle(hipmod, oracle, gpu, n, P, cplx):
    rng = np.random.default_rng(n + P)
    ang = np.arange(P) * np.pi / P
    if cplx:
        s = (rng.standard_normal((P, n, n)) + 1j * rng.standard_normal((P, n, n)))
        sd = hipmod.to_device(s.astype(np.complex64))
        vol = hipmod.zeros((n, n, n), np.complex64)
    else:
        s = rng.standard_normal((P, n, n))
        sd = hipmod.to_device(s.astype(np.float32))
        vol = hipmod.zeros((n, n, n), np.float32)
    angd = hipmod.to_device(ang.astype(np.float32))
    with hipmod.HipImpl(n, n, n, 1) as h:
        if cplx:
            assert h.Complex2DTo3DInterpolation(sd, vol, angd, n, n, P, n)
        else:
            assert h.Real2DTo3DInterpolation(sd, vol, angd, n, n, P, n)
        h.Sync()
    ref = oracle.slice_merge(s, ang.astype(np.float32).astype(np.float64), complex_data=cplx)
    got = vol.numpy()
    # the footprint geometry is FP64 on both sides, so also the voxels at the
    # edge of the covered set (tiny total weight) agree
    _, w = oracle.slice_scatter(np.ones((P, n, n)), ang.astype(np.float32).astype(np.float64),
                                complex_data=False)
    assert relerr(got, ref) < 1e-5
    edge = (w > 0) & (w < 1e-2)
    if edge.any():
        assert np.abs(got[edge] - ref[edge]).max() <= 1e-5 * np.abs(ref).max()
    assert np.all(got[w == 0] == 0)


@pytest.mark.parametrize("n,P", [(16, 5), (64, 180), (128, 180)])
def test_extract_matches_oracle_and_adjoint(hipmod, oracle, gpu, n, P):
    rng = np.random.default_rng(3 * n + P)
    ang = (np.arange(P) * np.pi / P).astype(np.float32)
    v = (rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n)))
    out = hipmod.zeros((P, n, n), np.complex64)
    with hipmod.HipImpl(n, n, n, 1) as h:
        assert h.slice_extract(hipmod.to_device(v.astype(np.complex64)), out, hipmod.to_device(ang), n, P)
        h.Sync()
    ref = oracle.slice_extract(v, ang.astype(np.float64))
    assert relerr(out.numpy(), ref) < 1e-5


def test_merge_rejects_non_cube(hipmod, gpu):
    d = hipmod.zeros(16, np.float32)
    with hipmod.HipImpl(16, 16, 8, 1) as h:
        with pytest.raises(hipmod.CpraHipError) as e:
            h.Real2DTo3DInterpolation(d, d, d, 16, 16, 4, 8)
        assert e.value.code == hipmod.ERR_UNSUPPORTED
