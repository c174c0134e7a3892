"""numpy/ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module.  The product (the HIP library and its
Python/C++ front ends) never does.

The functions take and return numpy arrays; complex arrays are complex128
(FP64 oracle) unless the function name ends in ``_f32``, in which case they
are complex64 and follow the reference's single-precision pointwise kernels
bit for bit where the reference semantics are exact (masks, permutations).
Every function cites the C routine it wraps; the C routines cite the
reference lines they restate.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libcpra_oracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with gcc (oracle/Makefile)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_P = C.c_void_p
_I64 = C.c_int64
_U64 = C.c_uint64


class HioParams(C.Structure):
    _fields_ = [
        ("iters", C.c_int64),
        ("beta", C.c_double),
        ("real_positive", C.c_int),
        ("dc_skip_zero", C.c_int),
        ("sw_every", C.c_int64),
        ("sw_sigma0", C.c_double),
        ("sw_sigma1", C.c_double),
        ("sw_frac", C.c_double),
    ]


def _declare(L):
    sig = {
        "orc_fft_lines": [_P, _I64, _I64, _I64, _I64, C.c_int, C.c_double],
        "orc_dft2": [_P, _I64, _I64, _I64, C.c_int],
        "orc_dft3": [_P, _I64, _I64, _I64, _I64, C.c_int],
        "orc_forward2d": [_P, _I64, _I64, _I64],
        "orc_backward2d": [_P, _I64, _I64, _I64],
        "orc_forward3d": [_P, _I64, _I64, _I64, _I64],
        "orc_backward3d": [_P, _I64, _I64, _I64, _I64],
        "orc_fftshift2d": [_P, C.c_size_t, _I64, _I64, _I64],
        "orc_fftshift3d": [_P, C.c_size_t, _I64, _I64, _I64, _I64],
        "orc_space_constraint_f32": [_P, _P, _U64, _U64],
        "orc_space_constraint_pos_f32": [_P, _P, _U64, _U64],
        "orc_data_constraint_f32": [_P, _P, _U64, _U64, C.c_int],
        "orc_data_constraint_cplx_f32": [_P, _P, _U64, _U64],
        "orc_merge_add_f32": [_P, _P, C.c_float, C.c_float, _U64],
        "orc_normalize_f32": [_P, C.c_float, _U64],
        "orc_space_constraint_f64": [_P, _P, _U64, _U64],
        "orc_data_constraint_f64": [_P, _P, _U64, _U64, C.c_int],
        "orc_merge_add_f64": [_P, _P, C.c_double, C.c_double, _U64],
        "orc_dm_ref": [_P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, C.c_double, _I64],
        "orc_final_projection": [_P, _P, _P, _I64, _I64, _I64, _I64, C.c_double],
        "orc_dm_ref3d": [_P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I64, C.c_double, _I64],
        "orc_final_projection3d": [_P, _P, _P, _I64, _I64, _I64, _I64, _I64, C.c_double],
        "orc_dm_textbook": [_P, _P, _P, _P, _I64, _I64, _I64, _I64, C.c_double, C.c_int, _I64],
        "orc_hio": [_P, _P, _P, _I64, _I64, _I64, _I64, C.POINTER(HioParams)],
        "orc_hio_step": [_P, _P, _P, _I64, _I64, _I64, _I64, C.c_double, C.c_int, C.c_int],
        "orc_shrinkwrap": [_P, _P, _I64, _I64, _I64, C.c_double, C.c_double, _P],
        "orc_rfactor": [_P, _P, _I64, _I64, _I64, _I64, _P],
        "orc_converge_error": [_P, _P, _P, _I64, _I64, _I64, _I64],
        "orc_slice_extract": [_P, _I64, _P, _I64, _P, C.c_int],
        "orc_slice_scatter": [_P, _I64, _P, _I64, _P, _P, C.c_int],
        "orc_slice_merge": [_P, _I64, _P, _I64, _P, C.c_int],
        "orc_amp_loss": [_P, _P, _I64, _I64, _I64, _I64, _P],
        "orc_amp_grad": [_P, _P, _I64, _I64, _I64, _I64, _P],
        "orc_amp_loss_3d": [_P, _P, _I64, _P],
        "orc_amp_grad_3d": [_P, _P, _I64, _P],
        "orc_num_threads": [],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def num_threads() -> int:
    return lib().orc_num_threads()


# ---- FFT -------------------------------------------------------------------

def fft_lines(x, n, stride, dist, count, sign, scale=1.0):
    x = _c128(x).copy()
    lib().orc_fft_lines(_ptr(x), n, stride, dist, count, sign, scale)
    return x


def forward2d(x):
    """orc_forward2d on [B][M][N] (or [M][N])."""
    x = _c128(x).copy()
    M, N = x.shape[-2:]
    B = x.size // (M * N)
    lib().orc_forward2d(_ptr(x), M, N, B)
    return x


def backward2d(x):
    x = _c128(x).copy()
    M, N = x.shape[-2:]
    B = x.size // (M * N)
    lib().orc_backward2d(_ptr(x), M, N, B)
    return x


def forward3d(x):
    """orc_forward3d on [B][L][M][N] (or [L][M][N])."""
    x = _c128(x).copy()
    L, M, N = x.shape[-3:]
    B = x.size // (L * M * N)
    lib().orc_forward3d(_ptr(x), M, N, L, B)
    return x


def backward3d(x):
    x = _c128(x).copy()
    L, M, N = x.shape[-3:]
    B = x.size // (L * M * N)
    lib().orc_backward3d(_ptr(x), M, N, L, B)
    return x


def fftshift2d(x, M, N, B):
    x = np.ascontiguousarray(x).copy()
    lib().orc_fftshift2d(_ptr(x), x.itemsize, M, N, B)
    return x


def fftshift3d(x, M, N, L, B):
    x = np.ascontiguousarray(x).copy()
    lib().orc_fftshift3d(_ptr(x), x.itemsize, M, N, L, B)
    return x


# ---- pointwise (reference FP32 semantics) ------------------------------------

def space_constraint_f32(x, S, batch):
    x = np.ascontiguousarray(x, dtype=np.complex64).copy()
    S = np.ascontiguousarray(S, dtype=np.float32)
    ok = lib().orc_space_constraint_f32(_ptr(x), _ptr(S), x.size, batch)
    return x, bool(ok)


def space_constraint_pos_f32(x, S, batch):
    x = np.ascontiguousarray(x, dtype=np.complex64).copy()
    S = np.ascontiguousarray(S, dtype=np.float32)
    ok = lib().orc_space_constraint_pos_f32(_ptr(x), _ptr(S), x.size, batch)
    return x, bool(ok)


def data_constraint_f32(x, A, batch, skip_zero=True):
    x = np.ascontiguousarray(x, dtype=np.complex64).copy()
    A = np.ascontiguousarray(A, dtype=np.float32)
    ok = lib().orc_data_constraint_f32(_ptr(x), _ptr(A), x.size, batch, int(skip_zero))
    return x, bool(ok)


def data_constraint_cplx_f32(x, Cc, batch):
    x = np.ascontiguousarray(x, dtype=np.complex64).copy()
    Cc = np.ascontiguousarray(Cc, dtype=np.complex64)
    ok = lib().orc_data_constraint_cplx_f32(_ptr(x), _ptr(Cc), x.size, batch)
    return x, bool(ok)


def merge_add_f32(src, dst, alpha, beta):
    src = np.ascontiguousarray(src, dtype=np.complex64)
    dst = np.ascontiguousarray(dst, dtype=np.complex64).copy()
    ok = lib().orc_merge_add_f32(_ptr(src), _ptr(dst), alpha, beta, dst.size)
    return dst, bool(ok)


def normalize_f32(x, norm):
    x = np.ascontiguousarray(x, dtype=np.complex64).copy()
    ok = lib().orc_normalize_f32(_ptr(x), norm, x.size)
    return x, bool(ok)


# ---- compositions (FP64) ---------------------------------------------------

def dm_ref(psi, t1, t2, A, S, beta, iters):
    psi, t1, t2 = _c128(psi).copy(), _c128(t1).copy(), _c128(t2).copy()
    A, S = _f64(A), _f64(S)
    M, N = psi.shape[-2:]
    B = psi.size // (M * N)
    ab = A.size // (M * N)
    ok = lib().orc_dm_ref(_ptr(psi), _ptr(t1), _ptr(t2), _ptr(A), _ptr(S), M, N, B, ab,
                          beta, iters)
    assert ok
    return psi, t1, t2


def final_projection(psi, A, S, beta):
    psi = _c128(psi).copy()
    A, S = _f64(A), _f64(S)
    M, N = psi.shape[-2:]
    B = psi.size // (M * N)
    ok = lib().orc_final_projection(_ptr(psi), _ptr(A), _ptr(S), M, N, B, A.size // (M * N),
                                    beta)
    assert ok
    return psi


def dm_ref3d(psi, t1, t2, A, S, beta, iters):
    """orc_dm_ref3d on volumes [B][L][M][N]; A, S: [amp_batch][L][M][N]."""
    psi, t1, t2 = _c128(psi).copy(), _c128(t1).copy(), _c128(t2).copy()
    A, S = _f64(A), _f64(S)
    L, M, N = psi.shape[-3:]
    vx = L * M * N
    ok = lib().orc_dm_ref3d(_ptr(psi), _ptr(t1), _ptr(t2), _ptr(A), _ptr(S), M, N, L,
                            psi.size // vx, A.size // vx, beta, iters)
    assert ok
    return psi, t1, t2


def final_projection3d(psi, A, S, beta):
    psi = _c128(psi).copy()
    A, S = _f64(A), _f64(S)
    L, M, N = psi.shape[-3:]
    vx = L * M * N
    ok = lib().orc_final_projection3d(_ptr(psi), _ptr(A), _ptr(S), M, N, L, psi.size // vx,
                                      A.size // vx, beta)
    assert ok
    return psi


def conv_driver(A, S, init, iters, beta):
    """FP64 restatement of one batch of the reference's conventional 3D
    reconstruction (src/CONV_MKL.cpp:11-101).  A, S: [L][M][N] data and space
    constraints (broadcast over the batch); init: psi [n][L][M][N].  `iters`
    difference-map iterations with t1 = t2 = psi at the start (:33-34) and
    carried after (:47-63); the convergence error ||F3 psi_before - F3
    psi_after|| / ||F3 psi_after|| of the last iteration (:40-45, 65-74), over
    the whole volume; the final projection (:78-88).  Returns (psi, err[n])."""
    psi = _c128(init).copy()
    n = psi.shape[0]
    L, M, N = psi.shape[-3:]
    t1, t2 = psi.copy(), psi.copy()
    err = np.zeros(n)
    if iters >= 1:
        psi, t1, t2 = dm_ref3d(psi, t1, t2, A, S, beta, iters - 1)
        old = psi.copy()
        psi, t1, t2 = dm_ref3d(psi, t1, t2, A, S, beta, 1)
        err, _ = converge_error(forward3d(old), forward3d(psi), M, N, L, n)
    psi = final_projection3d(psi, A, S, beta)
    return psi, err


def dm_textbook(psi, A, beta, iters, S=None, mask=None, space_mode=1):
    psi = _c128(psi).copy()
    A = _f64(A)
    M, N = psi.shape[-2:]
    B = psi.size // (M * N)
    Sp = _f64(S) if S is not None else None
    mp = np.ascontiguousarray(mask, dtype=np.uint8) if mask is not None else None
    ok = lib().orc_dm_textbook(_ptr(psi), _ptr(A), _ptr(Sp) if Sp is not None else None,
                               _ptr(mp) if mp is not None else None, M, N, B,
                               A.size // (M * N), beta, space_mode, iters)
    assert ok
    return psi


def hio(x, A, support, iters, beta=0.9, real_positive=True, dc_skip_zero=True,
        sw_every=0, sw_sigma0=3.0, sw_sigma1=1.5, sw_frac=0.1):
    x = _c128(x).copy()
    A = _f64(A)
    s = np.ascontiguousarray(support, dtype=np.uint8).copy()
    M, N = x.shape[-2:]
    B = x.size // (M * N)
    p = HioParams(iters, beta, int(real_positive), int(dc_skip_zero), sw_every, sw_sigma0,
                  sw_sigma1, sw_frac)
    ok = lib().orc_hio(_ptr(x), _ptr(A), _ptr(s), M, N, B, A.size // (M * N), C.byref(p))
    assert ok
    return x, s


def hio_step(x, A, support, beta=0.9, real_positive=True, dc_skip_zero=True):
    x = _c128(x).copy()
    A = _f64(A)
    s = np.ascontiguousarray(support, dtype=np.uint8)
    M, N = x.shape[-2:]
    B = x.size // (M * N)
    ok = lib().orc_hio_step(_ptr(x), _ptr(A), _ptr(s), M, N, B, A.size // (M * N), beta,
                            int(real_positive), int(dc_skip_zero))
    assert ok
    return x


def shrinkwrap(x, sigma, frac, want_blur=False):
    x = _c128(x)
    M, N = x.shape[-2:]
    B = x.size // (M * N)
    s = np.zeros(x.shape, dtype=np.uint8)
    blur = np.zeros(x.shape, dtype=np.float64) if want_blur else None
    lib().orc_shrinkwrap(_ptr(x), _ptr(s), M, N, B, sigma, frac,
                         _ptr(blur) if blur is not None else None)
    return (s, blur) if want_blur else s


def rfactor(x, A):
    x = _c128(x)
    A = _f64(A)
    M, N = x.shape[-2:]
    B = x.size // (M * N)
    r = np.zeros(B, dtype=np.float64)
    lib().orc_rfactor(_ptr(x), _ptr(A), M, N, B, A.size // (M * N), _ptr(r))
    return r


def converge_error(old, new, m, n, l, batch):
    old = _c128(old).copy()
    new = _c128(new)
    err = np.zeros(batch, dtype=np.float64)
    lib().orc_converge_error(_ptr(old), _ptr(new), _ptr(err), m, n, l, batch)
    return err, old


def slice_extract(vol, angles, complex_data=True):
    vol = _c128(vol) if complex_data else _f64(vol)
    n = vol.shape[-1]
    ang = _f64(angles)
    P = ang.size
    out = np.zeros((P, n, n), dtype=vol.dtype)
    lib().orc_slice_extract(_ptr(vol), n, _ptr(ang), P, _ptr(out), int(complex_data))
    return out


def slice_scatter(slices, angles, complex_data=True):
    slices = _c128(slices) if complex_data else _f64(slices)
    n = slices.shape[-1]
    ang = _f64(angles)
    vol = np.zeros((n, n, n), dtype=slices.dtype)
    w = np.zeros((n, n, n), dtype=np.float64)
    lib().orc_slice_scatter(_ptr(slices), n, _ptr(ang), ang.size, _ptr(vol), _ptr(w),
                            int(complex_data))
    return vol, w


def slice_merge(slices, angles, complex_data=True):
    slices = _c128(slices) if complex_data else _f64(slices)
    n = slices.shape[-1]
    ang = _f64(angles)
    vol = np.zeros((n, n, n), dtype=slices.dtype)
    lib().orc_slice_merge(_ptr(slices), n, _ptr(ang), ang.size, _ptr(vol), int(complex_data))
    return vol


def amp_loss(x, A):
    x = _c128(x)
    A = _f64(A)
    M, N = x.shape[-2:]
    B = x.size // (M * N)
    out = np.zeros(B, dtype=np.float64)
    lib().orc_amp_loss(_ptr(x), _ptr(A), M, N, B, A.size // (M * N), _ptr(out))
    return out


def amp_grad(x, A):
    x = _c128(x)
    A = _f64(A)
    M, N = x.shape[-2:]
    B = x.size // (M * N)
    g = np.zeros_like(x)
    lib().orc_amp_grad(_ptr(x), _ptr(A), M, N, B, A.size // (M * N), _ptr(g))
    return g


def amp_loss_3d(x, A):
    x = _c128(x)
    A = _f64(A)
    out = np.zeros(1, dtype=np.float64)
    lib().orc_amp_loss_3d(_ptr(x), _ptr(A), x.shape[-1], _ptr(out))
    return float(out[0])


def amp_grad_3d(x, A):
    x = _c128(x)
    A = _f64(A)
    g = np.zeros_like(x)
    lib().orc_amp_grad_3d(_ptr(x), _ptr(A), x.shape[-1], _ptr(g))
    return g


def carousel3d(x, A, support2d, angles, support3d, outer, inner=1, beta=0.9, dc_skip_zero=True):
    """FP64 restatement of one 3D carousel run (the batched carousel: the
    reference's Jacobi-form sweep, src/CPRA_MKL.cpp:58-124, with the 3D
    consistency step): per outer iteration, `inner` HIO steps on every
    projection (orc_hio_step, real-positive), F = Forward2D(x), V = slice merge
    (orc_slice_merge, complex), v = Backward3D(V), v <- max(Re v, 0) inside
    support3d and 0 elsewhere, V = Forward3D(v), F = slice extraction,
    x = Backward2D(F).  Returns (x, v) with v the constrained real-space
    object of the last iteration.  Composed from functions each pinned on
    its own."""
    x = _c128(x).copy()
    s3 = np.asarray(support3d) != 0
    v = None
    for _ in range(outer):
        for _ in range(inner):
            x = hio_step(x, A, support2d, beta, True, dc_skip_zero)
        V = slice_merge(forward2d(x), angles, True)
        w = backward3d(V)
        v = np.where(s3 & (w.real > 0), w.real, 0.0).astype(np.complex128)
        x = backward2d(slice_extract(forward3d(v), angles, True))
    return x, v


def carousel_driver(A, S, init, epi, iters, preit, beta):
    """FP64 restatement of one batch of the reference's carousel
    reconstruction (src/CPRA_MKL.cpp:10-154).  A, S: [P][M][N] data and space
    constraints; init: psi [P][n][M][N].  Step A (:35-55): `preit` DM
    iterations (orc_dm_ref) on projection 0 with t1 = t2 = psi_0; Step B
    (:58-124): `epi` sequential sweeps, each projection warm-started from its
    neighbours (:63-76), t1 = t2 = psi_p, `iters` DM iterations, and in the
    last sweep the convergence error ||F psi_before - F psi_after|| / ||F
    psi_after|| of the last iteration (:85-91, 112-121); final projection
    (:126-140).  Returns (psi [P][n][M][N], err [P][n])."""
    psi = _c128(init).copy()
    P, n = psi.shape[:2]
    M, N = psi.shape[-2:]
    err = np.zeros((P, n))

    def dm(p, k, t1, t2):
        if k <= 0:
            return t1, t2
        psi[p], t1, t2 = dm_ref(psi[p], t1, t2, A[p:p + 1], S[p:p + 1], beta, k)
        return t1, t2

    dm(0, preit, psi[0].copy(), psi[0].copy())
    for e in range(epi):
        for p in range(P):
            if p == P - 1:
                psi[p] = 0.5 * psi[p - 1] + 0.5 * psi[p]
            elif p == 0:
                psi[p] = 0.5 * psi[p + 1] + 0.5 * psi[p]
            else:
                psi[p] = 0.5 * psi[p - 1]
                psi[p] = 0.5 * psi[p + 1] + psi[p]
            t1, t2 = psi[p].copy(), psi[p].copy()
            if e == epi - 1 and iters >= 1:
                t1, t2 = dm(p, iters - 1, t1, t2)
                old = psi[p].copy()
                dm(p, 1, t1, t2)
                err[p], _ = converge_error(forward2d(old), forward2d(psi[p]), M, N, 1, n)
            else:
                dm(p, iters, t1, t2)
    for p in range(P):
        psi[p] = final_projection(psi[p], A[p:p + 1], S[p:p + 1], beta)
    return psi, err
