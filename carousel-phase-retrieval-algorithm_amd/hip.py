"""Python front end of libcpra_hip.so (ctypes; no torch types cross the ABI).

``HipImpl`` mirrors the reference's backend interface ``CPRA::CpraImpl<T>``
(include/CPRA_Impl.hpp:18-63): same method names, argument meaning and bool
results (``True`` = the reference's ``true``, ``False`` = its ``false`` for a
well-formed no-op such as ``num == 0`` or ``alpha == 0``).  Errors that the
reference would throw or abort on raise ``CpraHipError``.

Arguments that the C++ interface takes as pointers accept either a torch
tensor (its ``data_ptr()`` is passed; the tensor must live on the GPU) or a
raw integer device address, or a ``DeviceArray`` (device memory from the
library's own allocator, used by the tests and benchmarks so that no second
HIP runtime is loaded).  With ``use_torch_stream=True`` the handle launches
on torch's current stream so that work is ordered with torch operations.

The library must be built (``build.py``); nothing here falls back to a CPU
implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcpra_hip.so")

OK = 0
FALSE = 1
ERR_ARG = -1
ERR_UNSUPPORTED = -2
ERR_NOMEM = -3
ERR_HIP = -4
ERR_NODEV = -5

F32 = 0
F64 = 1

SPACE_MKL = 0
SPACE_POSITIVE = 1


class CpraHipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class DmParams(C.Structure):
    _fields_ = [
        ("iters", C.c_int32),
        ("beta", C.c_float),
        ("textbook", C.c_int32),
        ("space_mode", C.c_int32),
        ("dc_skip_zero", C.c_int32),
        ("use_graph", C.c_int32),
    ]


class SizeClass(C.Structure):
    _fields_ = [
        ("n", C.c_uint32),
        ("count", C.c_uint32),
        ("psi", C.c_void_p),
        ("amp", C.c_void_p),
        ("support", C.c_void_p),
    ]


MIXED_FUSED, MIXED_PER_CLASS, MIXED_PER_PATTERN = 0, 1, 2


class CarouselParams(C.Structure):
    _fields_ = [
        ("outer_iters", C.c_int32),
        ("inner_iters", C.c_int32),
        ("beta", C.c_float),
        ("dc_skip_zero", C.c_int32),
        ("use_graph", C.c_int32),
    ]


class HioParams(C.Structure):
    _fields_ = [
        ("iters", C.c_int32),
        ("beta", C.c_float),
        ("real_positive", C.c_int32),
        ("dc_skip_zero", C.c_int32),
        ("error_reduction", C.c_int32),
        ("sw_every", C.c_int32),
        ("sw_sigma0", C.c_float),
        ("sw_sigma1", C.c_float),
        ("sw_frac", C.c_float),
        ("use_graph", C.c_int32),
    ]


_lib = None

_P = C.c_void_p
_U64 = C.c_uint64
_H = C.c_void_p

_SIGS = {
    "cpra_hip_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, _U64, _U64, _U64, _U64]),
    "cpra_hip_destroy": (C.c_int, [_H]),
    "cpra_hip_last_error": (C.c_char_p, [_H]),
    "cpra_hip_version": (C.c_char_p, []),
    "cpra_hip_device_count": (C.c_int, []),
    "cpra_hip_set_stream": (C.c_int, [_H, _P]),
    "cpra_hip_get_stream": (C.c_void_p, [_H]),
    "cpra_hip_malloc": (C.c_int, [_H, C.POINTER(C.c_void_p), _U64, C.c_int]),
    "cpra_hip_free": (C.c_int, [_H, _P]),
    "cpra_hip_memcpy": (C.c_int, [_H, _P, _P, _U64]),
    "cpra_hip_sync": (C.c_int, [_H]),
    "cpra_hip_initialize": (C.c_int, [_H, _P, _U64, _U64]),
    "cpra_hip_forward2d": (C.c_int, [_H, _P]),
    "cpra_hip_backward2d": (C.c_int, [_H, _P]),
    "cpra_hip_forward3d": (C.c_int, [_H, _P]),
    "cpra_hip_backward3d": (C.c_int, [_H, _P]),
    "cpra_hip_fftshift2d": (C.c_int, [_H, _P, _U64, _U64, _U64]),
    "cpra_hip_fftshift3d": (C.c_int, [_H, _P, _U64, _U64, _U64, _U64]),
    "cpra_hip_space_constraint": (C.c_int, [_H, _P, _P, _U64, _U64, C.c_int]),
    "cpra_hip_data_constraint_real": (C.c_int, [_H, _P, _P, _U64, _U64, C.c_int]),
    "cpra_hip_data_constraint_complex": (C.c_int, [_H, _P, _P, _U64, _U64]),
    "cpra_hip_merge_add": (C.c_int, [_H, _P, _P, C.c_double, C.c_double, _U64]),
    "cpra_hip_normalize": (C.c_int, [_H, _P, C.c_double, _U64]),
    "cpra_hip_converge_error": (C.c_int, [_H, _P, _P, _P, _U64, _U64, _U64, _U64]),
    "cpra_hip_hio_default_params": (None, [C.POINTER(HioParams)]),
    "cpra_hip_hio_run": (C.c_int, [_H, _P, _P, _U64, _P, C.POINTER(HioParams), _P]),
    "cpra_hip_shrinkwrap": (C.c_int, [_H, _P, _P, C.c_float, C.c_float, _P]),
    "cpra_hip_dm_default_params": (None, [C.POINTER(DmParams)]),
    "cpra_hip_dm_run": (C.c_int, [_H, _P, _P, _P, _P, _U64, _P, C.POINTER(DmParams)]),
    "cpra_hip_dm_run_noalloc": (C.c_int, [_H, _P, _P, _P, _P, _U64, _P, C.POINTER(DmParams)]),
    "cpra_hip_dm3d_run": (C.c_int, [_H, _P, _P, _P, _P, _U64, _P, C.POINTER(DmParams)]),
    "cpra_hip_graph_begin": (C.c_int, [_H]),
    "cpra_hip_graph_end": (C.c_int, [_H, C.POINTER(C.c_uint64)]),
    "cpra_hip_graph_launch": (C.c_int, [_H, C.c_uint64]),
    "cpra_hip_graph_destroy": (C.c_int, [_H, C.c_uint64]),
    "cpra_hip_rfactor": (C.c_int, [_H, _P, _P, _U64, _P]),
    "cpra_hip_support_from_constraint": (C.c_int, [_H, _P, _P, _U64]),
    "cpra_hip_real2d_to_3d": (C.c_int, [_H, _P, _P, _P, _U64, _U64, _U64, _U64]),
    "cpra_hip_complex2d_to_3d": (C.c_int, [_H, _P, _P, _P, _U64, _U64, _U64, _U64]),
    "cpra_hip_slice_extract": (C.c_int, [_H, _P, _P, _P, _U64, _U64, C.c_int]),
    "cpra_hip_set_kernel_timing": (C.c_int, [_H, C.c_int]),
    "cpra_hip_autotune_2d": (C.c_int, [_H, _P, _P, _U64, _P, C.c_int, C.c_char_p, _U64]),
    "cpra_hip_set_tuning": (C.c_int, [_H, C.c_char_p]),
    "cpra_hip_get_tuning": (C.c_int, [_H, C.c_char_p, _U64]),
    "cpra_hip_set_tuning_cache": (C.c_int, [C.c_char_p, C.c_int]),
    "cpra_hip_get_kernel_times": (C.c_int, [_H, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                            C.c_int]),
    "cpra_hip_amp_loss": (C.c_int, [_H, _P, _P, _U64, _P]),
    "cpra_hip_amp_grad": (C.c_int, [_H, _P, _P, _U64, _P, _P]),
    "cpra_hip_amp_loss_3d": (C.c_int, [_H, _P, _P, _U64, _P]),
    "cpra_hip_amp_grad_3d": (C.c_int, [_H, _P, _P, _U64, _P, _P]),
    "cpra_hip_amp_gd_run": (C.c_int, [_H, _P, _P, _U64, C.c_int, C.c_float, _P]),
    "cpra_hip_amp_gd_run_3d": (C.c_int, [_H, _P, _P, _U64, C.c_int, C.c_float, _P]),
    "cpra_hip_carousel_default_params": (None, [C.POINTER(CarouselParams)]),
    "cpra_hip_dm_mixed_run": (C.c_int, [_H, C.POINTER(SizeClass), C.c_int, C.POINTER(DmParams), C.c_int]),
    "cpra_hip_carousel3d_run": (C.c_int, [_H, _P, _P, _U64, _P, _P, _P, _P,
                                          C.POINTER(CarouselParams), _P]),
}


def exported_symbols() -> Sequence[str]:
    return tuple(_SIGS)


def lib() -> C.CDLL:
    """Load libcpra_hip.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with build.py (there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class DeviceArray:
    """Device buffer owned through the library's own allocator
    (cpra_hip_malloc), with numpy upload/download.  Lets tests and benchmarks
    run without importing torch, so only one HIP runtime lives in the
    process."""

    def __init__(self, shape, dtype):
        import numpy as np

        self.shape = tuple(int(s) for s in (shape if hasattr(shape, "__len__") else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = C.c_void_p()
        rc = lib().cpra_hip_malloc(None, C.byref(p), max(self.nbytes, 1), 0)
        if rc < 0:
            raise CpraHipError(rc, lib().cpra_hip_last_error(None).decode())
        self.ptr = p.value

    @property
    def size(self):
        import numpy as np

        return int(np.prod(self.shape, dtype=np.int64))

    def numel(self):
        return self.size

    def is_complex(self):
        return self.dtype.kind == "c"

    def data_ptr(self):
        return self.ptr

    def upload(self, a):
        import numpy as np

        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        rc = lib().cpra_hip_memcpy(None, self.ptr, a.ctypes.data, self.nbytes)
        if rc < 0:
            raise CpraHipError(rc, lib().cpra_hip_last_error(None).decode())
        return self

    def numpy(self):
        import numpy as np

        out = np.empty(self.shape, dtype=self.dtype)
        rc = lib().cpra_hip_memcpy(None, out.ctypes.data, self.ptr, self.nbytes)
        if rc < 0:
            raise CpraHipError(rc, lib().cpra_hip_last_error(None).decode())
        return out

    def free(self):
        if self.ptr:
            lib().cpra_hip_free(None, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def to_device(a) -> "DeviceArray":
    import numpy as np

    a = np.ascontiguousarray(a)
    return DeviceArray(a.shape, a.dtype).upload(a)


def zeros(shape, dtype) -> "DeviceArray":
    import numpy as np

    return to_device(np.zeros(shape, dtype=dtype))


def _addr(x) -> Optional[int]:
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if isinstance(x, DeviceArray):
        return x.ptr
    if hasattr(x, "data_ptr"):
        if hasattr(x, "is_cuda") and not x.is_cuda:
            raise CpraHipError(ERR_ARG, "tensor must be on the GPU")
        if hasattr(x, "is_contiguous") and not x.is_contiguous():
            raise CpraHipError(ERR_ARG, "tensor must be contiguous")
        return x.data_ptr()
    raise TypeError(f"cannot pass {type(x)} as a device pointer")


def default_hio_params(**kw) -> HioParams:
    p = HioParams()
    lib().cpra_hip_hio_default_params(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v)
    return p


class HipImpl:
    """One handle of the HIP backend; geometry fixed at construction like
    ``NewCpraImpl<T>(IMPL_TYPE::HIP, m, n, l, batch_size)``."""

    def __init__(self, m: int, n: int, l: int = 1, batch_size: int = 1, dtype: str = "float32",
                 stream=None, use_torch_stream: bool = False):
        self.m, self.n, self.l, self.batch = int(m), int(n), int(l), int(batch_size)
        self.dtype = dtype
        code = F64 if dtype in ("float64", "double") else F32
        h = C.c_void_p()
        rc = lib().cpra_hip_create(C.byref(h), code, self.m, self.n, self.l, self.batch)
        if rc < 0:
            raise CpraHipError(rc, lib().cpra_hip_last_error(None).decode())
        self._h = h
        if stream is not None:
            self.set_stream(stream)
        elif use_torch_stream:
            import torch

            self.set_stream(torch.cuda.current_stream().cuda_stream)

    # -- plumbing ---------------------------------------------------------
    def _check(self, rc: int) -> bool:
        if rc < 0:
            raise CpraHipError(rc, lib().cpra_hip_last_error(self._h).decode())
        return rc == OK

    def close(self):
        if getattr(self, "_h", None):
            lib().cpra_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def set_stream(self, stream) -> bool:
        s = stream if isinstance(stream, int) else getattr(stream, "cuda_stream", stream)
        return self._check(lib().cpra_hip_set_stream(self._h, s))

    # -- CpraImpl<T> mirror (include/CPRA_Impl.hpp:23-61) -------------------
    def Initialize(self, data, num: int, seed: int = 0) -> bool:
        return self._check(lib().cpra_hip_initialize(self._h, _addr(data), num, seed))

    def Forward2D(self, x) -> bool:
        return self._check(lib().cpra_hip_forward2d(self._h, _addr(x)))

    def Backward2D(self, x) -> bool:
        return self._check(lib().cpra_hip_backward2d(self._h, _addr(x)))

    def FFTShift2D(self, x, M: int, N: int, Batch: int) -> bool:
        return self._check(lib().cpra_hip_fftshift2d(self._h, _addr(x), M, N, Batch))

    def Forward3D(self, x) -> bool:
        return self._check(lib().cpra_hip_forward3d(self._h, _addr(x)))

    def Backward3D(self, x) -> bool:
        return self._check(lib().cpra_hip_backward3d(self._h, _addr(x)))

    def FFTShift3D(self, x, M: int, N: int, L: int, Batch: int) -> bool:
        return self._check(lib().cpra_hip_fftshift3d(self._h, _addr(x), M, N, L, Batch))

    def SpaceConstraint(self, x, S, num: int, batch_size: int, mode: int = SPACE_MKL) -> bool:
        return self._check(
            lib().cpra_hip_space_constraint(self._h, _addr(x), _addr(S), num, batch_size, mode))

    def DataConstraint(self, x, A, num: int, batch_size: int, skip_zero: bool = True) -> bool:
        """Real amplitude (float A) or complex (complex A) data constraint,
        dispatched on A's dtype like the two C++ overloads."""
        if hasattr(A, "is_complex") and A.is_complex():
            return self._check(lib().cpra_hip_data_constraint_complex(
                self._h, _addr(x), _addr(A), num, batch_size))
        return self._check(lib().cpra_hip_data_constraint_real(
            self._h, _addr(x), _addr(A), num, batch_size, int(skip_zero)))

    def MergeAddData(self, src, dst, alpha: float, beta: float, num: int) -> bool:
        return self._check(
            lib().cpra_hip_merge_add(self._h, _addr(src), _addr(dst), alpha, beta, num))

    def Normalization(self, x, norm: float, num: int) -> bool:
        return self._check(lib().cpra_hip_normalize(self._h, _addr(x), norm, num))

    def Real2DTo3DInterpolation(self, slices, vol, angles, m: int, n: int, p: int, l: int) -> bool:
        return self._check(lib().cpra_hip_real2d_to_3d(self._h, _addr(slices), _addr(vol),
                                                       _addr(angles), m, n, p, l))

    def Complex2DTo3DInterpolation(self, slices, vol, angles, m: int, n: int, p: int,
                                   l: int) -> bool:
        return self._check(lib().cpra_hip_complex2d_to_3d(self._h, _addr(slices), _addr(vol),
                                                          _addr(angles), m, n, p, l))

    def slice_extract(self, vol, slices, angles, n: int, p: int, complex_data=True) -> bool:
        return self._check(lib().cpra_hip_slice_extract(self._h, _addr(vol), _addr(slices),
                                                        _addr(angles), n, p, int(complex_data)))

    def ConvergeError(self, old, new, m: int, n: int, l: int = 1, batch_size: int = 1):
        """Returns (ok, errors) with errors a numpy array of batch_size."""
        import numpy as np
        err = np.zeros(batch_size, dtype=np.float64 if self.dtype in ("float64", "double")
                       else np.float32)
        ok = self._check(lib().cpra_hip_converge_error(
            self._h, _addr(old), _addr(new), err.ctypes.data, m, n, l, batch_size))
        return ok, err

    def Memcpy(self, dst, src, nbytes: int) -> bool:
        return self._check(lib().cpra_hip_memcpy(self._h, _addr(dst), _addr(src), nbytes))

    def Sync(self) -> bool:
        return self._check(lib().cpra_hip_sync(self._h))

    # ---- graph capture of the handle's stream ---------------------------
    def graph_begin(self) -> bool:
        return self._check(lib().cpra_hip_graph_begin(self._h))

    def graph_end(self) -> int:
        gid = C.c_uint64(0)
        self._check(lib().cpra_hip_graph_end(self._h, C.byref(gid)))
        return gid.value

    def graph_launch(self, graph_id: int) -> bool:
        return self._check(lib().cpra_hip_graph_launch(self._h, graph_id))

    def graph_destroy(self, graph_id: int) -> bool:
        return self._check(lib().cpra_hip_graph_destroy(self._h, graph_id))

    # -- fused engines (include/cpra_hip_engine.h) -------------------------
    def hio_run(self, x, amp, support, params: Optional[HioParams] = None, amp_batch=None,
                want_rfactor: bool = False, **kw):
        import numpy as np
        p = params if params is not None else default_hio_params(**kw)
        ab = amp_batch if amp_batch is not None else amp.numel() // (self.m * self.n)
        rf = np.zeros(self.batch, dtype=np.float32) if want_rfactor else None
        self._check(lib().cpra_hip_hio_run(self._h, _addr(x), _addr(amp), ab, _addr(support),
                                           C.byref(p), rf.ctypes.data if rf is not None else None))
        return rf

    def dm_run(self, psi, amp, support, t1=None, t2=None, amp_batch=None, noalloc=False,
               **kw) -> bool:
        """cpra_hip_dm_run, or with noalloc=True cpra_hip_dm_run_noalloc (the
        iteration state in the library's static scratch array)."""
        p = DmParams()
        lib().cpra_hip_dm_default_params(C.byref(p))
        for k, v in kw.items():
            if not hasattr(p, k):
                raise AttributeError(k)
            setattr(p, k, v)
        ab = amp_batch if amp_batch is not None else amp.numel() // (self.m * self.n)
        fn = lib().cpra_hip_dm_run_noalloc if noalloc else lib().cpra_hip_dm_run
        return self._check(fn(self._h, _addr(psi), _addr(t1), _addr(t2), _addr(amp), ab,
                              _addr(support), C.byref(p)))

    def dm3d_run(self, psi, amp, support, t1=None, t2=None, amp_batch=None, **kw) -> bool:
        """Difference map on volumes [batch][l][m][n] (cpra_hip_dm3d_run)."""
        p = DmParams()
        lib().cpra_hip_dm_default_params(C.byref(p))
        for k, v in kw.items():
            if not hasattr(p, k):
                raise AttributeError(k)
            setattr(p, k, v)
        ab = amp_batch if amp_batch is not None else amp.numel() // (self.m * self.n * self.l)
        return self._check(lib().cpra_hip_dm3d_run(self._h, _addr(psi), _addr(t1), _addr(t2),
                                                   _addr(amp), ab, _addr(support), C.byref(p)))

    def carousel3d_run(self, x, amp, support2d, angles, support3d, vol, amp_batch=None,
                       want_rfactor: bool = False, **kw):
        """3D carousel reconstruction (cpra_hip_carousel3d_run): x [P][n][n]
        projections in/out, vol [n][n][n] receives the object.  Keyword
        arguments set cpra_hip_carousel_params fields.  Returns the final
        R-factors (P floats) when want_rfactor."""
        import numpy as np
        p = CarouselParams()
        lib().cpra_hip_carousel_default_params(C.byref(p))
        for k, v in kw.items():
            if not hasattr(p, k):
                raise AttributeError(k)
            setattr(p, k, v)
        ab = amp_batch if amp_batch is not None else amp.numel() // (self.m * self.n)
        rf = np.zeros(self.batch, dtype=np.float32) if want_rfactor else None
        self._check(lib().cpra_hip_carousel3d_run(
            self._h, _addr(x), _addr(amp), ab, _addr(support2d), _addr(angles), _addr(support3d),
            _addr(vol), C.byref(p), rf.ctypes.data if want_rfactor else None))
        return rf

    def dm_mixed_run(self, classes, launch_mode: int = MIXED_FUSED, **kw) -> bool:
        """Textbook difference map over size classes (cpra_hip_dm_mixed_run).
        classes: sequence of (psi, amp, support) device arrays [count][n][n]."""
        arr = (SizeClass * len(classes))()
        for i, (psi, amp, sup) in enumerate(classes):
            n = psi.shape[-1]
            arr[i] = SizeClass(n, psi.numel() // (n * n), _addr(psi), _addr(amp), _addr(sup))
        p = DmParams()
        lib().cpra_hip_dm_default_params(C.byref(p))
        p.textbook = 1
        for k, v in kw.items():
            if not hasattr(p, k):
                raise AttributeError(k)
            setattr(p, k, v)
        return self._check(lib().cpra_hip_dm_mixed_run(self._h, arr, len(classes), C.byref(p),
                                                       launch_mode))

    def shrinkwrap(self, x, support, sigma: float, frac: float, blur_out=None) -> bool:
        return self._check(lib().cpra_hip_shrinkwrap(self._h, _addr(x), _addr(support), sigma,
                                                     frac, _addr(blur_out)))

    def rfactor(self, x, amp, amp_batch=None):
        import numpy as np
        ab = amp_batch if amp_batch is not None else amp.numel() // (self.m * self.n)
        r = np.zeros(self.batch, dtype=np.float32)
        self._check(lib().cpra_hip_rfactor(self._h, _addr(x), _addr(amp), ab, r.ctypes.data))
        return r

    def autotune(self, x, amp, support, iters: int = 10, amp_batch=None) -> str:
        """Times the engine's column-kernel variants on this data, keeps the
        fastest for this handle and returns the report string."""
        ab = amp_batch if amp_batch is not None else amp.numel() // (self.m * self.n)
        buf = C.create_string_buffer(1024)
        self._check(lib().cpra_hip_autotune_2d(self._h, _addr(x), _addr(amp), ab,
                                               _addr(support), iters, buf, 1024))
        return buf.value.decode()

    def set_tuning(self, spec: str) -> bool:
        return self._check(lib().cpra_hip_set_tuning(self._h, spec.encode()))

    def get_tuning(self) -> dict:
        buf = C.create_string_buffer(512)
        self._check(lib().cpra_hip_get_tuning(self._h, buf, 512))
        return dict(kv.split("=") for kv in buf.value.decode().strip(";").split(";"))

    def set_kernel_timing(self, enable: bool) -> bool:
        return self._check(lib().cpra_hip_set_kernel_timing(self._h, int(enable)))

    def kernel_times(self):
        """{kind: (total_ms, launches)} per fused-kernel kind."""
        names = ("column", "row", "shrinkwrap", "gd", "merge", "volume", "extract", "proj_transpose",
                 "mixed_iteration", "pencil")
        ms = (C.c_double * len(names))()
        cnt = (C.c_int64 * len(names))()
        self._check(lib().cpra_hip_get_kernel_times(self._h, ms, cnt, len(names)))
        return {k: (ms[i], cnt[i]) for i, k in enumerate(names)}

    def support_from_constraint(self, S, mask, num: int) -> bool:
        return self._check(lib().cpra_hip_support_from_constraint(self._h, _addr(S), _addr(mask),
                                                                  num))

    # ---- amplitude loss L = sum (|F x| - A)^2 and its adjoint -------------
    def _amp_batch(self, amp, amp_batch, vol: bool):
        if amp_batch is not None:
            return amp_batch
        per = self.m * self.n * (self.l if vol else 1)
        return amp.numel() // per

    def amp_loss(self, x, amp, amp_batch=None, volume: bool = False):
        """Per-pattern loss (batch doubles) of x under amplitudes amp."""
        import numpy as np
        out = np.zeros(self.batch, dtype=np.float64)
        fn = lib().cpra_hip_amp_loss_3d if volume else lib().cpra_hip_amp_loss
        self._check(fn(self._h, _addr(x), _addr(amp), self._amp_batch(amp, amp_batch, volume),
                       out.ctypes.data))
        return out

    def amp_grad(self, x, amp, grad, amp_batch=None, volume: bool = False):
        """grad <- dL/d(re x) + i dL/d(im x); returns the per-pattern loss."""
        import numpy as np
        out = np.zeros(self.batch, dtype=np.float64)
        fn = lib().cpra_hip_amp_grad_3d if volume else lib().cpra_hip_amp_grad
        self._check(fn(self._h, _addr(x), _addr(amp), self._amp_batch(amp, amp_batch, volume),
                       _addr(grad), out.ctypes.data))
        return out

    def amp_gd_run(self, x, amp, iters: int, mu: float, amp_batch=None, volume: bool = False,
                   want_loss: bool = True):
        """iters fused gradient-descent steps on x in place; returns the loss
        of the final iterate (or None)."""
        import numpy as np
        out = np.zeros(self.batch, dtype=np.float64) if want_loss else None
        fn = lib().cpra_hip_amp_gd_run_3d if volume else lib().cpra_hip_amp_gd_run
        self._check(fn(self._h, _addr(x), _addr(amp), self._amp_batch(amp, amp_batch, volume),
                       int(iters), float(mu), out.ctypes.data if want_loss else None))
        return out


def set_tuning_cache(path, refresh: bool = False) -> None:
    """The autotuner's persistent JSON cache for this process (None: off)."""
    rc = lib().cpra_hip_set_tuning_cache(path.encode() if path else None, int(refresh))
    if rc < 0:
        raise CpraHipError(rc, lib().cpra_hip_last_error(None).decode())


def device_count() -> int:
    return lib().cpra_hip_device_count()


def version() -> str:
    return lib().cpra_hip_version().decode()
