"""ctypes front end of oracle/mkl_path.c: the reference's MKL backend in FP32
over libmkl_rt (TEST INFRASTRUCTURE AND CPU BASELINE ONLY).

Only tests/ and the ``cpu_baseline`` leg of ``bench.py`` use this module; the
HIP library and its front ends never do.  ``available()`` is False where
libmkl_rt cannot be loaded; callers then fall back or skip.
"""
from __future__ import annotations

import ctypes as C
import os
import statistics
import subprocess
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmkl_path.so")
_lib = None
_ok = None

_P = C.c_void_p
_I64 = C.c_int64
_U64 = C.c_uint64
_F = C.c_float


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            subprocess.run(["make", "-s", "-C", _HERE, "libmkl_path.so"], check=True)
        L = C.CDLL(_LIB_PATH)
        sig = {
            "mklp_load": (C.c_int, [C.c_char_p]),
            "mklp_last_error": (C.c_char_p, []),
            "mklp_library_path": (C.c_char_p, []),
            "mklp_set_threads": (None, [C.c_int]),
            "mklp_max_threads": (C.c_int, []),
            "mklp_create": (_P, [_I64, _I64, _I64]),
            "mklp_destroy": (None, [_P]),
            "mklp_fftshift2d": (None, [_P, _I64, _I64, _I64]),
            "mklp_forward2d": (C.c_int, [_P, _P]),
            "mklp_backward2d": (C.c_int, [_P, _P]),
            "mklp_data_constraint": (None, [_P, _P, _U64, _U64]),
            "mklp_merge_add": (C.c_int, [_P, _P, _F, _F, _U64]),
            "mklp_hio": (C.c_int, [_P, _P, _P, _I64, _P, _I64, _F, C.c_int, _I64, _F, _F, _F, _P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def available() -> bool:
    global _ok
    if _ok is None:
        _ok = bool(lib().mklp_load(os.environ.get("CPRA_MKL_RT", "").encode() or None))
    return _ok


def load_error() -> str:
    return lib().mklp_last_error().decode()


def library_path() -> str:
    return lib().mklp_library_path().decode()


def set_threads(n: int):
    lib().mklp_set_threads(int(n))


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.complex64)


def _err(what):
    raise RuntimeError(f"{what}: {lib().mklp_last_error().decode()}")


class Plan:
    """The reference's 2D descriptor for `batch` patterns of m x n."""

    def __init__(self, m: int, n: int, batch: int):
        if not available():
            _err("libmkl_rt")
        self.m, self.n, self.batch = m, n, batch
        self.h = lib().mklp_create(m, n, batch)
        if not self.h:
            _err("mklp_create")

    def close(self):
        if self.h:
            lib().mklp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def forward2d(self, x):
        x = _c64(x).copy()
        assert x.size == self.m * self.n * self.batch
        if not lib().mklp_forward2d(self.h, x.ctypes.data):
            _err("forward2d")
        return x

    def backward2d(self, x):
        x = _c64(x).copy()
        assert x.size == self.m * self.n * self.batch
        if not lib().mklp_backward2d(self.h, x.ctypes.data):
            _err("backward2d")
        return x

    def hio(self, x, A, support, iters, beta=0.9, real_positive=True, sw_every=0, sw_sigma0=3.0,
            sw_sigma1=1.5, sw_frac=0.1):
        """HIO + shrinkwrap on the plan's batch; returns (x, support)."""
        x = _c64(x).copy()
        A = np.ascontiguousarray(A, dtype=np.float32)
        s = np.ascontiguousarray(support, dtype=np.uint8).copy()
        assert x.size == self.m * self.n * self.batch and s.size == x.size
        work = np.empty(x.size, np.complex64)
        ok = lib().mklp_hio(self.h, x.ctypes.data, A.ctypes.data, A.size // (self.m * self.n),
                            s.ctypes.data, iters, beta, int(real_positive), sw_every, sw_sigma0,
                            sw_sigma1, sw_frac, work.ctypes.data)
        if not ok:
            _err("hio")
        return x, s


def fftshift2d(x, M, N, B):
    x = _c64(x).copy()
    lib().mklp_fftshift2d(x.ctypes.data, M, N, B)
    return x


def data_constraint(x, A, batch_size):
    x = _c64(x).copy()
    A = np.ascontiguousarray(A, dtype=np.float32)
    lib().mklp_data_constraint(x.ctypes.data, A.ctypes.data, x.size, batch_size)
    return x


def merge_add(src, dst, alpha, beta):
    src = _c64(src)
    dst = _c64(dst).copy()
    ok = lib().mklp_merge_add(src.ctypes.data, dst.ctypes.data, alpha, beta, dst.size)
    return dst, bool(ok)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def physical_cores() -> int:
    """Physical cores this process may run on (SMT siblings counted once)."""
    try:
        cpus = sorted(os.sched_getaffinity(0))
    except AttributeError:
        cpus = list(range(os.cpu_count() or 1))
    seen = set()
    for c in cpus:
        try:
            with open(f"/sys/devices/system/cpu/cpu{c}/topology/thread_siblings_list") as f:
                seen.add(f.read().strip())
        except OSError:
            seen.add(str(c))
    return max(1, len(seen))


def time_hio_schedule(x, A, support, threads, sched, min_seconds=10.0, reps=3, warmup_iters=2):
    """The reference's CPU path timed on the HIO + shrinkwrap schedule:
    `warmup_iters` untimed iterations, then reps runs of the whole schedule
    (repeated until at least `min_seconds` and at least 20 iterations have
    run), each from the same state; returns the median pattern-iterations/s
    and the per-rep figures."""
    B = x.shape[0]
    n = x.shape[-1]
    set_threads(threads)
    plan = Plan(n, n, B)
    kw = dict(beta=sched["beta"], real_positive=bool(sched["real_positive"]),
              sw_every=sched["sw_every"], sw_sigma0=sched["sw_sigma0"],
              sw_sigma1=sched["sw_sigma1"], sw_frac=sched["sw_frac"])
    plan.hio(x, A, support, warmup_iters, **{**kw, "sw_every": 0})
    iters = sched["iters"]
    rates, runs = [], []
    for _ in range(reps):
        done, t0 = 0, time.perf_counter()
        while True:
            plan.hio(x, A, support, iters, **kw)
            done += iters
            el = time.perf_counter() - t0
            if el >= min_seconds and done >= 20:
                break
        rates.append(B * done / el)
        runs.append(done // iters)
    plan.close()
    return statistics.median(rates), rates, runs
