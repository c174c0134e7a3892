"""C-ABI checks that need no GPU: the library loads, exports every symbol
the headers declare, and rejects bad arguments before any HIP call."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", f) for f in ("cpra_hip.h", "cpra_hip_engine.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(cpra_hip_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_functions():
    assert len(declared_functions()) >= 25


def test_library_exports_every_declared_symbol(hipmod):
    L = C.CDLL(hipmod.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    # and the Python front end binds exactly the declared set
    assert set(hipmod.exported_symbols()) == set(declared_functions())


def test_null_handle_is_rejected(hipmod):
    L = hipmod.lib()
    assert L.cpra_hip_forward2d(None, None) == hipmod.ERR_ARG
    assert L.cpra_hip_merge_add(None, None, None, 1.0, 1.0, 4) == hipmod.ERR_ARG
    assert L.cpra_hip_sync(None) == hipmod.ERR_ARG
    assert L.cpra_hip_destroy(None) == hipmod.ERR_ARG
    assert L.cpra_hip_hio_run(None, None, None, 1, None, None, None) == hipmod.ERR_ARG
    assert b"NULL" in L.cpra_hip_last_error(None) or L.cpra_hip_last_error(None) is not None


def test_create_validates_before_touching_the_gpu(hipmod):
    L = hipmod.lib()
    h = C.c_void_p()
    assert L.cpra_hip_create(C.byref(h), 7, 8, 8, 1, 1) == hipmod.ERR_ARG
    assert b"dtype" in L.cpra_hip_last_error(None)
    assert L.cpra_hip_create(C.byref(h), 0, 0, 8, 1, 1) == hipmod.ERR_ARG
    assert L.cpra_hip_create(None, 0, 8, 8, 1, 1) == hipmod.ERR_ARG


def test_create_without_gpu_fails_cleanly(hipmod):
    if hipmod.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(hipmod.CpraHipError) as e:
        hipmod.HipImpl(8, 8, 1, 1)
    assert e.value.code == hipmod.ERR_NODEV


def test_hio_params_layout_matches_header(hipmod):
    # the ctypes mirror must match the C struct (10 x 4 bytes)
    assert C.sizeof(hipmod.HioParams) == 40
    p = hipmod.default_hio_params()
    assert p.iters == 200 and abs(p.beta - 0.9) < 1e-7 and p.sw_every == 20


def test_graph_capture_entry_points_validate(hipmod):
    L = hipmod.lib()
    gid = C.c_uint64(0)
    assert L.cpra_hip_graph_begin(None) == hipmod.ERR_ARG
    assert L.cpra_hip_graph_end(None, C.byref(gid)) == hipmod.ERR_ARG
    assert L.cpra_hip_graph_launch(None, 1) == hipmod.ERR_ARG
    assert L.cpra_hip_graph_destroy(None, 1) == hipmod.ERR_ARG
