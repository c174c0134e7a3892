"""HipImpl<T> (include/HIP_Impl.hpp), the C++ drop-in for the reference's
CpraImpl<T>, exercised by tests/cpp/test_hipimpl.cpp: the reference's own
unit tests (test/unitest_mkl.cpp, test/unitest_cuda.cu) and a short run of
its DM composition (src/CPRA_MKL.cpp:94-110) through the virtual interface."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "bin", "test_hipimpl")


def _exe():
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/bin/test_hipimpl missing: run __graft_entry__.build()")
    return EXE


def test_header_compiles_without_hip_headers(tmp_path):
    # HIP_Impl.hpp must be usable from plain g++ (no hip/*.h on the path)
    src = tmp_path / "t.cpp"
    src.write_text('#include "HIP_Impl.hpp"\n'
                   'int main(){ CPRA::CpraImpl<float>* p = nullptr; (void)p; return 0; }\n')
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                    str(src)], check=True)


def test_without_gpu_constructor_throws(hipmod):
    if hipmod.device_count() > 0:
        pytest.skip("a GPU is visible")
    r = subprocess.run([_exe()], capture_output=True, text=True)
    assert r.returncode == 2 and "no GPU" in r.stdout


@pytest.mark.gpu
def test_reference_unit_tests_on_hip(gpu):
    r = subprocess.run([_exe()], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED" in r.stdout
