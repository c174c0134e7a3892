"""The reference's facade defaults on the HIP backend: managed buffers
(NewHIPAllocator(true), what the reference's NewAllocator(type) returns with
its default gpu_managed = true once the HIP case is added) filled and read by
host code, the DM body and final projection through HipImpl<float>
(tests/cpp/test_managed_facade.cpp), against the FP64 oracle's orc_dm_ref,
orc_converge_error and orc_final_projection on the same inputs.

Also, without a GPU: HIP_Impl.hpp compiled inside the reference tree
(CPRA_HIP_IN_REFERENCE_TREE against the reference's own include/), and the
backend's registration (IMPL_TYPE::HIP in CPRA_Impl.hpp, its case in
Impl_Factory.hpp and in Allocator.hpp's NewAllocator) applied to a scratch
copy of the reference's headers, compiled with Cpra<float, IMPL_TYPE::HIP>
used as the reference's programs use it.  Both skip where the reference is
absent."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "bin", "test_managed_facade")
REF_INCLUDE = "/root/reference/include"


@pytest.mark.gpu
@pytest.mark.parametrize("M,N,B,preit", [(64, 64, 2, 3), (96, 128, 3, 2)])
def test_managed_buffers_through_the_facade(oracle, gpu, tmp_path, M, N, B, preit):
    assert os.path.exists(EXE), "tests/cpp/bin/test_managed_facade missing: run the build"
    rng = np.random.default_rng(M + N + B)
    yy, xx = np.mgrid[0:M, 0:N]
    disc = (((yy - M / 2) / (M / 5)) ** 2 + ((xx - N / 2) / (N / 5)) ** 2 <= 1).astype(np.float32)
    A = np.abs(np.fft.fftshift(np.fft.fft2(rng.random((M, N)) * disc))).astype(np.float32)
    S = np.where(disc > 0, 1.0, rng.random((M, N)) * 0.8).astype(np.float32)
    psi0 = (rng.standard_normal((B, M, N)) + 1j * rng.standard_normal((B, M, N))).astype(np.complex64)
    A.tofile(tmp_path / "A.bin")
    S.tofile(tmp_path / "S.bin")
    np.ascontiguousarray(psi0.real).tofile(tmp_path / "psi_re.bin")
    np.ascontiguousarray(psi0.imag).tofile(tmp_path / "psi_im.bin")
    print(f"managed facade M={M} N={N} B={B} preit={preit}", flush=True)
    r = subprocess.run([EXE, str(tmp_path), str(M), str(N), str(B), str(preit), "0.9"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    psi = np.fromfile(tmp_path / "psi.bin", np.complex64).reshape(B, M, N)
    real = np.fromfile(tmp_path / "real.bin", np.float32).reshape(B, M, N)
    err = np.fromfile(tmp_path / "err.bin", np.float32)
    # FP64 oracle: preit - 1 iterations, the last one with its error, then
    # the final projection
    p, t1, t2 = oracle.dm_ref(psi0, psi0, psi0, A[None], S[None], 0.9, preit - 1)
    old = p.copy()
    p, _, _ = oracle.dm_ref(p, t1, t2, A[None], S[None], 0.9, 1)
    e_ref, _ = oracle.converge_error(oracle.forward2d(old), oracle.forward2d(p), M, N, 1, B)
    p = oracle.final_projection(p, A[None], S[None], 0.9)
    rel = np.linalg.norm(psi - p) / np.linalg.norm(p)
    print(f"  psi rel L2 {rel:.2e}, err {err} vs {e_ref}")
    assert rel < 2e-5 * (preit + 1)
    assert np.array_equal(real, psi.real)
    assert np.allclose(err, e_ref, rtol=1e-4)


def _need_reference():
    if not os.path.isdir(REF_INCLUDE):
        pytest.skip("the reference tree is not present")


def test_hip_impl_compiles_inside_the_reference_tree(tmp_path):
    """HIP_Impl.hpp with CPRA_HIP_IN_REFERENCE_TREE takes CpraImpl<T>,
    AllocatorInterface and IMPL_TYPE from the reference's own headers."""
    _need_reference()
    src = tmp_path / "t.cpp"
    src.write_text('#include "HIP_Impl.hpp"\n#include "Allocator.hpp"\n#include "HIP_Allocator.hpp"\n'
                   "int main(){\n"
                   "  std::unique_ptr<CPRA::CpraImpl<float>> a = CPRA::NewHIPImpl<float>(8, 8, 1, 1);\n"
                   "  std::unique_ptr<CPRA::CpraImpl<double>> b = CPRA::NewHIPImpl<double>(8, 8, 1, 1);\n"
                   "  std::unique_ptr<CPRA::AllocatorInterface> m = CPRA::NewHIPAllocator(true);\n"
                   "  return 0;\n}\n")
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-DCPRA_HIP_IN_REFERENCE_TREE",
                    "-I", REF_INCLUDE, "-I", os.path.join(ROOT, "include"), str(src)], check=True)


def _apply_integration_recipe(inc):
    """The backend's registration in three reference headers (the HIP
    enumerator, the factory case, the allocator case honouring
    gpu_managed), applied to a scratch copy of the reference's include/."""
    p = inc / "CPRA_Impl.hpp"
    s = p.read_text()
    s2 = s.replace("enum IMPL_TYPE {CUDA, MKL};", "enum IMPL_TYPE {CUDA, MKL, HIP};")
    assert s2 != s
    p.write_text(s2)
    p = inc / "Impl_Factory.hpp"
    s = p.read_text()
    s2 = s.replace('#ifdef HAS_MKL\n#include "MKL_Impl.hpp"\n#endif',
                   '#ifdef HAS_MKL\n#include "MKL_Impl.hpp"\n#endif\n\n#ifdef HAS_HIP\n'
                   '#include "HIP_Impl.hpp"\n#endif')
    s2, k = re.subn(r"(\n\s*default:\n\s*throw std::invalid_argument\(\"Wrong impl type!)",
                    "\n#ifdef HAS_HIP\n        case IMPL_TYPE::HIP:\n"
                    "            return NewHIPImpl<T>(m, n, l, batch_size);\n#endif\\1", s2)
    assert k == 1 and "HIP_Impl.hpp" in s2
    p.write_text(s2)
    p = inc / "Allocator.hpp"
    s = p.read_text()
    s = s.replace('#include "CPRA_Impl.hpp"\n',
                  '#include "CPRA_Impl.hpp"\n#ifdef HAS_HIP\n#include "HIP_Impl.hpp"\n#endif\n', 1)
    s += '\n#ifdef HAS_HIP\n#include "HIP_Allocator.hpp"\n#endif\n'
    s2, k = re.subn(r"(\n\s*default:\n\s*throw std::invalid_argument\(\"Wrong allocator type!)",
                    "\n#ifdef HAS_HIP\n        case IMPL_TYPE::HIP:\n"
                    "            return NewHIPAllocator(gpu_managed);\n#endif\\1", s)
    assert k == 1
    p.write_text(s2)


def test_integration_recipe_compiles_with_the_reference_facade(tmp_path):
    """Cpra<float, IMPL_TYPE::HIP> used as src/CPRA_MKL.cpp uses Cpra<float,
    MKL>: allocate (managed by default), ReadMatrixFromFile, RealToComplex,
    the DM primitives through impl_, ConvergeError, ComplexToReal,
    WriteMatrixToFile.  Compiled (not run) against the reference's headers
    with the recipe applied; the reference tree itself is not modified."""
    _need_reference()
    inc = tmp_path / "include"
    shutil.copytree(REF_INCLUDE, inc)
    _apply_integration_recipe(inc)
    src = tmp_path / "driver.cpp"
    src.write_text(r'''
#include "CPRA.hpp"
int main() {
    using C = std::complex<float>;
    CPRA::Cpra<float, CPRA::IMPL_TYPE::HIP> c(64, 64, 1, 2);
    const uint64_t n = 64 * 64 * 2;
    float* A = (float*)c.allocate(sizeof(float) * 64 * 64);
    float* S = (float*)c.allocate(sizeof(float) * 64 * 64);
    float* re = (float*)c.allocate(sizeof(float) * n);
    C* psi = (C*)c.allocate(sizeof(C) * n);
    C* t = (C*)c.allocate(sizeof(C) * n);
    float* err = (float*)c.allocate(sizeof(float) * 2);
    c.ReadMatrixFromFile("A.bin", A, 64, 64, 1);
    c.ReadMatrixFromFile("S.bin", S, 64, 64, 1);
    c.ReadMatrixFromFile("x.bin", re, 64, 64, 2);
    c.RealToComplex(re, psi, n);
    c.impl_->Memcpy(t, psi, sizeof(C) * n);
    c.impl_->Forward2D(t);
    c.impl_->DataConstraint(t, A, n, 2);
    c.impl_->Backward2D(t);
    c.impl_->MergeAddData(psi, t, -1.0f / 0.9f, 1.0f + 1.0f / 0.9f, n);
    c.impl_->SpaceConstraint(t, S, n, 2);
    c.impl_->ConvergeError(psi, t, err, 64, 64, 1, 2);
    c.impl_->Sync();
    c.ComplexToReal(t, re, n);
    c.WriteMatrixToFile("out.bin", re, 64, 64, 2);
    for (void* p : {(void*)A, (void*)S, (void*)re, (void*)psi, (void*)t, (void*)err}) c.deallocate(p);
    return 0;
}
''')
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-DHAS_HIP", "-DCPRA_HIP_IN_REFERENCE_TREE",
                    "-I", str(inc),
                    "-I", os.path.join(ROOT, "include"), str(src)], check=True)
