"""Host emulation of the gfx950 FFT building blocks (no GPU needed).

tests/cpp/fft_core_emu.cpp runs the device code paths of csrc/fft_core.hpp
on the CPU, one emulated thread after another between the points where the
kernels synchronise, for every radix plan the kernels use, and compares with
a direct long-double DFT."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def emu_binary(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("emu") / "fft_core_emu"
    src = os.path.join(ROOT, "tests", "cpp", "fft_core_emu.cpp")
    subprocess.run([hipcc, "-x", "hip", "--offload-host-only", "-O1", "-std=c++17", "-o",
                    str(out), src], check=True, capture_output=True)
    return str(out)


def test_all_plans_match_direct_dft(emu_binary):
    r = subprocess.run([emu_binary], check=True, capture_output=True, text=True)
    rows = [ln.split() for ln in r.stdout.strip().splitlines()]
    assert len(rows) >= 60
    for n, e, d, dt, err, *name in rows:
        tol = 4e-7 if dt == "f32" else 1e-15
        assert float(err) < tol, (n, e, d, dt, err, name)
