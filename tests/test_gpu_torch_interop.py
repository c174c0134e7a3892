"""The library used from a process that also runs torch on the GPU (torch's
wheel bundles its own HIP runtime, which the library then shares).  Runs in a
child process so a failure cannot take the test session down."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys, numpy as np, torch
sys.path.insert(0, sys.argv[1])
import __graft_entry__ as ge
ge.load_package()
from cpra_amd import hip
assert torch.cuda.is_available()
x = (np.random.default_rng(0).standard_normal((2, 64, 64)) * (1 + 0.5j)).astype(np.complex64)
t = torch.from_numpy(x).cuda()
with hip.HipImpl(64, 64, 1, 2, use_torch_stream=True) as h:
    assert h.Forward2D(t)
    torch.cuda.synchronize()
ref = np.fft.fftshift(np.fft.fft2(np.fft.fftshift(x.astype(complex), axes=(1, 2))), axes=(1, 2))
err = np.linalg.norm(t.cpu().numpy() - ref) / np.linalg.norm(ref)
print("interop rel err", err)
assert err < 1e-5
"""


def test_library_shares_torch_runtime(gpu):
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr[-2000:]
