"""Verification harness: the library's amplitude loss and adjoint wrapped as a
torch.autograd.Function, checked with torch.autograd.gradcheck in FP64.

Run as a child process (``python tests/amp_autograd_harness.py <repo root>``)
by tests/test_gpu_ampgrad.py: torch brings its own HIP runtime, which the
library then shares (code object v5), and keeping torch out of the pytest
process keeps the other GPU tests on a single runtime.

The loss is L_b = sum_k (|F x_b|_k - A_k)^2 (F the centred unnormalised DFT).
forward saves only x and A; backward recomputes F x from x through
cpra_hip_amp_grad (no transform of x is kept between the passes), which is
the recompute-in-backward scheme used for 3D volumes.
"""
import sys

import numpy as np
import torch


def main(root: str) -> int:
    sys.path.insert(0, root)
    import __graft_entry__ as ge

    ge.load_package()
    from cpra_amd import hip

    assert torch.cuda.is_available(), "harness needs the GPU"
    dev = torch.device("cuda:0")

    class AmpLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, A, handle, volume):
            x = x.contiguous()
            out = handle.amp_loss(x, A, volume=volume)
            ctx.save_for_backward(x, A)
            ctx.handle, ctx.volume = handle, volume
            return torch.from_numpy(out).to(dev)

        @staticmethod
        def backward(ctx, gout):
            x, A = ctx.saved_tensors
            g = torch.empty_like(x)
            ctx.handle.amp_grad(x, A, g, volume=ctx.volume)
            torch.cuda.synchronize()
            shape = (-1,) + (1,) * (x.dim() - 1)
            # torch's complex convention: dL/d(re) + i dL/d(im), scaled by the
            # (real) upstream gradient of each pattern's loss
            return g * gout.reshape(shape).to(g.dtype), None, None, None

    rng = np.random.default_rng(7)
    results = []
    # 2D: 16 x 64^2 FP64 (central differences, h = 1e-6)
    B, n = 16, 64
    obj = np.zeros((B, n, n))
    yy, xx = np.mgrid[:n, :n]
    for b in range(B):
        r = rng.uniform(8, 12)
        obj[b] = ((yy - n / 2) ** 2 + (xx - n / 2) ** 2 <= r * r) * rng.random((n, n))
    F = np.fft.fftshift(np.fft.fft2(np.fft.fftshift(obj, axes=(1, 2))), axes=(1, 2))
    A = torch.from_numpy(np.abs(F)).to(dev)
    x0 = obj + 0.05 * (rng.standard_normal((B, n, n)) + 1j * rng.standard_normal((B, n, n)))
    x = torch.from_numpy(x0).to(dev).requires_grad_(True)
    with hip.HipImpl(n, n, 1, B, dtype="float64", use_torch_stream=True) as h:
        ok = torch.autograd.gradcheck(lambda t: AmpLoss.apply(t, A, h, False), (x,), eps=1e-6,
                                      atol=1e-5, rtol=1e-6, fast_mode=True)
        results.append(("2d 16x64^2 gradcheck", ok))
        # the same FD by hand at coordinates shared by all patterns: one pair
        # of loss calls gives B finite differences
        loss = AmpLoss.apply(x, A, h, False)
        (gx,) = torch.autograd.grad(loss.sum(), x)
        gx = gx.detach().cpu().numpy()
        worst = 0.0
        for _ in range(8):
            r, c = rng.integers(n), rng.integers(n)
            for unit in (1.0, 1j):
                xp, xm = x0.copy(), x0.copy()
                xp[:, r, c] += 1e-6 * unit
                xm[:, r, c] -= 1e-6 * unit
                lp = h.amp_loss(torch.from_numpy(xp).to(dev), A)
                lm = h.amp_loss(torch.from_numpy(xm).to(dev), A)
                fd = (lp - lm) / 2e-6
                an = gx[:, r, c].real if unit == 1.0 else gx[:, r, c].imag
                scale = np.abs(gx).reshape(B, -1).max(axis=1)
                worst = max(worst, float((np.abs(fd - an) / scale).max()))
        print(f"2d FD worst error relative to max|grad|: {worst:.3e}")
        results.append(("2d FD <= 1e-6", worst <= 1e-6))
    # 3D: one 32^3 FP64 volume
    n3 = 32
    zz, yy3, xx3 = np.mgrid[:n3, :n3, :n3]
    vol = (((zz - 16) / 6.0) ** 2 + ((yy3 - 16) / 9.0) ** 2 + ((xx3 - 16) / 5.0) ** 2 <= 1) * 0.7
    A3 = torch.from_numpy(np.abs(np.fft.fftshift(np.fft.fftn(np.fft.fftshift(vol))))[None]).to(dev)
    x3 = torch.from_numpy((vol + 0.05 * rng.standard_normal(vol.shape))[None] + 0j).to(dev)
    x3.requires_grad_(True)
    with hip.HipImpl(n3, n3, n3, 1, dtype="float64", use_torch_stream=True) as h3:
        ok3 = torch.autograd.gradcheck(lambda t: AmpLoss.apply(t, A3, h3, True), (x3,), eps=1e-6,
                                       atol=1e-5, rtol=1e-6, fast_mode=True)
        results.append(("3d 32^3 gradcheck", ok3))
    for name, ok in results:
        print(f"{name}: {'ok' if ok else 'FAILED'}")
    return 0 if all(ok for _, ok in results) else 1


if __name__ == "__main__":
    sys.exit(main(sys.argv[1]))
