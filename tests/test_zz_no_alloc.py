"""cpra_hip_dm_run_noalloc (the difference map with its iteration state in
the library's static scratch array, for callers that cannot allocate) against
cpra_hip_dm_run through the C ABI: identical bits on two calls in a row, the
second with a smaller batch (so the array still holds the first call's
contents), for the reference variant and the textbook map, and a clear
refusal of a batch beyond the array."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def starts(amp, B, seed):
    rng = np.random.default_rng(seed)
    n = amp.shape[-1]
    x = np.stack([np.fft.fftshift(np.fft.ifft2(np.fft.fftshift(
        amp[b % amp.shape[0]] * np.exp(2j * np.pi * rng.random((n, n)))))) for b in range(B)])
    return x.astype(np.complex64)


def run_pair(hipmod, n, B, iters, textbook, seed):
    from cpra_amd import synth

    data = synth.make_2d_batch(1, n, n * 0.12, n * 0.18, 1e6, config=700 + n)
    amp = hipmod.to_device(data["amp"])
    mask = hipmod.to_device(np.tile((data["obj"] > 0).astype(np.uint8), (B, 1, 1)))
    init = [starts(data["amp"], B, seed + k) for k in range(3)]
    kw = dict(iters=iters, beta=0.9 if not textbook else 0.8, textbook=textbook,
              space_mode=0 if not textbook else 1)
    out = []
    for noalloc in (False, True):
        psi, t1, t2 = (hipmod.to_device(a) for a in init)
        with hipmod.HipImpl(n, n, 1, B) as h:
            assert h.dm_run(psi, amp, mask, t1=t1, t2=t2, noalloc=noalloc, **kw)
            h.Sync()
        out.append([a.numpy() for a in (psi, t1, t2)])
    return out


@pytest.mark.parametrize("textbook", [0, 1])
def test_noalloc_matches_dm_run_twice_smaller_second(hipmod, gpu, textbook):
    # first call: 8 patterns of 256^2 (16 MiB of scratch), then 2 of 128^2
    for n, B, iters, seed in ((256, 8, 20, 10), (128, 2, 7, 20)):
        print(f"dm_run_noalloc textbook={textbook} n={n} B={B} iters={iters}", flush=True)
        ref, got = run_pair(hipmod, n, B, iters, textbook, seed)
        for a, b in zip(ref, got):
            assert np.array_equal(a, b)


def test_noalloc_refuses_a_batch_beyond_the_scratch(hipmod, gpu):
    n = 512
    B = (64 << 20) // (32 * n * n) + 1          # one pattern more than fits
    print(f"dm_run_noalloc refusal n={n} B={B}", flush=True)
    with hipmod.HipImpl(n, n, 1, B) as h:
        psi = hipmod.zeros((B, n, n), np.complex64)
        amp = hipmod.zeros((1, n, n), np.float32)
        mask = hipmod.zeros((B, n, n), np.uint8)
        with pytest.raises(hipmod.CpraHipError) as e:
            h.dm_run(psi, amp, mask, t1=psi, t2=psi, noalloc=True, iters=1)
        assert "scratch" in str(e.value)
