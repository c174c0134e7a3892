"""Mixed-size batch (cpra_hip_dm_mixed_run): the textbook difference map over
size classes of 64^2..384^2 patterns in one launch pair per iteration.

The fused, per-class and per-pattern launch modes run the same kernel bodies
and must agree bit for bit, and with the single-size engine
(cpra_hip_dm_run, textbook = 1) on each class; single steps are checked
against the FP64 oracle's textbook map (orc_dm_textbook)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def synth(pkg):
    from cpra_amd import synth

    return synth


def make_classes(synth, spec, seed):
    out = []
    for i, (n, count) in enumerate(spec):
        d = synth.make_2d_batch(count, n, 0.12 * n, 0.18 * n, 1e6, config=300 + seed + i)
        out.append({"n": n, "psi": d["x0"], "amp": d["amp"],
                    "support": (d["obj"] > 0).astype(np.uint8)})
    return out


def run_mixed(hipmod, classes, mode, iters, use_graph, space_mode=1):
    dev = [(hipmod.to_device(c["psi"]), hipmod.to_device(c["amp"]), hipmod.to_device(c["support"]))
           for c in classes]
    with hipmod.HipImpl(64, 64, 1, 1) as h:
        assert h.dm_mixed_run(dev, launch_mode=mode, iters=iters, beta=0.8, space_mode=space_mode,
                              use_graph=use_graph)
        h.Sync()
    return [d[0].numpy() for d in dev]


SPEC = [(128, 3), (192, 2), (256, 2), (384, 1), (64, 4)]


@pytest.mark.parametrize("iters,use_graph", [(1, 0), (4, 1), (5, 0)])
def test_launch_modes_agree_bitwise(hipmod, gpu, synth, iters, use_graph):
    classes = make_classes(synth, SPEC, 1)
    fused = run_mixed(hipmod, classes, hipmod.MIXED_FUSED, iters, use_graph)
    per_class = run_mixed(hipmod, classes, hipmod.MIXED_PER_CLASS, iters, use_graph)
    per_pattern = run_mixed(hipmod, classes, hipmod.MIXED_PER_PATTERN, iters, use_graph)
    for f, c, p in zip(fused, per_class, per_pattern):
        assert np.array_equal(f, c)
        assert np.array_equal(f, p)


def test_mixed_matches_single_size_engine(hipmod, gpu, synth):
    classes = make_classes(synth, SPEC, 2)
    fused = run_mixed(hipmod, classes, hipmod.MIXED_FUSED, 6, 1)
    for c, got in zip(classes, fused):
        n, count = c["n"], c["psi"].shape[0]
        psi = hipmod.to_device(c["psi"])
        with hipmod.HipImpl(n, n, 1, count) as h:
            assert h.dm_run(psi, hipmod.to_device(c["amp"]), hipmod.to_device(c["support"]),
                            iters=6, beta=0.8, textbook=1, space_mode=1, use_graph=1)
            h.Sync()
        assert np.array_equal(got, psi.numpy()), n


@pytest.mark.parametrize("space_mode", [1, 2])
def test_mixed_step_matches_oracle(hipmod, oracle, gpu, synth, space_mode):
    classes = make_classes(synth, [(128, 2), (192, 1), (256, 2), (384, 1)], 3)
    got = run_mixed(hipmod, classes, hipmod.MIXED_FUSED, 1, 0, space_mode)
    for c, g in zip(classes, got):
        ref = oracle.dm_textbook(c["psi"], c["amp"], 0.8, 1, mask=c["support"], space_mode=space_mode)
        assert relerr(g, ref) < 1e-5, c["n"]


def test_mixed_rejects_bad_input(hipmod, gpu):
    z = hipmod.zeros((1, 96, 96), np.complex64)
    a = hipmod.zeros((1, 96, 96), np.float32)
    s = hipmod.zeros((1, 96, 96), np.uint8)
    with hipmod.HipImpl(64, 64, 1, 1) as h:
        with pytest.raises(hipmod.CpraHipError):
            h.dm_mixed_run([(z, a, s)], iters=1)                        # 96 not a fast size
        z2 = hipmod.zeros((1, 64, 64), np.complex64)
        a2 = hipmod.zeros((1, 64, 64), np.float32)
        s2 = hipmod.zeros((1, 64, 64), np.uint8)
        with pytest.raises(hipmod.CpraHipError):
            h.dm_mixed_run([(z2, a2, s2)], iters=1, textbook=0)       # textbook map only
        with pytest.raises(hipmod.CpraHipError):
            h.dm_mixed_run([(z2, a2, s2)] * 9, iters=1)                 # more than 8 classes
        with pytest.raises(hipmod.CpraHipError):
            h.dm_mixed_run([(z2, a2, s2)], launch_mode=7, iters=1)
