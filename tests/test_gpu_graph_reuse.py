"""Engines keep their iteration graph on the handle across calls and re-point
it at each call's buffers and parameters (hipGraphExecUpdate).  Two calls on
one handle with different buffers and parameters must each give the bits of
an eager run on a fresh handle, and the second call must not instantiate a
new graph.  Every buffer stays alive until the end, so that a replay with
the first call's pointers would show as wrong bits rather than a fault."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def synth(pkg):
    from cpra_amd import synth

    return synth


def captures(h):
    return int(h.get_tuning()["graph_captures"])


def test_carousel_graph_reused_across_calls(hipmod, gpu, synth):
    n, P = 64, 24
    ds = [synth.make_carousel(n, P, config=80 + k) for k in range(2)]
    betas = (0.9, 0.75)
    keep, got = [], []
    with hipmod.HipImpl(n, n, 1, P) as h:
        # 15 outer iterations: three eager, one replay of the ten-iteration
        # body, two eager
        for k, (d, b) in enumerate(zip(ds, betas)):
            dev = {key: hipmod.to_device(d[key]) for key in ("x0", "amp", "support2d", "angles", "support3d")}
            vol = hipmod.zeros((n, n, n), np.complex64)
            keep.append((dev, vol))
            h.carousel3d_run(dev["x0"], dev["amp"], dev["support2d"], dev["angles"],
                             dev["support3d"], vol, outer_iters=15, inner_iters=1, beta=b, use_graph=1)
            h.Sync()
            got.append((dev["x0"].numpy(), vol.numpy()))
            if k == 0:
                c0 = captures(h)
        assert captures(h) == c0
        # the same buffers and parameters again (fresh start data): the kept
        # graph replays as it is, without a capture
        dev, vol = keep[1]
        dev["x0"].upload(ds[1]["x0"])
        h.carousel3d_run(dev["x0"], dev["amp"], dev["support2d"], dev["angles"],
                         dev["support3d"], vol, outer_iters=15, inner_iters=1, beta=betas[1], use_graph=1)
        h.Sync()
        assert np.array_equal(dev["x0"].numpy(), got[1][0])
        assert np.array_equal(vol.numpy(), got[1][1])
        assert captures(h) == c0
    for d, b, (x, v) in zip(ds, betas, got):
        dev = {key: hipmod.to_device(d[key]) for key in ("x0", "amp", "support2d", "angles", "support3d")}
        vol = hipmod.zeros((n, n, n), np.complex64)
        with hipmod.HipImpl(n, n, 1, P) as h:
            h.carousel3d_run(dev["x0"], dev["amp"], dev["support2d"], dev["angles"],
                             dev["support3d"], vol, outer_iters=15, inner_iters=1, beta=b, use_graph=0)
            h.Sync()
        assert np.array_equal(x, dev["x0"].numpy())
        assert np.array_equal(v, vol.numpy())


def _dm_inputs(synth, n, B, config):
    d = synth.make_2d_batch(1, n, n * 0.12, n * 0.18, 1e6, config=config)
    rng = np.random.default_rng(config)
    init = [(d["x0"][0] * (1 + 0.1 * rng.standard_normal((B, n, n)))).astype(np.complex64)
            for _ in range(3)]
    mask = np.tile((d["obj"] > 0).astype(np.uint8), (B, 1, 1))
    return init, d["amp"], mask


@pytest.mark.parametrize("textbook", [0, 1])
def test_dm_graph_reused_across_calls(hipmod, gpu, synth, textbook):
    n, B = 128, 2
    runs = [(_dm_inputs(synth, n, B, 90), 0.9, 20), (_dm_inputs(synth, n, B, 91), 0.8, 35)]
    keep, got = [], []
    with hipmod.HipImpl(n, n, 1, B) as h:
        for k, ((init, amp, mask), beta, iters) in enumerate(runs):
            bufs = [hipmod.to_device(a) for a in init] + [hipmod.to_device(amp), hipmod.to_device(mask)]
            keep.append(bufs)
            psi, t1, t2, a, m = bufs
            assert h.dm_run(psi, a, m, t1=t1, t2=t2, iters=iters, beta=beta, textbook=textbook,
                            space_mode=0 if not textbook else 1, use_graph=1)
            h.Sync()
            got.append([b.numpy() for b in (psi, t1, t2)])
            if k == 0:
                c0 = captures(h)
        assert captures(h) == c0
    for ((init, amp, mask), beta, iters), g in zip(runs, got):
        psi, t1, t2 = (hipmod.to_device(a) for a in init)
        with hipmod.HipImpl(n, n, 1, B) as h:
            assert h.dm_run(psi, hipmod.to_device(amp), hipmod.to_device(mask), t1=t1, t2=t2,
                            iters=iters, beta=beta, textbook=textbook,
                            space_mode=0 if not textbook else 1, use_graph=0)
            h.Sync()
        for a, b in zip(g, (psi, t1, t2)):
            assert np.array_equal(a, b.numpy())


def test_mixed_graph_reused_across_calls(hipmod, gpu, synth):
    def classes(seed):
        out = []
        for i, (n, count) in enumerate([(128, 2), (192, 1), (64, 3)]):
            d = synth.make_2d_batch(count, n, 0.12 * n, 0.18 * n, 1e6, config=400 + seed + i)
            out.append((d["x0"], d["amp"], (d["obj"] > 0).astype(np.uint8)))
        return out

    sets = [classes(0), classes(10)]
    keep, got = [], []
    with hipmod.HipImpl(64, 64, 1, 1) as h:
        for k, (cl, beta) in enumerate(zip(sets, (0.8, 0.7))):
            dev = [tuple(hipmod.to_device(a) for a in c) for c in cl]
            keep.append(dev)
            assert h.dm_mixed_run(dev, iters=6, beta=beta, space_mode=1, use_graph=1)
            h.Sync()
            got.append([c[0].numpy() for c in dev])
            if k == 0:
                c0 = captures(h)
        assert captures(h) == c0
    for cl, beta, g in zip(sets, (0.8, 0.7), got):
        dev = [tuple(hipmod.to_device(a) for a in c) for c in cl]
        with hipmod.HipImpl(64, 64, 1, 1) as h:
            assert h.dm_mixed_run(dev, iters=6, beta=beta, space_mode=1, use_graph=0)
            h.Sync()
        for a, c in zip(g, dev):
            assert np.array_equal(a, c[0].numpy())


def test_dm3d_graph_reused_across_calls(hipmod, gpu):
    n = 64
    rng = np.random.default_rng(12)
    runs = []
    for beta in (0.9, 0.8):
        psi0 = (rng.standard_normal((1, n, n, n)) + 1j * rng.standard_normal((1, n, n, n))).astype(np.complex64)
        A = rng.random((1, n, n, n)).astype(np.float32)
        mask = (rng.random((1, n, n, n)) < 0.3).astype(np.uint8)
        runs.append((psi0, A, mask, beta))
    keep, got = [], []
    with hipmod.HipImpl(n, n, n, 1) as h:
        for k, (psi0, A, mask, beta) in enumerate(runs):
            bufs = [hipmod.to_device(psi0) for _ in range(3)] + [hipmod.to_device(A), hipmod.to_device(mask)]
            keep.append(bufs)
            psi, t1, t2, a, m = bufs
            assert h.dm3d_run(psi, a, m, t1=t1, t2=t2, iters=5, beta=beta, textbook=0, space_mode=0,
                              use_graph=1)
            h.Sync()
            got.append([b.numpy() for b in (psi, t1, t2)])
            if k == 0:
                c0 = captures(h)
        assert captures(h) == c0
    for (psi0, A, mask, beta), g in zip(runs, got):
        psi, t1, t2 = (hipmod.to_device(psi0) for _ in range(3))
        with hipmod.HipImpl(n, n, n, 1) as h:
            assert h.dm3d_run(psi, hipmod.to_device(A), hipmod.to_device(mask), t1=t1, t2=t2, iters=5,
                              beta=beta, textbook=0, space_mode=0, use_graph=0)
            h.Sync()
        for a, b in zip(g, (psi, t1, t2)):
            assert np.array_equal(a, b.numpy())


def test_hio_chunk_graph_follows_the_row_plan(hipmod, gpu, synth):
    """The HIO chunk graph kept by a handle is keyed on the row kernels'
    radix plan as well: after a change of plan on the same buffers, the run
    must give the new plan's bits (those of an eager run on a fresh handle
    with that plan), not replay the old plan's graph."""
    n, B = 256, 2
    d = synth.make_2d_batch(B, n, 0.12 * n, 0.18 * n, 1e6, config=31)
    sched = dict(synth.HIO_SCHEDULE)
    sched.pop("iters")
    sched["sw_every"] = 5
    x = hipmod.to_device(d["x0"])
    amp = hipmod.to_device(d["amp"])
    sup = hipmod.to_device(d["support"])
    got = {}
    with hipmod.HipImpl(n, n, 1, B) as h:
        for pv in (0, 1, 0):
            h.set_tuning(f"row_plan_{n}={pv};")
            x.upload(d["x0"])
            sup.upload(d["support"])
            h.hio_run(x, amp, sup, iters=10, use_graph=1, **sched)
            h.Sync()
            got.setdefault(pv, []).append(x.numpy())
    assert np.array_equal(got[0][0], got[0][1])
    for pv in (0, 1):
        xe = hipmod.to_device(d["x0"])
        se = hipmod.to_device(d["support"])
        with hipmod.HipImpl(n, n, 1, B) as h:
            h.set_tuning(f"row_plan_{n}={pv};")
            h.hio_run(xe, amp, se, iters=10, use_graph=0, **sched)
            h.Sync()
        assert np.array_equal(got[pv][0], xe.numpy()), pv
    assert not np.array_equal(got[0][0], got[1][0])
