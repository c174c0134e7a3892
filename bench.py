"""Benchmark: batched shrinkwrap-HIO phase retrieval on one MI355X.

Workload (the headline configuration): 512 patterns of 256x256 complex64,
each from its own seeded disc object (radii U[30, 50] px, Poisson noise at
1e6 photons), HIO beta = 0.9 with real-positive support, shrinkwrap every 20
iterations (sigma 3.0 -> 1.5 px, threshold 0.1 x max), the iteration body
captured in a hipGraph.  One step = one HIO iteration of the whole batch
(512 pattern-iterations), shrinkwrap updates included; the K timed steps
are iterations 1..K of the schedule from the seeded start, replaying the
chunk graph that the warmup captured.  Inputs are resident in HBM before
the timed region.

Prints one JSON line (rank 0) with:
  value      batched HIO iterations per second over the K timed steps;
  roofline   the dominant fused kernel: algorithmic bytes per launch / its
             average duration (HIP events on the stream it runs on, from a
             separate eager run of the same workload), against 8 TB/s;
  cpu_baseline  the reference's CPU path (its MKL backend restated in FP32
             over libmkl_rt, one pass per primitive) on 32 of the patterns,
             the same schedule, on this host's cores and on 8 threads;
  cpu_baseline_fp64_oracle  the FP64 CPU oracle on 16 of the patterns;
  fp64_check two of the patterns against the FP64 oracle: one iteration
             from the same state (relative L2), and the R-factor after the
             full 200-iteration schedule next to the oracle's own spread
             under FP32-level input perturbations (HIO with positivity and
             shrinkwrap is dynamically unstable).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# ROCm's graph packet capture (on by default) replays kernel nodes with a
# per-node cost that eager launches do not have: the 180 x 128^2 carousel's
# graph ran 4 % below its eager launches with it, level with them without it,
# and the 512 x 256^2 HIO graph runs at the eager rate either way.  The
# runtime reads this once, when HIP initialises, so it is set before the
# library loads; a caller's own setting wins.
os.environ.setdefault("DEBUG_CLR_GRAPH_PACKET_CAPTURE", "0")

HBM_PEAK_GBS = 8000.0
METRIC = "phase-retrieval iters/s (512×256² batch, 128³ carousel); % HBM peak"

# algorithmic bytes per pixel of each fused kernel (complex64 = 8 B,
# amplitude float = 4 B, support u8 = 1 B)
COL_BYTES_PER_PX = 8 + 4 + 8        # read Y, read A, write Z
ROW_BYTES_PER_PX = 8 + 8 + 0.125 + 8  # read Z, read Y, read the support (one bit per pixel), write Y
HIO_ALGO_BYTES_PER_PX = 21          # x read + write, A, support (one iteration)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(kernel: str, tuning: dict):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of known traffic), taken
    with the same column variant, patterns per launch (the chunk within a sub-batch)
    and row mode as `tuning`
    (on the initial support of the workload, as the event-timed iterations)."""
    path = os.path.join(ROOT, "profiles", "hio_512x256sq_pmc.md")
    if not os.path.exists(path):
        return None
    # patterns per launch: the chunk within a sub-batch of the split
    try:
        ppl = min(int(tuning.get("chunk")), int(tuning.get("batch", 512)) // max(int(tuning.get("split", 1)), 1))
    except (TypeError, ValueError):
        return None
    want = f"col_256={tuning.get('col_256')};ppl={ppl};rows_sparse_used={tuning.get('rows_sparse_used')}"
    for line in open(path):
        parts = line.split()
        if (len(parts) == 4 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel
                and parts[2] == want):
            return int(parts[3])
    return None


def cpu_baseline_oracle(data, n_patterns=16):
    """The FP64 oracle on the same 200-iteration HIO + shrinkwrap schedule,
    on `n_patterns` of the patterns, scaled to 512-pattern batched
    iterations."""
    import numpy as np

    import __graft_entry__ as ge

    o = ge.load_oracle()
    from cpra_amd import synth

    sched = dict(synth.HIO_SCHEDULE)
    iters = sched.pop("iters")
    x = data["x0"][:n_patterns].astype(np.complex128)
    A = data["amp"][:n_patterns].astype(np.float64)
    s = data["support"][:n_patterns]
    t0 = time.perf_counter()
    o.hio(x, A, s, iters, **sched)
    dt = time.perf_counter() - t0
    pit_per_s = n_patterns * iters / dt
    return {
        "value": pit_per_s / 512.0,
        "unit": "iters/s",
        "cores": min(o.num_threads(), n_patterns),
        "kind": "port",
        "sample": f"FP64 oracle, the 200-iteration HIO + shrinkwrap schedule on {n_patterns} of the "
                  f"512 patterns ({dt:.1f} s), scaled to 512-pattern batched iterations",
        "pattern_iters_per_s": pit_per_s,
    }


_MKL_CHILD = r"""
import json, os, sys
sys.path.insert(0, sys.argv[1])
sys.path.insert(0, os.path.join(sys.argv[1], "oracle"))
import __graft_entry__ as ge
ge.load_package()
from cpra_amd import synth
import mkl_path as mp
n_patterns, threads, kind, min_s = int(sys.argv[2]), int(sys.argv[3]), sys.argv[4], float(sys.argv[5])
if not mp.available():
    print(json.dumps({"error": mp.load_error()}))
    sys.exit(0)
if kind == "8starts":
    data = synth.multistart_256()
else:
    data = synth.make_2d_batch(n_patterns, 256, 30.0, 50.0, 1e6, starts_per_pattern=1, config=1)
sched = dict(synth.HIO_SCHEDULE)
med, rates, runs = mp.time_hio_schedule(data["x0"], data["amp"], data["support"], threads, sched,
                                        min_seconds=min_s)
print(json.dumps({"pattern_iters_per_s": med, "reps": rates, "schedules_per_rep": runs,
                  "library": mp.library_path(), "cpu_model": mp.cpu_model(),
                  "physical_cores": mp.physical_cores()}))
"""


def _mkl_path_runs(n_patterns, kind, min_seconds):
    """{threads: record} of the MKL-path child on all cores and on 8 threads;
    None when libmkl_rt cannot be loaded, {"error": ...} when a child fails."""
    import subprocess

    try:
        import mkl_path as mp  # noqa: F401
    except ImportError:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import mkl_path as mp
    cores = mp.physical_cores()
    env_threads = os.environ.get("OMP_NUM_THREADS", "")
    all_threads = min(cores, int(env_threads)) if env_threads.isdigit() else cores
    out = {}
    for threads in sorted({all_threads, 8}, reverse=True):
        r = subprocess.run([sys.executable, "-c", _MKL_CHILD, ROOT, str(n_patterns), str(threads), kind,
                            str(min_seconds)],
                           capture_output=True, text=True, timeout=900,
                           env={**os.environ, "HIP_VISIBLE_DEVICES": "", "CUDA_VISIBLE_DEVICES": "",
                                "OMP_NUM_THREADS": str(threads)})
        if r.returncode != 0:
            return {"error": r.stderr[-400:]}
        rec = json.loads(r.stdout.strip().splitlines()[-1])
        if "error" in rec:
            return None
        out[threads] = rec
    return out


def cpu_baseline_mkl_path(n_patterns=32):
    """The reference's CPU path: its MKL backend restated in FP32 over
    libmkl_rt (oracle/mkl_path.c: MKL DFTI with the reference's descriptor,
    its shift and constraint loops under OpenMP), running the same
    200-iteration HIO + shrinkwrap schedule one pass per primitive, on the
    first `n_patterns` of the 512 patterns (the same seeded data).  Protocol:
    2 untimed warm-up iterations, then 3 repetitions of at least 10 s and 20
    iterations each, median; on all of this host's cores available to the
    run (OMP_NUM_THREADS when set) and on 8 threads, the reference
    benchmark's setting.  Runs in a child process that never touches the
    GPU.  Returns None when libmkl_rt cannot be loaded."""
    out = _mkl_path_runs(n_patterns, "batch", 10.0)
    if out is None or "error" in out:
        return out
    # the faster of the two thread counts is the baseline (on a shared host
    # the run on every core can be the slower one)
    best = max(out, key=lambda t: out[t]["pattern_iters_per_s"])
    top = out[best]
    res = {
        "value": top["pattern_iters_per_s"] / 512.0,
        "unit": "iters/s",
        "cores": best,
        "kind": "port",
        "sample": f"the reference's MKL backend in FP32 over {top['library']} (DFTI + OpenMP, one "
                  f"pass per primitive), the 200-iteration HIO + shrinkwrap schedule on {n_patterns} "
                  f"of the 512 patterns, median of 3 runs of >= 10 s after 2 warm-up iterations, "
                  f"{best} threads (the faster of {sorted(out)}); scaled to 512-pattern batched "
                  "iterations",
        "cpu_model": top["cpu_model"],
        "physical_cores_visible": top["physical_cores"],
        "pattern_iters_per_s": top["pattern_iters_per_s"],
        "reps_pattern_iters_per_s": top["reps"],
    }
    res["by_threads"] = {str(t): {"value": r["pattern_iters_per_s"] / 512.0, "unit": "iters/s",
                                  "pattern_iters_per_s": r["pattern_iters_per_s"],
                                  "reps_pattern_iters_per_s": r["reps"]} for t, r in out.items()}
    return res


_MKL_SNIPPET = r"""
import json, sys, time
import numpy as np
import torch
sys.path.insert(0, sys.argv[1])
import __graft_entry__ as ge
ge.load_package()
from cpra_amd import synth
n_patterns, iters = int(sys.argv[2]), int(sys.argv[3])
data = synth.make_2d_batch(n_patterns, 256, 30.0, 50.0, 1e6, starts_per_pattern=1, config=1)
x = torch.from_numpy(data["x0"].copy())
A = torch.from_numpy(data["amp"].copy())
s = torch.from_numpy(data["support"].astype(bool))
"""


def cpu_baseline_torch(n_patterns=16, iters=4):
    """Fallback where libmkl_rt cannot be loaded: HIO on the host with torch's
    CPU FFT (MKL-backed), FP32, without shrinkwrap.  Runs in a child process
    so that torch's bundled HIP runtime never shares a process with the
    benchmark's."""
    import subprocess

    code = _MKL_SNIPPET + _MKL_BODY
    r = subprocess.run([sys.executable, "-c", code, ROOT, str(n_patterns), str(iters)],
                       capture_output=True, text=True, timeout=600,
                       env={**os.environ, "HIP_VISIBLE_DEVICES": "", "CUDA_VISIBLE_DEVICES": ""})
    if r.returncode != 0:
        return {"error": r.stderr[-400:]}
    return json.loads(r.stdout.strip().splitlines()[-1])


_MKL_BODY = r"""

def step(x):
    X = torch.fft.fftshift(torch.fft.fft2(torch.fft.fftshift(x, dim=(-2, -1))), dim=(-2, -1))
    m = X.abs()
    Xp = torch.where((A > 0) & (m > 0), X * (A / torch.where(m > 0, m, 1)), X)
    xp = torch.fft.fftshift(torch.fft.ifft2(torch.fft.fftshift(Xp, dim=(-2, -1))),
                            dim=(-2, -1))
    ok = s & (xp.real >= 0)
    return torch.where(ok, torch.complex(xp.real, torch.zeros_like(xp.real)), x - 0.9 * xp)

step(x)
t0 = time.perf_counter()
for _ in range(iters):
    x = step(x)
dt = time.perf_counter() - t0
print(json.dumps({
    "value": n_patterns * iters / dt / 512.0,
    "unit": "iters/s",
    "cores": torch.get_num_threads(),
    "kind": "port",
    "sample": f"fallback: torch's CPU FFT (MKL-backed), FP32, {n_patterns} patterns x {iters} HIO "
              "iterations without shrinkwrap",
}))
"""


def kernel_entry(ms, cnt, iters, bytes_per_iter, launches_per_iter=1):
    """One per-kernel roofline record from event-timed eager launches."""
    per_iter = ms / max(iters, 1)
    ach = bytes_per_iter / (per_iter * 1e-3) / 1e9 if per_iter > 0 else 0.0
    return {"avg_ms": ms / max(cnt, 1), "launches": cnt, "ms_per_iteration": per_iter,
            "bytes_per_launch": int(bytes_per_iter // max(launches_per_iter, 1)),
            "achieved_GBs": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4)}


def bench_carousel(hip, synth, ge, n=128, P=180, outer=50, warmup=15, timing_iters=10, fp64=True,
                   tuning=""):
    """3D carousel: 180 projections of 128^2 (12 random ellipsoids in 128^3,
    Poisson noise at 1e7 photons), 50 outer iterations of [1 HIO iteration on
    every projection, merge, 3D constraint, re-projection]; graph-captured
    outer body.  `tuning`: a fixed tuning string for the run (profiler A/B
    runs).  Returns a result record (value = outer iterations / s)."""
    import numpy as np

    t0 = time.perf_counter()
    d = synth.make_carousel(n, P, config=2)
    log(f"bench[carousel]: data ready in {time.perf_counter() - t0:.1f} s")
    dev = {k: hip.to_device(d[k]) for k in ("x0", "amp", "support2d", "angles", "support3d")}
    vol = hip.zeros((n, n, n), np.complex64)
    h = hip.HipImpl(n, n, 1, P)
    if tuning:
        h.set_tuning(tuning)
    run = lambda k, g=1: h.carousel3d_run(dev["x0"], dev["amp"], dev["support2d"], dev["angles"],
                                          dev["support3d"], vol, outer_iters=k, inner_iters=1,
                                          beta=0.9, use_graph=g)
    run(warmup)
    run(warmup, 0)
    h.Sync()

    def timed(g):
        # each run from the same start (uploaded outside the timed region)
        dev["x0"].upload(d["x0"])
        h.Sync()
        t0 = time.perf_counter()
        run(outer, g)
        h.Sync()
        return time.perf_counter() - t0

    # graph replays (the value) and eager launches (beside it), alternating,
    # three runs each; the median of each
    runs = {1: [], 0: []}
    for _ in range(3):
        for g in (0, 1):
            runs[g].append(timed(g))
    el = sorted(runs[1])[1]
    its = outer / el
    its_eager = outer / sorted(runs[0])[1]
    log(f"bench[carousel]: {outer} outer iterations in {el * 1e3:.2f} ms -> {its:.1f} it/s "
        f"(graph runs {[round(outer / t, 1) for t in runs[1]]})")
    log(f"bench[carousel]: eager launches -> {its_eager:.1f} it/s "
        f"(runs {[round(outer / t, 1) for t in runs[0]]})")
    h.set_kernel_timing(True)
    run(timing_iters, 0)
    h.Sync()
    kt = h.kernel_times()
    h.set_kernel_timing(False)
    px2, vx = P * n * n, n ** 3
    K = timing_iters
    planes = n <= 128 and not (tuning and "plane=0" in tuning)
    spec = {  # kind -> algorithmic bytes per outer iteration, launches per iteration
        "column": (px2 * COL_BYTES_PER_PX, 1),
        "row": (px2 * ROW_BYTES_PER_PX, 1),
        # the row-spectral projections into the merge's layout [P][v][u]
        "proj_transpose": (px2 * 16, 1),
        "merge": (px2 * 8 + vx * 8, 1),                # gather r (slices [P][v][u]), volume w
    }
    if planes:
        # per z-plane: IDFT_xy, the constraint (support as bits), DFT_xy and
        # the plane's rows of every slice: read the volume, write the slices
        spec["volume"] = (vx * 8 + vx / 8 + px2 * 8, 1)
    else:
        # x and y passes around the constraint kernel (support as bytes),
        # then the re-projection
        spec["volume"] = (4 * vx * 16 + vx * 17, 5)
        spec["extract"] = (px2 * 8 + vx * 8, 1)
    # (the transpose runs only where the row pass cannot store the merge's
    # layout itself)
    kern = {k: kernel_entry(kt[k][0], kt[k][1], timing_iters, b, l) for k, (b, l) in spec.items()
            if kt[k][1] > 0}
    dom = max(kern, key=lambda k: kern[k]["ms_per_iteration"])
    # one HIO iteration, the merge (slices r, volume w) and the re-projection
    # (volume r, slices w), the constraint's support as bits
    algo = px2 * HIO_ALGO_BYTES_PER_PX + (px2 * 8 + vx * 8) * 2 + vx // 8
    rec = {
        "metric": "carousel outer iters/s (180 x 128^2 projections, 128^3 volume)",
        "value": round(its, 2), "unit": "iters/s", "steps": outer, "warmup": warmup,
        "ms_per_step": round(1e3 * el / outer, 4),
        "config": {"workload": f"3D carousel, {P} projections of {n}x{n} complex64, {n}^3 volume, "
                               "1 HIO iteration (beta 0.9) + merge + 3D constraint + "
                               "re-projection per outer iteration", "graph": True},
        "value_eager_launches": round(its_eager, 2),
        "algorithmic_bytes_per_iteration": algo,
        "hbm_frac_algorithmic": round(algo * its / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": kern[dom]["achieved_GBs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": kern[dom]["frac"],
                     "kernels": kern},
    }
    if fp64:
        o = ge.load_oracle()
        x1 = hip.to_device(d["x0"])
        with hip.HipImpl(n, n, 1, P) as h1:
            h1.carousel3d_run(x1, dev["amp"], dev["support2d"], dev["angles"], dev["support3d"], vol,
                              outer_iters=1, use_graph=0)
            h1.Sync()
        xr, vr = o.carousel3d(d["x0"], d["amp"].astype(np.float64), d["support2d"],
                              d["angles"].astype(np.float64), d["support3d"], 1)
        rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
        rec["fp64_check"] = {"outer_iters": 1, "projections_rel_l2": rel(x1.numpy(), xr),
                             "volume_rel_l2": rel(vol.numpy(), vr)}
    h.close()
    return rec


def bench_gd2d(hip, synth, ge, P=256, n=256, steps=500, warmup=20, timing_iters=50, fp64=True):
    """Amplitude-loss gradient descent: 256 patterns of 256^2 (objects as the
    batched HIO workload), 500 steps of x <- x - mu grad with mu = 0.5/max(I),
    one fused pass per step."""
    import numpy as np

    t0 = time.perf_counter()
    d = synth.make_2d_batch(P, n, 30.0 * n / 256, 50.0 * n / 256, 1e6, starts_per_pattern=1, config=4)
    log(f"bench[gd2d]: data ready in {time.perf_counter() - t0:.1f} s")
    A32 = d["amp"]
    mu = 0.5 / float((A32.astype(np.float64) ** 2).max())
    # start: a random real guess inside the autocorrelation support (a start
    # with the measured magnitudes, like the HIO one, is already a minimiser)
    rng = np.random.default_rng(4000)
    x0 = (d["support"] * rng.random(d["support"].shape)).astype(np.complex64)
    x = hip.to_device(x0)
    amp = hip.to_device(A32)
    h = hip.HipImpl(n, n, 1, P)
    L0 = h.amp_loss(x, amp)
    h.amp_gd_run(x, amp, warmup, mu, want_loss=False)
    h.Sync()
    t0 = time.perf_counter()
    h.amp_gd_run(x, amp, steps, mu, want_loss=False)
    h.Sync()
    el = time.perf_counter() - t0
    its = steps / el
    log(f"bench[gd2d]: {steps} steps in {el * 1e3:.1f} ms -> {its:.1f} steps/s")
    h.set_kernel_timing(True)
    h.amp_gd_run(x, amp, timing_iters, mu, want_loss=False)
    h.Sync()
    ms, cnt = h.kernel_times()["gd"]
    h.set_kernel_timing(False)
    px = P * n * n
    k = kernel_entry(ms, cnt, timing_iters, px * 20)
    rec = {
        "metric": "amplitude-loss GD steps/s (256 x 256^2)",
        "value": round(its, 2), "unit": "steps/s", "steps": steps, "warmup": warmup,
        "ms_per_step": round(1e3 * el / steps, 4),
        "config": {"workload": f"gradient descent on sum(|F x| - A)^2, {P} x {n}x{n} complex64, "
                               f"mu = 0.5/max(I)", "patterns": P, "size": n},
        "hbm_frac_algorithmic_20Bpx": round(px * 20 * its / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": {"bound": "hbm", "kernel": f"k_gd_cols<{n}>", "achieved": k["achieved_GBs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": k["frac"], "kernels": {"gd": k}},
    }
    if fp64:
        o = ge.load_oracle()
        kk, ns = 2, 100
        xs = hip.to_device(x0[:kk])
        with hip.HipImpl(n, n, 1, kk) as hs:
            L = hs.amp_gd_run(xs, hip.to_device(A32[:kk]), ns, mu)
        xr = x0[:kk].astype(np.complex128)
        Ar = A32[:kk].astype(np.float64)
        for _ in range(ns):
            xr = xr - np.float32(mu) * o.amp_grad(xr, Ar)
        Lr = o.amp_loss(xr, Ar)
        rec["fp64_check"] = {"patterns": kk, "steps": ns,
                             "loss_start": [float(v) for v in L0[:kk]],
                             "loss_gpu": [float(v) for v in L], "loss_fp64": [float(v) for v in Lr],
                             "x_rel_l2": float(np.linalg.norm(xs.numpy() - xr) / np.linalg.norm(xr)),
                             "loss_rel_err": float(np.max(np.abs(L - Lr) / Lr))}
    h.close()
    return rec


def bench_gd3d(hip, synth, ge, n=256, steps=500, warmup=10, timing_iters=20, fp64=True):
    """Amplitude-loss gradient descent on one 256^3 volume (12 random
    ellipsoids), 500 steps, mu = 0.5/max(I); the gradient is recomputed from x
    at every step (no stored transforms)."""
    import numpy as np

    t0 = time.perf_counter()
    obj, A32 = synth.ellipsoid_amplitude_volume(n, config=4)
    # start: a random real guess inside the ball of radius 64 that holds the
    # object
    rng = np.random.default_rng(4004)
    k = np.arange(n) - n / 2
    ball = (k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2) <= 64.0 ** 2
    x0 = (ball * rng.random((n, n, n)) * float(obj[obj > 0].mean())).astype(np.complex64)
    del ball
    log(f"bench[gd3d]: data ready in {time.perf_counter() - t0:.1f} s")
    mu = 0.5 / float((A32.astype(np.float64) ** 2).max())
    x = hip.to_device(x0[None])
    amp = hip.to_device(A32[None])
    h = hip.HipImpl(n, n, n, 1)
    L0 = h.amp_loss(x, amp, volume=True)
    h.amp_gd_run(x, amp, warmup, mu, volume=True, want_loss=False)
    h.Sync()
    t0 = time.perf_counter()
    h.amp_gd_run(x, amp, steps, mu, volume=True, want_loss=False)
    h.Sync()
    el = time.perf_counter() - t0
    its = steps / el
    log(f"bench[gd3d]: {steps} steps in {el * 1e3:.1f} ms -> {its:.1f} steps/s")
    h.set_kernel_timing(True)
    h.amp_gd_run(x, amp, timing_iters, mu, volume=True, want_loss=False)
    h.Sync()
    ms, cnt = h.kernel_times()["gd"]
    h.set_kernel_timing(False)
    vx = n ** 3
    k = kernel_entry(ms, cnt, timing_iters, vx * 20)
    rec = {
        "metric": "amplitude-loss GD steps/s (256^3 volume)",
        "value": round(its, 2), "unit": "steps/s", "steps": steps, "warmup": warmup,
        "ms_per_step": round(1e3 * el / steps, 4),
        "config": {"workload": f"gradient descent on sum(|F x| - A)^2, one {n}^3 complex64 volume, "
                               "mu = 0.5/max(I), gradient recomputed every step", "size": n},
        "hbm_frac_algorithmic_20Bpx": round(vx * 20 * its / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": {"bound": "hbm", "kernel": f"k_gd_cols<{n}> (z-lines)", "achieved": k["achieved_GBs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": k["frac"], "kernels": {"gd": k}},
    }
    if fp64:
        o = ge.load_oracle()
        ns = 2
        xs = hip.to_device(x0[None])
        L = h.amp_gd_run(xs, amp, ns, mu, volume=True)
        xr = x0.astype(np.complex128)
        Ar = A32.astype(np.float64)
        for _ in range(ns):
            xr = xr - np.float32(mu) * o.amp_grad_3d(xr, Ar)
        Lr = o.amp_loss_3d(xr, Ar)
        rec["fp64_check"] = {"steps": ns, "loss_start": float(L0[0]), "loss_gpu": float(L[0]),
                             "loss_fp64": float(Lr),
                             "x_rel_l2": float(np.linalg.norm(xs.numpy()[0] - xr) / np.linalg.norm(xr)),
                             "loss_rel_err": float(abs(L[0] - Lr) / Lr)}
    h.close()
    return rec


def bench_conv3d(hip, synth, ge, n=512, iters=20, warmup=3, timing_iters=3, prim_iters=3,
                 fp64=True, check_n=128, tuning=""):
    """The reference's conventional 3D reconstruction at its benchmark scale
    (src/CONV_MKL.cpp:47-63 at M = N = L = 512, src/benchmark_mkl.cpp:7-13):
    difference-map iterations (t1, t2 carried) on one 512^3 complex64 volume,
    through the fused engine (cpra_hip_dm3d_run, graph-replayed) and, for
    comparison, through the reference's own sequence of primitive calls.
    Inputs are uniform random, as in the reference benchmark (Initialize); the
    support is the reference SpaceConstraint mask of a uniform float
    constraint (S >= 0.9)."""
    import numpy as np

    vox = n ** 3
    c64, f32, u8 = np.complex64, np.float32, np.uint8
    psi, t1, t2 = (hip.DeviceArray((1, n, n, n), c64) for _ in range(3))
    A = hip.DeviceArray((1, n, n, n), f32)
    S = hip.DeviceArray((1, n, n, n), f32)
    mask = hip.DeviceArray((1, n, n, n), u8)
    h = hip.HipImpl(n, n, n, 1)
    if tuning:
        h.set_tuning(tuning)
    for i, (arr, num) in enumerate(((psi, 2 * vox), (A, vox), (S, vox))):
        h.Initialize(arr, num, seed=5000 + i)
    h.support_from_constraint(S, mask, vox)
    h.Memcpy(t1, psi, 8 * vox)
    h.Memcpy(t2, psi, 8 * vox)
    kw = dict(beta=0.9, textbook=0, space_mode=0, dc_skip_zero=1)
    h.dm3d_run(psi, A, mask, t1=t1, t2=t2, iters=warmup, use_graph=1, **kw)
    h.Sync()
    t0 = time.perf_counter()
    h.dm3d_run(psi, A, mask, t1=t1, t2=t2, iters=iters, use_graph=1, **kw)
    h.Sync()
    el = time.perf_counter() - t0
    its = iters / el
    log(f"bench[conv3d]: {iters} iterations at {n}^3 in {el * 1e3:.1f} ms -> {its:.1f} it/s")
    h.set_kernel_timing(True)
    h.dm3d_run(psi, A, mask, t1=t1, t2=t2, iters=timing_iters, use_graph=0, **kw)
    h.Sync()
    kt = h.kernel_times()
    h.set_kernel_timing(False)
    # algorithmic bytes per launch kind and iteration: z lines over [Y1; Y2]
    # (read 8 + amplitude 4 + write 8, twice), two y-line passes over
    # [Y1; Y2] (read + write 16, twice), x rows (read Z1, Z2, psi, support;
    # write psi, Y1, Y2: 49)
    spec = {"column": (vox * 40, 1), "pencil": (vox * 64, 2), "row": (vox * 49, 1)}
    kern = {k: kernel_entry(kt[k][0], kt[k][1], timing_iters, b, l) for k, (b, l) in spec.items()}
    dom = max(kern, key=lambda k: kern[k]["ms_per_iteration"])
    # the reference's call sequence (12 primitive calls per iteration)
    t0 = time.perf_counter()
    for _ in range(prim_iters):
        h.Forward3D(t1)
        h.DataConstraint(t1, A, vox, 1)
        h.Backward3D(t1)
        h.MergeAddData(psi, t1, -1 / 0.9, 1 + 1 / 0.9, vox)
        h.SpaceConstraint(t1, S, vox, 1)
        h.SpaceConstraint(t2, S, vox, 1)
        h.MergeAddData(psi, t2, 1 / 0.9, 1 - 1 / 0.9, vox)
        h.Forward3D(t2)
        h.DataConstraint(t2, A, vox, 1)
        h.Backward3D(t2)
        h.MergeAddData(t1, psi, 0.9, 1.0, vox)
        h.MergeAddData(t2, psi, -0.9, 1.0, vox)
    h.Sync()
    prim_its = prim_iters / (time.perf_counter() - t0)
    log(f"bench[conv3d]: primitive sequence {prim_its:.1f} it/s")
    for a in (psi, t1, t2, A, S, mask):
        a.free()
    h.close()
    rec = {
        "metric": f"conventional 3D difference-map iters/s ({n}^3 complex64)",
        "value": round(its, 2), "unit": "iters/s", "steps": iters, "warmup": warmup,
        "ms_per_step": round(1e3 * el / iters, 4),
        "config": {"workload": f"difference map (reference variant, beta 0.9) on one {n}^3 complex64 "
                               "volume, uniform random inputs as in the reference benchmark",
                   "size": n, "graph": True},
        "primitive_sequence_iters_per_s": round(prim_its, 2),
        "speedup_vs_primitive_sequence": round(its / prim_its, 2),
        "hbm_frac_algorithmic_53Bpx": round(vox * 53 * its / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": kern[dom]["achieved_GBs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": kern[dom]["frac"],
                     "kernels": kern},
    }
    if fp64:
        # one iteration from the same random state at check_n^3 against the
        # FP64 oracle (orc_dm_ref3d)
        o = ge.load_oracle()
        m = check_n
        rng = np.random.default_rng(6)
        shp = (1, m, m, m)
        p0 = (rng.random(shp) + 1j * rng.random(shp)).astype(c64)
        q0 = (rng.random(shp) + 1j * rng.random(shp)).astype(c64)
        A0 = rng.random(shp).astype(f32)
        S0 = rng.random(shp).astype(f32)
        hs = hip.HipImpl(m, m, m, 1)
        dp, d1, d2 = hip.to_device(p0), hip.to_device(p0), hip.to_device(q0)
        mk = hip.zeros(shp, u8)
        hs.support_from_constraint(hip.to_device(S0), mk, m ** 3)
        hs.dm3d_run(dp, hip.to_device(A0), mk, t1=d1, t2=d2, iters=1, use_graph=0, **kw)
        hs.Sync()
        rp, r1, r2 = o.dm_ref3d(p0, p0, q0, A0, S0, 0.9, 1)
        rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))  # noqa: E731
        rec["fp64_check"] = {"size": m, "iters": 1, "psi_rel_l2": rel(dp.numpy(), rp),
                             "t1_rel_l2": rel(d1.numpy(), r1), "t2_rel_l2": rel(d2.numpy(), r2)}
        hs.close()
        # CPU baseline: the oracle's 3D map at check_n^3 for about 10 s,
        # scaled by the voxel count to n^3
        k, t0 = 0, time.perf_counter()
        p, q1, q2 = p0, p0, q0
        while time.perf_counter() - t0 < 10.0:
            p, q1, q2 = o.dm_ref3d(p, q1, q2, A0, S0, 0.9, 1)
            k += 1
        dt = time.perf_counter() - t0
        rec["cpu_baseline"] = {
            "value": k / dt * (m ** 3) / vox, "unit": "iters/s", "cores": o.num_threads(),
            "kind": "port",
            "sample": f"FP64 oracle 3D difference map, {k} iterations at {m}^3 in {dt:.1f} s, scaled "
                      f"by the voxel ratio to {n}^3 (FFT cost grows as V log V, so this "
                      "overstates the CPU rate by about a quarter)"}
    return rec


def bench_hio_8starts(hip, synth, ge, n=256, reps=5, fp64=True):
    """The CPU-runnable configuration: 8 random starts of one 256^2 pattern
    (disc radius 40 +- 5 px, 1e6 photons), the 200-iteration HIO + shrinkwrap
    schedule (beta 0.9, every 20, sigma 3.0 -> 1.5, threshold 0.1).  GPU: the
    whole schedule from the seeded start with graph replays, median of `reps`
    runs after one untimed run; beside it, the reference's MKL backend
    restated in FP32 over libmkl_rt (oracle/mkl_path.c) on the same 8 starts
    and schedule: 2 warm-up iterations, 3 repetitions of at least 2 s (whole
    schedules), median, on all cores and on 8 threads.  Both count batched
    iterations (one iteration of all 8 starts).  The GPU run's parity at this
    size is a test matter (HIO steps and schedule against the FP64 oracle);
    `fp64` is accepted for the common signature."""
    import statistics

    import numpy as np

    d = synth.multistart_256(n)
    B = d["x0"].shape[0]
    sched = dict(synth.HIO_SCHEDULE)
    iters = sched.pop("iters")
    x = hip.to_device(d["x0"])
    amp = hip.to_device(d["amp"])
    sup = hip.to_device(d["support"])
    times = []
    with hip.HipImpl(n, n, 1, B) as h:
        for r in range(reps + 1):
            x.upload(d["x0"])
            sup.upload(d["support"])
            h.Sync()
            t0 = time.perf_counter()
            h.hio_run(x, amp, sup, iters=iters, use_graph=1, **sched)
            h.Sync()
            if r:
                times.append(time.perf_counter() - t0)
        rf = h.rfactor(x, amp)
    el = statistics.median(times)
    its = iters / el
    log(f"bench[8starts]: {iters} iterations of {B} starts in {el * 1e3:.2f} ms -> {its:.1f} it/s")
    rec = {
        "metric": "batched HIO iters/s (8 starts of one 256^2 pattern, 200-iteration shrinkwrap schedule)",
        "value": round(its, 2), "unit": "iters/s", "steps": iters, "warmup": iters,
        "ms_per_step": round(1e3 * el / iters, 4),
        "config": {"workload": f"2D shrinkwrap-HIO, {B} random starts of one {n}x{n} pattern, complex64, "
                               "beta 0.9, shrinkwrap every 20 (sigma 3.0->1.5, thr 0.1)",
                   "graph": True, "runs": [round(iters / t, 1) for t in times]},
        "rfactor_final": [round(float(v), 6) for v in rf] if rf is not None else None,
    }
    out = _mkl_path_runs(B, "8starts", 2.0)
    if out is not None and "error" not in out:
        best = max(out, key=lambda t: out[t]["pattern_iters_per_s"])
        top = out[best]
        rec["cpu_baseline"] = {
            "value": top["pattern_iters_per_s"] / B, "unit": "iters/s", "cores": best, "kind": "port",
            "sample": f"the reference's MKL backend in FP32 over {top['library']} (DFTI + OpenMP, one "
                      f"pass per primitive), the same {B} starts and 200-iteration schedule, median of 3 "
                      f"runs of >= 2 s after 2 warm-up iterations, {best} threads (the faster of "
                      f"{sorted(out)})",
            "cpu_model": top["cpu_model"],
            "by_threads": {str(t): round(r["pattern_iters_per_s"] / B, 2) for t, r in out.items()},
        }
        rec["gpu_over_cpu"] = round(its / rec["cpu_baseline"]["value"], 1)
    else:
        rec["cpu_baseline"] = {"error": out}
    return rec


def bench_ref_benchmark(hip, synth, ge, epi=10, iters=10, fp64=True):
    """The reference's own benchmark program (src/benchmark_mkl.cpp) restated
    over HipImpl (bin/cpra_benchmark), run as its README documents it
    ("benchmark 10 10": 10 episodes of 10 iterations at M = N = L = 512,
    P = 288), with the fused engine and with the reference's primitive
    sequence, plus the conventional mode with 10 iterations.  The variant
    that records Step B's sweep as a graph and replays it (--graph) is timed
    beside them.  Value: the program's own headline, the estimated time for
    100 initial guesses."""
    import re
    import subprocess

    exe = os.path.join(ROOT, "carousel-phase-retrieval-algorithm_amd", "bin", "cpra_benchmark")

    def run(*args):
        r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=900)
        if r.returncode != 0:
            raise RuntimeError(f"cpra_benchmark {' '.join(args)} failed: {r.stderr[-400:]}")
        t = [float(x) for x in re.findall(r"^elapsed time: ([0-9.eE+-]+)s", r.stdout, re.M)]
        c = re.search(r"Estimated step C elapsed time: ([0-9.eE+-]+)s", r.stdout)
        tot = re.search(r"in total: ([0-9.eE+-]+)s", r.stdout)
        return {"step_A_s": t[0], "step_B_s": t[1], "step_D_s": t[2],
                "step_C_estimate_s": float(c.group(1)), "total_estimate_s": float(tot.group(1))}

    fused = run(str(epi), str(iters))
    log(f"bench[ref_benchmark]: fused {fused}")
    prim = run("--primitives", str(epi), str(iters))
    log(f"bench[ref_benchmark]: primitives {prim}")
    graphed = run("--graph", str(epi), str(iters))
    log(f"bench[ref_benchmark]: graph-replayed sweep {graphed}")
    r = subprocess.run([exe, str(iters)], capture_output=True, text=True, timeout=900)
    conv = re.search(r"^elapsed time: ([0-9.eE+-]+)s", r.stdout, re.M)
    return {
        "metric": "reference benchmark: estimated time for 100 initial guesses (s)",
        "value": fused["total_estimate_s"], "unit": "s", "higher_is_better": False,
        "config": {"workload": f"bin/cpra_benchmark {epi} {iters}: M = N = L = 512, P = 288, "
                               "1000 pre-iterations, one 512^3 3D iteration (src/benchmark_mkl.cpp)"},
        "fused_engine": fused,
        "fused_engine_graph_replayed_sweep": graphed,
        "primitive_sequence": prim,
        "conventional_mode_s": float(conv.group(1)) if conv else None,
    }


def bench_mixed(hip, synth, ge, sizes=(128, 192, 256, 384), per_class=256, iters=300, warmup=5,
                timing_iters=5, pattern_iters=3, fp64=True):
    """Mixed-size batch: 256 patterns each at 128^2, 192^2, 256^2 and 384^2
    (built as the single-pattern workload, autocorrelation supports), Elser's
    difference map with beta = 0.8 and positivity for 300 iterations, one
    column and one row launch per iteration for the whole batch (graph
    replays).  Also timed: each size class alone, one launch pair per class,
    and one launch triple per pattern (the naive form, eager, a few
    iterations)."""
    import numpy as np

    t0 = time.perf_counter()
    data = {}
    for n in sizes:
        d = synth.make_2d_batch(per_class, n, 30.0 * n / 256, 50.0 * n / 256, 1e6,
                                starts_per_pattern=1, config=3)
        data[n] = d
    log(f"bench[mixed]: data ready in {time.perf_counter() - t0:.1f} s")
    dev = {n: (hip.to_device(data[n]["x0"]), hip.to_device(data[n]["amp"]),
               hip.to_device(data[n]["support"])) for n in sizes}
    classes = [dev[n] for n in sizes]
    h = hip.HipImpl(64, 64, 1, 1)
    kw = dict(beta=0.8, space_mode=1, textbook=1)

    def timed(cls, mode, k, graph=1):
        h.Sync()
        t = time.perf_counter()
        h.dm_mixed_run(cls, launch_mode=mode, iters=k, use_graph=graph, **kw)
        h.Sync()
        return time.perf_counter() - t

    timed(classes, hip.MIXED_FUSED, warmup)
    el = timed(classes, hip.MIXED_FUSED, iters)
    its = iters / el
    log(f"bench[mixed]: fused {iters} iterations in {el:.3f} s -> {its:.1f} it/s")
    el_cls = timed(classes, hip.MIXED_PER_CLASS, iters)
    el_pat = timed(classes, hip.MIXED_PER_PATTERN, pattern_iters, graph=0)
    per_size = {}
    for n in sizes:
        e = timed([dev[n]], hip.MIXED_FUSED, iters)
        per_size[f"{n}x{n}"] = {"iters_per_s": round(iters / e, 1),
                                "pattern_iters_per_s": round(per_class * iters / e, 1)}
    h.set_kernel_timing(True)
    h.dm_mixed_run(classes, launch_mode=hip.MIXED_FUSED, iters=timing_iters, use_graph=0, **kw)
    h.Sync()
    kt = h.kernel_times()
    h.set_kernel_timing(False)
    px = per_class * sum(n * n for n in sizes)
    # two-pass textbook map: the column pass reads Y1, Y2 (16) and A (4) and
    # writes V, W (16); the row pass reads V, W (16) and the support (1) and
    # writes Y1, Y2 (16)
    kern = {"k_mixed_cols": kernel_entry(kt["column"][0], kt["column"][1], timing_iters, px * 36),
            "k_mixed_dm_rows": kernel_entry(kt["row"][0], kt["row"][1], timing_iters, px * 33)}
    dom = max(kern, key=lambda k: kern[k]["ms_per_iteration"])
    rec = {
        "metric": "mixed-size DM iters/s (256 each of 128^2, 192^2, 256^2, 384^2)",
        "value": round(its, 2), "unit": "iters/s", "steps": iters, "warmup": warmup,
        "ms_per_step": round(1e3 * el / iters, 4),
        "config": {"workload": f"textbook difference map beta 0.8 with positivity, {per_class} patterns "
                               f"each of {', '.join(f'{n}^2' for n in sizes)} complex64, one column and "
                               "one row launch per iteration (graph replays)", "patterns": per_class * len(sizes)},
        "pattern_iters_per_s": round(per_class * len(sizes) * its, 1),
        "per_size_class": per_size,
        "one_launch_per_class_iters_per_s": round(iters / el_cls, 2),
        "one_launch_per_pattern_iters_per_s": round(pattern_iters / el_pat, 2),
        "speedup_vs_launch_per_pattern": round((iters / el) / (pattern_iters / el_pat), 1),
        "hbm_frac_algorithmic_21Bpx": round(px * 21 * its / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": kern[dom]["achieved_GBs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": kern[dom]["frac"], "kernels": kern},
    }
    if fp64:
        o = ge.load_oracle()
        errs = {}
        with hip.HipImpl(64, 64, 1, 1) as hs:
            for n in sizes:
                x0 = data[n]["x0"][:2]
                xs = (hip.to_device(x0), hip.to_device(data[n]["amp"][:2]), hip.to_device(data[n]["support"][:2]))
                hs.dm_mixed_run([xs], iters=1, use_graph=0, **kw)
                hs.Sync()
                ref = o.dm_textbook(x0, data[n]["amp"][:2], 0.8, 1, mask=data[n]["support"][:2], space_mode=1)
                errs[f"{n}x{n}"] = float(np.linalg.norm(xs[0].numpy() - ref) / np.linalg.norm(ref))
        rec["fp64_check"] = {"patterns_per_class": 2, "iters": 1, "rel_l2": errs}
    h.close()
    return rec


WORKLOADS = ("hio2d", "hio2d_8starts", "carousel3d", "gd2d", "gd3d", "mixed", "conv3d", "ref_benchmark")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--patterns", type=int, default=512)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--workload", choices=WORKLOADS, default="hio2d",
                    help="hio2d (default): the headline line, which also carries the carousel "
                         "and gradient-descent workloads under 'workloads'; the others print "
                         "their own line only")
    ap.add_argument("--no-extra", action="store_true",
                    help="headline only (no nested carousel / gradient-descent runs)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp64-check", action="store_true")
    ap.add_argument("--timing-iters", type=int, default=20)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tuning", default="",
                    help="apply this autotune spec (e.g. 'col_256=4;rows_sparse=0;chunk=256;') "
                         "instead of autotuning, so that profiler runs repeat one choice")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one MergeAddData (dst = 1*src + 0*dst) over 2^26 complex64 "
                         "first: a kernel of known traffic (16 B read, 8 B written per "
                         "element) to scale PMC FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; every box here has one GPU, "
                                   "see the sharding design for more", "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np

    import __graft_entry__ as ge

    ge.load_package()
    from cpra_amd import hip, synth

    assert hip.device_count() > 0, "bench.py needs a GPU"
    fp64 = not args.no_fp64_check
    if args.calibrate_copy:
        ncal = 1 << 26
        a_ = hip.zeros(ncal, np.complex64)
        b_ = hip.zeros(ncal, np.complex64)
        with hip.HipImpl(1, 1, 1, 1) as hc:
            hc.MergeAddData(a_, b_, 1.0, 0.0, ncal)
            hc.Sync()
        a_.free()
        b_.free()
        log("bench: copy calibration kernel done (2^26 complex64)")
    if args.workload != "hio2d":
        fn = {"hio2d_8starts": bench_hio_8starts, "carousel3d": bench_carousel, "gd2d": bench_gd2d,
              "gd3d": bench_gd3d,
              "mixed": bench_mixed, "conv3d": bench_conv3d,
              "ref_benchmark": bench_ref_benchmark}[args.workload]
        extra = {"tuning": args.tuning} if args.workload in ("carousel3d", "conv3d") and args.tuning else {}
        rec = fn(hip, synth, ge, fp64=fp64, **extra)
        rec = {"n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "c64", "data": "synthetic", **rec}
        print(json.dumps(rec), flush=True)
        return
    n, P = args.size, args.patterns
    log(f"bench: generating {P} x {n}^2 patterns")
    t0 = time.perf_counter()
    data = synth.make_2d_batch(P, n, 30.0 * n / 256, 50.0 * n / 256, 1e6, starts_per_pattern=1,
                               config=1)
    log(f"bench: data ready in {time.perf_counter() - t0:.1f} s")
    x = hip.to_device(data["x0"])
    amp = hip.to_device(data["amp"])
    sup = hip.to_device(data["support"])
    sched = dict(synth.HIO_SCHEDULE)
    sched.pop("iters")
    h = hip.HipImpl(n, n, 1, P)

    # autotune the kernel variants on this workload (restores x, support);
    # the pick goes to the autotuner's JSON cache (timed afresh every run)
    import tempfile

    tune_cache = os.path.join(tempfile.gettempdir(), "cpra_hip_tuning.json")
    hip.set_tuning_cache(tune_cache, refresh=True)
    if args.tuning:
        h.set_tuning(args.tuning)
        tune = f"fixed: {args.tuning}"
    else:
        tune = "" if args.no_autotune else h.autotune(x, amp, sup, iters=10)
    log(f"bench: autotune {tune}")

    # warmup: at least one whole shrinkwrap chunk from the seeded start, so
    # that the chunk's graph (sw_every iterations) is captured here and the
    # timed run only replays it
    warm = max(args.warmup, sched["sw_every"])
    h.hio_run(x, amp, sup, iters=warm, use_graph=1, **sched)
    h.Sync()
    captures0 = int(h.get_tuning()["graph_captures"])
    log(f"bench: warmup done ({warm} iterations, {captures0} graph captures so far)")

    def timed_run(k, g=1):
        # the timed steps are iterations 1..k of the schedule: the state goes
        # back to the seeded start first (outside the timed region), since
        # the work per iteration follows the support, which the shrinkwrap
        # updates change (the first update grows it to every row)
        x.upload(data["x0"])
        sup.upload(data["support"])
        h.Sync()
        t0 = time.perf_counter()
        h.hio_run(x, amp, sup, iters=k, use_graph=g, **sched)
        h.Sync()
        return time.perf_counter() - t0

    elapsed = timed_run(args.steps)
    its = args.steps / elapsed
    captures1 = int(h.get_tuning()["graph_captures"])
    log(f"bench: {args.steps} iterations in {elapsed:.3f} s -> {its:.1f} it/s "
        f"({captures1 - captures0} graph captures in the timed run)")
    # the same steps with eager launches, beside the value
    its_eager = args.steps / timed_run(args.steps, 0)
    log(f"bench: eager launches -> {its_eager:.1f} it/s")
    full_sched = None
    if args.steps != 200:
        # the whole 200-iteration schedule, for reference beside the value
        el200 = timed_run(200)
        full_sched = round(200 / el200, 2)
        log(f"bench: the 200-iteration schedule -> {full_sched} it/s")

    # per-kernel durations with HIP events on the engine's stream (eager
    # launches of the same workload from the initial state): the column
    # kernel, whose algorithmic bytes do not depend on the support, over the
    # whole timed schedule; the row kernel over its first shrinkwrap interval,
    # where the support (and with it the row mode and the rows the row kernel
    # processes) is the initial one and known
    def event_times(iters):
        xt = hip.to_device(data["x0"])
        st_ = hip.to_device(data["support"])
        h.set_kernel_timing(True)
        h.hio_run(xt, amp, st_, iters=iters, use_graph=0, **sched)
        h.Sync()
        kt_ = h.kernel_times()
        h.set_kernel_timing(False)
        xt.free()
        st_.free()
        return kt_
    kt = event_times(args.timing_iters)
    # the row mode of that first interval (the whole schedule ends in another)
    tuning = h.get_tuning()
    kt_sched = event_times(args.steps)
    px = P * n * n
    # sparse rows: the row kernel reads and writes only the rows that hold a
    # support pixel (the column kernel finishes the others)
    row_frac = float(data["support"].reshape(P, n, n).any(axis=2).mean())
    sparse_used = tuning.get("rows_sparse_used") == "1"
    row_bpp = ROW_BYTES_PER_PX * (row_frac if sparse_used else 1.0)
    kern = {}
    for name, bpp, src, iters in (("column", COL_BYTES_PER_PX, kt_sched, args.steps),
                                  ("row", row_bpp, kt, args.timing_iters)):
        ms, cnt = src[name]
        per_iter = ms / iters                        # all launches of one iteration
        per_launch = ms / max(cnt, 1)
        ach = px * bpp / (per_iter * 1e-3) / 1e9
        kern[name] = {"avg_ms": per_launch, "launches": cnt, "ms_per_iteration": per_iter,
                      "bytes_per_launch": int(px * bpp * iters // max(cnt, 1)),
                      "achieved_GBs": ach, "frac": ach / HBM_PEAK_GBS,
                      "timed_iterations": iters}
    sw_ms, sw_n = kt_sched["shrinkwrap"]
    dom = max(kern, key=lambda k: kern[k]["ms_per_iteration"])
    roof = {"bound": "hbm", "achieved": round(kern[dom]["achieved_GBs"], 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(kern[dom]["achieved_GBs"] / HBM_PEAK_GBS, 4),
            "traffic": None, "kernel": f"k_eng_{'cols' if dom == 'column' else 'rows_hio'}<{n}>",
            "kernels": kern,
            "shrinkwrap_update_ms": sw_ms / max(sw_n, 1),
            "rows": {"sparse": sparse_used, "support_row_fraction": round(row_frac, 4)},
            "tuning": tuning}
    # counters are per launch of the 512 x 256^2 workload at a given tuning
    if P == 512 and n == 256:
        tb = measured_traffic(roof["kernel"], tuning)
        if tb is not None:
            roof["traffic"] = tb
            roof["traffic_source"] = ("profiles/hio_512x256sq_pmc.md (rocprofv3 FETCH_SIZE + "
                                      "WRITE_SIZE, same column variant, chunk and row mode)")
    iter_algo_frac = px * HIO_ALGO_BYTES_PER_PX * its / 1e9 / HBM_PEAK_GBS
    h.close()

    result = {
        "metric": METRIC,
        "value": round(its, 2),
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": f"batched 2D shrinkwrap-HIO, {P} x {n}x{n} complex64, "
                               f"beta 0.9, shrinkwrap every 20 (sigma 3.0->1.5, thr 0.1)",
                   "patterns": P, "size": n, "graph": True,
                   "graph_captures_in_timed_run": captures1 - captures0,
                   "warmup_iterations_run": warm,
                   "timed_iterations": f"1..{args.steps} of the schedule, from the seeded start"},
        "schedule_200_iters_per_s": full_sched if full_sched is not None else round(its, 2),
        "value_eager_launches": round(its_eager, 2),
        "pattern_iters_per_s": round(its * P, 1),
        "hbm_frac_algorithmic_21Bpx": round(iter_algo_frac, 4),
        "roofline": roof,
        "autotune": tune,
        "autotune_cache": tune_cache,
    }

    if not args.no_fp64_check:
        o = ge.load_oracle()
        k = 2
        # (1) one iteration from the GPU's current state (after the timed run)
        full = x.numpy()
        prev = full[:k].copy()
        xs = hip.to_device(prev)
        ss = hip.to_device(sup.numpy()[:k])
        as_ = hip.to_device(data["amp"][:k])
        with hip.HipImpl(n, n, 1, k) as hs:
            hs.hio_run(xs, as_, ss, iters=1, use_graph=0, **{**sched, "sw_every": 0})
            hs.Sync()
        step_ref = o.hio_step(prev, data["amp"][:k], ss.numpy(), 0.9, True, True)
        pm = o.hio_step(prev, data["amp"][:k], np.ones_like(ss.numpy()), 0.0, False, True)
        near = (ss.numpy() > 0) & (np.abs(pm.real) < 1e-5 * np.abs(pm).max())
        got = xs.numpy()
        step_err = float(np.linalg.norm((got - step_ref)[~near]) / np.linalg.norm(step_ref[~near]))
        # (2) the full 200-iteration schedule from the initial state
        xs = hip.to_device(data["x0"][:k])
        ss = hip.to_device(data["support"][:k])
        with hip.HipImpl(n, n, 1, k) as hs:
            rf = hs.hio_run(xs, as_, ss, iters=200, use_graph=1, want_rfactor=True, **sched)
        ref, _ = o.hio(data["x0"][:k], data["amp"][:k], data["support"][:k], 200, **sched)
        rref = o.rfactor(ref, data["amp"][:k])
        rng = np.random.default_rng(7)
        xp = data["x0"][:k].astype(complex) * (1 + 6e-8 * rng.standard_normal((k, n, n)))
        pert, _ = o.hio(xp, data["amp"][:k], data["support"][:k], 200, **sched)
        rpert = o.rfactor(pert, data["amp"][:k])
        result["fp64_check"] = {
            "patterns": k,
            "per_iteration_rel_l2": step_err,
            "per_iteration_boundary_pixels_excluded": int(near.sum()),
            "iters": 200,
            "rfactor_gpu": [round(float(v), 6) for v in rf],
            "rfactor_fp64": [round(float(v), 6) for v in rref],
            "rfactor_fp64_perturbed_6e-8": [round(float(v), 6) for v in rpert],
            "rfactor_rel_err": float(np.max(np.abs(rf - rref) / rref)),
            "rfactor_fp64_self_spread": float(np.max(np.abs(rpert - rref) / rref)),
        }
        log(f"bench: fp64 check {result['fp64_check']}")

    if not args.no_extra:
        result["workloads"] = {}
        for name, fn in (("hio2d_8starts", bench_hio_8starts), ("carousel3d", bench_carousel),
                         ("gd2d", bench_gd2d), ("gd3d", bench_gd3d),
                         ("mixed", bench_mixed), ("conv3d", bench_conv3d),
                         ("ref_benchmark", bench_ref_benchmark)):
            log(f"bench: workload {name}")
            result["workloads"][name] = fn(hip, synth, ge, fp64=fp64)

    if not args.no_cpu:
        log("bench: CPU baseline (the reference's MKL path)")
        mkl = cpu_baseline_mkl_path()
        if mkl is None or "error" in mkl:
            log(f"bench: MKL path unavailable ({mkl}); torch CPU FFT instead")
            result["cpu_baseline"] = {**cpu_baseline_torch(), "mkl_path_error": mkl}
        else:
            result["cpu_baseline"] = mkl
        log(f"bench: cpu baseline {result['cpu_baseline']}")
        log("bench: CPU baseline (FP64 oracle)")
        result["cpu_baseline_fp64_oracle"] = cpu_baseline_oracle(data, 16)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
