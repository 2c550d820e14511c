"""Benchmark of the fused PBD fluid step on one MI355X.

Workload: the 1,048,576-particle jittered lattice (128 x 64 x 128, spacing
0.005, h = 0.0125, 5 PBD iterations, dt = 1e-3, scene [0,1]^3), inputs resident
in HBM. A step is one PositionBasedFluidModel::step (neighbor search,
integration, 5 density/lambda/displacement iterations, velocity update).

Prints one JSON line: particle-steps/s over the timed steps, the roofline
position of the dominant kernel (per-pass HIP-event timing, on the stream the
kernels run on) and the CPU baseline (the oracle on this host's cores, a
bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "particle-steps/s at 1M PBD fluid; batched vs serial; bwd/fwd time ratio"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3      # MI355X published FP32 vector peak (packed issue)
FP32_NONPACKED_TFLOPS = 78.6  # the same without packed FP32, which the library's build disables

# Algorithmic traffic and flops (fp32, the reference's pass structure; neighbor
# data count as on-chip, so no list bytes):
#   bytes per particle: sort/bin 60, integrate 48, per iteration
#   density+lambda 32 and displacement 40, velocity update 48; 12 per grid cell
#   (the reference grid GridHash::setSpace builds, nx*ny*nz cells).
#   flops per neighbor pair (P = sum of neighbor-set sizes): 9 for the distance
#   test of the neighbor pass, 33 in density+lambda, 24 in displacement; 30 per
#   particle for the per-particle terms.
BYTES_PASS_A = 32
BYTES_PASS_B = 40
BYTES_NEIGHBORS = 16  # reads x^n once
FLOPS_PAIR_A = 33
FLOPS_PAIR_B = 24
FLOPS_PAIR_NBR = 9
TRAFFIC_PROFILE = ROOT / "profiles" / "pbf_step_lattice_1m_hbm_traffic.md"


def algorithmic_bytes(n, ncell, iters):
    """B_alg = (60 + 48 + 72 K + 48) N + 12 C bytes per step."""
    return (60 + 48 + (BYTES_PASS_A + BYTES_PASS_B) * iters + 48) * n + 12 * ncell


def algorithmic_flops(n, pairs, iters):
    """F_alg = (9 + 57 K) P + 30 N flops per step (true neighbor pairs only)."""
    return (FLOPS_PAIR_NBR + (FLOPS_PAIR_A + FLOPS_PAIR_B) * iters) * pairs + 30 * n


def kernel_counts(kernel, n, pairs):
    """(algorithmic bytes, flops) of one launch of `kernel` over n particles
    with `pairs` neighbor pairs."""
    return {"k_density_lambda": (BYTES_PASS_A * n, FLOPS_PAIR_A * pairs),
            "k_displacement": (BYTES_PASS_B * n, FLOPS_PAIR_B * pairs),
            "k_neighbors": (BYTES_NEIGHBORS * n, FLOPS_PAIR_NBR * pairs)}[kernel]


def roofline_of(bytes_, flops, seconds):
    """Position of `bytes_` and `flops` done in `seconds` against the HBM and
    FP32 peaks: the bound is the larger fraction."""
    hbm = bytes_ / seconds / (HBM_PEAK_GBS * 1e9)
    valu = flops / seconds / (FP32_PEAK_TFLOPS * 1e12)
    if valu >= hbm:
        return {"bound": "valu", "achieved": round(flops / seconds / 1e12, 3), "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(valu, 4), "hbm_frac": round(hbm, 4), "valu_frac": round(valu, 4)}
    return {"bound": "hbm", "achieved": round(bytes_ / seconds / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(hbm, 4), "hbm_frac": round(hbm, 4), "valu_frac": round(valu, 4)}


def profiled_traffic(kernel, block=None, persistent=False):
    """(HBM bytes per dispatch, profiled kernel name) of `kernel` from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload
    (copy-calibrated): the instance with the given workgroup size and launch
    form when profiled, else the kernel's only profiled instance; (None, None)
    when absent."""
    try:
        for line in TRAFFIC_PROFILE.read_text().splitlines():
            if line.startswith("traffic_json: "):
                table = json.loads(line[len("traffic_json: "):])
                names = [kernel + ("_p" if persistent else "") + f"<{block}>"] if block else []
                names += [k for k in table if k == kernel or k.startswith(kernel + "<")
                          or k.startswith(kernel + "_p<")]
                for name in names:
                    if name in table:
                        return table[name].get("bytes"), name
    except OSError:
        pass
    return None, None


def achievable_peaks(torch, pkg_dir, reps=10):
    """Achievable rates on this GPU from the probes of libpbf_peaks.so: a float4
    copy of 1 GiB (read + write bytes per second) and chains of FP32 fused
    multiply-adds (non-packed, as the library is built), each timed with HIP
    events on the stream it runs on."""
    import ctypes as C
    lib = C.CDLL(str(Path(pkg_dir) / "_build" / "libpbf_peaks.so"))
    lib.pbf_peak_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_longlong, C.c_void_p]
    lib.pbf_peak_fma.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    sh = C.c_void_p(stream.cuda_stream)
    n4 = 1 << 26  # 1 GiB of float4
    a = torch.ones(4 * n4, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    blocks, iters = 256 * 32, 2048
    f = torch.empty(blocks * 256, dtype=torch.float32, device=dev)

    def timed(launch):
        for _ in range(2):
            if launch() != 0:
                raise RuntimeError("peak probe launch failed")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            launch()
        e1.record(stream)
        torch.cuda.synchronize(dev)
        return e0.elapsed_time(e1) / 1e3 / reps

    t_copy = timed(lambda: lib.pbf_peak_copy(C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), n4, sh))
    t_fma = timed(lambda: lib.pbf_peak_fma(C.c_void_p(f.data_ptr()), blocks, iters, sh))
    if not torch.equal(a, b):
        raise RuntimeError("peak copy probe gave a wrong copy")
    del a, b, f
    return {"hbm_copy_gbs": round(2 * 16 * n4 / t_copy / 1e9, 1),
            "fp32_fma_tflops": round(16.0 * iters * blocks * 256 / t_fma / 1e12, 2),
            "probe": "float4 copy of 1 GiB, four per thread; 8 independent FMA chains per thread, "
                     "2,048 rounds, 2M threads"}


def progress(msg):
    """One line per finished phase on stderr (the JSON result stays alone on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_rate(x, steps, threads, warm=1):
    """Particle-steps/s of the oracle on `threads` host threads: `warm` untimed
    steps, then `steps` timed steps. Returns (rate, threads used, seconds)."""
    import __graft_entry__ as G
    O = G.load_oracle()
    used = O.set_threads(threads)
    p = O.make_params()
    v = np.zeros_like(x)
    if warm:
        x, v, _, _ = O.step(p, x, v, warm)
    t0 = time.perf_counter()
    O.step(p, x, v, steps)
    dt = time.perf_counter() - t0
    return len(x) * steps / dt, used, dt


def cpu_baseline(x, steps, threads, runs=3):
    """The oracle (C++ restatement of the reference path, OpenMP over
    particles) on this host: `runs` runs, each one warm-up step and then
    `steps` timed steps on `threads` threads (median), and one step on one
    thread."""
    rates, secs = [], []
    for _ in range(runs):
        rate, used, dt = oracle_rate(x, steps, threads)
        rates.append(rate)
        secs.append(dt)
    progress(f"cpu baseline: {runs} x {steps} steps on {used} threads, {sum(secs):.1f} s")
    rate1, _, dt1 = oracle_rate(x, 1, 1, warm=0)
    progress(f"cpu baseline: 1 step on 1 thread in {dt1:.1f} s")
    return {"value": float(np.median(rates)), "unit": "particle-steps/s", "cores": used, "kind": "port",
            "sample": f"{len(x)}-particle lattice (the benchmark workload): median of {runs} runs of {steps} steps, "
                      f"each after 1 warm-up step ({', '.join(f'{t:.2f}' for t in secs)} s)",
            "runs": [round(r, 1) for r in rates],
            "one_thread": {"value": rate1, "sample": f"1 step, {dt1:.1f} s"}, "host": host_cpu(),
            "threads_rule": "OMP_NUM_THREADS when set (the process's CPU share), else all CPUs of its affinity"}


TUNE = True  # autotune the launch settings of every measured context (--no-autotune: the defaults)
FIXED_TUNING = None  # a launch setting given with --tuning, used instead


def tune_summary(r):
    return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}


def gpu_rate(torch, fluid, x_np, steps, warmup, runs=3, params=None):
    """Median over `runs` runs, each from the same initial state: `warmup`
    untimed steps and one more step, then `steps` steps in one call timed with
    HIP events on the stream. Before them (untimed) the autotuner picks the
    context's launch settings on the initial state; ctx.tune_report holds its
    report. Returns (median ms per step, all ms per step, context, x, v)."""
    dev = torch.device("cuda:0")
    ctx = fluid.Context(dev, **(params or {}))
    ctx.reserve(len(x_np))
    ctx.tune_report = None
    if FIXED_TUNING:
        ctx.set_tuning(**FIXED_TUNING)
        ctx.tune_report = {**FIXED_TUNING, "fixed": True}
    elif TUNE:
        x0 = torch.from_numpy(x_np).to(dev)
        ctx.tune_report = tune_summary(ctx.autotune(x0, torch.zeros_like(x0), steps=min(max(steps, 2), 5)))
        del x0
    stream = torch.cuda.current_stream(dev)
    times = []
    for _ in range(runs):
        x = torch.from_numpy(x_np).to(dev)
        v = torch.zeros_like(x)
        ctx.step(x, v, None, max(warmup, 1))
        torch.cuda.synchronize()
        ctx.step(x, v, None, 1)  # reservations grown after the warm-up are in place before the timer starts
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        ctx.step(x, v, None, steps)
        e1.record(stream)
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1) / steps)
    return float(np.median(times)), times, ctx, x, v


def profiled_run(torch, fluid, ctx, x_np, steps, warmup):
    """The timed run again from the same initial state (`warmup` + 1 untimed
    steps, then `steps` steps in one call) with HIP events around every pass;
    then the same trajectory as single-step calls for the neighbor-pair count
    of each timed step. Returns ({pass: (ms per launch, launches)}, pairs per
    timed step)."""
    dev = torch.device("cuda:0")
    x = torch.from_numpy(x_np).to(dev)
    v = torch.zeros_like(x)
    ctx.step(x, v, None, max(warmup, 1))
    ctx.step(x, v, None, 1)
    torch.cuda.synchronize()
    ctx.set_profiling(True)
    ctx.step(x, v, None, steps)
    prof = ctx.profile()
    ctx.set_profiling(False)
    pass_ms = {k: (ms / max(cnt, 1), cnt) for k, (ms, cnt) in prof.items()}
    x = torch.from_numpy(x_np).to(dev)
    v = torch.zeros_like(x)
    ctx.step(x, v, None, max(warmup, 1))
    ctx.step(x, v, None, 1)
    pairs = []
    for _ in range(steps):
        ctx.step(x, v, None, 1)
        pairs.append(int(ctx.stats()["neighbor_pairs"]))
    return pass_ms, pairs


def host_cpu():
    """Model, sockets, physical cores and the CPUs this process may run on
    (lscpu where present, /proc/cpuinfo otherwise)."""
    info = {"model": cpu_model(), "affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        kv = {k.strip(): v.strip() for k, v in (ln.split(":", 1) for ln in out.splitlines() if ":" in ln)}
        info["model"] = kv.get("Model name", info["model"])
        for key, name in (("Socket(s)", "sockets"), ("Core(s) per socket", "cores_per_socket"),
                          ("Thread(s) per core", "threads_per_core"), ("CPU(s)", "cpus")):
            if key in kv and kv[key].isdigit():
                info[name] = int(kv[key])
        if "sockets" in info and "cores_per_socket" in info:
            info["physical_cores"] = info["sockets"] * info["cores_per_socket"]
    except (OSError, ValueError):
        pass
    return info


def baseline_threads():
    """Host threads for the CPU baseline: OMP_NUM_THREADS when the environment
    sets it (the process's share of the host), else every CPU this process may
    run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return min(int(env), len(os.sched_getaffinity(0)))
    return len(os.sched_getaffinity(0))


def config1(torch, fluid, scene, threads):
    """The 32,768-particle dam-break lattice, 100 steps: GPU (median of 3) and
    the oracle on the host (100 steps on all threads, 10 on one)."""
    x_np = scene.dam_break_32k()
    ms, runs, ctx, _, _ = gpu_rate(torch, fluid, x_np, 100, 5)
    replays = ctx.stats()["graph_replays"]
    rate, used, dt = oracle_rate(x_np, 100, threads)
    rate1, _, dt1 = oracle_rate(x_np, 10, 1)
    progress(f"config 1: GPU {ms:.4f} ms/step; CPU {dt:.1f} s for 100 steps on {used} threads")
    n = len(x_np)
    eager, _, _, _, _ = gpu_rate(torch, fluid, x_np, 100, 5, params={"flags": 4})
    progress(f"config 1: eager launches {eager:.4f} ms/step")
    return {"workload": "32,768-particle dam-break lattice at (0.10, 0.50, 0.10), seed 1; 100 steps",
            "gpu": {"value": round(n / (ms / 1e3), 1), "unit": "particle-steps/s", "ms_per_step": round(ms, 4),
                    "runs_ms_per_step": [round(t, 4) for t in runs], "graph_replays": replays,
                    "eager_ms_per_step": round(eager, 4), "tuning": ctx.tune_report},
            "cpu": {"value": round(rate, 1), "cores": used, "seconds": round(dt, 2),
                    "one_thread": round(rate1, 1)}}


def config2b(torch, fluid, scene, steps=200, dense_steps=20):
    """8 Gaussian blobs x 131,072 particles (sigma 0.05, seed 3): GPU rate over
    `steps` steps after 2 (median of 3), the same over the first `dense_steps`
    of them (the dense phase before the blobs fly apart), neighbor-count mean
    and maximum from the library, and the count histogram of a 20,000-particle
    sample of the input (k-d tree, exact distances)."""
    from scipy.spatial import cKDTree
    x_np = scene.blobs_1m()
    ms_d, runs_d, _, _, _ = gpu_rate(torch, fluid, x_np, dense_steps, 2)
    ms, runs, ctx, _, _ = gpu_rate(torch, fluid, x_np, steps, 2)
    st = ctx.stats()
    n = len(x_np)
    tree = cKDTree(x_np.astype(np.float64))
    pick = np.random.default_rng(3).choice(n, 20000, replace=False)
    counts = np.array([len(r) for r in tree.query_ball_point(x_np[pick].astype(np.float64), 0.0125)])
    edges = [1, 10, 30, 60, 100, 200, 300, 400, 500, 600, 800, 1200]
    hist, _ = np.histogram(counts, bins=edges + [10 ** 9])
    progress(f"config 2b: {ms:.3f} ms/step over {steps}, {ms_d:.3f} over the first {dense_steps}; "
             f"mean {st['neighbor_pairs'] / n:.1f} neighbors")
    return {"workload": "8 Gaussian blobs x 131,072 = 1,048,576 particles, sigma 0.05, seed 3; "
                        f"{steps} steps after 2", "particles": n,
            "value": round(n / (ms / 1e3), 1), "unit": "particle-steps/s", "ms_per_step": round(ms, 4),
            "runs_ms_per_step": [round(t, 4) for t in runs], "tuning": ctx.tune_report,
            "out_of_grid_last_step": int(st["out_of_grid"]),
            "dense_phase": {"steps": dense_steps, "value": round(n / (ms_d / 1e3), 1), "ms_per_step": round(ms_d, 4),
                            "runs_ms_per_step": [round(t, 4) for t in runs_d]},
            "neighbors_mean_last_step": round(st["neighbor_pairs"] / n, 1),
            "neighbors_max_last_step_slots": int(st["list_slots_needed"] // max(64 * ((n + 63) // 64), 1)),
            "input_sample_histogram": {"edges": edges, "counts": hist.tolist(),
                                       "mean": round(float(counts.mean()), 1), "max": int(counts.max())}}


def shard_share(torch, fluid, scene, steps=200, warmup=5):
    """One GPU's share of the 1M lattice split into 8 slabs along x: 16 of the
    128 columns (131,072 owned particles) plus 3 ghost columns on each side
    (h = 2.5 spacings), 180,224 particles, stepped at that shape. The halo
    exchange itself (ghost x before each density pass, ghost lambda before each
    displacement pass) is not run."""
    x = scene.jittered_lattice((22, 64, 128), (0.18 + 53 * 0.005, 0.40, 0.18), 0.005, 0.1, seed=2)
    ms, runs, ctx, xd, vd = gpu_rate(torch, fluid, x, steps, warmup)
    owned = 16 * 64 * 128
    progress(f"shard share: {len(x)} particles, {ms:.4f} ms/step")
    return {"workload": "one of 8 x-slabs of the 1M lattice: 16 owned + 2 x 3 ghost columns (x 64 x 128)",
            "particles": len(x), "owned": owned, "steps": steps, "ms_per_step": round(ms, 4),
            "runs_ms_per_step": [round(t, 4) for t in runs], "tuning": ctx.tune_report,
            "owned_particle_steps_per_s": round(owned / (ms / 1e3), 1)}


def batched_vs_serial(torch, fluid, scene, steps, serial_steps):
    """256 independent lattice scenes, n_s log-uniform in [2048, 65536] (seed
    4), each in its own [0,1]^3: all scenes in one pbf_batch_step call per
    `steps` steps, against a loop of one pbf_step per scene and step."""
    b = scene.batched_lattice_scenes(nscenes=256, nmin=2048, nmax=65536, seed=4)
    dev = torch.device("cuda:0")
    n = int(b.offsets[-1])
    off = torch.from_numpy(b.offsets.astype(np.int32)).to(dev)
    stream = torch.cuda.current_stream(dev)

    def timed(fn, k):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        fn(k)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    xb = torch.from_numpy(b.pos).to(dev)
    vb = torch.zeros_like(xb)
    rb = torch.empty(n, dtype=torch.float32, device=dev)
    cb = fluid.Context(dev)
    cb.reserve(n, 256)
    tuning = tune_summary(cb.autotune(xb, vb, off, steps=3)) if TUNE else None
    cb.batch_step(xb, vb, off, rb, 3)  # also captures the replayed middle step
    ms_b = timed(lambda k: cb.batch_step(xb, vb, off, rb, k), steps)
    cells = cb.stats()["cells_used"]
    progress(f"batched: {steps} steps of {n} particles in 256 scenes, {ms_b:.1f} ms")

    xs = torch.from_numpy(b.pos).to(dev)
    vs = torch.zeros_like(xs)
    rs = torch.empty(n, dtype=torch.float32, device=dev)
    cs = fluid.Context(dev)
    cs.reserve(int(b.counts.max()))
    bounds = [(int(a), int(e)) for a, e in zip(b.offsets[:-1], b.offsets[1:])]

    def serial(k):
        for _ in range(k):
            for a, e in bounds:
                cs.step(xs[a:e], vs[a:e], rs[a:e], 1)

    serial(1)
    ms_s = timed(serial, serial_steps)
    progress(f"serial: {serial_steps} steps, {ms_s:.1f} ms")
    vb_ = n * steps / (ms_b / 1e3)
    vs_ = n * serial_steps / (ms_s / 1e3)
    return {"workload": "256 lattice scenes, n_s log-uniform [2048, 65536], seed 4", "scenes": 256,
            "particles": n, "batched": {"value": round(vb_, 1), "unit": "particle-steps/s", "steps": steps,
                                        "ms_per_step": round(ms_b / steps, 4), "cells_used": cells,
                                        "tuning": tuning},
            "serial": {"value": round(vs_, 1), "unit": "particle-steps/s", "steps": serial_steps,
                       "ms_per_step": round(ms_s / serial_steps, 4),
                       "method": "one pbf_step call per scene and step (the plain per-scene launch loop)"},
            "speedup": round(vb_ / vs_, 3)}


def node_density(torch, x, cW, h, nodes=32, scene_of=None, nscenes=1):
    """rho_g = sum_j m m_factor W(|x_g - x_j|, h) on the cell-centre nodes of
    [0,1]^3 (spacing 1/32 > 2h: a particle reaches at most its nearest node),
    one grid per scene."""
    idx = torch.clamp(torch.floor(x * nodes), 0, nodes - 1)
    d = x - (idx + 0.5) / nodes
    r = torch.sqrt((d * d).sum(1) + 1e-30)
    t = torch.clamp(1.0 - r / h, min=0.0)
    flat = (idx[:, 0] + nodes * (idx[:, 1] + nodes * idx[:, 2])).long()
    if scene_of is not None:
        flat = flat + scene_of * nodes ** 3
    out = torch.zeros(nscenes * nodes ** 3, dtype=x.dtype, device=x.device)
    return out.index_add(0, flat, cW * t ** 3)


def differentiable(torch, fluid, scene, nsteps=50, every=10):
    """65,536-particle jittered block (64 x 32 x 32 at (0.3, 0.5, 0.3), seed 5),
    50 steps, checkpoint every 10: L = |mean(x_50) - c*|^2 + sum_g (rho_g -
    rho*_g)^2 on 32^3 nodes, targets from a run with seeded random v0 (seed 6);
    gradient w.r.t. v0. Forward = rollout + loss, backward = loss backward +
    rollout adjoint, each timed with HIP events on the stream the library runs
    on."""
    dev = torch.device("cuda:0")
    x0 = torch.from_numpy(scene.block_64k()).to(dev)
    n = x0.shape[0]
    ctx = fluid.Context(dev)
    ctx.reserve(n)
    st = ctx.stats()
    cW = float(st["density_factor"]) * float(st["particle_mass"]) * 15.0 / (np.pi * 0.0125 ** 3)
    vt = torch.from_numpy((0.2 * (scene.uniform01(6, 3 * n).reshape(n, 3) - 0.5)).astype(np.float32)).to(dev)
    xt, vt2 = x0.clone(), vt.clone()
    ctx.step(xt, vt2, None, nsteps)
    c_star = xt.mean(0)
    rho_star = node_density(torch, xt, cW, 0.0125)
    stream = torch.cuda.current_stream(dev)

    def run(dtype):
        xs = x0.to(dtype)
        v0 = torch.zeros_like(xs, requires_grad=True)
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        torch.cuda.synchronize()
        e0.record(stream)
        xN, vN, _ = fluid.rollout(ctx, xs, v0, nsteps, every)
        m = xN.mean(0) - c_star.to(dtype)
        L = (m * m).sum() + ((node_density(torch, xN, cW, 0.0125) - rho_star.to(dtype)) ** 2).sum()
        e1.record(stream)
        L.backward()
        e2.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1), e1.elapsed_time(e2), float(L), v0.grad

    run(torch.float32)
    f_ms, b_ms, L, g = run(torch.float32)
    progress(f"differentiable: fwd {f_ms:.2f} ms, bwd {b_ms:.2f} ms")
    run(torch.float64)
    f64_f, f64_b, L64, g64 = run(torch.float64)
    progress(f"differentiable, double path: fwd {f64_f:.2f} ms, bwd {f64_b:.2f} ms")
    rel = float((g.double() - g64).norm() / g64.norm())
    return {"workload": "65,536-particle block, 50 steps, checkpoint every 10, loss centroid + 32^3 density L2",
            "particles": n, "steps": nsteps, "checkpoint_every": every, "fwd_ms": round(f_ms, 3),
            "bwd_ms": round(b_ms, 3), "bwd_fwd_ratio": round(b_ms / f_ms, 3), "loss": L,
            "grad_norm": float(g.norm()),
            "f64": {"fwd_ms": round(f64_f, 3), "bwd_ms": round(f64_b, 3), "bwd_fwd_ratio": round(f64_b / f64_f, 3),
                    "loss": L64, "grad_norm": float(g64.norm()), "fp32_vs_f64_grad_rel": round(rel, 6)}}


def batched_inverse(torch, fluid, scene, adam_steps=20, nsteps=30, every=10, lr=0.05):
    """64 scenes, n_s log-uniform in [8192, 32768] (seed 8), per-scene initial
    velocity field optimised by Adam against per-scene targets (centroid and
    32^3 density field of a run from a seeded random v0); every Adam step is a
    30-step batched rollout and its adjoint."""
    dev = torch.device("cuda:0")
    b = scene.batched_lattice_scenes(nscenes=64, nmin=8192, nmax=32768, seed=8)
    n = int(b.offsets[-1])
    off = torch.from_numpy(b.offsets.astype(np.int32)).to(dev)
    sid = torch.from_numpy(np.repeat(np.arange(64), b.counts)).to(dev)
    counts = torch.from_numpy(b.counts.astype(np.float32)).to(dev)
    x0 = torch.from_numpy(b.pos).to(dev)
    ctx = fluid.Context(dev)
    ctx.reserve(n, 64)
    st = ctx.stats()
    cW = float(st["density_factor"]) * float(st["particle_mass"]) * 15.0 / (np.pi * 0.0125 ** 3)

    def centroids(x):
        return torch.zeros(64, 3, device=dev).index_add(0, sid, x) / counts[:, None]

    vt = torch.from_numpy((0.2 * (scene.uniform01(9, 3 * n).reshape(n, 3) - 0.5)).astype(np.float32)).to(dev)
    xt = x0.clone()
    ctx.batch_step(xt, vt, off, None, nsteps)
    c_star = centroids(xt)
    rho_star = node_density(torch, xt, cW, 0.0125, scene_of=sid, nscenes=64)
    v0 = torch.zeros_like(x0, requires_grad=True)
    opt = torch.optim.Adam([v0], lr=lr)
    losses = []
    stream = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for it in range(adam_steps + 1):
        if it == 1:  # the first step also sizes the context's reservations
            torch.cuda.synchronize()
            e0.record(stream)
        opt.zero_grad()
        xN, _, _ = fluid.rollout(ctx, x0, v0, nsteps, every, off)
        m = centroids(xN) - c_star
        L = (m * m).sum() + ((node_density(torch, xN, cW, 0.0125, scene_of=sid, nscenes=64) - rho_star) ** 2).sum()
        L.backward()
        opt.step()
        losses.append(float(L))
        if it % 5 == 0:
            progress(f"batched inverse: Adam step {it}, loss {losses[-1]:.6g}")
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / adam_steps
    grad = inverse_gradient_check(torch, fluid, ctx, b, off, sid, counts, cW, nsteps, every, c_star, rho_star)
    return {"workload": "64 scenes, n_s log-uniform [8192, 32768], seed 8; 30-step rollout + adjoint per Adam step",
            "scenes": 64, "particles": n, "adam_steps": adam_steps, "ms_per_adam_step": round(ms, 3),
            "loss_first": losses[0], "loss_last": losses[-1], "gradient_check": grad}


def inverse_gradient_check(torch, fluid, ctx, b, off, sid, counts, cW, nsteps, every, c_star, rho_star,
                           nscenes_fd=8, ncomp=16, eps=1e-6):
    """Gradient of the batched inverse loss at v0 = 0 against FP64 central
    differences on `nscenes_fd` scenes x `ncomp` seeded components, for the
    fp32 path and the double path. The differences run on the library's
    double forward (one batched run perturbs one component in each chosen
    scene; scenes are independent, so the per-scene losses separate them)."""
    dev = torch.device("cuda:0")
    n = int(b.offsets[-1])
    x0 = torch.from_numpy(b.pos).to(dev)

    def scene_losses(x):
        c = torch.zeros(64, 3, dtype=x.dtype, device=dev).index_add(0, sid, x) / counts[:, None].to(x.dtype)
        m = c - c_star.to(x.dtype)
        r = (node_density(torch, x, cW, 0.0125, scene_of=sid, nscenes=64) - rho_star.to(x.dtype)) ** 2
        return (m * m).sum(1) + r.reshape(64, -1).sum(1)

    grads, times = {}, {}
    stream = torch.cuda.current_stream(dev)
    for name, dt in (("fp32", torch.float32), ("f64", torch.float64)):
        v0 = torch.zeros((n, 3), dtype=dt, device=dev, requires_grad=True)
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        torch.cuda.synchronize()
        e0.record(stream)
        xN, _, _ = fluid.rollout(ctx, x0.to(dt), v0, nsteps, every, off)
        L = scene_losses(xN).sum()
        e1.record(stream)
        L.backward()
        e2.record(stream)
        torch.cuda.synchronize()
        grads[name] = v0.grad.double()
        times[name] = (e0.elapsed_time(e1), e1.elapsed_time(e2))
    rng = np.random.default_rng(10)
    pick = [int(s) for s in rng.choice(64, nscenes_fd, replace=False)]
    comps = {s: [(int(b.offsets[s]) + int(rng.integers(b.counts[s])), int(rng.integers(3))) for _ in range(ncomp)]
             for s in pick}
    fd = {s: [] for s in pick}
    for k in range(ncomp):
        ls = []
        for sign in (1.0, -1.0):
            vp = torch.zeros((n, 3), dtype=torch.float64, device=dev)
            for s in pick:
                i, c = comps[s][k]
                vp[i, c] = sign * eps
            x = x0.double().clone()
            ctx.rollout_forward(x, vp, None, nsteps, nsteps, off)
            ls.append(scene_losses(x))
        for s in pick:
            fd[s].append(float((ls[0][s] - ls[1][s]) / (2 * eps)))
    want = np.concatenate([fd[s] for s in pick])
    out = {"sample": f"{nscenes_fd} scenes x {ncomp} components of v0, central differences (eps {eps}) of the "
                     "double forward", "scenes": pick}
    for name in ("fp32", "f64"):
        g = grads[name].cpu().numpy()
        got = np.array([g[i, c] for s in pick for i, c in comps[s]])
        out[name] = {"rel_err": float(np.linalg.norm(got - want) / np.linalg.norm(want)),
                     "max_comp_err": float(np.abs(got - want).max() / np.abs(want).max()),
                     "fwd_ms": round(times[name][0], 3), "bwd_ms": round(times[name][1], 3)}
    progress(f"batched inverse gradient vs FP64 FD: fp32 {out['fp32']['rel_err']:.3e}, "
             f"double {out['f64']['rel_err']:.3e}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = all host cores")
    ap.add_argument("--runs", type=int, default=3, help="timed runs from the same initial state (median)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the 32k (config 1) and blob (config 2b) lines")
    ap.add_argument("--workload", default="lattice_1m", choices=["lattice_1m", "blobs_1m", "dam_break_32k"],
                    help="the main line's workload (the default is the benchmark's; the others are for profiles)")
    ap.add_argument("--batched-steps", type=int, default=100, help="steps of the 256-scene batched comparison")
    ap.add_argument("--serial-steps", type=int, default=10, help="steps of its serial per-scene loop")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiable and batched-inverse workloads")
    ap.add_argument("--no-peaks", action="store_true", help="skip the achievable copy / FMA rate probes")
    ap.add_argument("--no-autotune", action="store_true", help="measure with the default launch settings")
    ap.add_argument("--tuning", default="", metavar="BLOCK,STAGE,PERSISTENT[,LANES]",
                    help="measure with this launch setting instead of tuning (e.g. 128,1,0 for profiled runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 256 MiB device copy (reference kernel for FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    global TUNE, FIXED_TUNING
    TUNE = not args.no_autotune and not args.tuning
    if args.tuning:
        FIXED_TUNING = dict(zip(("pass_block", "stage_min_particles", "persistent", "pass_lanes"),
                                (int(t) for t in args.tuning.split(","))))
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU path is a design, not built",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import __graft_entry__ as G
    pkg = G.load_package()
    from importlib import import_module
    fluid = import_module(G.PKG_NAME + ".fluid")

    x_np = getattr(pkg.scene, args.workload)()
    n = len(x_np)
    progress(f"{args.workload}: {n} particles generated")
    dev = torch.device("cuda:0")
    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)  # 256 MiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    params = fluid.Context(dev).params()
    ms_step, runs, ctx, x, v = gpu_rate(torch, fluid, x_np, args.steps, args.warmup, args.runs)
    rho = torch.empty(n, dtype=torch.float32, device=dev)
    progress(f"{args.workload}: {args.steps} steps x {args.runs} runs, median {ms_step:.4f} ms/step")
    value = n / (ms_step / 1e3)

    stats = ctx.stats()
    cand = stats["candidates"]
    ncell_compact = int(stats["cells_used"])  # compact cells the sort touches (occupied window + bucket)
    ncell = int(np.prod(stats["grid_dims"]))   # the reference grid (GridHash::setSpace)
    iters = params["iterations"]

    # Per-pass timing of the timed region: one more run from the same initial
    # state with the same warm-up and steps, HIP events around every pass on
    # the kernels' stream; the neighbor-pair count of each of its steps from
    # single-step calls along the same trajectory (bit-identical to one call).
    pass_ms, pairs_per_step = profiled_run(torch, fluid, ctx, x_np, args.steps, args.warmup)
    pairs = float(np.mean(pairs_per_step))
    progress(f"per-pass profile done: mean P {pairs:.4g} over the timed steps")
    dom = max(("density_lambda", "displacement", "neighbors"), key=lambda k: pass_ms[k][0] * pass_ms[k][1])
    t_dom = pass_ms[dom][0] / 1e3
    kname = {"density_lambda": "k_density_lambda", "displacement": "k_displacement", "neighbors": "k_neighbors"}[dom]
    kb, kf = kernel_counts(kname, n, pairs)
    roofline = roofline_of(kb, kf, t_dom)
    tune = ctx.tuning()
    pair_lanes = ctx.stats()["pass_lanes"] == 2  # the lane-pair form of the passes ran
    traffic, traffic_kernel = profiled_traffic(kname + ("2" if pair_lanes and kname != "k_neighbors" else ""),
                                               tune["pass_block"] if kname != "k_neighbors" else None,
                                               bool(tune["persistent"]) and kname != "k_neighbors")
    roofline.update({
        "kernel": kname,
        "traffic": traffic,
        "traffic_source": TRAFFIC_PROFILE.name,
        "traffic_kernel": traffic_kernel,
        "algorithmic_bytes": int(kb), "algorithmic_flops": int(kf),
        "valu_frac_nonpacked": round(kf / t_dom / (FP32_NONPACKED_TFLOPS * 1e12), 4),
        "avg_launch_us": round(pass_ms[dom][0] * 1e3, 2), "launches": pass_ms[dom][1],
        "neighbor_pairs_mean": int(round(pairs)),
        "counting": "bytes 32 N (density+lambda) / 40 N (displacement) / 16 N (neighbors), no list bytes; "
                    "flops 33 P / 24 P / 9 P; P = mean neighbor pairs over the timed steps",
    })
    if roofline["traffic"]:  # measured HBM bytes per launch over the launch time
        roofline["traffic_gbs"] = round(roofline["traffic"] / t_dom / 1e9, 1)
        roofline["traffic_over_algorithmic"] = round(roofline["traffic"] / kb, 2)
    if not args.no_peaks:
        peaks = achievable_peaks(torch, Path(pkg.__file__).parent)
        peaks["hbm_frac"] = round(kb / t_dom / 1e9 / peaks["hbm_copy_gbs"], 4)
        peaks["valu_frac"] = round(kf / t_dom / 1e12 / peaks["fp32_fma_tflops"], 4)
        roofline["achievable"] = peaks
        progress(f"achievable: copy {peaks['hbm_copy_gbs']} GB/s, FP32 FMA {peaks['fp32_fma_tflops']} TFLOP/s")
    b_alg = algorithmic_bytes(n, ncell, iters)
    f_alg = algorithmic_flops(n, pairs, iters)
    step_s = ms_step / 1e3
    whole_step = {"bytes_alg": int(b_alg), "flops_alg": int(f_alg), **roofline_of(b_alg, f_alg, step_s)}

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": args.workload, "particles": n, "smoothing_length": params["smoothing_length"],
                   "spacing": params["sampling_distance"], "iterations": iters, "dt": params["dt"],
                   "grid_cells": ncell, "compact_cells": ncell_compact, "neighbor_pairs_mean": int(round(pairs)),
                   "neighbor_pairs_last": int(pairs_per_step[-1]), "candidate_pairs_last": cand},
        "roofline": roofline,
        "passes_ms": {k: round(v[0], 4) for k, v in pass_ms.items()},
        "passes_launches": {k: v[1] for k, v in pass_ms.items()},
        "step_roofline": whole_step,
        "runs_ms_per_step": [round(t, 4) for t in runs],
        "tuning": ctx.tune_report,
    }
    if not args.no_configs:
        out["shard_share"] = shard_share(torch, fluid, pkg.scene)
    if not args.no_batched:
        del x, v, rho, ctx
        out["batched_vs_serial"] = batched_vs_serial(torch, fluid, pkg.scene, args.batched_steps, args.serial_steps)
    if not args.no_diff:
        out["differentiable"] = differentiable(torch, fluid, pkg.scene)
        out["batched_inverse"] = batched_inverse(torch, fluid, pkg.scene)
    threads = args.cpu_threads or baseline_threads()
    if not args.no_configs:
        out["config1"] = config1(torch, fluid, pkg.scene, threads)
        out["config2b"] = config2b(torch, fluid, pkg.scene)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(x_np, args.cpu_steps, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
