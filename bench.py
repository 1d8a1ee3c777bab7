"""Benchmark of the MI355X Stam frame step.

Headline workload (``value``): the batched mix -- 512 independent scenes, N
drawn from {64,128,256,512}, viscosity and diffusion log-uniform in
[1e-6,1e-3], 0-4 rectangle obstacles, 20 Gauss-Seidel sweeps per lin_solve,
dt 0.1, exact (wavefront, bit-for-bit) mode -- advanced by one fused launch
per frame, K frames captured as one HIP graph.  A step is one frame of the
whole batch.  value = sum(N^2) * K / elapsed (interior cells; the ring is not
counted).

Also reported beside it: the single-grid drop-in on the same mix (one scene
at a time), the N=1024 single grid, the red-black legs, snapshots, serving,
the HBM roofline of the frame kernel and the CPU leg.  The CPU leg ("cpu") is
the only part of this script that runs the oracle: it checks the red-black
legs' GPU outputs against the f64 oracle (their vs_fp64 and Poisson residual
fields) and times the oracle on this host as cpu_baseline.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--scenes 512]
"""
import argparse
import concurrent.futures as cf
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

import numpy as np  # noqa: E402

from equilibrium_amd import _abi, scenes  # noqa: E402

METRIC = "cell-updates/s (N=4096 single + batched mix); p50/p99 frame latency; HBM GB/s"
# GPU outputs the CPU leg checks against the oracle (every oracle call of this
# script runs in that leg)
CHECKS = {}
HBM_PEAK_GBS = 8000.0


def b_alg(iters):
    """Algorithmic HBM bytes per cell-update (one interior cell, one frame):
    5 lin_solves x iters sweeps x 12 B (read x, read x0, write x) + 108 B for
    divergence/p-init (2x16), gradient (2x20), advect u,v (16) and advect
    density with the scratch copy (20)."""
    return 60 * iters + 108


def profiled_traffic(kernel, shape=None):
    """HBM bytes per launch measured by the committed rocprofv3 PMC runs, for
    the launch shape (batch: "max_waves:order:schedule", red-black:
    "band:sweeps") the run used; None when no profile
    was taken at that shape.  Returns (bytes, profile entry)."""
    try:
        with open(os.path.join(REPO, "profiles", "traffic.json")) as f:
            ent = json.load(f)[kernel]
    except (OSError, KeyError, ValueError):
        return None, None
    if "by_shape" in ent:
        ent = ent["by_shape"].get(shape)
    elif ent.get("shape") != shape:
        ent = None
    if ent is None:
        return None, None
    return ent.get("traffic_bytes_per_launch", ent.get("traffic_bytes_per_frame")), ent


def bench_bandwidth(gib=8, blocks_per_cu=(4, 8), reps=5):
    """Practical HBM ceiling: hand-written streaming read / write / copy
    kernels (libeqf_bwprobe.so) on a working set of `gib` GiB, far above the
    256 MB MALL; the best of the launch shapes tried."""
    import ctypes
    path = os.path.join(REPO, "equilibrium_amd", "libeqf_bwprobe.so")
    L = ctypes.CDLL(path)
    L.eqf_bwprobe_run.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_double)]
    best = {"read": 0.0, "write": 0.0, "copy": 0.0}
    for bpc in blocks_per_cu:
        o = (ctypes.c_double * 3)()
        rc = L.eqf_bwprobe_run(0, gib << 30, bpc, reps, o)
        if rc != 0:
            raise RuntimeError(f"bandwidth probe failed: HIP error {rc}")
        for k, v in zip(("read", "write", "copy"), o):
            best[k] = max(best[k], v)
    return {"unit": "GB/s", "working_set_gib": gib, **best,
            "kernels": "float4 non-temporal streams, 8 accesses in flight per thread, grid-stride"}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def device_sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


SCHED_NAMES = {0: "auto", 1: "wide", 2: "ring", 3: "fine"}  # eqf_batch_set_schedule


def make_scenes(scs):
    hs = []
    for sc in scs:
        h = _abi.Scene(scenes.scene_params(sc))
        scenes.apply_to_scene(h, sc)
        hs.append(h)
    for h in hs:
        h.sync()
    return hs


def bench_batch(scs, steps, warmup, tuning=None):
    from equilibrium_amd import tune

    hs = make_scenes(scs)
    if tuning:
        (w, o, sch), timings = (tuning + (_abi.SCHED_AUTO,))[:3], {}
        log(f"batch: max waves {w}, order {o}, schedule {SCHED_NAMES[sch]} (given)")
    else:
        w, o, sch, timings = tune.tune_batch(hs)
        log(f"batch: autotuned max waves {w}, order {o}, schedule {SCHED_NAMES[sch]}: {timings}")
    b = _abi.Batch(hs)
    b.set_tuning(w, o)
    b.set_schedule(sch)
    if warmup:
        b.step(warmup)
        b.sync()
    device_sync()
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    device_sync()
    wall = time.perf_counter() - t0
    ev_ms = b.last_ms()
    b.close()
    # one scene per size, checked bit for bit against the oracle in the CPU leg
    picks = {}
    for i, sc in enumerate(scs):
        picks.setdefault(sc["size"], i)
    CHECKS["batch"] = ([(scs[i], {name: hs[i].download(fid) for fid, name in enumerate(
        ("density", "scratch", "vx", "vy", "vx0", "vy0"))}) for i in sorted(picks.values())], warmup + steps)
    for h in hs:
        h.close()
    return wall, ev_ms, {"max_waves": w, "order": "largest first" if o == 0 else "given",
                         "schedule": SCHED_NAMES[sch], "shape": f"{w}:{o}:{sch}", "trial_ms": timings}


def check_batch_exact(samples, frames, threads):
    """The timed batch's scenes (one per size) against the oracle after all
    the frames the batch ran: every field, bit for bit."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from helpers import oracle_state

    def one(item):
        sc, got = item
        st = oracle_state(sc)
        st.run(sc["dt"], sc["diffusion"], sc["viscosity"], sc["iters"], frames,
               src_d=sc["src_d"], src_u=sc["src_u"], src_v=sc["src_v"])
        bad = {k: int(np.sum(got[k].view(np.uint32) != st.f[k].view(np.uint32))) for k in got}
        return {"N": sc["size"] - 2, "cells_differing": sum(bad.values())}

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(min(threads, len(samples))) as ex:
        res = list(ex.map(one, samples))
    return {"frames": frames, "scenes": res, "bit_exact": all(r["cells_differing"] == 0 for r in res),
            "oracle_seconds": time.perf_counter() - t0}


def bench_batch_redblack(scs, steps, warmup, check_frames=2, band=None):
    """The same mix in red-black mode: every frame kernel covers all 512
    scenes (band passes over all scenes' tiles with register temporal
    blocking), band shape and sweeps per launch autotuned; accuracy vs the f64
    red-black oracle on one scene per size, over the first frames."""
    from equilibrium_amd import tune

    def mk(sc):
        h = _abi.Scene(scenes.scene_params(sc, _abi.GS_REDBLACK))
        scenes.apply_to_scene(h, sc)
        return h

    hs = [mk(sc) for sc in scs]
    for h in hs:
        h.sync()
    if band:
        tile, tsw = band
        log(f"red-black batch: band {tune.TILES[tile]} x {tsw} sweeps/launch (given)")
    else:
        tile, tsw, timings = tune.tune_batch_band(hs)
        log(f"red-black batch: autotuned band {tune.TILES[tile]} x {tsw} sweeps/launch: {timings}")
    b = _abi.Batch(hs)
    b.set_band(tile, tsw)
    if warmup:
        b.step(warmup)
        b.sync()
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    wall = time.perf_counter() - t0
    ev_ms = b.last_ms()
    b.close()
    for h in hs:
        h.close()
    cells = sum((sc["size"] - 2) ** 2 for sc in scs)
    res = {"value": cells * steps / wall, "unit": "cell-updates/s", "ms_per_frame": wall * 1e3 / steps,
           "launch_ms_per_frame": ev_ms / steps, "band_waves_rows": list(tune.TILES[tile]),
           "sweeps_per_launch": tsw, "streaming_model_frac": cells * steps / wall * b_alg(scs[0]["iters"]) / (
               HBM_PEAK_GBS * 1e9)}
    traffic, ent = profiled_traffic("rb_frame_mix", f"{tile}:{tsw}")
    if traffic and len(scs) == 512:
        gbs = traffic / (wall / steps) / 1e9
        res["measured_hbm"] = {"bytes_per_frame": traffic, "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": gbs / HBM_PEAK_GBS,
                               "source": f"{ent['source']} at band:sweeps {tile}:{tsw} (rocprofv3 --pmc FETCH_SIZE, "
                                         "WRITE_SIZE per kernel x launches per frame)"}
    else:
        res["measured_hbm"] = None
    if check_frames > 0:
        # GPU fields after each of the first frames of one scene per size,
        # checked against the f64 oracle in the CPU leg
        got = {}
        for N in sorted({sc["size"] - 2 for sc in scs}):
            sc = [x for x in scs if x["size"] - 2 == N][0]
            g = mk(sc)
            frames = []
            for _ in range(check_frames):
                g.step(1)
                g.sync()
                frames.append({name: g.download(fid) for name, fid in
                               (("density", _abi.DENSITY), ("vx", _abi.VX), ("vy", _abi.VY))})
            got[N] = (sc, frames)
            g.close()
        CHECKS["rbmix"] = (res, check_frames, got)
    return res


def _rel(a, b):
    a = a.astype(np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def check_batch_redblack(res, check_frames, got, threads=16):
    """vs_fp64 of the red-black batch leg: the GPU fields after each of the
    first frames against the f64 red-black oracle, beside the reference's own
    f32 step (lexicographic order) against its f64 version on the same scene
    and frames.  Norm-wise relative error per field and frame."""
    import oracle

    def states(sc):
        out = []
        for dtype, order in ((np.float64, oracle.RB), (np.float32, oracle.LEX), (np.float64, oracle.LEX)):
            st = oracle.State.new(sc["size"]).astype(dtype)
            for r in sc["rects"]:
                st.fill_rect(*r)
            out.append((st, order))
        return out

    def one(N):
        sc, frames = got[N]
        sts = states(sc)
        ours, ref = [], []
        for fields in frames:
            for st, order in sts:
                st.run(sc["dt"], sc["diffusion"], sc["viscosity"], sc["iters"], 1, order=order,
                       src_d=sc["src_d"], src_u=sc["src_u"], src_v=sc["src_v"])
            rb64, lex32, lex64 = (st for st, _ in sts)
            ours.append(max(_rel(fields[k], rb64.f[k]) for k in fields))
            ref.append(max(_rel(lex32.f[k], lex64.f[k]) for k in fields))
        return N, ours, ref

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(min(threads, len(got))) as ex:
        rows = list(ex.map(one, sorted(got)))
    res["vs_fp64"] = {
        "frames": check_frames,
        "metric": "max over density, vx, vy of ||gpu - f64|| / ||f64|| (norm-wise), per frame",
        "per_frame": {f"N={N}": [float(f"{x:.3g}") for x in ours] for N, ours, _ in rows},
        "reference_f32_vs_fp64_per_frame": {f"N={N}": [float(f"{x:.3g}") for x in ref] for N, _, ref in rows},
        "worst_rel_norm": max(max(o) for _, o, _ in rows),
        "worst_reference_f32_rel_norm": max(max(r) for _, _, r in rows),
        "oracle_seconds": time.perf_counter() - t0}


def bench_single_dropin(scs, frames):
    """The plain drop-in: scenes advanced one at a time with eqf_step."""
    hs = make_scenes(scs)
    for h in hs:  # warm-up: one frame each
        h.step(1)
        h.sync()
    device_sync()
    t0 = time.perf_counter()
    for h in hs:
        h.step(frames)
        h.sync()
    wall = time.perf_counter() - t0
    for h in hs:
        h.close()
    return wall


def cpu_single(sc, frames=3):
    """The oracle's exact (lexicographic f32) step on one grid on one core, as
    the reference runs it (one thread, renderer.rs:126-128): one warm-up
    frame, then `frames` frames timed one by one."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from helpers import oracle_state

    st = oracle_state(sc)

    def frame():
        st.run(sc["dt"], sc["diffusion"], sc["viscosity"], sc["iters"], 1,
               src_d=sc["src_d"], src_u=sc["src_u"], src_v=sc["src_v"])

    frame()
    times = []
    for _ in range(frames):
        t0 = time.perf_counter()
        frame()
        times.append(time.perf_counter() - t0)
    N = sc["size"] - 2
    med = sorted(times)[len(times) // 2]
    return {"value": N * N / med, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "median_frame_s": med, "total_s": sum(times), "gs_sweeps": sc["iters"],
            "sample": f"N={N} exact step, 1 warm-up + {frames} timed frames on one core"}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scs, seconds, threads):
    """The CPU oracle (C restatement of the reference step, -O2, no FMA) on the
    same scenes, one scene per thread: one untimed warm-up frame of the mix,
    then whole frames until about `seconds` of wall time have passed (at
    least one frame)."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle  # noqa: F401
    from helpers import oracle_state

    states = [oracle_state(sc) for sc in scs]

    def run(i):
        sc = scs[i]
        states[i].run(sc["dt"], sc["diffusion"], sc["viscosity"], sc["iters"], 1,
                      src_d=sc["src_d"], src_u=sc["src_u"], src_v=sc["src_v"])

    os.environ.setdefault("OMP_NUM_THREADS", "1")
    frames = 0
    with cf.ThreadPoolExecutor(threads) as ex:
        list(ex.map(run, range(len(scs))))  # warm-up frame, untimed
        t0 = time.perf_counter()
        while True:
            list(ex.map(run, range(len(scs))))
            frames += 1
            wall = time.perf_counter() - t0
            if wall * (frames + 1) / frames > seconds:  # the next frame would overrun
                break
    cells = sum((sc["size"] - 2) ** 2 for sc in scs) * frames
    return cells / wall, wall, frames


def bench_redblack_4096(steps, check_frames, N=4096, tiling=None):
    """One N=4096 grid, 40 sweeps, red-black mode with register temporal
    blocking; autotuned band shape (or `tiling` = (band, sweeps per launch),
    e.g. for profiler runs, where the trial frames time differently);
    accuracy vs the f64 red-black oracle on the first frames; pressure
    Poisson residual of the last projection."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from equilibrium_amd import tune

    sc = scenes.redblack_scene(N=N)
    t_setup = time.perf_counter()
    h = _abi.Scene(scenes.scene_params(sc, _abi.GS_REDBLACK))
    scenes.apply_to_scene(h, sc)
    h.sync()
    log(f"rb N={N}: scene ready in {time.perf_counter() - t_setup:.1f} s")
    if tiling:
        tile, tsw = tiling
        log(f"rb N={N}: band {tune.TILES[tile]} x {tsw} sweeps/launch (given)")
    else:
        probe = h.clone()
        tile, tsw, timings = tune.tune_redblack(probe, frames=1)
        probe.close()
        log(f"rb N={N}: autotuned band {tune.TILES[tile]} x {tsw} sweeps/launch: {timings}")
    h.set_tuning(tile, tsw)
    h.step(1)
    h.sync()
    t0 = time.perf_counter()
    h.step(steps)
    h.sync()
    wall = time.perf_counter() - t0
    rate = N * N * steps / wall
    res = {"value": rate, "unit": "cell-updates/s", "ms_per_frame": wall * 1e3 / steps,
           "frames_timed": steps, "band_waves_rows": list(tune.TILES[tile]), "sweeps_per_launch": tsw,
           "gs_sweeps": sc["iters"], "alg_bytes_per_cell_update": b_alg(sc["iters"]),
           # streaming-model rate: k_rb_band keeps its region in registers for
           # several sweeps, so this can exceed 1 and is no HBM fraction
           "streaming_model_frac": rate * b_alg(sc["iters"]) / (HBM_PEAK_GBS * 1e9)}
    shape = f"{tile}:{tsw}"
    traffic, ent = profiled_traffic("rb_frame_n4096", shape)
    if traffic and N == 4096:
        gbs = traffic / (wall / steps) / 1e9
        res["measured_hbm"] = {"bytes_per_frame": traffic, "achieved": gbs, "peak": HBM_PEAK_GBS,
                               "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                               "source": f"{ent['source']} at band:sweeps {shape} (rocprofv3 --pmc FETCH_SIZE, "
                                         "WRITE_SIZE per kernel x launches per frame)"}
    else:
        res["measured_hbm"] = None
        res["measured_hbm_note"] = f"no rocprofv3 profile at band:sweeps {shape}"
    h.close()
    if check_frames > 0:
        # GPU fields after each of the first frames and the last projection's
        # pressure and divergence, checked against the f64 oracle in the CPU leg
        g = _abi.Scene(scenes.scene_params(sc, _abi.GS_REDBLACK))
        scenes.apply_to_scene(g, sc)
        g.set_tuning(tile, tsw)
        frames = []
        for f in range(check_frames):
            g.step(1)
            g.sync()
            frames.append({name: g.download(fid) for name, fid in
                           (("density", _abi.DENSITY), ("vx", _abi.VX), ("vy", _abi.VY))})
        p_gpu, div_gpu = g.download(_abi.VX0), g.download(_abi.VY0)
        g.close()
        CHECKS["rb4096"] = (res, sc, frames, p_gpu, div_gpu)
    return res


def check_redblack_4096(res, sc, frames, p_gpu, div_gpu):
    """vs_fp64 and poisson_residual of the N=4096 leg: per-frame max-abs and
    relative error of density and velocity against the f64 red-black oracle,
    and the pressure Poisson residual of the GPU's last projection beside the
    f64 oracle's."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle
    from helpers import oracle_state

    N = sc["size"] - 2
    st = oracle_state(sc).astype(np.float64)
    srcd = np.ascontiguousarray(sc["src_d"], np.float64)
    errs = []
    t_or = time.perf_counter()
    for f, fields in enumerate(frames):
        st.f["density"] += srcd
        st.f["scratch"] += srcd
        st.step_par_f64(sc["dt"], sc["diffusion"], sc["viscosity"], sc["iters"])
        e = {}
        for name, got in fields.items():
            got = got.astype(np.float64)
            want = st.f[name]
            e[name] = {"max_abs": float(f"{np.max(np.abs(got - want)):.3g}"),
                       "rel_norm": float(f"{np.linalg.norm(got - want) / max(np.linalg.norm(want), 1e-300):.3g}")}
        errs.append(e)
        frames[f] = None  # free the frame's fields
        log(f"rb N={N} frame {f}: {e}")
    mx, rr, rd = oracle.poisson_residual(p_gpu.astype(np.float64), div_gpu.astype(np.float64), N + 2)
    mx64, rr64, rd64 = oracle.poisson_residual(st.f["vx0"], st.f["vy0"], N + 2)
    res["vs_fp64"] = {"frames": len(frames), "per_frame": errs,
                      "worst_rel_norm": max(max(x["rel_norm"] for x in e.values()) for e in errs),
                      "oracle_seconds": time.perf_counter() - t_or}
    res["poisson_residual"] = {"gpu_max_abs": mx, "gpu_rms": rr, "div_rms": rd,
                               "gpu_rms_over_div_rms": rr / max(rd, 1e-300),
                               "fp64_rms_over_div_rms": rr64 / max(rd64, 1e-300)}


def bench_snapshots(frames, depth=2, N=4096, tiling=None):
    """Per-frame hand-over of the state to a consumer on the host (the
    reference clones the Fluid into the renderer's channel after every step):
    the N=4096 red-black scene advanced frame by frame (a) with no readback,
    (b) with a synchronous density download per frame, (c) through a snapshot
    ring of `depth` pinned slots carrying density and the rendered RGBA image,
    whose copies run behind the following steps."""
    sc = scenes.redblack_scene(N=N)
    h = _abi.Scene(scenes.scene_params(sc, _abi.GS_REDBLACK))
    scenes.apply_to_scene(h, sc)
    if tiling:
        h.set_tuning(*tiling)
    h.step(1)
    h.sync()

    def timed(body):
        t0 = time.perf_counter()
        body()
        h.sync()
        return (time.perf_counter() - t0) * 1e3 / frames

    def plain():
        for _ in range(frames):
            h.step(1)

    def sync_download():
        for _ in range(frames):
            h.step(1)
            h.sync()
            h.download(_abi.DENSITY)

    snap = _abi.Snapshot(h, depth, with_rgba=True)
    handed = []

    def ring():
        for f in range(frames):
            if f >= depth:
                fr, d, img = snap.acquire(copy=False)
                handed.append(fr)
                snap.release()
            h.step(1)
            snap.capture(f)
        for _ in range(min(depth, frames)):
            fr, d, img = snap.acquire(copy=False)
            handed.append(fr)
            snap.release()

    ms = {"no_readback": timed(plain), "sync_download_density": timed(sync_download),
          "snapshot_ring_density_rgba": timed(ring)}
    snap.close()
    h.close()
    assert handed == list(range(frames)), handed
    mb = (N + 2) ** 2 * 8 / 1e6
    return {"unit": "ms/frame", "frames_timed": frames, "depth": depth,
            "bytes_per_snapshot_mb": mb, **ms,
            "ring_overhead_frac": ms["snapshot_ring_density_rgba"] / ms["no_readback"] - 1.0,
            "sync_overhead_frac": ms["sync_download_density"] / ms["no_readback"] - 1.0}


WAVEFRONT_MAX_N = 1024  # exact mode holds up to 16 strips of 64 rows per workgroup


def bench_serving(rates, seconds, budget_gb=200.0, n_max=2048, buckets=16):
    """Poisson arrivals of scene requests (N log-uniform in [64, n_max],
    frames U(50,500), sources and obstacles as in small_scene), admitted
    against an HBM budget; each rate runs `seconds` of arrivals, then the
    remaining requests are cancelled.  Scenes up to N=1024 run in exact
    wavefront mode (one fused launch per tick), larger ones in red-black mode
    (launch groups beside it).  p50/p99 frame latency = tick durations
    weighted by resident scenes.  To keep the client on the Poisson schedule,
    N is rounded to one of `buckets` log-spaced sizes whose scenes (sources,
    obstacles) are generated before the clock starts and lent to the server
    (read in place at admission)."""
    import math
    sizes = sorted({int(round(64 * (n_max / 64) ** (k / (buckets - 1)))) for k in range(buckets)})
    pool = {N: scenes.stam_scene(N, scenes.derive_seed(40, N), n_rects=1) for N in sizes}
    logs = [math.log(N) for N in sizes]
    out = []
    for rate in rates:
        reqs = scenes.serving_requests(rate, seconds, n_max=n_max)
        srv = _abi.Server(int(budget_gb * (1 << 30)))
        t0 = time.perf_counter()
        tickets = []
        lag = 0.0
        for (ta, N, frames, seed) in reqs:
            while time.perf_counter() - t0 < ta:
                time.sleep(0.0002)
            lag = max(lag, time.perf_counter() - t0 - ta)
            nb = sizes[min(range(len(sizes)), key=lambda k: abs(logs[k] - math.log(N)))]
            mode = _abi.GS_WAVEFRONT if nb <= WAVEFRONT_MAX_N else _abi.GS_REDBLACK
            tickets.append(srv.submit(pool[nb], nframes=frames, gs_mode=mode, borrow=True))
        t_end = time.perf_counter() - t0
        time.sleep(0.5)
        st = srv.stats()
        classes = {}
        for name, lo, hi in (("N<=64", 0, 66), ("N<=128", 67, 130), ("N<=256", 131, 258), ("N<=512", 259, 514),
                             ("N<=1024", 515, 1026), ("N>1024 (red-black)", 1027, 1 << 31)):
            c = srv.stats(lo, hi)
            classes[name] = {"p50_frame_ms": round(c["p50_frame_ms"], 3), "p99_frame_ms": round(c["p99_frame_ms"], 3),
                             "p50_queue_ms": round(c["p50_queue_ms"], 3), "p99_queue_ms": round(c["p99_queue_ms"], 3),
                             "scene_frames": c["scene_frames"]}
        for t in tickets:
            srv.release(t)
        srv.wait_idle(120000)
        r = {"rate_per_s": rate, "requests": len(reqs), "window_s": t_end,
             "client_max_lag_s": lag,
             "cell_updates_per_s": st["cell_updates"] / max(st["busy_ms"] * 1e-3, 1e-9),
             "p50_frame_ms": st["p50_frame_ms"], "p99_frame_ms": st["p99_frame_ms"],
             "p50_queue_ms": st["p50_queue_ms"], "p99_queue_ms": st["p99_queue_ms"],
             "scene_frames": st["scene_frames"], "ticks": st["ticks"],
             "max_resident": st["max_resident"], "peak_gb": st["peak_bytes"] / 2 ** 30,
             "red_black_share": sum(1 for r in reqs if r[1] > WAVEFRONT_MAX_N) / max(1, len(reqs)),
             "by_size": classes}
        log(f"serve rate {rate}/s: {r}")
        srv.close()
        out.append(r)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scenes", type=int, default=512)
    ap.add_argument("--no-extras", action="store_true", help="headline batch only (for profiling)")
    ap.add_argument("--legs", type=str, default="bw,batch,rbmix,single,n1024,rb4096,snap,serve,cpu",
                    help="comma list of legs to run")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a 1 GiB device copy first (HBM counter calibration for rocprofv3 --pmc)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--rb-frames", type=int, default=10)
    ap.add_argument("--rb-check-frames", type=int, default=20)
    ap.add_argument("--batch-tuning", type=str, default="",
                    help="max_waves:order[:schedule] for the batch instead of autotuning (e.g. 16:0:2; "
                         "schedule 0 auto, 1 wide, 2 ring, 3 fine)")
    ap.add_argument("--rb-tiling", type=str, default="",
                    help="band:sweeps for the N=4096 red-black grid instead of autotuning (e.g. 0:5)")
    ap.add_argument("--rb-batch-band", type=str, default="",
                    help="band:sweeps for the red-black batched mix instead of autotuning (e.g. 3:6)")
    ap.add_argument("--snap-frames", type=int, default=20)
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--serve-rates", type=str, default="50,100,200,500,1000,2000")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    legs = set(args.legs.split(","))
    if args.no_extras:
        legs = {"batch"}
    if args.calibrate_copy:
        import torch
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("calibration copy: 3 x 1 GiB read + 1 GiB write (torch copy kernel)")
    mix = scenes.batched_mix(n=args.scenes, frames=args.steps)
    out = {"metric": METRIC, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded scenes generated on the host)"}
    value = None
    bw = None
    if "bw" in legs:
        bw = bench_bandwidth()
        out["hbm_ceiling"] = bw
        log(f"HBM ceiling (streaming probe): {bw}")
    if "batch" in legs:
        iters = mix[0]["iters"]
        cells = sum((sc["size"] - 2) ** 2 for sc in mix)
        log(f"batched mix: {len(mix)} scenes, {cells} interior cells, {iters} sweeps")
        given = tuple(int(x) for x in args.batch_tuning.split(":")) if args.batch_tuning else None
        wall, ev_ms, tuning = bench_batch(mix, args.steps, args.warmup, given)
        value = cells * args.steps / wall
        ms_per_step = wall * 1e3 / args.steps
        launch_ms = ev_ms / args.steps  # HIP events on the batch stream around the K launches
        alg_bytes = cells * b_alg(iters)
        achieved = alg_bytes / (launch_ms * 1e-3) / 1e9
        log(f"batch: {value:.4g} cell-updates/s, {ms_per_step:.3f} ms/frame, kernel {launch_ms:.3f} ms")

        out.update({
            "metric": METRIC,
            "value": value,
            "unit": "cell-updates/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": ms_per_step,
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f32",
            "data": "synthetic (seeded scenes generated on the host)",
            "config": {
                "workload": f"batched mix: {len(mix)} scenes, N uniform in {{64,128,256,512}}, "
                            f"visc/diff log-U[1e-6,1e-3], 0-4 rectangles, {iters} GS sweeps, dt 0.1, "
                            "exact wavefront mode, one fused launch per frame (HIP graph)",
                "scenes": len(mix),
                "interior_cells": cells,
                "gs_sweeps": iters,
                "mode": "wavefront, bit-exact: the CPU leg checks one timed scene per size against the "
                        "oracle after all frames (config.bit_exact_check)",
                "batch_tuning": tuning,
            },
            "roofline": {
                "bound": "hbm",
                "kernel": "eqf_frames_cyclic (fused frame step, one workgroup per scene)",
                "frac_basis": "algorithmic bytes (B_alg per cell-update) / launch time; traffic_frac: "
                              "measured HBM bytes (rocprofv3) / launch time",
                "achieved": achieved,
                "peak": HBM_PEAK_GBS,
                "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS,
                "alg_bytes_per_launch": alg_bytes,
                "launch_ms": launch_ms,
            },
        })
        shape = tuning["shape"]
        traffic, ent = profiled_traffic("eqf_frames_cyclic", shape)
        out["roofline"]["traffic"] = traffic
        out["roofline"]["traffic_frac"] = (traffic / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS) if traffic else None
        out["roofline"]["traffic_source"] = (
            f"{ent['source']} at launch shape {shape} (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE, scaled by a copy "
            "of known size)" if ent else f"no rocprofv3 profile at launch shape {shape}")
        if bw:
            out["roofline"]["practical_peak"] = bw["copy"]
            out["roofline"]["frac_of_practical"] = achieved / bw["copy"]

    if "rbmix" in legs:
        # the batched mix in red-black mode (<= 1e-5 vs FP64 instead of bit-exact)
        band = tuple(int(x) for x in args.rb_batch_band.split(":")) if args.rb_batch_band else None
        out["batched_mix_redblack"] = bench_batch_redblack(mix, args.steps, args.warmup, args.rb_check_frames, band)
        log(f"red-black batch: {out['batched_mix_redblack']}")

    if "single" in legs:
        # single-grid drop-in on a stratified sample of the same mix
        sample = mix[::8]
        sframes = 2
        swall = bench_single_dropin(sample, sframes)
        scells = sum((sc["size"] - 2) ** 2 for sc in sample) * sframes
        single = scells / swall
        out["single_grid_dropin"] = {"value": single, "unit": "cell-updates/s",
                                     "sample": f"{len(sample)} scenes of the mix x {sframes} frames, "
                                               "one eqf_step per scene"}
        if value:
            out["batched_over_single"] = value / single
        log(f"single drop-in: {single:.4g} cell-updates/s")

    if "n1024" in legs:
        # N=128 single grid (the reference's own CPU-runnable case), exact mode, all 100 frames
        c0 = scenes.small_scene()
        h = make_scenes([c0])[0]
        h.step(1)
        h.sync()
        t0 = time.perf_counter()
        h.step(c0["frames"])
        h.sync()
        w0 = time.perf_counter() - t0
        out["n128_single_wavefront"] = {"value": 128 * 128 * c0["frames"] / w0, "unit": "cell-updates/s",
                                        "ms_per_frame": w0 * 1e3 / c0["frames"], "frames_timed": c0["frames"]}
        h.close()
        log(f"N=128 single: {out['n128_single_wavefront']}")
        # N=1024 single grid, exact mode
        c1 = scenes.single_1024()
        h = make_scenes([c1])[0]
        h.step(1)
        h.sync()
        nf = 5
        t0 = time.perf_counter()
        h.step(nf)
        h.sync()
        w1 = time.perf_counter() - t0
        out["n1024_single_wavefront"] = {"value": 1024 * 1024 * nf / w1, "unit": "cell-updates/s",
                                         "ms_per_frame": w1 * 1e3 / nf}
        h.close()
        log(f"N=1024 single: {1024 * 1024 * nf / w1:.4g} cell-updates/s")

    if "rb4096" in legs:
        # N=4096 single grid, red-black mode
        tiling = tuple(int(x) for x in args.rb_tiling.split(":")) if args.rb_tiling else None
        out["n4096_single_redblack"] = bench_redblack_4096(args.rb_frames, args.rb_check_frames, tiling=tiling)
        log(f"N=4096 red-black: {out['n4096_single_redblack']['value']:.4g} cell-updates/s")

    if "snap" in legs:
        # per-frame snapshots of the N=4096 red-black scene (density + RGBA)
        rb = out.get("n4096_single_redblack")
        tiling = None
        if rb:
            from equilibrium_amd import tune
            tiling = ([k for k, v in tune.TILES.items() if list(v) == rb["band_waves_rows"]][0],
                      rb["sweeps_per_launch"])
        out["n4096_frame_snapshots"] = bench_snapshots(args.snap_frames, tiling=tiling)
        log(f"N=4096 snapshots: {out['n4096_frame_snapshots']}")

    if "serve" in legs:
        # serving with Poisson arrivals
        rates = [float(x) for x in args.serve_rates.split(",") if x]
        out["serving"] = bench_serving(rates, args.serve_seconds)

    if "cpu" in legs:
        # The CPU leg: the oracle's accuracy checks of the red-black legs, then
        # the CPU baseline (oracle on this host, bounded sample)
        if "rbmix" in CHECKS:
            check_batch_redblack(*CHECKS["rbmix"])
            log(f"red-black batch vs fp64: {out['batched_mix_redblack']['vs_fp64']}")
        if "rb4096" in CHECKS:
            check_redblack_4096(*CHECKS["rb4096"])
        threads = len(os.sched_getaffinity(0))  # every core this process may run on
        if "batch" in CHECKS:
            out["config"]["bit_exact_check"] = check_batch_exact(*CHECKS["batch"], threads)
            log(f"batch bit-exact check: {out['config']['bit_exact_check']}")
        rate, wcpu, cframes = cpu_baseline(mix, args.cpu_seconds, threads)
        out["cpu_baseline"] = {"value": rate, "unit": "cell-updates/s", "cores": threads,
                               "kind": "port", "cpu_model": cpu_model(), "nproc": os.cpu_count(),
                               "cores_note": "all cores in this process's CPU affinity set",
                               "sample": f"all {len(mix)} scenes of the mix x {cframes} frame(s) after one untimed "
                                         f"warm-up frame, one scene per thread, {wcpu:.1f} s"}
        log(f"cpu oracle: {rate:.4g} cell-updates/s on {threads} threads ({wcpu:.1f} s)")
        out["cpu_single_n1024"] = cpu_single(scenes.single_1024())
        log(f"cpu oracle N=1024 on one core: {out['cpu_single_n1024']}")
        out["cpu_single_n128"] = cpu_single(scenes.small_scene(), frames=10)
        log(f"cpu oracle N=128 on one core: {out['cpu_single_n128']}")

    if "cpu4096" in legs:
        # The N=4096 grid on one core, as the reference would run it: ~30 s per
        # frame, so opt-in (not in the default legs); GPU frame time beside it
        # from the rb4096 leg when that ran
        c = cpu_single(scenes.redblack_scene(), frames=3)
        if "n4096_single_redblack" in out:
            c["gpu_over_cpu"] = out["n4096_single_redblack"]["value"] / c["value"]
        out["cpu_single_n4096"] = c
        log(f"cpu oracle N=4096 on one core: {c}")

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
