"""GPU parity at the benchmark workloads themselves, through the C ABI.

* N=1024 exact scene (20 sweeps, 3 rectangles, 4 sources): all 200 frames
  bit-exact.  The GPU trajectory is cut into segments of 20 frames; the
  oracle runs each segment from the GPU's state at its start (segments in
  parallel threads) and must land on the GPU's state at its end, bit for bit.
  The first segment starts from the oracle's own initial state, so by
  induction the whole 200-frame GPU run equals the oracle's.
* Batched mix (N in {64,128,256,512}): a stratified sample of 64 scenes of
  the 512-scene mix (16 per size) in one exact batch, at both launch shapes
  the tuner chooses between for a mix holding N=512 scenes (8 and 16 waves
  per workgroup), 2 frames, bit-exact.
* N=4096 red-black scene (40 sweeps, 16 rectangles, curl-noise velocity): one
  frame bit-exact against the red-black f32 oracle, then three frames against
  the f64 red-black oracle (norm-wise relative error of density and velocity
  at most 1e-5 per frame) and the pressure Poisson residual of the last
  projection beside the f64 oracle's.
* Red-black accuracy on the mix: the first scene of each size of the
  512-scene mix in red-black mode, 4 frames, each frame's norm-wise relative
  error against the f64 red-black oracle at most max(1e-5, 2x the error of
  the reference's own f32 step against its f64 version on the same scene).
  The scenes are chaotic: both errors grow ~5x per frame and reach 0.1-0.6
  by frame 20, so the bar is per frame and relative to the reference's f32.
* Serving: one second of Poisson arrivals at 200 requests/s with N
  log-uniform in [64, 2048] (exact mode up to N=1024, red-black above), frames
  capped at 6; every result equals its solo run bit for bit.
"""
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from equilibrium_amd import _abi, scenes
from helpers import FIELD_NAMES, assert_bit_equal, assert_state_equal, download_all, gpu_scene, oracle_state

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def _case(label):
    print(f"[case] {label}", flush=True)


def _state_from(fields, mask, S):
    st = oracle.State(S)
    for k in FIELD_NAMES:
        st.f[k] = np.array(fields[k], np.float32)
    st.mask = mask.copy()
    return st


def test_single_1024_all_200_frames_bit_exact():
    sc = scenes.single_1024()
    S, K, segs = sc["size"], 20, 10
    assert K * segs == sc["frames"]
    _case(f"N=1024 exact, {segs} segments x {K} frames")
    s = gpu_scene(sc)
    s.sync()
    ckpt = [download_all(s)]
    st0 = oracle_state(sc)
    assert_state_equal(ckpt[0], st0, "N=1024 initial state")
    for _ in range(segs):
        s.step(K)
        s.sync()
        ckpt.append(download_all(s))
    s.close()
    mask = st0.mask

    def segment(k):
        st = _state_from(ckpt[k], mask, S)
        st.run(sc["dt"], sc["diffusion"], sc["viscosity"], sc["iters"], K,
               src_d=sc["src_d"], src_u=sc["src_u"], src_v=sc["src_v"])
        for name in FIELD_NAMES:
            g = ckpt[k + 1][name].view(np.uint32)
            w = st.f[name].view(np.uint32)
            if not np.array_equal(g, w):
                return f"frames {k * K}..{(k + 1) * K}: field {name}: {int(np.sum(g != w))} cells differ"
        return None

    with ThreadPoolExecutor(max_workers=min(THREADS, segs)) as ex:
        errs = [e for e in ex.map(segment, range(segs)) if e]
    assert not errs, errs


def _mix_sample(per_size=16, frames=2):
    mix = scenes.batched_mix(frames=frames)
    out = []
    for N in (64, 128, 256, 512):
        out += [sc for sc in mix if sc["size"] == N + 2][:per_size]
    return out


@pytest.mark.parametrize("max_waves,order,sched", [(8, _abi.BATCH_LARGEST_FIRST, _abi.SCHED_WIDE),
                                                   (16, _abi.BATCH_GIVEN_ORDER, _abi.SCHED_WIDE),
                                                   (16, _abi.BATCH_LARGEST_FIRST, _abi.SCHED_RING),
                                                   (16, _abi.BATCH_LARGEST_FIRST, _abi.SCHED_FINE)])
def test_batched_mix_sample_bit_exact(max_waves, order, sched):
    sample = _mix_sample()
    assert len(sample) == 64 and sum(sc["size"] == 514 for sc in sample) == 16
    _case(f"batched mix sample: 64 scenes, max_waves={max_waves} order={order} schedule={sched}, 2 frames")
    hs = [gpu_scene(sc) for sc in sample]
    b = _abi.Batch(hs)
    b.set_tuning(max_waves, order)
    b.set_schedule(sched)
    b.step(2)
    b.sync()
    got = [download_all(h) for h in hs]
    b.close()
    for h in hs:
        h.close()

    def check(i):
        sc = sample[i]
        st = oracle_state(sc)
        st.run(sc["dt"], sc["diffusion"], sc["viscosity"], sc["iters"], 2,
               src_d=sc["src_d"], src_u=sc["src_u"], src_v=sc["src_v"])
        for name in FIELD_NAMES:
            g, w = got[i][name].view(np.uint32), st.f[name].view(np.uint32)
            if not np.array_equal(g, w):
                return f"scene {i} N={sc['size'] - 2}: field {name}: {int(np.sum(g != w))} cells differ"
        return None

    with ThreadPoolExecutor(max_workers=THREADS) as ex:
        errs = [e for e in ex.map(check, range(len(sample))) if e]
    assert not errs, errs


def _rel(got, want):
    got = got.astype(np.float64)
    return float(np.linalg.norm(got - want) / max(np.linalg.norm(want), 1e-300))


def test_redblack_4096_bit_exact_and_fp64_error():
    sc = scenes.redblack_scene()
    S = sc["size"]
    it = sc["iters"]
    dt, diff, visc = np.float32(sc["dt"]), np.float32(sc["diffusion"]), np.float32(sc["viscosity"])
    _case("N=4096 red-black: 1 frame vs f32 red-black oracle")
    g = gpu_scene(sc, _abi.GS_REDBLACK)
    g.set_tuning(5, 5)   # band 8 waves x 10 rows, 5 sweeps per launch (the default)
    g.step(1)
    g.sync()
    got1 = download_all(g)
    st32 = oracle_state(sc)
    st32.f["density"] += sc["src_d"]
    st32.f["scratch"] += sc["src_d"]
    st32.step_par(dt, diff, visc, it)
    assert_state_equal(got1, st32, "N=4096 red-black frame 1")
    del st32

    _case("N=4096 red-black: 3 frames vs f64 red-black oracle")
    st64 = oracle_state(sc).astype(np.float64)
    srcd = sc["src_d"].astype(np.float64)
    worst = 0.0
    for f in range(3):
        if f > 0:
            g.step(1)
            g.sync()
        st64.f["density"] += srcd
        st64.f["scratch"] += srcd
        st64.step_par(float(dt), float(diff), float(visc), it)
        errs = {k: _rel(g.download(fid), st64.f[k])
                for k, fid in (("density", _abi.DENSITY), ("vx", _abi.VX), ("vy", _abi.VY))}
        print(f"[fp64] frame {f}: {errs}", flush=True)
        worst = max(worst, *errs.values())
        assert max(errs.values()) <= 1e-5, (f, errs)
    p, div = g.download(_abi.VX0), g.download(_abi.VY0)
    g.close()
    mx, rr, rd = oracle.poisson_residual(p.astype(np.float64), div.astype(np.float64), S)
    mx64, rr64, rd64 = oracle.poisson_residual(st64.f["vx0"], st64.f["vy0"], S)
    print(f"[poisson] gpu rms/div rms {rr / rd:.6f}, f64 {rr64 / rd64:.6f}, worst rel {worst:.3g}", flush=True)
    assert abs(rr / rd - rr64 / rd64) <= 1e-3 * (rr64 / rd64)


def test_redblack_mix_accuracy_against_fp64_beside_reference_f32():
    mix = scenes.batched_mix(frames=4)
    picks = [[sc for sc in mix if sc["size"] - 2 == N][0] for N in (64, 128, 256, 512)]
    nf = 4
    _case(f"red-black mix accuracy: N in 64..512, {nf} frames vs f64 beside the reference's f32")
    got = []
    for sc in picks:
        h = gpu_scene(sc, _abi.GS_REDBLACK)
        frames = []
        for _ in range(nf):
            h.step(1)
            h.sync()
            frames.append({k: h.download(fid) for k, fid in (("density", _abi.DENSITY), ("vx", _abi.VX),
                                                             ("vy", _abi.VY))})
        h.close()
        got.append(frames)

    def one(i):
        sc = picks[i]
        sts = []
        for dtype, order in ((np.float64, oracle.RB), (np.float32, oracle.LEX), (np.float64, oracle.LEX)):
            st = oracle_state(sc).astype(dtype)
            sts.append((st, order))
        rows = []
        for fields in got[i]:
            for st, order in sts:
                st.run(sc["dt"], sc["diffusion"], sc["viscosity"], sc["iters"], 1, order=order,
                       src_d=sc["src_d"], src_u=sc["src_u"], src_v=sc["src_v"])
            rb64, lex32, lex64 = (st for st, _ in sts)
            ours = max(_rel(fields[k], rb64.f[k]) for k in fields)
            ref = max(_rel(lex32.f[k], lex64.f[k]) for k in fields)
            rows.append((ours, ref))
        return rows

    with ThreadPoolExecutor(max_workers=len(picks)) as ex:
        res = list(ex.map(one, range(len(picks))))
    bad = []
    for sc, rows in zip(picks, res):
        N = sc["size"] - 2
        print(f"[fp64] N={N}: ours {[f'{o:.3g}' for o, _ in rows]} reference f32 {[f'{r:.3g}' for _, r in rows]}",
              flush=True)
        bad += [(N, f, o, r) for f, (o, r) in enumerate(rows) if o > max(1e-5, 2.0 * r)]
    assert not bad, bad


def test_serving_poisson_arrivals_equal_solo_runs():
    reqs = scenes.serving_requests(200, 1.0, n_max=2048)
    assert len(reqs) > 100
    cap = 6
    _case(f"serving: {len(reqs)} Poisson arrivals at 200/s, N up to 2048, frames <= {cap}")
    scs = []
    for (_, N, frames, seed) in reqs:
        sc = scenes.stam_scene(N, seed, frames=min(frames, cap), n_rects=1)
        mode = _abi.GS_WAVEFRONT if N <= 1024 else _abi.GS_REDBLACK
        scs.append((sc, mode))
    assert any(m == _abi.GS_REDBLACK for _, m in scs) and max(sc["size"] for sc, _ in scs) > 1500
    srv = _abi.Server(hbm_budget_bytes=200 << 30)
    t0 = time.perf_counter()
    tickets = []
    for (ta, _, _, _), (sc, mode) in zip(reqs, scs):
        while time.perf_counter() - t0 < ta:
            time.sleep(0.0002)
        tickets.append(srv.submit(sc, gs_mode=mode, borrow=True))
    srv.wait_idle(120000)
    st = srv.stats()
    print(f"[serving] {st}", flush=True)
    assert st["max_resident"] > 1
    assert st["scene_frames"] == sum(sc["frames"] for sc, _ in scs)
    for (sc, mode), t in zip(scs, tickets):
        assert srv.poll(t).state == _abi.REQ_DONE
        h = gpu_scene(sc, mode)
        h.step(sc["frames"])
        h.sync()
        assert_bit_equal(srv.result(t), h.download(_abi.DENSITY), f"N={sc['size'] - 2} mode={mode}")
        h.close()
    srv.close()
