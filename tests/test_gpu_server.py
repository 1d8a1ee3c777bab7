"""Serving loop on the GPU: continuous batching with an HBM budget.

Each result must be the same alone and in the mix (bit-exact, it is the exact
wavefront mode), requests that arrive while others are resident are admitted
between ticks, and a cancelled scene does not disturb the others."""
import time

import numpy as np
import pytest

from equilibrium_amd import _abi, scenes
from helpers import assert_bit_equal, gpu_scene, run_oracle

pytestmark = pytest.mark.gpu


def _case(label):
    print(f"[case] {label}", flush=True)


def _solo(sc):
    s = gpu_scene(sc)
    s.step(sc["frames"])
    s.sync()
    d = s.download(_abi.DENSITY)
    u = s.download(_abi.VX)
    s.close()
    return d, u


def _mix():
    out = []
    for i, (N, fr) in enumerate([(64, 7), (200, 3), (30, 9), (128, 5), (96, 4), (150, 6), (8, 5), (70, 8)]):
        out.append(scenes.stam_scene(N, scenes.derive_seed(30, i), iters=4, frames=fr, n_rects=i % 4,
                                     viscosity=10 ** (-3 - i % 3), diffusion=10 ** (-5 + i % 2)))
    return out


def test_server_results_equal_solo_runs():
    _case("server mix")
    mix = _mix()
    srv = _abi.Server(hbm_budget_bytes=8 << 30)
    tickets = []
    for i, sc in enumerate(mix):
        tickets.append(srv.submit(sc, keep_velocity=True))
        if i % 3 == 2:
            time.sleep(0.002)  # later arrivals join a running batch
    srv.wait_idle(60000)
    st = srv.stats()
    for sc, t in zip(mix, tickets):
        s = srv.poll(t)
        assert s.state == _abi.REQ_DONE and s.frames_done == sc["frames"]
        d, u = _solo(sc)
        assert_bit_equal(srv.result(t, _abi.DENSITY), d, f"density N={sc['size'] - 2}")
        assert_bit_equal(srv.result(t, _abi.VX), u, f"vx N={sc['size'] - 2}")
    o = run_oracle(mix[1])
    assert_bit_equal(srv.result(tickets[1]), o.f["density"], "vs oracle")
    assert st["scene_frames"] == sum(sc["frames"] for sc in mix)
    assert st["p99_frame_ms"] >= st["p50_frame_ms"] > 0
    srv.close()


def test_server_budget_limits_residency():
    _case("server budget")
    mix = _mix()
    # room for roughly one mid-size scene at a time
    srv = _abi.Server(hbm_budget_bytes=3 << 20)
    tickets = [srv.submit(sc) for sc in mix if sc["size"] <= 132]
    small = [sc for sc in mix if sc["size"] <= 132]
    srv.wait_idle(60000)
    st = srv.stats()
    assert st["peak_bytes"] <= 3 << 20
    for sc, t in zip(small, tickets):
        assert srv.poll(t).state == _abi.REQ_DONE
        assert_bit_equal(srv.result(t), _solo(sc)[0], f"budget N={sc['size'] - 2}")
    srv.close()


def test_server_rejects_oversized_and_cancels():
    _case("server cancel")
    mix = _mix()
    srv = _abi.Server(hbm_budget_bytes=1 << 20)
    big = scenes.stam_scene(600, scenes.derive_seed(31), iters=2, frames=2)
    with pytest.raises(_abi.EqfError):
        srv.submit(big)
    srv.close()
    srv = _abi.Server(hbm_budget_bytes=8 << 30)
    long_sc = scenes.stam_scene(128, scenes.derive_seed(32), iters=4, frames=10 ** 6)
    t_long = srv.submit(long_sc)
    ts = [srv.submit(sc) for sc in mix[:4]]
    time.sleep(0.2)
    srv.release(t_long)  # cancels the resident request and forgets its ticket once retired
    t_queued = srv.submit(long_sc)
    srv.release(t_queued)
    srv.wait_idle(60000)
    for t in (t_long, t_queued):
        with pytest.raises(_abi.EqfError) as ei:
            srv.poll(t)
        assert ei.value.code == _abi.EQF_ENOTFOUND
    for sc, t in zip(mix[:4], ts):
        assert_bit_equal(srv.result(t), _solo(sc)[0], f"cancel N={sc['size'] - 2}")
        srv.release(t)
        with pytest.raises(_abi.EqfError):
            srv.poll(t)
    srv.close()


def test_server_red_black_requests_beside_wavefront():
    """Red-black requests (one past the wavefront's size limit) are served
    beside the fused launch, as launch groups (one per pair of iteration
    counts, members leaving as they finish); each result equals its solo
    red-black run bit for bit, and a small one equals the red-black oracle."""
    import oracle

    _case("server red-black mix")
    wf = _mix()[:4]
    rbs = [scenes.stam_scene(N, scenes.derive_seed(31, N), iters=it, frames=fr, n_rects=2)
           for N, fr, it in ((150, 4, 6), (1100, 2, 6), (300, 3, 6), (64, 5, 9), (200, 2, 9))]
    srv = _abi.Server(hbm_budget_bytes=8 << 30)
    tw = [srv.submit(sc) for sc in wf]
    tr = [srv.submit(sc, gs_mode=_abi.GS_REDBLACK) for sc in rbs]
    srv.wait_idle(120000)
    for sc, t in zip(wf, tw):
        assert srv.poll(t).state == _abi.REQ_DONE
        assert_bit_equal(srv.result(t), _solo(sc)[0], f"wavefront N={sc['size'] - 2}")
    for sc, t in zip(rbs, tr):
        st = srv.poll(t)
        assert st.state == _abi.REQ_DONE and st.frames_done == sc["frames"]
        s = gpu_scene(sc, _abi.GS_REDBLACK)
        s.step(sc["frames"])
        s.sync()
        assert_bit_equal(srv.result(t), s.download(_abi.DENSITY), f"red-black N={sc['size'] - 2}")
        s.close()
    assert_bit_equal(srv.result(tr[0]), run_oracle(rbs[0], order=oracle.RB).f["density"], "red-black vs oracle")
    assert srv.stats()["scene_frames"] == sum(sc["frames"] for sc in wf + rbs)
    srv.close()


def test_server_large_lane_runs_smaller_scenes_several_frames_per_tick():
    """Wavefront scenes of 5-16 strips share a lane whose tick lasts one frame
    of its longest frame chain (N=1000 here); the N=300 and N=520 scenes run
    several frames per tick, never past their request (frame counts that are
    not multiples of the frames per tick).  Every result equals its solo run
    bit for bit, and their frame latency is below the N=1000 scene's."""
    _case("server large lane, several frames per tick")
    specs = [(1000, 3), (300, 7), (520, 13), (300, 2)]
    mix = [scenes.stam_scene(N, scenes.derive_seed(35, i), iters=6, frames=fr, n_rects=1)
           for i, (N, fr) in enumerate(specs)]
    srv = _abi.Server(hbm_budget_bytes=16 << 30)
    tickets = [srv.submit(sc) for sc in mix]
    srv.wait_idle(120000)
    st = srv.stats()
    assert st["scene_frames"] == sum(fr for _, fr in specs)
    for sc, t in zip(mix, tickets):
        s = srv.poll(t)
        assert s.state == _abi.REQ_DONE and s.frames_done == sc["frames"]
        d, _ = _solo(sc)
        assert_bit_equal(srv.result(t, _abi.DENSITY), d, f"density N={sc['size'] - 2}")
    small = srv.stats(min_size=302, max_size=522)
    big = srv.stats(min_size=1002, max_size=1002)
    print(f"[serving] p50 frame N=300/520 {small['p50_frame_ms']:.2f} ms, N=1000 {big['p50_frame_ms']:.2f} ms",
          flush=True)
    assert 0 < small["p50_frame_ms"] < big["p50_frame_ms"]
    srv.close()


def test_server_churn_results_equal_solo_runs():
    """Rounds of submissions while earlier requests are resident: finished
    tickets are released (their records freed) and new requests take their
    place, in both modes and with one red-black iteration count changing
    between rounds, so the fused launch's scene list and the red-black launch
    groups change membership every few ticks.  Every result equals its solo
    run bit for bit."""
    _case("server churn")
    rng = np.random.default_rng(33)
    srv = _abi.Server(hbm_budget_bytes=8 << 30)
    pending = {}
    k = 0
    for rnd in range(4):
        for _ in range(12):
            N = int(rng.integers(8, 220))
            rb = bool(rng.random() < 0.35)
            sc = scenes.stam_scene(N, scenes.derive_seed(34, k), iters=4 + (rnd % 2 if rb else 0),
                                   frames=int(rng.integers(1, 7)), n_rects=k % 3)
            mode = _abi.GS_REDBLACK if rb else _abi.GS_WAVEFRONT
            pending[srv.submit(sc, gs_mode=mode)] = (sc, mode)
            k += 1
        time.sleep(0.05)
        for t in list(pending):  # collect what has finished, free its record
            st = srv.poll(t)
            if st.state == _abi.REQ_DONE:
                sc, mode = pending.pop(t)
                s = gpu_scene(sc, mode)
                s.step(sc["frames"])
                s.sync()
                assert_bit_equal(srv.result(t), s.download(_abi.DENSITY), f"churn N={sc['size'] - 2} mode={mode}")
                s.close()
                srv.release(t)
    srv.wait_idle(120000)
    for t, (sc, mode) in pending.items():
        assert srv.poll(t).state == _abi.REQ_DONE
        s = gpu_scene(sc, mode)
        s.step(sc["frames"])
        s.sync()
        assert_bit_equal(srv.result(t), s.download(_abi.DENSITY), f"churn N={sc['size'] - 2} mode={mode}")
        s.close()
    srv.close()
