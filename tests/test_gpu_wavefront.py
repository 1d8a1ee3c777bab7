"""Exact (wavefront) mode on the GPU: bit-for-bit equality with the CPU oracle.

Every test drives libeqfluid.so through its C ABI and compares all six fields
(density, scratch, vx, vy, vx0, vy0, ring included) as raw f32 bits.
"""
import numpy as np
import pytest

import oracle
from equilibrium_amd import _abi, scenes
from helpers import (FIELD_NAMES, assert_bit_equal, assert_state_equal, download_all, gpu_scene,
                     oracle_state, run_oracle)

pytestmark = pytest.mark.gpu


def _case(label):
    print(f"[case] {label}", flush=True)


@pytest.mark.parametrize("size,times", [(3, 1), (12, 1), (20, 1), (21, 2), (22, 1), (130, 1), (128, 2)])
def test_initial_state_matches_reference_init(size, times):
    _case(f"init size={size} times={times}")
    p = _abi.default_params()
    p.size = size
    p.init_times = times
    s = _abi.Scene(p)
    st = oracle.State.new(size, times)
    assert_state_equal(download_all(s), st, f"init {size}")
    assert np.array_equal(s.get_mask(), st.mask)


@pytest.mark.parametrize("N,iters,frames,rects", [
    (8, 4, 3, 1),      # serial path (N < 11)
    (16, 5, 3, 1),
    (32, 6, 4, 2),
    (62, 7, 3, 3),
    (64, 20, 2, 2),
    (100, 6, 2, 3),    # two strips, partial second strip
    (130, 5, 2, 4),
    (198, 4, 2, 5),    # four strips
])
def test_scene_bit_exact(N, iters, frames, rects):
    _case(f"scene N={N} iters={iters} frames={frames}")
    sc = scenes.stam_scene(N, scenes.derive_seed(10, N), iters=iters, frames=frames, n_rects=rects,
                           diffusion=1e-4, viscosity=1e-3)
    s = gpu_scene(sc)
    s.step(frames)
    s.sync()
    st = run_oracle(sc)
    assert_state_equal(download_all(s), st, f"N={N}")


def test_obstacles_across_strip_boundaries():
    """Walls on both sides of the row-64/65 strip boundary exercise the
    cross-strip AdjustColumn fix-up and the hand-off of raw last-row values."""
    N = 150
    _case(f"strip boundary N={N}")
    sc = scenes.stam_scene(N, scenes.derive_seed(11), iters=6, frames=3, n_rects=0,
                           diffusion=1e-4, viscosity=1e-3)
    sc["rects"] = [(10, 60, 30, 65), (40, 65, 60, 70), (70, 63, 72, 67), (80, 129, 100, 131),
                   (110, 64, 111, 66), (120, 1, 125, 64)]
    s = gpu_scene(sc)
    s.step(sc["frames"])
    s.sync()
    st = run_oracle(sc)
    assert_state_equal(download_all(s), st, "strip boundary")


def test_advect_break_trigger():
    """Strong negative velocity makes the back-trace hit the far edge, so the
    reference's `break` (fluid.rs:420-423) fires on many rows."""
    N = 96
    _case(f"break N={N}")
    sc = scenes.stam_scene(N, scenes.derive_seed(12), iters=5, frames=3, n_rects=2)
    S = N + 2
    u = np.full(S * S, -40.0, np.float32)
    v = np.full(S * S, 3.0, np.float32)
    v[(np.arange(S * S) // S) % 7 == 0] = -50.0
    sc["init_u"], sc["init_v"] = u, v
    s = gpu_scene(sc)
    s.step(sc["frames"])
    s.sync()
    st = run_oracle(sc)
    assert_state_equal(download_all(s), st, "break")


def test_zero_iterations_and_default_state():
    """frames = 0 sweeps: lin_solve does nothing, not even set_boundaries."""
    _case("iters=0")
    sc = scenes.stam_scene(40, scenes.derive_seed(13), iters=0, frames=2, n_rects=1)
    s = gpu_scene(sc)
    s.step(2)
    s.sync()
    st = run_oracle(sc)
    assert_state_equal(download_all(s), st, "iters=0")


def test_small_scene_full_bit_exact():
    """The N=128 scene (20 sweeps) for all its 100 frames."""
    sc = scenes.small_scene()
    _case("small_scene N=128 100 frames")
    s = gpu_scene(sc)
    s.step(sc["frames"])
    s.sync()
    st = run_oracle(sc)
    assert_state_equal(download_all(s), st, "small_scene")


def test_single_1024_first_frames_bit_exact():
    """The N=1024 scene (16 strips, 3 rectangles, 4 sources), first
    2 frames compared with the oracle."""
    sc = scenes.single_1024()
    _case("single_1024 N=1024 2 frames")
    s = gpu_scene(sc)
    s.step(2)
    s.sync()
    st = run_oracle(sc, frames=2)
    assert_state_equal(download_all(s), st, "single_1024")


def test_step_in_pieces_equals_one_call():
    sc = scenes.stam_scene(70, scenes.derive_seed(14), iters=5, frames=4, n_rects=2)
    _case("pieces N=70")
    a = gpu_scene(sc)
    a.step(4)
    b = gpu_scene(sc)
    for _ in range(4):
        b.step(1)
    a.sync()
    b.sync()
    fa, fb = download_all(a), download_all(b)
    for n in FIELD_NAMES:
        assert_bit_equal(fa[n], fb[n], n)


def test_clone_and_resume():
    """Clone mid-run and continue: the copy follows the original bit for bit,
    and a download/upload round trip resumes exactly (checkpoint/resume)."""
    sc = scenes.stam_scene(80, scenes.derive_seed(15), iters=5, frames=4, n_rects=2)
    _case("clone N=80")
    a = gpu_scene(sc)
    a.step(2)
    b = a.clone()
    c = gpu_scene(sc)
    mid = download_all(a)
    for i, n in enumerate(FIELD_NAMES):
        c.upload(i, mid[n])
    for x in (a, b, c):
        x.step(2)
        x.sync()
    fa, fb, fc = download_all(a), download_all(b), download_all(c)
    for n in FIELD_NAMES:
        assert_bit_equal(fb[n], fa[n], "clone " + n)
        assert_bit_equal(fc[n], fa[n], "resume " + n)
    st = run_oracle(sc)
    assert_state_equal(fa, st, "clone vs oracle")


def test_add_density_velocity_cells():
    """add_density / add_velocity with clamped indices (fluid.rs:120-131)."""
    _case("add cell")
    p = _abi.default_params()
    p.size = 30
    s = _abi.Scene(p)
    s.add_density(3, 4, 2.5)
    s.add_density(100, 4, 1.0)   # clamped to x = 29
    s.add_velocity(7, 7, 0.5, -0.25)
    s.sync()
    st = oracle.State.new(30)
    st.f["density"][3 + 4 * 30] += np.float32(2.5)
    st.f["scratch"][3 + 4 * 30] += np.float32(2.5)
    st.f["density"][29 + 4 * 30] += np.float32(1.0)
    st.f["scratch"][29 + 4 * 30] += np.float32(1.0)
    st.f["vx"][7 + 7 * 30] += np.float32(0.5)
    st.f["vy"][7 + 7 * 30] += np.float32(-0.25)
    assert_state_equal(download_all(s), st, "add cell")


def test_batch_mixed_sizes_equals_solo():
    """A batch of scenes of different sizes advanced by one launch per frame:
    each scene equals its solo run and the oracle."""
    _case("batch mixed")
    scs = [scenes.stam_scene(N, scenes.derive_seed(16, i), iters=4, frames=3, n_rects=i % 3,
                             viscosity=10 ** (-3 - i % 3), diffusion=10 ** (-4 - i % 2))
           for i, N in enumerate([64, 30, 128, 100, 8, 190, 64])]
    handles = [gpu_scene(sc) for sc in scs]
    b = _abi.Batch(handles)
    b.step(3)
    b.sync()
    b.close()
    for sc, h in zip(scs, handles):
        st = run_oracle(sc)
        assert_state_equal(download_all(h), st, f"batch N={sc['size'] - 2}")


@pytest.mark.parametrize("mode", [_abi.GS_WAVEFRONT, _abi.GS_REDBLACK])
def test_add_cell_touches_one_cell_of_its_fields(mode):
    """eqf_add_cell: EQF_DENSITY adds to density and scratch like add_density;
    EQF_VX / EQF_VY add to that component only (the other keeps its -0);
    indices clamp like idx! (fluid.rs:31-35)."""
    N = 70
    S = N + 2
    s = gpu_scene(scenes.stam_scene(N, scenes.derive_seed(60), iters=3, frames=1, n_rects=1), mode)
    neg0 = np.full(S * S, -0.0, np.float32)
    s.upload(_abi.VX, neg0)
    s.upload(_abi.VY, neg0)
    before = download_all(s)
    s.add_cell(_abi.DENSITY, 5, 9, 2.5)
    s.add_cell(_abi.VX, 10, 3, 1.25)
    s.add_cell(_abi.VY, 10 ** 6, 4, -0.75)  # x clamps to S-1
    s.sync()
    after = download_all(s)
    want = {k: v.copy() for k, v in before.items()}
    for name in ("density", "scratch"):
        want[name][5 + 9 * S] = np.float32(want[name][5 + 9 * S] + np.float32(2.5))
    want["vx"][10 + 3 * S] = np.float32(np.float32(-0.0) + np.float32(1.25))
    want["vy"][(S - 1) + 4 * S] = np.float32(np.float32(-0.0) + np.float32(-0.75))
    for name in FIELD_NAMES:
        assert_bit_equal(after[name], want[name], name)
    with pytest.raises(_abi.EqfError):
        s.add_cell(_abi.VX0, 1, 1, 1.0)
    s.close()


def test_add_noise_rotation_closed_forms():
    """add_noise (fluid.rs:575-599) with the random inputs supplied: at 0, 90
    and 180 degrees the rotated end point has a closed form (the centre is
    (size/2, size/2)); twice it is added to the centre cell's velocity."""
    _case("add_noise")
    size = 40
    c = size // 2
    for ang, (rx, ry) in ((0.0, (31, 7)), (90.0, (31, 7)), (180.0, (12, 25))):
        s = _abi.Scene(_noise_params(size))
        u0, v0 = s.download(_abi.VX), s.download(_abi.VY)
        s.add_noise(ang, rx, ry)
        s.sync()
        du = s.download(_abi.VX) - u0
        dv = s.download(_abi.VY) - v0
        dx, dy = rx - c, ry - c
        want = {0.0: (rx, ry), 90.0: (c - dy, c + dx), 180.0: (c - dx, c - dy)}[ang]
        k = c + c * size
        assert np.count_nonzero(du) <= 1 and np.count_nonzero(dv) <= 1
        assert du[k] == pytest.approx(2.0 * want[0], abs=1e-4), (ang, du[k], want)
        assert dv[k] == pytest.approx(2.0 * want[1], abs=1e-4), (ang, dv[k], want)
        s.close()


def test_add_noise_seeded_point_is_deterministic():
    from equilibrium_amd.fluid import Fluid, FluidConfigs, SimulationConfigs

    _case("add_noise seeded")
    outs = []
    for _ in range(2):
        f = Fluid(FluidConfigs(), SimulationConfigs(0.1, 4, 50))
        rng = scenes.SplitMix64(1234)
        for _ in range(3):
            f.add_noise(37.0, rng=rng)
            f.step()
        outs.append((f.velocities_x, f.velocities_y))
    assert_bit_equal(outs[0][0], outs[1][0], "vx")
    assert_bit_equal(outs[0][1], outs[1][1], "vy")


def _noise_params(size):
    p = _abi.default_params()
    p.size = size
    return p


@pytest.mark.parametrize("sched", [_abi.SCHED_WIDE, _abi.SCHED_RING, _abi.SCHED_FINE])
def test_batch_tuning_keeps_results(sched):
    """Every launch shape the batch autotuner may pick (waves per workgroup,
    workgroup order, wavefront schedule) gives the oracle's bits.  With the
    ring schedule and 16 waves the scenes run 1 (N=40), 2 (N=70), 3 (N=130)
    and 4 (N=200) sweep streams per strip, chaining x0 through 0 to 3 rings."""
    _case(f"batch tuning, schedule {sched}")
    scs = [scenes.stam_scene(N, scenes.derive_seed(60, N), iters=5, frames=2, n_rects=2)
           for N in (40, 130, 200, 70)]
    want = [run_oracle(sc, frames=2) for sc in scs]
    for w, o in ((16, 0), (8, 1), (4, 0), (4, 1)):
        hs = [gpu_scene(sc) for sc in scs]
        b = _abi.Batch(hs)
        b.set_tuning(w, o)
        b.set_schedule(sched)
        b.step(2)
        b.sync()
        b.close()
        for sc, h, st in zip(scs, hs, want):
            assert_state_equal(download_all(h), st, f"tuning {w}:{o}:{sched} N={sc['size'] - 2}")
            h.close()
    hs = [gpu_scene(sc) for sc in scs]
    b = _abi.Batch(hs)
    with pytest.raises(_abi.EqfError):
        b.set_tuning(2, 0)  # N=200 needs 4 strips
    b.close()
