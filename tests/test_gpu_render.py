"""Rendering and per-frame snapshots on the GPU.

* eqf_render_rgba == the oracle's render_image rule (renderer_helpers.rs:
  143-165) byte for byte, on stepped scenes of both solver modes and on
  uploaded densities holding every special value of `f32 as u8`.
* Snapshot rings: each capture holds exactly the state after its frame (the
  density of a scene stepped in lockstep and downloaded synchronously) and
  its rendering, with captures overlapping the following steps; ring
  bookkeeping errors (full ring, empty ring) are reported, not hidden.
"""
import math

import numpy as np
import pytest

import oracle
from equilibrium_amd import _abi, scenes
from equilibrium_amd.fluid import CurrentSimulation, Fluid, FluidConfigs, RenderingListener, SimulationConfigs
from helpers import assert_bit_equal, gpu_scene, run_oracle

pytestmark = pytest.mark.gpu


def _case(label):
    print(f"[case] {label}", flush=True)


def _oracle_image(scene, colors=None):
    c = colors or _abi.make_colors()
    return oracle.render_rgba(scene.download(_abi.DENSITY), scene.get_mask(), scene.size,
                              list(c.fluid), list(c.world), list(c.obstacle))


@pytest.mark.parametrize("mode", [_abi.GS_WAVEFRONT, _abi.GS_REDBLACK])
@pytest.mark.parametrize("N", [30, 126, 301])
def test_render_stepped_scene(mode, N):
    _case(f"render mode={mode} N={N}")
    sc = scenes.stam_scene(N, scenes.derive_seed(40, N), iters=6, frames=3, n_rects=3)
    s = gpu_scene(sc, mode)
    s.step(sc["frames"])
    s.sync()
    img = s.render_rgba()
    np.testing.assert_array_equal(img, _oracle_image(s))
    if mode == _abi.GS_WAVEFRONT:
        # the density rendered is the reference's, bit for bit
        assert_bit_equal(s.download(_abi.DENSITY), run_oracle(sc).f["density"], "density")
    assert (img == np.array(oracle.OBSTACLE_COLOR, np.uint8)).all(1).sum() == int((s.get_mask() == 0).sum())


def test_render_special_values_and_colours():
    N = 62
    S = N + 2
    _case(f"render special values S={S}")
    sc = scenes.stam_scene(N, scenes.derive_seed(41), iters=2, frames=1, n_rects=2)
    s = gpu_scene(sc)
    rng = np.random.default_rng(41)
    special = np.array([0.0, -0.0, 0.5, 0.999, 1.0, 1.3, 254.99, 255.0, 300.0, 1e-30, -3.0, np.nan, np.inf, -np.inf,
                        np.float32(255.0) / np.float32(208.0),
                        np.nextafter(np.float32(255.0) / np.float32(208.0), np.float32(0))], np.float32)
    d = rng.choice(special, S * S).astype(np.float32)
    mix = rng.random(S * S) < 0.3
    d[mix] = rng.uniform(-2, 400, int(mix.sum())).astype(np.float32)
    s.upload(_abi.DENSITY, d)
    np.testing.assert_array_equal(s.render_rgba(), _oracle_image(s))
    c = _abi.make_colors((0, 7, 9, 11), (1, 2, 3, 4), (200, 100, 50, 25))
    np.testing.assert_array_equal(s.render_rgba(c), _oracle_image(s, c))


def test_render_after_mask_change():
    """A mask edited after the last step (not yet seen by any step) renders
    with the new mask."""
    N = 40
    _case("render after fill_rect")
    sc = scenes.stam_scene(N, scenes.derive_seed(42), iters=3, frames=2, n_rects=1)
    s = gpu_scene(sc)
    s.step(2)
    s.fill_rect(5, 5, 20, 12)
    np.testing.assert_array_equal(s.render_rgba(), _oracle_image(s))


@pytest.mark.parametrize("mode", [_abi.GS_WAVEFRONT, _abi.GS_REDBLACK])
@pytest.mark.parametrize("depth,N", [(1, 190), (2, 190), (3, 190), (2, 61)])
def test_snapshot_ring_matches_lockstep_downloads(mode, depth, N):
    """N=61: size*size = 3969 is odd, so the RGBA image starts at a rounded
    offset in the slot and the render kernel's last thread takes a tail."""
    frames = 7
    _case(f"snapshot mode={mode} depth={depth} N={N}")
    sc = scenes.stam_scene(N, scenes.derive_seed(43, N), iters=8, frames=frames, n_rects=2)
    a = gpu_scene(sc, mode)
    b = gpu_scene(sc, mode)
    snap = _abi.Snapshot(a, depth, with_rgba=True)
    want_d, want_img = [], []
    got = []
    for f in range(frames):
        if f >= depth:
            got.append(snap.acquire())
            snap.release()
        a.step(1)
        snap.capture(100 + f)
        b.step(1)
        b.sync()
        want_d.append(b.download(_abi.DENSITY))
        want_img.append(_oracle_image(b))
    while len(got) < frames:
        got.append(snap.acquire())
        snap.release()
    snap.close()
    for f, (fr, dens, img) in enumerate(got):
        assert fr == 100 + f
        assert_bit_equal(dens, want_d[f], f"frame {f}")
        np.testing.assert_array_equal(img, want_img[f])


def test_snapshot_density_only_and_ring_errors():
    N = 50
    _case("snapshot errors")
    sc = scenes.stam_scene(N, scenes.derive_seed(44), iters=3, frames=2, n_rects=1)
    a = gpu_scene(sc)
    snap = _abi.Snapshot(a, 2, with_rgba=False)
    with pytest.raises(_abi.EqfError) as ei:
        snap.acquire()
    assert ei.value.code == _abi.EQF_EINVAL
    with pytest.raises(_abi.EqfError):
        snap.release()
    a.step(1)
    snap.capture(0)
    a.step(1)
    snap.capture(1)
    with pytest.raises(_abi.EqfError) as ei:
        snap.capture(2)
    assert ei.value.code == _abi.EQF_EBUSY
    fr, dens, img = snap.acquire(copy=False)
    assert fr == 0 and img is None and dens.shape == (a.size * a.size,)
    snap.release()
    snap.capture(2)  # the freed slot takes the next capture
    a.sync()
    cur = a.download(_abi.DENSITY)
    snap.release()
    fr, dens, _ = snap.acquire()
    assert fr == 2
    assert_bit_equal(dens, cur, "frame 2")
    snap.close()
    with pytest.raises(_abi.EqfError):
        _abi.Snapshot(a, 0)
    with pytest.raises(_abi.EqfError):
        _abi.Snapshot(a, 9)


def test_current_simulation_snapshots_match_simulate():
    """simulate_snapshots hands over the frames simulate's callback sees."""
    _case("CurrentSimulation.simulate_snapshots")
    sim_cfg = SimulationConfigs(0.02, 5, 128)
    want = []
    ref = CurrentSimulation(Fluid(FluidConfigs(), sim_cfg, _init_times=2))
    listener = RenderingListener()
    ref.simulate(lambda i, f: want.append((f.density, listener.render_image(f))))
    got = list(CurrentSimulation(Fluid(FluidConfigs(), SimulationConfigs(0.02, 5, 128), _init_times=2))
               .simulate_snapshots(depth=2, listener=listener))
    assert [g[0] for g in got] == list(range(5))
    for (fr, dens, img), (wd, wi) in zip(got, want):
        assert_bit_equal(dens, wd, f"frame {fr}")
        np.testing.assert_array_equal(img, wi)
        assert img.shape == (128, 128, 4)
    assert not math.isnan(float(got[-1][1].sum()))
