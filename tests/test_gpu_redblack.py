"""Red-black mode on the GPU.

* Bit-for-bit equality with the red-black f32 CPU oracle for every band shape
  and every sweeps-per-launch setting (temporal blocking must not change a
  single bit: halo cells are recomputed with the same operations).
* Accuracy against the red-black f64 oracle: norm-wise relative error of
  density and velocity and the pressure Poisson residual.
"""
import numpy as np
import pytest

import oracle
from equilibrium_amd import _abi, scenes
from helpers import FIELD_NAMES, assert_state_equal, download_all, gpu_scene, oracle_state, run_oracle

pytestmark = pytest.mark.gpu


def _case(label):
    print(f"[case] {label}", flush=True)


@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("tsw", [1, 2, 4, 5, 6, 8])
@pytest.mark.parametrize("N", [150, 151, 330])
def test_redblack_bit_exact_all_tilings(tile, tsw, N):
    """N=150: two tiles per dimension (one at each edge); N=151: the same with
    an odd size (the grid's last column is the first of a lane's pair, its
    8-byte load straddles the edge and its store is a single); N=330: three
    or more tiles in most settings, so interior tiles with halos on both sides
    too."""
    _case(f"rb tile={tile} T={tsw} N={N}")
    sc = scenes.stam_scene(N, scenes.derive_seed(20, N), iters=7, frames=3, n_rects=4,
                           diffusion=1e-4, viscosity=1e-3)
    s = gpu_scene(sc, _abi.GS_REDBLACK)
    s.set_tuning(tile, tsw)
    s.step(sc["frames"])
    s.sync()
    st = run_oracle(sc, order=oracle.RB)
    assert_state_equal(download_all(s), st, f"rb tile={tile} T={tsw}")


@pytest.mark.parametrize("N,iters,frames", [(1, 2, 2), (2, 3, 2), (3, 3, 2), (8, 3, 2), (30, 5, 3), (31, 4, 2),
                                           (33, 4, 2), (64, 20, 2), (201, 9, 2), (1100, 5, 1), (2047, 4, 1)])
def test_redblack_bit_exact_sizes(N, iters, frames):
    _case(f"rb N={N}")
    sc = scenes.stam_scene(N, scenes.derive_seed(21, N), iters=iters, frames=frames, n_rects=3)
    s = gpu_scene(sc, _abi.GS_REDBLACK)
    s.step(frames)
    s.sync()
    st = run_oracle(sc, order=oracle.RB)
    assert_state_equal(download_all(s), st, f"rb N={N}")


def test_redblack_break_trigger_bit_exact():
    N = 96
    _case("rb break")
    sc = scenes.stam_scene(N, scenes.derive_seed(22), iters=5, frames=3, n_rects=2)
    S = N + 2
    u = np.full(S * S, -40.0, np.float32)
    v = np.full(S * S, 3.0, np.float32)
    v[(np.arange(S * S) // S) % 7 == 0] = -50.0
    sc["init_u"], sc["init_v"] = u, v
    s = gpu_scene(sc, _abi.GS_REDBLACK)
    s.step(sc["frames"])
    s.sync()
    assert_state_equal(download_all(s), run_oracle(sc, order=oracle.RB), "rb break")


def rel_err(a, b):
    return float(np.linalg.norm(a.astype(np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def test_redblack_accuracy_vs_fp64():
    """N=254, 40 sweeps, curl-noise velocity, 8 frames: f32 GPU vs f64 CPU."""
    N = 254
    _case(f"rb fp64 N={N}")
    sc = scenes.redblack_scene(N=N)
    sc["frames"] = 8
    s = gpu_scene(sc, _abi.GS_REDBLACK)
    st = oracle_state(sc).astype(np.float64)
    worst = {}
    for f in range(sc["frames"]):
        s.step(1)
        s.sync()
        st.run(sc["dt"], sc["diffusion"], sc["viscosity"], sc["iters"], 1, order=oracle.RB,
               src_d=sc["src_d"])
        got = download_all(s)
        for k in ("density", "vx", "vy"):
            worst[k] = max(worst.get(k, 0.0), rel_err(got[k], st.f[k]))
    print("[rb-fp64] worst per-frame relative error:", worst, flush=True)
    for k, e in worst.items():
        assert e < 1e-5, (k, e)


def _rb_mix():
    # sizes from one strip to past the wavefront's 1026 limit; one iteration count
    return [scenes.stam_scene(N, scenes.derive_seed(23, N), iters=7, frames=2, n_rects=k % 4,
                              viscosity=10 ** (-3 - k % 3), diffusion=10 ** (-5 + k % 2))
            for k, N in enumerate((30, 150, 64, 1100, 257, 8))]


def test_redblack_batch_equals_solo_and_oracle():
    """A red-black batch runs every frame kernel over all its scenes; each
    scene's result equals its solo run and the red-black oracle bit for bit,
    for every band setting the autotuner may pick."""
    _case("rb batch")
    mix = _rb_mix()
    want = {}
    for k, sc in enumerate(mix):
        s = gpu_scene(sc, _abi.GS_REDBLACK)
        s.step(sc["frames"])
        s.sync()
        want[k] = download_all(s)
        s.close()
        if sc["size"] <= 152:
            assert_state_equal(want[k], run_oracle(sc, order=oracle.RB), f"rb solo N={sc['size'] - 2}")
    for tile, tsw in ((0, 5), (2, 8), (1, 2), (4, 6), (5, 5)):
        _case(f"rb batch band {tile}:{tsw}")
        hs = [gpu_scene(sc, _abi.GS_REDBLACK) for sc in mix]
        b = _abi.Batch(hs)
        b.set_band(tile, tsw)
        b.step(mix[0]["frames"])
        b.sync()
        b.close()
        for k, (sc, h) in enumerate(zip(mix, hs)):
            got = download_all(h)
            for name in FIELD_NAMES:
                assert np.array_equal(got[name].view(np.uint32), want[k][name].view(np.uint32)), \
                    (tile, tsw, sc["size"], name)
            h.close()


def test_redblack_batch_of_many_small_scenes():
    """48 scenes of sizes 1..200 in one red-black batch: work units of many
    scenes share every launch (tiles, rows, cell blocks straddling nothing),
    each scene equal to the red-black oracle bit for bit."""
    _case("rb batch small")
    rng = np.random.default_rng(24)
    sizes = [1, 2, 3, 5, 16, 31, 32, 33] + [int(x) for x in rng.integers(4, 200, 40)]
    mix = [scenes.stam_scene(N, scenes.derive_seed(25, k), iters=5, frames=2, n_rects=k % 3,
                             viscosity=10 ** (-3 - k % 3), diffusion=10 ** (-5 + k % 2))
           for k, N in enumerate(sizes)]
    hs = [gpu_scene(sc, _abi.GS_REDBLACK) for sc in mix]
    b = _abi.Batch(hs)
    b.step(2)
    b.sync()
    b.close()
    for sc, h in zip(mix, hs):
        assert_state_equal(download_all(h), run_oracle(sc, order=oracle.RB), f"rb batch N={sc['size'] - 2}")
        h.close()


def test_redblack_batch_rejects_mixed_modes_and_counts():
    _case("rb batch errors")
    a = gpu_scene(scenes.stam_scene(40, 1, iters=4, frames=1), _abi.GS_REDBLACK)
    b = gpu_scene(scenes.stam_scene(40, 2, iters=4, frames=1), _abi.GS_WAVEFRONT)
    c = gpu_scene(scenes.stam_scene(40, 3, iters=5, frames=1), _abi.GS_REDBLACK)
    with pytest.raises(_abi.EqfError):
        _abi.Batch([a, b])
    with pytest.raises(_abi.EqfError):
        _abi.Batch([a, c])
    bt = _abi.Batch([a])
    with pytest.raises(_abi.EqfError):
        bt.set_tuning(8, 0)  # wavefront knob
    with pytest.raises(_abi.EqfError):
        bt.set_band(0, 3)
    # a batch member's descriptor is held by the batch: no edits while in it
    pa = scenes.scene_params(scenes.stam_scene(40, 1, iters=4, frames=1), _abi.GS_REDBLACK)
    for edit in (lambda: a.fill_rect(2, 2, 5, 5), lambda: a.step(1), lambda: a.set_params(pa)):
        with pytest.raises(_abi.EqfError) as ei:
            edit()
        assert ei.value.code == _abi.EQF_EBUSY
    bt.close()
    a.fill_rect(2, 2, 5, 5)  # free again once the batch is closed
