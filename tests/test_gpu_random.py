"""Seeded random scenes, both solver modes, against the oracle bit for bit.

Sizes 3..330 (single strip and multi-strip, partial last strips, periods
padded to 16, sweep streams), 0-5 obstacles, random sweep counts, time steps
large enough to trigger advect's row `break`, viscosity and diffusion over
five decades, and frame sources on some fields only.
"""
import numpy as np
import pytest

import oracle
from equilibrium_amd import _abi, scenes
from helpers import assert_state_equal, download_all, gpu_scene, run_oracle

pytestmark = pytest.mark.gpu


def _random_scene(k):
    rng = scenes.SplitMix64(scenes.derive_seed(50, k))
    N = rng.randint(1, 330) if k % 2 == 0 else int(round(rng.log_uniform(1.0, 330.0)))
    sc = scenes.stam_scene(N, scenes.derive_seed(51, k), dt=rng.uniform(0.01, 0.6),
                           diffusion=rng.log_uniform(1e-7, 1e-2), viscosity=rng.log_uniform(1e-7, 1e-2),
                           iters=rng.randint(1, 12), frames=rng.randint(1, 3), n_rects=rng.randint(0, 5))
    drop = rng.randint(0, 3)  # 0: keep every source; 1..3: drop one field's source
    if drop:
        sc[("src_d", "src_u", "src_v")[drop - 1]] = None
    return sc


@pytest.mark.parametrize("k", range(16))
def test_random_scene_wavefront_bit_exact(k):
    sc = _random_scene(k)
    print(f"[case] random wavefront k={k} N={sc['size'] - 2} iters={sc['iters']} frames={sc['frames']} "
          f"dt={sc['dt']:.3f} rects={len(sc['rects'])}", flush=True)
    s = gpu_scene(sc)
    s.step(sc["frames"])
    s.sync()
    assert_state_equal(download_all(s), run_oracle(sc), f"random k={k}")


@pytest.mark.parametrize("k", range(16, 24))
def test_random_scene_redblack_bit_exact(k):
    sc = _random_scene(k)
    print(f"[case] random red-black k={k} N={sc['size'] - 2} iters={sc['iters']} frames={sc['frames']}",
          flush=True)
    s = gpu_scene(sc, _abi.GS_REDBLACK)
    s.step(sc["frames"])
    s.sync()
    assert_state_equal(download_all(s), run_oracle(sc, order=oracle.RB), f"random red-black k={k}")


def test_random_scenes_batched_equal_oracle():
    scs = [_random_scene(k) for k in range(24, 40)]
    hs = [gpu_scene(sc) for sc in scs]
    frames = 2
    b = _abi.Batch(hs)
    b.step(frames)
    b.sync()
    b.close()
    for k, (sc, h) in enumerate(zip(scs, hs)):
        assert_state_equal(download_all(h), run_oracle(sc, frames=frames), f"batched random k={k + 24}")


@pytest.mark.parametrize("sched", [_abi.SCHED_AUTO, _abi.SCHED_RING, _abi.SCHED_FINE])
def test_exact_batch_of_strip_boundary_sizes_equals_oracle(sched):
    """One fused launch over scenes at and around the strip (64 rows) and
    period (16 slots) boundaries, down to size 3, with sweep counts from 3 to
    6 and frame sources on some fields only; every scene equal to the oracle
    bit for bit, under the default schedule and the x0 ring schedule (sweep
    counts below the stream count cut the streams)."""
    sizes = [1, 2, 3, 5, 15, 16, 17, 62, 63, 64, 65, 66, 126, 127, 128, 129, 130, 191, 192, 193, 255, 257]
    scs = []
    for k, N in enumerate(sizes):
        sc = scenes.stam_scene(N, scenes.derive_seed(52, k), iters=3 + k % 4, frames=2, n_rects=k % 4,
                               viscosity=10 ** (-3 - k % 3), diffusion=10 ** (-5 + k % 2), dt=0.05 + 0.1 * (k % 3))
        if k % 3 == 1:
            sc["src_u"] = None
        if k % 5 == 2:
            sc["src_d"] = None
        scs.append(sc)
    print(f"[case] exact batch of {len(scs)} boundary sizes, schedule {sched}", flush=True)
    hs = [gpu_scene(sc) for sc in scs]
    b = _abi.Batch(hs)
    b.set_schedule(sched)
    b.step(2)
    b.sync()
    b.close()
    for sc, h in zip(scs, hs):
        assert_state_equal(download_all(h), run_oracle(sc, frames=2), f"exact batch N={sc['size'] - 2}")
        h.close()
