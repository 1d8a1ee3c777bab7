"""Parameter changes between steps, and scene calls interleaved with batch
steps, on the GPU against the oracle."""
import numpy as np
import pytest

import oracle
from equilibrium_amd import _abi, scenes
from helpers import FIELD_NAMES, assert_state_equal, download_all, gpu_scene, oracle_state

pytestmark = pytest.mark.gpu


def _case(label):
    print(f"[case] {label}", flush=True)


def _run(st, sc, frames, dt=None, order=oracle.LEX):
    st.run(sc["dt"] if dt is None else dt, sc["diffusion"], sc["viscosity"], sc["iters"], frames, order=order,
           src_d=sc["src_d"], src_u=sc["src_u"], src_v=sc["src_v"])


@pytest.mark.parametrize("mode", [_abi.GS_REDBLACK, _abi.GS_WAVEFRONT])
def test_delta_t_change_with_zero_viscosity_and_diffusion(mode):
    """With viscosity = diffusion = 0 a new delta_t changes only the advection
    back-trace (dt0 = dt*(S-2), fluid.rs:390); the next frames must use it."""
    _case(f"delta_t change, mode {mode}")
    sc = scenes.stam_scene(90, scenes.derive_seed(60), iters=6, frames=3, n_rects=2, viscosity=0.0, diffusion=0.0)
    s = gpu_scene(sc, mode)
    s.step(1)
    p = scenes.scene_params(sc, mode)
    p.delta_t = 0.05
    s.set_params(p)
    s.step(2)
    s.sync()
    order = oracle.RB if mode == _abi.GS_REDBLACK else oracle.LEX
    st = oracle_state(sc)
    _run(st, sc, 1, order=order)
    _run(st, sc, 2, dt=0.05, order=order)
    assert_state_equal(download_all(s), st, f"delta_t change mode {mode}")
    s.close()


@pytest.mark.parametrize("mode", [_abi.GS_WAVEFRONT, _abi.GS_REDBLACK])
def test_scene_calls_between_batch_steps(mode):
    """The reference's frame loop (add sources, then step) on batched scenes:
    adds, uploads and downloads issued on member scenes between batch steps
    run after the previous step and before the next one, with no explicit
    synchronisation by the caller."""
    _case(f"scene calls between batch steps, mode {mode}")
    order = oracle.RB if mode == _abi.GS_REDBLACK else oracle.LEX
    scs = [scenes.stam_scene(N, scenes.derive_seed(61, N), iters=5, frames=3, n_rects=1) for N in (40, 72, 130)]
    hs = [gpu_scene(sc, mode) for sc in scs]
    b = _abi.Batch(hs)
    sts = [oracle_state(sc) for sc in scs]
    rng = np.random.default_rng(62)
    b.step(1)
    early = hs[1].download(_abi.DENSITY)   # right after the queued step, no sync
    for h, st, sc in zip(hs, sts, scs):
        _run(st, sc, 1, order=order)
    np.testing.assert_array_equal(early.view(np.uint32), sts[1].f["density"].view(np.uint32))
    # between steps: one-cell adds and a whole-field upload
    S0 = scs[0]["size"]
    hs[0].add_density(5, 7, 3.5)
    sts[0].f["density"][5 + 7 * S0] += np.float32(3.5)
    sts[0].f["scratch"][5 + 7 * S0] += np.float32(3.5)
    S2 = scs[2]["size"]
    vx = rng.standard_normal(S2 * S2).astype(np.float32)
    hs[2].upload(_abi.VX, vx)
    sts[2].f["vx"][:] = vx
    hs[1].add_velocity(9, 9, -1.25, 0.5)
    S1 = scs[1]["size"]
    sts[1].f["vx"][9 + 9 * S1] += np.float32(-1.25)
    sts[1].f["vy"][9 + 9 * S1] += np.float32(0.5)
    b.step(2)
    for st, sc in zip(sts, scs):
        _run(st, sc, 2, order=order)
    got = [download_all(h) for h in hs]   # ordered after the batch's steps
    for i, (g, st) in enumerate(zip(got, sts)):
        assert_state_equal(g, st, f"batch member {i} mode {mode}")
    b.close()
    for h in hs:
        h.close()


def test_device_key_names_the_gpu():
    """eqf_ctx_device_key: "<name>/<gfx arch>/<CU count>CU", the tuning caches' key."""
    _case("device key")
    key = _abi.default_context().device_key()
    print(f"[device] {key}", flush=True)
    name, arch, cus = key.rsplit("/", 2)
    assert name and arch.startswith("gfx950") and cus.endswith("CU") and int(cus[:-2]) > 0
