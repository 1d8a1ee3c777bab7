"""Shared helpers: run a synthetic scene on the GPU library and on the oracle."""
import numpy as np

import oracle
from equilibrium_amd import _abi, scenes

FIELD_NAMES = oracle.FIELDS  # density, scratch, vx, vy, vx0, vy0 (eqf_field order)


def oracle_state(sc):
    st = oracle.State.new(sc["size"])
    for r in sc["rects"]:
        st.fill_rect(*r)
    if sc.get("init_u") is not None:
        st.f["vx"][:] = sc["init_u"]
        st.f["vy"][:] = sc["init_v"]
    return st


def run_oracle(sc, frames=None, order=oracle.LEX, dtype=np.float32, st=None):
    if st is None:
        st = oracle_state(sc)
    if dtype != np.float32:
        st = st.astype(dtype)
    st.run(sc["dt"], sc["diffusion"], sc["viscosity"], sc["iters"],
           sc["frames"] if frames is None else frames, order=order,
           src_d=sc["src_d"], src_u=sc["src_u"], src_v=sc["src_v"])
    return st


def gpu_scene(sc, gs_mode=_abi.GS_WAVEFRONT):
    s = _abi.Scene(scenes.scene_params(sc, gs_mode))
    scenes.apply_to_scene(s, sc)
    return s


def download_all(scene):
    return {name: scene.download(i) for i, name in enumerate(FIELD_NAMES)}


def assert_bit_equal(got, want, label=""):
    g = np.asarray(got, np.float32).view(np.uint32)
    w = np.asarray(want, np.float32).view(np.uint32)
    if not np.array_equal(g, w):
        bad = np.flatnonzero(g != w)
        i = int(bad[0])
        raise AssertionError(
            f"{label}: {bad.size} cells differ; first at idx {i}: got {got[i]!r} want {want[i]!r}")


def assert_state_equal(fields, st, label=""):
    for name in FIELD_NAMES:
        assert_bit_equal(fields[name], st.f[name], f"{label} field {name}")
