"""Golden fixtures (tests/golden/*.npz): stored whole-frame f32 bits of tiny
scenes.  The oracle must reproduce them (CPU); the HIP library must too (GPU)."""
import glob
import os

import numpy as np
import pytest

import oracle
from helpers import FIELD_NAMES, assert_bit_equal

FIXTURES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def _load(path):
    z = np.load(path, allow_pickle=False)
    return {k: z[k] for k in z.files}


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_oracle_reproduces_golden(path):
    g = _load(path)
    S = int(g["size"])
    st = oracle.State(S)
    for k in FIELD_NAMES:
        st.f[k] = g["init_" + k].copy()
    st.mask = g["mask"].copy()
    st.run(float(g["dt"]), float(g["diffusion"]), float(g["viscosity"]), int(g["iters"]),
           int(g["frames"]), src_d=g["src_d"], src_u=g["src_u"], src_v=g["src_v"])
    for k in FIELD_NAMES:
        assert_bit_equal(st.f[k], g["out_" + k], f"{os.path.basename(path)} {k}")


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_gpu_wavefront_reproduces_golden(path):
    from equilibrium_amd import _abi
    g = _load(path)
    S = int(g["size"])
    print(f"[case] golden {os.path.basename(path)}", flush=True)
    p = _abi.default_params()
    p.size = S
    p.delta_t = float(g["dt"])
    p.diffusion = float(g["diffusion"])
    p.viscosity = float(g["viscosity"])
    p.gs_iters = int(g["iters"])
    p.pressure_iters = int(g["iters"])
    s = _abi.Scene(p)
    s.set_mask(g["mask"])
    for i, k in enumerate(FIELD_NAMES):
        s.upload(i, g["init_" + k])
    s.set_frame_source(_abi.DENSITY, g["src_d"])
    s.set_frame_source(_abi.VX, g["src_u"])
    s.set_frame_source(_abi.VY, g["src_v"])
    s.step(int(g["frames"]))
    s.sync()
    for i, k in enumerate(FIELD_NAMES):
        assert_bit_equal(s.download(i), g["out_" + k], f"{os.path.basename(path)} {k}")
