"""Writes the golden fixtures of tiny scenes (tests/golden/*.npz).

The reference crate cannot be built or run here (Rust, no toolchain, crates.io
dependencies not vendored), so these vectors come from the CPU oracle, whose
own correctness is pinned by tests/test_oracle.py.  They freeze the exact f32
bits of whole frames (size 10, 18, 34, with obstacles, sources and a velocity
that triggers advect's row break) so that the oracle and the GPU path are both
checked against stored data.  Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

import oracle  # noqa: E402
from equilibrium_amd import scenes  # noqa: E402

CASES = [
    # name, N, iters, frames, rects, break velocity
    ("s10_f3", 8, 4, 3, 1, False),
    ("s18_f2_break", 16, 5, 2, 2, True),
    ("s34_f5", 32, 6, 5, 3, False),
    ("s34_f1_break", 32, 3, 1, 2, True),
]


def build_case(N, iters, frames, rects, brk):
    sc = scenes.stam_scene(N, scenes.derive_seed(77, N, frames), iters=iters, frames=frames,
                           n_rects=rects, diffusion=1e-3, viscosity=1e-3)
    if brk:
        S = N + 2
        u = np.full(S * S, 1.0, np.float32)
        u[(np.arange(S * S) // S) % 3 == 0] = -30.0
        v = np.full(S * S, 0.5, np.float32)
        v[(np.arange(S * S) % S) == 5] = -40.0
        sc["init_u"], sc["init_v"] = u, v
    return sc


def main():
    for name, N, iters, frames, rects, brk in CASES:
        sc = build_case(N, iters, frames, rects, brk)
        st = oracle.State.new(sc["size"])
        for r in sc["rects"]:
            st.fill_rect(*r)
        if sc["init_u"] is not None:
            st.f["vx"][:] = sc["init_u"]
            st.f["vy"][:] = sc["init_v"]
        mask = st.mask.copy()
        init = {k: st.f[k].copy() for k in oracle.FIELDS}
        st.run(sc["dt"], sc["diffusion"], sc["viscosity"], iters, frames,
               src_d=sc["src_d"], src_u=sc["src_u"], src_v=sc["src_v"])
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"), size=sc["size"], dt=np.float32(sc["dt"]),
            diffusion=np.float32(sc["diffusion"]), viscosity=np.float32(sc["viscosity"]),
            iters=iters, frames=frames, rects=np.array(sc["rects"], np.int64).reshape(-1, 4),
            mask=mask, src_d=sc["src_d"], src_u=sc["src_u"], src_v=sc["src_v"],
            **{"init_" + k: v for k, v in init.items()},
            **{"out_" + k: st.f[k] for k in oracle.FIELDS})
        print("wrote", name)


if __name__ == "__main__":
    main()
