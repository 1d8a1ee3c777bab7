"""Known answers for the oracle's whole `step` and `diffuse` (fluid.rs:276-298,
437-524), derived by hand, plus mutation checks showing that they catch the
plausible mistakes in the step's buffer roles.

`step` hands buffers between its stages (fluid.rs:438-523):

    1 diffuse(Row,     x=U0, x0=U,  visc)   GS guess = U0's stale content
    2 diffuse(Column,  x=V0, x0=V,  visc)
    3 project(u=U0, v=V0, p=U,  div=V)
    4 advect(Row,      d=U,  d0=U0, u=U0, v=V0)
    5 advect(Column,   d=V,  d0=V0, u=U0, v=V0)
    6 project(u=U,  v=V,  p=U0, div=V0)
    7 diffuse(Passive, x=Sc, x0=D,  diff)   GS guess = Sc (D's copy of last frame)
    8 advect(Passive,  d=D,  d0=Sc, u=U,  v=V)
    9 Sc := D

Two tiny scenes are worked through by hand below.  Every value is a dyadic
fraction with few bits (or one IEEE rounding of a quotient by 6, written as
such), so the f32 results are exact and the comparisons are bit-for-bit on
values.

* S = 4 with one sweep per solve.  Every interior cell touches the ring, so
  Row / Column set_boundaries overwrite the whole interior of a velocity
  field with negated ring values; the only interior values that survive
  come through the buffers that serve as p or div (Passive), and through
  density.  The scene catches a GS guess taken from x0, swapped p/div in
  either project, a dropped `S := D`, and `a` computed with S instead of S-2.
* S = 6 with zero sweeps (lin_solve does nothing, not even set_boundaries).
  Columns 2 and 3 of a Row field survive set_boundaries, so the velocity
  advection's gathers are visible.  The scene catches advect reading the
  wrong velocity buffers, and the project swaps again.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from oracle import PASSIVE

F32 = np.float32


def _grid(S, fn):
    """Row-major field with value fn(i, j) at cell (i, j)."""
    return np.array([fn(i, j) for j in range(S) for i in range(S)], F32)


def _expand2(v11, v21, v12, v22):
    """A 4x4 field after Passive set_boundaries with every cell fluid: ring
    cells copy their interior neighbour and each corner averages two equal
    ring cells, so cell (i, j) holds interior (clamp(i,1,2), clamp(j,1,2))."""
    v = {(1, 1): v11, (2, 1): v21, (1, 2): v12, (2, 2): v22}
    return _grid(4, lambda i, j: v[(min(max(i, 1), 2), min(max(j, 1), 2))])


def _mismatches(got, want, name):
    got = np.asarray(got, F32)
    want = np.asarray(want, F32)
    bad = np.flatnonzero(got != want)
    if bad.size == 0:
        return []
    k = int(bad[0])
    return [f"{name}: {bad.size} cells differ, first idx {k}: got {got[k]!r}, want {want[k]!r}"]


# ------------------------------------------------------------ S = 4, F = 1

def s4_one_sweep_errors():
    """S = 4, one GS sweep in all five solves, dt = 0.25, diffusion = 0.25,
    viscosity = 0.25, every interior cell fluid.

    Start: Fluid::new(4) (U = V = 1 everywhere, U0 = V0 = D = Sc = 0; size
    4 gets no density block because size/2-10 wraps), then U0(0,1) =
    U0(3,1) = 4 (left and right ring of row 1) and D(1,1) = 1.

    Derivation.  sf = S-2 = 2, dt0 = dt*sf = 0.5; density diffuse
    a = dt*k*sf*sf = 0.25, c = 1+4a = 2, 1/c = 0.5; project a = 1, 1/c = 0.25.
    Row set_boundaries at S = 4 sets (1,j) = -(0,j) and (2,j) = -(3,j);
    Column sets (i,1) = -(i,0) and (i,2) = -(i,3).

    1-2. U0 interior = -ring: row 1 -> (-4, -4), row 2 -> (-0, -0); V0
         interior -0.  (The GS values are overwritten.)
    3.  div1 = (-0.5*(((u(i+1)-u(i-1)) + v(j+1)) - v(j-1)))/S:
        (1,1): (-0.5*(-4-4))/4 = 1;  (2,1): (-0.5*(4-(-4)))/4 = -1;
        (1,2), (2,2): 0.  One pressure sweep from p = 0 in lexicographic
        order: p(1,1) = 1/4; p(2,1) = (-1 + p(1,1))/4 = -0.1875;
        p(1,2) = (0 + p(1,1))/4 = 0.0625; p(2,2) = (0 + (p(1,2)+p(2,1)))/4
        = -0.03125.  U := p1 (Passive ring), V := div1 (Passive ring).  The
        gradient on U0, V0 is overwritten by their set_boundaries.
    4-5. Advect writes U's and V's interior, then set_boundaries replaces it
        with the negated ring: U interior = -p1, V interior = -div1.
    6.  div2 from U, V: (1,1): -0.5*((0.1875-0.25) + 0 - 1)/4 = 0.1328125;
        (2,1): -0.5*((-0.1875+0.25) + 0 + 1)/4 = -0.1328125;
        (1,2): -0.5*((0.03125-0.0625) + 0 + 1)/4 = -0.12109375;
        (2,2): -0.5*((-0.03125+0.0625) + 0 - 1)/4 = 0.12109375.
        p2: (1,1) = 0.1328125/4 = 0.033203125;
        (2,1) = (-0.1328125 + 0.033203125)/4 = -0.02490234375;
        (1,2) = (-0.12109375 + 0.033203125)/4 = -0.02197265625;
        (2,2) = (0.12109375 + (-0.02197265625 - 0.02490234375))/4 = 0.0185546875.
        U0 := p2, V0 := div2; U and V keep -p1 / -div1 (set_boundaries).
    7.  Density sweep from Sc = 0 with rhs D: (1,1) = 1*0.5 = 0.5;
        (2,1) = (0.25*0.5)*0.5 = 0.0625; (1,2) = 0.0625;
        (2,2) = (0.25*(0.0625+0.0625))*0.5 = 0.015625; Passive ring.
    8.  Back-traces x = i - dt0*U, y = j - dt0*V:
        (1,1) -> (1.125, 1.5): 0.875*(0.5*0.5 + 0.5*0.0625)
                 + 0.125*(0.5*0.0625 + 0.5*0.015625) = 0.2509765625
        (2,1) -> (1.90625, 0.5): 0.09375*0.5 + 0.90625*0.0625 = 0.103515625
        (1,2) -> (1.03125, 2): 0.96875*0.0625 + 0.03125*0.015625 = 0.06103515625
        (2,2) -> (1.984375, 2): 0.015625*0.0625 + 0.984375*0.015625 = 0.016357421875
    9.  Sc := D.
    """
    S = 4
    st = oracle.State.new(S)
    st.f["vx0"][0 + 1 * S] = 4.0
    st.f["vx0"][3 + 1 * S] = 4.0
    st.f["density"][1 + 1 * S] = 1.0
    st.run(0.25, 0.25, 0.25, 1, 1, press_iters=1)

    p1 = (0.25, -0.1875, 0.0625, -0.03125)
    div1 = (1.0, -1.0, 0.0, 0.0)
    u = _expand2(*p1)
    v = _expand2(*div1)
    for (i, j), q in zip(((1, 1), (2, 1), (1, 2), (2, 2)), range(4)):
        u[i + j * S] = -p1[q]
        v[i + j * S] = -div1[q]
    d = _expand2(0.2509765625, 0.103515625, 0.06103515625, 0.016357421875)
    want = {
        "vx": u,
        "vy": v,
        "vx0": _expand2(0.033203125, -0.02490234375, -0.02197265625, 0.0185546875),
        "vy0": _expand2(0.1328125, -0.1328125, -0.12109375, 0.12109375),
        "density": d,
        "scratch": d,
    }
    errs = []
    for k, w in want.items():
        errs += _mismatches(st.f[k], w, "S=4 " + k)
    return errs


# ------------------------------------------------------------ S = 6, F = 0

def s6_no_sweep_errors():
    """S = 6, zero GS sweeps (lin_solve runs no sweep and no
    set_boundaries), dt = 0.25, all interior cells fluid.

    Start: Fluid::new(6) (U = V = 1, U0 = V0 = D = Sc = 0), then U0 = 0.75
    on interior columns 2 and 3 (0 elsewhere) and Sc(i, j) = i on every cell.

    Derivation.  dt0 = dt*(S-2) = 1.
    1-2. No-ops: U0, V0 keep their content.
    3.  div1 reads U0 as set: u(i+1)-u(i-1) = 0.75, 0.75, -0.75, -0.75 for
        columns 1..4 and the v terms are 0, so div1 = (-0.5*du)/6 =
        -0.0625, -0.0625, 0.0625, 0.0625 (exact), on every row.  V := div1
        with its Passive ring; U := p1 = 0 everywhere (ring copies, corners).
        U0 and V0 keep their values (set_boundaries writes +-0).
    4.  U from U0 at (i - U0, j): column 2 -> x = 1.25:
        0.75*U0(1) + 0.25*U0(2) = 0.1875; column 3 -> x = 2.25:
        0.75*0.75 + 0.25*0.75 = 0.75.  Row set_boundaries then sets columns
        1 and 4 to -(U's ring) = -0.
    5.  V gathers zeros from V0; Column set_boundaries sets row 1 to
        -(V's top ring) = -div1(i,1) and row 4 to -div1(i,4):
        0.0625, 0.0625, -0.0625, -0.0625.
    6.  div2 = (-0.5*s)/6 with s = du - q on rows 1 and 3 and du + q on rows
        2 and 4, where du = u(i+1)-u(i-1) of the new U (0.1875, 0.75,
        -0.1875, -0.75) and q = div1 of the column; V0 := div2 (Passive
        ring), U0 := p2 = 0 everywhere.
    7.  No-op.
    8.  Sc(i,j) = i is linear in x, so the bilinear gather returns the
        back-traced x: columns 1..4 -> 1, 2 - 0.1875 = 1.8125, 2.25, 4 (the
        y weights sum to 1 exactly).  Passive ring copies, corners.
    9.  Sc := D.
    """
    S = 6
    st = oracle.State.new(S)
    u0 = st.f["vx0"]
    for j in range(1, S - 1):
        u0[2 + j * S] = 0.75
        u0[3 + j * S] = 0.75
    st.f["scratch"][:] = _grid(S, lambda i, j: i)
    st.run(0.25, 0.5, 0.5, 0, 1, press_iters=0)

    interior = lambda i, j: 0 < i < S - 1 and 0 < j < S - 1
    ucol = {2: 0.1875, 3: 0.75}
    u = _grid(S, lambda i, j: ucol.get(i, 0.0) if interior(i, j) else 0.0)
    q = {1: -0.0625, 2: -0.0625, 3: 0.0625, 4: 0.0625}
    qc = lambda i: q[min(max(i, 1), 4)]

    def v_at(i, j):
        if j in (0, S - 1):
            return qc(i)                 # Passive ring of div1 (corners too)
        if i in (0, S - 1):
            return qc(i)
        return -q[i] if j in (1, S - 2) else 0.0

    du = {1: 0.1875, 2: 0.75, 3: -0.1875, 4: -0.75}
    div2 = {(i, j): F32(F32(-0.5) * F32(du[i] - q[i] if j % 2 == 1 else du[i] + q[i])) / F32(6)
            for i in range(1, S - 1) for j in range(1, S - 1)}
    cl = lambda t: min(max(t, 1), S - 2)
    dcol = {0: 1.0, 1: 1.0, 2: 1.8125, 3: 2.25, 4: 4.0, 5: 4.0}
    d = _grid(S, lambda i, j: dcol[i])
    want = {
        "vx": u,
        "vy": _grid(S, v_at),
        "vx0": np.zeros(S * S, F32),
        "vy0": _grid(S, lambda i, j: div2[(cl(i), cl(j))]),
        "density": d,
        "scratch": d,
    }
    errs = []
    for k, w in want.items():
        errs += _mismatches(st.f[k], w, "S=6 " + k)
    return errs


# ------------------------------------------------------ diffuse coefficient

def diffuse_coefficient_errors():
    """`diffuse` (fluid.rs:286-294): a = ((dt*k)*(S-2))*(S-2), c = 1+4a.

    S = 6, dt = k = 0.125: a = 0.015625*16 = 0.25, c = 2, 1/c = 0.5.  Guess
    0, rhs 1 at (1,1), Passive mode.
    One sweep: (1,1) = 1*0.5 = 0.5; (2,1) = 0.25*0.5*0.5 = 0.0625;
    (3,1) = 0.25*0.0625*0.5 = 0.0078125; (1,2) = 0.0625;
    (2,2) = 0.25*(0.0625+0.0625)*0.5 = 0.015625; ring (0,1) = (1,0) = 0.5.
    Two sweeps: the ring copies feed the second sweep, (1,1) =
    (1 + 0.25*(((0.0625 + 0.5) + 0.0625) + 0.5))*0.5 = 0.640625.
    """
    S = 6
    errs = []
    m = np.full(S * S, oracle.FLUID, np.uint8)
    for t in range(S):
        m[t] = m[t + (S - 1) * S] = m[t * S] = m[S - 1 + t * S] = oracle.WALL
    for iters, cells in ((1, {(1, 1): 0.5, (2, 1): 0.0625, (3, 1): 0.0078125, (1, 2): 0.0625,
                              (2, 2): 0.015625, (0, 1): 0.5, (1, 0): 0.5}),
                         (2, {(1, 1): 0.640625})):
        x = np.zeros(S * S, F32)
        x0 = np.zeros(S * S, F32)
        x0[1 + S] = 1.0
        oracle.diffuse(PASSIVE, x, x0, F32(0.125), S, F32(0.125), iters, m)
        for (i, j), w in cells.items():
            if x[i + j * S] != F32(w):
                errs.append(f"diffuse F={iters} ({i},{j}): got {x[i + j * S]!r}, want {w}")
    return errs


def test_step_s4_one_sweep_known_answer():
    errs = s4_one_sweep_errors()
    assert not errs, errs


def test_step_s6_no_sweep_known_answer():
    errs = s6_no_sweep_errors()
    assert not errs, errs


def test_diffuse_coefficient_known_answer():
    errs = diffuse_coefficient_errors()
    assert not errs, errs


# ---------------------------------------------------------------- mutations
# Each entry plants one mistake into a scratch copy of the oracle's source
# (text replacements in eqf_oracle_impl.h); the known answers above must
# reject every one of them.

IMPL = "eqf_oracle_impl.h"
MUTANTS = {
    # diffuse starts from a copy of its right-hand side, not the stale buffer
    "guess_from_rhs": [("    T sf = (T)(S - 2);\n    T a = dt * k * sf * sf;\n",
                        "    T sf = (T)(S - 2);\n    T a = dt * k * sf * sf;\n"
                        "    memcpy(x, x0, (size_t)(S * S) * sizeof(T));\n")],
    # diffuse of u with x and x0 exchanged
    "diffuse_u_roles_swapped": [("(ORA_ROW, U0, U, viscosity", "(ORA_ROW, U, U0, viscosity")],
    # velocity advection traced with U, V instead of U0, V0
    "advect_uv_velocity_buffers": [("(ORA_ROW, U, U0, U0, V0,", "(ORA_ROW, U, U0, U, V,"),
                                   ("(ORA_COL, V, V0, U0, V0,", "(ORA_COL, V, V0, U, V,")],
    # p and div exchanged in the first / second project
    "project1_p_div_swapped": [("ora_project)(U0, V0, U, V,", "ora_project)(U0, V0, V, U,")],
    "project2_p_div_swapped": [("ora_project)(U, V, U0, V0,", "ora_project)(U, V, V0, U0,")],
    # S := D dropped
    "no_scratch_copy": [("    memcpy(Sc, D, (size_t)(S * S) * sizeof(T));\n", "")],
    # diffuse coefficient with the ring-inclusive size
    "diffuse_a_uses_size": [("    T sf = (T)(S - 2);\n    T a = dt * k", "    T sf = (T)S;\n    T a = dt * k")],
    # c = 1 + 2a
    "diffuse_c_wrong": [("(T)1.0 + (T)4.0 * a", "(T)1.0 + (T)2.0 * a")],
    # density advected with the pre-projection velocities
    "density_advect_u0v0": [("(ORA_PASSIVE, D, Sc, U, V,", "(ORA_PASSIVE, D, Sc, U0, V0,")],
    # back-trace scaled by the ring-inclusive size
    "advect_dt0_uses_size": [("T dt0 = dt * (T)(S - 2);", "T dt0 = dt * (T)S;")],
    # divergence and gradient scaled by the interior width
    "project_scale_interior": [("    T sz = (T)S;\n", "    T sz = (T)(S - 2);\n")],
}


def _build_mutant(tmp, name, edits):
    src_dir = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    d = os.path.join(tmp, name)
    os.makedirs(d)
    for f in ("eqf_oracle.c", IMPL, "eqf_oracle_par.h"):
        shutil.copy(os.path.join(src_dir, f), d)
    path = os.path.join(d, IMPL)
    text = open(path).read()
    for old, new in edits:
        assert text.count(old) == 1, f"mutant {name}: pattern not found once: {old!r}"
        text = text.replace(old, new)
    open(path, "w").write(text)
    so = os.path.join(d, "liboracle_mut.so")
    subprocess.run(["gcc", "-O1", "-fPIC", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
                    "-shared", "-o", so, os.path.join(d, "eqf_oracle.c"), "-lm"], check=True)
    return oracle.load_library(so)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_known_answers_reject_planted_mistake(name, tmp_path):
    handle = _build_mutant(str(tmp_path), name, MUTANTS[name])
    with oracle.using(handle):
        errs = s4_one_sweep_errors() + s6_no_sweep_errors() + diffuse_coefficient_errors()
    assert errs, f"mutant {name} passes every known answer"
