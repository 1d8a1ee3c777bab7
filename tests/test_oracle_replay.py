"""Brute-force pins of the oracle's set_boundaries, lexicographic
lin_solve, project, advect and whole frames: a naive per-cell replay of the
reference's loops in pure Python (one np.float32 operation at a time;
fluid.rs:133-190, 276-325, 329-372, 375-432 and 437-524) on tiny grids with
random interior walls, against the C oracle bit for bit.

The replay follows the reference literally: set_boundaries visits every cell
(j outer, i inner), skips walls, reads clamped neighbours, writes in place,
and averages the corners after the scan; lin_solve updates every interior
cell (walls included) in lexicographic order with the reference's sum order
and 1/c, then applies set_boundaries after every sweep.  A wrong index, a
swapped precedence (right wins in AdjustRow, up wins in AdjustColumn) or an
update order that reads other neighbours' new values in the oracle's loops
fails it.
"""
import numpy as np
import pytest

import oracle

WALL, FLUID = 0, 1
ROW, COL, PASSIVE = 0, 1, 2
F32 = np.float32


def _mask(S, rng, wall_frac):
    m = np.full(S * S, FLUID, np.uint8)
    for j in range(S):
        for i in range(S):
            if i in (0, S - 1) or j in (0, S - 1) or rng.random() < wall_frac:
                m[i + j * S] = WALL
    return m


def replay_set_bnd(mode, x, m, S):
    """fluid.rs:133-190 (manage_single_cell_boundary over every cell, then the corners)."""
    def k(i, j):
        return i + j * S

    def cl(v):
        return min(max(v, 0), S - 1)

    for j in range(S):
        for i in range(S):
            if m[k(i, j)] == WALL:
                continue
            up, down = (i, cl(j - 1)), (i, cl(j + 1))
            left, right = (cl(i - 1), j), (cl(i + 1), j)
            if mode == ROW:
                if m[k(*left)] == WALL:
                    x[k(i, j)] = -x[k(*left)]
                if m[k(*right)] == WALL:
                    x[k(i, j)] = -x[k(*right)]
            elif mode == COL:
                if m[k(*down)] == WALL:
                    x[k(i, j)] = -x[k(*down)]
                if m[k(*up)] == WALL:
                    x[k(i, j)] = -x[k(*up)]
            else:
                x[k(i, 0)] = x[k(i, 1)]
                x[k(i, S - 1)] = x[k(i, S - 2)]
                x[k(0, j)] = x[k(1, j)]
                x[k(S - 1, j)] = x[k(S - 2, j)]
    h = F32(0.5)
    x[k(0, 0)] = h * (x[k(1, 0)] + x[k(0, 1)])
    x[k(0, S - 1)] = h * (x[k(1, S - 1)] + x[k(0, S - 2)])
    x[k(S - 1, 0)] = h * (x[k(S - 2, 0)] + x[k(S - 1, 1)])
    x[k(S - 1, S - 1)] = h * (x[k(S - 2, S - 1)] + x[k(S - 1, S - 2)])


def replay_lin_solve(mode, x, x0, a, c, m, S, iters):
    """fluid.rs:301-325."""
    a = F32(a)
    cr = F32(1.0) / F32(c)
    for _ in range(iters):
        for j in range(1, S - 1):
            for i in range(1, S - 1):
                t = i + j * S
                s = ((x[t + 1] + x[t - 1]) + x[t + S]) + x[t - S]
                x[t] = (x0[t] + a * s) * cr
        replay_set_bnd(mode, x, m, S)


def _same(got, want):
    return np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("mode", [ROW, COL, PASSIVE])
def test_set_bnd_equals_naive_scan_on_random_masks(mode):
    rng = np.random.default_rng(70 + mode)
    for case in range(40):
        S = int(rng.integers(3, 10))
        m = _mask(S, rng, wall_frac=(0.0, 0.15, 0.35, 0.6)[case % 4])
        x = rng.standard_normal(S * S).astype(F32)
        want = x.copy()
        replay_set_bnd(mode, want, m, S)
        got = x.copy()
        oracle.set_bnd(mode, got, S, m)
        assert _same(got, want), (mode, case, S)


@pytest.mark.parametrize("mode", [ROW, COL, PASSIVE])
def test_lexicographic_lin_solve_equals_naive_replay(mode):
    rng = np.random.default_rng(80 + mode)
    for case in range(24):
        S = int(rng.integers(3, 9))
        m = _mask(S, rng, wall_frac=(0.0, 0.2, 0.4)[case % 3])
        x = rng.standard_normal(S * S).astype(F32)
        x0 = rng.standard_normal(S * S).astype(F32)
        a = F32(rng.uniform(0.05, 3.0))
        c = F32(1.0) + F32(4.0) * a
        iters = int(rng.integers(1, 5))
        want = x.copy()
        replay_lin_solve(mode, want, x0, a, c, m, S, iters)
        got = x.copy()
        oracle.lin_solve(mode, got, x0, a, c, S, iters, m, order=oracle.LEX)
        assert _same(got, want), (mode, case, S, iters)


def test_replay_rejects_a_planted_order_change():
    """The replay distinguishes the reference's lexicographic order from a
    reversed scan (new right and lower neighbours instead of new left and
    upper ones), so an order slip in the oracle would show.  (A column-major
    scan would not: with a 5-point stencil it reads the same new and old
    neighbours as the row-major one, which is what lets the anti-diagonal
    wavefront reproduce the reference bit for bit.)"""
    rng = np.random.default_rng(90)
    S = 7
    m = _mask(S, rng, wall_frac=0.0)
    x = rng.standard_normal(S * S).astype(F32)
    x0 = rng.standard_normal(S * S).astype(F32)
    a, c = F32(1.0), F32(5.0)
    lex = x.copy()
    replay_lin_solve(PASSIVE, lex, x0, a, c, m, S, 1)
    rev = x.copy()
    cr = F32(1.0) / c
    for j in range(S - 2, 0, -1):
        for i in range(S - 2, 0, -1):
            t = i + j * S
            rev[t] = (x0[t] + a * (((rev[t + 1] + rev[t - 1]) + rev[t + S]) + rev[t - S])) * cr
    replay_set_bnd(PASSIVE, rev, m, S)
    assert not _same(lex, rev)
    colmajor = x.copy()
    for i in range(1, S - 1):
        for j in range(1, S - 1):
            t = i + j * S
            colmajor[t] = (x0[t] + a * (((colmajor[t + 1] + colmajor[t - 1]) + colmajor[t + S]) + colmajor[t - S])) * cr
    replay_set_bnd(PASSIVE, colmajor, m, S)
    assert _same(lex, colmajor)


def replay_project(u, v, p, div, m, S, iters):
    """fluid.rs:329-372."""
    sf = F32(S)
    for j in range(1, S - 1):
        for i in range(1, S - 1):
            t = i + j * S
            div[t] = (F32(-0.5) * (((u[t + 1] - u[t - 1]) + v[t + S]) - v[t - S])) / sf
            p[t] = F32(0.0)
    replay_set_bnd(PASSIVE, div, m, S)
    replay_set_bnd(PASSIVE, p, m, S)
    replay_lin_solve(PASSIVE, p, div, F32(1.0), F32(4.0), m, S, iters)
    for j in range(1, S - 1):
        for i in range(1, S - 1):
            t = i + j * S
            u[t] = u[t] - (F32(0.5) * (p[t + 1] - p[t - 1])) * sf
            v[t] = v[t] - (F32(0.5) * (p[t + S] - p[t - S])) * sf
    replay_set_bnd(ROW, u, m, S)
    replay_set_bnd(COL, v, m, S)


def replay_advect(mode, d, d0, u, v, dt, m, S):
    """fluid.rs:375-432, the row `break` included."""
    dtx = F32(dt) * F32(S - 2)
    sf = F32(S)
    for j in range(1, S - 1):
        for i in range(1, S - 1):
            t = i + j * S
            x = F32(i) - dtx * u[t]
            y = F32(j) - dtx * v[t]
            x = min(max(x, F32(0.5)), sf - F32(1.0))  # f32::clamp (no NaN in these inputs)
            y = min(max(y, F32(0.5)), sf - F32(1.0))
            i0 = np.floor(x)
            i1 = i0 + F32(1.0)
            j0 = np.floor(y)
            j1 = j0 + F32(1.0)
            s1 = x - i0
            s0 = F32(1.0) - s1
            t1 = y - j0
            t0 = F32(1.0) - t1
            if i1 >= sf or j1 >= sf:
                d[t] = d[t - 1]
                break
            a0, a1, b0, b1 = int(i0), int(i1), int(j0), int(j1)
            d[t] = s0 * (t0 * d0[a0 + b0 * S] + t1 * d0[a0 + b1 * S]) + \
                s1 * (t0 * d0[a1 + b0 * S] + t1 * d0[a1 + b1 * S])
    replay_set_bnd(mode, d, m, S)


def test_project_equals_naive_replay():
    rng = np.random.default_rng(100)
    for case in range(24):
        S = int(rng.integers(4, 9))
        m = _mask(S, rng, wall_frac=(0.0, 0.2)[case % 2])
        u = rng.standard_normal(S * S).astype(F32)
        v = rng.standard_normal(S * S).astype(F32)
        p = rng.standard_normal(S * S).astype(F32)
        div = rng.standard_normal(S * S).astype(F32)
        iters = int(rng.integers(1, 5))
        want = [a.copy() for a in (u, v, p, div)]
        replay_project(*want, m, S, iters)
        got = [a.copy() for a in (u, v, p, div)]
        oracle.project(*got, S, iters, m, order=oracle.LEX)
        for g, w, name in zip(got, want, ("u", "v", "p", "div")):
            assert _same(g, w), (case, S, name)


@pytest.mark.parametrize("mode", [ROW, COL, PASSIVE])
def test_advect_equals_naive_replay_with_breaks(mode):
    rng = np.random.default_rng(110 + mode)
    breaks = 0
    for case in range(30):
        S = int(rng.integers(4, 10))
        m = _mask(S, rng, wall_frac=(0.0, 0.25)[case % 2])
        scale = (0.5, 3.0, 12.0)[case % 3]  # large velocities back-trace past the last column
        u = (scale * rng.standard_normal(S * S)).astype(F32)
        v = (scale * rng.standard_normal(S * S)).astype(F32)
        d = rng.standard_normal(S * S).astype(F32)
        d0 = rng.standard_normal(S * S).astype(F32)
        dt = F32(rng.uniform(0.02, 0.3))
        want = d.copy()
        replay_advect(mode, want, d0, u, v, dt, m, S)
        got = d.copy()
        oracle.advect(mode, got, d0, u, v, S, dt, m)
        assert _same(got, want), (mode, case, S)
        sf = F32(S)
        for j in range(1, S - 1):  # count rows whose back-trace hits the clamp's upper end
            for i in range(1, S - 1):
                t = i + j * S
                x = min(max(F32(i) - F32(dt) * F32(S - 2) * u[t], F32(0.5)), sf - F32(1.0))
                y = min(max(F32(j) - F32(dt) * F32(S - 2) * v[t], F32(0.5)), sf - F32(1.0))
                if np.floor(x) + F32(1.0) >= sf or np.floor(y) + F32(1.0) >= sf:
                    breaks += 1
                    break
    assert breaks > 10  # the row break is exercised


def replay_diffuse(mode, x, x0, k, dt, m, S, iters):
    """fluid.rs:276-298: a = dt * k * (S-2) * (S-2), c = 1 + 4a."""
    sf = F32(S - 2)
    a = ((F32(dt) * F32(k)) * sf) * sf
    replay_lin_solve(mode, x, x0, a, F32(1.0) + F32(4.0) * a, m, S, iters)


def replay_frame(f, m, S, dt, diff, visc, iters, sd, su, sv):
    """The frame loop's add-then-step (renderer_helpers.rs:55-59: sources
    added as add_density / add_velocity do, fluid.rs:120-131), then step
    (fluid.rs:437-524) with its buffer roles."""
    D, Sc, U, V, U0, V0 = (f[k] for k in oracle.FIELDS)
    D += sd
    Sc += sd
    U += su
    V += sv
    replay_diffuse(ROW, U0, U, visc, dt, m, S, iters)
    replay_diffuse(COL, V0, V, visc, dt, m, S, iters)
    replay_project(U0, V0, U, V, m, S, iters)
    replay_advect(ROW, U, U0, U0, V0, dt, m, S)
    replay_advect(COL, V, V0, U0, V0, dt, m, S)
    replay_project(U, V, U0, V0, m, S, iters)
    replay_diffuse(PASSIVE, Sc, D, diff, dt, m, S, iters)
    replay_advect(PASSIVE, D, Sc, U, V, dt, m, S)
    Sc[:] = D


def test_whole_frames_equal_naive_replay():
    """Three frames with sources, obstacles and random fields: the oracle's
    step (its buffer roles included) equals the replay composed from the
    component replays above, all six fields bit for bit."""
    rng = np.random.default_rng(120)
    for case in range(6):
        S = int(rng.integers(6, 11))
        st = oracle.State.new(S)
        st.mask[:] = _mask(S, rng, wall_frac=(0.0, 0.15)[case % 2])
        for k in oracle.FIELDS:
            st.f[k][:] = rng.standard_normal(S * S).astype(F32)
        sd = rng.standard_normal(S * S).astype(F32)
        su = (2.0 * rng.standard_normal(S * S)).astype(F32)
        sv = (2.0 * rng.standard_normal(S * S)).astype(F32)
        dt, diff, visc = F32(rng.uniform(0.05, 0.2)), F32(rng.uniform(0, 1e-2)), F32(rng.uniform(0, 1e-2))
        iters = int(rng.integers(1, 5))
        want = {k: st.f[k].copy() for k in oracle.FIELDS}
        for _ in range(3):
            replay_frame(want, st.mask, S, dt, diff, visc, iters, sd, su, sv)
        st.run(dt, diff, visc, iters, 3, src_d=sd, src_u=su, src_v=sv)
        for k in oracle.FIELDS:
            assert _same(st.f[k], want[k]), (case, S, k)


def replay_init(S, times):
    """Fluid::new (times=1) / Fluid::default (times=2): init_velocities,
    init_density, init_walls (fluid.rs:84-117, 527-571, 602-606) through
    add_velocity / add_density (fluid.rs:120-131), with the reference's u32
    range bounds (S/2 - 10 wraps below 20: an empty range) and idx!'s
    clamped coordinates (fluid.rs:31-35)."""
    f = {k: np.zeros(S * S, F32) for k in oracle.FIELDS}
    mask = np.full(S * S, FLUID, np.uint8)

    def idx(x, y):
        return min(max(x, 0), S - 1) + min(max(y, 0), S - 1) * S

    for _ in range(times):
        for j in range(S):
            for i in range(S):
                f["vx"][idx(i, j)] += F32(1.0)
                f["vy"][idx(i, j)] += F32(1.0)
        for j in range(S):
            for i in range(S):
                f["density"][idx(i, j)] += F32(0.0)
                f["scratch"][idx(i, j)] += F32(0.0)
        lo, hi = (S // 2 - 10) % (1 << 32), S // 2 + 10
        for j in range(lo, hi + 1):
            for i in range(lo, hi + 1):
                f["density"][idx(i, j)] += F32(0.9)
                f["scratch"][idx(i, j)] += F32(0.9)
        for i in range(S):
            mask[idx(i, 0)] = mask[idx(i, S - 1)] = WALL
        for j in range(S):
            mask[idx(0, j)] = mask[idx(S - 1, j)] = WALL
    return f, mask


@pytest.mark.parametrize("times", [1, 2])
def test_init_equals_naive_replay_for_many_sizes(times):
    for S in list(range(3, 44)) + [64, 128]:
        f, mask = replay_init(S, times)
        st = oracle.State.new(S, times)
        assert np.array_equal(st.mask, mask), S
        for k in oracle.FIELDS:
            assert _same(st.f[k], f[k]), (S, k)
