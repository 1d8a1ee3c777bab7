/*
 * Body of the CPU oracle, instantiated twice by eqf_oracle.c: once with
 * T=float (ORA_SFX=_f32) and once with T=double (ORA_SFX=_f64).
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates the reference's Stam step
 * (src/simulation/fluid.rs) as plain C so that tests can check the HIP path
 * against it.  Nothing in the product links or calls it.
 *
 * Every loop below follows the reference loop it cites, including its
 * evaluation order (left-associative sums, multiply by 1/c, divide by the
 * ring-inclusive size) because the f32 instantiation is the bit-exact
 * parity target of the wavefront GPU mode.
 */

#define ORA_CAT2(a, b) a##b
#define ORA_CAT(a, b) ORA_CAT2(a, b)
#define ORA_FN(name) ORA_CAT(name, ORA_SFX)

/* f32::clamp / f64::clamp: `if x < lo {x = lo}; if x > hi {x = hi}`; NaN passes
 * through unchanged (std clamp, used at fluid.rs:403-404). */
static inline T ORA_FN(ora_clamp)(T x, T lo, T hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}

/* `manage_single_cell_boundary` (fluid.rs:133-189) for one cell. */
static void ORA_FN(ora_bnd_cell)(int mode, T *x, int64_t S, const uint8_t *m,
                                 int64_t i, int64_t j) {
    if (m[IDX(i, j, S)] == ORA_WALL) return;
    int64_t up_y = clampi(j - 1, 0, S - 1);
    int64_t down_y = clampi(j + 1, 0, S - 1);
    int64_t left_x = clampi(i - 1, 0, S - 1);
    int64_t right_x = clampi(i + 1, 0, S - 1);
    switch (mode) {
    case ORA_ROW:
        if (m[IDX(left_x, j, S)] == ORA_WALL) x[IDX(i, j, S)] = -x[IDX(left_x, j, S)];
        if (m[IDX(right_x, j, S)] == ORA_WALL) x[IDX(i, j, S)] = -x[IDX(right_x, j, S)];
        break;
    case ORA_COL:
        if (m[IDX(i, down_y, S)] == ORA_WALL) x[IDX(i, j, S)] = -x[IDX(i, down_y, S)];
        if (m[IDX(i, up_y, S)] == ORA_WALL) x[IDX(i, j, S)] = -x[IDX(i, up_y, S)];
        break;
    default: /* Passive: only the edges (fluid.rs:179-187) */
        x[IDX(i, 0, S)] = x[IDX(i, 1, S)];
        x[IDX(i, S - 1, S)] = x[IDX(i, S - 2, S)];
        x[IDX(0, j, S)] = x[IDX(1, j, S)];
        x[IDX(S - 1, j, S)] = x[IDX(S - 2, j, S)];
        break;
    }
}

/* `set_boundaries` (fluid.rs:252-272): full scan, then the four corners. */
void ORA_FN(ora_set_bnd)(int mode, T *x, int64_t S, const uint8_t *m) {
    for (int64_t j = 0; j <= S - 1; j++)
        for (int64_t i = 0; i <= S - 1; i++) ORA_FN(ora_bnd_cell)(mode, x, S, m, i, j);
    x[IDX(0, 0, S)] = (T)0.5 * (x[IDX(1, 0, S)] + x[IDX(0, 1, S)]);
    x[IDX(0, S - 1, S)] = (T)0.5 * (x[IDX(1, S - 1, S)] + x[IDX(0, S - 2, S)]);
    x[IDX(S - 1, 0, S)] = (T)0.5 * (x[IDX(S - 2, 0, S)] + x[IDX(S - 1, 1, S)]);
    x[IDX(S - 1, S - 1, S)] = (T)0.5 * (x[IDX(S - 2, S - 1, S)] + x[IDX(S - 1, S - 2, S)]);
}

/* One Gauss-Seidel cell update, exactly as fluid.rs:315-320 evaluates it:
 * (x0 + a*(((x[i+1]+x[i-1])+x[j+1])+x[j-1])) * c_recip. */
static inline void ORA_FN(ora_gs_cell)(T *x, const T *x0, T a, T cr, int64_t S,
                                       int64_t i, int64_t j) {
    x[IDX(i, j, S)] = (x0[IDX(i, j, S)] +
                       a * (((x[IDX(i + 1, j, S)] + x[IDX(i - 1, j, S)]) + x[IDX(i, j + 1, S)]) +
                            x[IDX(i, j - 1, S)])) *
                      cr;
}

/* `lin_solve` (fluid.rs:301-325).  order == ORA_LEX is the reference's
 * lexicographic in-place sweep; order == ORA_RB updates the (i+j) even cells
 * of a sweep first and the odd cells second (same cell formula, same
 * set_boundaries after every full sweep). */
void ORA_FN(ora_lin_solve)(int mode, T *x, const T *x0, T a, T c, int64_t S, int64_t iters,
                           const uint8_t *m, int order) {
    T cr = (T)1.0 / c;
    for (int64_t k = 0; k < iters; k++) {
        if (order == ORA_RB) {
            for (int color = 0; color < 2; color++)
                for (int64_t j = 1; j < S - 1; j++)
                    for (int64_t i = 1 + ((j + 1 + color) & 1); i < S - 1; i += 2)
                        ORA_FN(ora_gs_cell)(x, x0, a, cr, S, i, j);
        } else {
            for (int64_t j = 1; j < S - 1; j++)
                for (int64_t i = 1; i < S - 1; i++) ORA_FN(ora_gs_cell)(x, x0, a, cr, S, i, j);
        }
        ORA_FN(ora_set_bnd)(mode, x, S, m);
    }
}

/* `diffuse` (fluid.rs:276-298): a = dt*k*(S-2)*(S-2), c = 1 + 4a. */
void ORA_FN(ora_diffuse)(int mode, T *x, const T *x0, T k, int64_t S, T dt, int64_t iters,
                         const uint8_t *m, int order) {
    T sf = (T)(S - 2);
    T a = dt * k * sf * sf;
    ORA_FN(ora_lin_solve)(mode, x, x0, a, (T)1.0 + (T)4.0 * a, S, iters, m, order);
}

/* `project` (fluid.rs:330-375). */
void ORA_FN(ora_project)(T *u, T *v, T *p, T *div, int64_t S, int64_t iters, const uint8_t *m,
                         int order) {
    T sz = (T)S;
    for (int64_t j = 1; j < S - 1; j++) {
        for (int64_t i = 1; i < S - 1; i++) {
            div[IDX(i, j, S)] = (T)(-0.5) *
                                (u[IDX(i + 1, j, S)] - u[IDX(i - 1, j, S)] + v[IDX(i, j + 1, S)] -
                                 v[IDX(i, j - 1, S)]) /
                                sz;
            p[IDX(i, j, S)] = (T)0.0;
        }
    }
    ORA_FN(ora_set_bnd)(ORA_PASSIVE, div, S, m);
    ORA_FN(ora_set_bnd)(ORA_PASSIVE, p, S, m);
    ORA_FN(ora_lin_solve)(ORA_PASSIVE, p, div, (T)1.0, (T)4.0, S, iters, m, order);
    for (int64_t j = 1; j < S - 1; j++) {
        for (int64_t i = 1; i < S - 1; i++) {
            u[IDX(i, j, S)] -= (T)0.5 * (p[IDX(i + 1, j, S)] - p[IDX(i - 1, j, S)]) * sz;
            v[IDX(i, j, S)] -= (T)0.5 * (p[IDX(i, j + 1, S)] - p[IDX(i, j - 1, S)]) * sz;
        }
    }
    ORA_FN(ora_set_bnd)(ORA_ROW, u, S, m);
    ORA_FN(ora_set_bnd)(ORA_COL, v, S, m);
}

/* `advect` (fluid.rs:378-432), including the `break` out of a row when the
 * back-traced point reaches the far edge (fluid.rs:420-423). */
void ORA_FN(ora_advect)(int mode, T *d, const T *d0, const T *u, const T *v, int64_t S, T dt,
                        const uint8_t *m) {
    T dt0 = dt * (T)(S - 2);
    T sf = (T)S;
    for (int64_t j = 1; j < S - 1; j++) {
        for (int64_t i = 1; i < S - 1; i++) {
            T x = (T)i - dt0 * u[IDX(i, j, S)];
            T y = (T)j - dt0 * v[IDX(i, j, S)];
            x = ORA_FN(ora_clamp)(x, (T)0.5, sf - (T)1.0);
            y = ORA_FN(ora_clamp)(y, (T)0.5, sf - (T)1.0);
            T i0 = ORA_FLOOR(x), i1 = i0 + (T)1.0;
            T j0 = ORA_FLOOR(y), j1 = j0 + (T)1.0;
            T s1 = x - i0, s0 = (T)1.0 - s1;
            T t1 = y - j0, t0 = (T)1.0 - t1;
            int64_t i0i = sat_u32((double)i0), i1i = sat_u32((double)i1);
            int64_t j0i = sat_u32((double)j0), j1i = sat_u32((double)j1);
            if (i1 >= sf || j1 >= sf) {
                d[IDX(i, j, S)] = d[IDX(i - 1, j, S)];
                break;
            }
            d[IDX(i, j, S)] = s0 * (t0 * d0[IDX(i0i, j0i, S)] + t1 * d0[IDX(i0i, j1i, S)]) +
                              s1 * (t0 * d0[IDX(i1i, j0i, S)] + t1 * d0[IDX(i1i, j1i, S)]);
        }
    }
    ORA_FN(ora_set_bnd)(mode, d, S, m);
}

/* `step` (fluid.rs:437-524) with the reference's buffer roles.  `diff_iters`
 * is used by the three diffuse calls and `press_iters` by the two project
 * calls; the reference passes `simulation_configs.frames` to all five. */
void ORA_FN(ora_step)(T *D, T *Sc, T *U, T *V, T *U0, T *V0, int64_t S, T dt, T diffusion,
                      T viscosity, int64_t diff_iters, int64_t press_iters, const uint8_t *m,
                      int order) {
    ORA_FN(ora_diffuse)(ORA_ROW, U0, U, viscosity, S, dt, diff_iters, m, order);
    ORA_FN(ora_diffuse)(ORA_COL, V0, V, viscosity, S, dt, diff_iters, m, order);
    ORA_FN(ora_project)(U0, V0, U, V, S, press_iters, m, order);
    ORA_FN(ora_advect)(ORA_ROW, U, U0, U0, V0, S, dt, m);
    ORA_FN(ora_advect)(ORA_COL, V, V0, U0, V0, S, dt, m);
    ORA_FN(ora_project)(U, V, U0, V0, S, press_iters, m, order);
    ORA_FN(ora_diffuse)(ORA_PASSIVE, Sc, D, diffusion, S, dt, diff_iters, m, order);
    ORA_FN(ora_advect)(ORA_PASSIVE, D, Sc, U, V, S, dt, m);
    memcpy(Sc, D, (size_t)(S * S) * sizeof(T));
}

/* Per-frame sources: `add_density(i, j, s[idx])` for every cell in row-major
 * order (density goes to D and to the scratch field, fluid.rs:120-124) and
 * `add_velocity(i, j, su[idx], sv[idx])` (fluid.rs:127-131).  Null arrays are
 * skipped. */
void ORA_FN(ora_add_sources)(T *D, T *Sc, T *U, T *V, const T *sd, const T *su, const T *sv,
                             int64_t S) {
    int64_t n = S * S;
    if (sd)
        for (int64_t c = 0; c < n; c++) {
            D[c] += sd[c];
            Sc[c] += sd[c];
        }
    if (su)
        for (int64_t c = 0; c < n; c++) U[c] += su[c];
    if (sv)
        for (int64_t c = 0; c < n; c++) V[c] += sv[c];
}

/* `nframes` frames of the reference frame loop (renderer_helpers.rs:54-59)
 * with the deterministic per-frame sources above in place of add_noise. */
void ORA_FN(ora_run)(T *D, T *Sc, T *U, T *V, T *U0, T *V0, int64_t S, T dt, T diffusion,
                     T viscosity, int64_t diff_iters, int64_t press_iters, const uint8_t *m,
                     int order, const T *sd, const T *su, const T *sv, int64_t nframes) {
    for (int64_t f = 0; f < nframes; f++) {
        ORA_FN(ora_add_sources)(D, Sc, U, V, sd, su, sv, S);
        ORA_FN(ora_step)(D, Sc, U, V, U0, V0, S, dt, diffusion, viscosity, diff_iters,
                         press_iters, m, order);
    }
}

#undef ORA_FN
#undef ORA_CAT
#undef ORA_CAT2
