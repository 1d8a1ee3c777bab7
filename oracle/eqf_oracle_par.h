/*
 * Parallel red-black step of the CPU oracle, instantiated by eqf_oracle.c
 * with T=float (ORA_SFX=_f32) and T=double (ORA_SFX=_f64).
 *
 * TEST INFRASTRUCTURE ONLY (see eqf_oracle.c).
 *
 * Same arithmetic, evaluation order and results as ora_step(..., ORA_RB) of
 * eqf_oracle_impl.h: every loop parallelised here writes disjoint cells and
 * reads only cells that the loop does not write (red cells read black
 * neighbours and vice versa; advect rows are independent; set_boundaries
 * phase 1 writes only fluid cells or ring cells and reads only wall or
 * interior cells).  Passive phase 1 is written per ring row/column ("copy if
 * that column/row holds a NoWall cell"), which is what the cell scan
 * produces.  tests/test_oracle.py checks equality with the sequential
 * red-black oracle for both types.
 */

#define ORA_CAT2(a, b) a##b
#define ORA_CAT(a, b) ORA_CAT2(a, b)
#define ORA_FN(name) ORA_CAT(name, ORA_SFX)

static void ORA_FN(par_set_bnd)(int mode, T *x, int64_t S, const uint8_t *m, const uint8_t *colf,
                                const uint8_t *rowf) {
    if (mode == ORA_PASSIVE) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 1; i < S - 1; i++) {
            if (colf[i]) {
                x[i] = x[i + S];
                x[i + (S - 1) * S] = x[i + (S - 2) * S];
            }
            if (rowf[i]) {
                x[i * S] = x[1 + i * S];
                x[S - 1 + i * S] = x[S - 2 + i * S];
            }
        }
    } else {
#pragma omp parallel for schedule(static)
        for (int64_t j = 1; j < S - 1; j++)
            for (int64_t i = 1; i < S - 1; i++) ORA_FN(ora_bnd_cell)(mode, x, S, m, i, j);
    }
    x[0] = (T)0.5 * (x[1] + x[S]);
    x[(S - 1) * S] = (T)0.5 * (x[1 + (S - 1) * S] + x[(S - 2) * S]);
    x[S - 1] = (T)0.5 * (x[S - 2] + x[S - 1 + S]);
    x[S - 1 + (S - 1) * S] = (T)0.5 * (x[S - 2 + (S - 1) * S] + x[S - 1 + (S - 2) * S]);
}

static void ORA_FN(par_lin_solve)(int mode, T *x, const T *x0, T a, T c, int64_t S, int64_t iters,
                                  const uint8_t *m, const uint8_t *colf, const uint8_t *rowf) {
    T cr = (T)1.0 / c;
    for (int64_t k = 0; k < iters; k++) {
        for (int color = 0; color < 2; color++) {
#pragma omp parallel for schedule(static)
            for (int64_t j = 1; j < S - 1; j++)
                for (int64_t i = 1 + ((j + 1 + color) & 1); i < S - 1; i += 2) {
                    size_t q = (size_t)(i + j * S);
                    x[q] = (x0[q] + a * (((x[q + 1] + x[q - 1]) + x[q + S]) + x[q - S])) * cr;
                }
        }
        ORA_FN(par_set_bnd)(mode, x, S, m, colf, rowf);
    }
}

static void ORA_FN(par_project)(T *u, T *v, T *p, T *div, int64_t S, int64_t iters, const uint8_t *m,
                                const uint8_t *colf, const uint8_t *rowf) {
    T sz = (T)S;
#pragma omp parallel for schedule(static)
    for (int64_t j = 1; j < S - 1; j++)
        for (int64_t i = 1; i < S - 1; i++) {
            size_t q = (size_t)(i + j * S);
            div[q] = (T)(-0.5) * (u[q + 1] - u[q - 1] + v[q + S] - v[q - S]) / sz;
            p[q] = (T)0.0;
        }
    ORA_FN(par_set_bnd)(ORA_PASSIVE, div, S, m, colf, rowf);
    ORA_FN(par_set_bnd)(ORA_PASSIVE, p, S, m, colf, rowf);
    ORA_FN(par_lin_solve)(ORA_PASSIVE, p, div, (T)1.0, (T)4.0, S, iters, m, colf, rowf);
#pragma omp parallel for schedule(static)
    for (int64_t j = 1; j < S - 1; j++)
        for (int64_t i = 1; i < S - 1; i++) {
            size_t q = (size_t)(i + j * S);
            u[q] -= (T)0.5 * (p[q + 1] - p[q - 1]) * sz;
            v[q] -= (T)0.5 * (p[q + S] - p[q - S]) * sz;
        }
    ORA_FN(par_set_bnd)(ORA_ROW, u, S, m, colf, rowf);
    ORA_FN(par_set_bnd)(ORA_COL, v, S, m, colf, rowf);
}

static void ORA_FN(par_advect)(int mode, T *d, const T *d0, const T *u, const T *v, int64_t S, T dt,
                               const uint8_t *m, const uint8_t *colf, const uint8_t *rowf) {
    T dt0 = dt * (T)(S - 2);
    T sf = (T)S;
#pragma omp parallel for schedule(static)
    for (int64_t j = 1; j < S - 1; j++) {
        for (int64_t i = 1; i < S - 1; i++) {
            T x = (T)i - dt0 * u[i + j * S];
            T y = (T)j - dt0 * v[i + j * S];
            x = ORA_FN(ora_clamp)(x, (T)0.5, sf - (T)1.0);
            y = ORA_FN(ora_clamp)(y, (T)0.5, sf - (T)1.0);
            T i0 = ORA_FLOOR(x), i1 = i0 + (T)1.0, j0 = ORA_FLOOR(y), j1 = j0 + (T)1.0;
            T s1 = x - i0, s0 = (T)1.0 - s1, t1 = y - j0, t0 = (T)1.0 - t1;
            int64_t i0i = sat_u32((double)i0), i1i = sat_u32((double)i1);
            int64_t j0i = sat_u32((double)j0), j1i = sat_u32((double)j1);
            if (i1 >= sf || j1 >= sf) {
                d[i + j * S] = d[i - 1 + j * S];
                break;
            }
            d[i + j * S] = s0 * (t0 * d0[IDX(i0i, j0i, S)] + t1 * d0[IDX(i0i, j1i, S)]) +
                           s1 * (t0 * d0[IDX(i1i, j0i, S)] + t1 * d0[IDX(i1i, j1i, S)]);
        }
    }
    ORA_FN(par_set_bnd)(mode, d, S, m, colf, rowf);
}

void ORA_FN(ora_step_par)(T *D, T *Sc, T *U, T *V, T *U0, T *V0, int64_t S, T dt, T diffusion,
                          T viscosity, int64_t diff_iters, int64_t press_iters, const uint8_t *m) {
    uint8_t *colf = (uint8_t *)calloc((size_t)S, 1);
    uint8_t *rowf = (uint8_t *)calloc((size_t)S, 1);
    for (int64_t j = 0; j < S; j++)
        for (int64_t i = 0; i < S; i++)
            if (m[i + j * S] != ORA_WALL) {
                colf[i] = 1;
                rowf[j] = 1;
            }
    T sf = (T)(S - 2);
    T av = dt * viscosity * sf * sf, ad = dt * diffusion * sf * sf;
    ORA_FN(par_lin_solve)(ORA_ROW, U0, U, av, (T)1.0 + (T)4.0 * av, S, diff_iters, m, colf, rowf);
    ORA_FN(par_lin_solve)(ORA_COL, V0, V, av, (T)1.0 + (T)4.0 * av, S, diff_iters, m, colf, rowf);
    ORA_FN(par_project)(U0, V0, U, V, S, press_iters, m, colf, rowf);
    ORA_FN(par_advect)(ORA_ROW, U, U0, U0, V0, S, dt, m, colf, rowf);
    ORA_FN(par_advect)(ORA_COL, V, V0, U0, V0, S, dt, m, colf, rowf);
    ORA_FN(par_project)(U, V, U0, V0, S, press_iters, m, colf, rowf);
    ORA_FN(par_lin_solve)(ORA_PASSIVE, Sc, D, ad, (T)1.0 + (T)4.0 * ad, S, diff_iters, m, colf, rowf);
    ORA_FN(par_advect)(ORA_PASSIVE, D, Sc, U, V, S, dt, m, colf, rowf);
    memcpy(Sc, D, (size_t)(S * S) * sizeof(T));
    free(colf);
    free(rowf);
}

#undef ORA_FN
#undef ORA_CAT
#undef ORA_CAT2
