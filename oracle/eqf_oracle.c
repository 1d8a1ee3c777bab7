/*
 * eqf_oracle.c -- CPU oracle for the Stam frame step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker or
 * as the timed CPU baseline.  The product path (libeqfluid.so) never links,
 * calls or falls back to it.
 *
 * It restates src/simulation/fluid.rs of the reference (a Rust crate that
 * cannot be built here: no rustc/cargo, crates.io dependencies not vendored)
 * in C, compiled with -ffp-contract=off -fno-fast-math so that f32 results are
 * those of rustc's IEEE f32 code (rustc never contracts a*b+c into an FMA).
 *
 * Functions (f32 and f64 instantiations, see eqf_oracle_impl.h):
 *   ora_set_bnd      fluid.rs:133-189, 252-272
 *   ora_lin_solve    fluid.rs:301-325 (+ red-black order, build extension)
 *   ora_diffuse      fluid.rs:276-298
 *   ora_project      fluid.rs:330-375
 *   ora_advect       fluid.rs:378-432
 *   ora_step         fluid.rs:437-524
 *   ora_add_sources  fluid.rs:120-131 applied to every cell
 *   ora_run          renderer_helpers.rs:54-59 frame loop
 *   ora_init         fluid.rs:93-110, 527-570, 602-606 (+ Default, 83-89)
 *   ora_fill_rect    fluid.rs:610-619
 *   ora_rect_valid   obstacle.rs:74-87
 *   ora_step_par_f32 / ora_step_par_f64
 *                    parallel (OpenMP) red-black step, identical results to
 *                    ora_step_f32/_f64(order=ORA_RB), for large grids.
 *   ora_render_rgba  renderer_helpers.rs:143-165 (RenderingListener::render_image
 *                    pixel rule, `f32 as u8` with Rust's saturating semantics)
 *
 * Parity pinning: see tests/test_oracle.py (closed forms, brute force,
 * reference known answers).  add_noise (fluid.rs:575-599) is not restated:
 * it depends on unseeded thread_rng and the noise/geo crates ("parity
 * unpinned", and disabled in every benchmark configuration).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORA_WALL 0  /* ContainerWall::DefaultWall (declared first, fluid.rs:14) */
#define ORA_FLUID 1 /* ContainerWall::NoWall (fluid.rs:16) */

enum { ORA_ROW = 0, ORA_COL = 1, ORA_PASSIVE = 2 }; /* Orientation, fluid.rs:22-29 */
enum { ORA_LEX = 0, ORA_RB = 1 };

static inline int64_t clampi(int64_t v, int64_t lo, int64_t hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

/* idx! macro (fluid.rs:31-35): clamp both coordinates into [0, size-1]. */
#define IDX(x, y, S) ((size_t)(clampi((x), 0, (S)-1) + clampi((y), 0, (S)-1) * (S)))

/* `f as u32` in Rust saturates: NaN and negatives give 0, too-large gives
 * u32::MAX (used at fluid.rs:417-418). */
static inline int64_t sat_u32(double v) {
    if (!(v > 0.0)) return 0;
    if (v >= 4294967295.0) return 4294967295LL;
    return (int64_t)v;
}

#define T float
#define ORA_SFX _f32
#define ORA_FLOOR floorf
#include "eqf_oracle_impl.h"
#undef T
#undef ORA_SFX
#undef ORA_FLOOR

#define T double
#define ORA_SFX _f64
#define ORA_FLOOR floor
#include "eqf_oracle_impl.h"
#undef T
#undef ORA_SFX
#undef ORA_FLOOR

/* ---------------------------------------------------------------- init -- */

/* `Fluid::new` (fluid.rs:93-110) runs `init` once; `Fluid::default`
 * (fluid.rs:83-89) runs it a second time.  `times` selects how many.
 * All fields start at 0.0 and every cell at NoWall (fluid.rs:96-103). */
void ora_init_f32(float *D, float *Sc, float *U, float *V, float *U0, float *V0, uint8_t *m,
                  uint32_t S, int times) {
    size_t n = (size_t)S * S;
    memset(D, 0, n * 4);
    memset(Sc, 0, n * 4);
    memset(U, 0, n * 4);
    memset(V, 0, n * 4);
    memset(U0, 0, n * 4);
    memset(V0, 0, n * 4);
    memset(m, ORA_FLUID, n);
    for (int t = 0; t < times; t++) {
        /* init_velocities (fluid.rs:542-548) */
        for (uint32_t j = 0; j < S; j++)
            for (uint32_t i = 0; i < S; i++) {
                size_t c = IDX((int64_t)i, (int64_t)j, (int64_t)S);
                U[c] += 1.0f;
                V[c] += 1.0f;
            }
        /* init_density (fluid.rs:527-539).  `size/2 - 10` is u32 arithmetic:
         * it wraps for size < 20 in a release build, leaving the block
         * empty; indices are clamped by idx!, so for size 20 and 21 the
         * last row/column of the block lands on the ring more than once. */
        for (uint32_t j = 0; j < S; j++)
            for (uint32_t i = 0; i < S; i++) {
                size_t c = IDX((int64_t)i, (int64_t)j, (int64_t)S);
                D[c] += 0.0f;
                Sc[c] += 0.0f;
            }
        uint32_t lo = (uint32_t)(S / 2u - 10u);
        uint32_t hi = S / 2u + 10u;
        for (uint64_t j = lo; j <= hi; j++)
            for (uint64_t i = lo; i <= hi; i++) {
                size_t c = IDX((int64_t)i, (int64_t)j, (int64_t)S);
                D[c] += 0.9f;
                Sc[c] += 0.9f;
            }
        /* init_walls (fluid.rs:552-570) */
        for (uint32_t i = 0; i < S; i++) {
            m[IDX((int64_t)i, 0, (int64_t)S)] = ORA_WALL;
            m[IDX((int64_t)i, (int64_t)S - 1, (int64_t)S)] = ORA_WALL;
        }
        for (uint32_t j = 0; j < S; j++) {
            m[IDX(0, (int64_t)j, (int64_t)S)] = ORA_WALL;
            m[IDX((int64_t)S - 1, (int64_t)j, (int64_t)S)] = ORA_WALL;
        }
    }
}

/* `fill_obstacle` (fluid.rs:610-619): half-open [x0,x1) x [y0,y1), each
 * index clamped by idx!. */
void ora_fill_rect(uint8_t *m, uint32_t S, int64_t x0, int64_t y0, int64_t x1, int64_t y1) {
    for (int64_t x = x0; x < x1; x++)
        for (int64_t y = y0; y < y1; y++) m[IDX(x, y, (int64_t)S)] = ORA_WALL;
}

/* `Rectangle::are_all_points_valid` (obstacle.rs:74-87). */
int ora_rect_valid(int64_t x0, int64_t y0, int64_t x1, int64_t y1, int64_t size) {
    return x0 != x1 && y0 != y1 && x0 < x1 && y0 < y1 && x0 < size && y0 < size && x1 < size &&
           y1 < size;
}

/* `Fluid::to_coordinate` (fluid.rs:112-117). */
void ora_to_coordinate(int64_t idx, int64_t size, int64_t *x, int64_t *y) {
    *x = idx % size;
    *y = idx / size;
}

int64_t ora_idx(int64_t x, int64_t y, int64_t size) { return (int64_t)IDX(x, y, size); }

/* ---------------------------------------- parallel red-black f32/f64 step --
 * ora_step_par_f32 / ora_step_par_f64: body in eqf_oracle_par.h. */
#define T float
#define ORA_SFX _f32
#define ORA_FLOOR floorf
#include "eqf_oracle_par.h"
#undef T
#undef ORA_SFX
#undef ORA_FLOOR

#define T double
#define ORA_SFX _f64
#define ORA_FLOOR floor
#include "eqf_oracle_par.h"
#undef T
#undef ORA_SFX
#undef ORA_FLOOR

/* Residual of the pressure Poisson system solved by project's lin_solve:
 * r = div + (sum of the 4 neighbours of p) - 4p over interior cells.
 * Returns max |r| and writes the RMS of r and of div. */
double ora_poisson_residual_f64(const double *p, const double *div, int64_t S, double *rms_r,
                                double *rms_div) {
    double mx = 0.0, s2 = 0.0, d2 = 0.0;
    for (int64_t j = 1; j < S - 1; j++)
        for (int64_t i = 1; i < S - 1; i++) {
            size_t q = (size_t)(i + j * S);
            double r = div[q] + (((p[q + 1] + p[q - 1]) + p[q + S]) + p[q - S]) - 4.0 * p[q];
            double ar = fabs(r);
            if (ar > mx) mx = ar;
            s2 += r * r;
            d2 += div[q] * div[q];
        }
    double n = (double)((S - 2) * (S - 2));
    if (rms_r) *rms_r = sqrt(s2 / n);
    if (rms_div) *rms_div = sqrt(d2 / n);
    return mx;
}

/* Rust `f as u8` for an f32 (saturating since Rust 1.45: NaN -> 0, values
 * below 0 -> 0, above 255 -> 255, otherwise truncation toward zero). */
static uint8_t rust_f32_as_u8(float f) {
  if (isnan(f) || f <= 0.0f) return 0;
  if (f >= 255.0f) return 255;
  return (uint8_t)(int)truncf(f);
}

/* renderer_helpers.rs:143-165: pixel (x, y) of the size x size image from
 * density[idx(x, y)] and cells_type[idx(x, y)]; rgba holds 4 bytes per pixel,
 * row-major (x + y*size), as image::ImageBuffer lays them out. */
void ora_render_rgba(const float *density, const uint8_t *mask, int64_t size, const uint8_t fluid[4],
                     const uint8_t world[4], const uint8_t obstacle[4], uint8_t *rgba) {
  for (int64_t y = 0; y < size; ++y)
    for (int64_t x = 0; x < size; ++x) {
      const int64_t e = (int64_t)IDX(x, y, size);
      uint8_t *px = rgba + 4 * (x + y * size);
      const float d = density[e];
      if (mask[e] == ORA_WALL) {
        memcpy(px, obstacle, 4);
      } else if (d != 0.0f && mask[e] == ORA_FLUID) {
        px[0] = rust_f32_as_u8(d * (float)fluid[0]);
        px[1] = fluid[1];
        px[2] = rust_f32_as_u8(d);
        px[3] = 1;
      } else {
        memcpy(px, world, 4);
      }
    }
}
