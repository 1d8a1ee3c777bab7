// eqf_layout.h -- scene layout shared by the host runtime and the kernels.
//
// A scene of `size` S = N + 2 (N interior rows/columns plus the 1-cell wall
// ring of the reference, fluid.rs:552-570) keeps its six fields resident in
// HBM.  Two layouts exist:
//
//  * CYCLIC (exact wavefront mode).  Interior rows are grouped in strips of 64
//    rows (strip s = rows 64s+1 .. 64s+64, lane r = row-in-strip).  With the
//    period P = N rounded up to a multiple of 16, cell (c, j), 1 <= c, j <= N,
//    s = (j-1)/64, r = (j-1)%64, is stored at slot t = (c-1+r) mod P of strip s:
//    offset (s*P + t)*64 + r (the P-N slots per lane that hold no column are
//    unused).  A wave that owns a strip walks the Gauss-Seidel anti-diagonal
//    wavefront (lane r at column t-r+1 at step t) and every step touches one
//    256-byte slot row of each operand: fully coalesced, no LDS staging; with
//    P a multiple of 16 a block of 16 steps never wraps, so all its addresses
//    are one base plus constants.  The ring follows the interior: top[c]
//    (row 0), bottom[c] (row S-1), left[j] (column 0), right[j] (column S-1),
//    4*S floats; corners live in top[0], top[S-1], bottom[0], bottom[S-1].
//
//  * ROWMAJOR (red-black mode).  Plain `x + y*S` like the reference
//    (fluid.rs:31-35), ring included.
//
// Field order follows the Fluid struct (fluid.rs:59-77).
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#define EQF_HD __host__ __device__
#else
#define EQF_HD
#endif

namespace eqf {

enum Field { F_D = 0, F_SC = 1, F_U = 2, F_V = 3, F_U0 = 4, F_V0 = 5, NFIELDS = 6 };
enum Mode { M_ROW = 0, M_COL = 1, M_PASSIVE = 2 };  // Orientation, fluid.rs:22-29
enum Layout { L_CYCLIC = 0, L_ROWMAJOR = 1 };

// Boundary facts of an interior cell, computed while the boundary codes are
// built (k_codes).
enum CodeBits : uint32_t {
  C_FLUID = 1u,     // cell is ContainerWall::NoWall
  C_LWALL = 2u,     // (c-1, j) is a wall
  C_RWALL = 4u,     // (c+1, j) is a wall
  C_DWALL = 8u,     // (c, j+1) is a wall ("down" in fluid.rs:146)
  C_UWALL = 16u,    // (c, j-1) is a wall ("up" in fluid.rs:145)
  C_FIXUP = 32u,    // cyclic only: (c, j-1) is fluid, sits in the strip above,
                    // (c, j) is a wall and (c, j-2) is not: the wave that
                    // computes (c, j) writes the AdjustColumn value of (c, j-1)
  C_COLF = 64u,     // column c holds at least one NoWall cell (Passive ring copy)
};

// Per-cell rule byte of a ROWMAJOR scene (one byte per cell, ring included,
// cell (x, y) at y * rm_code_stride(S) + x), read by the red-black sweeps:
//   bit 0     interior cell (updated by the sweeps)
//   bits 1-2  AdjustRow: 0 keep, 1 -(left), 2 -(right)  (right wins, fluid.rs:151-164)
//   bits 3-4  AdjustColumn: 0 keep, 1 -(up), 2 -(down)  (up wins, fluid.rs:165-177)
//   bits 5-7  Passive, ring cells whose row / column holds a NoWall cell
//             (fluid.rs:179-187): 0 none, 1 copy the left, 2 the right, 3 the
//             upper, 4 the lower neighbour
enum RuleBits : uint32_t {
  R_INTERIOR = 1u,
  R_ROW_SHIFT = 1, R_COL_SHIFT = 3, R_PAS_SHIFT = 5,
  R_NEG_FIRST = 1u, R_NEG_SECOND = 2u,  // left / up, right / down
  R_COPY_LEFT = 1u, R_COPY_RIGHT = 2u, R_COPY_UP = 3u, R_COPY_DOWN = 4u,
};

// Boundary selectors of a CYCLIC scene: two bits per interior cell and mode
// (array of mode m at byte m * cyc_code_bytes, cells at cyc_code_word), telling
// the wavefront what set_boundaries of that mode makes of the cell:
//   M_ROW:     0 keep, 1 -(left neighbour), 2 -(right neighbour)  (right wins, fluid.rs:151-164)
//   M_COL:     0 keep, 1 -(upper neighbour), 2 -(lower neighbour) (up wins, fluid.rs:165-177);
//              a lower neighbour in the next strip is left to SEL_FIXUP there;
//              3 (SEL_FIXUP): the cell is a wall at the top row of a strip and
//              the fluid cell above it takes -(this cell) (and nothing above that)
//   M_PASSIVE: 1 if the column holds a NoWall cell (ring copy, fluid.rs:179-187)
enum Sel : uint32_t { SEL_KEEP = 0, SEL_NEG1 = 1, SEL_NEG2 = 2, SEL_FIXUP = 3 };

struct SceneDesc {
  float* fld;            // NFIELDS fields, `fstride` floats apart
  const float* src;      // per-frame sources (density, vx, vy) or nullptr
  const uint8_t* code;   // cyclic: 3 selector arrays (Sel); row-major: RuleBits at x + y*rm_code_stride(S)
  const uint8_t* rowf;   // [S] row j holds a NoWall cell
  const int2* bnd;       // (dst, src) field offsets: ROW list, COL list, PASSIVE list
  int64_t fstride;       // floats per field
  int S, N, ns;          // size, interior size, strips
  int P;                 // cyclic period (N rounded up to a multiple of 16)
  int nbnd[3];           // list lengths per mode
  int diff_iters, press_iters;
  float a_visc, cr_visc; // diffuse coefficients for velocity (fluid.rs:286-294, 311)
  float a_diff, cr_diff; // diffuse coefficients for density
  float dt0;             // dt * (S-2)         (fluid.rs:390)
  float sf;              // (float)S           (fluid.rs:345, 396)
  float sfm1;            // (float)S - 1.0f    (fluid.rs:403)
  int has_src;           // frame sources to add: bit 0 density (and scratch), 1 vx, 2 vy
  int prio;              // wave issue priority 0..3 in a batch launch (largest scenes highest)
  int frame_mult;        // frames per launch frame (> 1: a small scene of a server lane takes
                         // several frames in one tick of the lane's largest); 0 or 1: one
  int stream_l2;         // exact mode: >= 1: the one-pass phases and the last sweep stream of
                         // a strip load and store non-temporally (fields too large to find in
                         // L2 again, or a launch shared with such scenes); 2: so do the x0
                         // loads of a strip's first stream under the ring schedule (launches
                         // of many large scenes)
};

// Exact-mode wavefront schedules (compile-time per kernel, so that every
// access offset is an immediate): prefetch depth gs_depth steps, blocks of
// GS_B steps (per-block set-up, one selector word), a workgroup barrier every
// BS steps (BS divides GS_B), and whether x0 goes from sweep stream to sweep
// stream through LDS (cyc_x0_ring).
//   GS_WIDE: BS = 16, x0 read from HBM by every stream (the fewest barriers:
//            lone scenes and small launches, issue-bound)
//   GS_RING: BS = 8, x0 handed on in LDS (a shorter stream lag that lets the
//            rings fit, and fewer HBM bytes: large launches, bound by the
//            memory system)
//   GS_FINE: BS = 4 and prefetch depth 4, x0 handed on in LDS (the shortest
//            stream lag: the most sweep streams per strip for small scenes)
enum GsSched { GS_WIDE = 0, GS_RING = 1, GS_FINE = 2 };
constexpr int GS_B = 16;
EQF_HD constexpr int gs_sync(int sched) { return sched == GS_FINE ? 4 : sched == GS_RING ? 8 : 16; }
EQF_HD constexpr int gs_depth(int sched) { return sched == GS_FINE ? 4 : 8; }
EQF_HD constexpr bool gs_has_ring(int sched) { return sched != GS_WIDE; }

EQF_HD constexpr int cyc_period(int N) { return (N + 15) & ~15; }

EQF_HD constexpr int round_up(int v, int q) { return (v + q - 1) / q * q; }
// Wavefront lags for blocks of B steps, barriers every BS steps and prefetch
// depth D <= BS, multiples of B so that every stream's blocks start on
// a multiple of B global steps.  A value handed from one wave to another
// must be read at least BS steps after it was stored (a barrier lies in
// between whatever the alignment), and reads are issued D steps before use:
//  * strip s+1 trails strip s by L >= 64 + BS + D steps: its lane 0 reads strip
//    s's raw last-row value of column c, stored at strip-local step c + 62,
//    at step c - 1 - D of its own;
//  * sweep stream j+1 of a strip trails stream j by LAG >= L - 62 + BS + D: its
//    last lane reads the strip below's post-boundary row-0 value of column c,
//    which stream j of that strip stores at its step c (AdjustRow /
//    AdjustColumn store a cell's final value one step after computing it),
//    at step c + 62 - D of its own (the AdjustColumn fix-up is stored a step
//    earlier).
EQF_HD constexpr int cyc_strip_lag(int B, int BS, int D) { return round_up(64 + BS + D, B); }
EQF_HD constexpr int cyc_stream_lag(int B, int BS, int D) {
  return round_up(cyc_strip_lag(B, BS, D) - 62 + BS + D, B);
}

// Exact mode keeps the strips' hand-off rows in LDS: 2*ns rows of P floats
// (top ring, raw last rows, post-boundary first rows, bottom ring) plus a
// sink row after them that absorbs the stores of lanes with nothing to hand
// off (all such lanes of a store write the same float: no bank conflict).
// Rows are P + 1 floats apart, so that lane 63's edge read (the strip below)
// and the other lanes' shared read (the strip above) fall in different banks.
EQF_HD constexpr int cyc_lds_row(int P) { return P + 1; }
EQF_HD constexpr int gs_strip_lag(int sched) { return cyc_strip_lag(GS_B, gs_sync(sched), gs_depth(sched)); }
EQF_HD constexpr int gs_stream_lag(int sched) { return cyc_stream_lag(GS_B, gs_sync(sched), gs_depth(sched)); }
// Sweep streams per strip of a scene run by `waves` waves (lin_solves with
// fewer sweeps than that use fewer).
EQF_HD constexpr int cyc_streams(int waves, int ns, int P, int sched) {
  int m = waves / (ns > 0 ? ns : 1);
  const int mp = P / gs_stream_lag(sched);
  m = m < mp ? m : mp;
  return m > 1 ? m : 1;
}
// x0 (a lin_solve's right-hand side, the same in every sweep) goes from a
// sweep stream of a strip to the next through a ring of R slot rows in LDS
// instead of being read from HBM again: stream j stores the 64 values of its
// step tau at entry tau mod R, and stream j+1 loads them D steps before its
// own step tau (its first D entries before its first barrier), LAG - D >= BS
// steps, so a barrier, later.  Stream j overwrites the entry at its step
// tau + R, which with R >= LAG + BS lies after a later barrier than either
// load.  R is a multiple of BS: the BS steps between two barriers never wrap.
EQF_HD constexpr int cyc_x0_ring(int B, int BS, int D) { return round_up(cyc_stream_lag(B, BS, D) + BS, BS); }
// LDS layout (floats): the hand-off rows and their sink row, then (GS_RING,
// at cyc_x0_base) a sink of GS_B slot rows for x0 stores nobody reads and
// rings of R slot rows for the pairs of consecutive streams of the strips,
// strip by strip, as many as fit in a CU's 160 KiB (cyc_x0_rings); streams
// past them read x0 from HBM.
constexpr int LDS_FLOATS_MAX = 160 * 1024 / 4;
EQF_HD constexpr int cyc_x0_base(int ns, int P) { return (2 * ns + 1) * cyc_lds_row(P); }
EQF_HD constexpr int cyc_x0_rings(int B, int BS, int D, int ns, int P, int m) {
  const int R = cyc_x0_ring(B, BS, D);
  const int fit = (LDS_FLOATS_MAX - cyc_x0_base(ns, P) - B * 64) / (R * 64);
  const int want = ns * (m - 1);
  return want < fit ? want : (fit > 0 ? fit : 0);
}
// The one-pass phases reuse the space for the advect trigger columns (one int
// per row: 64 per strip).
EQF_HD constexpr int cyc_lds_floats(int ns, int P, int waves, int sched) {
  int h = cyc_x0_base(ns, P);
  if (gs_has_ring(sched)) {
    const int BS = gs_sync(sched), D = gs_depth(sched);
    h += GS_B * 64 + cyc_x0_rings(GS_B, BS, D, ns, P, cyc_streams(waves, ns, P, sched)) *
                         cyc_x0_ring(GS_B, BS, D) * 64;
  }
  return h > 64 * ns ? h : 64 * ns;
}
// Every exact-mode launch (N up to 1024, 1..16 waves, either schedule) fits
// the 160 KiB of LDS of a CU (N = P - 15 .. P share the strip count and period).
constexpr bool cyc_lds_fits() {
  for (int sched = GS_WIDE; sched <= GS_FINE; ++sched)
    for (int w = 1; w <= 16; ++w)
      for (int P = 16; P <= 1024; P += 16)
        if (cyc_lds_floats((P + 63) / 64, P, w, sched) > LDS_FLOATS_MAX) return false;
  return true;
}
static_assert(cyc_lds_fits(), "an exact-mode launch would need more LDS than a CU has");

// Ring offset inside a cyclic field.
EQF_HD inline int64_t cyc_ring(int P, int ns) { return (int64_t)ns * P * 64; }

// Offset of cell (x, y), 0 <= x, y <= S-1, inside one field.
EQF_HD inline int64_t cyc_off(int x, int y, int S, int P, int ns) {
  const int64_t ro = cyc_ring(P, ns);
  if (y == 0) return ro + x;
  if (y == S - 1) return ro + S + x;
  if (x == 0) return ro + 2 * S + y;
  if (x == S - 1) return ro + 3 * S + y;
  const int j = y - 1, s = j >> 6, r = j & 63;
  int t = x - 1 + r;
  if (t >= P) t -= P;
  return ((int64_t)s * P + t) * 64 + r;
}

// Boundary selectors of a cyclic scene: two bits per interior cell and mode,
// the sixteen slots 16k..16k+15 of a lane packed in one 32-bit word, so a
// wave reads the selectors of a whole 16-step wavefront block with one load:
// slot t of lane r in strip s is bits 2*(t & 15) of word (s*P/16 + t/16)*64 + r
// of the mode's array (cyc_code_bytes bytes per mode).
EQF_HD inline int64_t cyc_code_bytes(int P, int ns) { return (int64_t)ns * P * 16; }
EQF_HD inline int64_t cyc_code_word(int x, int y, int P) {
  const int j = y - 1, s = j >> 6, r = j & 63;
  int t = x - 1 + r;
  if (t >= P) t -= P;
  return ((int64_t)s * (P >> 4) + (t >> 4)) * 64 + r;
}
EQF_HD inline int cyc_code_shift(int x, int y, int P) {
  const int r = (y - 1) & 63;
  int t = x - 1 + r;
  if (t >= P) t -= P;
  return 2 * (t & 15);
}

EQF_HD inline int64_t cyc_field_floats(int S, int P, int ns) {
  int64_t n = cyc_ring(P, ns) + 4 * (int64_t)S;
  return (n + 63) & ~(int64_t)63;
}

EQF_HD inline int64_t rm_field_floats(int S) {
  int64_t n = (int64_t)S * S;
  return (n + 63) & ~(int64_t)63;
}

// Row-major rule bytes (RuleBits): rows padded to an even length, so that a
// lane's column pair (even first column) is one 2-byte-aligned load for any S.
EQF_HD inline int rm_code_stride(int S) { return S + (S & 1); }
EQF_HD inline int64_t rm_code_bytes(int S) { return (int64_t)S * rm_code_stride(S); }

}  // namespace eqf
