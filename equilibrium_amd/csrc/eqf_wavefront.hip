// eqf_wavefront.hip -- exact (bit-for-bit) Stam frame step on gfx950.
//
// One workgroup advances one scene through whole frames (fluid.rs:437-524):
// every phase of the step runs inside the workgroup, separated by workgroup
// barriers, so a batch of scenes of any mix of sizes is ONE launch per frame
// (grid = number of scenes) and a single scene is the same kernel with grid 1.
//
// Fields use the CYCLIC strip layout (eqf_layout.h): wave w owns 64-row
// strips, lane r owns row 64s+r+1, and slot t of a strip holds column
// (t-r) mod P + 1 of every row (P = N rounded up to a multiple of 16; slots
// whose column would exceed N are padding).  The reference's lexicographic
// Gauss-Seidel (fluid.rs:312-324) is reproduced exactly by an anti-diagonal
// wavefront: at strip-local step tau lane r updates column (tau-r) mod P + 1
// of sweep (tau-r) div P (and idles on padding slots).  Its left/up neighbours are this sweep's values (own
// register / lane r-1 via DPP wave_shr), right/down neighbours are the
// previous sweep's values after set_boundaries (slot t+1 of the lane / lane
// r+1 via DPP wave_shl).  Every step reads and writes whole 256-byte slot
// rows: coalesced without LDS.  Consecutive sweeps run back to back and
// strip s+1 trails strip s by a fixed lag, so the schedule is static and
// only needs a workgroup barrier every B steps.  set_boundaries
// (fluid.rs:252-272) is fused: a cell's store is delayed by one step until
// the wall values it mirrors are known.
//
// Arithmetic is written in the reference's evaluation order and the file is
// compiled with -ffp-contract=off: rustc does not contract a*b+c, so the HIP
// compiler must not either.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "eqf_layout.h"

#define GAS __attribute__((address_space(1)))

namespace eqf {

typedef GAS float gf;
typedef GAS const float gcf;
typedef GAS const uint8_t gcu8;
typedef GAS const int gci;
typedef __attribute__((address_space(3))) float lf;  // LDS
typedef float f4 __attribute__((ext_vector_type(4)));  // elementwise IEEE adds, no contraction
typedef GAS f4 gf4;
typedef GAS const f4 gcf4;

// Lane shifts through DPP.  The lane without a source (lane 0 for wave_shr,
// lane 63 for wave_shl) keeps `edge` (bound_ctrl off).  A DPP move reads its
// source lane only if that lane is enabled in EXEC, so the move must run
// with the whole wave enabled: the empty asm pins it where it is written, out
// of any branch the compiler would otherwise build from a select around it.
__device__ __forceinline__ float from_prev_lane(float v, float edge = 0.0f) {  // lane i <- lane i-1 (wave_shr:1)
  float r = __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, edge),
                                                                  __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, false));
  asm volatile("" : "+v"(r));
  return r;
}
__device__ __forceinline__ float from_next_lane(float v, float edge = 0.0f) {  // lane i <- lane i+1 (wave_shl:1)
  float r = __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, edge),
                                                                  __builtin_bit_cast(int, v), 0x130, 0xf, 0xf, false));
  asm volatile("" : "+v"(r));
  return r;
}
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
// LDS pointer `bytes` past H.
__device__ __forceinline__ lf* lds_at(lf* H, int bytes) {
  return (lf*)((__attribute__((address_space(3))) char*)H + bytes);
}

typedef __amdgpu_buffer_rsrc_t rsrc_t;
// Raw buffer descriptor (gfx9 family: DATA_FORMAT 32 in word 3).
// Base and size are wave-uniform; readfirstlane keeps the descriptor in
// scalar registers even where the compiler computed them in vector registers.
__device__ __forceinline__ rsrc_t make_rsrc(GAS const void* base, uint32_t bytes) {
  const uint64_t a = (uint64_t)(uintptr_t)(const void*)base;
  const uint64_t au = (uint64_t)(uint32_t)uni((int)(uint32_t)a) | ((uint64_t)(uint32_t)uni((int)(a >> 32)) << 32);
  return __builtin_amdgcn_make_buffer_rsrc((void*)(uintptr_t)au, (short)0, uni((int)bytes), 0x00020000);
}
__device__ __forceinline__ float ldf(rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
// The same with a cache policy (the buffer instruction's aux bits; CP_NT
// marks the access non-temporal: its L2 line is the first to be evicted).
constexpr int CP_NT = 2;
template <int POL>
__device__ __forceinline__ float ldfp(rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, POL));
}
template <int POL>
__device__ __forceinline__ void stfp(float v, rsrc_t r, int voff, int soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, POL);
}
__device__ __forceinline__ uint32_t ldu32(rsrc_t r, int voff, int soff) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
}
__device__ __forceinline__ void stf(float v, rsrc_t r, int voff, int soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 0);
}

// f32::clamp (NaN passes through), fluid.rs:403-404
__device__ __forceinline__ float rclamp(float x, float lo, float hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}

// `f as u32` followed by idx!'s clamp into [0, S-1] (fluid.rs:417-418, 31-35)
__device__ __forceinline__ int sat_idx(float f, int S) {
  if (!(f > 0.0f)) return 0;
  if (f >= (float)(S - 1)) return S - 1;
  return (int)f;
}

__device__ __forceinline__ void corners(gf* ring, int S, int N) {  // fluid.rs:265-271
  ring[0] = 0.5f * (ring[1] + ring[2 * S + 1]);
  ring[S] = 0.5f * (ring[S + 1] + ring[2 * S + N]);
  ring[S - 1] = 0.5f * (ring[N] + ring[3 * S + 1]);
  ring[2 * S - 1] = 0.5f * (ring[S + N] + ring[3 * S + N]);
}

// set_boundaries as a list pass: x[dst] = -x[src] (ROW/COL) or x[dst] = x[src]
// (PASSIVE), then the four corners.  Up to two fields share one pass.
__device__ __forceinline__ void bnd_lists(const SceneDesc& sc, int mode0, gf* x0, int mode1, gf* x1) {
  const int S = sc.S, N = sc.N;
  const int ro = sc.P * 64 * sc.ns;
  const int off0 = mode0 == M_ROW ? 0 : mode0 == M_COL ? sc.nbnd[0] : sc.nbnd[0] + sc.nbnd[1];
  const int off1 = mode1 == M_ROW ? 0 : mode1 == M_COL ? sc.nbnd[0] : sc.nbnd[0] + sc.nbnd[1];
  const int n0 = sc.nbnd[mode0], n1 = x1 ? sc.nbnd[mode1] : 0;
  gci* bnd = (gci*)sc.bnd;
  const float sg0 = mode0 == M_PASSIVE ? 1.0f : -1.0f;
  const float sg1 = mode1 == M_PASSIVE ? 1.0f : -1.0f;
  for (int e = threadIdx.x; e < n0 + n1; e += blockDim.x) {
    const bool second = e >= n0;
    const int k = second ? off1 + e - n0 : off0 + e;
    gf* x = second ? x1 : x0;
    x[bnd[2 * k]] = (second ? sg1 : sg0) * x[bnd[2 * k + 1]];
  }
  __syncthreads();
  if (threadIdx.x == 0) corners(x0 + ro, S, N);
  if (threadIdx.x == 1 && x1) corners(x1 + ro, S, N);
  __syncthreads();
}

// Bits i of [lo, hi) within a block of B <= 16 steps.
__device__ __forceinline__ uint32_t step_bits(int lo, int hi, int B) {
  lo = min(max(lo, 0), B);
  hi = min(max(hi, 0), B);
  return ((1u << hi) - 1u) & ~((1u << lo) - 1u);
}

// ------------------------------------------------------------------------
// Gauss-Seidel wavefront (lin_solve, fluid.rs:301-325) with set_boundaries
// fused.  MODE selects the set_boundaries rule.  D = prefetch depth (steps),
// B = steps between workgroup barriers.
//
// Lane r of strip s works at position u = tau - r of the strip-local step
// tau = g - s*L: column (u mod P) + 1 of sweep u div P (idle when that column
// is past N: the P-N padding slots).  P and L are multiples of 16 and B
// divides 16, so a block of B steps never wraps: every stream of a block is
// one uniform slot base (soffset) plus a compile-time immediate, and the
// per-lane conditions of the block (active steps, first/last column) are
// bit masks and step indices computed once per block.
//
// Strips hand values to each other through H, in LDS (2*ns rows of P floats,
// indexed by the READER's slot so reads need no wrap):
//   row 0            top ring values for strip 0's lane 0 (copied at entry)
//   row 1+s          raw last-row values of strip s, read by lane 0 of s+1
//   row ns-1+s       post-boundary row-0 values of strip s, read by the last
//                    lane of strip s-1 in the next sweep
//   row 2ns-1        bottom ring values for the last strip's last lane
//
// Ordering argument.  A value stored at global step g is read by another
// wave through a load issued at a step >= g + B (a barrier lies in between:
// LDS hand-offs complete before it, and on gfx950 the waves of a workgroup
// share one in-order vector-memory path, so a global load issued after the
// barrier observes global stores issued before it).
// With L >= 64+B+D: strip s+1 lane 0 reads strip s's raw last-row value of a
// column L-63-D >= B steps after it was stored; strip s's last lane reads
// strip s+1's post-boundary row-0 value (and the AdjustColumn fix-up) of the
// previous sweep P+61-D-L >= B steps after it was stored, which holds for
// P >= 80 (any scene with more than one strip).  Within a wave a cell's next
// sweep load is issued P-2-D >= 0 steps after its store.
//
// PASSIVE: set_boundaries(Passive) copies the interior cell next to the ring
// into the ring (rows/columns holding a NoWall cell, fluid.rs:179-187).  The
// sweep reads such a ring value as `xself`, the previous sweep's value of the
// cell being updated, and the ring itself is written once after the last
// sweep by the list pass.
template <int MODE, int D, int B, int BS, bool XR>
__device__ __forceinline__ void gs_wavefront(const SceneDesc& sc, lf* __restrict__ H, gf* __restrict__ X,
                                             gcf* __restrict__ X0, float a, float cr, int iters) {
  static_assert((B == 16 || B == 8) && D <= B, "blocks of 8 or 16 steps (within one selector word), at least "
                                              "as deep as the prefetch");
  static_assert(BS >= D && B % BS == 0, "barriers every BS steps: BS divides the block, at least the prefetch");
  // the prefetch registers form a ring indexed by the step within the block
  // modulo D, which carries over from block to block only if D divides B
  static_assert(B % D == 0, "prefetch depth must divide the block length");
  constexpr int L = cyc_strip_lag(B, BS, D);    // strip lag, a multiple of B
  constexpr int LAG = cyc_stream_lag(B, BS, D);  // sweep lag between streams of a strip
  // within a strip, stream j+1 loads a slot stream j stored BS + D + 2 steps
  // earlier at least (one step of delayed store, one of prefetch offset)
  static_assert(LAG >= BS + D + 2, "stream lag too short for the prefetch");
  static_assert(L >= 64 + BS + D && LAG >= L - 62 + BS + D && L % B == 0 && LAG % B == 0,
                "hand-off lags too short");
  const int N = sc.N, S = sc.S, ns = sc.ns, P = sc.P;
  const int lane = threadIdx.x & 63;
  const int wave = uni(threadIdx.x >> 6);
  const int RO = ns * P * 64;
  const int HP = cyc_lds_row(P);  // H row stride (floats)
  gf* __restrict__ ring = X + RO;
  // Sweep streams: with m waves per strip, stream j of a strip runs sweeps
  // j, j+m, j+2m, ... back to back, LAG steps behind stream j-1 (stream 0
  // starts its next sweep P >= m*LAG steps after its last one, so every
  // stream stays LAG behind its predecessor).  Sweep k+1 thus reads sweep
  // k's values LAG steps after they were written, while they are still in
  // cache, and the strip's chain is iters*P/m steps instead of iters*P.
  int m = min(uni(blockDim.x >> 6) / ns, P / LAG);
  m = max(1, min(m, iters));
  const int nq0 = (iters + m - 1) / m;  // sweeps of stream 0
  const int G = (ns - 1) * L + (m - 1) * LAG + nq0 * P + 65;
  const int Gr = (G + B - 1) / B * B;

  const bool has_strip = wave < ns * m;
  const int s = has_strip ? wave % ns : 0;  // strip
  const int jst = has_strip ? wave / ns : 0;  // stream
  const int nq = (iters - jst + m - 1) / m;  // sweeps of this stream
  // The last of several streams reads values that no later stream of this
  // sweep round reads again (the previous stream's output, x0) and writes
  // values read only by stream 0's next sweep, about P steps later: its
  // interior blocks load and store non-temporally, so that L2 keeps the
  // lines the other streams are about to re-read (batched mix: 12.97 ->
  // 12.32 ms per frame); only where SceneDesc::stream_l2 says so.
  const bool nt_stream = m > 1 && jst == m - 1 && sc.stream_l2 != 0;
  // x0 hand-off rings (cyc_x0_ring): stream jst > 0 loads x0 from ring
  // s*(m-1) + jst-1 in LDS instead of HBM, and every stream but the last
  // stores the x0 of each step into its own ring s*(m-1) + jst (the others
  // into a sink that nobody reads), for the rings that fit (cyc_x0_rings;
  // the launch sized LDS for at least as many).  Ring entries are slot rows
  // indexed by the stream-local step modulo R, so writer and reader agree.
  constexpr int R = cyc_x0_ring(B, BS, D);
  static_assert(!XR || (R % BS == 0 && R >= LAG + BS && LAG - D >= BS && R >= B),
                "x0 ring too short or read too early");
  const int nrings = XR ? cyc_x0_rings(B, BS, D, ns, P, m) : 0;
  const bool rin = XR && has_strip && jst > 0 && s * (m - 1) + jst - 1 < nrings;
  const bool rout = XR && has_strip && jst < m - 1 && s * (m - 1) + jst < nrings;
  const int xsink = cyc_x0_base(ns, P);
  const int xring = xsink + B * 64 + (s * (m - 1) + jst) * R * 64;  // this stream's ring (if rout)
  const int xo_ring = rout ? xring : -1;           // < 0: stores go to the sink
  const int xi_ring = rin ? xring - R * 64 : -1;  // < 0: x0 from HBM
  const int j = 64 * s + lane + 1;
  const bool valid = has_strip && j <= N;
  const int rl = min(63, N - 1 - 64 * s);  // last valid lane of the strip
  const bool first = lane == 0;
  const bool last = lane == rl;
  const bool has_prev = s > 0;
  const bool has_next = s < ns - 1;
  const int jc = min(j, N);
  const float ringL = ring[2 * S + jc];
  const float ringR = ring[3 * S + jc];
  const bool rowf = ((gcu8*)sc.rowf)[jc] != 0;

  // Values from outside the strip for the edge lanes.  Lane 0 reads its up
  // neighbour from an H row (top ring or the strip above).  The last valid
  // lane reads its down neighbour from an H row when it is lane 63 (the
  // strip below's post row, seeded here with its entry row 0 for sweep 0, or
  // the bottom ring); otherwise (last strip, N not a multiple of 64) the
  // bottom ring is mirrored into the unused row below it, lane rl+1 of the
  // strip, and arrives through the ordinary lane shift.  Entry ring values
  // are what Passive sweeps need: they read the changing ring through xself.
  if (has_strip && jst == 0 && s == 0)
    for (int t = lane; t < P; t += 64) H[t] = t < N ? ring[t + 1] : 0.0f;
  if (has_strip && jst == 0 && s > 0)  // entry row 0 of this strip: the "previous sweep" of sweep 0
    for (int t = lane; t < P; t += 64) {
      int c = t - 63;  // column - 1 of the reading lane 63 of the strip above at its slot t
      c += c < 0 ? P : 0;
      H[(ns - 1 + s) * HP + t] = c < N ? X[(s * P + c) * 64] : 0.0f;
    }
  if (has_strip && jst == 0 && s == ns - 1)
    for (int t = lane; t < P; t += 64) {
      int c = t - rl;
      c += c < 0 ? P : 0;
      if (rl == 63) {
        H[(2 * ns - 1) * HP + t] = c < N ? ring[S + c + 1] : 0.0f;
      } else {  // row N+1 at lane rl+1: slot (c - 1 + rl + 1) mod P holds column c
        int cm = c - 1;
        cm += cm < 0 ? P : 0;
        X[(s * P + t) * 64 + rl + 1] = cm < N ? ring[S + cm + 1] : 0.0f;
      }
    }
  __syncthreads();

  const uint32_t fbytes = (uint32_t)sc.fstride * 4u;
  const rsrc_t rX = make_rsrc(X, fbytes);
  const rsrc_t rX0 = make_rsrc(X0, fbytes);
  const int CB = (int)cyc_code_bytes(P, ns);
  const rsrc_t rC = make_rsrc((gcu8*)sc.code + MODE * CB, (uint32_t)CB);  // this mode's selectors
  const int OOB = (int)0x80000000;  // voffset that drops a store / reads 0
  const int vl = lane * 4;
  const int vls = valid ? vl : OOB;                   // stores: rows past N stay untouched
  const bool last_e = last && lane == 63;             // down neighbour from H
  const bool mir = !has_next && rl < 63 && lane == rl + 1;  // holds the mirrored bottom ring
  const bool has_mirror = !has_next && rl < 63;              // (uniform) the strip has a mirrored row
  const int vmir = (last && has_mirror) ? (rl + 1) * 4 : OOB;  // Passive: lane rl refreshes the mirror
  const bool top = first && !has_prev;                // up neighbour is the top ring
  const bool bot = last && !has_next;                 // down neighbour is the bottom ring
  const int up_row = has_prev ? s : 0;                // raw row of the strip above, or the top ring
  const int dn_row = has_next ? ns + s : 2 * ns - 1;  // post row of the strip below, or the bottom ring
  const int e_row = (last_e ? dn_row : up_row) * HP;  // H row of this lane's edge value (floats)
  // Sink row for the hand-off stores of lanes with nothing to hand off: the
  // lanes of a store all write the same float of it (same-address writes of
  // a wave do not conflict; no lane reads it).
  const int TR = 2 * ns * HP;
  const int raw_row = (1 + s) * HP;                    // this strip's raw last row (if has_next)
  const int post_row = (ns - 1 + s) * HP;              // this strip's post row 0 (if has_prev)
  const bool raw_ok = last && valid && has_next;      // lane 63 hands its raw values down
  const bool post_ok = first && has_prev;             // lane 0 hands its final values up
  // Byte offsets of each lane's hand-off rows (the sink row for the others),
  // to which a block adds its uniform slot offsets.
  const int e_b = e_row * 4;
  const int raw_b = (raw_ok ? raw_row : TR) * 4;
  const int post_b = (post_ok ? post_row : TR) * 4;
  const int sb = uni(s * P * 256);                        // strip base in X (bytes)
  const int cb = uni(s * P * 16);                        // strip base in the code array (bytes)
  const int fxb = uni(has_prev ? sb - P * 256 + 252 : 0);  // lane 63 of the strip above (bytes)
  const int itP = uni(nq * P);  // positions of this stream
  const int o64 = uni(64 % P);

  float fxs[D], fx0[D], fe[D], fe2[D];  // fe / fe2: the edge value, as the up / down shift's edge operand
  uint32_t fcw, fcw_n;  // selector words of this block and of the next one

  // Stream j of strip s idles for its first (s*L + j*LAG)/B blocks (every
  // lane before its first sweep): those blocks are barriers only.  Its first
  // working block then starts at stream-local step tau = 0, slot 0; stream
  // position u = tau - lane covers sweep jst + m*(u div P), column u mod P.
  const int nblk = Gr / B;
  const int head = has_strip ? s * (L / B) + jst * (LAG / B) : nblk;
  int tau = 0;       // strip-local step of the block's first step
  int tb = 0;        // its slot (uniform)
  int tbp = P - B;   // previous block's slot
  int rbk = 0;       // x0 ring entry of the block's first step (tau mod R)

  float res_prev = 0.f, xl_prev = 0.f, xu_prev = 0.f, xd_prev = 0.f, xself = 0.f;
  uint32_t code_prev = 0;
  bool prev_active = false, cN_prev = false;

  // One block of B steps: EDGE blocks predicate every store on the lane's
  // step lying in the sweep range; interior blocks (every lane's steps and
  // the pending step before them inside the range) only meet padding slots
  // (columns past N, never read as cell values) and store unconditionally.
  int blk = 0;
#pragma nounroll
  for (; blk < head; ++blk)
#pragma unroll
    for (int k = 0; k < B / BS; ++k) __syncthreads();
  if (has_strip) {
#pragma unroll
    for (int i = 0; i < D; ++i) {  // prefetch the first D steps (slot i)
      fxs[i] = ldf(rX, vl + (i + 1) * 256, sb);
      fx0[i] = rin ? H[xi_ring + i * 64 + lane] : ldf(rX0, vl + i * 256, sb);
      fe[i] = H[e_row + i];
      fe2[i] = H[e_row + i];
    }
    fcw = ldu32(rC, vl, cb);  // slots 0..15
    // xself of the first step: what the step before would have loaded (slot
    // 0, lane 0's own column-1 cell, from the previous sweep when jst > 0)
    xself = ldf(rX, vl, sb);
#pragma nounroll
    while (blk < nblk) {
      if (tau >= 64 && tau + B <= itP) {
        // interior blocks, by where x0 comes from (XS: 0 HBM, 1 the ring)
        // and whether it is handed on (XW)
        if (nt_stream && (rin || !XR)) {
#pragma nounroll
          do {
            {
              constexpr bool EDGE = false;
              constexpr int POL = CP_NT, XS = XR ? 1 : 0;
              constexpr bool XW = false;
              constexpr int X0POL = POL;
#include "eqf_gs_block.inc"
            }
            ++blk;
          } while (tau + B <= itP);
        } else if (XR && nt_stream) {  // the last stream of a strip past the rings that fit
#pragma nounroll
          do {
            {
              constexpr bool EDGE = false;
              constexpr int POL = CP_NT, XS = 0;
              constexpr bool XW = false;
              constexpr int X0POL = POL;
#include "eqf_gs_block.inc"
            }
            ++blk;
          } while (tau + B <= itP);
        } else if (XR && rin && rout) {
#pragma nounroll
          do {
            {
              constexpr bool EDGE = false;
              constexpr int POL = 0, XS = 1;
              constexpr bool XW = true;
              constexpr int X0POL = POL;
#include "eqf_gs_block.inc"
            }
            ++blk;
          } while (tau + B <= itP);
        } else if (XR && rin) {
#pragma nounroll
          do {
            {
              constexpr bool EDGE = false;
              constexpr int POL = 0, XS = 1;
              constexpr bool XW = false;
              constexpr int X0POL = POL;
#include "eqf_gs_block.inc"
            }
            ++blk;
          } while (tau + B <= itP);
        } else if (XR && rout && sc.stream_l2 > 1) {
          // the first stream of a strip with a ring, in a launch of many
          // large scenes: its x0 loads are the lines' last uses
#pragma nounroll
          do {
            {
              constexpr bool EDGE = false;
              constexpr int POL = 0, XS = 0;
              constexpr bool XW = true;
              constexpr int X0POL = CP_NT;
#include "eqf_gs_block.inc"
            }
            ++blk;
          } while (tau + B <= itP);
        } else if (XR && rout) {
#pragma nounroll
          do {
            {
              constexpr bool EDGE = false;
              constexpr int POL = 0, XS = 0;
              constexpr bool XW = true;
              constexpr int X0POL = POL;
#include "eqf_gs_block.inc"
            }
            ++blk;
          } while (tau + B <= itP);
        } else {
#pragma nounroll
          do {
            {
              constexpr bool EDGE = false;
              constexpr int POL = 0, XS = 0;
              constexpr bool XW = false;
              constexpr int X0POL = POL;
#include "eqf_gs_block.inc"
            }
            ++blk;
          } while (tau + B <= itP);
        }
      } else if (tau < itP + 64) {
        {
          // both x0 sources, selected per wave; stores to the ring or the sink
          constexpr bool EDGE = true;
          constexpr int POL = 0, XS = XR ? 2 : 0;
          constexpr bool XW = XR;
          constexpr int X0POL = POL;
#include "eqf_gs_block.inc"
        }
        ++blk;
      } else {
#pragma unroll
        for (int k = 0; k < B / BS; ++k) __syncthreads();
        ++blk;
      }
    }
  }

  if constexpr (MODE == M_PASSIVE) {
    // set_boundaries(Passive) of the last sweep (its corners included)
    bnd_lists(sc, M_PASSIVE, X, M_PASSIVE, nullptr);
  } else {
    if (threadIdx.x == 0) corners(ring, S, N);
    __syncthreads();
  }
}

// Global load / store, non-temporal (the L2 line is the first to be
// evicted) where the enclosing phase's NT is set (SceneDesc::stream_l2):
// the phases then stream their fields past L2 and leave the lines of the
// sweeps' re-reads alone.
#define ntl(p) (NT ? __builtin_nontemporal_load(p) : *(p))
#define nts(p, v)                                    \
  do {                                               \
    if constexpr (NT) __builtin_nontemporal_store((v), (p)); \
    else *(p) = (v);                                 \
  } while (0)
// The one-pass phases below walk each strip in blocks of 16 slots (P is a
// multiple of 16, so a block never wraps): all loads of a block are issued
// before any of its values is used, so a wave keeps ~50 loads in flight
// instead of waiting out one memory latency per slot.

// Offset, from the lane's edge base, of the value beyond the strip edge at
// slot t: lane 0 reads lane 63 of the strip above at the same column (slot
// t+63) or the top ring; the other lanes read lane 0 of the strip below at
// the last lane's column (slot t-rl) or the bottom ring.  Lane 0 uses its
// value as its up neighbour; lane 63 (a full strip) or lane rl+1 (the last,
// partial strip) hands its value to the last lane rl through the lane shift.
__device__ __forceinline__ int edge_index(int t, int lane, int rl, bool has_prev, bool has_next, int N, int P,
                                          int S) {
  int tl = t - rl;
  tl += tl < 0 ? P : 0;
  int tu = t + 63;
  tu -= tu >= P ? P : 0;
  return lane == 0 ? (has_prev ? tu * 64 : min(t, N - 1) + 1) : (has_next ? tl * 64 : S + min(tl, N - 1) + 1);
}

// The one-pass phases deal blocks of slots to all the waves of the workgroup:
// item k covers strip k / nb, slots [(k % nb) * BS, +BS) (nb = P / BS), and
// wave w takes items w, w + W, w + 2W, ...  Adjacent items are adjacent
// blocks of one strip.  An item loads the slot before its block itself.

// project, first half (fluid.rs:339-349): div and p = 0 on interior cells,
// then set_boundaries(Passive) on both (fluid.rs:351-352).
template <bool NT>
__device__ __forceinline__ void div_phase(const SceneDesc& sc, gcf* __restrict__ u, gcf* __restrict__ v,
                                          gf* __restrict__ div, gf* __restrict__ p) {
  const int N = sc.N, S = sc.S, ns = sc.ns, P = sc.P;
  const int lane = threadIdx.x & 63;
  const int PS = P * 64;
  gcf* uring = u + ns * PS;
  gcf* vring = v + ns * PS;
  const float sf = sc.sf;
  const int nb = P / 16;
  for (int item = uni(threadIdx.x >> 6); item < ns * nb; item += (blockDim.x >> 6)) {
    const int s = item / nb;
    const int tb = (item - s * nb) * 16;
    const int j = 64 * s + lane + 1;
    const bool row_ok = j <= N;
    const int jc = min(j, N);
    const int rl = min(63, N - 1 - 64 * s);
    const bool has_prev = s > 0, has_next = s < ns - 1;
    gcf* us = u + s * PS + lane;
    gcf* vs = v + s * PS + lane;
    // lane 0 reads the strip above (its lane 63) or the top ring, every other
    // lane the strip below (its lane 0) or the bottom ring
    gcf* eb = lane == 0 ? (has_prev ? v + (s - 1) * PS + 63 : vring) : (has_next ? v + (s + 1) * PS : vring);
    gf* ds = div + s * PS + lane;
    gf* ps = p + s * PS + lane;
    const float uL = uring[2 * S + jc], uR = uring[3 * S + jc];
    const int lane_nm1 = (lane + N - 1) % P;
    const int tp = tb == 0 ? P - 1 : tb - 1;  // the slot before the block
    float ub[17], vb[17], e[16];
    const int tn = tb + 16 == P ? 0 : tb + 16;
    const float um1 = us[tp * 64], vm1 = vs[tp * 64];
#pragma unroll
    for (int k = 0; k < 17; ++k) {
      const int t = k < 16 ? tb + k : tn;
      ub[k] = ntl(&us[t * 64]);
      vb[k] = ntl(&vs[t * 64]);
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) e[k] = eb[edge_index(tb + k, lane, rl, has_prev, has_next, N, P, S)];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int t = tb + k;
      int c0 = t - lane;
      c0 += c0 < 0 ? P : 0;
      const float um = k == 0 ? um1 : ub[k - 1];
      const float vm = k == 0 ? vm1 : vb[k - 1];
      // neighbours (fluid.rs:342-344)
      const float uplus = t == lane_nm1 ? uR : ub[k + 1];
      const float uminus = t == lane ? uL : um;
      // below: lane+1's value; the last lane gets e, its own (lane 63) or
      // lane rl+1's (which loaded the value below the last row)
      const float vplus = from_next_lane(lane == rl + 1 ? e[k] : vb[k + 1], e[k]);
      const float vminus = from_prev_lane(vm, e[k]);  // lane 0: the value above the strip
      const float d = -0.5f * (uplus - uminus + vplus - vminus) / sf;
      if (row_ok && c0 < N) {
        nts(&ds[t * 64], d);
        nts(&ps[t * 64], 0.0f);
      }
    }
  }
  __syncthreads();
  bnd_lists(sc, M_PASSIVE, div, M_PASSIVE, p);
}

// project, second half (fluid.rs:364-374)
template <bool NT>
__device__ __forceinline__ void grad_phase(const SceneDesc& sc, gf* __restrict__ u, gf* __restrict__ v,
                                           gcf* __restrict__ p) {
  const int N = sc.N, S = sc.S, ns = sc.ns, P = sc.P;
  const int lane = threadIdx.x & 63;
  const int PS = P * 64;
  gcf* pring = p + ns * PS;
  const float sf = sc.sf;
  const int nb = P / 8;  // 8 slots keep an item within 128 registers
  for (int item = uni(threadIdx.x >> 6); item < ns * nb; item += (blockDim.x >> 6)) {
    const int s = item / nb;
    const int tb = (item - s * nb) * 8;
    const int j = 64 * s + lane + 1;
    const bool row_ok = j <= N;
    const int jc = min(j, N);
    const int rl = min(63, N - 1 - 64 * s);
    const bool has_prev = s > 0, has_next = s < ns - 1;
    gcf* ps = p + s * PS + lane;
    gcf* eb = lane == 0 ? (has_prev ? p + (s - 1) * PS + 63 : pring) : (has_next ? p + (s + 1) * PS : pring);
    gf* us = u + s * PS + lane;
    gf* vs = v + s * PS + lane;
    const float pL = pring[2 * S + jc], pR = pring[3 * S + jc];
    const int lane_nm1 = (lane + N - 1) % P;
    const int tp = tb == 0 ? P - 1 : tb - 1;
    float pb[9], e[8], ub[8], vb[8];
    const int tn = tb + 8 == P ? 0 : tb + 8;
    const float pm1 = ps[tp * 64];
#pragma unroll
    for (int k = 0; k < 9; ++k) pb[k] = ntl(&ps[(k < 8 ? tb + k : tn) * 64]);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int t = tb + k;
      e[k] = eb[edge_index(t, lane, rl, has_prev, has_next, N, P, S)];
      ub[k] = ntl(&us[t * 64]);
      vb[k] = ntl(&vs[t * 64]);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int t = tb + k;
      int c0 = t - lane;
      c0 += c0 < 0 ? P : 0;
      const float pm = k == 0 ? pm1 : pb[k - 1];
      const float pplus = t == lane_nm1 ? pR : pb[k + 1];
      const float pminus = t == lane ? pL : pm;
      const float pdn = from_next_lane(lane == rl + 1 ? e[k] : pb[k + 1], e[k]);
      const float pup = from_prev_lane(pm, e[k]);
      if (row_ok && c0 < N) {
        nts(&us[t * 64], ub[k] - 0.5f * (pplus - pminus) * sf);
        nts(&vs[t * 64], vb[k] - 0.5f * (pdn - pup) * sf);
      }
    }
  }
  __syncthreads();
  bnd_lists(sc, M_ROW, u, M_COL, v);
}

// Bilinear back-trace of advect (fluid.rs:400-428) for cell (c, j).
struct Trace {
  int i0, i1, j0, j1;
  float s0, s1, t0, t1;
  bool trig;
};
__device__ __forceinline__ Trace back_trace(int c, int j, float uu, float vv, float dt0, float sf, float sfm1,
                                            int S) {
  Trace r;
  float x = (float)c - dt0 * uu;
  float y = (float)j - dt0 * vv;
  x = rclamp(x, 0.5f, sfm1);
  y = rclamp(y, 0.5f, sfm1);
  const float i0 = floorf(x), i1 = i0 + 1.0f;
  const float j0 = floorf(y), j1 = j0 + 1.0f;
  r.s1 = x - i0;
  r.s0 = 1.0f - r.s1;
  r.t1 = y - j0;
  r.t0 = 1.0f - r.t1;
  r.trig = (i1 >= sf) || (j1 >= sf);
  r.i0 = sat_idx(i0, S);
  r.i1 = sat_idx(i1, S);
  r.j0 = sat_idx(j0, S);
  r.j1 = sat_idx(j1, S);
  return r;
}
__device__ __forceinline__ float gather(gcf* d0, const Trace& r, int S, int P, int ns) {
  const float a = d0[cyc_off(r.i0, r.j0, S, P, ns)];
  const float b = d0[cyc_off(r.i0, r.j1, S, P, ns)];
  const float cc = d0[cyc_off(r.i1, r.j0, S, P, ns)];
  const float d = d0[cyc_off(r.i1, r.j1, S, P, ns)];
  return r.s0 * (r.t0 * a + r.t1 * b) + r.s1 * (r.t0 * cc + r.t1 * d);
}

// advect (fluid.rs:378-432) of one or two fields that share the velocity
// (advect of u and v in step, fluid.rs:469-489), then set_boundaries.  The
// `break` at fluid.rs:420-423 is reproduced per row: cells left of the first
// trigger column get the bilinear value, the trigger cell copies its left
// neighbour (the ring cell when it is column 1), cells to the right keep the
// buffer's previous content.  Pass A (one wave per strip) finds each row's
// first trigger column and leaves it in LDS (`trig_lds`, one int per row);
// pass B deals blocks of 8 slots to all waves.
template <bool NT>
__device__ __forceinline__ void advect_phase(const SceneDesc& sc, lf* __restrict__ trig_lds, int mode0,
                                             gf* __restrict__ d0out, gcf* __restrict__ d0in, int mode1,
                                             gf* __restrict__ d1out, gcf* __restrict__ d1in, gcf* __restrict__ u,
                                             gcf* __restrict__ v) {
  const int N = sc.N, S = sc.S, ns = sc.ns, P = sc.P;
  const int lane = threadIdx.x & 63;
  const int PS = P * 64;
  const float dt0 = sc.dt0, sf = sc.sf, sfm1 = sc.sfm1;
  const bool two = d1out != nullptr;
  // pass A: first trigger column of every row
  for (int s = uni(threadIdx.x >> 6); s < ns; s += (blockDim.x >> 6)) {
    const int j = 64 * s + lane + 1;
    const int jc = min(j, N);
    gcf* us = u + s * PS + lane;
    gcf* vs = v + s * PS + lane;
    int trig = N + 1;
    for (int tb = 0; tb < P; tb += 16) {
      float ub[16], vb[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        ub[k] = us[(tb + k) * 64];
        vb[k] = vs[(tb + k) * 64];
      }
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        int c0 = tb + k - lane;
        c0 += c0 < 0 ? P : 0;
        const Trace r = back_trace(c0 + 1, jc, ub[k], vb[k], dt0, sf, sfm1, S);
        if (c0 < N && r.trig && c0 + 1 < trig) trig = c0 + 1;
      }
    }
    trig_lds[64 * s + lane] = __builtin_bit_cast(float, trig);
  }
  __syncthreads();
  // pass B
  const int ro = ns * PS;
  const int nb = P / 8;  // 8 slots: velocity loads, then 32-64 gathers in flight
  for (int item = uni(threadIdx.x >> 6); item < ns * nb; item += (blockDim.x >> 6)) {
    const int s = item / nb;
    const int tb = (item - s * nb) * 8;
    const int j = 64 * s + lane + 1;
    const bool row_ok = j <= N;
    const int jc = min(j, N);
    const int trig = __builtin_bit_cast(int, trig_lds[64 * s + lane]);
    gcf* us = u + s * PS + lane;
    gcf* vs = v + s * PS + lane;
    const float left0 = d0out[ro + 2 * S + jc];
    const float left1 = two ? d1out[ro + 2 * S + jc] : 0.0f;
    gf* o0 = d0out + s * PS + lane;
    gf* o1 = (two ? d1out : d0out) + s * PS + lane;
    const int tp = tb == 0 ? P - 1 : tb - 1;
    const float um1 = us[tp * 64], vm1 = vs[tp * 64];
    float ub[8], vb[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      ub[k] = us[(tb + k) * 64];
      vb[k] = vs[(tb + k) * 64];
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int t = tb + k;
      int c0 = t - lane;
      c0 += c0 < 0 ? P : 0;
      const int c = c0 + 1;
      // c < trig: this cell; c == trig > 1: the left neighbour's value
      const bool at = c == trig;
      const float uu = at ? (k == 0 ? um1 : ub[k - 1]) : ub[k];
      const float vv = at ? (k == 0 ? vm1 : vb[k - 1]) : vb[k];
      const Trace tr = back_trace(at ? c - 1 : c, j, uu, vv, dt0, sf, sfm1, S);
      const float g0 = gather(d0in, tr, S, P, ns);
      const float g1 = two ? gather(d1in, tr, S, P, ns) : 0.0f;
      if (row_ok && c <= N && c <= trig) {
        const bool ring1 = at && c == 1;  // trigger in column 1: copy the ring cell (fluid.rs:421)
        nts(&o0[t * 64], ring1 ? left0 : g0);
        if (two) nts(&o1[t * 64], ring1 ? left1 : g1);
      }
    }
  }
  __syncthreads();
  bnd_lists(sc, mode0, d0out, mode1, d1out);
}

// Frame sources: add_density (fluid.rs:122-123) and add_velocity
// (fluid.rs:129-130) applied to every cell, interior and ring.
template <bool NT>
__device__ __forceinline__ void add_sources_phase(const SceneDesc& sc, gf* f) {
  const int n4 = (int)(sc.fstride / 4);  // fstride is a multiple of 64
  gcf4* sd = (gcf4*)sc.src;
  gcf4* su = sd + n4;
  gcf4* sv = sd + 2 * n4;
  gf4* fd = (gf4*)f + F_D * n4;
  gf4* fs = (gf4*)f + F_SC * n4;
  gf4* fu = (gf4*)f + F_U * n4;
  gf4* fv = (gf4*)f + F_V * n4;
  const int m = sc.has_src;  // only the sources that were set (adding 0 would turn -0 into +0)
  for (int e = threadIdx.x; e < n4; e += blockDim.x) {
    if (m & 1) {
      const f4 a = ntl(&sd[e]);
      nts(&fd[e], ntl(&fd[e]) + a);
      nts(&fs[e], ntl(&fs[e]) + a);
    }
    if (m & 2) nts(&fu[e], ntl(&fu[e]) + ntl(&su[e]));
    if (m & 4) nts(&fv[e], ntl(&fv[e]) + ntl(&sv[e]));
  }
  __syncthreads();
}

// dst = src over the whole field (fstride is a multiple of 64 floats),
// 8 float4 loads in flight per thread.
template <bool NT>
__device__ __forceinline__ void copy_phase(const SceneDesc& sc, gf* __restrict__ dst, gcf* __restrict__ src) {
  const int n4 = (int)(sc.fstride / 4);
  gcf4* s4 = (gcf4*)src;
  gf4* d4 = (gf4*)dst;
  const int stride = blockDim.x;
  int e = threadIdx.x;
  for (; e + 7 * stride < n4; e += 8 * stride) {
    f4 r[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) r[k] = ntl(&s4[e + k * stride]);
#pragma unroll
    for (int k = 0; k < 8; ++k) nts(&d4[e + k * stride], r[k]);
  }
  for (; e < n4; e += stride) nts(&d4[e], ntl(&s4[e]));
  __syncthreads();
}

enum PhaseKind { PK_GS, PK_DIV, PK_GRAD, PK_ADV, PK_COPY };

// One full frame (fluid.rs:437-524) on one scene, with the buffer roles of the
// reference: U0/V0 receive the diffused velocity, U/V double as pressure and
// divergence of the first projection, U0/V0 as those of the second.  The
// phases run from a loop so that each phase kernel body exists once in the
// code object.
template <int D, int B, int BS, bool XR>
__device__ __forceinline__ void frame_cyclic(const SceneDesc& sc, lf* H) {
  gf* f = (gf*)sc.fld;
  const int n = (int)sc.fstride;
  gf* Dn = f + F_D * n;
  gf* Sc = f + F_SC * n;
  gf* U = f + F_U * n;
  gf* V = f + F_V * n;
  gf* U0 = f + F_U0 * n;
  gf* V0 = f + F_V0 * n;
  const bool big = sc.stream_l2 != 0;  // one-pass phases stream past L2
  if (sc.has_src) {
    if (big) add_sources_phase<true>(sc, f);
    else add_sources_phase<false>(sc, f);
  }
#pragma nounroll
  for (int ph = 0; ph < 12; ++ph) {
    int kind = PK_GS, mode = M_PASSIVE, mode1 = M_PASSIVE, iters = 0;
    gf *x = nullptr, *y = nullptr, *z = nullptr, *w = nullptr, *q = nullptr, *r = nullptr;
    float a = 1.0f, cr = 0.25f;
    switch (ph) {
      case 0: mode = M_ROW; x = U0; y = U; a = sc.a_visc; cr = sc.cr_visc; iters = sc.diff_iters; break;  // fluid.rs:438
      case 1: mode = M_COL; x = V0; y = V; a = sc.a_visc; cr = sc.cr_visc; iters = sc.diff_iters; break;  // fluid.rs:448
      case 2: kind = PK_DIV; x = U0; y = V0; z = V; w = U; break;                                         // fluid.rs:459
      case 3: x = U; y = V; iters = sc.press_iters; break;
      case 4: kind = PK_GRAD; x = U0; y = V0; z = U; break;
      case 5: kind = PK_ADV; mode = M_ROW; mode1 = M_COL; x = U; y = U0; z = V; w = V0; q = U0; r = V0; break;  // fluid.rs:469-489
      case 6: kind = PK_DIV; x = U; y = V; z = V0; w = U0; break;                                         // fluid.rs:491
      case 7: x = U0; y = V0; iters = sc.press_iters; break;
      case 8: kind = PK_GRAD; x = U; y = V; z = U0; break;
      case 9: x = Sc; y = Dn; a = sc.a_diff; cr = sc.cr_diff; iters = sc.diff_iters; break;              // fluid.rs:501
      case 10: kind = PK_ADV; x = Dn; y = Sc; q = U; r = V; break;                                       // fluid.rs:512
      default: kind = PK_COPY; x = Sc; y = Dn; break;                                                    // fluid.rs:523
    }
    if (kind == PK_GS) {
      if (iters > 0) {
        if (mode == M_ROW) gs_wavefront<M_ROW, D, B, BS, XR>(sc, H, x, y, a, cr, iters);
        else if (mode == M_COL) gs_wavefront<M_COL, D, B, BS, XR>(sc, H, x, y, a, cr, iters);
        else gs_wavefront<M_PASSIVE, D, B, BS, XR>(sc, H, x, y, a, cr, iters);
      }
    } else if (kind == PK_DIV) {
      if (big) div_phase<true>(sc, x, y, z, w);
      else div_phase<false>(sc, x, y, z, w);
    } else if (kind == PK_GRAD) {
      if (big) grad_phase<true>(sc, x, y, z);
      else grad_phase<false>(sc, x, y, z);
    } else if (kind == PK_ADV) {
      if (big) advect_phase<true>(sc, H, mode, x, y, mode1, z, w, q, r);
      else advect_phase<false>(sc, H, mode, x, y, mode1, z, w, q, r);
    } else {
      if (big) copy_phase<true>(sc, x, y);
      else copy_phase<false>(sc, x, y);
    }
  }
}

}  // namespace eqf

// grid = number of scenes, block = 64 * waves; scene b is descs[b].  D, B,
// BS and XR are the launch's schedule (GsSched).
template <int D, int B, int BS, bool XR, int MAXT>
__global__ void __launch_bounds__(MAXT) eqf_frames_cyclic(const eqf::SceneDesc* __restrict__ descs, int nframes) {
  extern __shared__ float eqf_lds[];  // cyc_lds_floats(ns, P, waves, schedule) floats, sized by the launch
  const eqf::SceneDesc& sc = descs[blockIdx.x];
  // The largest scenes of a batch set the launch's length: their waves win
  // issue arbitration against co-resident smaller scenes.
  switch (sc.prio) {
    case 3: __builtin_amdgcn_s_setprio(3); break;
    case 2: __builtin_amdgcn_s_setprio(2); break;
    case 1: __builtin_amdgcn_s_setprio(1); break;
    default: break;
  }
  const int nf = sc.frame_mult > 1 ? nframes * sc.frame_mult : nframes;
  for (int fr = 0; fr < nf; ++fr) eqf::frame_cyclic<D, B, BS, XR>(sc, (eqf::lf*)eqf_lds);
}

#define EQF_SCHED_KERNELS(S)                                                                              \
  template __global__ void                                                                               \
  eqf_frames_cyclic<eqf::gs_depth(S), eqf::GS_B, eqf::gs_sync(S), eqf::gs_has_ring(S), 1024>(           \
      const eqf::SceneDesc* __restrict__, int);                                                          \
  template __global__ void eqf_frames_cyclic<eqf::gs_depth(S), eqf::GS_B, eqf::gs_sync(S), eqf::gs_has_ring(S), 512>( \
      const eqf::SceneDesc* __restrict__, int);
EQF_SCHED_KERNELS(eqf::GS_WIDE)
EQF_SCHED_KERNELS(eqf::GS_RING)
EQF_SCHED_KERNELS(eqf::GS_FINE)

