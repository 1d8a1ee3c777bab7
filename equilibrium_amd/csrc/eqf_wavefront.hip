// eqf_wavefront.hip -- exact (bit-for-bit) Stam frame step on gfx950.
//
// One workgroup advances one scene through whole frames (fluid.rs:437-524):
// every phase of the step runs inside the workgroup, separated by workgroup
// barriers, so a batch of scenes of any mix of sizes is ONE launch per frame
// (grid = number of scenes) and a single scene is the same kernel with grid 1.
//
// Fields use the CYCLIC strip layout (eqf_layout.h): wave w owns 64-row
// strips, lane r owns row 64s+r+1, and slot t of a strip holds column
// (t-r) mod N + 1 of every row.  The reference's lexicographic Gauss-Seidel
// (fluid.rs:312-324) is reproduced exactly by an anti-diagonal wavefront:
// at strip-local step tau lane r updates column (tau-r) mod N + 1 of sweep
// (tau-r) div N.  Its left/up neighbours are this sweep's values (own
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

__device__ __forceinline__ float from_prev_lane(float v) {  // lane i <- lane i-1 (DPP wave_shr:1)
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_next_lane(float v) {  // lane i <- lane i+1 (DPP wave_shl:1)
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

typedef __amdgpu_buffer_rsrc_t rsrc_t;
// Raw buffer descriptor (gfx9 family: DATA_FORMAT 32 in word 3).
__device__ __forceinline__ rsrc_t make_rsrc(GAS const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float ldf(rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ uint32_t ldu8(rsrc_t r, int voff, int soff) {
  return (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(r, voff, soff, 0);
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

// ------------------------------------------------------------------------
// Gauss-Seidel wavefront (lin_solve, fluid.rs:301-325) with set_boundaries
// fused.  MODE selects the set_boundaries rule.  D = prefetch depth (steps),
// B = steps between workgroup barriers (a multiple of D).
//
// Ordering argument.  Lane r of strip s handles position u = tau - r, tau =
// g - s*L.  A value stored at global step g is read by another wave through
// a load issued at step >= g + B (a barrier lies in between; on gfx950 all
// waves of a workgroup share one in-order vector-memory path, so a load
// issued after the barrier observes stores issued before it).  With the lag
// L = 63+B+D:
//  * strip s+1 lane 0 reads strip s lane 63's raw value of the same column
//    L-63-D >= B steps after it was stored;
//  * strip s lane 63 reads strip s+1 lane 0's post-boundary value (and its
//    AdjustColumn fix-up) of the previous sweep N+62-D-L-1 >= B steps after
//    it was stored, which needs N >= 2(B+D)+2: true whenever a scene has more
//    than one strip (N > 64);
//  * within a wave, a cell's next-sweep load is issued N-2-D steps after its
//    store: N >= D+3 (smaller grids take the sequential path).
template <int MODE, int D, int B>
__device__ __forceinline__ void gs_wavefront(const SceneDesc& sc, gf* __restrict__ X, gcf* __restrict__ X0,
                                             float a, float cr, int iters) {
  const int N = sc.N, S = sc.S, ns = sc.ns;
  const int lane = threadIdx.x & 63;
  const int wave = uni(threadIdx.x >> 6);
  const int L = 63 + B + D;
  const int NS64 = N * 64;
  const int RO = ns * NS64;  // ring offset (floats)
  gf* __restrict__ ring = X + RO;
  const int G = (ns - 1) * L + 64 + iters * N + 1;
  const int Gr = (G + B - 1) / B * B;

  const bool has_strip = wave < ns;
  const int s = has_strip ? wave : 0;
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

  // Buffer descriptors over whole arrays; per-step addresses are a constant
  // per-lane VGPR offset plus a uniform SGPR offset.  Lanes that must not
  // store get an offset past the end of the buffer: the store is dropped.
  const uint32_t fbytes = (uint32_t)sc.fstride * 4u;
  const rsrc_t rX = make_rsrc(X, fbytes);
  const rsrc_t rX0 = make_rsrc(X0, fbytes);
  const rsrc_t rC = make_rsrc((gcu8*)sc.code, (uint32_t)RO);
  const rsrc_t rH = make_rsrc((gcf*)sc.hraw, (uint32_t)(ns * N) * 4u);
  const rsrc_t rU = has_prev ? rH : rX;
  const int sb = s * NS64 * 4;                                // strip base (bytes)
  const int ub = has_prev ? (s - 1) * N * 4 : (RO + 1) * 4;   // exu: hraw row above or top ring
  const int dnb = has_next ? (s + 1) * NS64 * 4 : (RO + S + 1) * 4;  // exd: strip below or bottom ring
  const int dns = has_next ? 8 : 2;                           // exd: slot stride 256 B or 4 B
  const int vl = lane * 4;
  const int OOB = (int)0x80000000;

  const int lane_m1 = lane == 0 ? N - 1 : lane - 1;  // slot at which this lane is in column N
  const unsigned itN = (unsigned)(iters * N);
  const int lastN = (iters - 1) * N;

  int tau = -s * L;                // uniform strip-local step
  int tcur = ((tau % N) + N) % N;  // uniform slot of tau
  int tld = (tcur + D) % N;        // uniform slot being prefetched

  float fxs[D], fx0[D], fexu[D], fexd[D];
  uint32_t fcode[D];

#define EQF_PREFETCH(SL, T)                                                     \
  {                                                                             \
    const int t_ = (T);                                                         \
    const int t1_ = t_ + 1 == N ? 0 : t_ + 1;                                   \
    fxs[SL] = ldf(rX, vl, sb + t1_ * 256);                                      \
    fx0[SL] = ldf(rX0, vl, sb + t_ * 256);                                      \
    fcode[SL] = ldu8(rC, lane, s * NS64 + t_ * 64);                             \
    fexu[SL] = ldf(rU, 0, ub + t_ * 4);                                         \
    int tl_ = t_ - rl;                                                          \
    tl_ += tl_ < 0 ? N : 0;                                                     \
    fexd[SL] = ldf(rX, 0, dnb + (tl_ << dns));                                  \
  }

  if (has_strip) {
    int t = tcur;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      EQF_PREFETCH(i, t);
      t = t + 1 == N ? 0 : t + 1;
    }
  }

  float res_prev = 0.f, xl_prev = 0.f, xu_prev = 0.f, xself = 0.f, exd_prev = 0.f;
  uint32_t code_prev = 0;
  bool prev_active = false;
  int tprev = 0;

#pragma nounroll
  for (int g0 = 0; g0 < Gr; g0 += B) {
    if (has_strip) {
#pragma unroll
      for (int i = 0; i < B; ++i) {
        const int sl = i % D;
        const float xs = fxs[sl], x0v = fx0[sl], exu = fexu[sl], exd = fexd[sl];
        const uint32_t code = fcode[sl];
        const int u = tau - lane;
        const bool active = (unsigned)u < itN;
        const bool c1 = tcur == lane;
        const bool cN = tcur == lane_m1;
        int tl = tcur - rl;
        tl += tl < 0 ? N : 0;  // slot index of the last lane's column - 1
        float xl, xr, eu, ed;
        if constexpr (MODE == M_PASSIVE) {
          // set_boundaries(Passive) of the previous sweep copied the interior
          // cell next to the ring into the ring (rows/columns that hold a
          // NoWall cell).  That cell is the one updated now: its value from
          // the previous sweep is xself.
          const bool kp = u >= N;
          const bool rc = kp && rowf;
          const bool cc = kp && (code & C_COLF);
          xl = c1 ? (rc ? xself : ringL) : res_prev;
          xr = cN ? (rc ? xself : ringR) : xs;
          eu = (!has_prev && cc) ? xself : exu;
          ed = (!has_next && cc) ? xself : exd;
        } else {
          xl = c1 ? ringL : res_prev;
          xr = cN ? ringR : xs;
          eu = exu;
          ed = exd;
        }
        const float dn = from_next_lane(xs);
        const float up = from_prev_lane(res_prev);
        const float xd = last ? ed : dn;
        const float xu = first ? eu : up;
        // fluid.rs:315-320
        const float res = (x0v + a * (((xr + xl) + xd) + xu)) * cr;

        if constexpr (MODE == M_PASSIVE) {
          // Passive set_boundaries never rewrites interior cells: store now.
          stf(res, rX, (active && valid) ? vl : OOB, sb + tcur * 256);
          if (active && u >= lastN) {  // final sweep: materialise the ring
            if (c1 && rowf && valid) ring[2 * S + j] = res;
            if (cN && rowf && valid) ring[3 * S + j] = res;
            if (first && s == 0 && (code & C_COLF)) ring[tcur + 1] = res;
            if (last && s == ns - 1 && (code & C_COLF) && valid) ring[S + tl + 1] = res;
          }
        } else {
          // Finalise the cell updated at the previous step (set_boundaries of
          // this sweep, fluid.rs:151-178): its wall neighbours are known now.
          float fin = res_prev;
          const bool fl = (code_prev & C_FLUID) != 0;
          if constexpr (MODE == M_ROW) {
            const float vR = c1 ? ringR : res;  // right neighbour of the pending cell
            fin = (fl && (code_prev & C_RWALL)) ? -vR : ((fl && (code_prev & C_LWALL)) ? -xl_prev : res_prev);
          } else {
            const float vD = from_next_lane(res);  // the lane below computed the pending column now
            const float vDn = last ? exd_prev : vD;
            const bool dok = !(last && has_next);  // otherwise the strip below writes it
            fin = (fl && (code_prev & C_UWALL)) ? -xu_prev
                                                 : ((fl && dok && (code_prev & C_DWALL)) ? -vDn : res_prev);
          }
          stf(fin, rX, (prev_active && valid) ? vl : OOB, sb + tprev * 256);
          if constexpr (MODE == M_COL) {
            // AdjustColumn value of the cell above (last row of the strip above)
            int tu = tcur + 63;
            tu -= tu >= N ? N : 0;
            const bool fix = first && has_prev && active && (code & C_FIXUP);
            stf(-res, rX, fix ? 0 : OOB, sb - NS64 * 4 + (tu * 64 + 63) * 4);
          }
          xl_prev = xl;
          xu_prev = xu;
          code_prev = code;
          exd_prev = exd;
          prev_active = active;
          tprev = tcur;
        }
        // raw last-row value for the strip below
        stf(res, rH, (last && has_next && active && valid) ? 0 : OOB, (s * N + tl) * 4);
        res_prev = res;
        xself = xs;
        ++tau;
        tcur = tcur + 1 == N ? 0 : tcur + 1;
        EQF_PREFETCH(sl, tld);
        tld = tld + 1 == N ? 0 : tld + 1;
      }
    }
    __syncthreads();
  }
#undef EQF_PREFETCH

  if (threadIdx.x == 0) corners(ring, S, N);
  __syncthreads();
}

// Sequential lin_solve for tiny grids (N < D+3): one thread runs the
// reference loops on the cyclic layout, set_boundaries through the lists.
__device__ __forceinline__ void gs_serial(const SceneDesc& sc, int mode, gf* X, gcf* X0, float a, float cr,
                                          int iters) {
  const int S = sc.S, N = sc.N, ns = sc.ns;
  gci* list = (gci*)sc.bnd + 2 * (mode == M_ROW ? 0 : mode == M_COL ? sc.nbnd[0] : sc.nbnd[0] + sc.nbnd[1]);
  const int nl = sc.nbnd[mode];
  const float sign = mode == M_PASSIVE ? 1.0f : -1.0f;
  gf* ring = X + cyc_ring(N, ns);
  if (threadIdx.x == 0) {
    for (int kk = 0; kk < iters; ++kk) {
      for (int y = 1; y <= N; ++y)
        for (int x = 1; x <= N; ++x) {
          X[cyc_off(x, y, S, N, ns)] =
              (X0[cyc_off(x, y, S, N, ns)] +
               a * (((X[cyc_off(x + 1, y, S, N, ns)] + X[cyc_off(x - 1, y, S, N, ns)]) + X[cyc_off(x, y + 1, S, N, ns)]) +
                    X[cyc_off(x, y - 1, S, N, ns)])) *
              cr;
        }
      for (int e = 0; e < nl; ++e) X[list[2 * e]] = sign * X[list[2 * e + 1]];
      corners(ring, S, N);
    }
  }
  __syncthreads();
}

// set_boundaries as a list pass: x[dst] = -x[src] (ROW/COL) or x[dst] = x[src]
// (PASSIVE), then the four corners.  Up to two fields share one pass.
__device__ __forceinline__ void bnd_lists(const SceneDesc& sc, int mode0, gf* x0, int mode1, gf* x1) {
  const int S = sc.S, N = sc.N;
  const int ro = N * 64 * sc.ns;
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

// project, first half (fluid.rs:339-349): div and p = 0 on interior cells,
// then set_boundaries(Passive) on both (fluid.rs:351-352).
__device__ __forceinline__ void div_phase(const SceneDesc& sc, gcf* __restrict__ u, gcf* __restrict__ v,
                                          gf* __restrict__ div, gf* __restrict__ p) {
  const int N = sc.N, S = sc.S, ns = sc.ns;
  const int lane = threadIdx.x & 63;
  const int NS64 = N * 64;
  gcf* uring = u + ns * NS64;
  gcf* vring = v + ns * NS64;
  const float sf = sc.sf;
  for (int s = uni(threadIdx.x >> 6); s < ns; s += (blockDim.x >> 6)) {
    const int j = 64 * s + lane + 1;
    const bool valid = j <= N;
    const int jc = min(j, N);
    const int rl = min(63, N - 1 - 64 * s);
    const bool first = lane == 0, last = lane == rl;
    const bool has_prev = s > 0, has_next = s < ns - 1;
    gcf* us = u + s * NS64;
    gcf* vs = v + s * NS64;
    gcf* vnext = v + (has_next ? s + 1 : s) * NS64;
    gcf* vprev = v + (has_prev ? s - 1 : s) * NS64;
    gf* ds = div + s * NS64;
    gf* ps = p + s * NS64;
    const float uL = uring[2 * S + jc], uR = uring[3 * S + jc];
    const int lane_m1 = lane == 0 ? N - 1 : lane - 1;
    float up_ = us[(N - 1) * 64 + lane], uc = us[lane];
    float vp_ = vs[(N - 1) * 64 + lane], vc = vs[lane];
    for (int t = 0; t < N; ++t) {
      const int t1 = t + 1 == N ? 0 : t + 1;
      const float un = us[t1 * 64 + lane];
      const float vn = vs[t1 * 64 + lane];
      int tl = t - rl;
      tl += tl < 0 ? N : 0;  // column of the last lane is tl+1
      int tu = t + 63;
      tu -= tu >= N ? N : 0;  // strip above, lane 63, column t+1
      gcf* pdn = has_next ? vnext + tl * 64 : vring + S + tl + 1;
      gcf* pup = has_prev ? vprev + tu * 64 + 63 : vring + t + 1;
      const float vdn_e = *pdn, vup_e = *pup;
      // neighbours (fluid.rs:342-344)
      const float uplus = t == lane_m1 ? uR : un;
      const float uminus = t == lane ? uL : up_;
      const float vdn_l = from_next_lane(vn);
      const float vup_l = from_prev_lane(vp_);
      const float vplus = last ? vdn_e : vdn_l;
      const float vminus = first ? vup_e : vup_l;
      const float d = -0.5f * (uplus - uminus + vplus - vminus) / sf;
      if (valid) {
        ds[t * 64 + lane] = d;
        ps[t * 64 + lane] = 0.0f;
      }
      up_ = uc;
      uc = un;
      vp_ = vc;
      vc = vn;
    }
  }
  __syncthreads();
  bnd_lists(sc, M_PASSIVE, div, M_PASSIVE, p);
}

// project, second half (fluid.rs:364-374): subtract the pressure gradient,
// then set_boundaries(AdjustRow, u) and (AdjustColumn, v).
__device__ __forceinline__ void grad_phase(const SceneDesc& sc, gf* __restrict__ u, gf* __restrict__ v,
                                           gcf* __restrict__ p) {
  const int N = sc.N, S = sc.S, ns = sc.ns;
  const int lane = threadIdx.x & 63;
  const int NS64 = N * 64;
  gcf* pring = p + ns * NS64;
  const float sf = sc.sf;
  for (int s = uni(threadIdx.x >> 6); s < ns; s += (blockDim.x >> 6)) {
    const int j = 64 * s + lane + 1;
    const bool valid = j <= N;
    const int jc = min(j, N);
    const int rl = min(63, N - 1 - 64 * s);
    const bool first = lane == 0, last = lane == rl;
    const bool has_prev = s > 0, has_next = s < ns - 1;
    gcf* ps = p + s * NS64;
    gcf* pnext = p + (has_next ? s + 1 : s) * NS64;
    gcf* pprev = p + (has_prev ? s - 1 : s) * NS64;
    gf* us = u + s * NS64;
    gf* vs = v + s * NS64;
    const float pL = pring[2 * S + jc], pR = pring[3 * S + jc];
    const int lane_m1 = lane == 0 ? N - 1 : lane - 1;
    float pp = ps[(N - 1) * 64 + lane], pc = ps[lane];
    for (int t = 0; t < N; ++t) {
      const int t1 = t + 1 == N ? 0 : t + 1;
      const float pn = ps[t1 * 64 + lane];
      int tl = t - rl;
      tl += tl < 0 ? N : 0;
      int tu = t + 63;
      tu -= tu >= N ? N : 0;
      gcf* qdn = has_next ? pnext + tl * 64 : pring + S + tl + 1;
      gcf* qup = has_prev ? pprev + tu * 64 + 63 : pring + t + 1;
      const float pdn_e = *qdn, pup_e = *qup;
      const float pplus = t == lane_m1 ? pR : pn;
      const float pminus = t == lane ? pL : pp;
      const float pdn_l = from_next_lane(pn);
      const float pup_l = from_prev_lane(pp);
      const float pdn = last ? pdn_e : pdn_l;
      const float pup = first ? pup_e : pup_l;
      if (valid) {
        const int o = t * 64 + lane;
        us[o] = us[o] - 0.5f * (pplus - pminus) * sf;
        vs[o] = vs[o] - 0.5f * (pdn - pup) * sf;
      }
      pp = pc;
      pc = pn;
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
__device__ __forceinline__ float gather(gcf* d0, const Trace& r, int S, int N, int ns) {
  const float a = d0[cyc_off(r.i0, r.j0, S, N, ns)];
  const float b = d0[cyc_off(r.i0, r.j1, S, N, ns)];
  const float cc = d0[cyc_off(r.i1, r.j0, S, N, ns)];
  const float d = d0[cyc_off(r.i1, r.j1, S, N, ns)];
  return r.s0 * (r.t0 * a + r.t1 * b) + r.s1 * (r.t0 * cc + r.t1 * d);
}

// advect (fluid.rs:378-432) of one or two fields that share the velocity
// (advect of u and v in step, fluid.rs:469-489), then set_boundaries.  The
// `break` at fluid.rs:420-423 is reproduced per row: cells left of the first
// trigger column get the bilinear value, the trigger cell copies its left
// neighbour (the ring cell when it is column 1), cells to the right keep the
// buffer's previous content.
__device__ __forceinline__ void advect_phase(const SceneDesc& sc, int mode0, gf* __restrict__ d0out,
                                             gcf* __restrict__ d0in, int mode1, gf* __restrict__ d1out,
                                             gcf* __restrict__ d1in, gcf* __restrict__ u, gcf* __restrict__ v) {
  const int N = sc.N, S = sc.S, ns = sc.ns;
  const int lane = threadIdx.x & 63;
  const int NS64 = N * 64;
  const float dt0 = sc.dt0, sf = sc.sf, sfm1 = sc.sfm1;
  const bool two = d1out != nullptr;
  for (int s = uni(threadIdx.x >> 6); s < ns; s += (blockDim.x >> 6)) {
    const int j = 64 * s + lane + 1;
    const bool valid = j <= N;
    const int jc = min(j, N);
    gcf* us = u + s * NS64;
    gcf* vs = v + s * NS64;
    // pass A: first trigger column of this row
    int trig = N + 1;
    for (int t = 0; t < N; ++t) {
      int c = t - lane;
      c += c < 0 ? N : 0;
      c += 1;
      const Trace r = back_trace(c, jc, us[t * 64 + lane], vs[t * 64 + lane], dt0, sf, sfm1, S);
      if (r.trig && c < trig) trig = c;
    }
    // pass B
    const int ro = ns * NS64;
    const float left0 = d0out[ro + 2 * S + jc];
    const float left1 = two ? d1out[ro + 2 * S + jc] : 0.0f;
    gf* o0 = d0out + s * NS64;
    gf* o1 = two ? d1out + s * NS64 : d0out;
    int tp = N - 1;
    for (int t = 0; t < N; ++t) {
      int c = t - lane;
      c += c < 0 ? N : 0;
      c += 1;
      if (valid && c <= trig) {
        const int o = t * 64 + lane;
        const bool copy_left = c == trig && c == 1;
        if (copy_left) {  // trigger in column 1: copy the ring cell (fluid.rs:421)
          o0[o] = left0;
          if (two) o1[o] = left1;
        } else {
          // c < trig: this cell; c == trig > 1: the left neighbour's value
          const int tt = c == trig ? tp : t;
          const int cc = c == trig ? c - 1 : c;
          const Trace r = back_trace(cc, j, us[tt * 64 + lane], vs[tt * 64 + lane], dt0, sf, sfm1, S);
          o0[o] = gather(d0in, r, S, N, ns);
          if (two) o1[o] = gather(d1in, r, S, N, ns);
        }
      }
      tp = t;
    }
  }
  __syncthreads();
  bnd_lists(sc, mode0, d0out, mode1, d1out);
}

// Frame sources: add_density (fluid.rs:122-123) and add_velocity
// (fluid.rs:129-130) applied to every cell, interior and ring.
__device__ __forceinline__ void add_sources_phase(const SceneDesc& sc, gf* f) {
  const int n = (int)sc.fstride;
  gcf* sd = (gcf*)sc.src;
  gcf* su = sd + n;
  gcf* sv = sd + 2 * n;
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    const float a = sd[e];
    f[F_D * n + e] += a;
    f[F_SC * n + e] += a;
    f[F_U * n + e] += su[e];
    f[F_V * n + e] += sv[e];
  }
  __syncthreads();
}

__device__ __forceinline__ void copy_phase(const SceneDesc& sc, gf* __restrict__ dst, gcf* __restrict__ src) {
  const int n = (int)sc.fstride;
  for (int e = threadIdx.x; e < n; e += blockDim.x) dst[e] = src[e];
  __syncthreads();
}

enum PhaseKind { PK_GS, PK_DIV, PK_GRAD, PK_ADV, PK_COPY };

// One full frame (fluid.rs:437-524) on one scene, with the buffer roles of the
// reference: U0/V0 receive the diffused velocity, U/V double as pressure and
// divergence of the first projection, U0/V0 as those of the second.  The
// phases run from a loop so that each phase kernel body exists once in the
// code object.
template <int D, int B>
__device__ __forceinline__ void frame_cyclic(const SceneDesc& sc) {
  gf* f = (gf*)sc.fld;
  const int n = (int)sc.fstride;
  gf* Dn = f + F_D * n;
  gf* Sc = f + F_SC * n;
  gf* U = f + F_U * n;
  gf* V = f + F_V * n;
  gf* U0 = f + F_U0 * n;
  gf* V0 = f + F_V0 * n;
  // The wavefront needs N >= D+3 (a cell's next-sweep load must be issued
  // after its store); more than one strip implies N > 64 >= 2*(B+D)+2.
  const bool serial = sc.N < D + 3;
  if (sc.has_src) add_sources_phase(sc, f);
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
        if (serial) gs_serial(sc, mode, x, y, a, cr, iters);
        else if (mode == M_ROW) gs_wavefront<M_ROW, D, B>(sc, x, y, a, cr, iters);
        else if (mode == M_COL) gs_wavefront<M_COL, D, B>(sc, x, y, a, cr, iters);
        else gs_wavefront<M_PASSIVE, D, B>(sc, x, y, a, cr, iters);
      }
    } else if (kind == PK_DIV) {
      div_phase(sc, x, y, z, w);
    } else if (kind == PK_GRAD) {
      grad_phase(sc, x, y, z);
    } else if (kind == PK_ADV) {
      advect_phase(sc, mode, x, y, mode1, z, w, q, r);
    } else {
      copy_phase(sc, x, y);
    }
  }
}

}  // namespace eqf

// grid = number of scenes, block = 64 * waves; scene b is descs[b].
template <int D, int B, int MAXT>
__global__ void __launch_bounds__(MAXT) eqf_frames_cyclic(const eqf::SceneDesc* __restrict__ descs, int nframes) {
  const eqf::SceneDesc& sc = descs[blockIdx.x];
  for (int fr = 0; fr < nframes; ++fr) eqf::frame_cyclic<D, B>(sc);
}

template __global__ void eqf_frames_cyclic<8, 16, 1024>(const eqf::SceneDesc* __restrict__, int);
template __global__ void eqf_frames_cyclic<8, 16, 512>(const eqf::SceneDesc* __restrict__, int);
