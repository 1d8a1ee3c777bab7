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
// r+1 via DPP wave_shl).  Consecutive sweeps are pipelined back to back and
// strip s+1 trails strip s by a fixed lag, so the schedule is static.
// set_boundaries (fluid.rs:252-272) is fused: a cell's store is delayed by
// one step until the wall values it mirrors are known.
//
// Arithmetic is written in the reference's evaluation order and the file is
// compiled with -ffp-contract=off: rustc does not contract a*b+c, so the HIP
// compiler must not either.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "eqf_layout.h"

namespace eqf {

__device__ __forceinline__ float from_prev_lane(float v) {  // lane i <- lane i-1
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_next_lane(float v) {  // lane i <- lane i+1
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

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

// ------------------------------------------------------------------------
// Gauss-Seidel wavefront (lin_solve, fluid.rs:301-325) with set_boundaries
// fused.  MODE selects the set_boundaries rule.  D = prefetch depth in steps,
// B = steps between workgroup barriers.  Requires N >= 2*(B+D)+1 when the
// scene has more than one strip and N >= D+3 always (checked by the host).
template <int MODE, int D, int B>
__device__ __noinline__ void gs_wavefront(const SceneDesc& sc, float* __restrict__ X, const float* __restrict__ X0,
                                          float a, float cr, int iters) {
  if (iters <= 0) return;
  const int N = sc.N, S = sc.S, ns = sc.ns;
  const int lane = threadIdx.x & 63;
  const int wave = uni(threadIdx.x >> 6);
  const int W = blockDim.x >> 6;
  const int L = 63 + B + D;  // lag between consecutive strips
  const int64_t NS64 = (int64_t)N * 64;
  float* __restrict__ ring = X + (int64_t)ns * NS64;
  const int G = (ns - 1) * L + 64 + iters * N + 1;
  const int Gr = (G + B - 1) / B * B;

  const bool has_strip = wave < ns;
  const int s = has_strip ? wave : 0;
  const int j = 64 * s + lane + 1;          // row of this lane
  const bool valid = has_strip && j <= N;
  const int rl = min(63, N - 1 - 64 * s);   // last valid lane of the strip
  const bool first = lane == 0;
  const bool last = lane == rl;
  const bool has_prev = s > 0;
  const bool has_next = s < ns - 1;
  const int jc = min(j, N);
  const float ringL = ring[2 * S + jc];
  const float ringR = ring[3 * S + jc];
  const bool rowf = sc.rowf[jc] != 0;
  const float* __restrict__ Xs = X + s * NS64;
  const float* __restrict__ X0s = X0 + s * NS64;
  const uint8_t* __restrict__ Cs = sc.code + s * NS64;
  const float* __restrict__ Xnext = X + (int64_t)(has_next ? s + 1 : s) * NS64;
  const float* __restrict__ hup = sc.hraw + (int64_t)(has_prev ? s - 1 : 0) * N;
  float* __restrict__ hdn = sc.hraw + (int64_t)s * N;

  // strip-local step at g = 0
  const int tau0 = -s * L;
  int tcur = ((tau0 % N) + N) % N;      // slot of this strip at the current step
  int tld = (tcur + D) % N;             // slot of the step being prefetched
  const int u0 = tau0 - lane;
  int k = (u0 >= 0) ? u0 / N : -((-u0 + N - 1) / N);
  int c = u0 - k * N + 1;

  float fxs[D], fx0[D], fexu[D], fexd[D];
  uint32_t fcode[D];

  auto prefetch = [&](int sl, int t) {
    int t1 = t + 1 == N ? 0 : t + 1;
    fxs[sl] = Xs[(int64_t)t1 * 64 + lane];
    fx0[sl] = X0s[(int64_t)t * 64 + lane];
    fcode[sl] = Cs[(int64_t)t * 64 + lane];
    fexu[sl] = has_prev ? hup[t] : ring[t + 1];
    int tl = t - rl;
    if (tl < 0) tl += N;
    fexd[sl] = has_next ? Xnext[(int64_t)tl * 64] : ring[S + tl + 1];
  };

  if (has_strip) {
    int t = tcur;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      prefetch(i, t);
      t = t + 1 == N ? 0 : t + 1;
    }
  }

  float res_prev = 0.f, xl_prev = 0.f, xu_prev = 0.f, xself = 0.f, exd_prev = 0.f;
  uint32_t code_prev = 0;
  bool prev_active = false;
  int tprev = 0;

  for (int g0 = 0; g0 < Gr; g0 += B) {
    if (has_strip) {
#pragma unroll
      for (int i = 0; i < B; ++i) {
        const int sl = i % D;
        const float xs = fxs[sl], x0v = fx0[sl], exu = fexu[sl], exd = fexd[sl];
        const uint32_t code = fcode[sl];
        const bool active = (k >= 0) && (k < iters);
        const bool c1 = c == 1, cN = c == N;
        float xl, xr, eu, ed;
        if constexpr (MODE == M_PASSIVE) {
          // After sweep k-1, set_boundaries(Passive) copied the interior cell
          // next to the ring into the ring (only for rows/columns holding a
          // NoWall cell).  That interior cell is the one being updated now, and
          // its previous-sweep value is xself.
          const bool kp = k > 0;
          const bool rc = kp && rowf;
          const bool cc = kp && (code & C_COLF);
          xl = c1 ? (rc ? xself : ringL) : res_prev;
          xr = cN ? (rc ? xself : ringR) : xs;
          eu = has_prev ? exu : (cc ? xself : exu);
          ed = has_next ? exd : (cc ? xself : exd);
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

        // Finalise the cell updated at the previous step (set_boundaries of
        // this sweep, fluid.rs:151-178): its wall neighbours are known now.
        float fin = res_prev;
        const bool fl = (code_prev & C_FLUID) != 0;
        if constexpr (MODE == M_ROW) {
          const float vR = c1 ? ringR : res;  // right neighbour of the pending cell
          fin = (fl && (code_prev & C_RWALL)) ? -vR : ((fl && (code_prev & C_LWALL)) ? -xl_prev : res_prev);
        } else if constexpr (MODE == M_COL) {
          const float vD = from_next_lane(res);  // lane below computed the pending column now
          const float vDn = last ? exd_prev : vD;
          const bool dok = !(last && has_next);  // else the strip below writes it
          fin = (fl && (code_prev & C_UWALL)) ? -xu_prev
                                               : ((fl && dok && (code_prev & C_DWALL)) ? -vDn : res_prev);
        }
        if (prev_active && valid) X[s * NS64 + (int64_t)tprev * 64 + lane] = fin;

        if (active && valid) {
          if constexpr (MODE == M_COL) {
            if (first && has_prev && (code & C_FIXUP)) {
              int tu = tcur + 63;
              if (tu >= N) tu -= N;
              X[(s - 1) * NS64 + (int64_t)tu * 64 + 63] = -res;
            }
          }
          if (last && has_next) hdn[c - 1] = res;
          if constexpr (MODE == M_PASSIVE) {
            if (k == iters - 1) {
              if (c1 && rowf) ring[2 * S + j] = res;
              if (cN && rowf) ring[3 * S + j] = res;
              if (first && s == 0 && (code & C_COLF)) ring[c] = res;
              if (last && s == ns - 1 && (code & C_COLF)) ring[S + c] = res;
            }
          }
        }

        xl_prev = xl;
        xu_prev = xu;
        res_prev = res;
        code_prev = code;
        xself = xs;
        exd_prev = exd;
        prev_active = active;
        ++c;
        if (c > N) {
          c = 1;
          ++k;
        }
        tprev = tcur;
        tcur = tcur + 1 == N ? 0 : tcur + 1;
        prefetch(sl, tld);
        tld = tld + 1 == N ? 0 : tld + 1;
      }
    }
    __syncthreads();
  }

  if (threadIdx.x == 0) {  // corners, fluid.rs:265-271
    ring[0] = 0.5f * (ring[1] + ring[2 * S + 1]);
    ring[S] = 0.5f * (ring[S + 1] + ring[2 * S + N]);
    ring[S - 1] = 0.5f * (ring[N] + ring[3 * S + 1]);
    ring[2 * S - 1] = 0.5f * (ring[S + N] + ring[3 * S + N]);
  }
  __syncthreads();
}

// Sequential lin_solve for tiny grids (N below the wavefront minimum): one
// thread runs the reference loops verbatim on the cyclic layout.
template <int MODE>
__device__ __noinline__ void gs_serial(const SceneDesc& sc, float* X, const float* X0, float a, float cr,
                                       int iters) {
  if (iters <= 0) return;
  const int S = sc.S, N = sc.N, ns = sc.ns;
  const int2* list = sc.bnd + (MODE == M_ROW ? 0 : MODE == M_COL ? sc.nbnd[0] : sc.nbnd[0] + sc.nbnd[1]);
  const int nl = sc.nbnd[MODE];
  const float sign = MODE == M_PASSIVE ? 1.0f : -1.0f;
  float* ring = X + cyc_ring(N, ns);
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
      for (int e = 0; e < nl; ++e) X[list[e].x] = sign * X[list[e].y];
      ring[0] = 0.5f * (ring[1] + ring[2 * S + 1]);
      ring[S] = 0.5f * (ring[S + 1] + ring[2 * S + N]);
      ring[S - 1] = 0.5f * (ring[N] + ring[3 * S + 1]);
      ring[2 * S - 1] = 0.5f * (ring[S + N] + ring[3 * S + N]);
    }
  }
  __syncthreads();
}

// set_boundaries as a list pass: x[dst] = -x[src] (ROW/COL) or x[dst] = x[src]
// (PASSIVE), then the four corners.  Up to two fields share one pass.
__device__ __noinline__ void bnd_lists(const SceneDesc& sc, int mode0, float* x0, int mode1, float* x1) {
  const int S = sc.S, N = sc.N;
  const int64_t ro = cyc_ring(N, sc.ns);
  const int off[3] = {0, sc.nbnd[0], sc.nbnd[0] + sc.nbnd[1]};
  for (int f = 0; f < 2; ++f) {
    const int mode = f == 0 ? mode0 : mode1;
    float* x = f == 0 ? x0 : x1;
    if (!x) continue;
    const int2* list = sc.bnd + off[mode];
    const int n = sc.nbnd[mode];
    const float sign = mode == M_PASSIVE ? 1.0f : -1.0f;
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
      const int2 p = list[e];
      x[p.x] = sign * x[p.y];
    }
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    float* x = threadIdx.x == 0 ? x0 : x1;
    if (x) {
      float* ring = x + ro;
      ring[0] = 0.5f * (ring[1] + ring[2 * S + 1]);
      ring[S] = 0.5f * (ring[S + 1] + ring[2 * S + N]);
      ring[S - 1] = 0.5f * (ring[N] + ring[3 * S + 1]);
      ring[2 * S - 1] = 0.5f * (ring[S + N] + ring[3 * S + N]);
    }
  }
  __syncthreads();
}

// Loop helper over (strip, lane) rows owned by the calling wave.
#define EQF_FOR_STRIPS(sc)                                            \
  for (int s = uni(threadIdx.x >> 6); s < (sc).ns; s += (blockDim.x >> 6))

// project, first half (fluid.rs:339-349): div and p = 0 on interior cells,
// then set_boundaries(Passive) on both (fluid.rs:351-352).
__device__ __noinline__ void div_phase(const SceneDesc& sc, const float* __restrict__ u, const float* __restrict__ v,
                                       float* __restrict__ div, float* __restrict__ p) {
  const int N = sc.N, S = sc.S, ns = sc.ns;
  const int lane = threadIdx.x & 63;
  const int64_t NS64 = (int64_t)N * 64;
  const float* uring = u + (int64_t)ns * NS64;
  const float* vring = v + (int64_t)ns * NS64;
  const float sf = sc.sf;
  EQF_FOR_STRIPS(sc) {
    const int j = 64 * s + lane + 1;
    const bool valid = j <= N;
    const int jc = min(j, N);
    const int rl = min(63, N - 1 - 64 * s);
    const bool first = lane == 0, last = lane == rl;
    const bool has_prev = s > 0, has_next = s < ns - 1;
    const float* us = u + s * NS64;
    const float* vs = v + s * NS64;
    const float uL = uring[2 * S + jc], uR = uring[3 * S + jc];
    float up_ = us[(int64_t)(N - 1) * 64 + lane], uc = us[lane];
    float vp_ = vs[(int64_t)(N - 1) * 64 + lane], vc = vs[lane];
    for (int t = 0; t < N; ++t) {
      const int t1 = t + 1 == N ? 0 : t + 1;
      const float un = us[(int64_t)t1 * 64 + lane];
      const float vn = vs[(int64_t)t1 * 64 + lane];
      int c = t - lane;
      if (c < 0) c += N;
      c += 1;
      // neighbours (fluid.rs:342-344)
      const float uplus = c == N ? uR : un;
      const float uminus = c == 1 ? uL : up_;
      const float vdn_l = from_next_lane(vn);
      const float vup_l = from_prev_lane(vp_);
      int tl = t - rl;
      if (tl < 0) tl += N;  // column of the last lane is tl+1
      const float vdn_e = has_next ? v[(s + 1) * NS64 + (int64_t)tl * 64] : vring[S + tl + 1];
      int tu = t + 63;
      if (tu >= N) tu -= N;  // strip above, lane 63, column t+1
      const float vup_e = has_prev ? v[(s - 1) * NS64 + (int64_t)tu * 64 + 63] : vring[t + 1];
      const float vplus = last ? vdn_e : vdn_l;
      const float vminus = first ? vup_e : vup_l;
      const float d = -0.5f * (uplus - uminus + vplus - vminus) / sf;
      if (valid) {
        div[s * NS64 + (int64_t)t * 64 + lane] = d;
        p[s * NS64 + (int64_t)t * 64 + lane] = 0.0f;
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
__device__ __noinline__ void grad_phase(const SceneDesc& sc, float* __restrict__ u, float* __restrict__ v,
                                        const float* __restrict__ p) {
  const int N = sc.N, S = sc.S, ns = sc.ns;
  const int lane = threadIdx.x & 63;
  const int64_t NS64 = (int64_t)N * 64;
  const float* pring = p + (int64_t)ns * NS64;
  const float sf = sc.sf;
  EQF_FOR_STRIPS(sc) {
    const int j = 64 * s + lane + 1;
    const bool valid = j <= N;
    const int jc = min(j, N);
    const int rl = min(63, N - 1 - 64 * s);
    const bool first = lane == 0, last = lane == rl;
    const bool has_prev = s > 0, has_next = s < ns - 1;
    const float* ps = p + s * NS64;
    float* us = u + s * NS64;
    float* vs = v + s * NS64;
    const float pL = pring[2 * S + jc], pR = pring[3 * S + jc];
    float pp = ps[(int64_t)(N - 1) * 64 + lane], pc = ps[lane];
    for (int t = 0; t < N; ++t) {
      const int t1 = t + 1 == N ? 0 : t + 1;
      const float pn = ps[(int64_t)t1 * 64 + lane];
      int c = t - lane;
      if (c < 0) c += N;
      c += 1;
      const float pplus = c == N ? pR : pn;
      const float pminus = c == 1 ? pL : pp;
      const float pdn_l = from_next_lane(pn);
      const float pup_l = from_prev_lane(pp);
      int tl = t - rl;
      if (tl < 0) tl += N;
      const float pdn_e = has_next ? p[(s + 1) * NS64 + (int64_t)tl * 64] : pring[S + tl + 1];
      int tu = t + 63;
      if (tu >= N) tu -= N;
      const float pup_e = has_prev ? p[(s - 1) * NS64 + (int64_t)tu * 64 + 63] : pring[t + 1];
      const float pdn = last ? pdn_e : pdn_l;
      const float pup = first ? pup_e : pup_l;
      if (valid) {
        const int64_t o = (int64_t)t * 64 + lane;
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
__device__ __forceinline__ Trace back_trace(int c, int j, float uu, float vv, const SceneDesc& sc) {
  Trace r;
  float x = (float)c - sc.dt0 * uu;
  float y = (float)j - sc.dt0 * vv;
  x = rclamp(x, 0.5f, sc.sfm1);
  y = rclamp(y, 0.5f, sc.sfm1);
  const float i0 = floorf(x), i1 = i0 + 1.0f;
  const float j0 = floorf(y), j1 = j0 + 1.0f;
  r.s1 = x - i0;
  r.s0 = 1.0f - r.s1;
  r.t1 = y - j0;
  r.t0 = 1.0f - r.t1;
  r.trig = (i1 >= sc.sf) || (j1 >= sc.sf);
  r.i0 = sat_idx(i0, sc.S);
  r.i1 = sat_idx(i1, sc.S);
  r.j0 = sat_idx(j0, sc.S);
  r.j1 = sat_idx(j1, sc.S);
  return r;
}
__device__ __forceinline__ float gather(const float* d0, const Trace& r, const SceneDesc& sc) {
  const int S = sc.S, N = sc.N, ns = sc.ns;
  const float a = d0[cyc_off(r.i0, r.j0, S, N, ns)];
  const float b = d0[cyc_off(r.i0, r.j1, S, N, ns)];
  const float cc = d0[cyc_off(r.i1, r.j0, S, N, ns)];
  const float d = d0[cyc_off(r.i1, r.j1, S, N, ns)];
  return r.s0 * (r.t0 * a + r.t1 * b) + r.s1 * (r.t0 * cc + r.t1 * d);
}

// advect (fluid.rs:378-432) of up to two fields that share the velocity
// (advect of u and v in step, fluid.rs:469-489), then set_boundaries.  The
// `break` at fluid.rs:420-423 is reproduced per row: cells left of the first
// trigger column get the bilinear value, the trigger cell copies its left
// neighbour (the ring cell when it is column 1), cells to the right keep the
// buffer's previous content.
__device__ __noinline__ void advect_phase(const SceneDesc& sc, int mode0, float* __restrict__ d0out,
                                          const float* __restrict__ d0in, int mode1, float* __restrict__ d1out,
                                          const float* __restrict__ d1in, const float* __restrict__ u,
                                          const float* __restrict__ v) {
  const int N = sc.N, S = sc.S, ns = sc.ns;
  const int lane = threadIdx.x & 63;
  const int64_t NS64 = (int64_t)N * 64;
  EQF_FOR_STRIPS(sc) {
    const int j = 64 * s + lane + 1;
    const bool valid = j <= N;
    const int jc = min(j, N);
    const float* us = u + s * NS64;
    const float* vs = v + s * NS64;
    // pass A: first trigger column of this row
    int trig = N + 1;
    for (int t = 0; t < N; ++t) {
      int c = t - lane;
      if (c < 0) c += N;
      c += 1;
      const Trace r = back_trace(c, jc, us[(int64_t)t * 64 + lane], vs[(int64_t)t * 64 + lane], sc);
      if (r.trig && c < trig) trig = c;
    }
    // pass B
    const int64_t ro = (int64_t)ns * NS64;
    const float left0 = d0out[ro + 2 * S + jc];
    const float left1 = d1out ? d1out[ro + 2 * S + jc] : 0.0f;
    int tp = N - 1;
    for (int t = 0; t < N; ++t) {
      int c = t - lane;
      if (c < 0) c += N;
      c += 1;
      if (!valid || c > trig) {
        tp = t;
        continue;
      }
      const int64_t o = s * NS64 + (int64_t)t * 64 + lane;
      if (c < trig) {
        const Trace r = back_trace(c, j, us[(int64_t)t * 64 + lane], vs[(int64_t)t * 64 + lane], sc);
        d0out[o] = gather(d0in, r, sc);
        if (d1out) d1out[o] = gather(d1in, r, sc);
      } else {  // c == trig: copy the left neighbour (fluid.rs:421)
        if (c == 1) {
          d0out[o] = left0;
          if (d1out) d1out[o] = left1;
        } else {
          const Trace r = back_trace(c - 1, j, us[(int64_t)tp * 64 + lane], vs[(int64_t)tp * 64 + lane], sc);
          d0out[o] = gather(d0in, r, sc);
          if (d1out) d1out[o] = gather(d1in, r, sc);
        }
      }
      tp = t;
    }
  }
  __syncthreads();
  bnd_lists(sc, mode0, d0out, mode1, d1out);
}

// Elementwise passes over whole field arrays (interior slots and ring).
__device__ __noinline__ void add_sources_phase(const SceneDesc& sc) {
  const int64_t n = sc.fstride;
  float* f = sc.fld;
  const float* sd = sc.src;
  const float* su = sc.src + n;
  const float* sv = sc.src + 2 * n;
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) {
    const float a = sd[e];
    f[F_D * n + e] += a;   // add_density, fluid.rs:122-123
    f[F_SC * n + e] += a;
    f[F_U * n + e] += su[e];  // add_velocity, fluid.rs:129-130
    f[F_V * n + e] += sv[e];
  }
  __syncthreads();
}

__device__ __noinline__ void copy_phase(const SceneDesc& sc, float* __restrict__ dst, const float* __restrict__ src) {
  const int64_t n = sc.fstride;
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) dst[e] = src[e];
  __syncthreads();
}

template <int D, int B>
__device__ void gs_any(int mode, const SceneDesc& sc, float* X, const float* X0, float a, float cr, int iters,
                       bool serial) {
  if (serial) {
    if (mode == M_ROW) gs_serial<M_ROW>(sc, X, X0, a, cr, iters);
    else if (mode == M_COL) gs_serial<M_COL>(sc, X, X0, a, cr, iters);
    else gs_serial<M_PASSIVE>(sc, X, X0, a, cr, iters);
  } else {
    if (mode == M_ROW) gs_wavefront<M_ROW, D, B>(sc, X, X0, a, cr, iters);
    else if (mode == M_COL) gs_wavefront<M_COL, D, B>(sc, X, X0, a, cr, iters);
    else gs_wavefront<M_PASSIVE, D, B>(sc, X, X0, a, cr, iters);
  }
}

// One full frame (fluid.rs:437-524) on one scene, with the buffer roles of the
// reference: U0/V0 receive the diffused velocity, U/V double as pressure and
// divergence of the first projection, U0/V0 as those of the second.
template <int D, int B>
__device__ void frame_cyclic(const SceneDesc& sc) {
  float* f = sc.fld;
  const int64_t n = sc.fstride;
  float* Dn = f + F_D * n;
  float* Sc = f + F_SC * n;
  float* U = f + F_U * n;
  float* V = f + F_V * n;
  float* U0 = f + F_U0 * n;
  float* V0 = f + F_V0 * n;
  // The wavefront needs N >= D+3 (a cell's next-sweep load must be issued
  // after its store); more than one strip implies N > 64 >= 2*(B+D)+1.
  const bool serial = sc.N < D + 3;
  if (sc.has_src) add_sources_phase(sc);
  gs_any<D, B>(M_ROW, sc, U0, U, sc.a_visc, sc.cr_visc, sc.diff_iters, serial);  // fluid.rs:438
  gs_any<D, B>(M_COL, sc, V0, V, sc.a_visc, sc.cr_visc, sc.diff_iters, serial);  // fluid.rs:448
  div_phase(sc, U0, V0, V, U);                                                    // fluid.rs:459
  gs_any<D, B>(M_PASSIVE, sc, U, V, 1.0f, 0.25f, sc.press_iters, serial);
  grad_phase(sc, U0, V0, U);
  advect_phase(sc, M_ROW, U, U0, M_COL, V, V0, U0, V0);                           // fluid.rs:469-489
  div_phase(sc, U, V, V0, U0);                                                    // fluid.rs:491
  gs_any<D, B>(M_PASSIVE, sc, U0, V0, 1.0f, 0.25f, sc.press_iters, serial);
  grad_phase(sc, U, V, U0);
  gs_any<D, B>(M_PASSIVE, sc, Sc, Dn, sc.a_diff, sc.cr_diff, sc.diff_iters, serial);  // fluid.rs:501
  advect_phase(sc, M_PASSIVE, Dn, Sc, M_PASSIVE, nullptr, nullptr, U, V);        // fluid.rs:512
  copy_phase(sc, Sc, Dn);                                                         // fluid.rs:523
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
