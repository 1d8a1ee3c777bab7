// eqf_redblack.hip -- red-black mode of the Stam frame step (row-major layout).
//
// Every lin_solve (fluid.rs:301-325) updates the (i+j)-even cells of a sweep
// first and the odd cells second, then applies set_boundaries, so all cells
// of one colour are independent and the whole grid is updated in parallel by
// many workgroups.  Sweeps are temporally blocked in registers (k_rb_band):
// a workgroup loads a region (its output tile plus a halo of 3T cells), runs
// T full sweeps (red, black, set_boundaries) on it, and writes the tile back,
// so x and x0 cross HBM once per T sweeps instead of once per sweep.  Each
// full sweep shrinks the valid region by 3 cells: red reads black neighbours
// (1), black reads the new red neighbours (2), and set_boundaries copies a
// wall neighbour (3).  A pass reads one buffer and writes another (ping-pong).
//
// The other phases (divergence, gradient, advection with its row `break`,
// set_boundaries lists, copy) are grid-wide kernels.  One frame is captured
// once as a HIP graph and replayed.  Results are identical to the red-black
// f32 CPU oracle bit for bit and are compared with its f64 version for
// accuracy.  Compiled with -ffp-contract=off like the exact mode.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "eqf_internal.h"



namespace eqf {
namespace rb {

__device__ __forceinline__ float rclamp(float x, float lo, float hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}
__device__ __forceinline__ int sat_idx(float f, int S) {
  if (!(f > 0.0f)) return 0;
  if (f >= (float)(S - 1)) return S - 1;
  return (int)f;
}

// Lane shifts (DPP wave_shr:1 / wave_shl:1); the lane without a source reads
// 0 (bound_ctrl).  The empty asm keeps each shift where it is written,
// executed by the whole wave (a shift moved under a divergent branch would
// read disabled lanes), and keeps selects on its result from being folded
// into the move.
__device__ __forceinline__ float from_prev_lane(float v) {
  float r = __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x138, 0xf, 0xf, true));
  asm volatile("" : "+v"(r));
  return r;
}
__device__ __forceinline__ float from_next_lane(float v) {
  float r = __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x130, 0xf, 0xf, true));
  asm volatile("" : "+v"(r));
  return r;
}

// set_boundaries rule of one cell in one mode (k_rb_band): low nibble of the
// per-cell byte, bit 4 = the cell is interior (updated by the sweeps).
enum BandSel : uint32_t {
  BS_KEEP = 0,
  BS_NEG_LEFT = 1, BS_NEG_RIGHT = 2,   // AdjustRow, right wins (fluid.rs:151-164)
  BS_NEG_UP = 3, BS_NEG_DOWN = 4,      // AdjustColumn, up wins (fluid.rs:165-177)
  BS_LEFT = 5, BS_RIGHT = 6, BS_UP = 7, BS_DOWN = 8,  // Passive ring copies (fluid.rs:179-187)
  BS_UPD = 16,
};

// Red-black Gauss-Seidel with set_boundaries, T full sweeps per launch, all
// in registers.  A workgroup of NW waves owns a region of 128 columns x
// NW*R rows: lane l holds columns 2l and 2l+1 of its wave's band of R rows
// (x, x0 and the cell's rule byte in registers).  Horizontal neighbours come
// from the lane's other column or the neighbouring lane (DPP), vertical ones
// from the band's next register row; only the band's first and last rows
// cross waves, through LDS (double-buffered, one barrier per exchange).
// Red cells have (x + y) even; the region origin is chosen so that row i of
// every band has its red cell in column (i & 1) of the lane.  Every full
// sweep shrinks the valid region by 3 cells (red reads black, black reads the
// new red, set_boundaries reads a neighbour), so the launch stores the
// (128 - 6T) x (NW*R - 6T) interior of its region; the region reads one
// buffer and the tile is written to the other (ping-pong).  Corners are set
// once after the lin_solve (nothing reads them during the sweeps).
template <int MODE, int T, int NW, int R>
__global__ void __launch_bounds__(64 * NW) k_rb_band(const SceneDesc* __restrict__ dp, const float* __restrict__ src,
                                                     float* __restrict__ dst, const float* __restrict__ x0g,
                                                     const uint8_t* __restrict__ code,
                                                     const uint8_t* __restrict__ rowf, float a, float cr,
                                                     int nsweeps) {
  constexpr int H = 3 * T, W = 128, RH = NW * R;
  constexpr int TX = W - 2 * H, TY = RH - 2 * H;
  static_assert(TX > 0 && TY > 0 && TX % 2 == 0 && TY % 2 == 0 && R % 2 == 0, "band geometry");
  // [buffer][first / last row][column of the lane's pair][wave][lane]: one
  // float per lane and store, so a wave's 64 accesses hit 64 distinct banks
  __shared__ float edg[2][2][2][NW][64];
  const SceneDesc& d = *dp;
  const int S = d.S, N = d.N;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // Tile order: workgroups are dealt to the 8 XCDs round-robin by linear id;
  // XCD k takes a contiguous run of row-major tiles instead, so tiles that
  // share halo columns run together on one L2 (halves the HBM fetch).
  int bx, by;
  {
    const int ntx = gridDim.x, nt = gridDim.x * gridDim.y;
    const int L = blockIdx.x + blockIdx.y * ntx;
    const int per = nt >> 3, rem = nt & 7, xcd = L & 7;
    const int t = xcd * per + min(xcd, rem) + (L >> 3);
    bx = t % ntx;
    by = t / ntx;
  }
  const int gx0 = bx * TX - H;  // region origin; gx0 + gy0 is even (TX, TY, R even)
  const int gyb = by * TY - H;
  const int gy0 = gyb + w * R;          // this wave's band
  const int gxa = gx0 + 2 * lane;
  float x[R][2], x0[R][2];
  uint32_t cd[R];
  // Loads first, all of them, from clamped in-grid addresses (cells outside
  // the grid are replaced by 0 afterwards), so that one wait covers every
  // load of the region.
  uint32_t craw[R][2];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int64_t row = (int64_t)min(max(gy0 + i, 0), S - 1) * S;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int64_t g = row + min(max(gxa + b, 0), S - 1);
      x[i][b] = src[g];
      x0[i][b] = x0g[g];
      if constexpr (MODE != M_PASSIVE) craw[i][b] = code[g];
    }
  }
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int gy = gy0 + i;
    uint32_t c2 = 0;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int gx = gxa + b;
      const bool in = gx >= 0 && gx < S && gy >= 0 && gy < S;
      if (!in) {
        x[i][b] = 0.0f;
        x0[i][b] = 0.0f;
      }
      uint32_t sel = 0;
      if (gx >= 1 && gx <= N && gy >= 1 && gy <= N) {
        sel = BS_UPD;
        if constexpr (MODE != M_PASSIVE) {
          const uint32_t c = craw[i][b];
          if (MODE == M_ROW && (c & C_FLUID)) sel |= (c & C_RWALL) ? BS_NEG_RIGHT : (c & C_LWALL) ? BS_NEG_LEFT : 0;
          if (MODE == M_COL && (c & C_FLUID)) sel |= (c & C_UWALL) ? BS_NEG_UP : (c & C_DWALL) ? BS_NEG_DOWN : 0;
        }
      } else if (MODE == M_PASSIVE && in) {
        // a ring cell of column i / row j is rewritten iff that column / row
        // has a NoWall cell (ring cells only: a rare branch)
        if (gy == 0 && gx >= 1 && gx <= N) sel = (code[S + gx] & C_COLF) ? BS_DOWN : 0;
        else if (gy == S - 1 && gx >= 1 && gx <= N) sel = (code[(int64_t)(S - 2) * S + gx] & C_COLF) ? BS_UP : 0;
        else if (gx == 0 && gy >= 1 && gy <= N) sel = rowf[gy] ? BS_RIGHT : 0;
        else if (gx == S - 1 && gy >= 1 && gy <= N) sel = rowf[gy] ? BS_LEFT : 0;
      }
      c2 |= sel << (8 * b);
    }
    cd[i] = c2;
  }
  // Wave-uniform shortcuts: a band whose cells are all interior needs no
  // update mask, and one without set_boundaries rules skips that phase.
  uint32_t allu = 0xffffffffu, anys = 0;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    allu &= cd[i];
    anys |= cd[i];
  }
  const bool full = __builtin_amdgcn_ballot_w64((allu & (BS_UPD | (BS_UPD << 8))) != (BS_UPD | (BS_UPD << 8))) == 0;
  const bool rules = __builtin_amdgcn_ballot_w64((anys & 0x0f0fu) != 0) != 0;

  int par = 0;
  for (int k = 0; k < nsweeps; ++k) {
#pragma unroll
    for (int ph = 0; ph < 3; ++ph) {  // red, black, set_boundaries
      // the rule bytes are loop-invariant; keeping the compilers from hoisting
      // every per-cell lane mask out of the sweep loop (into scalar registers)
#pragma unroll
      for (int i = 0; i < R; ++i) asm volatile("" : "+v"(cd[i]));
      float2 ab = make_float2(0.0f, 0.0f), be = make_float2(0.0f, 0.0f);  // rows above / below the band
      if (ph < 2 || MODE != M_ROW) {
        edg[par][0][0][w][lane] = x[0][0];
        edg[par][0][1][w][lane] = x[0][1];
        edg[par][1][0][w][lane] = x[R - 1][0];
        edg[par][1][1][w][lane] = x[R - 1][1];
        __syncthreads();
        if (w > 0) ab = make_float2(edg[par][1][0][w - 1][lane], edg[par][1][1][w - 1][lane]);
        if (w < NW - 1) be = make_float2(edg[par][0][0][w + 1][lane], edg[par][0][1][w + 1][lane]);
        par ^= 1;
      }
      if (ph < 2) {
        // one colour: fluid.rs:315-320 on every cell with (x + y) % 2 == ph
        if (full) {
#pragma unroll
          for (int i = 0; i < R; ++i) {
            const int b = (i + ph) & 1;
            const float left = b ? x[i][0] : from_prev_lane(x[i][1]);
            const float right = b ? from_next_lane(x[i][0]) : x[i][1];
            const float up = i > 0 ? x[i - 1][b] : (b ? ab.y : ab.x);
            const float down = i < R - 1 ? x[i + 1][b] : (b ? be.y : be.x);
            x[i][b] = (x0[i][b] + a * (((right + left) + down) + up)) * cr;
          }
        } else {
#pragma unroll
          for (int i = 0; i < R; ++i) {
            const int b = (i + ph) & 1;
            const float left = b ? x[i][0] : from_prev_lane(x[i][1]);
            const float right = b ? from_next_lane(x[i][0]) : x[i][1];
            const float up = i > 0 ? x[i - 1][b] : (b ? ab.y : ab.x);
            const float down = i < R - 1 ? x[i + 1][b] : (b ? be.y : be.x);
            float nv = (x0[i][b] + a * (((right + left) + down) + up)) * cr;
            asm volatile("" : "+v"(nv));  // a select, not a branch around the update
            x[i][b] = ((cd[i] >> (8 * b)) & BS_UPD) ? nv : x[i][b];
          }
        }
      } else if (rules) {
        // set_boundaries (fluid.rs:141-188): writes only NoWall or ring
        // cells and reads only walls or interior cells, so in place is safe
#pragma unroll
        for (int i = 0; i < R; ++i) {
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            const uint32_t sel = (cd[i] >> (8 * b)) & 15u;
            float v = x[i][b];
            if constexpr (MODE == M_ROW || MODE == M_PASSIVE) {
              const float left = b ? x[i][0] : from_prev_lane(x[i][1]);
              const float right = b ? from_next_lane(x[i][0]) : x[i][1];
              if constexpr (MODE == M_ROW) {
                v = sel == BS_NEG_RIGHT ? -right : (sel == BS_NEG_LEFT ? -left : v);
              } else {
                v = sel == BS_RIGHT ? right : (sel == BS_LEFT ? left : v);
              }
            }
            if constexpr (MODE == M_COL || MODE == M_PASSIVE) {
              const float up = i > 0 ? x[i - 1][b] : (b ? ab.y : ab.x);
              const float down = i < R - 1 ? x[i + 1][b] : (b ? be.y : be.x);
              if constexpr (MODE == M_COL) {
                v = sel == BS_NEG_UP ? -up : (sel == BS_NEG_DOWN ? -down : v);
              } else {
                v = sel == BS_DOWN ? down : (sel == BS_UP ? up : v);
              }
            }
            x[i][b] = v;
          }
        }
      }
    }
  }

#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int gy = gy0 + i;
    if (gy < gyb + H || gy >= gyb + H + TY || gy >= S) continue;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int gx = gxa + b;
      if (gx >= gx0 + H && gx < gx0 + H + TX && gx < S) dst[(int64_t)gy * S + gx] = x[i][b];
    }
  }
}

// project, first half (fluid.rs:339-349)
__global__ void k_div(const SceneDesc* __restrict__ dp, const float* __restrict__ u, const float* __restrict__ v,
                      float* __restrict__ div, float* __restrict__ p) {
  const SceneDesc& d = *dp;
  const int S = d.S, N = d.N;
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1;
  if (i > N) return;
  const int q = j * S + i;
  div[q] = -0.5f * (u[q + 1] - u[q - 1] + v[q + S] - v[q - S]) / d.sf;
  p[q] = 0.0f;
}

// project, second half (fluid.rs:364-371)
__global__ void k_grad(const SceneDesc* __restrict__ dp, float* __restrict__ u, float* __restrict__ v,
                       const float* __restrict__ p) {
  const SceneDesc& d = *dp;
  const int S = d.S, N = d.N;
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1;
  if (i > N) return;
  const int q = j * S + i;
  u[q] = u[q] - 0.5f * (p[q + 1] - p[q - 1]) * d.sf;
  v[q] = v[q] - 0.5f * (p[q + S] - p[q - S]) * d.sf;
}

struct Trace {
  int i0, i1, j0, j1;
  float s0, s1, t0, t1;
  bool trig;
};
__device__ __forceinline__ Trace back_trace(const SceneDesc& d, int c, int j, float uu, float vv) {
  Trace r;
  float x = (float)c - d.dt0 * uu;
  float y = (float)j - d.dt0 * vv;
  x = rclamp(x, 0.5f, d.sfm1);
  y = rclamp(y, 0.5f, d.sfm1);
  const float i0 = floorf(x), i1 = i0 + 1.0f, j0 = floorf(y), j1 = j0 + 1.0f;
  r.s1 = x - i0;
  r.s0 = 1.0f - r.s1;
  r.t1 = y - j0;
  r.t0 = 1.0f - r.t1;
  r.trig = (i1 >= d.sf) || (j1 >= d.sf);
  r.i0 = sat_idx(i0, d.S);
  r.i1 = sat_idx(i1, d.S);
  r.j0 = sat_idx(j0, d.S);
  r.j1 = sat_idx(j1, d.S);
  return r;
}
__device__ __forceinline__ float gather(const float* __restrict__ d0, const Trace& r, int S) {
  return r.s0 * (r.t0 * d0[r.j0 * S + r.i0] + r.t1 * d0[r.j1 * S + r.i0]) +
         r.s1 * (r.t0 * d0[r.j0 * S + r.i1] + r.t1 * d0[r.j1 * S + r.i1]);
}

// advect, pass A: first column of each row whose back-trace reaches the far
// edge (the `break` of fluid.rs:420-423); N+1 when none.  One workgroup per
// row; each thread tests 8 columns per round with all 16 loads in flight.
__global__ void __launch_bounds__(256) k_adv_trig(const SceneDesc* __restrict__ dp, const float* __restrict__ u,
                                                  const float* __restrict__ v, int* __restrict__ trig) {
  const SceneDesc& d = *dp;
  const int S = d.S, N = d.N;
  const int j = blockIdx.x + 1;
  const int64_t row = (int64_t)j * S;
  constexpr int K = 8;
  int t = N + 1;
  for (int base = 1; base <= N; base += 256 * K) {
    float uu[K], vv[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = min(base + k * 256 + (int)threadIdx.x, N);  // in-row address; tested below
      uu[k] = u[row + i];
      vv[k] = v[row + i];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = base + k * 256 + threadIdx.x;
      if (i <= N && back_trace(d, i, j, uu[k], vv[k]).trig) t = min(t, i);
    }
    // later rounds hold only larger columns: the block stops together after
    // the first round in which any column triggered
    if (__syncthreads_or(t <= N)) break;
  }
  __shared__ int red[256];
  red[threadIdx.x] = t;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = min(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) trig[j] = red[0];
}

// advect, pass B (fluid.rs:398-430) for one or two fields sharing u, v.
__global__ void k_adv(const SceneDesc* __restrict__ dp, float* __restrict__ o0, const float* __restrict__ i0,
                      float* __restrict__ o1, const float* __restrict__ i1, const float* __restrict__ u,
                      const float* __restrict__ v, const int* __restrict__ trig) {
  const SceneDesc& d = *dp;
  const int S = d.S, N = d.N;
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1;
  if (i > N) return;
  const int tr = trig[j];
  if (i > tr) return;  // right of the break: buffer keeps its content
  const int q = j * S + i;
  if (i == tr && i == 1) {  // copies the ring cell d(0, j)
    o0[q] = o0[q - 1];
    if (o1) o1[q] = o1[q - 1];
    return;
  }
  const int c = i == tr ? i - 1 : i;  // the trigger cell copies its left neighbour's value
  const Trace r = back_trace(d, c, j, u[j * S + c], v[j * S + c]);
  o0[q] = gather(i0, r, S);
  if (o1) o1[q] = gather(i1, r, S);
}

// set_boundaries as lists (built on the host from the mask), then corners.
__global__ void k_bnd(const SceneDesc* __restrict__ dp, int mode0, float* __restrict__ x0, int mode1,
                      float* __restrict__ x1) {
  const SceneDesc& d = *dp;
  const int off[3] = {0, d.nbnd[0], d.nbnd[0] + d.nbnd[1]};
  const int n0 = d.nbnd[mode0], n1 = x1 ? d.nbnd[mode1] : 0;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n0) {
    const int2 p = d.bnd[off[mode0] + e];
    x0[p.x] = (mode0 == M_PASSIVE ? 1.0f : -1.0f) * x0[p.y];
  } else if (e < n0 + n1) {
    const int2 p = d.bnd[off[mode1] + e - n0];
    x1[p.x] = (mode1 == M_PASSIVE ? 1.0f : -1.0f) * x1[p.y];
  }
}
__global__ void k_corners(const SceneDesc* __restrict__ dp, float* __restrict__ x0, float* __restrict__ x1) {
  const SceneDesc& d = *dp;
  const int S = d.S;
  float* x = threadIdx.x == 0 ? x0 : x1;
  if (!x) return;
  x[0] = 0.5f * (x[1] + x[S]);
  x[(S - 1) * S] = 0.5f * (x[1 + (S - 1) * S] + x[(S - 2) * S]);
  x[S - 1] = 0.5f * (x[S - 2] + x[S - 1 + S]);
  x[S * S - 1] = 0.5f * (x[S - 2 + (S - 1) * S] + x[S - 1 + (S - 2) * S]);
}

__global__ void k_sources(const SceneDesc* __restrict__ dp) {
  const SceneDesc& d = *dp;
  const int64_t n = d.fstride;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)d.S * d.S) return;
  const int m = d.has_src;  // only the sources that were set (adding 0 would turn -0 into +0)
  if (m & 1) {
    const float a = d.src[e];
    d.fld[F_D * n + e] += a;
    d.fld[F_SC * n + e] += a;
  }
  if (m & 2) d.fld[F_U * n + e] += d.src[n + e];
  if (m & 4) d.fld[F_V * n + e] += d.src[2 * n + e];
}

__global__ void k_copy(float* __restrict__ dst, const float* __restrict__ src, int64_t n) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) dst[e] = src[e];
}

}  // namespace rb

// ------------------------------------------------------------------ host --
struct RBWork {
  float* tmp = nullptr;   // ping-pong field
  int* trig = nullptr;    // per-row break column
  hipGraphExec_t graph = nullptr;
  int tile = 0;           // band shape index (kBands)
  int tsweeps = 4;        // sweeps per launch
  uint64_t key = 0;       // parameters the graph was built for
};

// Band shapes (waves per workgroup x rows per wave); the region is always 128
// columns wide.
struct BandCfg {
  int nw, r;
};
static const BandCfg kBands[] = {{8, 16}, {8, 32}, {16, 16}};

template <int MODE, int T, int NW, int R>
static void launch_band(hipStream_t st, const SceneDesc* dd, const float* src, float* dst, const float* x0,
                        const uint8_t* code, const uint8_t* rowf, float a, float cr, int nsw, int S) {
  constexpr int TX = 128 - 6 * T, TY = NW * R - 6 * T;
  dim3 grid((S + TX - 1) / TX, (S + TY - 1) / TY);
  hipLaunchKernelGGL((rb::k_rb_band<MODE, T, NW, R>), grid, dim3(64 * NW), 0, st, dd, src, dst, x0, code, rowf, a, cr,
                     nsw);
}

template <int MODE, int NW, int R>
static void launch_band_t(hipStream_t st, int T, const SceneDesc* dd, const float* src, float* dst,
                          const float* x0, const uint8_t* code, const uint8_t* rowf, float a, float cr, int nsw,
                          int S) {
  if (T == 1) launch_band<MODE, 1, NW, R>(st, dd, src, dst, x0, code, rowf, a, cr, nsw, S);
  else if (T == 2) launch_band<MODE, 2, NW, R>(st, dd, src, dst, x0, code, rowf, a, cr, nsw, S);
  else if (T == 4) launch_band<MODE, 4, NW, R>(st, dd, src, dst, x0, code, rowf, a, cr, nsw, S);
  else if (T == 5) launch_band<MODE, 5, NW, R>(st, dd, src, dst, x0, code, rowf, a, cr, nsw, S);
  else if (T == 6) launch_band<MODE, 6, NW, R>(st, dd, src, dst, x0, code, rowf, a, cr, nsw, S);
  else launch_band<MODE, 8, NW, R>(st, dd, src, dst, x0, code, rowf, a, cr, nsw, S);
}

template <int MODE>
static void launch_gs_mode(hipStream_t st, int tile, int T, const SceneDesc* dd, const float* src, float* dst,
                           const float* x0, const uint8_t* code, const uint8_t* rowf, float a, float cr, int nsw,
                           int S) {
  if (tile == 1) launch_band_t<MODE, 8, 32>(st, T, dd, src, dst, x0, code, rowf, a, cr, nsw, S);
  else if (tile == 2) launch_band_t<MODE, 16, 16>(st, T, dd, src, dst, x0, code, rowf, a, cr, nsw, S);
  else launch_band_t<MODE, 8, 16>(st, T, dd, src, dst, x0, code, rowf, a, cr, nsw, S);
}

// lin_solve in red-black order: ceil(iters/T) tile passes, ping-ponging
// between x and the scratch field; an odd pass count ends with a copy back.
static void rb_lin_solve(eqf_scene* s, RBWork* w, int mode, float* x, const float* x0, float a, float cr,
                         int iters) {
  if (iters <= 0) return;
  const int S = s->S, T = w->tsweeps;
  const SceneDesc* dd = s->d_desc;
  float* a_ = x;
  float* b_ = w->tmp;
  for (int done = 0; done < iters;) {
    const int nsw = std::min(T, iters - done);
    const uint8_t *cd = s->d_code, *rf = s->d_rowf;
    if (mode == M_ROW) launch_gs_mode<M_ROW>(s->stream, w->tile, T, dd, a_, b_, x0, cd, rf, a, cr, nsw, S);
    else if (mode == M_COL) launch_gs_mode<M_COL>(s->stream, w->tile, T, dd, a_, b_, x0, cd, rf, a, cr, nsw, S);
    else launch_gs_mode<M_PASSIVE>(s->stream, w->tile, T, dd, a_, b_, x0, cd, rf, a, cr, nsw, S);
    std::swap(a_, b_);
    done += nsw;
  }
  if (a_ != x) {
    const int64_t n = (int64_t)S * S;
    hipLaunchKernelGGL(rb::k_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s->stream, x, a_, n);
  }
  // corners of the last set_boundaries (fluid.rs:265-271)
  hipLaunchKernelGGL(rb::k_corners, dim3(1), dim3(2), 0, s->stream, s->d_desc, x, (float*)nullptr);
}

static void rb_bnd(eqf_scene* s, int mode0, float* x0, int mode1, float* x1) {
  const int n = s->nbnd[mode0] + (x1 ? s->nbnd[mode1] : 0);
  if (n > 0)
    hipLaunchKernelGGL(rb::k_bnd, dim3((n + 255) / 256), dim3(256), 0, s->stream, s->d_desc, mode0, x0, mode1, x1);
  hipLaunchKernelGGL(rb::k_corners, dim3(1), dim3(2), 0, s->stream, s->d_desc, x0, x1);
}

static void rb_frame(eqf_scene* s, RBWork* w) {
  const int S = s->S, N = s->N;
  const SceneDesc& d = s->desc;
  float* f = s->d_fld;
  const int64_t n = s->fstride;
  float *Dn = f + F_D * n, *Sc = f + F_SC * n, *U = f + F_U * n, *V = f + F_V * n, *U0 = f + F_U0 * n,
        *V0 = f + F_V0 * n;
  hipStream_t st = s->stream;
  const dim3 rows((N + 255) / 256, N);
  if (d.has_src)
    hipLaunchKernelGGL(rb::k_sources, dim3((unsigned)(((int64_t)S * S + 255) / 256)), dim3(256), 0, st, s->d_desc);
  rb_lin_solve(s, w, M_ROW, U0, U, d.a_visc, d.cr_visc, d.diff_iters);  // fluid.rs:438
  rb_lin_solve(s, w, M_COL, V0, V, d.a_visc, d.cr_visc, d.diff_iters);  // fluid.rs:448
  // project(U0, V0, p=U, div=V), fluid.rs:459
  hipLaunchKernelGGL(rb::k_div, rows, dim3(256), 0, st, s->d_desc, U0, V0, V, U);
  rb_bnd(s, M_PASSIVE, V, M_PASSIVE, U);
  rb_lin_solve(s, w, M_PASSIVE, U, V, 1.0f, 0.25f, d.press_iters);
  hipLaunchKernelGGL(rb::k_grad, rows, dim3(256), 0, st, s->d_desc, U0, V0, U);
  rb_bnd(s, M_ROW, U0, M_COL, V0);
  // advect U from U0 and V from V0 with (U0, V0), fluid.rs:469-489
  hipLaunchKernelGGL(rb::k_adv_trig, dim3(N), dim3(256), 0, st, s->d_desc, U0, V0, w->trig);
  hipLaunchKernelGGL(rb::k_adv, rows, dim3(256), 0, st, s->d_desc, U, U0, V, V0, U0, V0, w->trig);
  rb_bnd(s, M_ROW, U, M_COL, V);
  // project(U, V, p=U0, div=V0), fluid.rs:491
  hipLaunchKernelGGL(rb::k_div, rows, dim3(256), 0, st, s->d_desc, U, V, V0, U0);
  rb_bnd(s, M_PASSIVE, V0, M_PASSIVE, U0);
  rb_lin_solve(s, w, M_PASSIVE, U0, V0, 1.0f, 0.25f, d.press_iters);
  hipLaunchKernelGGL(rb::k_grad, rows, dim3(256), 0, st, s->d_desc, U, V, U0);
  rb_bnd(s, M_ROW, U, M_COL, V);
  // density, fluid.rs:501-523
  rb_lin_solve(s, w, M_PASSIVE, Sc, Dn, d.a_diff, d.cr_diff, d.diff_iters);
  hipLaunchKernelGGL(rb::k_adv_trig, dim3(N), dim3(256), 0, st, s->d_desc, U, V, w->trig);
  hipLaunchKernelGGL(rb::k_adv, rows, dim3(256), 0, st, s->d_desc, Dn, Sc, (float*)nullptr, (const float*)nullptr,
                     U, V, w->trig);
  rb_bnd(s, M_PASSIVE, Dn, M_PASSIVE, nullptr);
  const int64_t nn = (int64_t)S * S;
  hipLaunchKernelGGL(rb::k_copy, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, Sc, Dn, nn);
}

void* rb_create(eqf_scene* s, std::string& err) {
  auto* w = new RBWork;
  const size_t n = (size_t)s->S * s->S;
  hipError_t e = hipMalloc(&w->tmp, sizeof(float) * n);
  if (e == hipSuccess) e = hipMalloc(&w->trig, sizeof(int) * (size_t)s->S);
  if (e != hipSuccess) {
    err = std::string("red-black workspace: ") + hipGetErrorString(e);
    rb_destroy(w);
    return nullptr;
  }
  // Default tiling; the autotuner (eqf_scene_tune) may override it.
  w->tile = 0;
  w->tsweeps = 4;
  if (const char* t = getenv("EQF_RB_TSWEEPS")) {
    const int v = atoi(t);
    w->tsweeps = v >= 8 ? 8 : v >= 4 ? 4 : v >= 2 ? 2 : 1;
  }
  if (const char* t = getenv("EQF_RB_TILE")) w->tile = std::max(0, std::min(2, atoi(t)));
  return w;
}

void rb_destroy(void* p) {
  auto* w = static_cast<RBWork*>(p);
  if (!w) return;
  if (w->graph) (void)hipGraphExecDestroy(w->graph);
  (void)hipFree(w->tmp);
  (void)hipFree(w->trig);
  delete w;
}

int rb_set_tiling(eqf_scene* s, int tile, int tsweeps) {
  auto* w = static_cast<RBWork*>(s->rb);
  if (!w) return EQF_EINVAL;
  if (tile < 0 || tile > 2 || tsweeps < 1 || tsweeps > 8 || tsweeps == 3 || tsweeps == 7) return EQF_EINVAL;
  if (w->tile != tile || w->tsweeps != tsweeps) {
    w->tile = tile;
    w->tsweeps = tsweeps;
    w->key = 0;
  }
  return EQF_OK;
}

int rb_step(eqf_scene* s, int64_t nframes, std::string& err) {
  auto* w = static_cast<RBWork*>(s->rb);
  // Parameters are baked into the graph: rebuild when they change.
  const SceneDesc& d = s->desc;
  uint64_t key = 1469598103934665603ull;
  auto mix = [&key](uint64_t v) { key = (key ^ v) * 1099511628211ull; };
  mix((uint64_t)d.diff_iters);
  mix((uint64_t)d.press_iters);
  mix((uint64_t)__builtin_bit_cast(uint32_t, d.a_visc));
  mix((uint64_t)__builtin_bit_cast(uint32_t, d.a_diff));
  mix((uint64_t)d.has_src);
  mix((uint64_t)d.nbnd[0] | ((uint64_t)d.nbnd[1] << 21) | ((uint64_t)d.nbnd[2] << 42));
  mix((uint64_t)w->tile * 16 + (uint64_t)w->tsweeps);
  if (!w->graph || w->key != key) {
    if (w->graph) (void)hipGraphExecDestroy(w->graph);
    w->graph = nullptr;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      err = std::string("capture: ") + hipGetErrorString(e);
      return EQF_EDEVICE;
    }
    rb_frame(s, w);
    hipError_t le = hipGetLastError();
    e = hipStreamEndCapture(s->stream, &g);
    if (le != hipSuccess || e != hipSuccess) {
      err = std::string("red-black frame capture: ") + hipGetErrorString(le != hipSuccess ? le : e);
      if (g) (void)hipGraphDestroy(g);
      return EQF_EDEVICE;
    }
    e = hipGraphInstantiate(&w->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      err = std::string("graph instantiate: ") + hipGetErrorString(e);
      return EQF_EDEVICE;
    }
    w->key = key;
  }
  for (int64_t f = 0; f < nframes; ++f) {
    hipError_t e = hipGraphLaunch(w->graph, s->stream);
    if (e != hipSuccess) {
      err = std::string("graph launch: ") + hipGetErrorString(e);
      return EQF_EDEVICE;
    }
  }
  return EQF_OK;
}

}  // namespace eqf
