// eqf_redblack.hip -- red-black mode of the Stam frame step (row-major layout).
//
// Every lin_solve (fluid.rs:301-325) updates the (i+j)-even cells of a sweep
// first and the odd cells second, then applies set_boundaries, so all cells
// of one colour are independent and the whole grid is updated in parallel by
// many workgroups.  Sweeps are temporally blocked in registers (k_rb_band):
// a workgroup loads a region (its output tile plus a halo of 3T cells on the
// sides that face other tiles), runs T full sweeps (red, black,
// set_boundaries) on it, and writes the tile back, so x and x0 cross HBM once
// per T sweeps instead of once per sweep.  Each full sweep shrinks the valid
// region by 3 cells on every side with cells beyond it: red reads black
// neighbours (1), black reads the new red neighbours (2), and set_boundaries
// copies a wall neighbour (3).  A side on the grid's edge has nothing beyond
// it, so the first and last tiles of a row or column of tiles reach the edge
// with no halo there.  A pass reads one buffer and writes another (ping-pong).
//
// The other phases (divergence, gradient, advection with its row `break`,
// set_boundaries lists, copy) are grid-wide kernels.  One frame is captured
// once as a HIP graph and replayed.  Results are identical to the red-black
// f32 CPU oracle bit for bit and are compared with its f64 version for
// accuracy.  Compiled with -ffp-contract=off like the exact mode.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "eqf_internal.h"

#define GAS __attribute__((address_space(1)))



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
// Unpinned shifts for straight-line code with every lane active and no
// select on the result: the compiler may fold them into the add that uses them.
__device__ __forceinline__ float prev_lane_add(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float next_lane_add(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x130, 0xf, 0xf, true));
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

// One scene of a red-black launch group: a single scene, or every scene of
// a batch.  Each kernel below covers all scenes of its group in one launch.
// A workgroup finds its scene and its place in it with one load from a map
// of the launch's work units (band tile, block of CELL_BLOCK interior cells
// or of CELL_BLOCK cells ring included -> {scene, index in the scene}); the
// set_boundaries list kernel, whose threads straddle scenes, searches a
// prefix table instead.
struct RbScene {
  SceneDesc d;  // the scene's descriptor (a copy): fields, sizes, coefficients, rule bytes, lists
  float* tmp;   // its ping-pong field
  int* trig;    // its per-row break columns
};
// Field roles in kernel arguments: F_* (eqf_layout.h), the ping-pong field, or none.
constexpr int ROLE_TMP = NFIELDS;
constexpr int ROLE_NONE = -1;
// cells per workgroup of the whole-field kernels (256 threads x 4 cells)
constexpr int CELL_K = 4;
constexpr int CELL_BLOCK = 256 * CELL_K;

// The CELL_K interior cells of this thread in interior block `blk` of a
// scene: flat indices blk*CELL_BLOCK + k*256 + threadIdx.x over the N x N
// interior in row-major order, as 1-based (i, j).  Indices past N*N are
// flagged !ok and clamped to cell (N, N), so their loads stay in the grid.
__device__ __forceinline__ void interior_cells(int blk, int N, int (&ci)[CELL_K], int (&cj)[CELL_K],
                                               bool (&ok)[CELL_K]) {
  const int64_t idx0 = (int64_t)blk * CELL_BLOCK, nn = (int64_t)N * N;
  const int j0 = (int)(idx0 / N), i0 = (int)(idx0 - (int64_t)j0 * N);
  const float inv = 1.0f / (float)N;
#pragma unroll
  for (int k = 0; k < CELL_K; ++k) {
    const int off = i0 + k * 256 + (int)threadIdx.x;  // < N + CELL_BLOCK: exact in f32
    int q = (int)((float)off * inv);
    int r = off - q * N;
    if (r < 0) {  // the rounded reciprocal is off by at most one row
      --q;
      r += N;
    } else if (r >= N) {
      ++q;
      r -= N;
    }
    ok[k] = idx0 + k * 256 + threadIdx.x < nn;
    ci[k] = ok[k] ? r + 1 : N;
    cj[k] = ok[k] ? j0 + q + 1 : N;
  }
}
__device__ __forceinline__ float* role_ptr(const RbScene& s, int role) {
  return role == ROLE_TMP ? s.tmp : s.d.fld + (int64_t)role * s.d.fstride;
}
// The last scene j with start[j] <= g (start[0] = 0; scenes without work
// units have start[j] == start[j+1] and are skipped).
template <class I>
__device__ __forceinline__ int find_scene(const I* __restrict__ start, int n, I g) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
// Tiles along one dimension of S cells for regions of W cells and output
// tiles of TX: region b starts at b*TX; the first needs no halo before it
// (the grid's edge), the last must reach the far edge.
__host__ __device__ __forceinline__ int band_span(int S, int W, int TX) {
  return S <= W ? 1 : (S - W + TX - 1) / TX + 1;
}

// lin_solve coefficients (a, 1/c) of a scene: velocity diffusion, density
// diffusion, pressure (fluid.rs:286-294, 311, 356).
__device__ __forceinline__ void coefs(const SceneDesc& d, int coef, float& a, float& cr) {
  a = coef == 0 ? d.a_visc : coef == 1 ? d.a_diff : 1.0f;
  cr = coef == 0 ? d.cr_visc : coef == 1 ? d.cr_diff : 0.25f;
}

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
// new red, set_boundaries reads a neighbour) on each side that has cells
// beyond it, so the launch stores the (128 - 6T) x (NW*R - 6T) interior of its
// region, extended to the grid's edge on the sides where the region reaches
// it (tile b of a dimension: region origin b*TX, so the first region starts
// at the edge and the last one passes the far edge); the region reads one
// buffer and the tile is written to the other (ping-pong).  Corners are set
// once after the lin_solve (nothing reads them during the sweeps).
template <int MODE, int T, int NW, int R>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu((NW == 8 && R <= 10) ? 8 : (NW == 8 && R <= 14) ? 6 : (NW == 8 && R <= 18) ? 4 : 1))) k_rb_band(const RbScene* __restrict__ scs, const int2* __restrict__ tmap,
                                                     int nt, int src_role, int dst_role, int x0_role, int coef,
                                                     int nsweeps) {
  constexpr int H = 3 * T, W = 128, RH = NW * R;
  constexpr int TX = W - 2 * H, TY = RH - 2 * H;
  static_assert(TX > 0 && TY > 0 && TX % 2 == 0 && TY % 2 == 0 && R % 2 == 0, "band geometry");
  // [buffer][first / last row][column of the lane's pair][wave][lane]: one
  // float per lane and store, so a wave's 64 accesses hit 64 distinct banks
  __shared__ float edg[2][2][2][NW][64];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // Tile order: workgroups are dealt to the 8 XCDs round-robin by linear id;
  // XCD k takes a contiguous run of the row-major tiles (scene after scene)
  // instead, so tiles that share halo columns run together on one L2.
  int t;
  {
    const int L = blockIdx.x;  // nt workgroups, one per tile
    const int per = nt >> 3, rem = nt & 7, xcd = L & 7;
    t = __builtin_amdgcn_readfirstlane(xcd * per + min(xcd, rem) + (L >> 3));
  }
  const int2 tm = tmap[t];
  const int js = __builtin_amdgcn_readfirstlane(tm.x), tl = __builtin_amdgcn_readfirstlane(tm.y);
  const RbScene& sc = scs[js];
  const SceneDesc& d = sc.d;
  const int S = d.S, N = d.N;
  const int tiles_x = band_span(S, W, TX), tiles_y = band_span(S, RH, TY);
  const int bx = tl % tiles_x, by = tl / tiles_x;
  // Global-memory views: a row's base address is wave-uniform (scalar
  // registers), a lane's column a 32-bit offset.
  GAS const float* __restrict__ src = (GAS const float*)role_ptr(sc, src_role);
  GAS float* __restrict__ dst = (GAS float*)role_ptr(sc, dst_role);
  GAS const float* __restrict__ x0g = (GAS const float*)role_ptr(sc, x0_role);
  GAS const uint8_t* __restrict__ code = (GAS const uint8_t*)d.code;
  float a, cr;
  coefs(d, coef, a, cr);
  const int gx0 = bx * TX;  // region origin; gx0 + gy0 is even (TX, TY, R even)
  const int gyb = by * TY;
  const int gy0 = gyb + w * R;          // this wave's band
  const int gxa = gx0 + 2 * lane;
  // whether this lane's two columns are in the grid (cells outside it are
  // replaced by 0 after the loads)
  bool cin[2];
#pragma unroll
  for (int b = 0; b < 2; ++b) cin[b] = gxa + b < S;  // gxa >= 0
  float x[R][2], x0[R][2];
  uint32_t cd[R];
  // Loads first, all of them, so that one wait covers every load of the
  // region.  Addresses: the tile's first row as one wave-uniform base per
  // array, a row's cells as a 32-bit offset from it (rows clamped into the
  // grid).  A lane's column pair is one 8-byte load per field at
  // min(gxa, S - 2) (gxa >= 0 is even): a pair that straddles the far edge
  // (odd S, gxa = S - 1) reads (S - 2, S - 1) and takes column S - 1 from the
  // second half; pairs past the edge read in-grid cells that are zeroed below.
  // The pair's rule bytes are one 2-byte load (rule rows have an even
  // stride Sc, so min(gxa, Sc - 2) is even; the straddling pair reads column
  // S - 1 and the row's pad byte).
  const int Sc = rm_code_stride(S);
  const uint32_t pc = (uint32_t)min(gxa, S - 2), pcc = (uint32_t)min(gxa, Sc - 2);
  const bool hi = gxa > S - 2;
  GAS const char* sb = (GAS const char*)(src + (int64_t)gyb * S);  // gyb <= S - 1
  GAS const char* xb = (GAS const char*)(x0g + (int64_t)gyb * S);
  GAS const uint8_t* cb = code + (int64_t)gyb * Sc;
  typedef float f2 __attribute__((ext_vector_type(2), aligned(4)));
  uint32_t cw[R];  // rule bytes of the pair, first column low
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const uint32_t rr = (uint32_t)(min(gy0 + i, S - 1) - gyb);  // gy0 >= gyb
    const uint32_t eo = rr * (uint32_t)S + pc;
    const f2 xv = *(GAS const f2*)(sb + 4u * eo);
    const f2 x0v = *(GAS const f2*)(xb + 4u * eo);
    x[i][0] = hi ? xv.y : xv.x;
    x[i][1] = xv.y;
    x0[i][0] = hi ? x0v.y : x0v.x;
    x0[i][1] = x0v.y;
    cw[i] = *(GAS const uint16_t*)(cb + (rr * (uint32_t)Sc + pcc));
  }
  // The rule byte (eqf_layout.h RuleBits) of each cell, as this mode's
  // selector; cells outside the grid (clamped loads) get none.
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int gy = gy0 + i;
    const bool rin = gy >= 0 && gy < S;  // wave-uniform
    uint32_t c2 = 0;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const bool inside = rin && cin[b];
      x[i][b] = inside ? x[i][b] : 0.0f;
      x0[i][b] = inside ? x0[i][b] : 0.0f;
      const uint32_t c = (cw[i] >> (8 * b)) & 255u;
      uint32_t sel = (c & R_INTERIOR) ? (uint32_t)BS_UPD : 0u;
      if constexpr (MODE == M_ROW) sel |= (c >> R_ROW_SHIFT) & 3u;  // BS_NEG_LEFT / BS_NEG_RIGHT
      if constexpr (MODE == M_COL) {
        const uint32_t r = (c >> R_COL_SHIFT) & 3u;
        sel |= r ? r + (BS_NEG_UP - R_NEG_FIRST) : 0u;
      }
      if constexpr (MODE == M_PASSIVE) {
        const uint32_t r = (c >> R_PAS_SHIFT) & 7u;
        sel |= r ? r + (BS_LEFT - R_COPY_LEFT) : 0u;
      }
      c2 |= (inside ? sel : 0u) << (8 * b);
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
  // two rows' selectors per register (row 2k low half, 2k+1 high half)
  uint32_t cp[R / 2];
#pragma unroll
  for (int k = 0; k < R / 2; ++k) cp[k] = cd[2 * k] | (cd[2 * k + 1] << 16);
  const bool full = __builtin_amdgcn_ballot_w64((allu & (BS_UPD | (BS_UPD << 8))) != (BS_UPD | (BS_UPD << 8))) == 0;
  const bool rules = __builtin_amdgcn_ballot_w64((anys & 0x0f0fu) != 0) != 0;

  int par = 0;
  // nsweeps >= 1: a loop without an entry test keeps the x0 loads above in the
  // first block (a skippable loop lets them sink below the other loads' wait)
  int k = 0;
  do {
#pragma unroll
    for (int ph = 0; ph < 3; ++ph) {  // red, black, set_boundaries
      // the rule bytes are loop-invariant; keeping the compilers from hoisting
      // every per-cell lane mask out of the sweep loop (into scalar registers)
#pragma unroll
      for (int k = 0; k < R / 2; ++k) asm volatile("" : "+v"(cp[k]));
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
            const float left = b ? x[i][0] : prev_lane_add(x[i][1]);
            const float right = b ? next_lane_add(x[i][0]) : x[i][1];
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
            x[i][b] = (cp[i >> 1] & (BS_UPD << (16 * (i & 1) + 8 * b))) ? nv : x[i][b];
          }
        }
      } else if constexpr (MODE == M_PASSIVE) {
        // Passive set_boundaries (fluid.rs:179-187) writes ring cells only,
        // each from its interior neighbour (column 0 from column 1, column
        // S-1 from S-2, row 0 from row 1, row S-1 from S-2, where the rule
        // byte says its row / column holds a NoWall cell; corners have none),
        // so only the tiles on the grid's edge have work, at known places.
        if (bx == 0) {  // column 0: the first column of lane 0
#pragma unroll
          for (int i = 0; i < R; ++i) x[i][0] = ((cp[i >> 1] >> (16 * (i & 1))) & 15u) == BS_RIGHT ? x[i][1] : x[i][0];
        }
        if (bx == tiles_x - 1) {  // column S-1
          if (S & 1) {  // the first column of its lane: its left neighbour is in the lane before
#pragma unroll
            for (int i = 0; i < R; ++i) {
              const float l = from_prev_lane(x[i][1]);
              x[i][0] = ((cp[i >> 1] >> (16 * (i & 1))) & 15u) == BS_LEFT ? l : x[i][0];
            }
          } else {
#pragma unroll
            for (int i = 0; i < R; ++i) x[i][1] = ((cp[i >> 1] >> (16 * (i & 1) + 8)) & 15u) == BS_LEFT ? x[i][0] : x[i][1];
          }
        }
        if (gy0 == 0) {  // row 0 (R >= 2)
#pragma unroll
          for (int b = 0; b < 2; ++b) x[0][b] = ((cp[0] >> (8 * b)) & 15u) == BS_DOWN ? x[1][b] : x[0][b];
        }
        if (gy0 + R > S - 1) {  // row S-1 (only its cells carry BS_UP)
#pragma unroll
          for (int i = R - 1; i >= 0; --i) {
#pragma unroll
            for (int b = 0; b < 2; ++b) {
              const float up = i > 0 ? x[i - 1][b] : (b ? ab.y : ab.x);
              x[i][b] = ((cp[i >> 1] >> (16 * (i & 1) + 8 * b)) & 15u) == BS_UP ? up : x[i][b];
            }
          }
        }
      } else if (rules) {
        // set_boundaries AdjustRow / AdjustColumn (fluid.rs:141-178): writes
        // only NoWall cells and reads only walls, so in place is safe
#pragma unroll
        for (int i = 0; i < R; ++i) {
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            const uint32_t sel = (cp[i >> 1] >> (16 * (i & 1) + 8 * b)) & 15u;
            float v = x[i][b];
            if constexpr (MODE == M_ROW) {
              const float left = b ? x[i][0] : from_prev_lane(x[i][1]);
              const float right = b ? from_next_lane(x[i][0]) : x[i][1];
              v = sel == BS_NEG_RIGHT ? -right : (sel == BS_NEG_LEFT ? -left : v);
            } else {
              const float up = i > 0 ? x[i - 1][b] : (b ? ab.y : ab.x);
              const float down = i < R - 1 ? x[i + 1][b] : (b ? be.y : be.x);
              v = sel == BS_NEG_UP ? -up : (sel == BS_NEG_DOWN ? -down : v);
            }
            x[i][b] = v;
          }
        }
      }
    }
  } while (++k < nsweeps);

  // (columns recomputed here rather than kept live through the sweeps)
  // Output tile: [gx0 + H, gx0 + H + TX) x [gyb + H, gyb + H + TY), from the
  // grid's edge for the first tile of a dimension and to its far edge for
  // the last one.
  const int gxl = bx * TX + 2 * (int)(threadIdx.x & 63);
  const int xlo = bx == 0 ? 0 : bx * TX + H, xhi = bx == tiles_x - 1 ? S : bx * TX + H + TX;
  const int ylo = by == 0 ? 0 : gyb + H, yhi = by == tiles_y - 1 ? S : gyb + H + TY;
  bool cout[2];  // this lane's columns lie in the output tile
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int gx = gxl + b;
    cout[b] = gx >= xlo && gx < xhi;
  }
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int gy = gy0 + i;
    if (gy < ylo || gy >= yhi) continue;  // wave-uniform
    GAS char* dp = (GAS char*)(dst + (int64_t)gy * S);
    if (cout[0] && cout[1]) {  // one 8-byte store; only the tile's edge lanes store singles
      f2 v;
      v.x = x[i][0];
      v.y = x[i][1];
      *(GAS f2*)(dp + 4u * (uint32_t)gxl) = v;
    } else {
#pragma unroll
      for (int b = 0; b < 2; ++b)
        if (cout[b]) *(GAS float*)(dp + 4u * (uint32_t)(gxl + b)) = x[i][b];
    }
  }
}

// project, first half (fluid.rs:339-349).  imap: workgroup -> {scene, block
// of CELL_BLOCK interior cells}; every load of a thread's cells is issued
// before its stores.  Also resets every row's break column to N+1 for the
// advect that follows this projection (the cell in column 1 does it).
__global__ void __launch_bounds__(256) k_div(const RbScene* __restrict__ scs, const int2* __restrict__ imap,
                                             int u_role, int v_role, int div_role, int p_role) {
  const int2 im = imap[blockIdx.x];
  const RbScene& sc = scs[im.x];
  const int S = sc.d.S, N = sc.d.N;
  int ci[CELL_K], cj[CELL_K];
  bool ok[CELL_K];
  interior_cells(im.y, N, ci, cj, ok);
  const float* __restrict__ u = role_ptr(sc, u_role);
  const float* __restrict__ v = role_ptr(sc, v_role);
  float ur[CELL_K], ul[CELL_K], vd[CELL_K], vu[CELL_K];
#pragma unroll
  for (int k = 0; k < CELL_K; ++k) {
    const int q = cj[k] * S + ci[k];
    ur[k] = u[q + 1];
    ul[k] = u[q - 1];
    vd[k] = v[q + S];
    vu[k] = v[q - S];
  }
  float* __restrict__ dv = role_ptr(sc, div_role);
  float* __restrict__ p = role_ptr(sc, p_role);
  const float sf = sc.d.sf;
#pragma unroll
  for (int k = 0; k < CELL_K; ++k) {
    if (!ok[k]) continue;
    const int q = cj[k] * S + ci[k];
    dv[q] = -0.5f * (ur[k] - ul[k] + vd[k] - vu[k]) / sf;
    p[q] = 0.0f;
    if (ci[k] == 1) sc.trig[cj[k]] = N + 1;
  }
}

// project, second half (fluid.rs:364-371); imap as for k_div.
__global__ void __launch_bounds__(256) k_grad(const RbScene* __restrict__ scs, const int2* __restrict__ imap,
                                              int u_role, int v_role, int p_role) {
  const int2 im = imap[blockIdx.x];
  const RbScene& sc = scs[im.x];
  const int S = sc.d.S, N = sc.d.N;
  int ci[CELL_K], cj[CELL_K];
  bool ok[CELL_K];
  interior_cells(im.y, N, ci, cj, ok);
  float* __restrict__ u = role_ptr(sc, u_role);
  float* __restrict__ v = role_ptr(sc, v_role);
  const float* __restrict__ p = role_ptr(sc, p_role);
  float uq[CELL_K], vq[CELL_K], pr[CELL_K], pl[CELL_K], pd[CELL_K], pu[CELL_K];
#pragma unroll
  for (int k = 0; k < CELL_K; ++k) {
    const int q = cj[k] * S + ci[k];
    uq[k] = u[q];
    vq[k] = v[q];
    pr[k] = p[q + 1];
    pl[k] = p[q - 1];
    pd[k] = p[q + S];
    pu[k] = p[q - S];
  }
  const float sf = sc.d.sf;
#pragma unroll
  for (int k = 0; k < CELL_K; ++k) {
    if (!ok[k]) continue;
    const int q = cj[k] * S + ci[k];
    u[q] = uq[k] - 0.5f * (pr[k] - pl[k]) * sf;
    v[q] = vq[k] - 0.5f * (pd[k] - pu[k]) * sf;
  }
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
// edge (the `break` of fluid.rs:420-423), as the minimum of the triggering
// columns (k_div set every row to N+1 "none" before).  imap as for k_div.  A
// wave's lanes hold consecutive cells, so only the leftmost triggering cell
// of each row within a wave updates the row's minimum.
__global__ void __launch_bounds__(256) k_adv_trig(const RbScene* __restrict__ scs, const int2* __restrict__ imap,
                                                  int u_role, int v_role) {
  const int2 im = imap[blockIdx.x];
  const RbScene& sc = scs[im.x];
  const SceneDesc& d = sc.d;
  const int S = d.S, N = d.N;
  int ci[CELL_K], cj[CELL_K];
  bool ok[CELL_K];
  interior_cells(im.y, N, ci, cj, ok);
  const float* __restrict__ u = role_ptr(sc, u_role);
  const float* __restrict__ v = role_ptr(sc, v_role);
  float uu[CELL_K], vv[CELL_K];
#pragma unroll
  for (int k = 0; k < CELL_K; ++k) {
    const int q = cj[k] * S + ci[k];
    uu[k] = u[q];
    vv[k] = v[q];
  }
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < CELL_K; ++k) {
    const bool t = ok[k] && back_trace(d, ci[k], cj[k], uu[k], vv[k]).trig;
    const uint64_t m = __builtin_amdgcn_ballot_w64(t);
    if (t) {
      // lanes max(0, lane - (i-1)) .. lane-1 hold this row's cells left of this one
      const int first = max(0, lane - (ci[k] - 1));
      const uint64_t left = ((1ull << lane) - 1) & ~((1ull << first) - 1);
      if ((m & left) == 0) atomicMin(&sc.trig[cj[k]], ci[k]);
    }
  }
}

// advect, pass B (fluid.rs:398-430) for one or two fields sharing u, v
// (o1_role ROLE_NONE: one field); imap as for k_div.  The break columns and
// velocities of a thread's cells are loaded together, then every gather, then
// the stores.  Cells right of their row's break keep the buffer's content;
// their (clamped, in-grid) gathers are computed and dropped.
__global__ void __launch_bounds__(256) k_adv(const RbScene* __restrict__ scs, const int2* __restrict__ imap,
                                             int o0_role, int i0_role, int o1_role, int i1_role, int u_role,
                                             int v_role) {
  const int2 im = imap[blockIdx.x];
  const RbScene& sc = scs[im.x];
  const int S = sc.d.S, N = sc.d.N;
  int ci[CELL_K], cj[CELL_K];
  bool ok[CELL_K];
  interior_cells(im.y, N, ci, cj, ok);
  float* __restrict__ o0 = role_ptr(sc, o0_role);
  float* __restrict__ o1 = o1_role == ROLE_NONE ? nullptr : role_ptr(sc, o1_role);
  const float* __restrict__ d0 = role_ptr(sc, i0_role);
  const float* __restrict__ d1 = o1 ? role_ptr(sc, i1_role) : d0;
  const float* __restrict__ u = role_ptr(sc, u_role);
  const float* __restrict__ v = role_ptr(sc, v_role);
  int tr[CELL_K];
  float uu[CELL_K], vv[CELL_K];
#pragma unroll
  for (int k = 0; k < CELL_K; ++k) {
    const int q = cj[k] * S + ci[k];
    tr[k] = sc.trig[cj[k]];
    uu[k] = u[q];
    vv[k] = v[q];
  }
  float r0[CELL_K], r1[CELL_K];
#pragma unroll
  for (int k = 0; k < CELL_K; ++k) {
    const int i = ci[k], j = cj[k], row = j * S;
    int c = i;
    if (i == tr[k] && i > 1) {  // the trigger cell copies its left neighbour's value (at most one per row)
      c = i - 1;
      uu[k] = u[row + c];
      vv[k] = v[row + c];
    }
    const Trace t = back_trace(sc.d, c, j, uu[k], vv[k]);
    r0[k] = gather(d0, t, S);
    r1[k] = gather(d1, t, S);
    if (i == tr[k] && i == 1) {  // trigger in column 1: copies the ring cell d(0, j)
      r0[k] = o0[row];
      if (o1) r1[k] = o1[row];
    }
  }
#pragma unroll
  for (int k = 0; k < CELL_K; ++k) {
    if (!ok[k] || ci[k] > tr[k]) continue;  // right of the break: buffer keeps its content
    const int q = cj[k] * S + ci[k];
    o0[q] = r0[k];
    if (o1) o1[q] = r1[k];
  }
}

// set_boundaries as lists (built from the mask on the device): entry e of a
// scene's mode0 list then of its mode1 list (role1 ROLE_NONE: none); bstart =
// prefix of the scenes' entry counts for this (mode0, mode1).
__global__ void k_bnd(const RbScene* __restrict__ scs, const int* __restrict__ bstart, int nsc, int mode0,
                      int role0, int mode1, int role1) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= bstart[nsc]) return;
  const int js = find_scene(bstart, nsc, e);
  const RbScene& sc = scs[js];
  const SceneDesc& d = sc.d;
  const int off[3] = {0, d.nbnd[0], d.nbnd[0] + d.nbnd[1]};
  const int k = e - bstart[js];
  const bool first = k < d.nbnd[mode0];
  const int mode = first ? mode0 : mode1;
  float* x = role_ptr(sc, first ? role0 : role1);
  const int2 p = d.bnd[off[mode] + (first ? k : k - d.nbnd[mode0])];
  x[p.x] = (mode == M_PASSIVE ? 1.0f : -1.0f) * x[p.y];
}

// set_boundaries corners (fluid.rs:265-271): one thread per scene and field.
__global__ void k_corners(const RbScene* __restrict__ scs, int nsc, int role0, int role1) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int js = t >> 1;
  const int role = (t & 1) ? role1 : role0;
  if (js >= nsc || role == ROLE_NONE) return;
  const RbScene& sc = scs[js];
  const int S = sc.d.S;
  float* x = role_ptr(sc, role);
  x[0] = 0.5f * (x[1] + x[S]);
  x[(S - 1) * S] = 0.5f * (x[1 + (S - 1) * S] + x[(S - 2) * S]);
  x[S - 1] = 0.5f * (x[S - 2] + x[S - 1 + S]);
  x[S * S - 1] = 0.5f * (x[S - 2 + (S - 1) * S] + x[S - 1 + (S - 2) * S]);
}

// Frame sources (only the fields that were set; adding 0 would turn -0 into
// +0) over every cell of every scene.  cmap: workgroup -> {scene, block of
// CELL_BLOCK cells}; each thread loads its CELL_K cells (256 apart) before
// storing any.
__device__ __forceinline__ void add_field(float* __restrict__ f, const float* __restrict__ a, int64_t e0,
                                          int64_t cells) {
  float x[CELL_K], y[CELL_K];
#pragma unroll
  for (int k = 0; k < CELL_K; ++k) {
    const int64_t e = min(e0 + k * 256, cells - 1);
    x[k] = f[e];
    y[k] = a[e];
  }
#pragma unroll
  for (int k = 0; k < CELL_K; ++k)
    if (e0 + k * 256 < cells) f[e0 + k * 256] = x[k] + y[k];
}
__global__ void __launch_bounds__(256) k_sources(const RbScene* __restrict__ scs, const int2* __restrict__ cmap) {
  const int2 cm = cmap[blockIdx.x];
  const SceneDesc& d = scs[cm.x].d;
  const int m = d.has_src;
  if (!m) return;
  const int64_t n = d.fstride, cells = (int64_t)d.S * d.S;
  const int64_t e0 = (int64_t)cm.y * CELL_BLOCK + threadIdx.x;
  if (m & 1) {
    float a[CELL_K], x[CELL_K], y[CELL_K];
#pragma unroll
    for (int k = 0; k < CELL_K; ++k) {
      const int64_t e = min(e0 + k * 256, cells - 1);
      a[k] = d.src[e];
      x[k] = d.fld[F_D * n + e];
      y[k] = d.fld[F_SC * n + e];
    }
#pragma unroll
    for (int k = 0; k < CELL_K; ++k)
      if (e0 + k * 256 < cells) {
        d.fld[F_D * n + e0 + k * 256] = x[k] + a[k];
        d.fld[F_SC * n + e0 + k * 256] = y[k] + a[k];
      }
  }
  if (m & 2) add_field(d.fld + F_U * n, d.src + n, e0, cells);
  if (m & 4) add_field(d.fld + F_V * n, d.src + 2 * n, e0, cells);
}

// dst = src over every cell of every scene (cmap as for k_sources).
__global__ void __launch_bounds__(256) k_copy(const RbScene* __restrict__ scs, const int2* __restrict__ cmap,
                                              int dst_role, int src_role) {
  const int2 cm = cmap[blockIdx.x];
  const RbScene& sc = scs[cm.x];
  const int64_t cells = (int64_t)sc.d.S * sc.d.S;
  const int64_t e0 = (int64_t)cm.y * CELL_BLOCK + threadIdx.x;
  const float* __restrict__ src = role_ptr(sc, src_role);
  float* __restrict__ dst = role_ptr(sc, dst_role);
  float x[CELL_K];
#pragma unroll
  for (int k = 0; k < CELL_K; ++k) x[k] = src[min(e0 + k * 256, cells - 1)];
#pragma unroll
  for (int k = 0; k < CELL_K; ++k)
    if (e0 + k * 256 < cells) dst[e0 + k * 256] = x[k];
}

}  // namespace rb

// ------------------------------------------------------------------ host --
using rb::RbScene;
// A launch group: its scenes and the prefix tables of their work units, in
// one device buffer written before any frame is captured.
struct RbGroup {
  std::vector<eqf_scene*> scenes;  // largest first
  hipStream_t stream = nullptr;    // null: each frame runs on scenes[0]'s stream
  void* d_tab = nullptr;
  size_t tab_bytes = 0;
  std::vector<char> h_tab;  // host image of the tables
  RbScene* d_scs = nullptr;
  // work unit -> {scene, index in it}: band tiles, interior cell blocks, cell blocks
  int2 *d_tmap = nullptr, *d_imap = nullptr, *d_cmap = nullptr;
  int *d_bpp = nullptr, *d_brc = nullptr, *d_bp = nullptr;  // prefixes of list entries
  int n = 0, maxN = 0, tiles = 0, iblocks = 0, cblocks = 0, bpp = 0, brc = 0, bp = 0;
  bool any_src = false;
  int diff_iters = 0, press_iters = 0;
  int tile = 5, tsweeps = 5;  // band shape index (kBands) and sweeps per launch (untuned default)
  int tab_tile = -1, tab_T = -1;
  hipGraphExec_t graph = nullptr;
  int64_t graph_frames = 0;
  uint64_t key = 0;  // what the table and the graph were built for
};

struct RBWork {
  float* tmp = nullptr;  // ping-pong field
  int* trig = nullptr;   // per-row break columns
  RbGroup one;           // the scene alone (eqf_step)
};

// Band shapes (waves per workgroup x rows per wave); the region is always 128
// columns wide.
struct BandCfg {
  int nw, r;
};
static const BandCfg kBands[] = {{8, 16}, {8, 32}, {16, 16}, {8, 18}, {8, 14}, {8, 10}};

static int band_tiles(int S, int tile, int T) {
  const BandCfg b = kBands[tile];
  const int TX = 128 - 6 * T, TY = b.nw * b.r - 6 * T;
  return rb::band_span(S, 128, TX) * rb::band_span(S, b.nw * b.r, TY);
}

static uint64_t group_key(const RbGroup& g) {
  uint64_t key = 1469598103934665603ull;
  auto mix = [&key](uint64_t v) { key = (key ^ v) * 1099511628211ull; };
  for (const eqf_scene* s : g.scenes) {
    // The tables embed a copy of every descriptor, so every byte of it is
    // part of the key (dt0 included: with zero viscosity and diffusion a new
    // delta_t changes only dt0).  The descriptor is value-initialised, so
    // its padding bytes are zero.
    SceneDesc d = s->desc;
    d.prio = 0;
    uint64_t words[(sizeof(SceneDesc) + 7) / 8] = {};
    std::memcpy(words, &d, sizeof(SceneDesc));
    mix(s->uid);
    for (uint64_t w : words) mix(w);
  }
  mix((uint64_t)g.tile * 16 + (uint64_t)g.tsweeps);
  return key;
}

// (Re)writes the group's tables from its scenes' descriptors, on `st`.
static int group_tables(RbGroup& g, hipStream_t st, std::string& err) {
  const int n = (int)g.scenes.size();
  std::vector<RbScene> scs(n);
  std::vector<int2> tmap, imap, cmap;
  std::vector<int> bpp(n + 1, 0), brc(n + 1, 0), bp(n + 1, 0);
  g.any_src = false;
  g.maxN = 0;
  for (int i = 0; i < n; ++i) {
    eqf_scene* s = g.scenes[i];
    auto* w = static_cast<RBWork*>(s->rb);
    scs[i] = RbScene{s->desc, w->tmp, w->trig};
    const int nt = band_tiles(s->S, g.tile, g.tsweeps);
    for (int t = 0; t < nt; ++t) tmap.push_back(make_int2(i, t));
    const int64_t ni = ((int64_t)s->N * s->N + rb::CELL_BLOCK - 1) / rb::CELL_BLOCK;
    for (int64_t k = 0; k < ni; ++k) imap.push_back(make_int2(i, (int)k));
    const int64_t nb = ((int64_t)s->S * s->S + rb::CELL_BLOCK - 1) / rb::CELL_BLOCK;
    for (int64_t k = 0; k < nb; ++k) cmap.push_back(make_int2(i, (int)k));
    bpp[i + 1] = bpp[i] + 2 * s->nbnd[M_PASSIVE];
    brc[i + 1] = brc[i] + s->nbnd[M_ROW] + s->nbnd[M_COL];
    bp[i + 1] = bp[i] + s->nbnd[M_PASSIVE];
    g.any_src = g.any_src || s->desc.has_src;
    g.maxN = std::max(g.maxN, s->N);
  }
  if (tmap.size() > INT32_MAX / 2 || cmap.size() > INT32_MAX / 2) {  // imap < cmap
    err = "red-black launch group too large for one launch";
    return EQF_EINVAL;
  }
  g.n = n;
  g.tiles = (int)tmap.size();
  g.iblocks = (int)imap.size();
  g.cblocks = (int)cmap.size();
  g.bpp = bpp[n];
  g.brc = brc[n];
  g.bp = bp[n];
  auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
  const size_t b_scs = up(sizeof(RbScene) * n), b_i = up(sizeof(int) * (n + 1));
  const size_t b_t = up(sizeof(int2) * tmap.size());
  const size_t b_m = up(sizeof(int2) * imap.size()), b_c = up(sizeof(int2) * cmap.size());
  const size_t bytes = b_scs + 3 * b_i + b_t + b_m + b_c;
  if (bytes > g.tab_bytes) {  // with headroom: a server's group changes size often
    (void)hipFree(g.d_tab);
    g.d_tab = nullptr;
    g.tab_bytes = 0;
    const size_t cap = std::max(bytes, (size_t)4096) * 2;
    hipError_t e = hipMalloc(&g.d_tab, cap);
    if (e != hipSuccess) {
      err = std::string("red-black tables: ") + hipGetErrorString(e);
      return EQF_ENOMEM;
    }
    g.tab_bytes = cap;
  }
  g.h_tab.assign(bytes, 0);
  char* base = g.h_tab.data();
  size_t off = 0;
  auto put = [&](const void* src, size_t len, size_t slot) {
    if (len) std::memcpy(base + off, src, len);
    const size_t at = off;
    off += slot;
    return at;
  };
  const size_t o_scs = put(scs.data(), sizeof(RbScene) * n, b_scs);
  const size_t o_bpp = put(bpp.data(), sizeof(int) * (n + 1), b_i);
  const size_t o_brc = put(brc.data(), sizeof(int) * (n + 1), b_i);
  const size_t o_bp = put(bp.data(), sizeof(int) * (n + 1), b_i);
  const size_t o_t = put(tmap.data(), sizeof(int2) * tmap.size(), b_t);
  const size_t o_m = put(imap.data(), sizeof(int2) * imap.size(), b_m);
  const size_t o_c = put(cmap.data(), sizeof(int2) * cmap.size(), b_c);
  char* d = (char*)g.d_tab;
  g.d_scs = (RbScene*)(d + o_scs);
  g.d_bpp = (int*)(d + o_bpp);
  g.d_brc = (int*)(d + o_brc);
  g.d_bp = (int*)(d + o_bp);
  g.d_tmap = (int2*)(d + o_t);
  g.d_imap = (int2*)(d + o_m);
  g.d_cmap = (int2*)(d + o_c);
  // behind the frames already queued on `st` (they read the old tables)
  hipError_t e = hipMemcpyAsync(g.d_tab, base, bytes, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    err = std::string("red-black tables: ") + hipGetErrorString(e);
    return EQF_EDEVICE;
  }
  g.tab_tile = g.tile;
  g.tab_T = g.tsweeps;
  return EQF_OK;
}

template <int MODE, int T, int NW, int R>
static void launch_band(hipStream_t st, const RbGroup& g, int src, int dst, int x0, int coef, int nsw) {
  hipLaunchKernelGGL((rb::k_rb_band<MODE, T, NW, R>), dim3(g.tiles), dim3(64 * NW), 0, st, g.d_scs, g.d_tmap,
                     g.tiles, src, dst, x0, coef, nsw);
}

template <int MODE, int NW, int R>
static void launch_band_t(hipStream_t st, const RbGroup& g, int src, int dst, int x0, int coef, int nsw) {
  const int T = g.tsweeps;
  if (T == 1) launch_band<MODE, 1, NW, R>(st, g, src, dst, x0, coef, nsw);
  else if (T == 2) launch_band<MODE, 2, NW, R>(st, g, src, dst, x0, coef, nsw);
  else if (T == 4) launch_band<MODE, 4, NW, R>(st, g, src, dst, x0, coef, nsw);
  else if (T == 5) launch_band<MODE, 5, NW, R>(st, g, src, dst, x0, coef, nsw);
  else if (T == 6) launch_band<MODE, 6, NW, R>(st, g, src, dst, x0, coef, nsw);
  else launch_band<MODE, 8, NW, R>(st, g, src, dst, x0, coef, nsw);
}

template <int MODE>
static void launch_gs_mode(hipStream_t st, const RbGroup& g, int src, int dst, int x0, int coef, int nsw) {
  if (g.tile == 1) launch_band_t<MODE, 8, 32>(st, g, src, dst, x0, coef, nsw);
  else if (g.tile == 2) launch_band_t<MODE, 16, 16>(st, g, src, dst, x0, coef, nsw);
  else if (g.tile == 3) launch_band_t<MODE, 8, 18>(st, g, src, dst, x0, coef, nsw);
  else if (g.tile == 4) launch_band_t<MODE, 8, 14>(st, g, src, dst, x0, coef, nsw);
  else if (g.tile == 5) launch_band_t<MODE, 8, 10>(st, g, src, dst, x0, coef, nsw);
  else launch_band_t<MODE, 8, 16>(st, g, src, dst, x0, coef, nsw);
}

static void g_copy(hipStream_t st, const RbGroup& g, int dst, int src) {
  hipLaunchKernelGGL(rb::k_copy, dim3((unsigned)g.cblocks), dim3(256), 0, st, g.d_scs, g.d_cmap, dst, src);
}

// lin_solve in red-black order on every scene of the group (one iteration
// count): ceil(iters/T) tile passes ping-ponging between x and the scratch
// field; an odd pass count ends with a copy back; then, if `corners`, the
// corners of the last set_boundaries (fluid.rs:265-271).
static void g_lin_solve(hipStream_t st, const RbGroup& g, int mode, int x, int x0, int coef, int iters,
                        bool corners) {
  if (iters <= 0) return;
  int a_ = x, b_ = rb::ROLE_TMP;
  for (int done = 0; done < iters;) {
    const int nsw = std::min(g.tsweeps, iters - done);
    if (mode == M_ROW) launch_gs_mode<M_ROW>(st, g, a_, b_, x0, coef, nsw);
    else if (mode == M_COL) launch_gs_mode<M_COL>(st, g, a_, b_, x0, coef, nsw);
    else launch_gs_mode<M_PASSIVE>(st, g, a_, b_, x0, coef, nsw);
    std::swap(a_, b_);
    done += nsw;
  }
  if (a_ != x) g_copy(st, g, x, a_);
  if (corners)
    hipLaunchKernelGGL(rb::k_corners, dim3((unsigned)((2 * g.n + 255) / 256)), dim3(256), 0, st, g.d_scs, g.n, x,
                       rb::ROLE_NONE);
}

// set_boundaries lists of one or two fields, then the corners of the fields
// c0 and c1 (ROLE_NONE: none).
static void g_bnd(hipStream_t st, const RbGroup& g, int mode0, int x0, int mode1, int x1, int c0, int c1) {
  const int* tab = x1 == rb::ROLE_NONE ? g.d_bp : mode0 == M_PASSIVE ? g.d_bpp : g.d_brc;
  const int total = x1 == rb::ROLE_NONE ? g.bp : mode0 == M_PASSIVE ? g.bpp : g.brc;
  if (total > 0)
    hipLaunchKernelGGL(rb::k_bnd, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, g.d_scs, tab, g.n, mode0,
                       x0, mode1, x1);
  if (c0 != rb::ROLE_NONE || c1 != rb::ROLE_NONE)
    hipLaunchKernelGGL(rb::k_corners, dim3((unsigned)((2 * g.n + 255) / 256)), dim3(256), 0, st, g.d_scs, g.n, c0,
                       c1);
}

// One frame (fluid.rs:437-524) of every scene of the group.  All scenes share
// the iteration counts, so the sequence of launches is the same for all.
//
// Corners (fluid.rs:265-271) are read only by advect's bilinear gathers and
// are part of the frame's final state; no sweep, divergence or gradient reads
// them.  So a field's corners are computed only where a gather or the end of
// the frame reads them before the field's next set_boundaries rewrites them:
// U0 and V0 after the first gradient (advected from), V0 after the second
// divergence, U0 after the second pressure solve (or after that divergence
// when the solve has no sweeps), U and V after the second gradient, the
// density scratch after its diffusion, the density after its advection.
static void g_frame(hipStream_t st, const RbGroup& g) {
  using rb::ROLE_NONE;
  const dim3 cells((unsigned)g.iblocks);  // interior cell blocks
  if (g.any_src)
    hipLaunchKernelGGL(rb::k_sources, dim3((unsigned)g.cblocks), dim3(256), 0, st, g.d_scs, g.d_cmap);
  g_lin_solve(st, g, M_ROW, F_U0, F_U, 0, g.diff_iters, false);  // fluid.rs:438
  g_lin_solve(st, g, M_COL, F_V0, F_V, 0, g.diff_iters, false);  // fluid.rs:448
  // project(U0, V0, p=U, div=V), fluid.rs:459
  hipLaunchKernelGGL(rb::k_div, cells, dim3(256), 0, st, g.d_scs, g.d_imap, F_U0, F_V0, F_V, F_U);
  g_bnd(st, g, M_PASSIVE, F_V, M_PASSIVE, F_U, ROLE_NONE, ROLE_NONE);
  g_lin_solve(st, g, M_PASSIVE, F_U, F_V, 2, g.press_iters, false);
  hipLaunchKernelGGL(rb::k_grad, cells, dim3(256), 0, st, g.d_scs, g.d_imap, F_U0, F_V0, F_U);
  g_bnd(st, g, M_ROW, F_U0, M_COL, F_V0, F_U0, F_V0);
  // advect U from U0 and V from V0 with (U0, V0), fluid.rs:469-489
  hipLaunchKernelGGL(rb::k_adv_trig, cells, dim3(256), 0, st, g.d_scs, g.d_imap, F_U0, F_V0);
  hipLaunchKernelGGL(rb::k_adv, cells, dim3(256), 0, st, g.d_scs, g.d_imap, F_U, F_U0, F_V, F_V0, F_U0, F_V0);
  g_bnd(st, g, M_ROW, F_U, M_COL, F_V, ROLE_NONE, ROLE_NONE);
  // project(U, V, p=U0, div=V0), fluid.rs:491
  hipLaunchKernelGGL(rb::k_div, cells, dim3(256), 0, st, g.d_scs, g.d_imap, F_U, F_V, F_V0, F_U0);
  g_bnd(st, g, M_PASSIVE, F_V0, M_PASSIVE, F_U0, F_V0, g.press_iters > 0 ? ROLE_NONE : F_U0);
  g_lin_solve(st, g, M_PASSIVE, F_U0, F_V0, 2, g.press_iters, true);
  hipLaunchKernelGGL(rb::k_grad, cells, dim3(256), 0, st, g.d_scs, g.d_imap, F_U, F_V, F_U0);
  g_bnd(st, g, M_ROW, F_U, M_COL, F_V, F_U, F_V);
  // density, fluid.rs:501-523
  g_lin_solve(st, g, M_PASSIVE, F_SC, F_D, 1, g.diff_iters, true);
  hipLaunchKernelGGL(rb::k_adv_trig, cells, dim3(256), 0, st, g.d_scs, g.d_imap, F_U, F_V);
  hipLaunchKernelGGL(rb::k_adv, cells, dim3(256), 0, st, g.d_scs, g.d_imap, F_D, F_SC, ROLE_NONE, ROLE_NONE,
                     F_U, F_V);
  g_bnd(st, g, M_PASSIVE, F_D, M_PASSIVE, ROLE_NONE, F_D, ROLE_NONE);
  g_copy(st, g, F_SC, F_D);
}

static void group_release(RbGroup& g) {
  if (g.graph) (void)hipGraphExecDestroy(g.graph);
  g.graph = nullptr;
  (void)hipFree(g.d_tab);
  g.d_tab = nullptr;
  g.tab_bytes = 0;
}

// Advances every scene of the group by nframes on `st`: the frames are
// captured once as a graph (rebuilt when the scenes' parameters, lists or
// the tiling change) and replayed.
static int group_step(RbGroup& g, hipStream_t st, int64_t nframes, std::string& err) {
  if (nframes <= 0 || g.scenes.empty()) return EQF_OK;
  g.diff_iters = g.scenes[0]->desc.diff_iters;
  g.press_iters = g.scenes[0]->desc.press_iters;
  for (const eqf_scene* s : g.scenes)
    if (s->desc.diff_iters != g.diff_iters || s->desc.press_iters != g.press_iters) {
      err = "red-black scenes of one batch need the same iteration counts";
      return EQF_EINVAL;
    }
  const uint64_t key = group_key(g);
  if (!g.d_tab || key != g.key || g.tab_tile != g.tile || g.tab_T != g.tsweeps) {
    int rc = group_tables(g, st, err);
    if (rc) return rc;
    if (g.graph) (void)hipGraphExecDestroy(g.graph);
    g.graph = nullptr;
    g.key = key;
  }
  // one frame per graph, replayed nframes times (a frame is ~100 launches)
  const int64_t per = 1;
  if (!g.graph || g.graph_frames != per) {
    if (g.graph) (void)hipGraphExecDestroy(g.graph);
    g.graph = nullptr;
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      err = std::string("capture: ") + hipGetErrorString(e);
      return EQF_EDEVICE;
    }
    for (int64_t f = 0; f < per; ++f) g_frame(st, g);
    hipError_t le = hipGetLastError();
    e = hipStreamEndCapture(st, &gr);
    if (le != hipSuccess || e != hipSuccess) {
      err = std::string("red-black frame capture: ") + hipGetErrorString(le != hipSuccess ? le : e);
      if (gr) (void)hipGraphDestroy(gr);
      return EQF_EDEVICE;
    }
    e = hipGraphInstantiate(&g.graph, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) {
      err = std::string("graph instantiate: ") + hipGetErrorString(e);
      return EQF_EDEVICE;
    }
    g.graph_frames = per;
  }
  for (int64_t f = 0; f < nframes; f += per) {
    hipError_t e = hipGraphLaunch(g.graph, st);
    if (e != hipSuccess) {
      err = std::string("graph launch: ") + hipGetErrorString(e);
      return EQF_EDEVICE;
    }
  }
  return EQF_OK;
}

void* rb_create(eqf_scene* s, std::string& err) {
  auto* w = new RBWork;
  const size_t n = (size_t)s->S * s->S;
  hipError_t e = eqf_scene_malloc(s, (void**)&w->tmp, sizeof(float) * n);
  if (e == hipSuccess) e = eqf_scene_malloc(s, (void**)&w->trig, sizeof(int) * (size_t)s->S);
  if (e != hipSuccess) {
    err = std::string("red-black workspace: ") + hipGetErrorString(e);
    eqf_scene_mfree(s, w->tmp);
    eqf_scene_mfree(s, w->trig);
    delete w;
    return nullptr;
  }
  w->one.scenes = {s};
  w->one.tile = 5;  // the autotuner (equilibrium_amd/tune.py) may override these
  w->one.tsweeps = 5;
  return w;
}

void rb_destroy(eqf_scene* s) {
  auto* w = static_cast<RBWork*>(s->rb);
  if (!w) return;
  group_release(w->one);
  eqf_scene_mfree(s, w->tmp);
  eqf_scene_mfree(s, w->trig);
  delete w;
  s->rb = nullptr;
}

static bool tiling_ok(int tile, int tsweeps) {
  return tile >= 0 && tile <= 5 && tsweeps >= 1 && tsweeps <= 8 && tsweeps != 3 && tsweeps != 7;
}

int rb_set_tiling(eqf_scene* s, int tile, int tsweeps) {
  auto* w = static_cast<RBWork*>(s->rb);
  if (!w || !tiling_ok(tile, tsweeps)) return EQF_EINVAL;
  w->one.tile = tile;
  w->one.tsweeps = tsweeps;
  return EQF_OK;
}

int rb_step(eqf_scene* s, int64_t nframes, std::string& err) {
  auto* w = static_cast<RBWork*>(s->rb);
  return group_step(w->one, s->stream, nframes, err);
}

// Batches of red-black scenes (eqf_runtime.hip): one group over all of them.
void* rb_batch_create(const std::vector<eqf_scene*>& scenes, hipStream_t st, std::string& err) {
  for (const eqf_scene* s : scenes)
    if (s->desc.diff_iters != scenes[0]->desc.diff_iters || s->desc.press_iters != scenes[0]->desc.press_iters) {
      err = "red-black scenes of one batch need the same iteration counts";
      return nullptr;
    }
  auto* g = new RbGroup;
  g->scenes = scenes;
  g->stream = st;
  return g;
}
void rb_batch_destroy(void* p) {
  auto* g = static_cast<RbGroup*>(p);
  if (!g) return;
  group_release(*g);
  delete g;
}
int rb_batch_set_tiling(void* p, int tile, int tsweeps) {
  auto* g = static_cast<RbGroup*>(p);
  if (!g || !tiling_ok(tile, tsweeps)) return EQF_EINVAL;
  g->tile = tile;
  g->tsweeps = tsweeps;
  return EQF_OK;
}
int rb_batch_step(void* p, int64_t nframes, std::string& err) {
  auto* g = static_cast<RbGroup*>(p);
  return group_step(*g, g->stream, nframes, err);
}
void rb_batch_set_scenes(void* p, const std::vector<eqf_scene*>& scenes) {
  auto* g = static_cast<RbGroup*>(p);
  g->scenes = scenes;  // a new membership changes the key: tables and graph are rebuilt
}

}  // namespace eqf
