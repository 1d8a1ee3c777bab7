// eqf_setup.hip -- device-side scene setup: initial state, layout conversion
// of uploaded/downloaded fields and sources, and the boundary codes/lists
// derived from the obstacle mask.  Keeping these on the device makes
// admitting a scene cost a few small launches instead of host loops over
// every cell, which matters when the server admits hundreds of scenes/s.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "eqf_internal.h"

using namespace eqf;

namespace {

struct Geo {
  int S, N, ns, P, layout;
};

__device__ __forceinline__ int64_t off_of(const Geo& g, int x, int y) {
  return g.layout == L_CYCLIC ? cyc_off(x, y, g.S, g.P, g.ns) : (int64_t)x + (int64_t)y * g.S;
}

// Fluid::new / Fluid::default state (fluid.rs:93-110, 527-570) for sizes
// where the density block needs no index clamping (S >= 22): U = V = `times`
// additions of 1.0, D = S = `times` additions of 0.9 on [S/2-10, S/2+10]^2.
__global__ void k_init(Geo g, float* __restrict__ f, int64_t fs, int times) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)g.S * g.S) return;
  const int x = (int)(e % g.S), y = (int)(e / g.S);
  const int64_t o = off_of(g, x, y);
  float vel = 0.0f, den = 0.0f;
  const int lo = g.S / 2 - 10, hi = g.S / 2 + 10;
  const bool blk = x >= lo && x <= hi && y >= lo && y <= hi;
  for (int t = 0; t < times; ++t) {
    vel += 1.0f;
    den += blk ? 0.9f : 0.0f;
  }
  f[F_D * fs + o] = den;
  f[F_SC * fs + o] = den;
  f[F_U * fs + o] = vel;
  f[F_V * fs + o] = vel;
  f[F_U0 * fs + o] = 0.0f;
  f[F_V0 * fs + o] = 0.0f;
}

// Row-major <-> scene layout.  mode 0: dst = src; mode 1: dst += src (and
// dst2 += src when dst2 is set: add_density writes density and scratch).
__global__ void k_from_rm(Geo g, const float* __restrict__ rm, float* __restrict__ dst, float* __restrict__ dst2,
                          int mode) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)g.S * g.S) return;
  const int64_t o = off_of(g, (int)(e % g.S), (int)(e / g.S));
  const float v = rm[e];
  if (mode == 0) {
    dst[o] = v;
  } else {
    dst[o] += v;
    if (dst2) dst2[o] += v;
  }
}

__global__ void k_to_rm(Geo g, const float* __restrict__ src, float* __restrict__ rm) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)g.S * g.S) return;
  rm[e] = src[off_of(g, (int)(e % g.S), (int)(e / g.S))];
}

// rowf[j] / colf[i]: row / column holds a NoWall cell (Passive copy rule of
// fluid.rs:179-187).  A 2D grid of 256-column x 64-row tiles: each thread ORs
// its column over the tile's rows (coalesced row reads), each wave ORs every
// row over its 64 columns with a wave vote, and tiles combine through atomic
// ORs into zeroed 32-bit flags; k_rowcol_bytes turns them into the byte arrays.
__global__ void __launch_bounds__(256) k_rowcol_or(Geo g, const uint8_t* __restrict__ m, uint32_t* __restrict__ rowf32,
                                                   uint32_t* __restrict__ colf32) {
  const int S = g.S;
  const int x = blockIdx.x * 256 + threadIdx.x;
  const int y0 = blockIdx.y * 64;
  const int y1 = min(y0 + 64, S);
  uint32_t c = 0;
  for (int y = y0; y < y1; ++y) {  // uniform bounds: every lane takes part in the vote
    const bool f = x < S && m[(int64_t)y * S + x] != EQF_DEFAULT_WALL;
    c |= f ? 1u : 0u;
    if (__any(f) && (threadIdx.x & 63) == 0) atomicOr(&rowf32[y], 1u);
  }
  if (c) atomicOr(&colf32[x], 1u);
}

__global__ void k_rowcol_bytes(int S, const uint32_t* __restrict__ rowf32, const uint32_t* __restrict__ colf32,
                               uint8_t* __restrict__ rowf, uint8_t* __restrict__ colf) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= S) return;
  rowf[k] = rowf32[k] != 0;
  colf[k] = colf32[k] != 0;
}

// Per-cell boundary codes (eqf_layout.h CodeBits) and the set_boundaries
// lists: AdjustRow (right wall wins over left, fluid.rs:151-164),
// AdjustColumn (up wins over down, fluid.rs:165-177), Passive ring copies.
// fill == 0 only counts entries.
__global__ void k_codes(Geo g, const uint8_t* __restrict__ m, const uint8_t* __restrict__ rowf,
                        const uint8_t* __restrict__ colf, uint8_t* __restrict__ code, int* __restrict__ cnt,
                        int2* __restrict__ lrow, int2* __restrict__ lcol, int2* __restrict__ lpas, int fill) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int S = g.S, N = g.N;
  if (e >= (int64_t)S * S) return;
  const int x = (int)(e % S), y = (int)(e / S);
  auto wall = [&](int i, int j) { return m[(int64_t)j * S + i] == EQF_DEFAULT_WALL; };
  if (x >= 1 && x <= N && y >= 1 && y <= N) {
    uint32_t c = 0;
    const bool fl = !wall(x, y);
    if (fl) c |= C_FLUID;
    if (wall(x - 1, y)) c |= C_LWALL;
    if (wall(x + 1, y)) c |= C_RWALL;
    if (wall(x, y + 1)) c |= C_DWALL;
    if (wall(x, y - 1)) c |= C_UWALL;
    if (colf[x]) c |= C_COLF;
    if (fill && g.layout == L_ROWMAJOR) {
      uint32_t r = R_INTERIOR;
      if (fl) {
        r |= ((c & C_RWALL) ? R_NEG_SECOND : (c & C_LWALL) ? R_NEG_FIRST : 0u) << R_ROW_SHIFT;
        r |= ((c & C_UWALL) ? R_NEG_FIRST : (c & C_DWALL) ? R_NEG_SECOND : 0u) << R_COL_SHIFT;
      }
      code[(int64_t)y * rm_code_stride(S) + x] = (uint8_t)r;
    }
    if (fill && g.layout == L_CYCLIC) {
      const int64_t o = cyc_code_word(x, y, g.P), mstride = cyc_code_bytes(g.P, g.ns) / 4;
      const int sh = cyc_code_shift(x, y, g.P);
      const bool strip_bottom = (y & 63) == 0 && y < N;  // the row below is in the next strip
      uint8_t srow = SEL_KEEP, scol = SEL_KEEP;
      if (fl) {
        srow = (c & C_RWALL) ? SEL_NEG2 : (c & C_LWALL) ? SEL_NEG1 : SEL_KEEP;
        scol = (c & C_UWALL) ? SEL_NEG1 : ((c & C_DWALL) && !strip_bottom) ? SEL_NEG2 : SEL_KEEP;
      } else if (((y - 1) & 63) == 0 && y > 1 && !wall(x, y - 1) && !wall(x, y - 2)) {
        scol = SEL_FIXUP;
      }
      // the words start at zero; OR in the non-zero selectors
      uint32_t* cw = reinterpret_cast<uint32_t*>(code);
      if (srow) atomicOr(&cw[o], (uint32_t)srow << sh);
      if (scol) atomicOr(&cw[mstride + o], (uint32_t)scol << sh);
      if (colf[x]) atomicOr(&cw[2 * mstride + o], 1u << sh);
    }
    if (fl) {
      int sx = -1;
      if (wall(x + 1, y)) sx = x + 1;
      else if (wall(x - 1, y)) sx = x - 1;
      if (sx >= 0) {
        const int k = atomicAdd(&cnt[0], 1);
        if (fill) lrow[k] = make_int2((int)off_of(g, x, y), (int)off_of(g, sx, y));
      }
      int sy = -1;
      if (wall(x, y - 1)) sy = y - 1;
      else if (wall(x, y + 1)) sy = y + 1;
      if (sy >= 0) {
        const int k = atomicAdd(&cnt[1], 1);
        if (fill) lcol[k] = make_int2((int)off_of(g, x, y), (int)off_of(g, x, sy));
      }
    }
  } else if (!((x == 0 || x == S - 1) && (y == 0 || y == S - 1))) {  // non-corner ring cell
    int sx = x, sy = y;
    bool on = false;
    if (y == 0) { sy = 1; on = colf[x]; }
    else if (y == S - 1) { sy = S - 2; on = colf[x]; }
    else if (x == 0) { sx = 1; on = rowf[y]; }
    else { sx = S - 2; on = rowf[y]; }
    if (on) {
      const int k = atomicAdd(&cnt[2], 1);
      if (fill) lpas[k] = make_int2((int)off_of(g, x, y), (int)off_of(g, sx, sy));
      if (fill && g.layout == L_ROWMAJOR) {
        const uint32_t r = y == 0 ? R_COPY_DOWN : y == S - 1 ? R_COPY_UP : x == 0 ? R_COPY_RIGHT : R_COPY_LEFT;
        code[(int64_t)y * rm_code_stride(S) + x] = (uint8_t)(r << R_PAS_SHIFT);
      }
    }
  }
}

Geo geo_of(const eqf_scene* s) { return Geo{s->S, s->N, s->ns, s->P, s->layout}; }

unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

int eqf_dev_init_fields(eqf_scene* s, int times) {
  const int64_t n = (int64_t)s->S * s->S;
  hipLaunchKernelGGL(k_init, dim3(blocks_for(n)), dim3(256), 0, s->stream, geo_of(s), s->d_fld, s->fstride, times);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? EQF_OK : eqf_internal_error(EQF_EDEVICE, std::string("init: ") + hipGetErrorString(e));
}

static int ensure_stage(eqf_scene* s) {
  if (s->d_stage) return EQF_OK;
  hipError_t e = eqf_scene_malloc(s, (void**)&s->d_stage, sizeof(float) * (size_t)s->S * s->S + 256);
  return e == hipSuccess ? EQF_OK
                         : eqf_internal_error(EQF_ENOMEM, std::string("staging buffer: ") + hipGetErrorString(e));
}

int eqf_dev_upload(eqf_scene* s, float* dst, float* dst2, const float* host, int add) {
  const size_t n = (size_t)s->S * s->S;
  if (s->layout == L_ROWMAJOR && !add) {
    hipError_t e = hipMemcpyAsync(dst, host, sizeof(float) * n, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return e == hipSuccess ? EQF_OK : eqf_internal_error(EQF_EDEVICE, std::string("upload: ") + hipGetErrorString(e));
  }
  int rc = ensure_stage(s);
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(s->d_stage, host, sizeof(float) * n, hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_from_rm, dim3(blocks_for((int64_t)n)), dim3(256), 0, s->stream, geo_of(s), s->d_stage, dst,
                       dst2, add);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  return e == hipSuccess ? EQF_OK : eqf_internal_error(EQF_EDEVICE, std::string("upload: ") + hipGetErrorString(e));
}

int eqf_dev_to_rm_async(eqf_scene* s, const float* src, float* dst) {
  const size_t n = (size_t)s->S * s->S;
  hipError_t e;
  if (s->layout == L_ROWMAJOR) {
    e = hipMemcpyAsync(dst, src, sizeof(float) * n, hipMemcpyDeviceToDevice, s->stream);
  } else {
    hipLaunchKernelGGL(k_to_rm, dim3(blocks_for((int64_t)n)), dim3(256), 0, s->stream, geo_of(s), src, dst);
    e = hipGetLastError();
  }
  return e == hipSuccess ? EQF_OK : eqf_internal_error(EQF_EDEVICE, std::string("layout copy: ") + hipGetErrorString(e));
}

int eqf_dev_download(eqf_scene* s, const float* src, float* host) {
  const size_t n = (size_t)s->S * s->S;
  hipError_t e;
  if (s->layout == L_ROWMAJOR) {
    e = hipMemcpyAsync(host, src, sizeof(float) * n, hipMemcpyDeviceToHost, s->stream);
  } else {
    int rc = ensure_stage(s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_to_rm, dim3(blocks_for((int64_t)n)), dim3(256), 0, s->stream, geo_of(s), src, s->d_stage);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host, s->d_stage, sizeof(float) * n, hipMemcpyDeviceToHost, s->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  return e == hipSuccess ? EQF_OK : eqf_internal_error(EQF_EDEVICE, std::string("download: ") + hipGetErrorString(e));
}

void eqf_dev_trim(eqf_scene* s) {
  eqf_scene_mfree(s, s->d_stage);
  s->d_stage = nullptr;
  eqf_scene_mfree(s, s->d_mask);
  s->d_mask = nullptr;
}

// Layout of the scene's d_mask buffer: the mask (S*S bytes), colf (S bytes),
// then the 32-bit row/column flags (2*S) and the entry counters of the two
// k_codes passes (3 + 3), zeroed by one memset.
struct MaskBuf {
  size_t colf, flags, cnt, bytes;
};
static MaskBuf mask_buf(int S) {
  const size_t n = (size_t)S * S;
  auto up = [](size_t v) { return (v + 15) & ~(size_t)15; };
  MaskBuf b;
  b.colf = up(n);
  b.flags = up(b.colf + (size_t)S);
  b.cnt = b.flags + 8 * (size_t)S;
  b.bytes = b.cnt + 6 * sizeof(int);
  return b;
}

int eqf_dev_rebuild_bnd(eqf_scene* s) {
  const int S = s->S;
  const int64_t n = (int64_t)S * S;
  const MaskBuf mb = mask_buf(S);
  hipError_t e = hipSuccess;
  if (!s->d_mask) e = eqf_scene_malloc(s, (void**)&s->d_mask, mb.bytes);
  if (e != hipSuccess) return eqf_internal_error(EQF_ENOMEM, std::string("mask buffer: ") + hipGetErrorString(e));
  uint8_t* d_colf = s->d_mask + mb.colf;
  uint32_t* d_flags = (uint32_t*)(s->d_mask + mb.flags);
  int* d_cnt = (int*)(s->d_mask + mb.cnt);
  const Geo g = geo_of(s);
  e = hipMemcpyAsync(s->d_mask, s->mask.data(), (size_t)n, hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess) e = hipMemsetAsync(d_flags, 0, mb.bytes - mb.flags, s->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_rowcol_or, dim3((unsigned)((S + 255) / 256), (unsigned)((S + 63) / 64)), dim3(256), 0,
                       s->stream, g, s->d_mask, d_flags, d_flags + S);
    hipLaunchKernelGGL(k_rowcol_bytes, dim3(blocks_for(S)), dim3(256), 0, s->stream, S, d_flags, d_flags + S,
                       s->d_rowf, d_colf);
    hipLaunchKernelGGL(k_codes, dim3(blocks_for(n)), dim3(256), 0, s->stream, g, s->d_mask, s->d_rowf, d_colf,
                       s->d_code, d_cnt, (int2*)nullptr, (int2*)nullptr, (int2*)nullptr, 0);
    e = hipGetLastError();
  }
  int cnt[3] = {0, 0, 0};
  if (e == hipSuccess) e = hipMemcpyAsync(cnt, d_cnt, sizeof(cnt), hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return eqf_internal_error(EQF_EDEVICE, std::string("boundary codes: ") + hipGetErrorString(e));
  const size_t total = std::max<size_t>(1, (size_t)cnt[0] + cnt[1] + cnt[2]);
  if (total > s->bnd_cap) {
    eqf_scene_mfree(s, s->d_bnd);
    s->d_bnd = nullptr;
    e = eqf_scene_malloc(s, (void**)&s->d_bnd, sizeof(int2) * total);
    if (e != hipSuccess) return eqf_internal_error(EQF_ENOMEM, std::string("boundary lists: ") + hipGetErrorString(e));
    s->bnd_cap = total;
  }
  {
    const size_t ncode = s->layout == L_CYCLIC ? 3 * (size_t)cyc_code_bytes(s->P, s->ns) : (size_t)rm_code_bytes(S);
    e = hipMemsetAsync(s->d_code, 0, ncode, s->stream);
  }
  if (e == hipSuccess) {  // the fill pass counts into the second, still zero, counters
    hipLaunchKernelGGL(k_codes, dim3(blocks_for(n)), dim3(256), 0, s->stream, g, s->d_mask, s->d_rowf, d_colf,
                       s->d_code, d_cnt + 3, s->d_bnd, s->d_bnd + cnt[0], s->d_bnd + cnt[0] + cnt[1], 1);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return eqf_internal_error(EQF_EDEVICE, std::string("boundary lists: ") + hipGetErrorString(e));
  s->nbnd[0] = cnt[0];
  s->nbnd[1] = cnt[1];
  s->nbnd[2] = cnt[2];
  s->bnd_dirty = false;
  return EQF_OK;
}
