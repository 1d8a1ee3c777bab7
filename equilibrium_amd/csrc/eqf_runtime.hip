// eqf_runtime.hip -- host side of libeqfluid.so: the C ABI of include/eqfluid.h.
//
// Scene state lives on the device.  The host keeps only the parameters and the
// obstacle mask (static during a run, fluid.rs:610-619), from which it derives
// the per-cell boundary codes and the set_boundaries lists the kernels use.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "eqf_internal.h"

using namespace eqf;

template <int D, int B, int BS, bool XR, int MAXT>
__global__ void eqf_frames_cyclic(const eqf::SceneDesc* __restrict__ descs, int nframes);


// ----------------------------------------------------------------- errors --
static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
#define HIPCHK(expr)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(EQF_EDEVICE, std::string(#expr " failed: ") + hipGetErrorString(e_)); \
  } while (0)

extern "C" const char* eqf_last_error(void) { return g_err.c_str(); }

extern "C" void eqf_params_default(eqf_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->size = 128;       // configs.rs:19
  p->delta_t = 0.02f;  // configs.rs:17
  p->frames = 16;      // configs.rs:18
  p->diffusion = 0.0f; // configs.rs:53
  p->viscosity = 0.001f;  // configs.rs:54
  p->gs_iters = -1;
  p->pressure_iters = -1;
  p->gs_mode = EQF_GS_WAVEFRONT;
  p->init_times = 1;
}

static int64_t resolved_gs(const eqf_params& p) { return p.gs_iters < 0 ? p.frames : p.gs_iters; }
static int64_t resolved_press(const eqf_params& p) {
  return p.pressure_iters < 0 ? resolved_gs(p) : p.pressure_iters;
}

int64_t eqf_scene_off(const eqf_scene* s, int x, int y) {
  return s->layout == L_CYCLIC ? cyc_off(x, y, s->S, s->P, s->ns) : (int64_t)x + (int64_t)y * s->S;
}

static float* field_ptr(const eqf_scene* s, int f) { return s->d_fld + (int64_t)f * s->fstride; }

// row-major host array <-> device field (layout conversion on the device)
static int upload_rm(eqf_scene* s, float* dfield, const float* host) {
  return eqf_dev_upload(s, dfield, nullptr, host, 0);
}

static int download_rm(const eqf_scene* cs, const float* dfield, float* host) {
  return eqf_dev_download(const_cast<eqf_scene*>(cs), dfield, host);
}

// Reference initial state (fluid.rs:93-110, 527-570, 602-606), on the host.
static void host_init(int S, int times, std::vector<std::vector<float>>& F, std::vector<uint8_t>& m) {
  const size_t n = (size_t)S * S;
  F.assign(NFIELDS, std::vector<float>(n, 0.0f));
  m.assign(n, EQF_NO_WALL);
  auto idx = [S](int64_t x, int64_t y) {
    x = std::min<int64_t>(std::max<int64_t>(x, 0), S - 1);
    y = std::min<int64_t>(std::max<int64_t>(y, 0), S - 1);
    return (size_t)(x + y * S);
  };
  const uint32_t su = (uint32_t)S;
  for (int t = 0; t < times; ++t) {
    for (uint32_t j = 0; j < su; ++j)  // init_velocities
      for (uint32_t i = 0; i < su; ++i) {
        F[F_U][idx(i, j)] += 1.0f;
        F[F_V][idx(i, j)] += 1.0f;
      }
    for (uint32_t j = 0; j < su; ++j)  // init_density, first loop adds 0.0
      for (uint32_t i = 0; i < su; ++i) {
        F[F_D][idx(i, j)] += 0.0f;
        F[F_SC][idx(i, j)] += 0.0f;
      }
    // `size/2 - 10 ..= size/2 + 10` in u32: wraps (empty range) below size 20
    const uint32_t lo = su / 2u - 10u, hi = su / 2u + 10u;
    for (uint64_t j = lo; j <= hi; ++j)
      for (uint64_t i = lo; i <= hi; ++i) {
        F[F_D][idx((int64_t)i, (int64_t)j)] += 0.9f;
        F[F_SC][idx((int64_t)i, (int64_t)j)] += 0.9f;
      }
    for (int i = 0; i < S; ++i) {  // init_walls
      m[idx(i, 0)] = EQF_DEFAULT_WALL;
      m[idx(i, S - 1)] = EQF_DEFAULT_WALL;
    }
    for (int j = 0; j < S; ++j) {
      m[idx(0, j)] = EQF_DEFAULT_WALL;
      m[idx(S - 1, j)] = EQF_DEFAULT_WALL;
    }
  }
}

// Boundary codes and set_boundaries lists derived from the mask (built on
// the device from the host copy of the mask).
static int rebuild_bnd(eqf_scene* s) { return eqf_dev_rebuild_bnd(s); }

static void fill_desc(eqf_scene* s) {
  SceneDesc& d = s->desc;
  const eqf_params& p = s->p;
  d.fld = s->d_fld;
  d.src = s->d_src;
  d.code = s->d_code;
  d.rowf = s->d_rowf;
  d.bnd = s->d_bnd;
  d.fstride = s->fstride;
  d.S = s->S;
  d.N = s->N;
  d.ns = s->ns;
  d.P = s->P;
  for (int i = 0; i < 3; ++i) d.nbnd[i] = s->nbnd[i];
  d.diff_iters = (int)resolved_gs(p);
  d.press_iters = (int)resolved_press(p);
  // diffuse (fluid.rs:286-294) and lin_solve's c_recip (fluid.rs:311), in f32
  const float sf = (float)(s->S - 2);
  float a = p.delta_t * p.viscosity * sf * sf;
  d.a_visc = a;
  d.cr_visc = 1.0f / (1.0f + 4.0f * a);
  a = p.delta_t * p.diffusion * sf * sf;
  d.a_diff = a;
  d.cr_diff = 1.0f / (1.0f + 4.0f * a);
  d.dt0 = p.delta_t * (float)(s->S - 2);  // fluid.rs:390
  d.sf = (float)s->S;
  d.sfm1 = d.sf - 1.0f;
  d.has_src = s->d_src ? s->src_mask : 0;
  d.prio = 0;
  d.stream_l2 = eqf_stream_l2(s->P, 1);
}

static int sync_desc(eqf_scene* s) {
  if (s->bnd_dirty) {
    int rc = rebuild_bnd(s);
    if (rc) return rc;
  }
  fill_desc(s);
  HIPCHK(hipMemcpyAsync(s->d_desc, &s->desc, sizeof(SceneDesc), hipMemcpyHostToDevice, s->stream));
  return EQF_OK;
}

static int check_params(const eqf_params* p) {
  if (!p) return fail(EQF_EINVAL, "params is NULL");
  if (p->size < 3) return fail(EQF_EINVAL, "size must be >= 3 (one interior cell inside the wall ring)");
  if (p->size > 32768) return fail(EQF_EINVAL, "size must be <= 32768");
  if (p->gs_mode != EQF_GS_WAVEFRONT && p->gs_mode != EQF_GS_REDBLACK)
    return fail(EQF_EINVAL, "gs_mode must be EQF_GS_WAVEFRONT or EQF_GS_REDBLACK");
  if (p->init_times < 1 || p->init_times > 2) return fail(EQF_EINVAL, "init_times must be 1 (Fluid::new) or 2 (Fluid::default)");
  const int64_t it = resolved_gs(*p), pi = resolved_press(*p);
  if (it > (1 << 24) || pi > (1 << 24)) return fail(EQF_EINVAL, "iteration count too large");
  if (p->gs_mode == EQF_GS_WAVEFRONT) {
    const int N = (int)p->size - 2, ns = (N + 63) / 64;
    if (ns > 16)
      return fail(EQF_EINVAL, "wavefront mode supports size <= 1026 (16 strips of 64 rows per workgroup); "
                              "use EQF_GS_REDBLACK for larger grids");
    if ((int64_t)(ns - 1) * 87 + 64 + std::max(it, pi) * N + 64 > (int64_t)INT32_MAX)
      return fail(EQF_EINVAL, "iteration count too large for this size");
  }
  return EQF_OK;
}

extern "C" int eqf_ctx_create(int device, eqf_ctx** out) {
  if (!out) return fail(EQF_EINVAL, "out is NULL");
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(EQF_EINVAL, "no such device");
  HIPCHK(hipSetDevice(device));
  if (int rc = eqf_internal_scratch_init(device)) return rc;  // with the device current
  auto* c = new (std::nothrow) eqf_ctx;
  if (!c) return fail(EQF_ENOMEM, "host allocation failed");
  c->device = device;
  *out = c;
  return EQF_OK;
}

extern "C" int eqf_ctx_destroy(eqf_ctx* ctx) {
  delete ctx;
  return EQF_OK;
}

extern "C" int eqf_ctx_device_key(const eqf_ctx* ctx, char* buf, size_t len) {
  if (!ctx || !buf || len == 0) return fail(EQF_EINVAL, "ctx or buf is NULL, or len is 0");
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, ctx->device));
  // the properties' name can be empty (seen on MI355X): then the driver's
  // name, else a placeholder; the architecture and CU count follow anyway
  char name[256] = {0};
  std::snprintf(name, sizeof(name), "%s", prop.name);
  if (!name[0] && hipDeviceGetName(name, (int)sizeof(name), ctx->device) != hipSuccess) name[0] = 0;
  if (!name[0]) std::snprintf(name, sizeof(name), "gpu");
  const int n = snprintf(buf, len, "%s/%s/%dCU", name, prop.gcnArchName, prop.multiProcessorCount);
  if (n < 0 || (size_t)n >= len) {
    buf[0] = 0;
    return fail(EQF_EINVAL, "buffer too small for the device key");
  }
  return EQF_OK;
}

hipError_t eqf_scene_malloc(eqf_scene* s, void** p, size_t bytes) {
  if (s->pool) return hipMallocFromPoolAsync(p, bytes, s->pool, s->pool_stream);
  return hipMalloc(p, bytes);
}

void eqf_scene_mfree(eqf_scene* s, void* p) {
  if (!p) return;
  if (s->pool) (void)hipFreeAsync(p, s->pool_stream);
  else (void)hipFree(p);
}

static void scene_free(eqf_scene* s) {
  if (!s) return;
  if (s->rb) rb_destroy(s);
  if (s->ev_ops) (void)hipEventDestroy(s->ev_ops);
  eqf_scene_mfree(s, s->d_main);
  eqf_scene_mfree(s, s->d_src);
  eqf_scene_mfree(s, s->d_bnd);
  eqf_scene_mfree(s, s->d_stage);
  eqf_scene_mfree(s, s->d_mask);
  if (s->stream && s->owns_stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

static int scene_alloc(eqf_ctx* ctx, const eqf_params* p, eqf_scene** out, hipStream_t stream = nullptr,
                       hipMemPool_t pool = nullptr) {
  int rc = check_params(p);
  if (rc) return rc;
  HIPCHK(hipSetDevice(ctx->device));
  auto* s = new (std::nothrow) eqf_scene;
  if (!s) return fail(EQF_ENOMEM, "host allocation failed");
  static std::atomic<uint64_t> next_uid{1};
  s->uid = next_uid++;
  s->ctx = ctx;
  s->p = *p;
  s->S = (int)p->size;
  s->N = s->S - 2;
  s->ns = (s->N + 63) / 64;
  s->layout = p->gs_mode == EQF_GS_WAVEFRONT ? L_CYCLIC : L_ROWMAJOR;
  s->P = cyc_period(s->N);
  s->fstride = s->layout == L_CYCLIC ? cyc_field_floats(s->S, s->P, s->ns) : rm_field_floats(s->S);
  hipError_t e = hipSuccess;
  if (stream) {
    s->stream = stream;
    s->owns_stream = false;
    if (pool) {
      s->pool = pool;
      s->pool_stream = stream;
    }
  } else {
    e = hipStreamCreate(&s->stream);
  }
  if (e != hipSuccess) {
    scene_free(s);
    return fail(EQF_EDEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  const size_t ncode = s->layout == L_CYCLIC ? 3 * (size_t)cyc_code_bytes(s->P, s->ns) : (size_t)rm_code_bytes(s->S);
  auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
  const size_t b_fld = up(sizeof(float) * NFIELDS * (size_t)s->fstride);
  const size_t b_code = up(ncode), b_rowf = up((size_t)s->S);
  const size_t b_desc = up(sizeof(SceneDesc));
  e = eqf_scene_malloc(s, &s->d_main, b_fld + b_code + b_rowf + b_desc);
  if (e == hipSuccess) {
    char* base = (char*)s->d_main;
    s->d_fld = (float*)base;
    s->d_code = (uint8_t*)(base + b_fld);
    s->d_rowf = (uint8_t*)(base + b_fld + b_code);
    s->d_desc = (SceneDesc*)(base + b_fld + b_code + b_rowf);
    e = hipMemsetAsync(s->d_fld, 0, sizeof(float) * NFIELDS * s->fstride, s->stream);
  }
  if (e != hipSuccess) {
    scene_free(s);
    return fail(e == hipErrorOutOfMemory ? EQF_ENOMEM : EQF_EDEVICE,
                std::string("device allocation failed: ") + hipGetErrorString(e));
  }
  *out = s;
  return EQF_OK;
}

extern "C" int eqf_scene_create(eqf_ctx* ctx, const eqf_params* p, eqf_scene** out) {
  return eqf_internal_scene_create(ctx, p, nullptr, nullptr, out);
}

int eqf_internal_scene_create(eqf_ctx* ctx, const eqf_params* p, hipStream_t stream, hipMemPool_t pool,
                              eqf_scene** out) {
  if (!ctx || !out) return fail(EQF_EINVAL, "ctx/out is NULL");
  eqf_scene* s = nullptr;
  int rc = scene_alloc(ctx, p, &s, stream, pool);
  if (rc) return rc;
  if (s->S >= 22) {
    // No index clamping in init_density at this size: write the state on the
    // device; the host keeps only the mask (ring walls, init_walls).
    const int S = s->S;
    s->mask.assign((size_t)S * S, EQF_NO_WALL);
    for (int i = 0; i < S; ++i) {
      s->mask[i] = s->mask[(size_t)i + (size_t)(S - 1) * S] = EQF_DEFAULT_WALL;
      s->mask[(size_t)i * S] = s->mask[(size_t)(S - 1) + (size_t)i * S] = EQF_DEFAULT_WALL;
    }
    rc = eqf_dev_init_fields(s, p->init_times);
    if (!rc && hipStreamSynchronize(s->stream) != hipSuccess) rc = fail(EQF_EDEVICE, "scene init failed");
  } else {
    std::vector<std::vector<float>> F;
    host_init(s->S, p->init_times, F, s->mask);
    for (int f = 0; f < NFIELDS && !rc; ++f) rc = upload_rm(s, field_ptr(s, f), F[f].data());
  }
  if (!rc && s->layout == L_ROWMAJOR) {
    std::string err;
    s->rb = rb_create(s, err);
    if (!s->rb) rc = fail(EQF_ENOMEM, err);
  }
  if (rc) {
    scene_free(s);
    return rc;
  }
  s->bnd_dirty = true;
  *out = s;
  return EQF_OK;
}

extern "C" int eqf_scene_clone(const eqf_scene* src, eqf_scene** out) {
  if (!src || !out) return fail(EQF_EINVAL, "src/out is NULL");
  int rc = eqf_internal_scene_enter(const_cast<eqf_scene*>(src));
  if (rc) return rc;
  eqf_scene* s = nullptr;
  rc = scene_alloc(src->ctx, &src->p, &s);
  if (rc) return rc;
  s->mask = src->mask;
  s->bnd_dirty = true;
  hipError_t e = hipStreamSynchronize(src->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(s->d_fld, src->d_fld, sizeof(float) * NFIELDS * s->fstride, hipMemcpyDeviceToDevice,
                       s->stream);
  s->src_mask = src->src_mask;
  if (e == hipSuccess && src->d_src) {
    e = eqf_scene_malloc(s, (void**)&s->d_src, sizeof(float) * 3 * s->fstride);
    if (e == hipSuccess)
      e = hipMemcpyAsync(s->d_src, src->d_src, sizeof(float) * 3 * s->fstride, hipMemcpyDeviceToDevice, s->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) {
    scene_free(s);
    return fail(EQF_EDEVICE, std::string("clone: ") + hipGetErrorString(e));
  }
  if (s->layout == L_ROWMAJOR) {
    std::string err;
    s->rb = rb_create(s, err);
    if (!s->rb) {
      scene_free(s);
      return fail(EQF_ENOMEM, err);
    }
  }
  *out = s;
  return EQF_OK;
}

extern "C" int eqf_scene_destroy(eqf_scene* s) {
  if (!s) return EQF_OK;
  if (s->in_batch) return fail(EQF_EBUSY, "scene is part of a batch");
  (void)hipStreamSynchronize(s->stream);
  scene_free(s);
  return EQF_OK;
}

extern "C" int eqf_scene_get_params(const eqf_scene* s, eqf_params* out) {
  if (!s || !out) return fail(EQF_EINVAL, "NULL argument");
  *out = s->p;
  return EQF_OK;
}

extern "C" int eqf_scene_set_params(eqf_scene* s, const eqf_params* p) {
  if (!s || !p) return fail(EQF_EINVAL, "NULL argument");
  if (s->in_batch) return fail(EQF_EBUSY, "scene is part of a batch");  // the batch holds its descriptor
  if (p->size != s->p.size || p->gs_mode != s->p.gs_mode)
    return fail(EQF_EINVAL, "size and gs_mode are fixed for the lifetime of a scene");
  int rc = check_params(p);
  if (rc) return rc;
  s->p = *p;
  return EQF_OK;
}

extern "C" int eqf_rect_validate(int64_t x0, int64_t y0, int64_t x1, int64_t y1, uint32_t size) {
  const int64_t sz = size;
  const bool ok = x0 != x1 && y0 != y1 && x0 < x1 && y0 < y1 && x0 < sz && y0 < sz && x1 < sz && y1 < sz;
  return ok ? EQF_OK : fail(EQF_EINVAL, "Invalid input for Rectangle");
}

extern "C" int eqf_fill_rect(eqf_scene* s, int64_t x0, int64_t y0, int64_t x1, int64_t y1) {
  if (!s) return fail(EQF_EINVAL, "scene is NULL");
  if (s->in_batch) return fail(EQF_EBUSY, "scene is part of a batch");
  const int64_t S = s->S;
  auto cl = [S](int64_t v) { return std::min<int64_t>(std::max<int64_t>(v, 0), S - 1); };
  for (int64_t x = x0; x < x1; ++x)
    for (int64_t y = y0; y < y1; ++y) s->mask[(size_t)(cl(x) + cl(y) * S)] = EQF_DEFAULT_WALL;
  s->bnd_dirty = true;
  return EQF_OK;
}

extern "C" int eqf_set_mask(eqf_scene* s, const uint8_t* mask) {
  if (!s || !mask) return fail(EQF_EINVAL, "NULL argument");
  if (s->in_batch) return fail(EQF_EBUSY, "scene is part of a batch");
  const int S = s->S;
  for (int i = 0; i < S * S; ++i)
    if (mask[i] != EQF_DEFAULT_WALL && mask[i] != EQF_NO_WALL)
      return fail(EQF_EINVAL, "mask bytes must be EQF_DEFAULT_WALL (0) or EQF_NO_WALL (1)");
  for (int i = 0; i < S; ++i)
    if (mask[i] != EQF_DEFAULT_WALL || mask[i + (S - 1) * S] != EQF_DEFAULT_WALL || mask[i * S] != EQF_DEFAULT_WALL ||
        mask[S - 1 + i * S] != EQF_DEFAULT_WALL)
      return fail(EQF_EINVAL, "the outer ring of the mask must be EQF_DEFAULT_WALL (init_walls)");
  s->mask.assign(mask, mask + (size_t)S * S);
  s->bnd_dirty = true;
  return EQF_OK;
}

extern "C" int eqf_get_mask(const eqf_scene* s, uint8_t* mask) {
  if (!s || !mask) return fail(EQF_EINVAL, "NULL argument");
  std::memcpy(mask, s->mask.data(), s->mask.size());
  return EQF_OK;
}

__global__ void eqf_add_cell_kernel(float* a, float* b, float va, float vb) {
  a[0] += va;
  if (b) b[0] += vb;
}

__global__ void eqf_add_array_kernel(float* dst0, float* dst1, const float* src, int64_t n) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) {
    const float v = src[e];
    dst0[e] += v;
    if (dst1) dst1[e] += v;
  }
}

static int check_field(int f) {
  if (f < 0 || f >= NFIELDS) return fail(EQF_EINVAL, "field out of range");
  return EQF_OK;
}

static int64_t clamp_cell(const eqf_scene* s, uint32_t x, uint32_t y) {
  const uint32_t S = (uint32_t)s->S;
  return eqf_scene_off(s, (int)std::min(x, S - 1), (int)std::min(y, S - 1));
}

extern "C" int eqf_add_density(eqf_scene* s, uint32_t x, uint32_t y, float amount) {
  if (!s) return fail(EQF_EINVAL, "scene is NULL");
  if (int rc = eqf_internal_scene_enter(s)) return rc;
  const int64_t o = clamp_cell(s, x, y);
  hipLaunchKernelGGL(eqf_add_cell_kernel, dim3(1), dim3(1), 0, s->stream, field_ptr(s, F_D) + o,
                     field_ptr(s, F_SC) + o, amount, amount);
  HIPCHK(hipGetLastError());
  return EQF_OK;
}

extern "C" int eqf_add_velocity(eqf_scene* s, uint32_t x, uint32_t y, float ax, float ay) {
  if (!s) return fail(EQF_EINVAL, "scene is NULL");
  if (int rc = eqf_internal_scene_enter(s)) return rc;
  const int64_t o = clamp_cell(s, x, y);
  hipLaunchKernelGGL(eqf_add_cell_kernel, dim3(1), dim3(1), 0, s->stream, field_ptr(s, F_U) + o,
                     field_ptr(s, F_V) + o, ax, ay);
  HIPCHK(hipGetLastError());
  return EQF_OK;
}

extern "C" int eqf_add_cell(eqf_scene* s, int field, uint32_t x, uint32_t y, float amount) {
  if (!s) return fail(EQF_EINVAL, "scene is NULL");
  if (field == EQF_DENSITY) return eqf_add_density(s, x, y, amount);
  if (field != EQF_VX && field != EQF_VY) return fail(EQF_EINVAL, "field must be EQF_DENSITY, EQF_VX or EQF_VY");
  // one component only: adding 0 to the other would turn its -0 into +0
  if (int rc = eqf_internal_scene_enter(s)) return rc;
  const int64_t o = clamp_cell(s, x, y);
  hipLaunchKernelGGL(eqf_add_cell_kernel, dim3(1), dim3(1), 0, s->stream,
                     field_ptr(s, field == EQF_VX ? F_U : F_V) + o, (float*)nullptr, amount, 0.0f);
  HIPCHK(hipGetLastError());
  return EQF_OK;
}

extern "C" int eqf_add_noise(eqf_scene* s, float angle_degrees, uint32_t rand_x, uint32_t rand_y) {
  if (!s) return fail(EQF_EINVAL, "scene is NULL");
  // Rotate the line's end point (rand_x, rand_y) about the centre by
  // angle_degrees, all in f32 like the reference's f32 line string
  // (fluid.rs:586-590): geo 0.18 `rotate_around_point` (absent here; its
  // published rotation restated) turns degrees into radians with f32
  // `to_radians` (angle * (PI / 180)), takes f32 sin and cos, and maps
  // (x, y) to ((x - x0) cos - (y - y0) sin + x0, (x - x0) sin + (y - y0) cos + y0).
  const float c = (float)(s->S / 2);
  const float rad = angle_degrees * (3.14159265358979323846f / 180.0f);
  const float cs = std::cos(rad), sn = std::sin(rad);
  const float dx = (float)rand_x - c, dy = (float)rand_y - c;
  const float rx = (dx * cs - dy * sn) + c, ry = (dx * sn + dy * cs) + c;
  return eqf_add_velocity(s, (uint32_t)(s->S / 2), (uint32_t)(s->S / 2), rx * 2.0f, ry * 2.0f);
}

extern "C" int eqf_add_source(eqf_scene* s, int field, const float* amounts) {
  if (!s || !amounts) return fail(EQF_EINVAL, "NULL argument");
  if (field != EQF_DENSITY && field != EQF_VX && field != EQF_VY)
    return fail(EQF_EINVAL, "sources exist for EQF_DENSITY, EQF_VX and EQF_VY");
  if (int rc = eqf_internal_scene_enter(s)) return rc;
  return eqf_dev_upload(s, field_ptr(s, field), field == EQF_DENSITY ? field_ptr(s, F_SC) : nullptr, amounts, 1);
}

extern "C" int eqf_set_frame_source(eqf_scene* s, int field, const float* amounts) {
  if (!s) return fail(EQF_EINVAL, "scene is NULL");
  if (s->in_batch) return fail(EQF_EBUSY, "scene is part of a batch");
  if (field != EQF_DENSITY && field != EQF_VX && field != EQF_VY)
    return fail(EQF_EINVAL, "sources exist for EQF_DENSITY, EQF_VX and EQF_VY");
  const int slot = field == EQF_DENSITY ? 0 : field == EQF_VX ? 1 : 2;
  if (!s->d_src) {
    if (!amounts) return EQF_OK;
    HIPCHK(eqf_scene_malloc(s, (void**)&s->d_src, sizeof(float) * 3 * s->fstride));
    HIPCHK(hipMemsetAsync(s->d_src, 0, sizeof(float) * 3 * s->fstride, s->stream));
  }
  float* dst = s->d_src + (int64_t)slot * s->fstride;
  if (!amounts) {
    HIPCHK(hipMemsetAsync(dst, 0, sizeof(float) * s->fstride, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    s->src_mask &= ~(1 << slot);
    return EQF_OK;
  }
  const int rc = eqf_dev_upload(s, dst, nullptr, amounts, 0);
  if (!rc) s->src_mask |= 1 << slot;
  return rc;
}

extern "C" int eqf_upload(eqf_scene* s, int field, const float* host) {
  if (!s || !host) return fail(EQF_EINVAL, "NULL argument");
  int rc = check_field(field);
  if (!rc) rc = eqf_internal_scene_enter(s);
  if (rc) return rc;
  return upload_rm(s, field_ptr(s, field), host);
}

extern "C" int eqf_download(eqf_scene* s, int field, float* host) {
  if (!s || !host) return fail(EQF_EINVAL, "NULL argument");
  int rc = check_field(field);
  if (!rc) rc = eqf_internal_scene_enter(s);
  if (rc) return rc;
  return download_rm(s, field_ptr(s, field), host);
}

int eqf_scene_waves(int ns, int P, int sched) {
  return std::max(1, std::min(16, ns * cyc_streams(16, ns, P, sched)));
}

extern "C" int eqf_step(eqf_scene* s, int64_t nframes) {
  if (!s) return fail(EQF_EINVAL, "scene is NULL");
  if (s->in_batch) return fail(EQF_EBUSY, "scene is part of a batch");
  if (nframes < 0) return fail(EQF_EINVAL, "nframes < 0");
  if (nframes == 0) return EQF_OK;
  HIPCHK(hipSetDevice(s->ctx->device));
  int rc = sync_desc(s);
  if (rc) return rc;
  if (s->layout == L_ROWMAJOR) {
    std::string err;
    rc = rb_step(s, nframes, err);
    return rc ? fail(rc, err) : EQF_OK;
  }
  const int sched = eqf_pick_sched(s->ns, s->P, 16);
  const int W = eqf_scene_waves(s->ns, s->P, sched);
  while (nframes > 0) {
    const int nf = (int)std::min<int64_t>(nframes, 1 << 20);
    HIPCHK(eqf_internal_launch(s->d_desc, 1, W, cyc_lds_floats(s->ns, s->P, W, sched), nf, sched, s->stream));
    nframes -= nf;
  }
  return EQF_OK;
}

extern "C" int eqf_scene_set_tuning(eqf_scene* s, int tile, int sweeps_per_block) {
  if (!s) return fail(EQF_EINVAL, "scene is NULL");
  if (s->layout != L_ROWMAJOR) return fail(EQF_EINVAL, "tuning applies to red-black scenes");
  if (rb_set_tiling(s, tile, sweeps_per_block) != EQF_OK)
    return fail(EQF_EINVAL, "tile must be 0..5 and sweeps_per_block 1, 2, 4, 5, 6 or 8");
  return EQF_OK;
}

extern "C" int eqf_sync(eqf_scene* s) {
  if (!s) return fail(EQF_EINVAL, "scene is NULL");
  if (int rc = eqf_internal_scene_enter(s)) return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  return EQF_OK;
}

// ------------------------------------------------------------------ batch --
struct eqf_batch {
  eqf_ctx* ctx = nullptr;
  std::vector<eqf_scene*> given;   // the caller's order
  std::vector<eqf_scene*> scenes;  // workgroup order of the launch
  SceneDesc* d_descs = nullptr;
  int W = 1;
  int max_waves = 16;  // tuning: waves per workgroup at most (>= the largest strip count)
  int order = EQF_BATCH_LARGEST_FIRST;
  int sched_req = EQF_SCHED_AUTO;  // tuning: wavefront schedule (eqf_batch_set_schedule)
  int sched = GS_WIDE;             // the schedule in use (GsSched)
  void* rbg = nullptr;  // red-black batch: the launch group over all scenes (eqf_redblack.hip)
  hipStream_t stream = nullptr;
  hipGraphExec_t gexec = nullptr;
  int64_t graph_frames = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
};

// Workgroup order, descriptors and workgroup width of a batch from its
// tuning; uploads the descriptors on the batch stream.
static int batch_layout(eqf_batch* b) {
  b->scenes = b->given;
  if (b->order == EQF_BATCH_LARGEST_FIRST) {
    // their workgroups are the longest: dispatched first (longest-processing-
    // time order), they spread over the XCDs and set the launch's length
    std::stable_sort(b->scenes.begin(), b->scenes.end(), [](const eqf_scene* a, const eqf_scene* c) {
      return (int64_t)a->N * a->N * std::max(a->desc.diff_iters, 1) >
             (int64_t)c->N * c->N * std::max(c->desc.diff_iters, 1);
    });
  }
  const int n = (int)b->scenes.size();
  std::vector<SceneDesc> descs(n);
  int maxns = 1, maxw = 1, bigP = 16;  // bigP: the period of the largest scene (most strips)
  for (int i = 0; i < n; ++i) {
    descs[i] = b->scenes[i]->desc;
    const eqf_scene* s = b->scenes[i];
    if (s->ns > maxns || (s->ns == maxns && s->P > bigP)) bigP = s->P;
    maxns = std::max(maxns, s->ns);
  }
  b->sched = b->sched_req == EQF_SCHED_RING   ? GS_RING
             : b->sched_req == EQF_SCHED_WIDE ? GS_WIDE
             : b->sched_req == EQF_SCHED_FINE ? GS_FINE
                                              : eqf_pick_sched(maxns, bigP, b->max_waves);
  for (int i = 0; i < n; ++i) maxw = std::max(maxw, eqf_scene_waves(b->scenes[i]->ns, b->scenes[i]->P, b->sched));
  if (b->max_waves < maxns) return fail(EQF_EINVAL, "max_waves below the largest scene's strip count");
  int maxP = 0;
  for (int i = 0; i < n; ++i) maxP = std::max(maxP, b->scenes[i]->P);
  for (int i = 0; i < n; ++i) {
    descs[i].prio = eqf_batch_prio(b->scenes[i]->ns, maxns);
    descs[i].stream_l2 = eqf_stream_l2(maxP, n);
  }
  b->W = std::min(maxw, b->max_waves);
  hipError_t e =
      hipMemcpyAsync(b->d_descs, descs.data(), sizeof(SceneDesc) * n, hipMemcpyHostToDevice, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return fail(EQF_EDEVICE, std::string("batch layout: ") + hipGetErrorString(e));
  return EQF_OK;
}

extern "C" int eqf_batch_create(eqf_ctx* ctx, eqf_scene* const* scenes, int n, eqf_batch** out) {
  if (!ctx || !scenes || n <= 0 || !out) return fail(EQF_EINVAL, "bad arguments");
  {
    // a scene listed twice would be advanced twice per frame by concurrent workgroups
    std::vector<const eqf_scene*> sorted(scenes, scenes + n);
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
      return fail(EQF_EINVAL, "a scene appears more than once in the batch");
  }
  for (int i = 0; i < n; ++i) {
    if (!scenes[i]) return fail(EQF_EINVAL, "NULL scene");
    if (scenes[i]->layout != scenes[0]->layout)
      return fail(EQF_EINVAL, "a batch holds scenes of one mode (wavefront or red-black)");
    if (scenes[i]->in_batch) return fail(EQF_EBUSY, "scene already in a batch");
  }
  HIPCHK(hipSetDevice(ctx->device));
  for (int i = 0; i < n; ++i) {
    int rc = sync_desc(scenes[i]);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(scenes[i]->stream));
  }
  auto* b = new eqf_batch;
  b->ctx = ctx;
  b->given.assign(scenes, scenes + n);
  hipError_t e = hipStreamCreate(&b->stream);
  if (e == hipSuccess) e = hipMalloc(&b->d_descs, sizeof(SceneDesc) * n);
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  int rc = EQF_OK;
  if (e != hipSuccess) {
    rc = fail(EQF_EDEVICE, std::string("batch create: ") + hipGetErrorString(e));
  } else if (scenes[0]->layout == L_ROWMAJOR) {
    // red-black: every frame kernel covers all scenes, largest first
    b->scenes = b->given;
    std::stable_sort(b->scenes.begin(), b->scenes.end(),
                     [](const eqf_scene* a, const eqf_scene* c) { return a->N > c->N; });
    std::string err;
    b->rbg = rb_batch_create(b->scenes, b->stream, err);
    if (!b->rbg) rc = fail(EQF_EINVAL, err);
  } else {
    rc = batch_layout(b);
  }
  if (rc) {
    (void)hipFree(b->d_descs);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return rc;
  }
  for (auto* s : b->scenes) {
    s->in_batch = 1;
    s->batch = b;
    s->ops_pending = false;
  }
  *out = b;
  return EQF_OK;
}

extern "C" int eqf_batch_set_band(eqf_batch* b, int tile, int sweeps_per_block) {
  if (!b) return fail(EQF_EINVAL, "batch is NULL");
  if (!b->rbg) return fail(EQF_EINVAL, "band tiling applies to red-black batches");
  HIPCHK(hipSetDevice(b->ctx->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  if (rb_batch_set_tiling(b->rbg, tile, sweeps_per_block) != EQF_OK)
    return fail(EQF_EINVAL, "tile must be 0..5 and sweeps 1, 2, 4, 5, 6 or 8");
  return EQF_OK;
}

extern "C" int eqf_batch_set_schedule(eqf_batch* b, int sched) {
  if (!b) return fail(EQF_EINVAL, "batch is NULL");
  if (b->rbg) return fail(EQF_EINVAL, "red-black batches have no wavefront schedule");
  if (sched != EQF_SCHED_AUTO && sched != EQF_SCHED_WIDE && sched != EQF_SCHED_RING && sched != EQF_SCHED_FINE)
    return fail(EQF_EINVAL, "sched must be EQF_SCHED_AUTO, EQF_SCHED_WIDE, EQF_SCHED_RING or EQF_SCHED_FINE");
  HIPCHK(hipSetDevice(b->ctx->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  const int prev = b->sched_req;
  b->sched_req = sched;
  int rc = batch_layout(b);
  if (rc) {  // keep the previous schedule
    b->sched_req = prev;
    (void)batch_layout(b);
    return rc;
  }
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);  // the graph holds the launch
  b->gexec = nullptr;
  b->graph_frames = 0;
  return EQF_OK;
}

extern "C" int eqf_batch_set_tuning(eqf_batch* b, int max_waves, int order) {
  if (!b) return fail(EQF_EINVAL, "batch is NULL");
  if (b->rbg) return fail(EQF_EINVAL, "red-black batches are tuned with eqf_batch_set_band");
  if (max_waves < 1 || max_waves > 16) return fail(EQF_EINVAL, "max_waves must be 1..16");
  if (order != EQF_BATCH_LARGEST_FIRST && order != EQF_BATCH_GIVEN_ORDER) return fail(EQF_EINVAL, "bad order");
  HIPCHK(hipSetDevice(b->ctx->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  const int mw = b->max_waves, od = b->order;
  b->max_waves = max_waves;
  b->order = order;
  int rc = batch_layout(b);
  if (rc) {  // keep the previous tuning
    b->max_waves = mw;
    b->order = od;
    (void)batch_layout(b);
    return rc;
  }
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);  // the graph holds the launch shape
  b->gexec = nullptr;
  b->graph_frames = 0;
  return EQF_OK;
}

hipError_t eqf_internal_launch(const SceneDesc* d_descs, int n, int W, int lds_floats, int nframes, int sched,
                               hipStream_t st) {
  const size_t lds = sizeof(float) * (size_t)lds_floats;
#define EQF_LAUNCH_SCHED(S)                                                                               \
  do {                                                                                                    \
    if (W <= 8)                                                                                           \
      hipLaunchKernelGGL((eqf_frames_cyclic<gs_depth(S), GS_B, gs_sync(S), gs_has_ring(S), 512>), dim3(n), \
                         dim3(64 * W), lds, st, d_descs, nframes);                                        \
    else                                                                                                  \
      hipLaunchKernelGGL((eqf_frames_cyclic<gs_depth(S), GS_B, gs_sync(S), gs_has_ring(S), 1024>), dim3(n), \
                         dim3(64 * W), lds, st, d_descs, nframes);                                        \
  } while (0)
  if (sched == GS_FINE) EQF_LAUNCH_SCHED(GS_FINE);
  else if (sched == GS_RING) EQF_LAUNCH_SCHED(GS_RING);
  else EQF_LAUNCH_SCHED(GS_WIDE);
#undef EQF_LAUNCH_SCHED
  return hipGetLastError();
}

// Wavefront schedule of a launch whose largest scene (the longest frame
// chain) has maxns strips and period P, with at most max_waves waves per
// workgroup.  The ring schedules hand x0 between sweep streams in LDS (fewer
// vector-memory instructions and HBM bytes) and, with their shorter stream
// lags, run more streams per strip (P / 48 for GS_RING, P / 32 for GS_FINE,
// against P / 64), at the cost of a barrier every 8 or 4 steps.  The rule:
// GS_FINE when it gives that scene more streams than GS_RING (N 49..192),
// else GS_RING when it gives more than GS_WIDE or two or more streams and at
// least 8 waves (N 193..512), else GS_WIDE (N > 512: one stream per strip;
// N <= 48: one stream under every schedule).  Measured (ms per frame, 20 sweeps, 16 waves at
// most, wide / ring / fine, one box): single scenes N=64 0.93 / 0.96 / 0.62,
// N=100 1.73 / 1.12 / 0.92, N=128 1.14 / 1.22 / 0.86, N=150 1.56 / 1.34 /
// 1.21, N=192 1.53 / 1.34 / 1.33, N=256 2.12 / 2.01 / 2.04, N=320 3.20 / 3.01
// / 3.15, N=512 7.37 / 7.00 / 7.37, N=600 (one stream) 12.97 / 13.38 / -;
// batches: 130 N=64 0.95 / 0.97 / 0.63, 118 N=128 1.18 / 1.27 / 0.91, 130
// N=256 2.66 / 2.44 / 2.41, 391 scenes N <= 256 3.09 / 2.94 / 2.76, the
// 512-scene mix 12.23 / 10.94 / 10.73 (the tuner picks per batch), 121 N=512
// with 8 waves (one stream per strip) 12.06 / 12.36 / -.
int eqf_pick_sched(int maxns, int P, int max_waves) {
  const int mf = cyc_streams(max_waves, maxns, P, GS_FINE);
  const int mr = cyc_streams(max_waves, maxns, P, GS_RING);
  const int mw = cyc_streams(max_waves, maxns, P, GS_WIDE);
  if (mf > mr) return GS_FINE;
  return (mr > mw || (mr >= 2 && maxns * mr >= 8)) ? GS_RING : GS_WIDE;
}

// Non-temporal streaming (SceneDesc::stream_l2) for the n scenes of a launch
// whose largest scene has period maxP: fields of P > 256 (over 256 KB each)
// are not found in L2 again by the next phase, and smaller scenes sharing
// their launch leave L2 to the big scenes' sweep re-reads (batched mix 12.4
// -> 12.2 ms per frame, wide schedule); a launch of small scenes only keeps
// L2 for itself (391 scenes N <= 256: 3.27 -> 3.00 ms).  Level 2 also
// streams the x0 loads of ring writers (their lines' last uses): with many
// large scenes sharing L2 it pays (the 512-scene mix 11.49 -> 10.98 ms, 121
// N=512 9.99 -> 9.85), with few it costs (one N=512 6.93 -> 7.11, 32 N=512
// 7.38 -> 7.81).
int eqf_stream_l2(int maxP, int n) { return maxP > 256 ? (n >= 64 ? 2 : 1) : 0; }

int eqf_batch_prio(int ns, int maxns) {
  int p = 3;
  for (int m = maxns; p > 0 && 2 * ns <= m; m /= 2) --p;
  return p;
}

int eqf_internal_prepare(eqf_scene* s) { return sync_desc(s); }

int eqf_internal_scene_enter(eqf_scene* s) {
  if (!s->batch) return EQF_OK;
  // ev1 is recorded after every step of the batch (an event never recorded
  // counts as complete)
  HIPCHK(hipStreamWaitEvent(s->stream, s->batch->ev1, 0));
  s->ops_pending = true;
  return EQF_OK;
}
int eqf_internal_error(int code, const std::string& msg) { return fail(code, msg); }
int eqf_internal_check_params(const eqf_params* p) { return check_params(p); }

static hipError_t batch_launch(eqf_batch* b, int nframes_per_launch) {
  int lds = 0;
  for (auto* s : b->scenes) lds = std::max(lds, cyc_lds_floats(s->ns, s->P, b->W, b->sched));
  return eqf_internal_launch(b->d_descs, (int)b->scenes.size(), b->W, lds, nframes_per_launch, b->sched,
                             b->stream);
}

extern "C" int eqf_batch_step(eqf_batch* b, int64_t nframes) {
  if (!b) return fail(EQF_EINVAL, "batch is NULL");
  if (nframes < 0) return fail(EQF_EINVAL, "nframes < 0");
  if (nframes == 0) return EQF_OK;
  HIPCHK(hipSetDevice(b->ctx->device));
  // work the caller queued on member scenes since the last step (sources,
  // uploads) runs before this step
  for (eqf_scene* s : b->scenes) {
    if (!s->ops_pending) continue;
    if (!s->ev_ops) HIPCHK(hipEventCreateWithFlags(&s->ev_ops, hipEventDisableTiming));
    HIPCHK(hipEventRecord(s->ev_ops, s->stream));
    HIPCHK(hipStreamWaitEvent(b->stream, s->ev_ops, 0));
    s->ops_pending = false;
  }
  if (b->rbg) {  // red-black: the group's frames, captured once per nframes
    HIPCHK(hipEventRecord(b->ev0, b->stream));
    std::string err;
    const int rc = rb_batch_step(b->rbg, nframes, err);
    if (rc) return fail(rc, err);
    HIPCHK(hipEventRecord(b->ev1, b->stream));
    return EQF_OK;
  }
  // One launch per frame, captured once as a graph of `nframes` launches.
  if (!b->gexec || b->graph_frames != nframes) {
    if (b->gexec) {
      HIPCHK(hipStreamSynchronize(b->stream));  // the previous replay may still run
      HIPCHK(hipGraphExecDestroy(b->gexec));
    }
    b->gexec = nullptr;
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t le = hipSuccess;
    for (int64_t f = 0; f < nframes && le == hipSuccess; ++f) le = batch_launch(b, 1);
    hipError_t ce = hipStreamEndCapture(b->stream, &g);
    if (le != hipSuccess) return fail(EQF_EDEVICE, std::string("batch capture: ") + hipGetErrorString(le));
    HIPCHK(ce);
    HIPCHK(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
    b->graph_frames = nframes;
  }
  HIPCHK(hipEventRecord(b->ev0, b->stream));
  HIPCHK(hipGraphLaunch(b->gexec, b->stream));
  HIPCHK(hipEventRecord(b->ev1, b->stream));
  return EQF_OK;
}

extern "C" int eqf_batch_sync(eqf_batch* b) {
  if (!b) return fail(EQF_EINVAL, "batch is NULL");
  HIPCHK(hipStreamSynchronize(b->stream));
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, b->ev0, b->ev1) == hipSuccess) b->last_ms = ms;
  return EQF_OK;
}

extern "C" int eqf_batch_last_ms(const eqf_batch* b, double* ms) {
  if (!b || !ms) return fail(EQF_EINVAL, "NULL argument");
  *ms = b->last_ms;
  return EQF_OK;
}

extern "C" int eqf_batch_destroy(eqf_batch* b) {
  if (!b) return EQF_OK;
  (void)hipStreamSynchronize(b->stream);
  for (auto* s : b->scenes)  // work that waits on the batch's events ends before they go
    if (s->ops_pending) (void)hipStreamSynchronize(s->stream);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  rb_batch_destroy(b->rbg);
  for (auto* s : b->scenes) {
    s->in_batch = 0;
    s->batch = nullptr;
    s->ops_pending = false;
  }
  (void)hipFree(b->d_descs);
  (void)hipEventDestroy(b->ev0);
  (void)hipEventDestroy(b->ev1);
  (void)hipStreamDestroy(b->stream);
  delete b;
  return EQF_OK;
}
