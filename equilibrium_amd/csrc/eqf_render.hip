// eqf_render.hip -- the renderer's side of the path, on the device:
//  * eqf_render_rgba: RenderingListener::render_image's pixel rule
//    (renderer_helpers.rs:143-165) from the scene's density and mask;
//  * snapshot rings: CurrentSimulation::simulate hands the renderer a clone
//    of the Fluid after every step (renderer_helpers.rs:59-65).  A capture
//    converts density (and the RGBA image) on the scene's stream behind the
//    queued steps, then a copy stream moves it to pinned host memory, so the
//    transfer overlaps the following steps instead of stalling them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "eqf_internal.h"

using namespace eqf;

namespace {

// Rust `f as u8` (saturating: NaN -> 0, below 0 -> 0, above 255 -> 255,
// otherwise truncation toward zero).
__device__ __forceinline__ uint32_t rust_u8(float f) {
  if (!(f > 0.0f)) return 0;
  if (f >= 255.0f) return 255;
  return (uint32_t)f;
}

__device__ __forceinline__ uint32_t pixel(float dens, uint32_t cell, uint32_t fluid, uint32_t world,
                                          uint32_t obstacle) {
  if (cell == EQF_DEFAULT_WALL) return obstacle;
  // [(density * fluid.r as f32) as u8, fluid.g, density as u8, 1]
  if (dens != 0.0f && cell == EQF_NO_WALL)
    return rust_u8(dens * (float)(fluid & 0xffu)) | (fluid & 0xff00u) | (rust_u8(dens) << 16) | (1u << 24);
  return world;
}

// Four pixels per thread (16 B density, 4 B mask in, 16 B RGBA out):
// rgba[e] for cells e = x + y*S of the row-major density `d` and mask `m`
// (renderer_helpers.rs:143-165).  `d`, `m` and `rgba` are 16-byte aligned;
// the last thread takes the n % 4 tail one pixel at a time.
__global__ void k_render(int64_t n, const float* __restrict__ d, const uint8_t* __restrict__ m, uint32_t fluid,
                         uint32_t world, uint32_t obstacle, uint32_t* __restrict__ rgba) {
  const int64_t e0 = 4 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
  if (e0 >= n) return;
  if (e0 + 4 <= n) {
    const float4 dv = *reinterpret_cast<const float4*>(d + e0);
    const uint32_t mv = *reinterpret_cast<const uint32_t*>(m + e0);
    uint4 o;
    o.x = pixel(dv.x, mv & 0xffu, fluid, world, obstacle);
    o.y = pixel(dv.y, (mv >> 8) & 0xffu, fluid, world, obstacle);
    o.z = pixel(dv.z, (mv >> 16) & 0xffu, fluid, world, obstacle);
    o.w = pixel(dv.w, mv >> 24, fluid, world, obstacle);
    *reinterpret_cast<uint4*>(rgba + e0) = o;
  } else {
    for (int64_t e = e0; e < n; ++e) rgba[e] = pixel(d[e], m[e], fluid, world, obstacle);
  }
}

uint32_t pack(const uint8_t c[4]) {
  return (uint32_t)c[0] | ((uint32_t)c[1] << 8) | ((uint32_t)c[2] << 16) | ((uint32_t)c[3] << 24);
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

// RGBA follows the density in one staging buffer, at a 16-byte boundary.
size_t rgba_offset_floats(size_t n) { return (n + 3) & ~(size_t)3; }

}  // namespace

// Density of the scene as a row-major device array and its mask on the
// device (rebuilt if the mask changed since the boundary data was built).
static int render_inputs(eqf_scene* s, float* d_dens_rm) {
  int rc = eqf_internal_prepare(s);  // boundary data, incl. the device mask
  if (rc) return rc;
  if (!s->d_mask) {
    rc = eqf_dev_rebuild_bnd(s);
    if (rc) return rc;
  }
  return eqf_dev_to_rm_async(s, s->d_fld + (int64_t)F_D * s->fstride, d_dens_rm);
}

static int launch_render(eqf_scene* s, const float* d_dens_rm, const eqf_colors* c, uint8_t* d_rgba) {
  const int64_t n = (int64_t)s->S * s->S;
  hipLaunchKernelGGL(k_render, dim3(blocks_for((n + 3) / 4)), dim3(256), 0, s->stream, n, d_dens_rm, s->d_mask,
                     pack(c->fluid), pack(c->world), pack(c->obstacle), (uint32_t*)d_rgba);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? EQF_OK : eqf_internal_error(EQF_EDEVICE, std::string("render: ") + hipGetErrorString(e));
}

extern "C" void eqf_colors_default(eqf_colors* c) {
  if (!c) return;
  // configs.rs:56-57 (fluid_color, world_color); RenderingListener's obstacles_color
  // is egui's Color32::RED (renderer_helpers.rs:97-103)
  const uint8_t fl[4] = {208, 88, 157, 220}, wo[4] = {94, 146, 162, 128}, ob[4] = {255, 0, 0, 255};
  std::copy(fl, fl + 4, c->fluid);
  std::copy(wo, wo + 4, c->world);
  std::copy(ob, ob + 4, c->obstacle);
}

extern "C" int eqf_render_rgba(eqf_scene* s, const eqf_colors* c, uint8_t* rgba_host) {
  if (!s || !rgba_host) return eqf_internal_error(EQF_EINVAL, "bad arguments");
  if (int rc = eqf_internal_scene_enter(s)) return rc;
  eqf_colors def;
  if (!c) {
    eqf_colors_default(&def);
    c = &def;
  }
  hipError_t e = hipSetDevice(s->ctx->device);
  const size_t n = (size_t)s->S * s->S;
  float* d_dens = nullptr;
  uint8_t* d_rgba = nullptr;
  if (e == hipSuccess) e = hipMalloc(&d_dens, sizeof(float) * rgba_offset_floats(n) + 4 * n);
  if (e != hipSuccess) return eqf_internal_error(EQF_ENOMEM, std::string("render buffer: ") + hipGetErrorString(e));
  d_rgba = (uint8_t*)(d_dens + rgba_offset_floats(n));
  int rc = render_inputs(s, d_dens);
  if (!rc) rc = launch_render(s, d_dens, c, d_rgba);
  if (!rc) {
    e = hipMemcpyAsync(rgba_host, d_rgba, 4 * n, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) rc = eqf_internal_error(EQF_EDEVICE, std::string("render copy: ") + hipGetErrorString(e));
  }
  (void)hipFree(d_dens);
  return rc;
}

// ---------------------------------------------------------------- snapshots
struct eqf_snapshot {
  eqf_scene* scene = nullptr;
  int device = 0;
  int depth = 0, with_rgba = 0;
  eqf_colors colors{};
  hipStream_t copy = nullptr;
  struct Slot {
    float* d_dens = nullptr;  // device staging (row-major density, then RGBA)
    uint8_t* d_rgba = nullptr;
    float* h_dens = nullptr;  // pinned host copy
    uint8_t* h_rgba = nullptr;
    hipEvent_t ready = nullptr, done = nullptr;
    int64_t frame = 0;
  };
  std::vector<Slot> slots;
  int head = 0, count = 0;  // oldest capture, captures held
};

static void snapshot_free(eqf_snapshot* sn) {
  for (auto& sl : sn->slots) {
    (void)hipFree(sl.d_dens);
    (void)hipHostFree(sl.h_dens);
    if (sl.ready) (void)hipEventDestroy(sl.ready);
    if (sl.done) (void)hipEventDestroy(sl.done);
  }
  if (sn->copy) (void)hipStreamDestroy(sn->copy);
  delete sn;
}

extern "C" int eqf_snapshot_create(eqf_scene* s, int depth, int with_rgba, const eqf_colors* c, eqf_snapshot** out) {
  if (!s || !out || depth < 1 || depth > 8) return eqf_internal_error(EQF_EINVAL, "bad arguments (depth 1..8)");
  hipError_t e = hipSetDevice(s->ctx->device);
  auto* sn = new eqf_snapshot;
  sn->scene = s;
  sn->device = s->ctx->device;
  sn->depth = depth;
  sn->with_rgba = with_rgba ? 1 : 0;
  if (c) sn->colors = *c;
  else eqf_colors_default(&sn->colors);
  const size_t n = (size_t)s->S * s->S;
  const size_t bytes = sn->with_rgba ? sizeof(float) * rgba_offset_floats(n) + 4 * n : sizeof(float) * n;
  if (e == hipSuccess) e = hipStreamCreate(&sn->copy);
  sn->slots.resize(depth);
  for (auto& sl : sn->slots) {
    if (e == hipSuccess) e = hipMalloc(&sl.d_dens, bytes);
    if (e == hipSuccess) e = hipHostMalloc(&sl.h_dens, bytes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.ready, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
    if (e == hipSuccess && sn->with_rgba) {
      sl.d_rgba = (uint8_t*)(sl.d_dens + rgba_offset_floats(n));
      sl.h_rgba = (uint8_t*)(sl.h_dens + rgba_offset_floats(n));
    }
  }
  if (e != hipSuccess) {
    snapshot_free(sn);
    return eqf_internal_error(EQF_ENOMEM, std::string("snapshot ring: ") + hipGetErrorString(e));
  }
  *out = sn;
  return EQF_OK;
}

extern "C" int eqf_snapshot_capture(eqf_snapshot* sn, int64_t frame) {
  if (!sn) return eqf_internal_error(EQF_EINVAL, "snapshot is NULL");
  if (sn->count == sn->depth) return eqf_internal_error(EQF_EBUSY, "every snapshot slot holds an unreleased capture");
  eqf_scene* s = sn->scene;
  if (int rc = eqf_internal_scene_enter(s)) return rc;
  hipError_t e = hipSetDevice(s->ctx->device);
  if (e != hipSuccess) return eqf_internal_error(EQF_EDEVICE, hipGetErrorString(e));
  auto& sl = sn->slots[(sn->head + sn->count) % sn->depth];
  const size_t n = (size_t)s->S * s->S;
  // A slot released without an acquire may still be copying out: its next
  // contents wait for that copy.  (An event never recorded is complete.)
  e = hipStreamWaitEvent(s->stream, sl.done, 0);
  if (e != hipSuccess) return eqf_internal_error(EQF_EDEVICE, std::string("snapshot: ") + hipGetErrorString(e));
  int rc = render_inputs(s, sl.d_dens);
  if (!rc && sn->with_rgba) rc = launch_render(s, sl.d_dens, &sn->colors, sl.d_rgba);
  if (rc) return rc;
  const size_t bytes = sn->with_rgba ? sizeof(float) * rgba_offset_floats(n) + 4 * n : sizeof(float) * n;
  e = hipEventRecord(sl.ready, s->stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(sn->copy, sl.ready, 0);
  if (e == hipSuccess) e = hipMemcpyAsync(sl.h_dens, sl.d_dens, bytes, hipMemcpyDeviceToHost, sn->copy);
  if (e == hipSuccess) e = hipEventRecord(sl.done, sn->copy);
  if (e != hipSuccess) return eqf_internal_error(EQF_EDEVICE, std::string("snapshot: ") + hipGetErrorString(e));
  sl.frame = frame;
  ++sn->count;
  return EQF_OK;
}

extern "C" int eqf_snapshot_acquire(eqf_snapshot* sn, int64_t* frame, const float** density, const uint8_t** rgba) {
  if (!sn) return eqf_internal_error(EQF_EINVAL, "snapshot is NULL");
  if (sn->count == 0) return eqf_internal_error(EQF_EINVAL, "no capture to acquire");
  auto& sl = sn->slots[sn->head];
  hipError_t e = hipEventSynchronize(sl.done);
  if (e != hipSuccess) return eqf_internal_error(EQF_EDEVICE, std::string("snapshot wait: ") + hipGetErrorString(e));
  if (frame) *frame = sl.frame;
  if (density) *density = sl.h_dens;
  if (rgba) *rgba = sn->with_rgba ? sl.h_rgba : nullptr;
  return EQF_OK;
}

extern "C" int eqf_snapshot_release(eqf_snapshot* sn) {
  if (!sn) return eqf_internal_error(EQF_EINVAL, "snapshot is NULL");
  if (sn->count == 0) return eqf_internal_error(EQF_EINVAL, "no capture to release");
  sn->head = (sn->head + 1) % sn->depth;
  --sn->count;
  return EQF_OK;
}

extern "C" int eqf_snapshot_destroy(eqf_snapshot* sn) {
  if (!sn) return EQF_OK;
  // The copy stream's last copy waited for every capture's device work, so
  // draining it leaves no work on the slots (the scene itself is not touched).
  (void)hipSetDevice(sn->device);
  (void)hipStreamSynchronize(sn->copy);
  snapshot_free(sn);
  return EQF_OK;
}

// ------------------------------------------------- allocation-free render --
// eqf_render_rgba_noalloc stages the row-major density and the RGBA image in
// `scratch`, a fixed device array of EQF_NOALLOC_SCRATCH_BYTES that every GPU
// the library runs on holds until the process ends, instead of the device
// buffer eqf_render_rgba allocates and frees per call.  In host code
// `scratch` names a host-side stand-in, never the device memory: the device
// address comes from hipGetSymbolAddress, once per device with that device
// current (eqf_ctx_create calls eqf_internal_scratch_init), and only that
// address is used below.  The staging is fully rewritten by each call, so
// nothing in it must start at zero.
static __device__ __align__(256) unsigned char scratch[EQF_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;                  // guards g_scratch_ptr
void* g_scratch_ptr[kMaxDevices] = {};    // device address of `scratch`, per device
std::mutex g_scratch_use[kMaxDevices];    // one call at a time per device
}  // namespace

int eqf_internal_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices)
    return eqf_internal_error(EQF_EINVAL, "device index too large for the render scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return EQF_OK;
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e == hipSuccess && cur != device) e = hipSetDevice(device);
  void* p = nullptr;
  if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return eqf_internal_error(EQF_EDEVICE, std::string("render scratch address: ") + hipGetErrorString(e));
  g_scratch_ptr[device] = p;
  return EQF_OK;
}

extern "C" int eqf_render_rgba_noalloc(eqf_scene* s, const eqf_colors* c, uint8_t* rgba_host) {
  if (!s || !rgba_host) return eqf_internal_error(EQF_EINVAL, "bad arguments");
  const size_t n = (size_t)s->S * s->S;
  const size_t need = sizeof(float) * rgba_offset_floats(n) + 4 * n;
  if (need > (size_t)EQF_NOALLOC_SCRATCH_BYTES)
    return eqf_internal_error(EQF_EINVAL, "scene too large for eqf_render_rgba_noalloc (size <= " +
                                              std::to_string(EQF_NOALLOC_RENDER_MAX_SIZE) +
                                              "); use eqf_render_rgba");
  // The device mask is built by the first step after a mask change (it is
  // allocated then); this entry point only reads it.
  if (s->bnd_dirty || !s->d_mask)
    return eqf_internal_error(EQF_EINVAL, "the scene's mask changed since its last step (or it was never "
                                          "stepped): step it, or call eqf_render_rgba, first");
  const int dev = s->ctx->device;
  if (dev < 0 || dev >= kMaxDevices) return eqf_internal_error(EQF_EINVAL, "device index too large");
  hipError_t e = hipSetDevice(dev);
  if (e != hipSuccess) return eqf_internal_error(EQF_EDEVICE, hipGetErrorString(e));
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = g_scratch_ptr[dev];
  }
  if (!base) {  // a context created by an older copy of the library state
    int rc = eqf_internal_scratch_init(dev);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = g_scratch_ptr[dev];
  }
  eqf_colors def;
  if (!c) {
    eqf_colors_default(&def);
    c = &def;
  }
  std::lock_guard<std::mutex> use(g_scratch_use[dev]);
  int rc = eqf_internal_scene_enter(s);
  if (rc) return rc;
  float* d_dens = (float*)base;
  uint8_t* d_rgba = (uint8_t*)(d_dens + rgba_offset_floats(n));
  rc = eqf_dev_to_rm_async(s, s->d_fld + (int64_t)F_D * s->fstride, d_dens);
  if (!rc) rc = launch_render(s, d_dens, c, d_rgba);
  if (rc) return rc;
  e = hipMemcpyAsync(rgba_host, d_rgba, 4 * n, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);  // the next call may reuse the scratch
  if (e != hipSuccess) return eqf_internal_error(EQF_EDEVICE, std::string("render copy: ") + hipGetErrorString(e));
  return EQF_OK;
}
