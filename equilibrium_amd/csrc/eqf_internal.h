// eqf_internal.h -- host-side objects of libeqfluid.so shared by the
// runtime (eqf_runtime.hip) and the red-black driver (eqf_redblack.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/eqfluid.h"
#include "eqf_layout.h"

struct eqf_ctx {
  int device;
};

struct eqf_scene {
  eqf_ctx* ctx = nullptr;
  uint64_t uid = 0;  // unique per scene object of the process (never reused)
  eqf_params p{};
  int S = 0, N = 0, ns = 0, P = 0;
  int layout = eqf::L_CYCLIC;
  int64_t fstride = 0;
  std::vector<uint8_t> mask;  // row-major, host copy
  bool bnd_dirty = true;
  hipStream_t stream = nullptr;
  float* d_fld = nullptr;
  float* d_src = nullptr;
  int src_mask = 0;           // frame sources set: bit 0 density, 1 vx, 2 vy
  uint8_t* d_code = nullptr;
  uint8_t* d_rowf = nullptr;
  int2* d_bnd = nullptr;
  size_t bnd_cap = 0;
  int nbnd[3] = {0, 0, 0};
  void* d_main = nullptr;     // one allocation: fields, codes, rowf, hand-off rows, descriptor
  float* d_stage = nullptr;   // row-major staging for uploads/downloads (lazy)
  uint8_t* d_mask = nullptr;  // mask + scratch for the boundary build (lazy)
  eqf::SceneDesc* d_desc = nullptr;
  eqf::SceneDesc desc{};
  std::vector<float> hbuf;  // layout staging
  int in_batch = 0;
  // Batch membership: the batch advances the scene on its own stream.  Work
  // queued on the scene's stream meanwhile first waits for the batch's last
  // step (eqf_internal_scene_enter), and the batch's next step waits for it
  // (ev_ops, recorded when ops_pending).
  struct eqf_batch* batch = nullptr;
  hipEvent_t ev_ops = nullptr;
  bool ops_pending = false;
  bool owns_stream = true;
  void* rb = nullptr;       // red-black workspace (eqf_redblack.hip)
  // Scenes of a server allocate from the server's memory pool, ordered on
  // pool_stream (hipMallocFromPoolAsync / hipFreeAsync): freeing a retired
  // scene then waits for nothing else on the device, where hipFree would wait
  // for every stream.  Other scenes use hipMalloc / hipFree.
  hipMemPool_t pool = nullptr;
  hipStream_t pool_stream = nullptr;
};

// Device memory of a scene (see eqf_scene::pool).
hipError_t eqf_scene_malloc(eqf_scene* s, void** p, size_t bytes);
void eqf_scene_mfree(eqf_scene* s, void* p);

// Runtime services shared with the server (eqf_runtime.hip).
int eqf_internal_prepare(eqf_scene* s);  // boundary lists + device descriptor
// Issue priority of a scene among the scenes of one launch: 3 for the largest
// (strip count), then 2, 1, 0 for each halving.
int eqf_batch_prio(int ns, int maxns);
int eqf_stream_l2(int maxP, int n);
// Waves a wavefront-mode scene uses under schedule `sched` (GsSched): its
// strips times the sweep streams per strip (up to 16 waves, at most
// P / stream lag streams).
int eqf_scene_waves(int ns, int P, int sched);
// The schedule (GsSched) of a launch whose largest scene has maxns strips
// and period P, with at most max_waves waves per workgroup.
int eqf_pick_sched(int maxns, int P, int max_waves);
// waves / lds_floats: the largest eqf_scene_waves() and cyc_lds_floats() of
// the scenes under schedule `sched`.
hipError_t eqf_internal_launch(const eqf::SceneDesc* d_descs, int n, int waves, int lds_floats, int nframes,
                               int sched, hipStream_t st);
int eqf_internal_error(int code, const std::string& msg);
// Device address of eqf_render_rgba_noalloc's fixed scratch on `device`
// (eqf_render.hip), looked up once per device at context creation.
int eqf_internal_scratch_init(int device);
// Orders work about to be queued on the scene's stream after the last step of
// the scene's batch (no-op outside a batch).
int eqf_internal_scene_enter(eqf_scene* s);
int eqf_internal_check_params(const eqf_params* p);
// eqf_scene_create on a caller-owned stream (the server's admission stream),
// allocating from `pool` in that stream's order when `pool` is set.
int eqf_internal_scene_create(eqf_ctx* ctx, const eqf_params* p, hipStream_t stream, hipMemPool_t pool,
                              eqf_scene** out);

// Device-side setup (eqf_setup.hip).
int eqf_dev_init_fields(eqf_scene* s, int times);
int eqf_dev_upload(eqf_scene* s, float* dst, float* dst2, const float* host, int add);
int eqf_dev_download(eqf_scene* s, const float* src, float* host);
// src field (scene layout) -> row-major device array, queued on the scene's stream.
int eqf_dev_to_rm_async(eqf_scene* s, const float* src, float* dst);
int eqf_dev_rebuild_bnd(eqf_scene* s);
void eqf_dev_trim(eqf_scene* s);

namespace eqf {
// Red-black mode (eqf_redblack.hip).
void* rb_create(eqf_scene* s, std::string& err);
void rb_destroy(eqf_scene* s);
int rb_step(eqf_scene* s, int64_t nframes, std::string& err);
int rb_set_tiling(eqf_scene* s, int tile, int tsweeps);
// A batch of red-black scenes: every frame kernel covers all of them.
void* rb_batch_create(const std::vector<eqf_scene*>& scenes, hipStream_t st, std::string& err);
void rb_batch_destroy(void* g);
int rb_batch_set_tiling(void* g, int tile, int tsweeps);
int rb_batch_step(void* g, int64_t nframes, std::string& err);
void rb_batch_set_scenes(void* g, const std::vector<eqf_scene*>& scenes);
}  // namespace eqf
