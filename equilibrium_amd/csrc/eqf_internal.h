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
  eqf_params p{};
  int S = 0, N = 0, ns = 0;
  int layout = eqf::L_CYCLIC;
  int64_t fstride = 0;
  std::vector<uint8_t> mask;  // row-major, host copy
  bool bnd_dirty = true;
  hipStream_t stream = nullptr;
  float* d_fld = nullptr;
  float* d_src = nullptr;
  uint8_t* d_code = nullptr;
  uint8_t* d_rowf = nullptr;
  int2* d_bnd = nullptr;
  size_t bnd_cap = 0;
  int nbnd[3] = {0, 0, 0};
  float* d_hraw = nullptr;
  eqf::SceneDesc* d_desc = nullptr;
  eqf::SceneDesc desc{};
  std::vector<float> hbuf;  // layout staging
  int in_batch = 0;
  void* rb = nullptr;       // red-black workspace (eqf_redblack.hip)
};

// Runtime services shared with the server (eqf_runtime.hip).
int eqf_internal_prepare(eqf_scene* s);  // boundary lists + device descriptor
hipError_t eqf_internal_launch(const eqf::SceneDesc* d_descs, int n, int maxns, int nframes, hipStream_t st);
int eqf_internal_error(int code, const std::string& msg);
int eqf_internal_check_params(const eqf_params* p);

namespace eqf {
// Red-black mode (eqf_redblack.hip).
void* rb_create(eqf_scene* s, std::string& err);
void rb_destroy(void* rb);
int rb_step(eqf_scene* s, int64_t nframes, std::string& err);
int rb_set_tiling(eqf_scene* s, int tile, int tsweeps);
}  // namespace eqf
