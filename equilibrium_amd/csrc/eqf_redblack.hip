// eqf_redblack.hip -- red-black Gauss-Seidel mode (row-major layout).
#include <hip/hip_runtime.h>

#include "eqf_internal.h"

namespace eqf {
void* rb_create(eqf_scene*, std::string& err) {
  err = "red-black mode is not available in this build";
  return nullptr;
}
void rb_destroy(void*) {}
int rb_step(eqf_scene*, int64_t, std::string& err) {
  err = "red-black mode is not available in this build";
  return EQF_EINVAL;
}
}  // namespace eqf
