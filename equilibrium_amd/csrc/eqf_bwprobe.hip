// eqf_bwprobe.hip -- practical HBM bandwidth ceiling of the GPU, measured
// beside the stencil kernels (libeqf_bwprobe.so, loaded by bench.py only;
// not part of the simulation library).
//
// Three streaming kernels over a working set far above the 256 MB MALL, so
// that every byte crosses HBM: read (sum of float4 loads), write (float4
// stores) and copy (float4 load + store).  Each thread keeps U independent
// 16-byte accesses in flight per iteration; the grid is a multiple of the CU
// count with a grid-stride loop.  The host times R launches with HIP events
// and reports bytes moved / time.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

template <int U>
__global__ void __launch_bounds__(256) k_read(const f4* __restrict__ src, int64_t n4, float* __restrict__ sink) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  for (; e + (U - 1) * stride < n4; e += U * stride) {
    f4 r[U];
#pragma unroll
    for (int k = 0; k < U; ++k) r[k] = __builtin_nontemporal_load(src + e + k * stride);
#pragma unroll
    for (int k = 0; k < U; ++k) acc += r[k];
  }
  for (; e < n4; e += stride) acc += src[e];
  const float s = acc.x + acc.y + acc.z + acc.w;
  if (s == 1.2345f) sink[0] = s;  // keeps the loads; the source holds zeros
}

template <int U>
__global__ void __launch_bounds__(256) k_write(f4* __restrict__ dst, int64_t n4, float v) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const f4 w = {v, v, v, v};
  for (; e + (U - 1) * stride < n4; e += U * stride) {
#pragma unroll
    for (int k = 0; k < U; ++k) __builtin_nontemporal_store(w, dst + e + k * stride);
  }
  for (; e < n4; e += stride) dst[e] = w;
}

template <int U>
__global__ void __launch_bounds__(256) k_copy(const f4* __restrict__ src, f4* __restrict__ dst, int64_t n4) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; e + (U - 1) * stride < n4; e += U * stride) {
    f4 r[U];
#pragma unroll
    for (int k = 0; k < U; ++k) r[k] = __builtin_nontemporal_load(src + e + k * stride);
#pragma unroll
    for (int k = 0; k < U; ++k) __builtin_nontemporal_store(r[k], dst + e + k * stride);
  }
  for (; e < n4; e += stride) dst[e] = src[e];
}

}  // namespace

// Measures the three kernels on a `bytes`-sized working set (read: the whole
// buffer; write: the whole buffer; copy: half to half) with `blocks_per_cu`
// workgroups of 256 threads per CU and `reps` timed launches each (after one
// warm-up).  Writes GB/s (1e9 bytes/s) into out[0..2] = read, write, copy.
// Returns 0, or a HIP error code.
extern "C" int eqf_bwprobe_run(int device, uint64_t bytes, int blocks_per_cu, int reps, double* out) {
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return (int)e;
  int cus = 0;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  if (e != hipSuccess) return (int)e;
  bytes &= ~(uint64_t)4095;
  void* buf = nullptr;
  float* sink = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  e = hipMalloc(&buf, bytes);
  if (e == hipSuccess) e = hipMalloc(&sink, 256);
  if (e == hipSuccess) e = hipMemset(buf, 0, bytes);
  if (e == hipSuccess) e = hipEventCreate(&t0);
  if (e == hipSuccess) e = hipEventCreate(&t1);
  const dim3 grid((unsigned)(cus * blocks_per_cu)), blk(256);
  const int64_t n4 = (int64_t)(bytes / 16), h4 = n4 / 2;
  f4* a = (f4*)buf;
  for (int which = 0; which < 3 && e == hipSuccess; ++which) {
    float ms = 0.f;
    for (int r = 0; r <= reps && e == hipSuccess; ++r) {
      if (r == 1) e = hipEventRecord(t0, 0);
      if (which == 0) hipLaunchKernelGGL(k_read<8>, grid, blk, 0, 0, a, n4, sink);
      else if (which == 1) hipLaunchKernelGGL(k_write<8>, grid, blk, 0, 0, a, n4, 0.0f);
      else hipLaunchKernelGGL(k_copy<8>, grid, blk, 0, 0, a, a + h4, h4);
      if (e == hipSuccess) e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(t1, 0);
    if (e == hipSuccess) e = hipEventSynchronize(t1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
    const double moved = which == 2 ? 2.0 * 16.0 * (double)h4 : 16.0 * (double)n4;
    if (e == hipSuccess) out[which] = moved * reps / (ms * 1e-3) / 1e9;
  }
  if (t0) (void)hipEventDestroy(t0);
  if (t1) (void)hipEventDestroy(t1);
  (void)hipFree(sink);
  (void)hipFree(buf);
  return (int)e;
}
