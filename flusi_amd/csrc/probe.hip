// HBM stream probe: what plain copy, read-only and write-only streams reach
// on this device, the practical ceiling the passes' achieved bandwidth is
// compared with beside the nominal peak.  Measurement support for bench.py;
// not part of the solver's C ABI (include/flusi_hip.h).
//
// Each kernel streams 16-byte elements with non-temporal accesses in a
// grid-stride loop (every index checked against n), 65536 blocks of 256
// threads.
#include <hip/hip_runtime.h>

namespace {

using V4 = float __attribute__((ext_vector_type(4)));

__global__ void probe_copy_k(const V4* __restrict__ a, V4* __restrict__ b, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}

__global__ void probe_read_k(const V4* __restrict__ a, float* __restrict__ sink, long n) {
  float s = 0.f;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const V4 v = __builtin_nontemporal_load(a + i);
    s += (v.x + v.y) + (v.z + v.w);
  }
  if (s == -1.0f) sink[0] = s;  // never true for the zero-filled input; keeps the loads
}

__global__ void probe_write_k(V4* __restrict__ b, long n) {
  const V4 v = {1.f, 2.f, 3.f, 4.f};
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    __builtin_nontemporal_store(v, b + i);
}

}  // namespace

extern "C" {

// bytes: size of each of the two buffers (rounded down to 16 bytes); reps:
// timed launches of each kernel after one untimed one.  out[0..2]: copy
// (bytes read + written), read-only and write-only bandwidth in bytes/s.
// Returns 0, or the HIP error code.
int flusi_probe_hbm(long bytes, int reps, double* out) {
  const long n = bytes / 16;
  if (n <= 0 || reps <= 0 || !out) return 1;
  V4 *a = nullptr, *b = nullptr;
  float* sink = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&a, n * 16);
  if (e == hipSuccess) e = hipMalloc(&b, n * 16);
  if (e == hipSuccess) e = hipMalloc(&sink, 64);
  if (e == hipSuccess) e = hipMemset(a, 0, n * 16);
  if (e == hipSuccess) e = hipMemset(b, 0, n * 16);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  const dim3 grid(65536), block(256);
  for (int k = 0; k < 3 && e == hipSuccess; ++k) {
    auto launch = [&] {
      if (k == 0) hipLaunchKernelGGL(probe_copy_k, grid, block, 0, 0, a, b, n);
      else if (k == 1) hipLaunchKernelGGL(probe_read_k, grid, block, 0, 0, a, sink, n);
      else hipLaunchKernelGGL(probe_write_k, grid, block, 0, 0, b, n);
    };
    launch();
    e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipEventRecord(e0, 0);
    for (int r = 0; r < reps && e == hipSuccess; ++r) launch();
    if (e == hipSuccess) e = hipEventRecord(e1, 0);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess) out[k] = (k == 0 ? 2.0 : 1.0) * (double)(n * 16) * reps / (ms * 1e-3);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (sink) (void)hipFree(sink);
  return e == hipSuccess ? 0 : (int)e;
}

}  // extern "C"
