// HBM stream probe: what plain copy, read-only and write-only streams reach
// on this device, and a copy whose reads come in 128-byte segments from a
// strided side (the tile passes' pattern): the practical ceilings the passes'
// achieved bandwidth is compared with beside the nominal peak.  Measurement
// support for bench.py; not part of the solver's C ABI (include/flusi_hip.h).
//
// The stream kernels move U 16-byte elements per thread (each step of u
// covering the block's next blockDim.x elements) with non-temporal accesses,
// over a grid that covers the buffer once (every index checked against n).
// The fastest shape differs per stream kind (measured on MI355X: copy 6.8
// TB/s with 64-thread blocks and U = 1, where a grid-stride loop over 65536
// blocks reached 5.2; read-only 7.1 and write-only 5.9 TB/s with 256-thread
// blocks and U = 4 to 8, where U = 1 in 64-thread blocks reaches 4.9), so
// each kind is timed in the shapes of SHAPES and its best rate reported.
#include <algorithm>
#include <hip/hip_runtime.h>

namespace {

using V4 = float __attribute__((ext_vector_type(4)));
using V2 = float __attribute__((ext_vector_type(2)));

template <int U> __global__ __launch_bounds__(256) void probe_copy_k(const V4* __restrict__ a, V4* __restrict__ b, long n) {
  const long base = (long)blockIdx.x * blockDim.x * U + threadIdx.x;
  V4 v[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const long i = base + (long)u * blockDim.x;
    if (i < n) v[u] = __builtin_nontemporal_load(a + i);
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const long i = base + (long)u * blockDim.x;
    if (i < n) __builtin_nontemporal_store(v[u], b + i);
  }
}

template <int U> __global__ __launch_bounds__(256) void probe_read_k(const V4* __restrict__ a, float* __restrict__ sink, long n) {
  const long base = (long)blockIdx.x * blockDim.x * U + threadIdx.x;
  float s = 0.f;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const long i = base + (long)u * blockDim.x;
    if (i < n) {
      const V4 v = __builtin_nontemporal_load(a + i);
      s += (v.x + v.y) + (v.z + v.w);
    }
  }
  if (s == -1.0f) sink[0] = s;  // never true for the zero-filled input; keeps the loads
}

template <int U> __global__ __launch_bounds__(256) void probe_write_k(V4* __restrict__ b, long n) {
  const long base = (long)blockIdx.x * blockDim.x * U + threadIdx.x;
  const V4 v = {1.f, 2.f, 3.f, 4.f};
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const long i = base + (long)u * blockDim.x;
    if (i < n) __builtin_nontemporal_store(v, b + i);
  }
}

// (threads per block, elements per thread) of the stream kernels
struct Shape { int t, u; };
constexpr Shape SHAPES[] = {{64, 1}, {256, 1}, {256, 4}, {256, 8}};

template <int U> void launch_stream(int kind, int t, V4* a, V4* b, float* sink, long n) {
  const dim3 grid((unsigned)((n + (long)t * U - 1) / ((long)t * U))), block(t);
  if (kind == 0) hipLaunchKernelGGL(probe_copy_k<U>, grid, block, 0, 0, a, b, n);
  else if (kind == 1) hipLaunchKernelGGL(probe_read_k<U>, grid, block, 0, 0, a, sink, n);
  else hipLaunchKernelGGL(probe_write_k<U>, grid, block, 0, 0, b, n);
}
void launch_stream(int kind, Shape sh, V4* a, V4* b, float* sink, long n) {
  if (sh.u == 1) launch_stream<1>(kind, sh.t, a, b, sink, n);
  else if (sh.u == 4) launch_stream<4>(kind, sh.t, a, b, sink, n);
  else launch_stream<8>(kind, sh.t, a, b, sink, n);
}

// Segmented copy: a block of 256 threads reads a tile of 16 rows x 128 bytes
// (16 eight-byte elements per row; rows SEG_RS elements = 32 KB apart, tiles
// taken along a row first) and writes it contiguously.
constexpr long SEG_RS = 4096;
__global__ __launch_bounds__(256) void probe_segcopy_k(const V2* __restrict__ a, V2* __restrict__ b, long rows) {
  constexpr long cols = SEG_RS / 16;
  const long t = blockIdx.x;
  const long r0 = (t / cols) * 16, c0 = (t % cols) * 16;
  const int r = threadIdx.x / 16, c = threadIdx.x % 16;
  if (r0 + r < rows)
    __builtin_nontemporal_store(__builtin_nontemporal_load(a + (r0 + r) * SEG_RS + c0 + c), b + t * 256 + threadIdx.x);
}

}  // namespace

extern "C" {

// bytes: size of each of the two buffers (rounded down to 16 bytes); reps:
// timed launches of each kernel and shape after one untimed one.
// out[0..3]: copy (bytes read + written), read-only, write-only (each the
// best over SHAPES) and segmented-copy (bytes read + written) bandwidth in
// bytes/s.  Returns 0, or the HIP error code.
int flusi_probe_hbm(long bytes, int reps, double* out) {
  const long n = bytes / 16;
  const long seg_rows = (bytes / 8 / SEG_RS) / 16 * 16;  // whole 16-row tiles
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
  const dim3 seg_grid((unsigned)((seg_rows / 16) * (SEG_RS / 16))), seg_block(256);
  auto timed = [&](auto launch, double moved, double& best) {
    launch();
    e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipEventRecord(e0, 0);
    for (int r = 0; r < reps && e == hipSuccess; ++r) launch();
    if (e == hipSuccess) e = hipEventRecord(e1, 0);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess && ms > 0.f) best = std::max(best, moved * reps / (ms * 1e-3));
  };
  for (int k = 0; k < 4; ++k) out[k] = 0.0;
  for (int k = 0; k < 3 && e == hipSuccess; ++k)
    for (const Shape sh : SHAPES) {
      if (e != hipSuccess) break;
      timed([&] { launch_stream(k, sh, a, b, sink, n); }, (k == 0 ? 2.0 : 1.0) * (double)(n * 16), out[k]);
    }
  if (e == hipSuccess && seg_rows > 0)
    timed([&] {
      hipLaunchKernelGGL(probe_segcopy_k, seg_grid, seg_block, 0, 0, reinterpret_cast<const V2*>(a),
                         reinterpret_cast<V2*>(b), seg_rows);
    }, 2.0 * seg_rows * SEG_RS * 8, out[3]);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (sink) (void)hipFree(sink);
  return e == hipSuccess ? 0 : (int)e;
}

}  // extern "C"
