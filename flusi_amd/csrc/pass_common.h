// Helpers shared by the transform passes (passes.hip, zstep.hip).
#pragma once
#include "fft_core.h"
#include <stdexcept>
#include <string>

namespace flusi {

template <typename T> struct Tile { static constexpr int TL = 16; };
template <> struct Tile<double> { static constexpr int TL = 8; };

// threads of a tile block: one per radix-8 butterfly of the tile, 64..1024
__host__ __device__ constexpr int pick_nt(int n) {
  return (n / 8) >= 1024 ? 1024 : ((n / 8) <= 64 ? 64 : ((n / 8 + 63) / 64) * 64);
}

// Streamed complex elements (each read or written once per launch):
// non-temporal accesses keep them from crowding the caches.
template <typename T> __device__ __forceinline__ cplx<T> lds(const cplx<T>* p) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(cplx<T>, __builtin_nontemporal_load(reinterpret_cast<const double*>(p)));
  } else {
    using V = long long __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(cplx<T>, __builtin_nontemporal_load(reinterpret_cast<const V*>(p)));
  }
}
template <typename T> __device__ __forceinline__ void sts(cplx<T>* p, cplx<T> v) {
  if constexpr (sizeof(T) == 4) {
    __builtin_nontemporal_store(__builtin_bit_cast(double, v), reinterpret_cast<double*>(p));
  } else {
    using V = long long __attribute__((ext_vector_type(2)));
    __builtin_nontemporal_store(__builtin_bit_cast(V, v), reinterpret_cast<V*>(p));
  }
}

// the same through global-address-space pointers (see gptr)
template <typename T> __device__ __forceinline__ cplx<T> lds(gptr<const cplx<T>> p) {
  using W = typename GWord<T>::t;
  return __builtin_bit_cast(cplx<T>, __builtin_nontemporal_load(reinterpret_cast<gptr<const W>>(p)));
}
template <typename T> __device__ __forceinline__ void sts(gptr<cplx<T>> p, cplx<T> v) {
  using W = typename GWord<T>::t;
  __builtin_nontemporal_store(__builtin_bit_cast(W, v), reinterpret_cast<gptr<W>>(p));
}

template <typename C> __device__ __forceinline__ C conj_(C a) { a.y = -a.y; return a; }
template <typename T> __device__ __forceinline__ cplx<T> ik_(T k, cplx<T> v) { return mk<T>(-k * v.y, k * v.x); }

template <typename KF> static void set_smem(KF kernel, size_t bytes) {
  if (bytes > 65536) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) throw std::runtime_error(std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
  }
}

// -------------------------------------------------------------- wave/block sums
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = nanmax(v, __shfl_xor(v, o, 64));
  return v;
}

}  // namespace flusi
