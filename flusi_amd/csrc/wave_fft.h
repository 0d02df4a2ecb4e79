// In-wave line transforms shared by the y/z tile passes (tiles_wave.hip) and
// the z-forward step pass (zstep.hip).
//
// A line of L points (L = 64, 128, 256, 512) lives in the registers of
// G = L/8 lanes, lane l holding positions l + qG (q = 0..7) before and after
// the transform.  Radix-8/4 Stockham stages exchange through the line's own
// LDS row with wave-scope ordering only.
#pragma once
#include "fft_core.h"

namespace flusi {
namespace wavefft {

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// radix plan with 8 elements per lane: 512 = 8.8.8, 256 = 8.8.4, 128 = 8.4.4, 64 = 8.8
template <int L> struct TWPlan {
  static_assert(L == 64 || L == 128 || L == 256 || L == 512, "in-wave line length");
  static constexpr int G = L / 8;      // lanes per line
  static constexpr int LPW = 64 / G;   // lines per wave
  static constexpr int NST = (L == 64) ? 2 : 3;
  static constexpr __host__ __device__ int radix(int s) {
    return L == 256 ? (s == 2 ? 4 : 8) : (L == 128 ? (s == 0 ? 8 : 4) : 8);
  }
  static constexpr __host__ __device__ int ns(int s) {
    return s == 0 ? 1 : (s == 1 ? radix(0) : radix(0) * radix(1));
  }
};

template <int S, typename C> __device__ __forceinline__ C twm(C v, C w) {
  return (S < 0) ? cmul(v, w) : conjmul(v, w);
}

// Stage s: B = 8/R butterflies per lane, j = l + bG; input q of butterfly b
// is register b + qB (position j + q L/R); outputs go to
// (j/Ns) Ns R + j%Ns + q Ns, which for the last stage is again register b + qB.
template <typename T, int L, int S, int s, typename TWP>
__device__ __forceinline__ void twstage(cplx<T> (&v)[8], TWP tw, cplx<T>* ln, int l) {
  constexpr int G = TWPlan<L>::G;
  constexpr int R = TWPlan<L>::radix(s), Ns = TWPlan<L>::ns(s), B = 8 / R;
  constexpr bool last = (s == TWPlan<L>::NST - 1);
#pragma unroll
  for (int b = 0; b < B; ++b) {
    // unsigned: j % Ns and j / Ns become a mask and a shift (l is an opaque
    // copy whose sign the compiler cannot see)
    const unsigned j = (unsigned)l + b * G;
    cplx<T> x[R];
#pragma unroll
    for (int q = 0; q < R; ++q) x[q] = v[b + q * B];
    if (s > 0) {
      const int step = (int)(j % Ns) * (L / (Ns * R));
#pragma unroll
      for (int q = 1; q < R; ++q) x[q] = twm<S>(x[q], twv<T>(tw, q * step));
    }
    Bfly<R, S, T>::run(x);
    if (last) {
#pragma unroll
      for (int q = 0; q < R; ++q) v[b + q * B] = x[q];
    } else {
      const int base = (int)(j / Ns) * Ns * R + (int)(j % Ns);
#pragma unroll
      for (int q = 0; q < R; ++q) ln[lp<T>(base + q * Ns)] = x[q];
    }
  }
  if (!last) {
    wave_sync();
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = ln[lp<T>(l + r * G)];
    wave_sync();
  }
}

template <typename T, int L, int S, typename TWP>
__device__ __forceinline__ void twfft(cplx<T> (&v)[8], TWP tw, cplx<T>* ln, int l_) {
  int l = l_;
  asm volatile("" : "+v"(l));  // recompute addresses per transform (register pressure)
  twstage<T, L, S, 0>(v, tw, ln, l);
  twstage<T, L, S, 1>(v, tw, ln, l);
  if constexpr (TWPlan<L>::NST > 2) twstage<T, L, S, 2>(v, tw, ln, l);
}

}  // namespace wavefft
}  // namespace flusi
