// Device-side building blocks shared by every transform pass: complex helpers,
// radix-2/3/4/8 butterflies and an in-LDS Stockham FFT over a tile of lines.
//
// A tile holds NL lines of length L in LDS, line i starting at i*LS.  All
// threads of the workgroup cooperate: each Stockham stage loads the radix-r
// butterflies a thread owns into registers, synchronises, and stores them in
// auto-sorted order, so the transform runs in place in one LDS buffer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace flusi {

template <typename T> struct V2;
template <> struct V2<float> { using t = float2; };
template <> struct V2<double> { using t = double2; };
template <typename T> using cplx = typename V2<T>::t;

template <typename T> __device__ __forceinline__ cplx<T> mk(T x, T y) { cplx<T> r; r.x = x; r.y = y; return r; }
__device__ __forceinline__ float2 add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  C r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r;
}
template <typename C> __device__ __forceinline__ C conjmul(C a, C b) {  // a * conj(b)
  C r; r.x = a.x * b.x + a.y * b.y; r.y = a.y * b.x - a.x * b.y; return r;
}
// multiply by s*i (s = +1 or -1)
template <int S, typename C> __device__ __forceinline__ C mul_si(C a) {
  C r; if (S > 0) { r.x = -a.y; r.y = a.x; } else { r.x = a.y; r.y = -a.x; } return r;
}
template <typename C, typename R> __device__ __forceinline__ C scl(C a, R s) { C r; r.x = a.x * s; r.y = a.y * s; return r; }

// ---------------------------------------------------------------- radix plans
// L = product of radices (3 first, then 8s, then one 4 or 2).  Supported
// lengths are 2^a * 3^b with b <= 1 and L >= 2.
__host__ __device__ constexpr int plan_count(int L) {
  int n = 0;
  if (L % 3 == 0) { L /= 3; ++n; }
  while (L % 8 == 0 && L > 1) { L /= 8; ++n; }
  if (L == 4 || L == 2) ++n;
  return n;
}
__host__ __device__ constexpr int plan_radix(int L, int s) {
  int n = 0;
  if (L % 3 == 0) { if (n == s) return 3; L /= 3; ++n; }
  while (L % 8 == 0 && L > 1) { if (n == s) return 8; L /= 8; ++n; }
  if (L == 4) return 4;
  if (L == 2) return 2;
  return 1;
}
__host__ __device__ constexpr bool plan_ok(int L) {
  int m = L;
  if (m % 3 == 0) m /= 3;
  while (m % 2 == 0) m /= 2;
  return m == 1 && L >= 2;
}

// ---------------------------------------------------------------- butterflies
template <int R, int S, typename T> struct Bfly;

template <int S, typename T> struct Bfly<2, S, T> {
  static __device__ __forceinline__ void run(cplx<T>* a) {
    cplx<T> t = a[0]; a[0] = add(t, a[1]); a[1] = sub(t, a[1]);
  }
};
template <int S, typename T> struct Bfly<4, S, T> {
  static __device__ __forceinline__ void run(cplx<T>* a) {
    cplx<T> s02 = add(a[0], a[2]), d02 = sub(a[0], a[2]);
    cplx<T> s13 = add(a[1], a[3]), d13 = mul_si<S>(sub(a[1], a[3]));
    a[0] = add(s02, s13); a[2] = sub(s02, s13);
    a[1] = add(d02, d13); a[3] = sub(d02, d13);
  }
};
template <int S, typename T> struct Bfly<3, S, T> {
  static __device__ __forceinline__ void run(cplx<T>* a) {
    const T h = T(0.8660254037844386467637231707529362);
    cplx<T> t1 = add(a[1], a[2]);
    cplx<T> t2 = mul_si<S>(scl(sub(a[1], a[2]), h));
    cplx<T> m = sub(a[0], scl(t1, T(0.5)));
    a[0] = add(a[0], t1); a[1] = add(m, t2); a[2] = sub(m, t2);
  }
};
template <int S, typename T> struct Bfly<8, S, T> {
  static __device__ __forceinline__ void run(cplx<T>* a) {
    cplx<T> e[4] = {a[0], a[2], a[4], a[6]};
    cplx<T> o[4] = {a[1], a[3], a[5], a[7]};
    Bfly<4, S, T>::run(e);
    Bfly<4, S, T>::run(o);
    const T r = T(0.7071067811865475244008443621048490);
    // W8^1 = r(1 + S i), W8^2 = S i, W8^3 = r(-1 + S i)
    cplx<T> o1 = scl(add(o[1], mul_si<S>(o[1])), r);
    cplx<T> o2 = mul_si<S>(o[2]);
    cplx<T> o3 = scl(sub(mul_si<S>(o[3]), o[3]), r);
    a[0] = add(e[0], o[0]); a[4] = sub(e[0], o[0]);
    a[1] = add(e[1], o1);   a[5] = sub(e[1], o1);
    a[2] = add(e[2], o2);   a[6] = sub(e[2], o2);
    a[3] = add(e[3], o3);   a[7] = sub(e[3], o3);
  }
};

// ------------------------------------------------------------ Stockham stage
// One stage of radix R over NL lines of length L held in LDS.  Ns is the
// length of the already-transformed sub-sequences.  tw[m] = exp(-2 pi i m/L);
// the inverse (S=+1) uses its conjugate.
template <typename T, int L, int NL, int NT, int R, int S, int Ns>
__device__ __forceinline__ void stockham_stage(cplx<T>* buf, int LS, const cplx<T>* __restrict__ tw) {
  constexpr int NB = NL * (L / R);                 // butterflies in the tile
  constexpr int PER = (NB + NT - 1) / NT;          // per thread (upper bound)
  cplx<T> v[PER][R];
  const int tid = threadIdx.x;
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int b = tid + p * NT;
    if (NB % NT == 0 || b < NB) {
      const int line = b / (L / R);
      const int j = b - line * (L / R);
      const int k = j % Ns;
      cplx<T>* src = buf + line * LS + j;
#pragma unroll
      for (int q = 0; q < R; ++q) v[p][q] = src[q * (L / R)];
      if (Ns > 1) {
        const int step = k * (L / (Ns * R));
#pragma unroll
        for (int q = 1; q < R; ++q) {
          cplx<T> w = tw[q * step];
          v[p][q] = (S < 0) ? cmul(v[p][q], w) : conjmul(v[p][q], w);
        }
      }
      Bfly<R, S, T>::run(v[p]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int b = tid + p * NT;
    if (NB % NT == 0 || b < NB) {
      const int line = b / (L / R);
      const int j = b - line * (L / R);
      const int k = j % Ns;
      cplx<T>* dst = buf + line * LS + (j / Ns) * Ns * R + k;
#pragma unroll
      for (int q = 0; q < R; ++q) dst[q * Ns] = v[p][q];
    }
  }
  __syncthreads();
}

template <typename T, int L, int NL, int NT, int S, int STAGE, int NS, bool DONE = (STAGE >= plan_count(L))>
struct StockhamRun;
template <typename T, int L, int NL, int NT, int S, int STAGE, int NS>
struct StockhamRun<T, L, NL, NT, S, STAGE, NS, false> {
  static __device__ __forceinline__ void run(cplx<T>* buf, int LS, const cplx<T>* tw) {
    constexpr int R = plan_radix(L, STAGE);
    stockham_stage<T, L, NL, NT, R, S, NS>(buf, LS, tw);
    StockhamRun<T, L, NL, NT, S, STAGE + 1, NS * R>::run(buf, LS, tw);
  }
};
template <typename T, int L, int NL, int NT, int S, int STAGE, int NS>
struct StockhamRun<T, L, NL, NT, S, STAGE, NS, true> {
  static __device__ __forceinline__ void run(cplx<T>*, int, const cplx<T>*) {}
};

// Full in-place FFT of NL lines in LDS (caller synchronises before the call;
// the call ends with a barrier).  S = -1 forward (exp(-i...)), S = +1 inverse;
// both unnormalised.
template <typename T, int L, int NL, int NT, int S>
__device__ __forceinline__ void block_fft(cplx<T>* buf, int LS, const cplx<T>* __restrict__ tw) {
  static_assert(plan_ok(L), "unsupported FFT length");
  StockhamRun<T, L, NL, NT, S, 0, 1>::run(buf, LS, tw);
}

}  // namespace flusi
