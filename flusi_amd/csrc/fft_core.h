// Device-side building blocks shared by every transform pass: complex helpers,
// radix-2/3/4/8 butterflies, radix plans and Stockham stages over lines held
// in LDS.
//
// LDS addressing is padded: element a of a line sits at a + (a >> PS) (one
// spare slot per 16 complex<float> / 8 complex<double>, i.e. per bank row),
// and lines are LSP slots apart with LSP = 1 modulo a bank row, so that both
// the stride-R writes of early Stockham stages and the line-strided accesses
// of the transposing passes avoid bank conflicts.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace flusi {

template <typename T> struct V2;
template <> struct V2<float> { using t = float2; };
template <> struct V2<double> { using t = double2; };
template <typename T> using cplx = typename V2<T>::t;

template <typename T> __device__ __forceinline__ cplx<T> mk(T x, T y) { cplx<T> r; r.x = x; r.y = y; return r; }
// Maximum that keeps a NaN: fmax drops a NaN operand, while the reference's
// maxval over a field with NaN entries is NaN, which adjust_dt turns into
// abort 100011 (fluid_time_step.f90:947-950).  A NaN operand wins.
__device__ __forceinline__ double nanmax(double a, double b) { return (b > a || b != b) ? b : a; }
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

// ------------------------------------------------------------ global memory
// Pointers in the global address space.  A device pointer that reaches a
// kernel through an argument struct (with batch-member offsets applied) stays
// generic to the compiler, and generic ("flat") accesses count against the
// LDS wait counter as well as the vector-memory one: every LDS wait of an
// in-wave transform would then also wait for the pass's outstanding loads and
// stores.  Accesses through these types are global loads and stores.
template <typename X> using gptr = X __attribute__((address_space(1)))*;
template <typename T> struct GWord { using t = T __attribute__((ext_vector_type(2))); };
template <> struct GWord<float> { using t = double; };
// complex loads and stores (the vector types' own assignment does not bind
// across address spaces)
template <typename T> __device__ __forceinline__ cplx<T> gld(gptr<const cplx<T>> p) {
  return __builtin_bit_cast(cplx<T>, *reinterpret_cast<gptr<const typename GWord<T>::t>>(p));
}
template <typename T> __device__ __forceinline__ void gst(gptr<cplx<T>> p, cplx<T> v) {
  *reinterpret_cast<gptr<typename GWord<T>::t>>(p) = __builtin_bit_cast(typename GWord<T>::t, v);
}
// twiddle i from an LDS (or generic) table or from global memory
template <typename T> __device__ __forceinline__ cplx<T> twv(const cplx<T>* __restrict__ p, int i) { return p[i]; }
template <typename T> __device__ __forceinline__ cplx<T> twv(gptr<const cplx<T>> p, int i) { return gld<T>(p + i); }

// ---------------------------------------------------------------- LDS padding
template <typename T> struct Pad { static constexpr int PS = (sizeof(T) == 4) ? 4 : 3; };
template <typename T> __host__ __device__ constexpr int lp(int a) { return a + (a >> Pad<T>::PS); }
// slots of one line holding n elements, rounded up to 1 modulo a bank row
template <typename T> __host__ __device__ constexpr int line_stride(int n) {
  return (lp<T>(n - 1) + 1) + ((((1 - (lp<T>(n - 1) + 1)) % (1 << Pad<T>::PS)) + (1 << Pad<T>::PS)) % (1 << Pad<T>::PS));
}

// ---------------------------------------------------------------- radix plans
// PLAN_TILE: 3 first, then 8s, then one 4 or 2 (y/z passes, L >= 16).
// PLAN_XF:   4 first, then 3, then 8s, then one 4 or 2 (x pass forward).
// PLAN_XI:   PLAN_XF reversed (x pass inverse), so that the last inverse
//            stage and the first forward stage share their index sets.
enum { PLAN_TILE = 0, PLAN_XF = 1, PLAN_XI = 2 };

__host__ __device__ constexpr int plan_list(int kind, int L, int s, bool count) {
  int n = 0, m = L;
  if (kind == PLAN_XF && m % 4 == 0 && m > 4) { if (!count && n == s) return 4; m /= 4; ++n; }
  if (m % 3 == 0) { if (!count && n == s) return 3; m /= 3; ++n; }
  while (m % 8 == 0 && m > 1) { if (!count && n == s) return 8; m /= 8; ++n; }
  if (m == 4 || m == 2) { if (!count && n == s) return m; ++n; m = 1; }
  return count ? (m == 1 ? n : -1) : 1;
}
__host__ __device__ constexpr int plan_count_k(int kind, int L) {
  return plan_list(kind == PLAN_XI ? PLAN_XF : kind, L, 0, true);
}
__host__ __device__ constexpr int plan_radix_k(int kind, int L, int s) {
  return kind == PLAN_XI ? plan_list(PLAN_XF, L, plan_count_k(PLAN_XF, L) - 1 - s, false)
                         : plan_list(kind, L, s, false);
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

// Stockham stage of radix R: butterfly j (0 <= j < L/R) reads positions
// j + q L/R, multiplies input q by W_{Ns R}^{q k} (k = j mod Ns), applies the
// radix-R DFT and writes positions (j/Ns) Ns R + k + q Ns.  tw[m] = exp(-2 pi
// i m / L); the inverse (S = +1) uses its conjugate.
template <typename T, int L, int R, int S, int Ns, typename TWP>
__device__ __forceinline__ void twiddle_bfly(cplx<T>* v, int j, TWP tw) {
  if (Ns > 1) {
    // (unsigned: j mod Ns without the sign handling of int %)
    const int step = (int)((unsigned)j % Ns) * (L / (Ns * R));
#pragma unroll
    for (int q = 1; q < R; ++q) {
      const cplx<T> w = twv<T>(tw, q * step);
      v[q] = (S < 0) ? cmul(v[q], w) : conjmul(v[q], w);
    }
  }
  Bfly<R, S, T>::run(v);
}
template <int R, int Ns> __device__ __forceinline__ int stage_out(int j, int q) {
  const unsigned u = (unsigned)j;
  return (int)(u / Ns) * Ns * R + (int)(u % Ns) + q * Ns;
}

// One Stockham stage over NL lines (line i at i*LSP) held in LDS.  tid is
// the thread's index in the block; a caller that runs several transforms in
// one kernel can pass a copy the compiler cannot see through, so that the
// addresses and twiddles derived from it are recomputed per transform rather
// than kept live (in registers, or spilled) from the first to the last.
template <typename T, int L, int NL, int NT, int R, int S, int Ns, int LSP, typename TWP>
__device__ __forceinline__ void lds_stage(cplx<T>* buf, TWP tw, int tid) {
  constexpr int NB = NL * (L / R);
  constexpr int PER = (NB + NT - 1) / NT;
  cplx<T> v[PER][R];
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int b = tid + p * NT;
    if (NB % NT == 0 || b < NB) {
      const int line = (int)((unsigned)b / (L / R)), j = b - line * (L / R);
      const cplx<T>* src = buf + line * LSP;
#pragma unroll
      for (int q = 0; q < R; ++q) v[p][q] = src[lp<T>(j + q * (L / R))];
      twiddle_bfly<T, L, R, S, Ns>(v[p], j, tw);
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int b = tid + p * NT;
    if (NB % NT == 0 || b < NB) {
      const int line = (int)((unsigned)b / (L / R)), j = b - line * (L / R);
      cplx<T>* dst = buf + line * LSP;
#pragma unroll
      for (int q = 0; q < R; ++q) dst[lp<T>(stage_out<R, Ns>(j, q))] = v[p][q];
    }
  }
  __syncthreads();
}

// Stages [STAGE, LAST) of plan KIND; NS is the sub-transform length entering STAGE.
template <typename T, int L, int NL, int NT, int KIND, int S, int LSP, int STAGE, int LAST, int NS,
          bool DONE = (STAGE >= LAST)>
struct Stages;
template <typename T, int L, int NL, int NT, int KIND, int S, int LSP, int STAGE, int LAST, int NS>
struct Stages<T, L, NL, NT, KIND, S, LSP, STAGE, LAST, NS, false> {
  template <typename TWP>
  static __device__ __forceinline__ void run(cplx<T>* buf, TWP tw, int tid = threadIdx.x) {
    constexpr int R = plan_radix_k(KIND, L, STAGE);
    lds_stage<T, L, NL, NT, R, S, NS, LSP>(buf, tw, tid);
    Stages<T, L, NL, NT, KIND, S, LSP, STAGE + 1, LAST, NS * R>::run(buf, tw, tid);
  }
};
template <typename T, int L, int NL, int NT, int KIND, int S, int LSP, int STAGE, int LAST, int NS>
struct Stages<T, L, NL, NT, KIND, S, LSP, STAGE, LAST, NS, true> {
  template <typename TWP> static __device__ __forceinline__ void run(cplx<T>*, TWP, int = 0) {}
};

// product of the radices of stages [0, s) of a plan
__host__ __device__ constexpr int plan_prefix(int kind, int L, int s) {
  int p = 1;
  for (int i = 0; i < s; ++i) p *= plan_radix_k(kind, L, i);
  return p;
}

}  // namespace flusi
