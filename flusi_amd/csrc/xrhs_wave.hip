// x pass of the right-hand side with one wavefront per x line (nx/2 = 256),
// two per wave (128), four (64), five (48) or eight (32).
//
// Each line is a length-M = nx/2 complex transform (the real transform of
// length nx through the even/odd split).  A line lives in the registers of
// G = M/4 lanes, lane l holding positions l + qG (q = 0..3) before and after
// every transform.  The Stockham stages of radix 4 (one radix-2 stage for
// M = 128 and 32; radix 3, 4, 4 for M = 48) exchange data through a small
// per-wave LDS slot with wave-scope ordering only: waves never wait for each
// other, so a compute unit keeps as many lines in flight as its register
// file holds.  For M = 48 the five lines use 60 lanes; the other four work
// on a spare LDS row and store nothing.
//
// Per line (rhs.f90:147-330 cal_nlk_fsi, nonlinear and penalisation part):
//   load the six spectra V_x, V_y, V_z, w_x, A = dz u_x, B = dy u_x
//   form w_y = A - i kx V_z, w_z = i kx V_y - B            (basic_operators.f90:185-215)
//   C2R of u and w                                         (cof_p3dfft.f90:65-83)
//   nl = u x w - mask/eta (u - us) [- w/eta_sponge in the sponge]
//   R2C of nl [and of the sponge term]                      (cof_p3dfft.f90:33-56)
//   per block: max |u|^2 and the compensated sums of mask/eta (u - us)
#include "fft_core.h"
#include "flusi_kernels.h"
#include <cstdlib>
#include <stdexcept>
#include <string>

namespace flusi {

namespace {

constexpr int XW_NT = 256;  // threads of a block: four independent waves

template <typename C> __device__ __forceinline__ C conj_w(C a) { a.y = -a.y; return a; }
template <typename T> __device__ __forceinline__ cplx<T> ik_w(T k, cplx<T> v) { return mk<T>(-k * v.y, k * v.x); }

// orders this wave's LDS writes before its following reads (and the reads
// before the next writes); no other wave takes part
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wmax(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = nanmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Radix plan of the in-wave transform: radix 4 throughout, the last stage of
// M = 128 and M = 32 is radix 2 (two butterflies per lane).
template <int M> struct WPlan {
  static_assert(M == 32 || M == 64 || M == 128 || M == 256, "in-wave transform: M = 32, 64, 128 or 256");
  static constexpr int G = M / 4;
  static constexpr int NST = (M <= 64) ? 3 : 4;
  static constexpr __host__ __device__ int radix(int s) { return ((M == 128 && s == 3) || (M == 32 && s == 2)) ? 2 : 4; }
  // sub-transform length entering stage s (only the last stage can be radix 2)
  static constexpr __host__ __device__ int ns(int s) { return 1 << (2 * s); }
};

// lp(b + k) = lp(b) + lp(k) for the padded LDS accesses of the length-M
// transform: every stage's write base b with offsets q Ns, the lane's read
// base l with offsets q G and, with MIRROR, the mirrored base G - l with
// offsets q G.  Where it holds, one address per lane and stage serves all
// four elements through the LDS instruction's immediate offset.
template <typename T, int M> constexpr bool split_ok(bool mirror) {
  if constexpr (M == 48) {
    return false;  // the radix-3 plan addresses its rows itself
  } else {
    constexpr int G = M / 4;
    for (int l = 0; l < G; ++l) {
      for (int q = 0; q < 4; ++q) {
        if (lp<T>(l + q * G) != lp<T>(l) + lp<T>(q * G)) return false;
        if (mirror && lp<T>(G - l + q * G) != lp<T>(G - l) + lp<T>(q * G)) return false;
      }
      for (int s = 0; s + 1 < WPlan<M>::NST; ++s) {
        if (WPlan<M>::radix(s) != 4) return false;
        const int Ns = WPlan<M>::ns(s), b = (l / Ns) * Ns * 4 + l % Ns;
        for (int q = 0; q < 4; ++q)
          if (lp<T>(b + q * Ns) != lp<T>(b) + lp<T>(q * Ns)) return false;
      }
    }
    return true;
  }
}

template <int S, typename C> __device__ __forceinline__ C twm(C v, C w) {
  return (S < 0) ? cmul(v, w) : conjmul(v, w);
}

// Stage s of the in-wave Stockham transform.  Inputs: v[q] = x[l + qG].
// Outputs of butterfly j go to (j / Ns) Ns R + j % Ns + q Ns; all but the
// last stage pass them through the wave's LDS line `ln` and read back
// positions l + qG.  The last stage's outputs already sit at l + qG.
// tw: the block's LDS copy of exp(-2 pi i m / M), m < M.
template <typename T, int M, int S, int s>
__device__ __forceinline__ void wstage(cplx<T> (&v)[4], const cplx<T>* tw, cplx<T>* ln, int l) {
  // index arithmetic on the unsigned lane index: l % Ns and l / Ns are then
  // a mask and a shift (on a signed int the compiler has to allow l < 0,
  // which costs four instructions each)
  const unsigned ul = (unsigned)l;
  constexpr int G = WPlan<M>::G;
  constexpr int R = WPlan<M>::radix(s), Ns = WPlan<M>::ns(s);
  constexpr bool last = (s == WPlan<M>::NST - 1);
  if (R == 4) {
    if (s > 0) {
      const int step = (int)(ul % Ns) * (M / (Ns * 4));
#pragma unroll
      for (int q = 1; q < 4; ++q) v[q] = twm<S>(v[q], tw[q * step]);
    }
    Bfly<4, S, T>::run(v);
    if (!last) {
      const int base = (int)(ul / Ns) * Ns * 4 + (int)(ul % Ns);
      if constexpr (split_ok<T, M>(false)) {
        cplx<T>* wr = ln + lp<T>(base);
#pragma unroll
        for (int q = 0; q < 4; ++q) wr[lp<T>(q * Ns)] = v[q];
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) ln[lp<T>(base + q * Ns)] = v[q];
      }
    }
  } else {
    // butterflies j = l on (v0, v2) and j = l + G on (v1, v3)
    if (s > 0) {
      v[2] = twm<S>(v[2], tw[(int)(ul % Ns) * (M / (Ns * 2))]);
      v[3] = twm<S>(v[3], tw[(int)((ul + G) % Ns) * (M / (Ns * 2))]);
    }
    cplx<T> a0[2] = {v[0], v[2]}, a1[2] = {v[1], v[3]};
    Bfly<2, S, T>::run(a0);
    Bfly<2, S, T>::run(a1);
    if (!last) {
      const int b0 = (int)(ul / Ns) * Ns * 2 + (int)(ul % Ns), b1 = (int)((ul + G) / Ns) * Ns * 2 + (int)((ul + G) % Ns);
      ln[lp<T>(b0)] = a0[0]; ln[lp<T>(b0 + Ns)] = a0[1];
      ln[lp<T>(b1)] = a1[0]; ln[lp<T>(b1 + Ns)] = a1[1];
    } else {
      v[0] = a0[0]; v[2] = a0[1]; v[1] = a1[0]; v[3] = a1[1];
    }
  }
  if (!last) {
    wave_sync();
    if constexpr (split_ok<T, M>(false)) {
      const cplx<T>* rd = ln + lp<T>(l);
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = rd[lp<T>(q * G)];
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = ln[lp<T>(l + q * G)];
    }
    wave_sync();
  }
}

template <typename T, int M, int S>
__device__ __forceinline__ void wfft(cplx<T> (&v)[4], const cplx<T>* tw, cplx<T>* ln, int l_) {
  // recompute the LDS addresses in every transform instead of keeping a
  // few dozen of them live across the whole kernel
  int l = l_;
  asm volatile("" : "+v"(l));
  wstage<T, M, S, 0>(v, tw, ln, l);
  wstage<T, M, S, 1>(v, tw, ln, l);
  wstage<T, M, S, 2>(v, tw, ln, l);
  if constexpr (WPlan<M>::NST > 3) wstage<T, M, S, 3>(v, tw, ln, l);
}

// M = 48 = 3 x 4 x 4 (nx = 96).  Stage 0 (radix 3) has 16 butterflies for
// the 12 lanes of a line, so every stage reads its butterflies' inputs from
// the line's LDS row: butterfly j of stage 0 reads j + 16q and writes 3j + q
// (lanes 0..3 also take butterflies 12..15); stage 1 (radix 4, Ns = 3) reads
// j + 12q and writes (j/3) 12 + j%3 + 3q; stage 2 (radix 4, Ns = 12) reads
// j + 12q and leaves its outputs at j + 12q, the lane's own positions.
template <typename T, int S>
__device__ __forceinline__ void wfft48(cplx<T> (&v)[4], const cplx<T>* tw, cplx<T>* ln, int l_) {
  int l = l_;
  asm volatile("" : "+v"(l));
#pragma unroll
  for (int q = 0; q < 4; ++q) ln[lp<T>(l + 12 * q)] = v[q];
  wave_sync();
  const bool two = l < 4;
  cplx<T> a[3], b[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    a[q] = ln[lp<T>(l + 16 * q)];
    b[q] = ln[lp<T>((two ? 12 + l : l) + 16 * q)];
  }
  wave_sync();
  Bfly<3, S, T>::run(a);
#pragma unroll
  for (int q = 0; q < 3; ++q) ln[lp<T>(3 * l + q)] = a[q];
  if (two) {
    Bfly<3, S, T>::run(b);
#pragma unroll
    for (int q = 0; q < 3; ++q) ln[lp<T>(3 * (12 + l) + q)] = b[q];
  }
  wave_sync();
#pragma unroll
  for (int q = 0; q < 4; ++q) v[q] = ln[lp<T>(l + 12 * q)];
  wave_sync();
  {
    const int step = (l % 3) * 4;
#pragma unroll
    for (int q = 1; q < 4; ++q) v[q] = twm<S>(v[q], tw[q * step]);
    Bfly<4, S, T>::run(v);
    const int base = (l / 3) * 12 + (l % 3);
#pragma unroll
    for (int q = 0; q < 4; ++q) ln[lp<T>(base + 3 * q)] = v[q];
  }
  wave_sync();
#pragma unroll
  for (int q = 0; q < 4; ++q) v[q] = ln[lp<T>(l + 12 * q)];
  wave_sync();
#pragma unroll
  for (int q = 1; q < 4; ++q) v[q] = twm<S>(v[q], tw[q * l]);
  Bfly<4, S, T>::run(v);
}

template <typename T, int M, int S>
__device__ __forceinline__ void wtransform(cplx<T> (&v)[4], const cplx<T>* tw, cplx<T>* ln, int l) {
  if constexpr (M == 48) wfft48<T, S>(v, tw, ln, l);
  else wfft<T, M, S>(v, tw, ln, l);
}

// WPE: minimum waves per SIMD the register allocation must allow (0: none)
template <typename T, int M, bool SPONGE, bool STATS, int WPE>
__global__ __launch_bounds__(XW_NT) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1))) void x_rhs_wave_k(XPassArgs a) {
  to_member(a, blockIdx.y);
  constexpr int G = M / 4, LPW = 64 / G, NXH = M + 1;
  constexpr int LS = line_stride<T>(M + 1);
  constexpr int WPB = XW_NT / 64;
  constexpr int NFO = SPONGE ? 6 : 3;
  constexpr int ROWS = LPW + (LPW * G < 64 ? 1 : 0);  // a spare row for lanes beyond LPW lines
  __shared__ cplx<T> sh[WPB][ROWS * LS];
  __shared__ cplx<T> twl[M], twn[NXH];
  __shared__ double red[WPB][4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int l = lane % G;
  const int gl = (blockIdx.x * WPB + wv) * LPW + lane / G;
  const bool live = lane < LPW * G && gl < a.nlines;
  cplx<T>* ln = sh[wv] + (lane / G) * LS;

  // global-address-space views of the arguments (see gptr)
  const gptr<const cplx<T>> twM = (gptr<const cplx<T>>)a.twM;
  const gptr<const cplx<T>> twN = (gptr<const cplx<T>>)a.twN;
  const gptr<const cplx<T>> in = (gptr<const cplx<T>>)a.in;
  const gptr<cplx<T>> out = (gptr<cplx<T>>)a.out;
  const gptr<const T> mask = (gptr<const T>)a.mask;
  const gptr<const T> us = (gptr<const T>)a.us;

  for (int i = threadIdx.x; i < M; i += XW_NT) twl[i] = gld<T>(twM + i);
  for (int i = threadIdx.x; i < NXH; i += XW_NT) twn[i] = gld<T>(twN + i);

  // ---- loads: six spectra at k = l + qG and k = M
  cplx<T> X[6][4], XM[6];
  const long lb = (long)(live ? gl : 0) * a.prow;
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const gptr<const cplx<T>> p = in + f * a.fstride_in + lb;
#pragma unroll
    for (int q = 0; q < 4; ++q) X[f][q] = gld<T>(p + l + q * G);
    XM[f] = gld<T>(p + M);
  }

  // mask at the physical points 2n, 2n+1 (n = l + qG), read early; a line
  // whose flag says its mask is all zero (most lines of a flow around a
  // small body) reads the shared zero row instead, which stays in cache
  cplx<T> mk2[4];
  {
    const long ml = live ? gl : 0;
    const bool zero_line = a.mflag && ((gptr<const uint8_t>)a.mflag)[ml] == 0;
    const gptr<const T> mrow = zero_line ? (gptr<const T>)a.mzero : mask + ml * a.nx;
#pragma unroll
    for (int q = 0; q < 4; ++q) mk2[q] = gld<T>(reinterpret_cast<gptr<const cplx<T>>>(mrow + 2 * (l + q * G)));
  }

  __syncthreads();  // twiddle tables (the only block-wide wait besides the statistics)

  // ---- w_y = A - i kx V_z, w_z = i kx V_y - B
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const T kx = (T)(a.kx_scale * double(l + q * G));
    X[4][q] = sub(X[4][q], ik_w<T>(kx, X[2][q]));
    X[5][q] = sub(ik_w<T>(kx, X[1][q]), X[5][q]);
  }
  {
    const T kx = (T)(a.kx_scale * double(M));
    XM[4] = sub(XM[4], ik_w<T>(kx, XM[2]));
    XM[5] = sub(ik_w<T>(kx, XM[1]), XM[5]);
  }

  // ---- C2R of the six fields: z[n] = x[2n] + i x[2n+1] at n = l + qG
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    // one field at a time: keeps the scheduler from interleaving the
    // independent transforms (their temporaries would exhaust the registers)
    __builtin_amdgcn_sched_barrier(0);
    cplx<T> xb[4];
    if constexpr (split_ok<T, M>(true)) {
      // positions l + qG and their mirrors M - l - qG = (G - l) + (3 - q) G
      int lo = l;
      asm volatile("" : "+v"(lo));
      cplx<T>* rw = ln + lp<T>(lo);
      const cplx<T>* mr = ln + lp<T>(G - lo);
#pragma unroll
      for (int q = 0; q < 4; ++q) rw[lp<T>(q * G)] = X[f][q];
      if (lo == 0) ln[lp<T>(M)] = XM[f];
      wave_sync();
#pragma unroll
      for (int q = 0; q < 4; ++q) xb[q] = mr[lp<T>((3 - q) * G)];
      wave_sync();
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) ln[lp<T>(l + q * G)] = X[f][q];
      if (l == 0) ln[lp<T>(M)] = XM[f];
      wave_sync();
#pragma unroll
      for (int q = 0; q < 4; ++q) xb[q] = ln[lp<T>(M - l - q * G)];
      wave_sync();
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      cplx<T> xa = X[f][q], xm = xb[q];
      if (q == 0 && l == 0) { xa.y = T(0); xm.y = T(0); }
      const cplx<T> s1 = add(xa, conj_w(xm)), d1 = sub(xa, conj_w(xm));
      X[f][q] = add(s1, ik_w<T>(T(1), cmul(conj_w(twn[l + q * G]), d1)));
    }
    wtransform<T, M, +1>(X[f], twl, ln, l);
  }

  // ---- u x w - chi (u - us) [+ sponge]
  double umax2 = 0.0, fs[3] = {0.0, 0.0, 0.0}, fc[3] = {0.0, 0.0, 0.0};
  cplx<T> O[NFO][4];
  const int iy = live ? gl % a.ny : 0, iz = live ? gl / a.ny : 0;
  // A wave whose points all lie outside the mask (most waves of a flow
  // around a small body) has chi = 0 at every point, so that nl = u x w: it
  // skips the FP64 penalisation term (which is 0 there for finite u; only
  // the variant without statistics and sponge, whose other terms need every
  // point).  The skipped code also frees the registers for a fourth wave
  // per SIMD in the FP32 kernel.
  bool masked = false;
#pragma unroll
  for (int q = 0; q < 4; ++q) masked = masked || mk2[q].x != T(0) || mk2[q].y != T(0);
  if (!SPONGE && !STATS && !__any(masked)) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const cplx<T> u0 = X[0][q], u1 = X[1][q], u2 = X[2][q], w0 = X[3][q], w1 = X[4][q], w2 = X[5][q];
      O[0][q] = mk<T>(u1.x * w2.x - u2.x * w1.x, u1.y * w2.y - u2.y * w1.y);
      O[1][q] = mk<T>(u2.x * w0.x - u0.x * w2.x, u2.y * w0.y - u0.y * w2.y);
      O[2][q] = mk<T>(u0.x * w1.x - u1.x * w0.x, u0.y * w1.y - u1.y * w0.y);
    }
  } else
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    // one position at a time (FP64 temporaries of several would not fit)
    __builtin_amdgcn_sched_barrier(0);
    const int n = l + q * G;
    const long pt0 = (long)(live ? gl : 0) * a.nx + 2 * n;
    double nv[3][2], sv[3][2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      __builtin_amdgcn_sched_barrier(0);
      const double ux = h ? X[0][q].y : X[0][q].x, uy = h ? X[1][q].y : X[1][q].x, uz = h ? X[2][q].y : X[2][q].x;
      const double wx = h ? X[3][q].y : X[3][q].x, wy = h ? X[4][q].y : X[4][q].x, wz = h ? X[5][q].y : X[5][q].x;
      const long pt = pt0 + h;
      const double mkv = live ? (double)(h ? mk2[q].y : mk2[q].x) : 0.0;
      double dux = ux, duy = uy, duz = uz;
      if (mkv != 0.0) {
        dux -= (double)us[pt];
        duy -= (double)us[a.npts + pt];
        duz -= (double)us[2 * a.npts + pt];
      }
      const double chi = mkv * a.eps_inv;
      const double pv[3] = {chi * dux, chi * duy, chi * duz};
      if (STATS && live) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double t = fs[c] + pv[c];
          fc[c] += (fabs(fs[c]) >= fabs(pv[c])) ? (fs[c] - t) + pv[c] : (pv[c] - t) + fs[c];
          fs[c] = t;
        }
        umax2 = nanmax(umax2, ux * ux + uy * uy + uz * uz);
      }
      if constexpr (sizeof(T) == 4) {
        // FP32-mixed: u x w in FP32, the precision of the transforms that
        // produced u and w; the penalisation term stays FP64 (chi = mask/eta
        // is large at small eta).  Keeps the kernel at 3 waves per SIMD
        // without spilling.
        const float fx = h ? X[0][q].y : X[0][q].x, fy = h ? X[1][q].y : X[1][q].x, fz = h ? X[2][q].y : X[2][q].x;
        const float gx = h ? X[3][q].y : X[3][q].x, gy = h ? X[4][q].y : X[4][q].x, gz = h ? X[5][q].y : X[5][q].x;
        nv[0][h] = (double)(fy * gz - fz * gy) - pv[0];
        nv[1][h] = (double)(fz * gx - fx * gz) - pv[1];
        nv[2][h] = (double)(fx * gy - fy * gx) - pv[2];
      } else {
        nv[0][h] = uy * wz - uz * wy - pv[0];
        nv[1][h] = uz * wx - ux * wz - pv[1];
        nv[2][h] = ux * wy - uy * wx - pv[2];
      }
      if constexpr (SPONGE) {
        const int ix = 2 * n + h, th = a.sponge_thickness;
        bool in_sp = false;
        if (a.sponge == 1) {
          if (a.nx > 4 && (ix <= th - 1 || ix >= a.nx - th)) in_sp = true;
          if (iy <= th - 1 || iy >= a.ny - th) in_sp = true;
          if (iz <= th - 1 || iz >= a.nz - th) in_sp = true;
        } else if (a.sponge == 2) {
          in_sp = ix >= a.nx - th;
        }
        in_sp = in_sp && (mkv < a.sponge_mask_thr);
        sv[0][h] = in_sp ? -wx * a.sponge_eps_inv : 0.0;
        sv[1][h] = in_sp ? -wy * a.sponge_eps_inv : 0.0;
        sv[2][h] = in_sp ? -wz * a.sponge_eps_inv : 0.0;
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      O[c][q] = mk<T>((T)nv[c][0], (T)nv[c][1]);
      if constexpr (SPONGE) O[3 + c][q] = mk<T>((T)sv[c][0], (T)sv[c][1]);
    }
  }

  // ---- R2C of the outputs, split into the nx/2+1 entries
  const T half = T(0.5);
#pragma unroll
  for (int c = 0; c < NFO; ++c) {
    __builtin_amdgcn_sched_barrier(0);
    wtransform<T, M, -1>(O[c], twl, ln, l);
    cplx<T> zm[4];
    if constexpr (split_ok<T, M>(true)) {
      // mirrors M - n of n = l + qG, through (G - l) + (3 - q) G; n = 0 pairs with 0
      int lo = l;
      asm volatile("" : "+v"(lo));
      cplx<T>* rw = ln + lp<T>(lo);
      const cplx<T>* mr = ln + lp<T>(G - lo);
#pragma unroll
      for (int q = 0; q < 4; ++q) rw[lp<T>(q * G)] = O[c][q];
      wave_sync();
      zm[0] = (lo == 0 ? ln : mr + lp<T>(3 * G))[0];
#pragma unroll
      for (int q = 1; q < 4; ++q) zm[q] = mr[lp<T>((3 - q) * G)];
      wave_sync();
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) ln[lp<T>(l + q * G)] = O[c][q];
      wave_sync();
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int n = l + q * G;
        zm[q] = ln[lp<T>(n == 0 ? 0 : M - n)];
      }
      wave_sync();
    }
    if (live) {
      const gptr<cplx<T>> o = out + c * a.fstride_out + (long)gl * a.prow;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const cplx<T> zk = O[c][q];
        const cplx<T> s1 = add(zk, conj_w(zm[q])), d1 = sub(zk, conj_w(zm[q]));
        gst<T>(o + l + q * G, scl(sub(s1, ik_w<T>(T(1), cmul(twn[l + q * G], d1))), half));
      }
      if (l == 0) {
        // k = M: partner of k = 0 is itself
        const cplx<T> zk = O[c][0];
        const cplx<T> s2 = add(zk, conj_w(zk)), d2 = sub(zk, conj_w(zk));
        gst<T>(o + M, scl(sub(s2, ik_w<T>(T(1), cmul(twn[M], d2))), half));
      }
    }
  }

  if (STATS) {
    double v4[4] = {wmax(umax2), wsum(fs[0] + fc[0]), wsum(fs[1] + fc[1]), wsum(fs[2] + fc[2])};
    if (lane == 0) {
#pragma unroll
      for (int c = 0; c < 4; ++c) red[wv][c] = v4[c];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double r[4] = {red[0][0], red[0][1], red[0][2], red[0][3]};
      for (int w = 1; w < WPB; ++w) {
        r[0] = nanmax(r[0], red[w][0]);
        for (int c = 1; c < 4; ++c) r[c] += red[w][c];
      }
      for (int c = 0; c < 4; ++c) ((gptr<double>)a.partial)[(long)blockIdx.x * 4 + c] = r[c];
    }
  }
}

template <typename T, int M>
void run_wave(const XPassArgs& a, hipStream_t s) {
  constexpr int LPB = (XW_NT / 64) * (256 / M);  // lines per block
  const int blocks = (a.nlines + LPB - 1) / LPB;
  if (a.sponge) {
    if (a.want_stats) hipLaunchKernelGGL((x_rhs_wave_k<T, M, true, true, 0>), dim3(blocks, batch_n(a.bat)), dim3(XW_NT), 0, s, a);
    else hipLaunchKernelGGL((x_rhs_wave_k<T, M, true, false, 0>), dim3(blocks, batch_n(a.bat)), dim3(XW_NT), 0, s, a);
  } else if (a.want_stats) {
    if (a.wpe == 3) hipLaunchKernelGGL((x_rhs_wave_k<T, M, false, true, 3>), dim3(blocks, batch_n(a.bat)), dim3(XW_NT), 0, s, a);
    else hipLaunchKernelGGL((x_rhs_wave_k<T, M, false, true, 0>), dim3(blocks, batch_n(a.bat)), dim3(XW_NT), 0, s, a);
  } else {
    if (a.wpe == 3) hipLaunchKernelGGL((x_rhs_wave_k<T, M, false, false, 3>), dim3(blocks, batch_n(a.bat)), dim3(XW_NT), 0, s, a);
    else hipLaunchKernelGGL((x_rhs_wave_k<T, M, false, false, 0>), dim3(blocks, batch_n(a.bat)), dim3(XW_NT), 0, s, a);
  }
}

// One wavefront per x line of the region: the line's flag is 1 where any of
// its nx mask values is nonzero (NaN included).  The wavefront returns as a
// whole (w depends on the wave only).
template <typename T>
__global__ __launch_bounds__(256) void mask_flags_k(const T* __restrict__ mask, uint8_t* __restrict__ flag, int nx, int ny,
                                                    int y0, int ey, int z0, long nl) {
  const long w = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (w >= nl) return;
  const int lane = threadIdx.x & 63;
  const long line = (long)(z0 + (int)(w / ey)) * ny + (y0 + (int)(w % ey));
  const T* row = mask + line * nx;
  bool nzv = false;
  for (int i = lane; i < nx; i += 64) nzv = nzv || !(row[i] == T(0));
  const bool any = __any(nzv);
  if (lane == 0) flag[line] = any ? 1 : 0;
}

}  // namespace

template <typename T>
void launch_mask_flags(const void* mask, uint8_t* flag, int nx, int ny, int y0, int y1, int z0, int z1, hipStream_t s) {
  if (y1 < y0 || z1 < z0) return;
  const int ey = y1 - y0 + 1;
  const long nl = (long)ey * (z1 - z0 + 1);
  hipLaunchKernelGGL(mask_flags_k<T>, dim3((unsigned)((nl + 3) / 4)), dim3(256), 0, s, (const T*)mask, flag, nx, ny, y0,
                     ey, z0, nl);
}
template void launch_mask_flags<float>(const void*, uint8_t*, int, int, int, int, int, int, hipStream_t);
template void launch_mask_flags<double>(const void*, uint8_t*, int, int, int, int, int, int, hipStream_t);

bool x_rhs_wave_supported(int nx) { return nx == 64 || nx == 96 || nx == 128 || nx == 256 || nx == 512; }

int x_rhs_wave_blocks(int nx, int nlines) {
  const int lpb = (XW_NT / 64) * (256 / (nx / 2));
  return (nlines + lpb - 1) / lpb;
}

template <typename T>
void launch_x_rhs_wave(const XPassArgs& a, hipStream_t s) {
  switch (a.nx / 2) {
    case 32: run_wave<T, 32>(a, s); break;
    case 48: run_wave<T, 48>(a, s); break;
    case 64: run_wave<T, 64>(a, s); break;
    case 128: run_wave<T, 128>(a, s); break;
    case 256: run_wave<T, 256>(a, s); break;
    default: throw std::runtime_error("in-wave x pass: nx must be 64, 96, 128, 256 or 512");
  }
}
template void launch_x_rhs_wave<float>(const XPassArgs&, hipStream_t);
template void launch_x_rhs_wave<double>(const XPassArgs&, hipStream_t);

}  // namespace flusi
