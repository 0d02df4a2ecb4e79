// z-forward pass with the spectral epilogue of the right-hand side and the
// stepper update (rhs.f90:147-330 tail, :441-468, :575-615; sponge.f90:22-83;
// fluid_time_step.f90:519-832; integrals.f90:522-552).
//
// One workgroup transforms TL (kx, ky) lines of length nz for each of the 3
// (6 with the sponge) Q-layout fields; the last Stockham stage leaves every
// thread with the same (ky, kz) positions of all fields, where the epilogue
// runs (pointwise terms in the field precision; dealias predicate, integrating
// factors and Parseval sum in FP64) and writes the S-layout state.
#include "fft_core.h"
#include "flusi_kernels.h"
#include "pass_common.h"
#include "wave_fft.h"
#include <cstdlib>
#include <type_traits>
#include <stdexcept>
#include <string>

namespace flusi {

__device__ __forceinline__ double vis_factor(double dtv, double nu, double k2, bool alias) {
  if (dtv < 0.0 || alias) return 0.0;
  return exp(-dtv * nu * k2);
}

// Stepper modes are grouped so that each kernel carries only the update code
// of its own stepper (the epilogue is unrolled over a thread's elements).
enum { MG_RHS = 0, MG_AB2 = 1, MG_RK2 = 2, MG_RK4 = 3 };
__host__ __device__ constexpr int mode_group(int mode) {
  return (mode == SM_RHS_ONLY || mode == SM_PRESSURE) ? MG_RHS
       : (mode == SM_EULER || mode == SM_AB2) ? MG_AB2
       : (mode == SM_RK2_1 || mode == SM_RK2_2) ? MG_RK2 : MG_RK4;
}
// MG = MG_FIXED + mode fixes the stepper mode itself at compile time, for the
// operands the steppers always pass with it (fixed_mode_ok): the epilogue
// then has no branch on the mode, the diffusion switch or the aliasing of its
// operands, and a position's state loads are issued together.  With the mode
// known only at run time, each of those loads sat under its own branch and
// was waited for before the next one was issued.
constexpr int MG_FIXED = 16;
__host__ __device__ constexpr int mg_group(int mg) { return mg >= MG_FIXED ? mode_group(mg - MG_FIXED) : mg; }
// the stepper mode of a launch of kernel group MG
template <int MG> __device__ __forceinline__ int mg_mode(const ZStepArgs& a) {
  if constexpr (MG >= MG_FIXED) return MG - MG_FIXED;
  else return mg_group(MG) == MG_RHS ? SM_RHS_ONLY : (MG == MG_RK4 && a.mode < SM_RK4_1 ? SM_RK4_1 : a.mode);
}
template <int MG> __device__ __forceinline__ bool mg_final_stage(const ZStepArgs& a) {
  if constexpr (mg_group(MG) == MG_RHS) {
    return false;
  } else {
    const int m = mg_mode<MG>(a);
    return m == SM_EULER || m == SM_AB2 || m == SM_RK2_2 || m == SM_RK4_4;
  }
}

// A batch member's slab is reached through index offsets rather than shifted
// pointers (the kernel argument pointers then stay rematerialisable, which
// keeps the register budget of the unbatched kernel): mc in complex elements,
// md in doubles.
struct MOff { long mc, md; };
template <typename T> __device__ __forceinline__ MOff member_off(const ZStepArgs& a) {
  const long b = (long)blockIdx.z * a.bat.mstride;
  return MOff{b / (long)sizeof(cplx<T>), b / (long)sizeof(double)};
}

// Operands of the epilogue at one position for a fixed-mode kernel, loaded
// one position ahead (zpreload), so that a position's loads are in flight
// while the previous position computes and before its stores are issued
// (the vector-memory counter is in order: a wait for loads issued after a
// store also waits for the store).  st[0] = the diffusion operand (RK4) or
// u_n; st[1] = acc (RK4 stages 2-4), n_{-1} (AB2) or n_0 (RK2 stage 2);
// st[2] = u_n (RK4 stage 3).
template <typename T> struct ZPre {
  cplx<T> st[3][3];
  double ky, kz, kyt2, kzt2;
};
template <typename T, int MG>
__device__ __forceinline__ void zpreload(const ZStepArgs& a, const MOff& mo, int kx, int ky, int kzi, ZPre<T>& pr) {
  constexpr int m = MG - MG_FIXED;
  const long si = ((long)kx * a.ny + ky) * a.nz + kzi + mo.mc;
  pr.ky = a.ky[ky + mo.md];
  pr.kz = a.kz[kzi + mo.md];
  pr.kyt2 = a.kyt2[ky + mo.md];
  pr.kzt2 = a.kzt2[kzi + mo.md];
  const cplx<T>* s0 = reinterpret_cast<const cplx<T>*>(m >= SM_RK4_1 ? a.udiff : a.u);
  const cplx<T>* s1 = reinterpret_cast<const cplx<T>*>(m >= SM_RK4_2 ? a.acc : (m == SM_AB2 ? a.r1 : a.r0));
  constexpr bool has1 = m >= SM_RK4_2 || m == SM_AB2 || m == SM_RK2_2;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    pr.st[0][c] = lds<T>(s0 + c * a.fs + si);
    if constexpr (has1) pr.st[1][c] = lds<T>(s1 + c * a.fs + si);
    if constexpr (m == SM_RK4_3) pr.st[2][c] = lds<T>(reinterpret_cast<const cplx<T>*>(a.u) + c * a.fs + si);
  }
}

// Epilogue at one spectral point (kx, ky, kz): rv[f] is field f of the
// z-transformed right-hand side (unscaled).  Sponge curl (sponge.f90:22-83),
// dynamic mean flow, explicit diffusion (rhs.f90:441-468), projection
// (rhs.f90:575-615), dealiasing and the stepper update
// (fluid_time_step.f90:519-832), the zero-mode overwrite and the Parseval
// partial sum (integrals.f90:522-552).
//
// Arithmetic type R = T: the FP64 path computes in FP64 throughout; the
// FP32-mixed path computes the pointwise terms in FP32, the precision of its
// stored fields and transforms.  In both, the dealiasing predicate is the
// reference's FP64 comparison (so that the same modes are removed), the
// integrating factors exp(-nu k^2 dt) are evaluated in FP64, and the
// Parseval sum is accumulated in compensated FP64.
//
// FOLDED: rv[f] already holds the scaled right-hand side with the sponge term
// added (sponge_fold below), so that a kernel need not keep the sponge
// fields' values live until the epilogue.
template <typename T, bool SPONGE, int MG, bool FOLDED = false>
__device__ __forceinline__ void zepilogue(const ZStepArgs& a, const MOff& mo, int kx, int ky, int kzi, const cplx<T>* rv,
                                          double kxv_d, double kxt2, double wE,
                                          bool final_stage, double& esum, double& ecomp,
                                          const ZPre<T>* pre = nullptr) {
  using R = T;
  cplx<T>* U = reinterpret_cast<cplx<T>*>(a.u);
  cplx<T>* UST = reinterpret_cast<cplx<T>*>(a.ustage);
  const cplx<T>* UD = reinterpret_cast<const cplx<T>*>(a.udiff);
  cplx<T>* ACC = reinterpret_cast<cplx<T>*>(a.acc);
  cplx<T>* RG0 = reinterpret_cast<cplx<T>*>(a.r0);
  const cplx<T>* RG1 = reinterpret_cast<const cplx<T>*>(a.r1);
  cplx<T>* RO = reinterpret_cast<cplx<T>*>(a.rhs_out);
  const long si = ((long)kx * a.ny + ky) * a.nz + kzi + mo.mc;
  constexpr int G = mg_group(MG);
  constexpr bool FIX = MG >= MG_FIXED;
  // operands: preloaded (fixed-mode kernels) or read here; slot as in ZPre
  auto opnd = [&](const void* base, int slot, int c) -> cplx<T> {
    if constexpr (FIX) return pre->st[slot][c];
    else return lds<T>(reinterpret_cast<const cplx<T>*>(base) + c * a.fs + si);
  };
  const double kyv_d = FIX ? pre->ky : a.ky[ky + mo.md], kzv_d = FIX ? pre->kz : a.kz[kzi + mo.md];
  const double k2d = (kxv_d * kxv_d + kyv_d * kyv_d) + kzv_d * kzv_d;  // FP64: integrating factors
  const R kxv = (R)kxv_d, kyv = (R)kyv_d, kzv = (R)kzv_d, k2 = (R)k2d;
  const bool alias = ((kxt2 + (FIX ? pre->kyt2 : a.kyt2[ky + mo.md])) + (FIX ? pre->kzt2 : a.kzt2[kzi + mo.md])) >= 1.0;
  const int mode = mg_mode<MG>(a);
  // explicit diffusion: RK4 only, where the fixed-mode kernels assume it
  const bool ediff = FIX ? G == MG_RK4 : a.explicit_diffusion != 0;
  const R scale = (R)a.scale, nu = (R)a.nu, dt = (R)a.dt;
  R nr[3], ni[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    nr[c] = FOLDED ? (R)rv[c].x : (R)rv[c].x * scale;
    ni[c] = FOLDED ? (R)rv[c].y : (R)rv[c].y * scale;
  }
  if constexpr (SPONGE && !FOLDED) {
    // sponge.f90:22-83: i k x (w/k^2), zero at k = 0
    if (k2d != 0.0) {
      R sr[3], sm[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) { sr[c] = (R)rv[3 + c].x * scale / k2; sm[c] = (R)rv[3 + c].y * scale / k2; }
      const R cr0 = kyv * sr[2] - kzv * sr[1], ci0 = kyv * sm[2] - kzv * sm[1];
      const R cr1 = kzv * sr[0] - kxv * sr[2], ci1 = kzv * sm[0] - kxv * sm[2];
      const R cr2 = kxv * sr[1] - kyv * sr[0], ci2 = kxv * sm[1] - kyv * sm[0];
      nr[0] += -ci0; ni[0] += cr0;
      nr[1] += -ci1; ni[1] += cr1;
      nr[2] += -ci2; ni[2] += cr2;
    }
  }
  if (kx == 0 && ky == 0 && kzi == 0) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (a.meanflow_dynamic[c]) { nr[c] = (R)(-a.pen_sum[c + mo.md] * a.dynamic_scale); ni[c] = R(0); }
  }
  cplx<T> udv[3];  // the diffusion operand, reused when it is u itself (RK4 stage 1)
  if (ediff) {
    // rhs.f90:441-468
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      udv[c] = opnd(UD, 0, c);
      nr[c] -= nu * k2 * (R)udv[c].x;
      ni[c] -= nu * k2 * (R)udv[c].y;
    }
  }
  if constexpr (G == MG_RHS) {
    if (a.mode == SM_PRESSURE) {
      // rhs.f90:342-379 pressure: -i (k . nlk) / k^2, zero at k = 0
      R pr = R(0), pi = R(0);
      if (k2d != 0.0) {
        const R dr = (kxv * nr[0] + kyv * nr[1]) + kzv * nr[2];
        const R di = (kxv * ni[0] + kyv * ni[1]) + kzv * ni[2];
        pr = di / k2;
        pi = -dr / k2;
      }
      sts<T>(RO + si, mk<T>((T)pr, (T)pi));
      return;
    }
  }
  if (k2d != 0.0) {
    // rhs.f90:575-615
    const R qr = (kxv * nr[0] + kyv * nr[1] + kzv * nr[2]) / k2;
    const R qi = (kxv * ni[0] + kyv * ni[1] + kzv * ni[2]) / k2;
    nr[0] -= kxv * qr; ni[0] -= kxv * qi;
    nr[1] -= kyv * qr; ni[1] -= kyv * qi;
    nr[2] -= kzv * qr; ni[2] -= kzv * qi;
  }
  R unr[3] = {0, 0, 0}, uni[3] = {0, 0, 0};
  switch (mode) {
    case SM_RHS_ONLY:
      if constexpr (G == MG_RHS) {
#pragma unroll
        for (int c = 0; c < 3; ++c) sts<T>(RO + c * a.fs + si, mk<T>((T)nr[c], (T)ni[c]));
      }
      break;
    case SM_EULER: if constexpr (G == MG_AB2) {
      // fluid_time_step.f90:718-727
      const R E = (R)vis_factor(a.vis_dt_new, a.nu, k2d, a.dealias_on && alias);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const cplx<T> u0 = opnd(U, 0, c);
        unr[c] = ((R)u0.x + dt * nr[c]) * E;
        uni[c] = ((R)u0.y + dt * ni[c]) * E;
        sts<T>(U + c * a.fs + si, mk<T>((T)unr[c], (T)uni[c]));
        sts<T>(RG0 + c * a.fs + si, mk<T>((T)(nr[c] * E), (T)(ni[c] * E)));
      }
      break;
    }
    case SM_AB2: if constexpr (G == MG_AB2) {
      // fluid_time_step.f90:802-811
      const R E = (R)vis_factor(a.vis_dt_new, a.nu, k2d, a.dealias_on && alias);
      const R b10 = (R)a.b10, b11 = (R)a.b11;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const cplx<T> u0 = opnd(U, 0, c);
        const cplx<T> n1 = opnd(RG1, 1, c);
        unr[c] = ((R)u0.x + b10 * nr[c] + b11 * (R)n1.x) * E;
        uni[c] = ((R)u0.y + b10 * ni[c] + b11 * (R)n1.y) * E;
        sts<T>(U + c * a.fs + si, mk<T>((T)unr[c], (T)uni[c]));
        sts<T>(RG0 + c * a.fs + si, mk<T>((T)(nr[c] * E), (T)(ni[c] * E)));
      }
      break;
    }
    case SM_RK2_1: if constexpr (G == MG_RK2) {
      // fluid_time_step.f90:554-575 (N0 weighted with the IF of the previous dt)
      const R Eo = (R)vis_factor(a.vis_dt_old, a.nu, k2d, a.dealias_on && alias);
      const R En = (R)vis_factor(a.vis_dt_new, a.nu, k2d, a.dealias_on && alias);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const R n0r = nr[c] * Eo, n0i = ni[c] * Eo;
        const cplx<T> u0 = opnd(U, 0, c);
        sts<T>(RG0 + c * a.fs + si, mk<T>((T)n0r, (T)n0i));
        sts<T>(U + c * a.fs + si, mk<T>((T)((R)u0.x * En + dt * n0r), (T)((R)u0.y * En + dt * n0i)));
      }
      break;
    }
    case SM_RK2_2: if constexpr (G == MG_RK2) {
      // fluid_time_step.f90:594-597
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const cplx<T> u0 = opnd(U, 0, c);
        const cplx<T> n0 = opnd(RG0, 1, c);
        unr[c] = (R)u0.x + R(0.5) * dt * (-(R)n0.x + nr[c]);
        uni[c] = (R)u0.y + R(0.5) * dt * (-(R)n0.y + ni[c]);
      }
      break;
    }
    default: if constexpr (G == MG_RK4) {
      // RK4 (fluid_time_step.f90:636-669) with an accumulator register
      const R kill = alias ? R(0) : R(1);
      const R dt6 = (R)(a.dt / 6.0), dt3 = (R)(a.dt / 3.0);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const R kr = nr[c] * kill, ki = ni[c] * kill;
        if (mode == SM_RK4_1) {
          const cplx<T> u0 = (FIX || (a.explicit_diffusion && a.udiff == a.u)) ? udv[c] : opnd(U, 0, c);
          sts<T>(ACC + c * a.fs + si, mk<T>((T)((R)u0.x + dt6 * kr), (T)((R)u0.y + dt6 * ki)));
          sts<T>(UST + c * a.fs + si, mk<T>((T)((R)u0.x + R(0.5) * dt * kr), (T)((R)u0.y + R(0.5) * dt * ki)));
        } else if (mode == SM_RK4_2 || mode == SM_RK4_3) {
          const cplx<T> ac = opnd(ACC, 1, c);
          R u0r, u0i;
          if (mode == SM_RK4_2 && (FIX || (a.explicit_diffusion && a.udiff == a.ustage))) {
            // stage 1 left acc = u_n + dt/6 k1 and ustage = u_n + dt/2 k1 (the
            // diffusion operand held in udv), so u_n = acc + (acc - ustage)/2
            // without a read of u_n
            u0r = (R)ac.x + R(0.5) * ((R)ac.x - (R)udv[c].x);
            u0i = (R)ac.y + R(0.5) * ((R)ac.y - (R)udv[c].y);
          } else {
            const cplx<T> u0 = opnd(U, 2, c);
            u0r = (R)u0.x; u0i = (R)u0.y;
          }
          const R cst = (mode == SM_RK4_2) ? R(0.5) * dt : dt;
          sts<T>(ACC + c * a.fs + si, mk<T>((T)((R)ac.x + dt3 * kr), (T)((R)ac.y + dt3 * ki)));
          sts<T>(UST + c * a.fs + si, mk<T>((T)(u0r + cst * kr), (T)(u0i + cst * ki)));
        } else {
          const cplx<T> ac = opnd(ACC, 1, c);
          unr[c] = ((R)ac.x + dt6 * kr) * kill;
          uni[c] = ((R)ac.y + dt6 * ki) * kill;
        }
      }
      break;
    }
  }
  if (final_stage) {
    if (kx == 0 && ky == 0 && kzi == 0) {
      // fluid_time_step.f90:1066-1090 set_mean_flow
#pragma unroll
      for (int c = 0; c < 3; ++c)
        if (a.set_mean[c]) { unr[c] = (R)a.mean_value[c]; uni[c] = R(0); }
    }
    if (mode == SM_RK2_2 || mode == SM_RK4_4 || (kx == 0 && ky == 0 && kzi == 0)) {
#pragma unroll
      for (int c = 0; c < 3; ++c) sts<T>(U + c * a.fs + si, mk<T>((T)unr[c], (T)uni[c]));
    }
    // integrals.f90:522-552 Parseval sum of the stored (rounded) field
    double ev = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double xr = (double)(T)unr[c], xi = (double)(T)uni[c];
      ev += xr * xr + xi * xi;
    }
    ev *= wE;
    const double t = esum + ev;
    ecomp += (fabs(esum) >= fabs(ev)) ? (esum - t) + ev : (ev - t) + esum;
    esum = t;
  }
}

// sponge.f90:22-83 at (kx, ky, kz), for the kernels that transform the three
// sponge fields first: s0..s2 (their unscaled transforms) are replaced by the
// term added to fields 0..2, i k x (w/k^2) scaled, zero at k = 0 and at
// padding lines (ky >= ny).  The operations are those of zepilogue's own
// sponge branch.
template <typename T>
__device__ __forceinline__ void sponge_term(const ZStepArgs& a, const MOff& mo, double kxv_d, int ky, int kzi,
                                            cplx<T>& s0, cplx<T>& s1, cplx<T>& s2) {
  using R = T;
  double k2d = 0.0, kyv_d = 0.0, kzv_d = 0.0;
  if (ky < a.ny) {
    kyv_d = a.ky[ky + mo.md];
    kzv_d = a.kz[kzi + mo.md];
    k2d = (kxv_d * kxv_d + kyv_d * kyv_d) + kzv_d * kzv_d;
  }
  if (k2d == 0.0) {
    s0 = s1 = s2 = mk<T>(T(0), T(0));
    return;
  }
  const R kxv = (R)kxv_d, kyv = (R)kyv_d, kzv = (R)kzv_d, k2 = (R)k2d, scale = (R)a.scale;
  const R sr[3] = {(R)s0.x * scale / k2, (R)s1.x * scale / k2, (R)s2.x * scale / k2};
  const R sm[3] = {(R)s0.y * scale / k2, (R)s1.y * scale / k2, (R)s2.y * scale / k2};
  const R cr0 = kyv * sr[2] - kzv * sr[1], ci0 = kyv * sm[2] - kzv * sm[1];
  const R cr1 = kzv * sr[0] - kxv * sr[2], ci1 = kzv * sm[0] - kxv * sm[2];
  const R cr2 = kxv * sr[1] - kyv * sr[0], ci2 = kxv * sm[1] - kyv * sm[0];
  s0 = mk<T>((T)-ci0, (T)cr0);
  s1 = mk<T>((T)-ci1, (T)cr1);
  s2 = mk<T>((T)-ci2, (T)cr2);
}
// field value o (unscaled transform) scaled, plus its sponge term c.  The
// sum is pinned where it is formed: left free, the compiler sinks the add
// into the epilogue and keeps both operands live until then.
template <typename T> __device__ __forceinline__ cplx<T> sponge_fold(cplx<T> o, cplx<T> c, T scale) {
  T x = o.x * scale, y = o.y * scale;
  x += c.x;
  y += c.y;
  asm volatile("" : "+v"(x), "+v"(y));
  return mk<T>(x, y);
}

// Lines per workgroup of the block kernel.  FP32 at nz = 512 takes 8 lines
// (512 threads, 64-byte Q segments) and is held to 128 VGPRs, so that two
// workgroups share a CU and one's transform stages (LDS work between block
// barriers) overlap the other's epilogue (HBM reads and writes of the state);
// a 16-line tile at nz = 512 fills a CU with one workgroup.
template <typename T> __host__ __device__ constexpr int zb_tl(int L) {
  return (sizeof(T) == 4 && L >= 512) ? 8 : Tile<T>::TL;
}
template <typename T> __host__ __device__ constexpr int zb_wpe(int L) {
  return (sizeof(T) == 4 && L >= 512) ? 4 : 1;
}

// LDS of the block kernel: the transform tile, a second tile where field 0's
// transformed lines wait for the epilogue when both fit in a CU's 160 KB,
// and the twiddle table.  Parking the lines there instead of in registers
// keeps the kernel within its VGPR budget (without it, the 16-line FP32
// kernel at nz = 512 spills 21 VGPRs, about 1.3 GB of extra memory traffic
// per launch at 512^3).  The twiddles are read from LDS because the vector-
// memory counter is in order: a stage waiting for a twiddle load from
// global memory would also wait for the next field's operands, loaded just
// before it.
template <typename T, int L, int TL> struct ZBShape {
  static constexpr size_t tile = sizeof(cplx<T>) * TL * line_stride<T>(L);
  static constexpr size_t tws = sizeof(cplx<T>) * L;
  static constexpr bool park = 2 * tile + tws <= 160 * 1024;
  static constexpr size_t smem = (park ? 2 * tile : tile) + tws;
};

// EU > 0: the epilogue of at most EU positions is scheduled together (a
// scheduling barrier after every EU positions keeps the compiler from
// hoisting all positions' state loads)
template <typename T, int L, int TL, int NT, bool SPONGE, int MG, int EU>
__global__ __launch_bounds__(NT, zb_wpe<T>(L)) void zstep_k(ZStepArgs a) {
  const MOff mo = member_off<T>(a);
  constexpr int NSTG = plan_count_k(PLAN_TILE, L);
  constexpr int R0 = plan_radix_k(PLAN_TILE, L, 0);
  constexpr int RL = plan_radix_k(PLAN_TILE, L, NSTG - 1);
  constexpr int NSL = L / RL;
  constexpr int LSP = line_stride<T>(L);
  constexpr int NB0 = TL * (L / R0), PER0 = (NB0 + NT - 1) / NT;
  constexpr int NBL = TL * (L / RL), PERL = (NBL + NT - 1) / NT;
  constexpr int NF = SPONGE ? 6 : 3;
  extern __shared__ char smem_raw[];
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem_raw);
  __shared__ double red[NT / 64];
  const int tid = threadIdx.x;
  // XCD-aware tile order: workgroups are dealt round-robin to the 8 XCDs, so
  // workgroups x and x + 8 of a row run on one XCD one after the other; they
  // take neighbouring ky tiles, whose Q segments share 128-byte lines.
  int bx = blockIdx.x;
  if (TL * sizeof(cplx<T>) < 128 && gridDim.x % 8 == 0) bx = (bx % 8) * (gridDim.x / 8) + bx / 8;
  const int b0 = bx * TL;
  const int kx = blockIdx.y;
  constexpr bool PARK = ZBShape<T, L, TL>::park;
  // fixed-mode kernels load each epilogue position's operands one position
  // ahead (ZPre); without the sponge, field 2 then waits in the transform tile
  constexpr bool AHEAD = MG >= MG_FIXED;
  constexpr bool TILE2 = AHEAD && !SPONGE;
  // twiddles in LDS (first read after the block barrier that ends stage 0):
  // loaded into registers after the first field's operands are requested
  // and written before that barrier, so that the two memory latencies overlap
  cplx<T>* twl = buf + (PARK ? 2 : 1) * TL * LSP;
  const cplx<T>* tw = twl;
  constexpr int NTW = (L + NT - 1) / NT;
  cplx<T> twr[NTW];
  auto put_tw = [&] {
#pragma unroll
    for (int k = 0; k < NTW; ++k)
      if (tid + k * NT < L) twl[tid + k * NT] = twr[k];
  };
  const cplx<T>* nq = reinterpret_cast<const cplx<T>*>(a.nq);
  // park: field 0 by (ky line, kz), written and read back by the same thread
  cplx<T>* park = buf + TL * LSP;
  // r[f][p][q]: field f (0..2) at (ky = b0 + bb, kz = j + q L/RL) of
  // last-stage butterfly p.  With the sponge the three sponge fields are
  // transformed first into r, turned in place into their curl terms
  // i k x (w/k^2) (sponge.f90:22-83) per position, and each term is added to
  // its field as that field is transformed, so that no more than three
  // fields' values are live (six would spill at nz >= 384).
  cplx<T> r[3][PERL][RL];
  // field of the i-th transform
  constexpr auto fld = [](int i) { return SPONGE ? (i < 3 ? 3 + i : i - 3) : i; };
  // stage-0 operands of a field; the next field's are in flight while the
  // current one goes through its LDS stages
  cplx<T> v[PER0][R0];
  auto loadq = [&](int f) {
#pragma unroll
    for (int p = 0; p < PER0; ++p) {
      const int beta = tid + p * NT;
      if (NB0 % NT == 0 || beta < NB0) {
        const int bb = beta % TL, j = beta / TL;
        const bool ok = b0 + bb < a.ny;
#pragma unroll
        for (int q = 0; q < R0; ++q)
          v[p][q] = ok ? nq[mo.mc + f * a.fs + ((long)kx * a.nz + j + q * (L / R0)) * a.ny + b0 + bb] : mk<T>(T(0), T(0));
      }
    }
  };
  loadq(fld(0));
  {
    const cplx<T>* twg = reinterpret_cast<const cplx<T>*>(a.tw) + mo.mc;
#pragma unroll
    for (int k = 0; k < NTW; ++k)
      if (tid + k * NT < L) twr[k] = twg[tid + k * NT];
  }
  if constexpr (!SPONGE) {
#pragma unroll
    for (int f = 0; f < NF; ++f) {
#pragma unroll
      for (int p = 0; p < PER0; ++p) {
        const int beta = tid + p * NT;
        if (NB0 % NT == 0 || beta < NB0) {
          const int bb = beta % TL, j = beta / TL;
          Bfly<R0, -1, T>::run(v[p]);
#pragma unroll
          for (int q = 0; q < R0; ++q) buf[bb * LSP + lp<T>(j * R0 + q)] = v[p][q];
        }
      }
      if (f == 0) put_tw();
      __syncthreads();
      if (f + 1 < NF) loadq(f + 1);
      Stages<T, L, TL, NT, PLAN_TILE, -1, LSP, 1, NSTG - 1, R0>::run(buf, tw);
#pragma unroll
      for (int p = 0; p < PERL; ++p) {
        const int beta = tid + p * NT;
        if (NBL % NT == 0 || beta < NBL) {
          const int bb = beta / (L / RL), j = beta - bb * (L / RL);
#pragma unroll
          for (int q = 0; q < RL; ++q) r[f][p][q] = buf[bb * LSP + lp<T>(j + q * (L / RL))];
          twiddle_bfly<T, L, RL, -1, NSL>(r[f][p], j, tw);
          if (PARK && f == 0) {
#pragma unroll
            for (int q = 0; q < RL; ++q) park[bb * LSP + lp<T>(j + q * (L / RL))] = r[f][p][q];
          }
        }
      }
      __syncthreads();
    }
    if constexpr (TILE2) {
      // field 2 waits in the transform tile, free after the last barrier;
      // the registers go to the next position's operands
#pragma unroll
      for (int p = 0; p < PERL; ++p) {
        const int beta = tid + p * NT;
        if (NBL % NT == 0 || beta < NBL) {
          const int bb = beta / (L / RL), j = beta - bb * (L / RL);
#pragma unroll
          for (int q = 0; q < RL; ++q) buf[bb * LSP + lp<T>(j + q * (L / RL))] = r[2][p][q];
        }
      }
      asm volatile("" ::: "memory");  // read back below, not forwarded from registers
    }
  } else {
    // Transform number i as straight-line code per i (with a loop left to
    // the unroller the allocation spills at nz = 512), each on its own copy
    // of the thread index (see lds_stage).
    auto xform = [&](auto ic) {
      constexpr int i = decltype(ic)::value;
      constexpr int f = fld(i);
      constexpr int slot = i < 3 ? i : f;  // register slot of this transform's output
      int tl = tid;
      asm volatile("" : "+v"(tl));
#pragma unroll
      for (int p = 0; p < PER0; ++p) {
        const int beta = tl + p * NT;
        if (NB0 % NT == 0 || beta < NB0) {
          const int bb = beta % TL, j = beta / TL;
          Bfly<R0, -1, T>::run(v[p]);
#pragma unroll
          for (int q = 0; q < R0; ++q) buf[bb * LSP + lp<T>(j * R0 + q)] = v[p][q];
        }
      }
      if constexpr (i == 0) put_tw();
      __syncthreads();
      if constexpr (i + 1 < NF) loadq(fld(i + 1));
      Stages<T, L, TL, NT, PLAN_TILE, -1, LSP, 1, NSTG - 1, R0>::run(buf, tw, tl);
#pragma unroll
      for (int p = 0; p < PERL; ++p) {
        const int beta = tl + p * NT;
        if (NBL % NT == 0 || beta < NBL) {
          const int bb = beta / (L / RL), j = beta - bb * (L / RL);
          cplx<T> o[RL];
#pragma unroll
          for (int q = 0; q < RL; ++q) o[q] = buf[bb * LSP + lp<T>(j + q * (L / RL))];
          twiddle_bfly<T, L, RL, -1, NSL>(o, j, tw);
          // slot 0 (the first sponge field, then its term, then field 0)
          // waits in the park tile where the LDS allows
          cplx<T>* pk = park + bb * LSP;
          if constexpr (i >= 3) {
            // scale, and add the field's sponge term
#pragma unroll
            for (int q = 0; q < RL; ++q) {
              const cplx<T> c = (PARK && slot == 0) ? pk[lp<T>(j + q * (L / RL))] : r[slot][p][q];
              o[q] = sponge_fold<T>(o[q], c, (T)a.scale);
            }
          }
          if constexpr (PARK && slot == 0) {
#pragma unroll
            for (int q = 0; q < RL; ++q) pk[lp<T>(j + q * (L / RL))] = o[q];
          } else {
#pragma unroll
            for (int q = 0; q < RL; ++q) r[slot][p][q] = o[q];
          }
          if constexpr (i == 2) {
            // all three sponge fields are in: turn them into the terms
            const double kxs = a.kx[kx + mo.md];
#pragma unroll
            for (int q = 0; q < RL; ++q) {
              cplx<T> s0 = PARK ? pk[lp<T>(j + q * (L / RL))] : r[0][p][q];
              sponge_term<T>(a, mo, kxs, b0 + bb, j + q * (L / RL), s0, r[1][p][q], r[2][p][q]);
              if constexpr (PARK) pk[lp<T>(j + q * (L / RL))] = s0;
              else r[0][p][q] = s0;
              __builtin_amdgcn_sched_barrier(0);  // one position at a time
            }
          }
        }
      }
      __syncthreads();
    };
    xform(std::integral_constant<int, 0>{});
    xform(std::integral_constant<int, 1>{});
    xform(std::integral_constant<int, 2>{});
    xform(std::integral_constant<int, 3>{});
    xform(std::integral_constant<int, 4>{});
    xform(std::integral_constant<int, 5>{});
  }

  const double kxv = a.kx[kx + mo.md];
  const double kxt2 = a.kxt2[kx + mo.md];
  const double wE = (kx == 0 || kx == a.nx / 2) ? 0.5 : 1.0;
  double esum = 0.0, ecomp = 0.0;
  const bool final_stage = mg_final_stage<MG>(a);
  // position k = p RL + q of this thread: ky line bb, kz index kzi; false
  // where it lies beyond the tile or the grid
  constexpr int NPOS = PERL * RL;
  auto at = [&](int k, int& bb, int& kzi) -> bool {
    const int beta = tid + (k / RL) * NT;
    bb = beta / (L / RL);
    kzi = beta - bb * (L / RL) + (k % RL) * (L / RL);
    return (NBL % NT == 0 || beta < NBL) && b0 + bb < a.ny;
  };
  // operands of position k, read at a valid line for every thread (a
  // position that does not exist reads line b0 and is not used), so that the
  // loads need no branch
  ZPre<T> pre[2];
  auto preload = [&](int k, ZPre<T>& pr) {
    int bb, kzi;
    const bool ok = at(k, bb, kzi);
    zpreload<T, MG>(a, mo, kx, ok ? b0 + bb : b0, ok ? kzi : 0, pr);
  };
  if constexpr (AHEAD) preload(0, pre[0]);
#pragma unroll
  for (int k = 0; k < NPOS; ++k) {
    if constexpr (AHEAD)
      if (k + 1 < NPOS) preload(k + 1, pre[(k + 1) & 1]);
    int bb, kzi;
    if (!at(k, bb, kzi)) continue;
    const int p = k / RL, q = k % RL;
    cplx<T> rv[3];
#pragma unroll
    for (int f = 0; f < 3; ++f)
      rv[f] = (PARK && f == 0) ? park[bb * LSP + lp<T>(kzi)] : ((TILE2 && f == 2) ? buf[bb * LSP + lp<T>(kzi)] : r[f][p][q]);
    zepilogue<T, false, MG, SPONGE>(a, mo, kx, b0 + bb, kzi, rv, kxv, kxt2, wE, final_stage, esum, ecomp, &pre[k & 1]);
    if constexpr (EU > 0)
      if ((k + 1) % EU == 0) __builtin_amdgcn_sched_barrier(0);
  }
  if (final_stage && a.energy_partial) {
    double v = wave_sum(esum + ecomp);
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
      for (int w = 0; w < NT / 64; ++w) s += red[w];
      a.energy_partial[mo.md + (long)blockIdx.y * gridDim.x + bx] = s;
    }
  }
}

// The same pass with each z line transformed inside one wavefront (nz = 64,
// 128, 256, 512): per field, TL lines' 128-byte ky segments are staged
// through the LDS tile (one block-wide wait each way), every wave transforms
// its own line in registers, and the next field's loads are in flight
// meanwhile.  A lane ends with positions kz = l + qG of all fields of its
// line, where the epilogue runs.
template <typename T, int L> struct ZWShape {
  static constexpr int TL = Tile<T>::TL;
  static constexpr int NW = TL / wavefft::TWPlan<L>::LPW;
  static constexpr int NT = NW * 64;
  static constexpr int LS = line_stride<T>(L);
  static constexpr int K = TL * L / NT;
  static constexpr bool TWL = 2 * (TL * LS + L) * (int)sizeof(cplx<T>) <= 160 * 1024;
  static constexpr size_t smem = (size_t)(TL * LS + (TWL ? L : 0)) * sizeof(cplx<T>);
};

template <typename T, int L, bool SPONGE, int MG>
__global__ __launch_bounds__((ZWShape<T, L>::NT)) void zstep_wave_k(ZStepArgs a) {
  const MOff mo = member_off<T>(a);
  using namespace wavefft;
  using Sh = ZWShape<T, L>;
  constexpr int TL = Sh::TL, NT = Sh::NT, LS = Sh::LS, K = Sh::K;
  constexpr int G = TWPlan<L>::G;
  constexpr int NF = SPONGE ? 6 : 3;
  extern __shared__ char smem_raw[];
  cplx<T>* tile = reinterpret_cast<cplx<T>*>(smem_raw);
  __shared__ double red[NT / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int line = wv * TWPlan<L>::LPW + lane / G, l = lane % G;
  cplx<T>* ln = tile + line * LS;
  const int b0 = blockIdx.x * TL;
  const int kx = blockIdx.y;
  // twiddles: the block's LDS copy (made after the first field's loads are
  // issued, so that the two memory latencies overlap; first read after the
  // first block barrier), or global memory where it does not fit
  const cplx<T>* tw = Sh::TWL ? tile + TL * LS : reinterpret_cast<const cplx<T>*>(a.tw) + mo.mc;
  auto copy_tw = [&] {
    if constexpr (Sh::TWL) {
      cplx<T>* twl = tile + TL * LS;
      const cplx<T>* twg = reinterpret_cast<const cplx<T>*>(a.tw) + mo.mc;
      for (int i = tid; i < L; i += NT) twl[i] = twg[i];
    }
  };
  const cplx<T>* nq = reinterpret_cast<const cplx<T>*>(a.nq);
  cplx<T> x[K];
  auto load = [&](int f) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int e = tid + k * NT, bb = e % TL, aa = e / TL;
      const bool ok = b0 + bb < a.ny;
      x[k] = ok ? nq[mo.mc + f * a.fs + ((long)kx * a.nz + aa) * a.ny + b0 + bb] : mk<T>(T(0), T(0));
    }
  };
  cplx<T> r[3][8];
  if constexpr (!SPONGE) {
    load(0);
    copy_tw();
#pragma unroll
    for (int f = 0; f < NF; ++f) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int e = tid + k * NT, bb = e % TL, aa = e / TL;
        tile[bb * LS + lp<T>(aa)] = x[k];
      }
      __syncthreads();
      if (f + 1 < NF) load(f + 1);
#pragma unroll
      for (int q = 0; q < 8; ++q) r[f][q] = ln[lp<T>(l + q * G)];
      wave_sync();
      twfft<T, L, -1>(r[f], tw, ln, l);
      __syncthreads();
    }
  } else {
    // transform order, folding of the sponge terms and per-transform copies
    // of the thread index as in zstep_k
    constexpr auto fld = [](int i) { return i < 3 ? 3 + i : i - 3; };
    auto loadt = [&](int f, int t) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int e = t + k * NT, bb = e % TL, aa = e / TL;
        const bool ok = b0 + bb < a.ny;
        x[k] = ok ? nq[mo.mc + f * a.fs + ((long)kx * a.nz + aa) * a.ny + b0 + bb] : mk<T>(T(0), T(0));
      }
    };
    loadt(fld(0), tid);
    copy_tw();
    auto xform = [&](auto ic) {
      constexpr int i = decltype(ic)::value;
      constexpr int f = fld(i);
      int t = tid;
      asm volatile("" : "+v"(t));
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int e = t + k * NT, bb = e % TL, aa = e / TL;
        tile[bb * LS + lp<T>(aa)] = x[k];
      }
      __syncthreads();
      if constexpr (i + 1 < NF) loadt(fld(i + 1), t);
      cplx<T> o[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) o[q] = ln[lp<T>(l + q * G)];
      wave_sync();
      twfft<T, L, -1>(o, tw, ln, l);
      if constexpr (i < 3) {
#pragma unroll
        for (int q = 0; q < 8; ++q) r[i][q] = o[q];
        if constexpr (i == 2) {
          const double kxs = a.kx[kx + mo.md];
#pragma unroll
          for (int q = 0; q < 8; ++q) sponge_term<T>(a, mo, kxs, b0 + line, l + q * G, r[0][q], r[1][q], r[2][q]);
        }
      } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) r[f][q] = sponge_fold<T>(o[q], r[f][q], (T)a.scale);
      }
      __syncthreads();
    };
    xform(std::integral_constant<int, 0>{});
    xform(std::integral_constant<int, 1>{});
    xform(std::integral_constant<int, 2>{});
    xform(std::integral_constant<int, 3>{});
    xform(std::integral_constant<int, 4>{});
    xform(std::integral_constant<int, 5>{});
  }

  const double kxv = a.kx[kx + mo.md];
  const double kxt2 = a.kxt2[kx + mo.md];
  const double wE = (kx == 0 || kx == a.nx / 2) ? 0.5 : 1.0;
  double esum = 0.0, ecomp = 0.0;
  const bool final_stage = mg_final_stage<MG>(a);
  const int ky = b0 + line;
  if (ky < a.ny) {
    // fixed-mode kernels: operands loaded one position ahead (ZPre)
    constexpr bool AHEAD = MG >= MG_FIXED;
    ZPre<T> pre[2];
    if constexpr (AHEAD) zpreload<T, MG>(a, mo, kx, ky, l, pre[0]);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if constexpr (AHEAD)
        if (q + 1 < 8) zpreload<T, MG>(a, mo, kx, ky, l + (q + 1) * G, pre[(q + 1) & 1]);
      cplx<T> rv[3];
#pragma unroll
      for (int f = 0; f < 3; ++f) rv[f] = r[f][q];
      zepilogue<T, false, MG, SPONGE>(a, mo, kx, ky, l + q * G, rv, kxv, kxt2, wE, final_stage, esum, ecomp, &pre[q & 1]);
    }
  }
  if (final_stage && a.energy_partial) {
    double v = wave_sum(esum + ecomp);
    if (lane == 0) red[wv] = v;
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
      for (int w = 0; w < NT / 64; ++w) s += red[w];
      a.energy_partial[mo.md + (long)blockIdx.y * gridDim.x + blockIdx.x] = s;
    }
  }
}

// The in-wave kernel holds eight positions of three fields per lane; at
// nz = 512 that exceeds its register budget (it spilled 48-130 bytes per lane
// in every stepper family), so nz = 512 always takes the block kernel.
bool zstep_wave_supported(int nz) { return nz == 64 || nz == 128 || nz == 256; }

int zstep_blocks(int nz, int ny, int nxh, bool fp64, bool wave) {
  const int TL = (wave && zstep_wave_supported(nz)) ? (fp64 ? Tile<double>::TL : Tile<float>::TL)
                                               : (fp64 ? zb_tl<double>(nz) : zb_tl<float>(nz));
  return ((ny + TL - 1) / TL) * nxh;
}

template <typename T, int L, bool SP, int MG, int EU>
static void run_zb(const ZStepArgs& a, hipStream_t s) {
  constexpr int TL = zb_tl<T>(L);
  constexpr int NT = pick_nt(TL * L);
  const size_t smem = ZBShape<T, L, TL>::smem;
  static bool init = false;
  if (!init) { set_smem(zstep_k<T, L, TL, NT, SP, MG, EU>, smem); init = true; }
  dim3 grid((a.ny + TL - 1) / TL, a.nxh, batch_n(a.bat));
  hipLaunchKernelGGL((zstep_k<T, L, TL, NT, SP, MG, EU>), grid, dim3(NT), smem, s, a);
}

template <typename T, int L, bool SP, int MG>
static void run_z(const ZStepArgs& a, hipStream_t s) {
  if constexpr (L == 64 || L == 128 || L == 256) {
    if (a.wave) {
      using Sh = ZWShape<T, L>;
      static bool winit = false;
      if (!winit) { set_smem(zstep_wave_k<T, L, SP, MG>, Sh::smem); winit = true; }
      dim3 grid((a.ny + Sh::TL - 1) / Sh::TL, a.nxh, batch_n(a.bat));
      hipLaunchKernelGGL((zstep_wave_k<T, L, SP, MG>), grid, dim3(Sh::NT), Sh::smem, s, a);
      return;
    }
  }
  if constexpr (L >= 256 && MG < MG_FIXED) {
    switch (a.eu) {
      case 1: run_zb<T, L, SP, MG, 1>(a, s); return;
      case 2: run_zb<T, L, SP, MG, 2>(a, s); return;
      case 4: run_zb<T, L, SP, MG, 4>(a, s); return;
      default: break;
    }
  }
  run_zb<T, L, SP, MG, 0>(a, s);
}
// The operands Solver::fluid_timestep passes with each stepper mode
// (explicit diffusion with RK4 only; RK4's diffusion operand is u_n in stage
// 1 and the stage state afterwards), which the fixed-mode kernels assume.
static bool fixed_mode_ok(const ZStepArgs& a) {
  switch (a.mode) {
    case SM_RK4_1: return a.explicit_diffusion && a.udiff == a.u;
    case SM_RK4_2: case SM_RK4_3: case SM_RK4_4: return a.explicit_diffusion && a.udiff == a.ustage;
    case SM_EULER: case SM_AB2: case SM_RK2_1: case SM_RK2_2: return !a.explicit_diffusion;
    default: return false;
  }
}
template <typename T, int L, bool SP>
static void run_z_mg(const ZStepArgs& a, hipStream_t s) {
  // fixed-mode kernels: without the sponge, nz >= 64, the default epilogue
  // grouping (the other variants keep the run-time mode)
  if constexpr (!SP && L >= 64) {
    if (fixed_mode_ok(a) && (L < 256 || a.eu == 0 || (L == 256 && a.wave))) {
      switch (a.mode) {
        case SM_EULER: run_z<T, L, SP, MG_FIXED + SM_EULER>(a, s); return;
        case SM_AB2: run_z<T, L, SP, MG_FIXED + SM_AB2>(a, s); return;
        case SM_RK2_1: run_z<T, L, SP, MG_FIXED + SM_RK2_1>(a, s); return;
        case SM_RK2_2: run_z<T, L, SP, MG_FIXED + SM_RK2_2>(a, s); return;
        case SM_RK4_1: run_z<T, L, SP, MG_FIXED + SM_RK4_1>(a, s); return;
        case SM_RK4_2: run_z<T, L, SP, MG_FIXED + SM_RK4_2>(a, s); return;
        case SM_RK4_3: run_z<T, L, SP, MG_FIXED + SM_RK4_3>(a, s); return;
        case SM_RK4_4: run_z<T, L, SP, MG_FIXED + SM_RK4_4>(a, s); return;
        default: break;
      }
    }
  }
  switch (mode_group(a.mode)) {
    case MG_RHS: run_z<T, L, SP, MG_RHS>(a, s); break;
    case MG_AB2: run_z<T, L, SP, MG_AB2>(a, s); break;
    case MG_RK2: run_z<T, L, SP, MG_RK2>(a, s); break;
    default: run_z<T, L, SP, MG_RK4>(a, s); break;
  }
}
template <typename T, int L>
static void run_z_sp(const ZStepArgs& a, hipStream_t s) {
  if (a.sponge) run_z_mg<T, L, true>(a, s); else run_z_mg<T, L, false>(a, s);
}

template <typename T>
void launch_zstep(const ZStepArgs& a, hipStream_t s) {
  switch (a.nz) {
    case 16: run_z_sp<T, 16>(a, s); break;
    case 24: run_z_sp<T, 24>(a, s); break;
    case 32: run_z_sp<T, 32>(a, s); break;
    case 48: run_z_sp<T, 48>(a, s); break;
    case 64: run_z_sp<T, 64>(a, s); break;
    case 96: run_z_sp<T, 96>(a, s); break;
    case 128: run_z_sp<T, 128>(a, s); break;
    case 192: run_z_sp<T, 192>(a, s); break;
    case 256: run_z_sp<T, 256>(a, s); break;
    case 384: run_z_sp<T, 384>(a, s); break;
    case 512: run_z_sp<T, 512>(a, s); break;
    default: throw std::runtime_error("unsupported nz " + std::to_string(a.nz));
  }
}
template void launch_zstep<float>(const ZStepArgs&, hipStream_t);
template void launch_zstep<double>(const ZStepArgs&, hipStream_t);

}  // namespace flusi
