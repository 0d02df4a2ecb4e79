// Pencil-FFT passes of the penalised pseudospectral right-hand side.
//
// One RHS evaluation (rhs.f90:147-330 cal_nlk_fsi + :575-615 add_grad_pressure)
// runs as five passes over HBM, every pointwise operation fused into one of
// them (layouts: S = spectral state (kx,ky,kz) kz fastest, Q = (kx,z,ky) ky
// fastest, P = (z,y,kx) kx fastest):
//
//   ZINV  S->Q  inverse z FFT of u, and of i kz u_x, i kz u_y (curl pieces)
//   YINV  Q->P  inverse y FFT; forms w_x = i ky U_z - D_y and i ky U_x
//   XRHS  P->P  C2R of u and w, u x w - chi/eta (u - us) [+ sponge], R2C
//   YFWD  P->Q  forward y FFT
//   ZSTEP Q->S  forward z FFT, 1/N, sponge curl, projection, explicit
//               diffusion, dealiasing and the AB2/RK2/RK4 register update
//
// Each pass reads whole contiguous lines on one side and 128-byte segments of
// TL lines on the other, staging the tile through LDS.
#include "fft_core.h"
#include "flusi_kernels.h"
#include "pass_common.h"
#include <cstdlib>
#include <stdexcept>
#include <string>

namespace flusi {

// ======================================================== generic tile pass
// One workgroup transforms TL lines of length L for each job: the first
// Stockham stage is computed straight from the global loads, the middle
// stages in LDS and the last stage straight into the global stores.  The
// thread <-> butterfly mapping of the first (last) stage follows the input
// (output) layout: consecutive threads walk the contiguous axis, or the TL
// lines of a 128-byte segment.
//
// A job's operands are read once into registers and give one or two
// transforms (e.g. u and i k u).  The next job's loads are issued as soon as
// the current operands have been consumed, so they are in flight while the
// LDS stages of the current transform run.
template <typename T, int L, int TL, int NT, int S, bool CONTIG_IN, bool PF>
__global__ __launch_bounds__(NT) void tile_pass_k(TilePassArgs a) {
  constexpr int NSTG = plan_count_k(PLAN_TILE, L);
  static_assert(NSTG >= 2, "tile pass needs at least two stages");
  constexpr int R0 = plan_radix_k(PLAN_TILE, L, 0);
  constexpr int RL = plan_radix_k(PLAN_TILE, L, NSTG - 1);
  constexpr int LSP = line_stride<T>(L);
  constexpr int NB0 = TL * (L / R0), PER0 = (NB0 + NT - 1) / NT;
  constexpr int NBL = TL * (L / RL), PERL = (NBL + NT - 1) / NT;
  extern __shared__ char smem_raw[];
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int b0 = blockIdx.x * TL;
  const long c = blockIdx.y;
  const long moff = (long)blockIdx.z * a.bat.mstride;  // batch member's slab
  // global-address-space views of the arguments (see gptr)
  const gptr<const cplx<T>> tw = (gptr<const cplx<T>>)mshift(a.tw, moff);
  const T sc = (T)a.scale;

  cplx<T> r0[PER0][R0], r1[PER0][R0];
  // operand `which` (0: in0, prefetched; 1: in1, read when its job starts)
  auto load = [&](int jb, int which) {
    const gptr<const cplx<T>> in = (gptr<const cplx<T>>)mshift(which ? a.jobs[jb].in1 : a.jobs[jb].in0, moff);
#pragma unroll
    for (int p = 0; p < PER0; ++p) {
      const int beta = tid + p * NT;
      if (NB0 % NT == 0 || beta < NB0) {
        int bb, j;
        if (CONTIG_IN) { bb = beta / (L / R0); j = beta - bb * (L / R0); }
        else { bb = beta % TL; j = beta / TL; }
        const bool ok = b0 + bb < a.nb;
        // the contiguous side has unit stride (checked on the host)
        const long base = CONTIG_IN ? c * a.sin.c + (long)(b0 + bb) * a.sin.b + j
                                    : c * a.sin.c + (long)(b0 + bb) + (long)j * a.sin.a;
        const long step = CONTIG_IN ? (long)(L / R0) : (long)(L / R0) * a.sin.a;
#pragma unroll
        for (int q = 0; q < R0; ++q) {
          const cplx<T> x = ok ? gld<T>(in + base + q * step) : mk<T>(T(0), T(0));
          if (which) r1[p][q] = x;
          else r0[p][q] = x;
        }
      }
    }
  };

  if (PF) load(0, 0);
  for (int jb = 0; jb < a.njobs; ++jb) {
    const int nout = a.jobs[jb].out2 ? 2 : 1;
    if (!PF) load(jb, 0);
    if (a.jobs[jb].in1) load(jb, 1);
    for (int o = 0; o < nout; ++o) {
      const int op = o ? a.jobs[jb].op2 : a.jobs[jb].op;
      const gptr<cplx<T>> out = (gptr<cplx<T>>)mshift(o ? a.jobs[jb].out2 : a.jobs[jb].out, moff);
      // FP32 at L >= 384: a copy of the thread index per transform, so that
      // the LDS addresses and twiddles are recomputed in each iteration of
      // the job loop instead of being kept live (spilled) across it
      int tl = tid;
      asm volatile("" : "+v"(tl));
#pragma unroll
      for (int p = 0; p < PER0; ++p) {
        const int beta = tl + p * NT;
        if (NB0 % NT == 0 || beta < NB0) {
          int bb, j;
          if (CONTIG_IN) { bb = beta / (L / R0); j = beta - bb * (L / R0); }
          else { bb = beta % TL; j = beta / TL; }
          cplx<T> v[R0];
          if (op == OP_COPY) {
#pragma unroll
            for (int q = 0; q < R0; ++q) v[q] = r0[p][q];
          } else {
            // k_a = kscale * wrap(a), wrap(a) = a - L for a >= L/2; the empty
            // asm keeps the wavenumbers from being hoisted out of the job loop
            // (they would hold R0 registers for the whole kernel)
            int jj = j;
            asm volatile("" : "+v"(jj));
#pragma unroll
            for (int q = 0; q < R0; ++q) {
              const int aa = jj + q * (L / R0);
              const T k = (T)(a.kscale * double(aa < L / 2 ? aa : aa - L));
              v[q] = ik_<T>(k, r0[p][q]);
            }
            if (op == OP_DERIV_MINUS) {
#pragma unroll
              for (int q = 0; q < R0; ++q) v[q] = sub(v[q], r1[p][q]);
            }
          }
          Bfly<R0, S, T>::run(v);
#pragma unroll
          for (int q = 0; q < R0; ++q) buf[bb * LSP + lp<T>(j * R0 + q)] = v[q];
        }
      }
      if (PF && o == nout - 1 && jb + 1 < a.njobs) load(jb + 1, 0);
      __syncthreads();
      Stages<T, L, TL, NT, PLAN_TILE, S, LSP, 1, NSTG - 1, R0>::run(buf, tw, tl);
      constexpr int NSL = L / RL;
#pragma unroll
      for (int p = 0; p < PERL; ++p) {
        const int beta = tl + p * NT;
        if (NBL % NT == 0 || beta < NBL) {
          int bb, j;
          if (!CONTIG_IN) { bb = beta / (L / RL); j = beta - bb * (L / RL); }
          else { bb = beta % TL; j = beta / TL; }
          cplx<T> w[RL];
#pragma unroll
          for (int q = 0; q < RL; ++q) w[q] = buf[bb * LSP + lp<T>(j + q * (L / RL))];
          twiddle_bfly<T, L, RL, S, NSL>(w, j, tw);
          // segment-side stores may cover the row padding (zeros; see nb_store)
          if (b0 + bb < (CONTIG_IN && a.nb_store > a.nb ? a.nb_store : a.nb)) {
            const long base = !CONTIG_IN ? c * a.sout.c + (long)(b0 + bb) * a.sout.b + j
                                         : c * a.sout.c + (long)(b0 + bb) + (long)j * a.sout.a;
            const long step = !CONTIG_IN ? (long)(L / RL) : (long)(L / RL) * a.sout.a;
#pragma unroll
            for (int q = 0; q < RL; ++q) gst<T>(out + base + q * step, mk<T>(w[q].x * sc, w[q].y * sc));
          }
        }
      }
      __syncthreads();
    }
  }
}

template <typename T, int L, int S, bool CI, bool PF>
static void run_tile_pf(const TilePassArgs& a, hipStream_t s) {
  constexpr int TL = Tile<T>::TL;
  constexpr int NT = pick_nt(TL * L);
  const size_t smem = sizeof(cplx<T>) * TL * line_stride<T>(L);
  static bool init = false;
  if (!init) { set_smem(tile_pass_k<T, L, TL, NT, S, CI, PF>, smem); init = true; }
  dim3 grid((a.nb + TL - 1) / TL, a.nc, batch_n(a.bat));
  hipLaunchKernelGGL((tile_pass_k<T, L, TL, NT, S, CI, PF>), grid, dim3(NT), smem, s, a);
}
template <typename T, int L, int S, bool CI>
static void run_tile(const TilePassArgs& a, hipStream_t s) {
  if (a.pf) run_tile_pf<T, L, S, CI, true>(a, s);
  else run_tile_pf<T, L, S, CI, false>(a, s);
}

template <typename T, int L>
static void run_tile_dir(const TilePassArgs& a, int dir, bool ci, hipStream_t s) {
  if (dir > 0) { if (ci) run_tile<T, L, +1, true>(a, s); else run_tile<T, L, +1, false>(a, s); }
  else { if (ci) run_tile<T, L, -1, true>(a, s); else run_tile<T, L, -1, false>(a, s); }
}

bool x_length_supported(int m) {
  switch (m) {
    case 8: case 16: case 24: case 32: case 48: case 64: case 96: case 128: case 192: case 256: return true;
    default: return false;
  }
}

bool fft_length_supported(int n) {
  switch (n) {
    case 16: case 24: case 32: case 48: case 64: case 96: case 128:
    case 192: case 256: case 384: case 512: return true;
    default: return false;
  }
}

template <typename T>
void launch_tile_pass(const TilePassArgs& a, int L, int dir, bool ci, hipStream_t s) {
  // contiguous input lines -> output tile axis contiguous, and vice versa
  const bool unit = ci ? (a.sin.a == 1 && a.sout.b == 1) : (a.sin.b == 1 && a.sout.a == 1);
  if (!unit) throw std::runtime_error("tile pass: the contiguous axes must have unit stride");
  if (a.njobs < 1 || a.njobs > 6) throw std::runtime_error("tile pass: 1..6 jobs");
  switch (L) {
    case 16: run_tile_dir<T, 16>(a, dir, ci, s); break;
    case 24: run_tile_dir<T, 24>(a, dir, ci, s); break;
    case 32: run_tile_dir<T, 32>(a, dir, ci, s); break;
    case 48: run_tile_dir<T, 48>(a, dir, ci, s); break;
    case 64: run_tile_dir<T, 64>(a, dir, ci, s); break;
    case 96: run_tile_dir<T, 96>(a, dir, ci, s); break;
    case 128: run_tile_dir<T, 128>(a, dir, ci, s); break;
    case 192: run_tile_dir<T, 192>(a, dir, ci, s); break;
    case 256: run_tile_dir<T, 256>(a, dir, ci, s); break;
    case 384: run_tile_dir<T, 384>(a, dir, ci, s); break;
    case 512: run_tile_dir<T, 512>(a, dir, ci, s); break;
    default: throw std::runtime_error("unsupported FFT length " + std::to_string(L));
  }
}
template void launch_tile_pass<float>(const TilePassArgs&, int, int, bool, hipStream_t);
template void launch_tile_pass<double>(const TilePassArgs&, int, int, bool, hipStream_t);

// ======================================================== x pass (real axis)
template <typename T> struct XTile { static constexpr int W = 1024; };
template <> struct XTile<double> { static constexpr int W = 512; };
__host__ __device__ constexpr int xtlx(int M, int W) { return (W / M) < 1 ? 1 : ((W / M) > 16 ? 16 : (W / M)); }
// threads of an x-pass block: one per butterfly of the fused stage, 64..256
__host__ __device__ constexpr int x_threads(int nb) { return nb >= 256 ? 256 : (nb <= 64 ? 64 : ((nb + 63) / 64) * 64); }

// z[m] = x[2m] + i x[2m+1] is transformed with a half-length complex FFT;
// the split/merge steps below follow from the even/odd decomposition of the
// length-nx DFT.  The imaginary parts of the kx = 0 and kx = nx/2 entries are
// ignored on input, as a C2R transform does.
//
// Structure of the RHS mode (per workgroup: TLX x-lines, all fields):
//   A  spectral lines -> LDS
//   B  form u, w spectra (w_y, w_z need i kx), pack, inverse stage 0 (regs)
//   C  inverse middle stages (LDS)
//   D  inverse last stage -> u x w - chi (u - us) [+ sponge] -> forward
//      stage 0, all in registers (PLAN_XI ends with the radix PLAN_XF starts with)
//   E  forward remaining stages (LDS), split into nx/2+1 entries, store
template <typename T, int M, int TLX, int NT, int MODE, int NFI, int NFO, bool SPONGE>
__global__ __launch_bounds__(NT) void x_pass_k(XPassArgs a) {
  to_member(a, blockIdx.y);
  constexpr int NST = plan_count_k(PLAN_XF, M);
  static_assert(NST >= 2, "x pass needs at least two stages");
  constexpr int RI0 = plan_radix_k(PLAN_XI, M, 0);   // first inverse radix
  constexpr int RM = plan_radix_k(PLAN_XF, M, 0);    // last inverse = first forward radix
  constexpr int LSP = line_stride<T>(M + 1);
  constexpr int NP = M / 2 + 1;
  constexpr int nxh = M + 1;
  extern __shared__ char smem_raw[];
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem_raw);
  __shared__ double red[NT / 64][4];
  const int tid = threadIdx.x;
  const int l0 = blockIdx.x * TLX;
  // global-address-space views of the arguments (see gptr)
  const gptr<const cplx<T>> twM = (gptr<const cplx<T>>)a.twM;
  const gptr<const cplx<T>> twN = (gptr<const cplx<T>>)a.twN;
  const gptr<const double> kxg = (gptr<const double>)a.kx;
  auto L_ = [&](int f, int line) { return buf + (f * TLX + line) * LSP; };

  if (MODE != XM_FWD) {
    // ---- A
    const gptr<const cplx<T>> in = (gptr<const cplx<T>>)a.in;
#pragma unroll
    for (int f = 0; f < NFI; ++f)
      for (int e = tid; e < TLX * nxh; e += NT) {
        const int line = e / nxh, k = e - line * nxh;
        const int gl = l0 + line;
        cplx<T> v = mk<T>(T(0), T(0));
        if (gl < a.nlines) v = gld<T>(in + f * a.fstride_in + (long)gl * a.prow + k);
        L_(f, line)[lp<T>(k)] = v;
      }
    __syncthreads();
    // ---- B: butterfly (f, line, j) of inverse stage 0 reads Z at j + q M/RI0
    constexpr int NB = NFI * TLX * (M / RI0), PER = (NB + NT - 1) / NT;
    cplx<T> v[PER][RI0];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int beta = tid + p * NT;
      if (NB % NT == 0 || beta < NB) {
        const int f = beta / (TLX * (M / RI0));
        const int rem = beta - f * (TLX * (M / RI0));
        const int line = rem / (M / RI0), j = rem - line * (M / RI0);
#pragma unroll
        for (int q = 0; q < RI0; ++q) {
          const int k = j + q * (M / RI0), k2 = M - k;
          cplx<T> xa, xb;
          if ((MODE == XM_RHS || MODE == XM_INV_UW) && f == 4) {
            // w_y = dz u_x - dx u_z = A - i kx V_z
            xa = sub(L_(4, line)[lp<T>(k)], ik_<T>((T)kxg[k], L_(2, line)[lp<T>(k)]));
            xb = sub(L_(4, line)[lp<T>(k2)], ik_<T>((T)kxg[k2], L_(2, line)[lp<T>(k2)]));
          } else if ((MODE == XM_RHS || MODE == XM_INV_UW) && f == 5) {
            // w_z = dx u_y - dy u_x = i kx V_y - B
            xa = sub(ik_<T>((T)kxg[k], L_(1, line)[lp<T>(k)]), L_(5, line)[lp<T>(k)]);
            xb = sub(ik_<T>((T)kxg[k2], L_(1, line)[lp<T>(k2)]), L_(5, line)[lp<T>(k2)]);
          } else {
            xa = L_(f, line)[lp<T>(k)];
            xb = L_(f, line)[lp<T>(k2)];
          }
          if (k == 0) { xa.y = T(0); xb.y = T(0); }
          const cplx<T> ea = conj_(twv<T>(twN, k));
          const cplx<T> s1 = add(xa, conj_(xb)), d1 = sub(xa, conj_(xb));
          v[p][q] = add(s1, ik_<T>(T(1), cmul(ea, d1)));
        }
        Bfly<RI0, +1, T>::run(v[p]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int beta = tid + p * NT;
      if (NB % NT == 0 || beta < NB) {
        const int f = beta / (TLX * (M / RI0));
        const int rem = beta - f * (TLX * (M / RI0));
        const int line = rem / (M / RI0), j = rem - line * (M / RI0);
#pragma unroll
        for (int q = 0; q < RI0; ++q) L_(f, line)[lp<T>(j * RI0 + q)] = v[p][q];
      }
    }
    __syncthreads();
    // ---- C
    Stages<T, M, NFI * TLX, NT, PLAN_XI, +1, LSP, 1, NST - 1, RI0>::run(buf, twM);
  }

  constexpr int NSD = M / RM;  // sub-transform length entering the last inverse stage

  if (MODE == XM_INV_UW || MODE == XM_INV) {
    // inverse last stage straight to the physical field
    const gptr<cplx<T>> out = (gptr<cplx<T>>)a.out;
    constexpr int NB = NFI * TLX * (M / RM), PER = (NB + NT - 1) / NT;
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int beta = tid + p * NT;
      if (NB % NT == 0 || beta < NB) {
        const int f = beta / (TLX * (M / RM));
        const int rem = beta - f * (TLX * (M / RM));
        const int line = rem / (M / RM), j = rem - line * (M / RM);
        cplx<T> w[RM];
#pragma unroll
        for (int q = 0; q < RM; ++q) w[q] = L_(f, line)[lp<T>(j + q * (M / RM))];
        twiddle_bfly<T, M, RM, +1, NSD>(w, j, twM);
        const int gl = l0 + line;
        if (gl < a.nlines) {
#pragma unroll
          for (int q = 0; q < RM; ++q) gst<T>(out + (f * a.fstride_out + (long)gl * a.nx) / 2 + j + q * (M / RM), w[q]);
        }
      }
    }
    return;
  }

  if (MODE == XM_FWD) {
    // physical pairs straight into forward stage 0
    const gptr<const cplx<T>> in = (gptr<const cplx<T>>)a.in;
    constexpr int NB = NFI * TLX * (M / RM), PER = (NB + NT - 1) / NT;
    cplx<T> v[PER][RM];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int beta = tid + p * NT;
      if (NB % NT == 0 || beta < NB) {
        const int f = beta / (TLX * (M / RM));
        const int rem = beta - f * (TLX * (M / RM));
        const int line = rem / (M / RM), j = rem - line * (M / RM);
        const int gl = l0 + line;
#pragma unroll
        for (int q = 0; q < RM; ++q)
          v[p][q] = (gl < a.nlines) ? gld<T>(in + (f * a.fstride_in + (long)gl * a.nx) / 2 + j + q * (M / RM)) : mk<T>(T(0), T(0));
        Bfly<RM, -1, T>::run(v[p]);
#pragma unroll
        for (int q = 0; q < RM; ++q) L_(f, line)[lp<T>(j * RM + q)] = v[p][q];
      }
    }
    __syncthreads();
  }

  if (MODE == XM_RHS) {
    // ---- D
    const gptr<const T> mask = (gptr<const T>)a.mask;
    const gptr<const T> us = (gptr<const T>)a.us;
    double umax2 = 0.0, fs[3] = {0.0, 0.0, 0.0}, fc[3] = {0.0, 0.0, 0.0};
    constexpr int NB = TLX * (M / RM), PER = (NB + NT - 1) / NT;
    cplx<T> nout[PER][NFO][RM];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int beta = tid + p * NT;
      if (NB % NT == 0 || beta < NB) {
        const int line = beta / (M / RM), j = beta - line * (M / RM);
        const int gl = l0 + line;
        cplx<T> z[6][RM];
#pragma unroll
        for (int f = 0; f < 6; ++f) {
#pragma unroll
          for (int q = 0; q < RM; ++q) z[f][q] = L_(f, line)[lp<T>(j + q * (M / RM))];
          twiddle_bfly<T, M, RM, +1, NSD>(z[f], j, twM);
        }
        const bool live = gl < a.nlines;
        const int iy = live ? gl % a.ny : 0, iz = live ? gl / a.ny : 0;
#pragma unroll
        for (int q = 0; q < RM; ++q) {
          const int pos = j + q * (M / RM);
          double nv[3][2], sv[3][2];
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const double ux = h ? z[0][q].y : z[0][q].x, uy = h ? z[1][q].y : z[1][q].x, uz = h ? z[2][q].y : z[2][q].x;
            const double wx = h ? z[3][q].y : z[3][q].x, wy = h ? z[4][q].y : z[4][q].x, wz = h ? z[5][q].y : z[5][q].x;
            const long pt = (long)gl * a.nx + 2 * pos + h;
            const double mk_ = live ? (double)mask[pt] : 0.0;
            double dux = ux, duy = uy, duz = uz;
            if (mk_ != 0.0) {
              dux -= (double)us[pt];
              duy -= (double)us[a.npts + pt];
              duz -= (double)us[2 * a.npts + pt];
            }
            const double chi = mk_ * a.eps_inv;
            const double pv[3] = {chi * dux, chi * duy, chi * duz};
            if (live) {
#pragma unroll
              for (int c = 0; c < 3; ++c) {
                const double t = fs[c] + pv[c];
                fc[c] += (fabs(fs[c]) >= fabs(pv[c])) ? (fs[c] - t) + pv[c] : (pv[c] - t) + fs[c];
                fs[c] = t;
              }
              umax2 = nanmax(umax2, ux * ux + uy * uy + uz * uz);
            }
            nv[0][h] = uy * wz - uz * wy - pv[0];
            nv[1][h] = uz * wx - ux * wz - pv[1];
            nv[2][h] = ux * wy - uy * wx - pv[2];
            if constexpr (SPONGE) {
              const int ix = 2 * pos + h, th = a.sponge_thickness;
              bool in_sp = false;
              if (a.sponge == 1) {
                if (a.nx > 4 && (ix <= th - 1 || ix >= a.nx - th)) in_sp = true;
                if (iy <= th - 1 || iy >= a.ny - th) in_sp = true;
                if (iz <= th - 1 || iz >= a.nz - th) in_sp = true;
              } else if (a.sponge == 2) {
                in_sp = ix >= a.nx - th;
              }
              in_sp = in_sp && (mk_ < a.sponge_mask_thr);
              sv[0][h] = in_sp ? -wx * a.sponge_eps_inv : 0.0;
              sv[1][h] = in_sp ? -wy * a.sponge_eps_inv : 0.0;
              sv[2][h] = in_sp ? -wz * a.sponge_eps_inv : 0.0;
            }
          }
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            nout[p][c][q] = mk<T>((T)nv[c][0], (T)nv[c][1]);
            if constexpr (SPONGE) nout[p][3 + c][q] = mk<T>((T)sv[c][0], (T)sv[c][1]);
          }
        }
#pragma unroll
        for (int c = 0; c < NFO; ++c) Bfly<RM, -1, T>::run(nout[p][c]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int beta = tid + p * NT;
      if (NB % NT == 0 || beta < NB) {
        const int line = beta / (M / RM), j = beta - line * (M / RM);
#pragma unroll
        for (int c = 0; c < NFO; ++c)
#pragma unroll
          for (int q = 0; q < RM; ++q) L_(c, line)[lp<T>(j * RM + q)] = nout[p][c][q];
      }
    }
    if (a.want_stats) {
      double v4[4] = {umax2, fs[0] + fc[0], fs[1] + fc[1], fs[2] + fc[2]};
      v4[0] = wave_max(v4[0]);
#pragma unroll
      for (int c = 1; c < 4; ++c) v4[c] = wave_sum(v4[c]);
      if ((tid & 63) == 0) {
#pragma unroll
        for (int c = 0; c < 4; ++c) red[tid >> 6][c] = v4[c];
      }
      __syncthreads();
      if (tid == 0) {
        double r[4] = {red[0][0], red[0][1], red[0][2], red[0][3]};
        for (int w = 1; w < NT / 64; ++w) {
          r[0] = nanmax(r[0], red[w][0]);
          for (int c = 1; c < 4; ++c) r[c] += red[w][c];
        }
        for (int c = 0; c < 4; ++c) ((gptr<double>)a.partial)[(long)blockIdx.x * 4 + c] = r[c];
      }
    } else {
      __syncthreads();
    }
  }

  // ---- E: remaining forward stages, then split into the nx/2+1 entries
  Stages<T, M, NFO * TLX, NT, PLAN_XF, -1, LSP, 1, NST, RM>::run(buf, twM);
  const gptr<cplx<T>> out = (gptr<cplx<T>>)a.out;
  const T half = T(0.5);
  for (int e = tid; e < TLX * NP; e += NT) {
    const int line = e / NP, k = e - line * NP, k2 = M - k;
    const int gl = l0 + line;
    if (gl >= a.nlines) continue;
    const cplx<T> wa = twv<T>(twN, k), wb = twv<T>(twN, k2);
#pragma unroll
    for (int f = 0; f < NFO; ++f) {
      const cplx<T> zk = L_(f, line)[lp<T>(k)];
      const cplx<T> zm = L_(f, line)[lp<T>(k2 == M ? 0 : k2)];
      const cplx<T> s1 = add(zk, conj_(zm)), d1 = sub(zk, conj_(zm));
      const gptr<cplx<T>> o = out + f * a.fstride_out + (long)gl * a.prow;
      gst<T>(o + k, scl(sub(s1, ik_<T>(T(1), cmul(wa, d1))), half));
      if (k2 != k) {
        const cplx<T> s2 = add(zm, conj_(zk)), d2 = sub(zm, conj_(zk));
        gst<T>(o + k2, scl(sub(s2, ik_<T>(T(1), cmul(wb, d2))), half));
      }
    }
  }
}

int x_pass_tlx(int nx, bool fp64) { return xtlx(nx / 2, fp64 ? 512 : 1024); }
int x_pass_blocks(int nx, int nlines, bool fp64) {
  const int t = x_pass_tlx(nx, fp64);
  return (nlines + t - 1) / t;
}

template <typename T, int M, int MODE, int NFI, int NFO, bool SP>
static void run_x(const XPassArgs& a, hipStream_t s) {
  constexpr int TLX = xtlx(M, XTile<T>::W);
  constexpr int NT = x_threads(TLX * (M / plan_radix_k(PLAN_XF, M, 0)));
  constexpr int NF = NFI > NFO ? NFI : NFO;
  const size_t smem = sizeof(cplx<T>) * (size_t)NF * TLX * line_stride<T>(M + 1);
  static bool init = false;
  if (!init) { set_smem(x_pass_k<T, M, TLX, NT, MODE, NFI, NFO, SP>, smem); init = true; }
  const int blocks = (a.nlines + TLX - 1) / TLX;
  hipLaunchKernelGGL((x_pass_k<T, M, TLX, NT, MODE, NFI, NFO, SP>), dim3(blocks, batch_n(a.bat)), dim3(NT), smem, s, a);
}

template <typename T, int M>
static void run_x_mode(const XPassArgs& a, int mode, hipStream_t s) {
  switch (mode) {
    case XM_RHS:
      if (a.sponge) run_x<T, M, XM_RHS, 6, 6, true>(a, s);
      else run_x<T, M, XM_RHS, 6, 3, false>(a, s);
      break;
    case XM_INV_UW: run_x<T, M, XM_INV_UW, 6, 6, false>(a, s); break;
    case XM_INV:
      if (a.nf == 1) run_x<T, M, XM_INV, 1, 1, false>(a, s);
      else if (a.nf == 3) run_x<T, M, XM_INV, 3, 3, false>(a, s);
      else throw std::runtime_error("x inverse pass: nf must be 1 or 3");
      break;
    case XM_FWD:
      if (a.nf == 1) run_x<T, M, XM_FWD, 1, 1, false>(a, s);
      else if (a.nf == 3) run_x<T, M, XM_FWD, 3, 3, false>(a, s);
      else throw std::runtime_error("x forward pass: nf must be 1 or 3");
      break;
    default: throw std::runtime_error("bad x pass mode");
  }
}

template <typename T>
void launch_x_pass(const XPassArgs& a, int mode, hipStream_t s) {
  switch (a.nx / 2) {
    case 8: run_x_mode<T, 8>(a, mode, s); break;
    case 16: run_x_mode<T, 16>(a, mode, s); break;
    case 24: run_x_mode<T, 24>(a, mode, s); break;
    case 32: run_x_mode<T, 32>(a, mode, s); break;
    case 48: run_x_mode<T, 48>(a, mode, s); break;
    case 64: run_x_mode<T, 64>(a, mode, s); break;
    case 96: run_x_mode<T, 96>(a, mode, s); break;
    case 128: run_x_mode<T, 128>(a, mode, s); break;
    case 192: run_x_mode<T, 192>(a, mode, s); break;
    case 256: run_x_mode<T, 256>(a, mode, s); break;
    default: throw std::runtime_error("unsupported nx " + std::to_string(a.nx));
  }
}
template void launch_x_pass<float>(const XPassArgs&, int, hipStream_t);
template void launch_x_pass<double>(const XPassArgs&, int, hipStream_t);


}  // namespace flusi
