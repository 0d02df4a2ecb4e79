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
#include <stdexcept>
#include <string>

namespace flusi {

template <typename T> struct Tile { static constexpr int TL = 16; };
template <> struct Tile<double> { static constexpr int TL = 8; };

__host__ __device__ constexpr int pick_nt(int n) {
  return n % 512 == 0 ? 512 : (n % 256 == 0 ? 256 : (n % 128 == 0 ? 128 : 64));
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
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// ======================================================== generic tile pass
template <typename T, int L, int TL, int NT, int S, bool CONTIG_IN>
__global__ __launch_bounds__(NT) void tile_pass_k(TilePassArgs a) {
  constexpr int LS = L + 1;
  constexpr int EPT = TL * L / NT;
  static_assert(TL * L % NT == 0, "tile must divide the block");
  extern __shared__ char smem_raw[];
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int b0 = blockIdx.x * TL;
  const long c = blockIdx.y;
  const cplx<T>* tw = reinterpret_cast<const cplx<T>*>(a.tw);
  const T sc = (T)a.scale;
  for (int j = 0; j < a.njobs; ++j) {
    const cplx<T>* in0 = reinterpret_cast<const cplx<T>*>(a.jobs[j].in0);
    const cplx<T>* in1 = reinterpret_cast<const cplx<T>*>(a.jobs[j].in1);
    cplx<T>* out = reinterpret_cast<cplx<T>*>(a.jobs[j].out);
    const int op = a.jobs[j].op;
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const int e = tid + i * NT;
      int aa, bb;
      if (CONTIG_IN) { aa = e % L; bb = e / L; } else { bb = e % TL; aa = e / TL; }
      cplx<T> v = mk<T>(T(0), T(0));
      if (b0 + bb < a.nb) {
        const long idx = c * a.sin.c + (long)(b0 + bb) * a.sin.b + (long)aa * a.sin.a;
        v = in0[idx];
        if (op != OP_COPY) {
          v = ik_<T>((T)a.kline[aa], v);
          if (op == OP_DERIV_MINUS) v = sub(v, in1[idx]);
        }
      }
      buf[bb * LS + aa] = v;
    }
    __syncthreads();
    block_fft<T, L, TL, NT, S>(buf, LS, tw);
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const int e = tid + i * NT;
      int aa, bb;
      if (CONTIG_IN) { bb = e % TL; aa = e / TL; } else { aa = e % L; bb = e / L; }
      if (b0 + bb < a.nb) {
        const long idx = c * a.sout.c + (long)(b0 + bb) * a.sout.b + (long)aa * a.sout.a;
        const cplx<T> v = buf[bb * LS + aa];
        out[idx] = mk<T>(v.x * sc, v.y * sc);
      }
    }
    __syncthreads();
  }
}

template <typename T, int L, int S, bool CI>
static void run_tile(const TilePassArgs& a, hipStream_t s) {
  constexpr int TL = Tile<T>::TL;
  constexpr int NT = pick_nt(TL * L);
  const size_t smem = sizeof(cplx<T>) * TL * (L + 1);
  static bool init = false;
  if (!init) { set_smem(tile_pass_k<T, L, TL, NT, S, CI>, smem); init = true; }
  dim3 grid((a.nb + TL - 1) / TL, a.nc);
  hipLaunchKernelGGL((tile_pass_k<T, L, TL, NT, S, CI>), grid, dim3(NT), smem, s, a);
}

template <typename T, int L>
static void run_tile_dir(const TilePassArgs& a, int dir, bool ci, hipStream_t s) {
  if (dir > 0) { if (ci) run_tile<T, L, +1, true>(a, s); else run_tile<T, L, +1, false>(a, s); }
  else { if (ci) run_tile<T, L, -1, true>(a, s); else run_tile<T, L, -1, false>(a, s); }
}

bool fft_length_supported(int n) {
  switch (n) {
    case 8: case 16: case 24: case 32: case 48: case 64: case 96: case 128:
    case 192: case 256: case 384: case 512: return true;
    default: return false;
  }
}

template <typename T>
void launch_tile_pass(const TilePassArgs& a, int L, int dir, bool ci, hipStream_t s) {
  switch (L) {
    case 8: run_tile_dir<T, 8>(a, dir, ci, s); break;
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

// z[m] = x[2m] + i x[2m+1] is transformed with a half-length complex FFT;
// the split/merge steps below follow from the even/odd decomposition of the
// length-nx DFT.  The imaginary parts of the kx = 0 and kx = nx/2 entries are
// ignored on input, as a C2R transform does.
template <typename T, int M, int TLX, int NT, int MODE, int NFI, int NFO, bool SPONGE>
__global__ __launch_bounds__(NT) void x_pass_k(XPassArgs a) {
  constexpr int LSX = (M + 1) | 1;
  constexpr int NP = M / 2 + 1;
  constexpr int nxh = M + 1;
  extern __shared__ char smem_raw[];
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem_raw);
  __shared__ double red[NT / 64][4];
  const int tid = threadIdx.x;
  const int l0 = blockIdx.x * TLX;
  const cplx<T>* twM = reinterpret_cast<const cplx<T>*>(a.twM);
  const cplx<T>* twN = reinterpret_cast<const cplx<T>*>(a.twN);

  if (MODE != XM_FWD) {
    const cplx<T>* in = reinterpret_cast<const cplx<T>*>(a.in);
#pragma unroll
    for (int f = 0; f < NFI; ++f)
      for (int e = tid; e < TLX * nxh; e += NT) {
        const int line = e / nxh, k = e - line * nxh;
        const int gl = l0 + line;
        cplx<T> v = mk<T>(T(0), T(0));
        if (gl < a.nlines) v = in[f * a.fstride_in + (long)gl * nxh + k];
        buf[(f * TLX + line) * LSX + k] = v;
      }
    __syncthreads();
    for (int e = tid; e < TLX * NP; e += NT) {
      const int line = e / NP, k = e - line * NP, k2 = M - k;
      cplx<T> xa[NFI], xb[NFI];
#pragma unroll
      for (int f = 0; f < NFI; ++f) {
        xa[f] = buf[(f * TLX + line) * LSX + k];
        xb[f] = buf[(f * TLX + line) * LSX + k2];
      }
      if (MODE == XM_RHS || MODE == XM_INV_UW) {
        // fields: 0 Vx, 1 Vy, 2 Vz, 3 Wx, 4 A = dz u_x, 5 B = dy u_x
        const T ka = (T)a.kx[k], kb = (T)a.kx[k2];
        const cplx<T> wya = sub(xa[4], ik_<T>(ka, xa[2])), wza = sub(ik_<T>(ka, xa[1]), xa[5]);
        const cplx<T> wyb = sub(xb[4], ik_<T>(kb, xb[2])), wzb = sub(ik_<T>(kb, xb[1]), xb[5]);
        xa[4] = wya; xa[5] = wza; xb[4] = wyb; xb[5] = wzb;
      }
      if (k == 0) {
#pragma unroll
        for (int f = 0; f < NFI; ++f) { xa[f].y = T(0); xb[f].y = T(0); }
      }
      const cplx<T> ea = conj_(twN[k]), eb = conj_(twN[k2]);
#pragma unroll
      for (int f = 0; f < NFI; ++f) {
        const cplx<T> s1 = add(xa[f], conj_(xb[f])), d1 = sub(xa[f], conj_(xb[f]));
        const cplx<T> za = add(s1, ik_<T>(T(1), cmul(ea, d1)));
        buf[(f * TLX + line) * LSX + k] = za;
        if (k2 < M && k2 != k) {
          const cplx<T> s2 = add(xb[f], conj_(xa[f])), d2 = sub(xb[f], conj_(xa[f]));
          buf[(f * TLX + line) * LSX + k2] = add(s2, ik_<T>(T(1), cmul(eb, d2)));
        }
      }
    }
    __syncthreads();
    block_fft<T, M, NFI * TLX, NT, +1>(buf, LSX, twM);
  }

  if (MODE == XM_INV_UW || MODE == XM_INV) {
    cplx<T>* out = reinterpret_cast<cplx<T>*>(a.out);
#pragma unroll
    for (int f = 0; f < NFI; ++f)
      for (int e = tid; e < TLX * M; e += NT) {
        const int line = e / M, m = e - line * M;
        const int gl = l0 + line;
        if (gl < a.nlines)
          out[(f * a.fstride_out + (long)gl * a.nx) / 2 + m] = buf[(f * TLX + line) * LSX + m];
      }
    return;
  }

  if (MODE == XM_FWD) {
    const cplx<T>* in = reinterpret_cast<const cplx<T>*>(a.in);
#pragma unroll
    for (int f = 0; f < NFI; ++f)
      for (int e = tid; e < TLX * M; e += NT) {
        const int line = e / M, m = e - line * M;
        const int gl = l0 + line;
        cplx<T> v = mk<T>(T(0), T(0));
        if (gl < a.nlines) v = in[(f * a.fstride_in + (long)gl * a.nx) / 2 + m];
        buf[(f * TLX + line) * LSX + m] = v;
      }
    __syncthreads();
  }

  if (MODE == XM_RHS) {
    const T* mask = reinterpret_cast<const T*>(a.mask);
    const T* us = reinterpret_cast<const T*>(a.us);
    double umax2 = 0.0, fs[3] = {0.0, 0.0, 0.0}, fc[3] = {0.0, 0.0, 0.0};
    for (int e = tid; e < TLX * M; e += NT) {
      const int line = e / M, m = e - line * M;
      const int gl = l0 + line;
      if (gl >= a.nlines) continue;
      cplx<T> z[6];
#pragma unroll
      for (int f = 0; f < 6; ++f) z[f] = buf[(f * TLX + line) * LSX + m];
      const long p0 = (long)gl * a.nx + 2 * m;
      const int iy = gl % a.ny, iz = gl / a.ny;
      double nout[3][2], sout[3][2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const double ux = h ? z[0].y : z[0].x, uy = h ? z[1].y : z[1].x, uz = h ? z[2].y : z[2].x;
        const double wx = h ? z[3].y : z[3].x, wy = h ? z[4].y : z[4].x, wz = h ? z[5].y : z[5].x;
        const double mk_ = (double)mask[p0 + h];
        double dux = ux, duy = uy, duz = uz;
        if (mk_ != 0.0) {
          dux -= (double)us[p0 + h];
          duy -= (double)us[a.npts + p0 + h];
          duz -= (double)us[2 * a.npts + p0 + h];
        }
        const double chi = mk_ * a.eps_inv;
        const double px = chi * dux, py = chi * duy, pz = chi * duz;
        // Neumaier-compensated running sums of the penalisation term
        const double pv[3] = {px, py, pz};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double t = fs[c] + pv[c];
          fc[c] += (fabs(fs[c]) >= fabs(pv[c])) ? (fs[c] - t) + pv[c] : (pv[c] - t) + fs[c];
          fs[c] = t;
        }
        umax2 = fmax(umax2, ux * ux + uy * uy + uz * uz);
        nout[0][h] = uy * wz - uz * wy - px;
        nout[1][h] = uz * wx - ux * wz - py;
        nout[2][h] = ux * wy - uy * wx - pz;
        if (SPONGE) {
          const int ix = 2 * m + h;
          bool in_sp = false;
          const int th = a.sponge_thickness;
          if (a.sponge == 1) {
            if (a.nx > 4 && (ix <= th - 1 || ix >= a.nx - th)) in_sp = true;
            if (iy <= th - 1 || iy >= a.ny - th) in_sp = true;
            if (iz <= th - 1 || iz >= a.nz - th) in_sp = true;
          } else if (a.sponge == 2) {
            in_sp = ix >= a.nx - th;
          }
          in_sp = in_sp && (mk_ < a.sponge_mask_thr);
          sout[0][h] = in_sp ? -wx * a.sponge_eps_inv : 0.0;
          sout[1][h] = in_sp ? -wy * a.sponge_eps_inv : 0.0;
          sout[2][h] = in_sp ? -wz * a.sponge_eps_inv : 0.0;
        }
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        buf[(c * TLX + line) * LSX + m] = mk<T>((T)nout[c][0], (T)nout[c][1]);
        if (SPONGE) buf[((3 + c) * TLX + line) * LSX + m] = mk<T>((T)sout[c][0], (T)sout[c][1]);
      }
    }
    // block reduction of the per-thread statistics
    double v4[4] = {umax2, fs[0] + fc[0], fs[1] + fc[1], fs[2] + fc[2]};
    v4[0] = wave_max(v4[0]);
#pragma unroll
    for (int c = 1; c < 4; ++c) v4[c] = wave_sum(v4[c]);
    const int wv = tid >> 6;
    if ((tid & 63) == 0) {
#pragma unroll
      for (int c = 0; c < 4; ++c) red[wv][c] = v4[c];
    }
    __syncthreads();
    if (tid == 0) {
      double r[4] = {red[0][0], red[0][1], red[0][2], red[0][3]};
      for (int w = 1; w < NT / 64; ++w) {
        r[0] = fmax(r[0], red[w][0]);
        for (int c = 1; c < 4; ++c) r[c] += red[w][c];
      }
      for (int c = 0; c < 4; ++c) a.partial[(long)blockIdx.x * 4 + c] = r[c];
    }
  }

  // forward half-length FFT and split into the nx/2+1 spectral entries
  block_fft<T, M, NFO * TLX, NT, -1>(buf, LSX, twM);
  cplx<T>* out = reinterpret_cast<cplx<T>*>(a.out);
  const T half = T(0.5);
  for (int e = tid; e < TLX * NP; e += NT) {
    const int line = e / NP, k = e - line * NP, k2 = M - k;
    const int gl = l0 + line;
    if (gl >= a.nlines) continue;
    const cplx<T> wa = twN[k], wb = twN[k2];
#pragma unroll
    for (int f = 0; f < NFO; ++f) {
      const cplx<T> zk = buf[(f * TLX + line) * LSX + k];
      const cplx<T> zm = buf[(f * TLX + line) * LSX + (k2 == M ? 0 : k2)];
      const cplx<T> s1 = add(zk, conj_(zm)), d1 = sub(zk, conj_(zm));
      const cplx<T> xk = scl(sub(s1, ik_<T>(T(1), cmul(wa, d1))), half);
      cplx<T>* o = out + f * a.fstride_out + (long)gl * nxh;
      o[k] = xk;
      if (k2 != k) {
        const cplx<T> s2 = add(zm, conj_(zk)), d2 = sub(zm, conj_(zk));
        o[k2] = scl(sub(s2, ik_<T>(T(1), cmul(wb, d2))), half);
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
  constexpr int NT = 256;
  constexpr int NF = NFI > NFO ? NFI : NFO;
  const size_t smem = sizeof(cplx<T>) * (size_t)NF * TLX * ((M + 1) | 1);
  static bool init = false;
  if (!init) { set_smem(x_pass_k<T, M, TLX, NT, MODE, NFI, NFO, SP>, smem); init = true; }
  const int blocks = (a.nlines + TLX - 1) / TLX;
  hipLaunchKernelGGL((x_pass_k<T, M, TLX, NT, MODE, NFI, NFO, SP>), dim3(blocks), dim3(NT), smem, s, a);
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

// ======================================================== z forward + epilogue
__device__ __forceinline__ double vis_factor(double dtv, double nu, double k2, bool alias) {
  if (dtv < 0.0 || alias) return 0.0;
  return exp(-dtv * nu * k2);
}

template <typename T, int L, int TL, int NT, bool SPONGE>
__global__ __launch_bounds__(NT) void zstep_k(ZStepArgs a) {
  constexpr int LS = L + 1;
  constexpr int EPT = TL * L / NT;
  constexpr int NF = SPONGE ? 6 : 3;
  extern __shared__ char smem_raw[];
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem_raw);
  __shared__ double red[NT / 64];
  const int tid = threadIdx.x;
  const int b0 = blockIdx.x * TL;
  const int kx = blockIdx.y;
  const cplx<T>* tw = reinterpret_cast<const cplx<T>*>(a.tw);
  const cplx<T>* nq = reinterpret_cast<const cplx<T>*>(a.nq);
  cplx<T> r[NF][EPT];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const int e = tid + i * NT;
      const int bb = e % TL, aa = e / TL;
      cplx<T> v = mk<T>(T(0), T(0));
      if (b0 + bb < a.ny) v = nq[f * a.fs + ((long)kx * a.nz + aa) * a.ny + b0 + bb];
      buf[bb * LS + aa] = v;
    }
    __syncthreads();
    block_fft<T, L, TL, NT, -1>(buf, LS, tw);
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const int e = tid + i * NT;
      const int aa = e % L, bb = e / L;
      r[f][i] = buf[bb * LS + aa];
    }
    __syncthreads();
  }

  const double kxv = a.kx[kx];
  const double kxt2 = a.kxt2[kx];
  const double wE = (kx == 0 || kx == a.nx / 2) ? 0.5 : 1.0;
  double esum = 0.0, ecomp = 0.0;
  cplx<T>* U = reinterpret_cast<cplx<T>*>(a.u);
  cplx<T>* UST = reinterpret_cast<cplx<T>*>(a.ustage);
  const cplx<T>* UD = reinterpret_cast<const cplx<T>*>(a.udiff);
  cplx<T>* ACC = reinterpret_cast<cplx<T>*>(a.acc);
  cplx<T>* R0 = reinterpret_cast<cplx<T>*>(a.r0);
  const cplx<T>* R1 = reinterpret_cast<const cplx<T>*>(a.r1);
  cplx<T>* RO = reinterpret_cast<cplx<T>*>(a.rhs_out);
  const bool final_stage = a.mode == SM_EULER || a.mode == SM_AB2 || a.mode == SM_RK2_2 || a.mode == SM_RK4_4;
#pragma unroll
  for (int i = 0; i < EPT; ++i) {
    const int e = tid + i * NT;
    const int kzi = e % L, bb = e / L;
    const int ky = b0 + bb;
    if (ky >= a.ny) continue;
    const long si = ((long)kx * a.ny + ky) * a.nz + kzi;
    const double kyv = a.ky[ky], kzv = a.kz[kzi];
    const double k2 = (kxv * kxv + kyv * kyv) + kzv * kzv;
    const bool alias = ((kxt2 + a.kyt2[ky]) + a.kzt2[kzi]) >= 1.0;
    double nr[3], ni[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) { nr[c] = (double)r[c][i].x * a.scale; ni[c] = (double)r[c][i].y * a.scale; }
    if (SPONGE) {
      // sponge.f90:22-83: i k x (w/k^2), zero at k = 0
      if (k2 != 0.0) {
        double sr[3], sm[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) { sr[c] = (double)r[3 + c][i].x * a.scale / k2; sm[c] = (double)r[3 + c][i].y * a.scale / k2; }
        const double cr0 = kyv * sr[2] - kzv * sr[1], ci0 = kyv * sm[2] - kzv * sm[1];
        const double cr1 = kzv * sr[0] - kxv * sr[2], ci1 = kzv * sm[0] - kxv * sm[2];
        const double cr2 = kxv * sr[1] - kyv * sr[0], ci2 = kxv * sm[1] - kyv * sm[0];
        nr[0] += -ci0; ni[0] += cr0;
        nr[1] += -ci1; ni[1] += cr1;
        nr[2] += -ci2; ni[2] += cr2;
      }
    }
    if (kx == 0 && ky == 0 && kzi == 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c)
        if (a.meanflow_dynamic[c]) { nr[c] = -a.pen_sum[c] * a.dynamic_scale; ni[c] = 0.0; }
    }
    if (a.explicit_diffusion) {
      // rhs.f90:441-468
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const cplx<T> us_ = UD[c * a.fs + si];
        nr[c] -= a.nu * k2 * (double)us_.x;
        ni[c] -= a.nu * k2 * (double)us_.y;
      }
    }
    if (k2 != 0.0) {
      // rhs.f90:575-615
      const double qr = (kxv * nr[0] + kyv * nr[1] + kzv * nr[2]) / k2;
      const double qi = (kxv * ni[0] + kyv * ni[1] + kzv * ni[2]) / k2;
      nr[0] -= kxv * qr; ni[0] -= kxv * qi;
      nr[1] -= kyv * qr; ni[1] -= kyv * qi;
      nr[2] -= kzv * qr; ni[2] -= kzv * qi;
    }
    double unr[3] = {0, 0, 0}, uni[3] = {0, 0, 0};
    switch (a.mode) {
      case SM_RHS_ONLY:
#pragma unroll
        for (int c = 0; c < 3; ++c) RO[c * a.fs + si] = mk<T>((T)nr[c], (T)ni[c]);
        break;
      case SM_EULER: {
        // fluid_time_step.f90:718-727
        const double E = vis_factor(a.vis_dt_new, a.nu, k2, a.dealias_on && alias);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const cplx<T> u0 = U[c * a.fs + si];
          unr[c] = ((double)u0.x + a.dt * nr[c]) * E;
          uni[c] = ((double)u0.y + a.dt * ni[c]) * E;
          U[c * a.fs + si] = mk<T>((T)unr[c], (T)uni[c]);
          R0[c * a.fs + si] = mk<T>((T)(nr[c] * E), (T)(ni[c] * E));
        }
        break;
      }
      case SM_AB2: {
        // fluid_time_step.f90:802-811
        const double E = vis_factor(a.vis_dt_new, a.nu, k2, a.dealias_on && alias);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const cplx<T> u0 = U[c * a.fs + si];
          const cplx<T> n1 = R1[c * a.fs + si];
          unr[c] = ((double)u0.x + a.b10 * nr[c] + a.b11 * (double)n1.x) * E;
          uni[c] = ((double)u0.y + a.b10 * ni[c] + a.b11 * (double)n1.y) * E;
          U[c * a.fs + si] = mk<T>((T)unr[c], (T)uni[c]);
          R0[c * a.fs + si] = mk<T>((T)(nr[c] * E), (T)(ni[c] * E));
        }
        break;
      }
      case SM_RK2_1: {
        // fluid_time_step.f90:554-575 (N0 weighted with the IF of the previous dt)
        const double Eo = vis_factor(a.vis_dt_old, a.nu, k2, a.dealias_on && alias);
        const double En = vis_factor(a.vis_dt_new, a.nu, k2, a.dealias_on && alias);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double n0r = nr[c] * Eo, n0i = ni[c] * Eo;
          const cplx<T> u0 = U[c * a.fs + si];
          R0[c * a.fs + si] = mk<T>((T)n0r, (T)n0i);
          U[c * a.fs + si] = mk<T>((T)((double)u0.x * En + a.dt * n0r), (T)((double)u0.y * En + a.dt * n0i));
        }
        break;
      }
      case SM_RK2_2: {
        // fluid_time_step.f90:594-597
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const cplx<T> u0 = U[c * a.fs + si];
          const cplx<T> n0 = R0[c * a.fs + si];
          unr[c] = (double)u0.x + 0.5 * a.dt * (-(double)n0.x + nr[c]);
          uni[c] = (double)u0.y + 0.5 * a.dt * (-(double)n0.y + ni[c]);
        }
        break;
      }
      default: {
        // RK4 (fluid_time_step.f90:636-669) with an accumulator register
        const double kill = alias ? 0.0 : 1.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double kr = nr[c] * kill, ki = ni[c] * kill;
          if (a.mode == SM_RK4_1) {
            const cplx<T> u0 = U[c * a.fs + si];
            ACC[c * a.fs + si] = mk<T>((T)((double)u0.x + (a.dt / 6.0) * kr), (T)((double)u0.y + (a.dt / 6.0) * ki));
            UST[c * a.fs + si] = mk<T>((T)((double)u0.x + 0.5 * a.dt * kr), (T)((double)u0.y + 0.5 * a.dt * ki));
          } else if (a.mode == SM_RK4_2 || a.mode == SM_RK4_3) {
            const cplx<T> u0 = U[c * a.fs + si];
            const cplx<T> ac = ACC[c * a.fs + si];
            const double cst = (a.mode == SM_RK4_2) ? 0.5 * a.dt : a.dt;
            ACC[c * a.fs + si] = mk<T>((T)((double)ac.x + (a.dt / 3.0) * kr), (T)((double)ac.y + (a.dt / 3.0) * ki));
            UST[c * a.fs + si] = mk<T>((T)((double)u0.x + cst * kr), (T)((double)u0.y + cst * ki));
          } else {
            const cplx<T> ac = ACC[c * a.fs + si];
            unr[c] = ((double)ac.x + (a.dt / 6.0) * kr) * kill;
            uni[c] = ((double)ac.y + (a.dt / 6.0) * ki) * kill;
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
          if (a.set_mean[c]) { unr[c] = a.mean_value[c]; uni[c] = 0.0; }
      }
      if (a.mode == SM_RK2_2 || a.mode == SM_RK4_4 || (kx == 0 && ky == 0 && kzi == 0)) {
#pragma unroll
        for (int c = 0; c < 3; ++c) U[c * a.fs + si] = mk<T>((T)unr[c], (T)uni[c]);
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
  if (final_stage && a.energy_partial) {
    double v = wave_sum(esum + ecomp);
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
      for (int w = 0; w < NT / 64; ++w) s += red[w];
      a.energy_partial[(long)blockIdx.y * gridDim.x + blockIdx.x] = s;
    }
  }
}

int zstep_blocks(int nz, int ny, int nxh, bool fp64) {
  const int TL = fp64 ? Tile<double>::TL : Tile<float>::TL;
  return ((ny + TL - 1) / TL) * nxh;
}

template <typename T, int L, bool SP>
static void run_z(const ZStepArgs& a, hipStream_t s) {
  constexpr int TL = Tile<T>::TL;
  constexpr int NT = pick_nt(TL * L);
  const size_t smem = sizeof(cplx<T>) * TL * (L + 1);
  static bool init = false;
  if (!init) { set_smem(zstep_k<T, L, TL, NT, SP>, smem); init = true; }
  dim3 grid((a.ny + TL - 1) / TL, a.nxh);
  hipLaunchKernelGGL((zstep_k<T, L, TL, NT, SP>), grid, dim3(NT), smem, s, a);
}
template <typename T, int L>
static void run_z_sp(const ZStepArgs& a, hipStream_t s) {
  if (a.sponge) run_z<T, L, true>(a, s); else run_z<T, L, false>(a, s);
}

template <typename T>
void launch_zstep(const ZStepArgs& a, hipStream_t s) {
  switch (a.nz) {
    case 8: run_z_sp<T, 8>(a, s); break;
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
