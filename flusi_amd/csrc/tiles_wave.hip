// y and z tile passes (ZINV, YINV, YFWD, plain z forward) with every line
// transformed inside one wavefront, for line lengths 64, 128, 256 and 512.
//
// A workgroup owns TL lines (16 FP32 / 8 FP64: 128-byte segments on the
// strided side) and one LDS row per line.  A line of L points lives in the
// registers of G = L/8 lanes, lane l holding positions l + qG (q = 0..7)
// before and after the transform; the radix-8/4 Stockham stages exchange
// through the line's own LDS row with wave-scope ordering only.  The block
// waits twice per transform: once to turn the tile around between the
// contiguous side (a wave's own line) and the 128-byte segments of TL lines,
// and once before the tile is reused.  The next job's loads are issued as
// soon as the current operands are consumed.
#include "fft_core.h"
#include "flusi_kernels.h"
#include "pass_common.h"
#include "wave_fft.h"
#include <type_traits>
#include <stdexcept>
#include <string>

namespace flusi {

namespace {

using namespace wavefft;

template <typename T> struct TWTile { static constexpr int TL = 16; };
template <> struct TWTile<double> { static constexpr int TL = 8; };

template <typename T, int L> struct TWShape {
  static constexpr int TL = TWTile<T>::TL;
  static constexpr int NW = TL / TWPlan<L>::LPW;     // waves per block
  static constexpr int NT = NW * 64;
  static constexpr int LS = line_stride<T>(L);
  static constexpr int K = TL * L / NT;              // tile elements per thread (8)
  // twiddles in LDS when two blocks still fit in 160 KB
  static constexpr bool TWL = 2 * (TL * LS + L) * (int)sizeof(cplx<T>) <= 160 * 1024;
  static constexpr size_t smem = (size_t)(TL * LS + (TWL ? L : 0)) * sizeof(cplx<T>);
};

template <typename T, int L>
__device__ __forceinline__ cplx<T> apply_op(int op, double kscale, int aa, cplx<T> x0, cplx<T> x1) {
  if (op == OP_COPY) return x0;
  const T k = (T)(kscale * double(aa < L / 2 ? aa : aa - L));
  cplx<T> x = mk<T>(-k * x0.y, k * x0.x);
  if (op == OP_DERIV_MINUS) x = sub(x, x1);
  return x;
}

// SF: every store of the grid lands on a row below the store bound (the host
// checks it), so that the stores need no guard.  Loads and stores then carry
// no branch, and the wait for the next job's prefetched operands does not
// have to cover the current job's stores (with a guard, the compiler cannot
// count the stores in flight and waits for all of them).
template <typename T, int L, int S, bool CONTIG_IN, bool SF>
__global__ __launch_bounds__((TWShape<T, L>::NT)) void tile_wave_k(TilePassArgs a) {
  using Sh = TWShape<T, L>;
  constexpr int TL = Sh::TL, NT = Sh::NT, LS = Sh::LS, K = Sh::K;
  constexpr int G = TWPlan<L>::G;
  extern __shared__ char smem_raw[];
  cplx<T>* tile = reinterpret_cast<cplx<T>*>(smem_raw);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int line = wv * TWPlan<L>::LPW + lane / G, l = lane % G;
  cplx<T>* ln = tile + line * LS;
  const int b0 = blockIdx.x * TL;
  const long c = blockIdx.y;
  const int nbs = a.nb_store > a.nb ? a.nb_store : a.nb;  // lines >= nb hold zeros
  const long moff = (long)blockIdx.z * a.bat.mstride;  // batch member's slab
  const T sc = (T)a.scale;
  // twiddles: the block's LDS copy, or global memory where it does not fit
  std::conditional_t<Sh::TWL, const cplx<T>*, gptr<const cplx<T>>> tw;
  if constexpr (Sh::TWL) {
    cplx<T>* twl = tile + TL * LS;
    const gptr<const cplx<T>> twg = (gptr<const cplx<T>>)mshift(a.tw, moff);
    for (int i = tid; i < L; i += NT) twl[i] = gld<T>(twg + i);
    tw = twl;
    __syncthreads();
  } else {
    tw = (gptr<const cplx<T>>)mshift(a.tw, moff);
  }

  // operands: CONTIG_IN - the lane's 8 positions of its own line;
  // otherwise - the thread's K tile elements (bb = e % TL, a = e / TL)
  // (lines beyond nb read line nb - 1, unused: zeros are taken where the
  // operands are consumed, so that neither the loads nor a select right
  // after them needs a branch or a wait)
  cplx<T> r0[8], r1[8];
  auto load = [&](int jb, int which) {
    // global-address-space views (see gptr): the prefetched loads then do
    // not hold up the transform's LDS waits
    const gptr<const cplx<T>> in = (gptr<const cplx<T>>)mshift(which ? a.jobs[jb].in1 : a.jobs[jb].in0, moff);
    if (CONTIG_IN) {
      const bool ok = b0 + line < a.nb;
      const long base = c * a.sin.c + (long)(ok ? b0 + line : a.nb - 1) * a.sin.b + l;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const cplx<T> x = lds<T>(in + base + q * G);
        if (which) r1[q] = x; else r0[q] = x;
      }
    } else {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int e = tid + k * NT, bb = e % TL, aa = e / TL;
        const bool ok = b0 + bb < a.nb;
        const cplx<T> x = lds<T>(in + c * a.sin.c + (long)(ok ? b0 + bb : a.nb - 1) + (long)aa * a.sin.a);
        if (which) r1[k] = x; else r0[k] = x;
      }
    }
  };

  // a job's operands (both inputs) are loaded while the previous job is
  // transformed and before its stores: a wait for them then never covers
  // stores in flight
  auto load_job = [&](int jb) {
    load(jb, 0);
    if (a.jobs[jb].in1) load(jb, 1);
  };
  load_job(0);
  for (int jb = 0; jb < a.njobs; ++jb) {
    const int nout = a.jobs[jb].out2 ? 2 : 1;
    for (int o = 0; o < nout; ++o) {
      const int op = o ? a.jobs[jb].op2 : a.jobs[jb].op;
      const gptr<cplx<T>> out = (gptr<cplx<T>>)mshift(o ? a.jobs[jb].out2 : a.jobs[jb].out, moff);
      const bool pf = (o == nout - 1) && (jb + 1 < a.njobs);
      cplx<T> v[8];
      if (CONTIG_IN) {
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = apply_op<T, L>(op, a.kscale, l + q * G, r0[q], r1[q]);
        if (b0 + line >= a.nb) {
#pragma unroll
          for (int q = 0; q < 8; ++q) v[q] = mk<T>(T(0), T(0));
        }
        if (pf) load_job(jb + 1);
        twfft<T, L, S>(v, tw, ln, l);
        // turn the tile: natural order into the line's row, then 128-byte
        // segments of TL lines out
#pragma unroll
        for (int q = 0; q < 8; ++q) ln[lp<T>(l + q * G)] = v[q];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const int e = tid + k * NT, bb = e % TL, aa = e / TL;
          const cplx<T> x = tile[bb * LS + lp<T>(aa)];
          if (SF || b0 + bb < nbs) sts<T>(out + c * a.sout.c + (long)(b0 + bb) + (long)aa * a.sout.a, mk<T>(x.x * sc, x.y * sc));
        }
        __syncthreads();
      } else {
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const int e = tid + k * NT, bb = e % TL, aa = e / TL;
          const cplx<T> x = apply_op<T, L>(op, a.kscale, aa, r0[k], r1[k]);
          tile[bb * LS + lp<T>(aa)] = b0 + bb < a.nb ? x : mk<T>(T(0), T(0));
        }
        __syncthreads();
        if (pf) load_job(jb + 1);
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = ln[lp<T>(l + q * G)];
        wave_sync();
        twfft<T, L, S>(v, tw, ln, l);
        if (SF || b0 + line < a.nb) {
          const long base = c * a.sout.c + (long)(b0 + line) * a.sout.b + l;
#pragma unroll
          for (int q = 0; q < 8; ++q) sts<T>(out + base + q * G, mk<T>(v[q].x * sc, v[q].y * sc));
        }
        __syncthreads();
      }
    }
  }
}

template <typename T, int L, int S, bool CI, bool SF>
void run_tw_sf(const TilePassArgs& a, hipStream_t s) {
  using Sh = TWShape<T, L>;
  static bool init = false;
  if (!init) { set_smem(tile_wave_k<T, L, S, CI, SF>, Sh::smem); init = true; }
  dim3 grid((a.nb + Sh::TL - 1) / Sh::TL, a.nc, batch_n(a.bat));
  hipLaunchKernelGGL((tile_wave_k<T, L, S, CI, SF>), grid, dim3(Sh::NT), Sh::smem, s, a);
}
template <typename T, int L, int S, bool CI>
void run_tw(const TilePassArgs& a, hipStream_t s) {
  // stores unguarded when the grid's last tile ends at or below the store
  // bound (nb_store rows for the segment-side stores, nb for the line stores)
  const int tl = TWShape<T, L>::TL;
  const long rows = (long)((a.nb + tl - 1) / tl) * tl;
  const long bound = CI ? (a.nb_store > a.nb ? a.nb_store : a.nb) : a.nb;
  if (rows <= bound) run_tw_sf<T, L, S, CI, true>(a, s);
  else run_tw_sf<T, L, S, CI, false>(a, s);
}

template <typename T, int L>
void run_tw_dir(const TilePassArgs& a, int dir, bool ci, hipStream_t s) {
  if (dir > 0) { if (ci) run_tw<T, L, +1, true>(a, s); else run_tw<T, L, +1, false>(a, s); }
  else { if (ci) run_tw<T, L, -1, true>(a, s); else run_tw<T, L, -1, false>(a, s); }
}

}  // namespace

bool tile_wave_supported(int L) { return L == 64 || L == 128 || L == 256 || L == 512; }

template <typename T>
void launch_tile_wave(const TilePassArgs& a, int L, int dir, bool ci, hipStream_t s) {
  const bool unit = ci ? (a.sin.a == 1 && a.sout.b == 1) : (a.sin.b == 1 && a.sout.a == 1);
  if (!unit) throw std::runtime_error("tile pass: the contiguous axes must have unit stride");
  if (a.njobs < 1 || a.njobs > 6) throw std::runtime_error("tile pass: 1..6 jobs");
  switch (L) {
    case 64: run_tw_dir<T, 64>(a, dir, ci, s); break;
    case 128: run_tw_dir<T, 128>(a, dir, ci, s); break;
    case 256: run_tw_dir<T, 256>(a, dir, ci, s); break;
    case 512: run_tw_dir<T, 512>(a, dir, ci, s); break;
    default: throw std::runtime_error("in-wave tile pass: unsupported length " + std::to_string(L));
  }
}
template void launch_tile_wave<float>(const TilePassArgs&, int, int, bool, hipStream_t);
template void launch_tile_wave<double>(const TilePassArgs&, int, int, bool, hipStream_t);

}  // namespace flusi
