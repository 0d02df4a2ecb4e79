// Kernel argument blocks and launcher declarations shared by the transform
// passes (passes.hip), the field kernels (fields.hip) and the solver host code.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace flusi {

// Element offsets of a 3-D complex array walked as (a = FFT axis,
// b = tile axis, c = outer axis).
struct Strides { long a, b, c; };

// Batched launches.  The members of a batch (simulations of one grid and one
// parameter set) keep identical buffer layouts in slabs `mstride` bytes
// apart.  A batched launch adds one grid dimension over its nmem members and
// shifts every device pointer of the argument block by member * mstride;
// nmem <= 1 is a plain launch.
template <typename P> __host__ __device__ __forceinline__ P* mshift(P* p, long off) {
  return p ? reinterpret_cast<P*>(reinterpret_cast<uintptr_t>(p) + off) : p;
}
struct Batch {
  int nmem;      // members in the launch (0 or 1: not batched)
  long mstride;  // bytes between two members' slabs (a multiple of 256)
};
__host__ __device__ __forceinline__ int batch_n(const Batch& b) { return b.nmem > 1 ? b.nmem : 1; }

enum TileOp { OP_COPY = 0, OP_DERIV = 1, OP_DERIV_MINUS = 2 };

// One job reads its operands once and writes one or two transforms of them.
struct TileJob {
  const void* in0;   // complex field, layout given by Strides in
  const void* in1;   // second operand of OP_DERIV_MINUS (nullptr: not read)
  void* out;
  int op;            // out = in0 | i k_a in0 | i k_a in0 - in1
  void* out2;        // second output of the same operands (nullptr: none)
  int op2;
};

struct TilePassArgs {
  TileJob jobs[6];
  int njobs;
  int nb, nc;          // extents of the tile axis and the outer axis
  int nb_store;        // in-wave passes: lines the segment-side store covers (>= nb; lines
                       // beyond nb are row padding no kernel reads, stored as zeros so that
                       // the last tile writes whole 128-byte segments); 0 = nb
  Strides sin, sout;
  double kscale;       // k_a = kscale * wrap(a): wavenumbers along the FFT axis
  double scale;        // applied to every stored value
  const void* tw;      // L twiddles exp(-2 pi i m / L), element type of the pass
  int pf;              // block passes: load the next job's operand while the current one is transformed
  Batch bat;           // grid z: member
};

// x pass modes
enum XMode {
  XM_RHS = 0,       // 6 spectral fields -> u x w - chi (u - us) [+ sponge] -> 3 [6] spectral
  XM_INV_UW = 1,    // 6 spectral fields -> u (3) and w (3) in physical space
  XM_INV = 2,       // nf spectral fields -> nf physical fields
  XM_FWD = 3,       // nf physical fields -> nf spectral fields
};

struct XPassArgs {
  const void* in;      // nf spectral (P layout) or physical fields, field stride fstride_in
  void* out;           // outputs, field stride fstride_out
  long fstride_in, fstride_out;
  int nf;              // fields for XM_INV / XM_FWD
  int nlines;          // ny*nz
  int nx, ny, nxh;
  long prow;           // row stride of the spectral (P-layout) side, >= nxh
  const double* kx;    // nxh wavenumbers
  double kx_scale;     // kx = kx_scale * k
  int wpe;             // one-wave-per-line RHS pass: register budget for 3 waves per SIMD (3) or none (0)
  const void* twM;     // M = nx/2 twiddles
  const void* twN;     // nxh twiddles exp(-2 pi i k / nx)
  // RHS extras
  const void* mask;    // N reals (element type of the pass)
  const void* us;      // 3N reals, component stride N
  const uint8_t* mflag;  // per x line (ny*nz): 0 where the line's mask is all zero; null: read every line
  const void* mzero;   // nx zeros of the element type, read instead of a line whose flag is 0
  long npts;           // N
  double eps_inv;      // 1/eta
  int sponge;          // 0 none, 1 cavity, 2 outlet_x
  int sponge_thickness;
  double sponge_eps_inv;
  double sponge_mask_thr;
  int nz;
  double* partial;     // per block: umax^2, sum chi(u-us) x,y,z (4 doubles)
  int want_stats;      // write the per-block statistics (adaptive dt / dynamic mean flow)
  Batch bat;           // grid y: member
};

// XPassArgs with its pointers moved to batch member m
__host__ __device__ __forceinline__ void to_member(XPassArgs& a, int m) {
  const long o = (long)m * a.bat.mstride;
  a.in = mshift(a.in, o); a.out = mshift(a.out, o); a.kx = mshift(a.kx, o);
  a.twM = mshift(a.twM, o); a.twN = mshift(a.twN, o);
  a.mask = mshift(a.mask, o); a.us = mshift(a.us, o); a.partial = mshift(a.partial, o);
}

// z-forward pass with the fused spectral epilogue (projection, explicit
// diffusion, dealiasing, stepper update, zero-mode forcing, Parseval sum)
enum StepMode {
  SM_RHS_ONLY = 0,   // out0 = nlk (projected)
  SM_EULER = 1,      // AB2 start-up Euler step
  SM_AB2 = 2,
  SM_RK2_1 = 3,
  SM_RK2_2 = 4,
  SM_RK4_1 = 5,
  SM_RK4_2 = 6,
  SM_RK4_3 = 7,
  SM_RK4_4 = 8,
  SM_PRESSURE = 9,   // out0 = -i k.nlk / k^2 of the unprojected right-hand side (1 field)
};

struct ZStepArgs {
  const void* nq;        // 3 (6 with sponge) Q-layout fields, field stride fs
  long fs;
  int nx, ny, nz, nxh;
  const double* kx; const double* ky; const double* kz;
  const double* kxt2; const double* kyt2; const double* kzt2;  // dealias predicate terms
  const void* tw;        // nz twiddles
  double scale;          // 1/(nx ny nz)
  int mode;
  int sponge;
  int dealias_on;        // iDealias
  double nu;
  // state (S layout, 3 fields each, field stride fs)
  void* u;               // uk (AB2/RK2), u_n (RK4)
  const void* udiff;     // field the explicit diffusion term is taken from (stage input)
  void* ustage;          // RK4 next-stage input (written)
  void* acc;             // RK4 accumulator
  void* r0;              // AB2: nlk(n0) out; RK2: N0 (out in stage 1, in stage 2)
  const void* r1;        // AB2: nlk(n1) in
  void* rhs_out;         // SM_RHS_ONLY output
  double dt, b10, b11;
  double vis_dt_new;     // IF exponent dt (<0: factor is zero)
  double vis_dt_old;     // RK2 stale IF (<0: zero)
  int explicit_diffusion;
  // zero mode forcing
  int meanflow_dynamic[3];
  const double* pen_sum; // device sum chi(u-us) over the grid (3 doubles)
  double dynamic_scale;  // dx dy dz * startup / m_fluid (negated applied)
  int set_mean[3];       // post-step zero-mode overwrite (fixed/sinusoidal)
  double mean_value[3];
  double* energy_partial;  // per block Parseval partial sum (final stages only)
  int wave;              // 1: transform each z line inside one wavefront (nz = 64..512)
  int eu;                // block kernel, nz >= 256: epilogue positions scheduled together (0 = all)
  Batch bat;             // grid z: member
};

// launchers (passes.hip)
template <typename T> void launch_tile_pass(const TilePassArgs& a, int L, int dir, bool contig_in, hipStream_t s);
// the same passes with each line transformed inside one wavefront (tiles_wave.hip)
template <typename T> void launch_tile_wave(const TilePassArgs& a, int L, int dir, bool contig_in, hipStream_t s);
bool tile_wave_supported(int L);
bool zstep_wave_supported(int nz);  // in-wave z step (else the block kernel)
template <typename T> void launch_x_pass(const XPassArgs& a, int mode, hipStream_t s);
template <typename T> void launch_zstep(const ZStepArgs& a, hipStream_t s);
int x_pass_blocks(int nx, int nlines, bool fp64);
// the right-hand-side x pass with one wavefront per line (xrhs_wave.hip)
template <typename T> void launch_x_rhs_wave(const XPassArgs& a, hipStream_t s);
bool x_rhs_wave_supported(int nx);
int x_rhs_wave_blocks(int nx, int nlines);
// per x line (iy, iz) of [y0, y1] x [z0, z1]: flag[iz ny + iy] = 1 where any
// mask value of the line is nonzero, else 0 (xrhs_wave.hip)
template <typename T>
void launch_mask_flags(const void* mask, uint8_t* flag, int nx, int ny, int y0, int y1, int z0, int z1, hipStream_t s);
// energy partial sums one z-step launch writes per member (the block kernel
// or, wave = true, the in-wave kernel)
int zstep_blocks(int nz, int ny, int nxh, bool fp64, bool wave);
bool fft_length_supported(int n);   // y/z line lengths
bool x_length_supported(int m);     // nx/2 of the real x axis

// reductions (fields.hip)
// out[v] = combine over rows of partial[row*nv + v]; kind[v]: 0 = Neumaier sum, 1 = max, 2 = min
void launch_reduce_rows(const double* partial, int nrows, int nv, const int* kinds_dev, double* mid, double* out, hipStream_t s);
size_t reduce_rows_scratch(int nv);  // doubles of `mid` needed for nv values

}  // namespace flusi
