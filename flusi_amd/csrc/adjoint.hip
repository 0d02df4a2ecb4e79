// Kernels of the discrete adjoint of the penalized right-hand side and the
// RK4 stepper (no counterpart in the reference; the forward operators they
// transpose are rhs.f90:147-330 cal_nlk_fsi, :441-468 add_explicit_diffusion,
// :575-615 add_grad_pressure and basic_operators.f90:952-1066 dealias).
//
// With the state u-hat (3 spectral fields, S layout), the right-hand side is
//   N(u-hat) = K P [ F g(u, w) - nu k^2 u-hat ],  g = u x w - chi/eta (u - us),
// u = I u-hat, w = I (i k x u-hat), F the forward transform (times 1/N), I the
// unnormalised C2R inverse, P the projection and K the dealiasing mask.  In
// the real inner product of the half spectrum, <a, b> = sum Re(conj(a) b):
//   F^T phi-hat = (1/N) I (w_x phi-hat), w_x = 1 at kx = 0, nx/2, else 1/2,
//   I^T f      = N c_x F f,             c_x = 1 at kx = 0, nx/2, else 2,
// P, K and i k x are symmetric.  So, for an adjoint lambda:
//   mu = P K lambda,  phi = F^T mu,
//   a_u = w x phi - chi/eta phi,  a_w = phi x u   (pointwise, physical),
//   N_u^T lambda = -nu k^2 mu + N c_x (F a_u + i k x F a_w).
#include "adjoint.h"
#include "fft_core.h"
#include "pass_common.h"

namespace flusi {

namespace {

struct ModeIdx { int kx, ky, kz; };
__device__ __forceinline__ ModeIdx mode_of(const SpecGeom& g, long i) {
  ModeIdx m;
  m.kz = (int)(i % g.nz);
  const long r = i / g.nz;
  m.ky = (int)(r % g.ny);
  m.kx = (int)(r / g.ny);
  return m;
}

template <typename T>
__global__ void adj_pre_k(SpecGeom g, double nu, double invN, const cplx<T>* __restrict__ lam, cplx<T>* __restrict__ W,
                          cplx<T>* __restrict__ MW) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.nc) return;
  const ModeIdx m = mode_of(g, i);
  const double kx = g.kx[m.kx], ky = g.ky[m.ky], kz = g.kz[m.kz];
  const double k2 = (kx * kx + ky * ky) + kz * kz;
  const bool alias = ((g.kxt2[m.kx] + g.kyt2[m.ky]) + g.kzt2[m.kz]) >= 1.0;
  double re[3], im[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const cplx<T> v = lam[c * g.nc + i];
    re[c] = (g.dealias_on && alias) ? 0.0 : (double)v.x;
    im[c] = (g.dealias_on && alias) ? 0.0 : (double)v.y;
  }
  if (k2 != 0.0) {
    const double qr = (kx * re[0] + ky * re[1] + kz * re[2]) / k2;
    const double qi = (kx * im[0] + ky * im[1] + kz * im[2]) / k2;
    re[0] -= kx * qr; im[0] -= kx * qi;
    re[1] -= ky * qr; im[1] -= ky * qi;
    re[2] -= kz * qr; im[2] -= kz * qi;
  }
  const double wx = (m.kx == 0 || m.kx == g.nx / 2) ? invN : 0.5 * invN;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    W[c * g.nc + i] = mk<T>((T)(-nu * k2 * re[c]), (T)(-nu * k2 * im[c]));
    MW[c * g.nc + i] = mk<T>((T)(wx * re[c]), (T)(wx * im[c]));
  }
}

template <typename T>
__global__ void adj_post_k(SpecGeom g, double s, const cplx<T>* __restrict__ A, const cplx<T>* __restrict__ B,
                           cplx<T>* __restrict__ W) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.nc) return;
  const ModeIdx m = mode_of(g, i);
  const double cx = ((m.kx == 0 || m.kx == g.nx / 2) ? 1.0 : 2.0) * s;
  double ar[3], ai[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) { const cplx<T> v = A[c * g.nc + i]; ar[c] = v.x; ai[c] = v.y; }
  if (B) {
    const double k[3] = {g.kx[m.kx], g.ky[m.ky], g.kz[m.kz]};
    double br[3], bi[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) { const cplx<T> v = B[c * g.nc + i]; br[c] = v.x; bi[c] = v.y; }
    // i (k x b): (k x b) = (ky bz - kz by, kz bx - kx bz, kx by - ky bx); times i
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
      const double xr = k[c1] * br[c2] - k[c2] * br[c1], xi = k[c1] * bi[c2] - k[c2] * bi[c1];
      ar[c] += -xi;
      ai[c] += xr;
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const cplx<T> w = W[c * g.nc + i];
    W[c * g.nc + i] = mk<T>((T)((double)w.x + cx * ar[c]), (T)((double)w.y + cx * ai[c]));
  }
}

template <typename T>
__global__ void axpby_k(long n, double a, const T* __restrict__ x, double b, const T* __restrict__ y, T* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double yv = y ? (double)y[i] : 0.0;
  out[i] = (T)(a * (double)x[i] + b * yv);
}

template <typename T>
__global__ void adj_point_k(long npts, double eps_inv, T* __restrict__ phi, const T* __restrict__ u,
                            const T* __restrict__ w, const T* __restrict__ mask, T* __restrict__ aw) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  double f[3], uu[3], ww[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) { f[c] = phi[c * npts + p]; uu[c] = u[c * npts + p]; ww[c] = w[c * npts + p]; }
  const double chi = (double)mask[p] * eps_inv;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
    phi[c * npts + p] = (T)((ww[c1] * f[c2] - ww[c2] * f[c1]) - chi * f[c]);  // w x phi - chi phi
    aw[c * npts + p] = (T)(f[c1] * uu[c2] - f[c2] * uu[c1]);                    // phi x u
  }
}

constexpr int RED_NT = 256;

// the block's sum into partial[blockIdx.x * nv]
__device__ __forceinline__ void block_partial(double v, double* partial, int nv = 1) {
  __shared__ double red[RED_NT / 64];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < RED_NT / 64; ++k) s += red[k];
    partial[(long)blockIdx.x * nv] = s;
  }
}

// sum over the box's points of f . (dm (u - us) - m dus), f = phi or e_z
template <typename T>
__device__ __forceinline__ double theta_sum(long npts, const GridBox& box, const T* __restrict__ phi,
                                            const T* __restrict__ u, const T* __restrict__ us,
                                            const T* __restrict__ mask, const double* __restrict__ mP,
                                            const double* __restrict__ usP, const double* __restrict__ mM,
                                            const double* __restrict__ usM) {
  double acc = 0.0;
  const long ex = box.hi[0] - box.lo[0] + 1, ey = box.hi[1] - box.lo[1] + 1, ez = box.hi[2] - box.lo[2] + 1;
  const long nbox = (ex > 0 && ey > 0 && ez > 0) ? ex * ey * ez : 0;
  for (long q = (long)blockIdx.x * RED_NT + threadIdx.x; q < nbox; q += (long)gridDim.x * RED_NT) {
    const long p = ((box.lo[2] + q / (ex * ey)) * box.ny + box.lo[1] + (q / ex) % ey) * box.nx + box.lo[0] + q % ex;
    const double dm = mP[p] - mM[p];
    const double m = mask[p];
    if (dm == 0.0 && m == 0.0) continue;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (!phi && c != 2) continue;
      const double f = phi ? (double)phi[c * npts + p] : 1.0;
      const double dus = usP[c * npts + p] - usM[c * npts + p];
      acc += f * (dm * ((double)u[c * npts + p] - (double)us[c * npts + p]) - m * dus);
    }
  }
  return acc;
}

template <typename T>
__global__ __launch_bounds__(RED_NT) void theta_k(long npts, GridBox box, const T* __restrict__ phi,
                                                  const T* __restrict__ u, const T* __restrict__ us,
                                                  const T* __restrict__ mask, const double* __restrict__ mP,
                                                  const double* __restrict__ usP, const double* __restrict__ mM,
                                                  const double* __restrict__ usM, double* __restrict__ partial, int nv) {
  block_partial(theta_sum<T>(npts, box, phi, u, us, mask, mP, usP, mM, usM), partial, nv);
}

// the same for every control at once (blockIdx.y = control i): its drawings
// 2i (+h) and 2i+1 (-h) are DrawSlab records `stride` bytes apart from `slab`
template <typename T>
__global__ __launch_bounds__(RED_NT) void theta_batch_k(long npts, GridBox box, const T* __restrict__ phi,
                                                        const T* __restrict__ u, const T* __restrict__ us,
                                                        const T* __restrict__ mask, const char* __restrict__ slab,
                                                        long stride, double* __restrict__ partial, int nv) {
  const int i = blockIdx.y;
  const double* mP = reinterpret_cast<const double*>(slab + (2L * i) * stride);
  const double* mM = reinterpret_cast<const double*>(slab + (2L * i + 1) * stride);
  block_partial(theta_sum<T>(npts, box, phi, u, us, mask, mP, mP + npts, mM, mM + npts), partial + i, nv);
}

// sum over points of m (u_z - us_z)
template <typename T>
__global__ __launch_bounds__(RED_NT) void lift_k(long npts, const T* __restrict__ u, const T* __restrict__ us,
                                                 const T* __restrict__ mask, double* __restrict__ partial) {
  double acc = 0.0;
  for (long p = (long)blockIdx.x * RED_NT + threadIdx.x; p < npts; p += (long)gridDim.x * RED_NT) {
    const double m = mask[p];
    if (m != 0.0) acc += m * ((double)u[2 * npts + p] - (double)us[2 * npts + p]);
  }
  block_partial(acc, partial);
}

// physical source chi e_z (scaled), the objective's derivative in u
template <typename T>
__global__ void ez_mask_k(long npts, double s, const T* __restrict__ mask, T* __restrict__ out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  out[p] = T(0);
  out[npts + p] = T(0);
  out[2 * npts + p] = (T)(s * (double)mask[p]);
}

inline unsigned nblk(long n, int nt) { return (unsigned)((n + nt - 1) / nt); }

}  // namespace

int adjoint_reduce_blocks(long npts) {
  const long b = (npts + RED_NT - 1) / RED_NT;
  return (int)(b < 1024 ? b : 1024);
}

template <typename T>
void launch_adj_pre(const SpecGeom& g, double nu, double invN, const void* lam, void* W, void* MW, hipStream_t s) {
  hipLaunchKernelGGL(adj_pre_k<T>, dim3(nblk(g.nc, 256)), dim3(256), 0, s, g, nu, invN, (const cplx<T>*)lam,
                     (cplx<T>*)W, (cplx<T>*)MW);
}
template <typename T>
void launch_adj_post(const SpecGeom& g, double scale, const void* A, const void* B, void* W, hipStream_t s) {
  hipLaunchKernelGGL(adj_post_k<T>, dim3(nblk(g.nc, 256)), dim3(256), 0, s, g, scale, (const cplx<T>*)A,
                     (const cplx<T>*)B, (cplx<T>*)W);
}
template <typename T>
void launch_axpby(long n, double a, const void* x, double b, const void* y, void* out, hipStream_t s) {
  hipLaunchKernelGGL(axpby_k<T>, dim3(nblk(n, 256)), dim3(256), 0, s, n, a, (const T*)x, b, (const T*)y, (T*)out);
}
template <typename T>
void launch_adj_point(long npts, double eps_inv, void* phi, const void* u, const void* w, const void* mask, void* aw,
                      hipStream_t s) {
  hipLaunchKernelGGL(adj_point_k<T>, dim3(nblk(npts, 256)), dim3(256), 0, s, npts, eps_inv, (T*)phi, (const T*)u,
                     (const T*)w, (const T*)mask, (T*)aw);
}
template <typename T>
void launch_theta(long npts, const GridBox& box, const void* phi, const void* u, const void* us, const void* mask,
                  const double* mP, const double* usP, const double* mM, const double* usM, double* partial, int nv,
                  hipStream_t s) {
  const int nz = (int)(npts / ((long)box.nx * box.ny));
  for (int i = 0; i < 3; ++i) {
    const int n = i == 0 ? box.nx : (i == 1 ? box.ny : nz);
    if (box.lo[i] < 0 || box.hi[i] >= n) throw std::runtime_error("theta terms: box outside the grid");
  }
  hipLaunchKernelGGL(theta_k<T>, dim3(adjoint_reduce_blocks(npts)), dim3(RED_NT), 0, s, npts, box, (const T*)phi,
                     (const T*)u, (const T*)us, (const T*)mask, mP, usP, mM, usM, partial, nv);
}
template <typename T>
void launch_theta_batch(long npts, const GridBox& box, const void* phi, const void* u, const void* us,
                        const void* mask, const void* slab, long stride, int ncon, double* partial, hipStream_t s) {
  const int nz = (int)(npts / ((long)box.nx * box.ny));
  for (int i = 0; i < 3; ++i) {
    const int n = i == 0 ? box.nx : (i == 1 ? box.ny : nz);
    if (box.lo[i] < 0 || box.hi[i] >= n) throw std::runtime_error("theta terms: box outside the grid");
  }
  if (ncon < 1 || ncon > 65535) throw std::runtime_error("theta terms: bad control count");
  hipLaunchKernelGGL(theta_batch_k<T>, dim3(adjoint_reduce_blocks(npts), ncon), dim3(RED_NT), 0, s, npts, box,
                     (const T*)phi, (const T*)u, (const T*)us, (const T*)mask, (const char*)slab, stride, partial,
                     ncon);
}
template <typename T>
void launch_lift(long npts, const void* u, const void* us, const void* mask, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(lift_k<T>, dim3(adjoint_reduce_blocks(npts)), dim3(RED_NT), 0, s, npts, (const T*)u,
                     (const T*)us, (const T*)mask, partial);
}
template <typename T>
void launch_ez_mask(long npts, double scale, const void* mask, void* out, hipStream_t s) {
  hipLaunchKernelGGL(ez_mask_k<T>, dim3(nblk(npts, 256)), dim3(256), 0, s, npts, scale, (const T*)mask, (T*)out);
}

#define FLUSI_ADJ_INST(T)                                                                                          \
  template void launch_adj_pre<T>(const SpecGeom&, double, double, const void*, void*, void*, hipStream_t);       \
  template void launch_adj_post<T>(const SpecGeom&, double, const void*, const void*, void*, hipStream_t);        \
  template void launch_axpby<T>(long, double, const void*, double, const void*, void*, hipStream_t);              \
  template void launch_adj_point<T>(long, double, void*, const void*, const void*, const void*, void*,            \
                                    hipStream_t);                                                                 \
  template void launch_theta<T>(long, const GridBox&, const void*, const void*, const void*, const void*,       \
                                const double*, const double*, const double*, const double*, double*, int,         \
                                hipStream_t);                                                                     \
  template void launch_theta_batch<T>(long, const GridBox&, const void*, const void*, const void*, const void*,   \
                                      const void*, long, int, double*, hipStream_t);                              \
  template void launch_lift<T>(long, const void*, const void*, const void*, double*, hipStream_t);                \
  template void launch_ez_mask<T>(long, double, const void*, void*, hipStream_t);
FLUSI_ADJ_INST(float)
FLUSI_ADJ_INST(double)


}  // namespace flusi
