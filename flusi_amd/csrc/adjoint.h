// Launchers of the adjoint kernels (adjoint.hip).  Spectral fields are in the
// S layout (kz fastest, component stride nc), physical fields (ix fastest,
// component stride npts), element type T.
#pragma once
#include <hip/hip_runtime.h>

namespace flusi {

struct SpecGeom {
  int nx, ny, nz, nxh;
  long nc;
  const double *kx, *ky, *kz, *kxt2, *kyt2, *kzt2;  // wavenumbers and dealias terms (device)
  int dealias_on;
};

// W = -nu k^2 mu, MW = w_x/N mu with mu = P K lam (3 fields each)
template <typename T> void launch_adj_pre(const SpecGeom& g, double nu, double invN, const void* lam, void* W, void* MW,
                                          hipStream_t s);
// W += scale c_x (A + i k x B), B may be null
template <typename T> void launch_adj_post(const SpecGeom& g, double scale, const void* A, const void* B, void* W,
                                           hipStream_t s);
// out = a x + b y over n elements of T (y may be null)
template <typename T> void launch_axpby(long n, double a, const void* x, double b, const void* y, void* out, hipStream_t s);
// in place phi -> w x phi - chi/eta phi; aw = phi x u (3 physical fields each)
template <typename T> void launch_adj_point(long npts, double eps_inv, void* phi, const void* u, const void* w,
                                            const void* mask, void* aw, hipStream_t s);
// index box [lo, hi] of an nx x ny x nz grid (x fastest)
struct GridBox { int lo[3], hi[3]; int nx, ny; };
// per-block partial sums of f . ((mP - mM)(u - us) - mask (usP - usM)), f = phi (3
// fields) or e_z when phi is null, over the points of box (outside it the
// perturbed drawings agree), into partial[block * nv] for
// adjoint_reduce_blocks(npts) blocks.  The perturbed masks mP, mM and us
// fields usP, usM are FP64 in both precisions.
template <typename T> void launch_theta(long npts, const GridBox& box, const void* phi, const void* u, const void* us,
                                        const void* mask, const double* mP, const double* usP, const double* mM,
                                        const double* usM, double* partial, int nv, hipStream_t s);
// the same for ncon controls in one launch: control i's drawings 2i (+h) and
// 2i+1 (-h) lie at slab + (2i) stride and slab + (2i+1) stride, each an FP64
// mask (npts) followed by its us (3 npts); partial[block * ncon + i]
template <typename T> void launch_theta_batch(long npts, const GridBox& box, const void* phi, const void* u,
                                              const void* us, const void* mask, const void* slab, long stride,
                                              int ncon, double* partial, hipStream_t s);
// per-block partial sums of mask (u_z - us_z)
template <typename T> void launch_lift(long npts, const void* u, const void* us, const void* mask, double* partial,
                                       hipStream_t s);
// out = (0, 0, scale mask) (3 physical fields)
template <typename T> void launch_ez_mask(long npts, double scale, const void* mask, void* out, hipStream_t s);
int adjoint_reduce_blocks(long npts);

}  // namespace flusi
