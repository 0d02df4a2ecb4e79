// Point-wise field kernels (masks, initial conditions, spectral helpers,
// diagnostics) and their argument blocks.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "flusi_kernels.h"

namespace flusi {

struct GridGeom {
  int nx, ny, nz, nxh;
  long npts, ncplx;
  double dx, dy, dz, xl, yl, zl;
  int periodic;
};

// geometry/hyd/draw_sphere.f90:2-43
struct SphereArgs { double x0, y0, z0, radius, smoothing, safety, color_thr; };

// geometry/add_cavity.f90:8-111 (iCavity = yes | freestream)
struct CavityArgs { int size; double ux, uy, uz; };

// One insect mask update: module_insects.f90:338-399 Draw_Insect, executed
// point by point (every stage of the reference is local to a grid point).
struct InsectArgs {
  int full_reset;           // delete_old_mask: reset everything (else wing colours only)
  int draw_body;            // body drawn in this call (once for a tethered insect)
  int body_type;            // 1 = jerry ellipsoid + head/eyes, 2 = ellipsoid only, 3 = bumblebee
  double xc[3];             // body centre (global)
  double Mb[9];             // body rotation (row-major, global -> body)
  double b_body, L_body, smooth, safety;
  int nspheres;             // head and eyes (jerry)
  double sph_c[3][3], sph_r[3];
  int sph_lo[3][3], sph_hi[3][3];  // index bounding boxes [sphere][axis]
  // bumblebee legs, antennae and proboscis: ncaps capsule records of
  // CAPSULE_W doubles (layout at CapsuleField)
  int ncaps;
  const double* caps;
  int wing_shape;           // 1 = TwoEllipses, 2 = rectangular, 3 = Fourier outline
  int smooth_mode;          // differentiable variant (adjoint runs): continuous wing-root taper
                            // and us blended by part weights where parts overlap
  const double* fw_table;   // Fourier outline: radius table on nodes j*fw_dphi
  double fw_xc, fw_yc, fw_dphi, fw_bbox[6];
  double L_span, b_top, b_bot, thickness;
  int draw_wing[2];         // right, left
  double Mw[2][9];          // wing rotation (body -> wing)
  double piv[2][3];         // pivots in the body frame
  double rot[2][3];         // relative wing angular velocity in the wing frame
  int color_w[2];           // 3 (right), 2 (left)
  double vb[3];             // M_body * v_body
  double rot_body_b[3];
  int use_box;              // 1: skip points outside the index boxes below
  int box_lo[3][3], box_hi[3][3];  // [body, right wing, left wing][axis]
  // points visited: [reg_lo, reg_hi] per axis.  Outside it nothing is drawn
  // and no wing point of the previous call lies there (the whole grid on a
  // full reset or for periodic insects).
  int reg_lo[3], reg_hi[3];
  // points that can change: [clip_lo, clip_hi] per axis (the region above
  // when it is the whole grid only because the launch is batched or captured
  // with one fixed extent); clip = 0: every visited point
  int clip;
  int clip_lo[3], clip_hi[3];
};

// capsule record: segment end points (global, not periodized), radius, and
// the end spheres' centres wrapped into the box
enum CapsuleField { CAP_A = 0, CAP_B = 3, CAP_R = 6, CAP_SA = 7, CAP_SB = 10, CAPSULE_W = 13 };

// drag.f90:17-454 + integrals.f90:48-238 reductions over physical u, w, div
struct DiagArgs {
  const void* u;  const void* w;  const void* div;   // physical fields (pass element type)
  const void* mask; const void* us; const uint8_t* color;
  double eps_inv;
  double x0[3];                    // torque reference point
  double piv_l[3], piv_r[3];       // insect pivots (global frame)
  int insect;                      // colours 2/3 use the pivots
  int want_w;                      // 0: w and div are not read (cal_drag's sums only)
  double* partial;                 // per block DIAG_NV doubles
};
enum {
  DG_FORCE = 0,          // [4 colours][3]
  DG_TORQUE0 = 12,       // [4][3] about (x0,y0,z0)
  DG_TORQUE = 24,        // [4][3] about pivots for wing colours
  DG_POWER = 36,         // [3]
  DG_URES = 39,          // [4] min |u-us|^2 inside the mask
  DG_EK = 43,            // [3] sum u_i^2
  DG_EKF = 46,           // [3] sum ((1-mask) u_i)^2
  DG_DISS = 49,          // [3] sum w_i^2
  DG_DISSF = 52,         // [3]
  DG_DIVMAX = 55,        // max div
  DG_DIVMAXF = 56,       // max div (1-mask)
  DG_MASKVOL = 57,       // sum mask
  DG_MASKUS = 58,        // [3] sum mask us_i
  DIAG_NV = 61
};

template <typename T> void launch_sphere(const GridGeom& g, const SphereArgs& a, void* mask, void* us, uint8_t* color, hipStream_t s);
template <typename T> void launch_cavity(const GridGeom& g, const CavityArgs& a, void* mask, void* us, uint8_t* color, hipStream_t s);
// a: the arguments on the host (the launch region comes from them); a_dev:
// the same arguments in device memory, read by the kernel, one record per
// batch member (bat.nmem > 1: every member's region must equal a's)
template <typename T> void launch_insect(const GridGeom& g, const InsectArgs& a, const InsectArgs* a_dev, void* mask,
                                         void* us, uint8_t* color, hipStream_t s, const Batch& bat);
// p -= |u|^2 / 2 (u: 3 physical fields, p: 1)
template <typename T> void launch_pressure_phys(long npts, const void* u, void* p, hipStream_t s);
template <typename T> void launch_vortex_ring(const GridGeom& g, double x0, double y0, double z0, double r00, double a, double gamma0, void* vort, hipStream_t s);
// spectral helpers on S-layout fields (3 components, component stride ncplx)
template <typename T> void launch_biot_savart(const GridGeom& g, const double* kx, const double* ky, const double* kz, const void* vk, void* uk, hipStream_t s);
template <typename T> void launch_divergence(const GridGeom& g, const double* kx, const double* ky, const double* kz, const void* uk, void* dk, hipStream_t s);
template <typename T> void launch_set_mode0(const GridGeom& g, void* uk, int comp, double re, double im, hipStream_t s);
template <typename T> int diag_blocks(const GridGeom& g);
template <typename T> void launch_diag(const GridGeom& g, const DiagArgs& a, hipStream_t s);
// precision conversion host-layout <-> device arrays
template <typename T> void launch_from_f64(const double* src, void* dst, long n, hipStream_t s);
template <typename T> void launch_to_f64(const void* src, double* dst, long n, hipStream_t s);
void launch_zero(void* p, size_t bytes, hipStream_t s);
// integrals.f90:522-552 Parseval sums of 3 spectral fields: energy_k_blocks()
// partials (to be summed by launch_reduce_rows)
int energy_k_blocks();
template <typename T> void launch_energy_k(const GridGeom& g, const void* uk, double* partial, hipStream_t s);
// vis.f90:9-66 cal_vis into a device table of ncplx doubles
void launch_cal_vis(const GridGeom& g, const double* kx, const double* ky, const double* kz, const double* kxt2,
                    const double* kyt2, const double* kzt2, double dt, double nu, int dealias, double* vis,
                    hipStream_t s);
// out = |w| of 3 physical fields (component stride npts)
template <typename T> void launch_magnitude(long npts, const void* w, void* out, hipStream_t s);

}  // namespace flusi
