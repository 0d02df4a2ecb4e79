// Point-wise kernels: obstacle masks, initial vorticity, spectral helpers,
// the output-time diagnostics reduction and the deterministic row reducer.
#include "fields.h"
#include "flusi_kernels.h"
#include "fft_core.h"
#include <stdexcept>
#include <vector>

namespace flusi {

static constexpr double PI_D = 3.14159265358979323846264338327950288;

__device__ __forceinline__ int color_clamp(uint8_t c) { return c < 4 ? (int)c : 4; }

__device__ __forceinline__ double steps_d(double x, double t, double h) {
  // module_insects.f90:408-422 steps / mask.f90:68-90 smoothstep (cosine)
  if (x <= t - h) return 1.0;
  if ((t - h) < x && x < (t + h)) return 0.5 * (1.0 + cos((x - t + h) * PI_D / (2.0 * h)));
  return 0.0;
}

__device__ __forceinline__ void periodize_d(double* x, const GridGeom& g) {
  // vars.f90:468-484
  if (!g.periodic) return;
  const double box[3] = {g.xl, g.yl, g.zl};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double v = x[i];
    if (v < -box[i] / 2.0) x[i] = v + box[i];
    if (v > box[i] / 2.0) x[i] = v - box[i];
  }
}

__device__ __forceinline__ void matvec(const double* M, const double* v, double* o) {
  o[0] = M[0] * v[0] + M[1] * v[1] + M[2] * v[2];
  o[1] = M[3] * v[0] + M[4] * v[1] + M[5] * v[2];
  o[2] = M[6] * v[0] + M[7] * v[1] + M[8] * v[2];
}
__device__ __forceinline__ void matTvec(const double* M, const double* v, double* o) {
  o[0] = M[0] * v[0] + M[3] * v[1] + M[6] * v[2];
  o[1] = M[1] * v[0] + M[4] * v[1] + M[7] * v[2];
  o[2] = M[2] * v[0] + M[5] * v[1] + M[8] * v[2];
}

// ------------------------------------------------------------------ sphere
template <typename T>
__global__ void sphere_k(GridGeom g, SphereArgs a, T* mask, T* us, uint8_t* color) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.npts) return;
  const int ix = p % g.nx, iy = (p / g.nx) % g.ny, iz = p / ((long)g.nx * g.ny);
  const double x = ix * g.dx, y = iy * g.dy, z = iz * g.dz;
  const double R = sqrt((x - a.x0) * (x - a.x0) + (y - a.y0) * (y - a.y0) + (z - a.z0) * (z - a.z0));
  double m = 0.0; uint8_t c = 0;
  if (R <= a.radius + a.safety) {
    m = steps_d(R, a.radius, a.smoothing);
    if (m > a.color_thr) c = 1;
  }
  mask[p] = (T)m; color[p] = c;
  us[p] = T(0); us[g.npts + p] = T(0); us[2 * g.npts + p] = T(0);
}

template <typename T>
void launch_sphere(const GridGeom& g, const SphereArgs& a, void* mask, void* us, uint8_t* color, hipStream_t s) {
  const int nt = 256;
  hipLaunchKernelGGL(sphere_k<T>, dim3((g.npts + nt - 1) / nt), dim3(nt), 0, s, g, a, (T*)mask, (T*)us, color);
}

// ------------------------------------------------------------------ cavity
template <typename T>
__global__ void cavity_k(GridGeom g, CavityArgs a, T* mask, T* us, uint8_t* color) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.npts) return;
  const int ix = p % g.nx, iy = (p / g.nx) % g.ny, iz = p / ((long)g.nx * g.ny);
  const int n = a.size;
  if (ix <= n - 1 || ix >= g.nx - n || iy <= n - 1 || iy >= g.ny - n || iz <= n - 1 || iz >= g.nz - n) {
    mask[p] = T(1); color[p] = 0;
    us[p] = (T)a.ux; us[g.npts + p] = (T)a.uy; us[2 * g.npts + p] = (T)a.uz;
  }
}
template <typename T>
void launch_cavity(const GridGeom& g, const CavityArgs& a, void* mask, void* us, uint8_t* color, hipStream_t s) {
  const int nt = 256;
  hipLaunchKernelGGL(cavity_k<T>, dim3((g.npts + nt - 1) / nt), dim3(nt), 0, s, g, a, (T*)mask, (T*)us, color);
}

// ------------------------------------------------------------------ insect
// module_insects.f90:338-399 Draw_Insect with delete_old_mask (:844-883),
// draw_body_jerry (body_geometry.f90:511-595), drawsphere (:1657-1724),
// draw_wing_twoellipses / _rectangular (wings_geometry.f90:527-712) and the
// stage II body-motion transform (wings_geometry.f90:71-108).
// One wing's value mt at a point (wing frame xw): the mask keeps the largest
// part, whose colour and velocity win (module_insects.f90 draw_insect_wings);
// the differentiable variant also accumulates every wing's weighted velocity.
__device__ __forceinline__ void wing_point(const InsectArgs& a, int w, const double* xw, double mt, double& m,
                                           uint8_t& c, bool& changed, double* u_s, double& wsum, double* usw) {
  if (!(mt > 0.0)) return;
  const double* r = a.rot[w];
  const double v[3] = {r[1] * xw[2] - r[2] * xw[1], r[2] * xw[0] - r[0] * xw[2], r[0] * xw[1] - r[1] * xw[0]};
  double vw[3];
  matTvec(a.Mw[w], v, vw);
  if (a.smooth_mode) {
    wsum += mt;
#pragma unroll
    for (int i = 0; i < 3; ++i) usw[i] += mt * vw[i];
  }
  if (m < mt) {
    m = mt; c = (uint8_t)a.color_w[w]; changed = true;
#pragma unroll
    for (int i = 0; i < 3; ++i) u_s[i] = vw[i];
  }
}

template <typename T>
__global__ void insect_k(GridGeom g, const InsectArgs* __restrict__ ap, T* mask, T* us, uint8_t* color, int ez,
                         long mstride) {
  // the arguments live in device memory, so that a captured launch reads the
  // kinematics the host wrote for this stage.  Batch member m = blockIdx.z / ez
  // has its own record ap[m] and its fields m * mstride bytes further on; the
  // members share the region's extents.
  const int mem = blockIdx.z / ez;
  const InsectArgs& a = ap[mem];
  mask = mshift(mask, mem * mstride); us = mshift(us, mem * mstride); color = mshift(color, mem * mstride);
  // grid over the region: x in blocks of blockDim.x, y = blockIdx.y, z = blockIdx.z % ez
  const int ix = a.reg_lo[0] + blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = a.reg_lo[1] + blockIdx.y, iz = a.reg_lo[2] + (blockIdx.z - mem * ez);
  if (ix > a.reg_hi[0]) return;
  if (a.clip && (iy < a.clip_lo[1] || iy > a.clip_hi[1] || iz < a.clip_lo[2] || iz > a.clip_hi[2] ||
                 ix < a.clip_lo[0] || ix > a.clip_hi[0]))
    return;  // neither drawn now nor holding a wing point of the previous update
  const long p = ((long)iz * g.ny + iy) * g.nx + ix;
  const uint8_t c_old = color[p];
  const bool old_wing = (c_old == 2 || c_old == 3);
  double m;
  uint8_t c;
  bool mask_loaded = false;
  if (a.full_reset) { m = 0.0; c = 0; mask_loaded = true; }
  else if (old_wing) { m = 0.0; c = 0; mask_loaded = true; }
  else { m = 0.0; c = c_old; }
  double u_s[3] = {0.0, 0.0, 0.0};
  bool changed = a.full_reset || old_wing;
  auto in_box = [&](int k) {
    return !a.use_box || (ix >= a.box_lo[k][0] && ix <= a.box_hi[k][0] && iy >= a.box_lo[k][1] &&
                          iy <= a.box_hi[k][1] && iz >= a.box_lo[k][2] && iz <= a.box_hi[k][2]);
  };
  const bool any = (a.draw_body && (in_box(0) || a.nspheres > 0 || a.ncaps > 0)) || (a.draw_wing[0] && in_box(1)) ||
                   (a.draw_wing[1] && in_box(2));
  if (!any) {
    // nothing can be drawn here: only the deletion of old wing points remains
    if (changed) {
      mask[p] = T(0); color[p] = 0;
      us[p] = T(0); us[g.npts + p] = T(0); us[2 * g.npts + p] = T(0);
    }
    return;
  }
  double xg[3] = {ix * g.dx - a.xc[0], iy * g.dy - a.xc[1], iz * g.dz - a.xc[2]};
  periodize_d(xg, g);
  double xb[3];
  matvec(a.Mb, xg, xb);

  if (a.draw_body && in_box(0) && a.body_type == 3) {
    // body_geometry.f90:157-268 bumblebee: body along an arc centreline, head
    const double x = xb[0], y = xb[1], z = xb[2], sf = a.safety;
    const double rbc = 1.3, thbc1 = 112.0 * PI_D / 180.0, thbc2 = 53.0 * PI_D / 180.0;
    const double x0bc = 0.0782301255230126, z0bc = -1.26512552301255;
    const double s = (atan2(z - z0bc, -(x - x0bc)) - thbc1) / (thbc2 - thbc1);
    if (s >= -sf && s <= 1.075 + sf) {
      const double rr = sqrt((x - x0bc) * (x - x0bc) + (z - z0bc) * (z - z0bc));
      const double xcs = x0bc + (x - x0bc) * rbc / rr, zcs = z0bc + (z - z0bc) * rbc / rr;
      double s1 = 1.0 - (s + 0.08 * tanh(30.0 * s)) / (1.0 + 0.08 * tanh(30.0));
      s1 = (s1 + 0.04 * tanh(60.0 * s1)) / (1.0 + 0.04 * tanh(60.0));
      s1 = sin(1.2 * s1) / sin(1.2);
      s1 = copysign(pow(fabs(s1), 1.25), s1);
      const double x1 = 1.075 * s1;
      double R0 = 0.0, a_body = 1.0;
      if (x1 < 0.6333) {
        R0 = fmax(-1.2990 * (x1 * x1) + 0.9490 * x1 + 0.0267, 0.0);
        a_body = 1.0 - 0.07 * (x1 - 0.6333) * x1 / 0.0488;
      } else if (x1 >= 0.6333 && x1 <= 1.075) {
        R0 = fmax(-2.1667 * (x1 * x1) + 3.4661 * x1 - 1.2194, 0.0);
      }
      R0 = 1.2 * (1.0 + 0.6 * ((1.0 - s) * (1.0 - s))) * R0;
      const double az = a_body * (z - zcs);
      const double R = sqrt((x - xcs) * (x - xcs) + y * y + az * az);
      if (R < R0 + sf && R0 > 0.0) {
        if (!mask_loaded) { m = (double)mask[p]; mask_loaded = true; }
        m = fmax(steps_d(R, R0, a.smooth), m);
        c = 1; changed = true;
      }
    }
    const double a_head = 1.04, xx_head = 0.58125, zz_head = -0.1, dx_head = 0.5 * 0.2035, dz_head = 0.5 * 0.297;
    if (fabs(y) <= dz_head + sf && fabs(z - zz_head) <= dz_head + sf && fabs(x - xx_head) < dx_head + sf) {
      const double qx = (x - xx_head) / dx_head;
      const double R = sqrt((a_head * y) * (a_head * y) + (z - zz_head) * (z - zz_head));
      if (qx * qx <= 1.0) {
        const double R0 = dz_head * sqrt(1.0 - qx * qx);
        if (R < R0 + sf) {
          if (!mask_loaded) { m = (double)mask[p]; mask_loaded = true; }
          m = fmax(steps_d(R, R0, a.smooth), m);
          c = 1; changed = true;
        }
      }
    }
  }
  if (a.draw_body && in_box(0) && a.body_type != 3) {
    const double ab = a.L_body / 2.0, b = a.b_body;
    if (fabs(xb[1]) <= b + a.safety && fabs(xb[2]) <= b + a.safety && fabs(xb[0]) < a.L_body / 2 + a.safety) {
      const double rx = xb[0] / ab;
      if (rx * rx <= 1.0) {
        const double R = sqrt(xb[1] * xb[1] + xb[2] * xb[2]);
        const double R0 = sqrt(b * b * (1.0 - rx * rx));
        if (R < R0 + a.safety) {
          if (!mask_loaded) { m = (double)mask[p]; mask_loaded = true; }
          m = fmax(steps_d(R, R0, a.smooth), m);
          c = 1; changed = true;
        }
      }
    }
  }
  if (a.draw_body) {
    for (int sidx = 0; sidx < a.nspheres; ++sidx) {
      if (ix < a.sph_lo[sidx][0] || ix > a.sph_hi[sidx][0] || iy < a.sph_lo[sidx][1] || iy > a.sph_hi[sidx][1] ||
          iz < a.sph_lo[sidx][2] || iz > a.sph_hi[sidx][2]) continue;
      double xs[3] = {ix * g.dx - a.sph_c[sidx][0], iy * g.dy - a.sph_c[sidx][1], iz * g.dz - a.sph_c[sidx][2]};
      periodize_d(xs, g);
      const double R = sqrt(xs[0] * xs[0] + xs[1] * xs[1] + xs[2] * xs[2]);
      if (R <= a.sph_r[sidx] + a.safety) {
        if (!mask_loaded) { m = (double)mask[p]; mask_loaded = true; }
        const double tmp = steps_d(R, a.sph_r[sidx], a.smooth);
        if (tmp >= m) { m = tmp; c = 1; changed = true; }
      }
    }
  }

  if (a.draw_body && a.ncaps > 0) {
    // body_geometry.f90:1531-1647 draw_cylinder_new and its end spheres
    // (:1657-1724 drawsphere), in the reference's drawing order.  With a
    // periodic insect every capsule is drawn whole across the box boundary
    // (distances to the segment midpoint and to the end points periodized),
    // as the reference's insects_flow_BB_tethered_AB2_periodic fixture shows
    // (its mask volume equals the centred insect's to 1e-9).
    const double xa[3] = {ix * g.dx, iy * g.dy, iz * g.dz};
    for (int k = 0; k < a.ncaps; ++k) {
      const double* cp = a.caps + (long)k * CAPSULE_W;
      const double r = cp[CAP_R], rr0 = r + a.safety;
      double d[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) d[i] = xa[i] - 0.5 * (cp[CAP_A + i] + cp[CAP_B + i]);
      periodize_d(d, g);
      double cb2 = 0.0, rb2 = 0.0, u2 = 0.0, vp[3], ab[3], u[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const double cb = -d[i];
        const double rb = cp[CAP_A + i] - cp[CAP_B + i];
        cb2 += cb * cb; rb2 += rb * rb;
        ab[i] = g.periodic ? (d[i] + 0.5 * (cp[CAP_A + i] + cp[CAP_B + i])) - cp[CAP_A + i] : xa[i] - cp[CAP_A + i];
        u[i] = cp[CAP_B + i] - cp[CAP_A + i];
        u2 += u[i] * u[i];
      }
      if (cb2 < 0.25 * rb2) {
        vp[0] = ab[1] * u[2] - ab[2] * u[1];
        vp[1] = ab[2] * u[0] - ab[0] * u[2];
        vp[2] = ab[0] * u[1] - ab[1] * u[0];
        const double R = sqrt((vp[0] * vp[0] + vp[1] * vp[1] + vp[2] * vp[2]) / u2);
        if (R <= rr0) {
          const double t = steps_d(R, r, a.smooth);
          if (!mask_loaded) { m = (double)mask[p]; mask_loaded = true; }
          if (t >= m) { m = t; c = 1; changed = true; }
        }
      }
      for (int e = 0; e < 2; ++e) {
        const double* sc = cp + (e ? CAP_SB : CAP_SA);  // end point, wrapped into the box
        double xs[3] = {xa[0] - sc[0], xa[1] - sc[1], xa[2] - sc[2]};
        periodize_d(xs, g);
        const double R = sqrt(xs[0] * xs[0] + xs[1] * xs[1] + xs[2] * xs[2]);
        if (R <= rr0) {
          const double t = steps_d(R, r, a.smooth);
          if (!mask_loaded) { m = (double)mask[p]; mask_loaded = true; }
          if (t >= m) { m = t; c = 1; changed = true; }
        }
      }
    }
  }

  // differentiable variant: the body's value and the wings' weighted velocities
  const double mbody = m;
  double wsum = 0.0, usw[3] = {0.0, 0.0, 0.0};
  for (int w = 0; w < 2; ++w) {
    if (!a.draw_wing[w] || !in_box(1 + w)) continue;
    const double d[3] = {xb[0] - a.piv[w][0], xb[1] - a.piv[w][1], xb[2] - a.piv[w][2]};
    double xw[3];
    matvec(a.Mw[w], d, xw);
    if (a.wing_shape == 3) {
      // wings_geometry.f90:161-260 draw_wing_fourier (flat, constant thickness)
      const double sf = a.safety;
      const double* bb = a.fw_bbox;
      if (!(xw[0] >= bb[0] - sf && xw[0] <= bb[1] + sf && xw[1] >= bb[2] - sf && xw[1] <= bb[3] + sf &&
            xw[2] >= bb[4] - sf && xw[2] <= bb[5] + sf)) continue;
      double theta = atan2(xw[1] - a.fw_yc, xw[0] - a.fw_xc);
      theta = (theta + PI_D) / (2.0 * PI_D);
      const long j = (long)floor(theta / a.fw_dphi);
      const double R0 = a.fw_table[j] + ((theta - (double)j * a.fw_dphi) / a.fw_dphi) * (a.fw_table[j + 1] - a.fw_table[j]);
      const double R = sqrt((xw[0] - a.fw_xc) * (xw[0] - a.fw_xc) + (xw[1] - a.fw_yc) * (xw[1] - a.fw_yc));
      const double mt = steps_d(fabs(xw[2] - 0.0), 0.5 * a.thickness, a.smooth) * steps_d(R, R0, a.smooth);
      if (!mask_loaded) { m = (double)mask[p]; mask_loaded = true; }
      wing_point(a, w, xw, mt, m, c, changed, u_s, wsum, usw);
      continue;
    }
    if (!(xw[1] >= -a.safety && xw[1] <= a.L_span + a.safety)) continue;
    if (!(fabs(xw[2]) <= 0.5 * a.thickness + a.safety)) continue;
    double x_top, x_bot;
    if (a.wing_shape == 1) {
      const double ab = 0.5 * a.L_span;
      const double q = 1.0 - ((xw[1] - ab) * (xw[1] - ab)) / (ab * ab);
      if (q >= 0.0) { x_top = sqrt((a.b_top * a.b_top) * q); x_bot = -sqrt((a.b_bot * a.b_bot) * q); }
      else { x_top = 0.0; x_bot = 0.0; }
    } else {
      x_top = a.b_top; x_bot = -a.b_bot;
    }
    if (!(xw[0] > x_bot - a.safety && xw[0] < x_top + a.safety)) continue;
    // the reference's taper jumps from 1/2 to 1 at the root plane y = 0; the
    // differentiable variant takes the smaller of the two end tapers
    const double yt = a.smooth_mode ? fmin(steps_d(-xw[1], 0.0, a.smooth), steps_d(xw[1], a.L_span, a.smooth))
                    : (xw[1] < 0.0) ? steps_d(-xw[1], 0.0, a.smooth) : steps_d(xw[1], a.L_span, a.smooth);
    const double zt = steps_d(fabs(xw[2]), 0.5 * a.thickness, a.smooth);
    const double xt = (xw[0] < 0.0) ? steps_d(-xw[0], -x_bot, a.smooth) : steps_d(xw[0], x_top, a.smooth);
    const double mt = zt * yt * xt;
    if (!mask_loaded) { m = (double)mask[p]; mask_loaded = true; }
    wing_point(a, w, xw, mt, m, c, changed, u_s, wsum, usw);
  }
  const bool new_wing = (c == 2 || c == 3);
  const bool add_body_velocity = a.smooth_mode ? wsum > 0.0 : (new_wing && m > 0.0);
  if (add_body_velocity) {
    if (a.smooth_mode) {
      // us = sum_w m_w us_w / (m_body + sum_w m_w): continuous where parts overlap
      const double wn = 1.0 / (mbody + wsum);
#pragma unroll
      for (int i = 0; i < 3; ++i) u_s[i] = usw[i] * wn;
    }
    const double v[3] = {a.vb[0] + a.rot_body_b[1] * xb[2] - a.rot_body_b[2] * xb[1],
                         a.vb[1] + a.rot_body_b[2] * xb[0] - a.rot_body_b[0] * xb[2],
                         a.vb[2] + a.rot_body_b[0] * xb[1] - a.rot_body_b[1] * xb[0]};
    const double t[3] = {u_s[0] + v[0], u_s[1] + v[1], u_s[2] + v[2]};
    matTvec(a.Mb, t, u_s);
  }
  if (changed) { mask[p] = (T)m; color[p] = c; }
  if (a.full_reset || old_wing || new_wing) {
    us[p] = (T)u_s[0]; us[g.npts + p] = (T)u_s[1]; us[2 * g.npts + p] = (T)u_s[2];
  }
}

template <typename T>
void launch_insect(const GridGeom& g, const InsectArgs& a, const InsectArgs* a_dev, void* mask, void* us, uint8_t* color,
                   hipStream_t s, const Batch& bat) {
  const int nt = 256;
  if (a.reg_lo[0] > a.reg_hi[0] || a.reg_lo[1] > a.reg_hi[1] || a.reg_lo[2] > a.reg_hi[2]) return;  // nothing to update
  for (int i = 0; i < 3; ++i) {
    const int n = i == 0 ? g.nx : (i == 1 ? g.ny : g.nz);
    if (a.reg_lo[i] < 0 || a.reg_hi[i] >= n || a.reg_lo[i] > a.reg_hi[i])
      throw std::runtime_error("insect mask: region outside the grid");
  }
  const int ex = a.reg_hi[0] - a.reg_lo[0] + 1, ey = a.reg_hi[1] - a.reg_lo[1] + 1, ez = a.reg_hi[2] - a.reg_lo[2] + 1;
  const int nm = batch_n(bat);
  if ((long)ez * nm > 65535) throw std::runtime_error("insect mask: batch too large for one launch");
  hipLaunchKernelGGL(insect_k<T>, dim3((ex + nt - 1) / nt, ey, ez * nm), dim3(nt), 0, s, g, a_dev, (T*)mask, (T*)us, color,
                     ez, nm > 1 ? bat.mstride : 0L);
}

// ------------------------------------------------------------- vortex ring
// inicond/hyd/init_fields_fsi.f90:381-435 (vorticity in physical space)
template <typename T>
__global__ void vortex_ring_k(GridGeom g, double x0, double y0, double z0, double r00, double a, double gamma0, T* vort) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.npts) return;
  const int ix = p % g.nx, iy = (p / g.nx) % g.ny, iz = p / ((long)g.nx * g.ny);
  const double x = ix * g.dx, y = iy * g.dy, z = iz * g.dz;
  const double r = sqrt((y - y0) * (y - y0) + (z - z0) * (z - z0));
  const double om = (gamma0 / (PI_D * a * a)) * exp(-((x - x0) * (x - x0) + (r - r00) * (r - r00)) / (a * a));
  double wy = 0.0, wz = 0.0;
  if (fabs(r) > 1.0e-12) { wy = -om * ((z - z0) / r); wz = om * ((y - y0) / r); }
  vort[p] = T(0); vort[g.npts + p] = (T)wy; vort[2 * g.npts + p] = (T)wz;
}
template <typename T>
void launch_vortex_ring(const GridGeom& g, double x0, double y0, double z0, double r00, double a, double gamma0, void* vort, hipStream_t s) {
  const int nt = 256;
  hipLaunchKernelGGL(vortex_ring_k<T>, dim3((g.npts + nt - 1) / nt), dim3(nt), 0, s, g, x0, y0, z0, r00, a, gamma0, (T*)vort);
}

// --------------------------------------------------------- spectral helpers
template <typename T>
__global__ void biot_savart_k(GridGeom g, const double* kx, const double* ky, const double* kz,
                              const cplx<T>* vk, cplx<T>* uk) {
  // basic_operators.f90:1081-1126
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.ncplx) return;
  const int iz = p % g.nz, iy = (p / g.nz) % g.ny, ix = p / ((long)g.nz * g.ny);
  const double k1 = kx[ix], k2_ = ky[iy], k3 = kz[iz];
  const double kk = k1 * k1 + k2_ * k2_ + k3 * k3;
  const long n = g.ncplx;
  if (kk >= 1.0e-13) {
    double sr[3], si[3];
    for (int c = 0; c < 3; ++c) { sr[c] = (double)vk[c * n + p].x / kk; si[c] = (double)vk[c * n + p].y / kk; }
    const double cr[3] = {k2_ * sr[2] - k3 * sr[1], k3 * sr[0] - k1 * sr[2], k1 * sr[1] - k2_ * sr[0]};
    const double ci[3] = {k2_ * si[2] - k3 * si[1], k3 * si[0] - k1 * si[2], k1 * si[1] - k2_ * si[0]};
    for (int c = 0; c < 3; ++c) uk[c * n + p] = mk<T>((T)(-ci[c]), (T)cr[c]);
  } else {
    for (int c = 0; c < 3; ++c) uk[c * n + p] = mk<T>(T(0), T(0));
  }
}
template <typename T>
void launch_biot_savart(const GridGeom& g, const double* kx, const double* ky, const double* kz, const void* vk, void* uk, hipStream_t s) {
  const int nt = 256;
  hipLaunchKernelGGL(biot_savart_k<T>, dim3((g.ncplx + nt - 1) / nt), dim3(nt), 0, s, g, kx, ky, kz,
                     (const cplx<T>*)vk, (cplx<T>*)uk);
}

template <typename T>
__global__ void divergence_k(GridGeom g, const double* kx, const double* ky, const double* kz,
                             const cplx<T>* uk, cplx<T>* dk) {
  // basic_operators.f90:707-734
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.ncplx) return;
  const int iz = p % g.nz, iy = (p / g.nz) % g.ny, ix = p / ((long)g.nz * g.ny);
  const long n = g.ncplx;
  const double re = kx[ix] * (double)uk[p].x + ky[iy] * (double)uk[n + p].x + kz[iz] * (double)uk[2 * n + p].x;
  const double im = kx[ix] * (double)uk[p].y + ky[iy] * (double)uk[n + p].y + kz[iz] * (double)uk[2 * n + p].y;
  dk[p] = mk<T>((T)(-im), (T)re);
}
template <typename T>
void launch_divergence(const GridGeom& g, const double* kx, const double* ky, const double* kz, const void* uk, void* dk, hipStream_t s) {
  const int nt = 256;
  hipLaunchKernelGGL(divergence_k<T>, dim3((g.ncplx + nt - 1) / nt), dim3(nt), 0, s, g, kx, ky, kz,
                     (const cplx<T>*)uk, (cplx<T>*)dk);
}

template <typename T>
__global__ void set_mode0_k(cplx<T>* uk, long off, T re, T im) { uk[off] = mk<T>(re, im); }
template <typename T>
void launch_set_mode0(const GridGeom& g, void* uk, int comp, double re, double im, hipStream_t s) {
  hipLaunchKernelGGL(set_mode0_k<T>, dim3(1), dim3(1), 0, s, (cplx<T>*)uk, (long)comp * g.ncplx, (T)re, (T)im);
}

// --------------------------------------------------------------- diagnostics
static constexpr int DIAG_NT = 256;
static constexpr int DIAG_PPT = 8;  // points per thread

// Colour-resolved sums (force, torques, residual) only receive contributions
// where mask > 0: a wave adds them, reduced over its lanes in a fixed tree, to
// its own LDS row only for point batches that touch the mask.  The other sums
// stay in registers.  All reductions have a fixed order (reproducible).
constexpr int DIAG_NCOL = 40;  // DG_FORCE .. DG_TORQUE at 0..35, the 4 DG_URES minima at 36..39
static_assert(DG_FORCE == 0 && DG_TORQUE0 == 12 && DG_TORQUE == 24 && DG_POWER == 36, "colour slots");

__device__ __forceinline__ double wred(double x, int kind) {  // 0 sum, 1 max, 2 min
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double y = __shfl_xor(x, o, 64);
    x = kind == 2 ? fmin(x, y) : (kind == 1 ? nanmax(x, y) : x + y);
  }
  return x;
}

template <typename T>
__global__ __launch_bounds__(DIAG_NT) void diag_k(GridGeom g, DiagArgs a) {
  const T* u = (const T*)a.u; const T* w = (const T*)a.w; const T* dv = (const T*)a.div;
  const T* mask = (const T*)a.mask; const T* us = (const T*)a.us;
  const long n = g.npts;
  constexpr int NW = DIAG_NT / 64;
  __shared__ double colacc[NW][DIAG_NCOL];
  __shared__ double red[NW][DIAG_NV];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (lane < DIAG_NCOL) colacc[wv][lane] = (lane >= 36) ? 9.0e9 : 0.0;
  // per-thread sums of the colour-independent quantities
  double pw[3] = {0, 0, 0}, ek[3] = {0, 0, 0}, ekf[3] = {0, 0, 0}, ds[3] = {0, 0, 0}, dsf[3] = {0, 0, 0},
         mus[3] = {0, 0, 0};
  double dmax = -1.0e300, dmaxf = -1.0e300, mvol = 0.0;
  const long base = ((long)blockIdx.x * DIAG_NT) * DIAG_PPT + threadIdx.x;
  for (int q = 0; q < DIAG_PPT; ++q) {
    const long p = base + (long)q * DIAG_NT;
    const bool live = p < n;
    const long pp = live ? p : 0;
    const double m = live ? (double)mask[pp] : 0.0;
    const double uu[3] = {(double)u[pp], (double)u[n + pp], (double)u[2 * n + pp]};
    const double ww[3] = {a.want_w ? (double)w[pp] : 0.0, a.want_w ? (double)w[n + pp] : 0.0,
                          a.want_w ? (double)w[2 * n + pp] : 0.0};
    const double ss[3] = {(double)us[pp], (double)us[n + pp], (double)us[2 * n + pp]};
    const double du[3] = {uu[0] - ss[0], uu[1] - ss[1], uu[2] - ss[2]};
    const double pen[3] = {-m * du[0] * a.eps_inv, -m * du[1] * a.eps_inv, -m * du[2] * a.eps_inv};
    if (live) {
      const double fm = 1.0 - m;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        pw[i] += ss[i] * pen[i];
        ek[i] += uu[i] * uu[i];
        ekf[i] += (uu[i] * fm) * (uu[i] * fm);
        ds[i] += ww[i] * ww[i];
        dsf[i] += (ww[i] * fm) * (ww[i] * fm);
        mus[i] += m * ss[i];
      }
      const double d = a.want_w ? (double)dv[pp] : 0.0;
      dmax = nanmax(dmax, d);
      dmaxf = nanmax(dmaxf, d * fm);
      mvol += m;
    }
    if (__any(m > 0.0)) {
      const int ix = (int)(pp % g.nx), iy = (int)((pp / g.nx) % g.ny), iz = (int)(pp / ((long)g.nx * g.ny));
      const int col = live ? color_clamp(a.color[pp]) : 4;
      double lev[3] = {ix * g.dx - a.x0[0], iy * g.dy - a.x0[1], iz * g.dz - a.x0[2]};
      periodize_d(lev, g);
      const double t0[3] = {-(lev[1] * pen[2] - lev[2] * pen[1]), -(lev[2] * pen[0] - lev[0] * pen[2]),
                            -(lev[0] * pen[1] - lev[1] * pen[0])};
      if (a.insect && (col == 2 || col == 3)) {
        const double* pv = (col == 2) ? a.piv_l : a.piv_r;
        lev[0] -= pv[0]; lev[1] -= pv[1]; lev[2] -= pv[2];
        periodize_d(lev, g);
      }
      const double tp[3] = {-(lev[1] * pen[2] - lev[2] * pen[1]), -(lev[2] * pen[0] - lev[0] * pen[2]),
                            -(lev[0] * pen[1] - lev[1] * pen[0])};
      const double d2 = du[0] * du[0] + du[1] * du[1] + du[2] * du[2];
      for (int c = 0; c < 4; ++c) {
        const bool on = (col == c);
        if (!__any(on)) continue;
        double r[10];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          r[i] = wred(on ? -pen[i] : 0.0, 0);
          r[3 + i] = wred(on ? t0[i] : 0.0, 0);
          r[6 + i] = wred(on ? tp[i] : 0.0, 0);
        }
        r[9] = wred((on && m > 0.0) ? d2 : 9.0e9, 2);
        if (lane == 0) {
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            colacc[wv][DG_FORCE + 3 * c + i] += r[i];
            colacc[wv][DG_TORQUE0 + 3 * c + i] += r[3 + i];
            colacc[wv][DG_TORQUE + 3 * c + i] += r[6 + i];
          }
          colacc[wv][36 + c] = fmin(colacc[wv][36 + c], r[9]);  // residual slots 36..39
        }
      }
    }
  }
  // wave totals of the register sums
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    pw[i] = wred(pw[i], 0); ek[i] = wred(ek[i], 0); ekf[i] = wred(ekf[i], 0);
    ds[i] = wred(ds[i], 0); dsf[i] = wred(dsf[i], 0); mus[i] = wred(mus[i], 0);
  }
  dmax = wred(dmax, 1); dmaxf = wred(dmaxf, 1); mvol = wred(mvol, 0);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      red[wv][DG_POWER + i] = pw[i]; red[wv][DG_EK + i] = ek[i]; red[wv][DG_EKF + i] = ekf[i];
      red[wv][DG_DISS + i] = ds[i]; red[wv][DG_DISSF + i] = dsf[i]; red[wv][DG_MASKUS + i] = mus[i];
    }
    red[wv][DG_DIVMAX] = dmax; red[wv][DG_DIVMAXF] = dmaxf; red[wv][DG_MASKVOL] = mvol;
  }
  __syncthreads();
  if (tid < DIAG_NV) {
    const int v = tid;
    const bool colv = v < 36 || (v >= DG_URES && v < DG_URES + 4);
    const bool is_min = (v >= DG_URES && v < DG_URES + 4);
    const bool is_max = (v == DG_DIVMAX || v == DG_DIVMAXF);
    double r = colv ? colacc[0][v < 36 ? v : 36 + (v - DG_URES)] : red[0][v];
    for (int i = 1; i < NW; ++i) {
      const double x = colv ? colacc[i][v < 36 ? v : 36 + (v - DG_URES)] : red[i][v];
      r = is_min ? fmin(r, x) : (is_max ? nanmax(r, x) : r + x);
    }
    a.partial[(long)blockIdx.x * DIAG_NV + v] = r;
  }
}

template <typename T> int diag_blocks(const GridGeom& g) {
  const long per = (long)DIAG_NT * DIAG_PPT;
  return (int)((g.npts + per - 1) / per);
}
template <typename T>
void launch_diag(const GridGeom& g, const DiagArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(diag_k<T>, dim3(diag_blocks<T>(g)), dim3(DIAG_NT), 0, s, g, a);
}

// ------------------------------------------------------------ row reducer
// Combines per-block partial rows in a fixed order (bitwise reproducible):
// Neumaier-compensated sums, or max/min.  Phase 1: a grid of (chunk, value)
// blocks, each a fixed chunk of rows; phase 2: one thread per value walks the
// chunk results in order.
static constexpr int RED_CHUNKS = 64;
static constexpr int RED_NT = 256;

__device__ __forceinline__ void neu_add(double& s, double& c, double x) {
  const double t = s + x;
  c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
  s = t;
}

__global__ __launch_bounds__(RED_NT) void reduce_rows_p1(const double* partial, int nrows, int nv, const int* kinds, double* mid) {
  __shared__ double sv[RED_NT];
  __shared__ double sc[RED_NT];
  const int v = blockIdx.y, b = blockIdx.x, tid = threadIdx.x;
  const int kind = kinds[v];
  const int chunk = (nrows + RED_CHUNKS - 1) / RED_CHUNKS;
  const int r0 = b * chunk, r1 = min(nrows, r0 + chunk);
  double s = (kind == 1) ? -1.0e308 : (kind == 2 ? 1.0e308 : 0.0), comp = 0.0;
  for (int r = r0 + tid; r < r1; r += RED_NT) {
    const double x = partial[(long)r * nv + v];
    if (kind == 1) s = nanmax(s, x);
    else if (kind == 2) s = fmin(s, x);
    else neu_add(s, comp, x);
  }
  sv[tid] = s; sc[tid] = comp;
  __syncthreads();
  for (int w = RED_NT / 2; w > 0; w >>= 1) {
    if (tid < w) {
      if (kind == 1) sv[tid] = nanmax(sv[tid], sv[tid + w]);
      else if (kind == 2) sv[tid] = fmin(sv[tid], sv[tid + w]);
      else {
        double a = sv[tid], c = sc[tid] + sc[tid + w];
        neu_add(a, c, sv[tid + w]);
        sv[tid] = a; sc[tid] = c;
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    mid[(long)v * RED_CHUNKS * 2 + 2 * b] = sv[0];
    mid[(long)v * RED_CHUNKS * 2 + 2 * b + 1] = sc[0];
  }
}

__global__ void reduce_rows_p2(const double* mid, int nv, const int* kinds, double* out) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const int kind = kinds[v];
  double s = mid[(long)v * RED_CHUNKS * 2], c = mid[(long)v * RED_CHUNKS * 2 + 1];
  for (int b = 1; b < RED_CHUNKS; ++b) {
    const double x = mid[(long)v * RED_CHUNKS * 2 + 2 * b];
    if (kind == 1) s = nanmax(s, x);
    else if (kind == 2) s = fmin(s, x);
    else { c += mid[(long)v * RED_CHUNKS * 2 + 2 * b + 1]; neu_add(s, c, x); }
  }
  out[v] = (kind == 0) ? s + c : s;
}

size_t reduce_rows_scratch(int nv) { return (size_t)nv * RED_CHUNKS * 2; }

void launch_reduce_rows(const double* partial, int nrows, int nv, const int* kinds_dev, double* mid, double* out, hipStream_t s) {
  hipLaunchKernelGGL(reduce_rows_p1, dim3(RED_CHUNKS, nv), dim3(RED_NT), 0, s, partial, nrows, nv, kinds_dev, mid);
  hipLaunchKernelGGL(reduce_rows_p2, dim3((nv + 63) / 64), dim3(64), 0, s, mid, nv, kinds_dev, out);
}

// ------------------------------------------------------------ conversions
template <typename T>
__global__ void from_f64_k(const double* src, T* dst, long n) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) dst[p] = (T)src[p];
}
template <typename T>
__global__ void to_f64_k(const T* src, double* dst, long n) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) dst[p] = (double)src[p];
}
template <typename T> void launch_from_f64(const double* src, void* dst, long n, hipStream_t s) {
  const int nt = 256;
  hipLaunchKernelGGL(from_f64_k<T>, dim3((n + nt - 1) / nt), dim3(nt), 0, s, src, (T*)dst, n);
}
template <typename T> void launch_to_f64(const void* src, double* dst, long n, hipStream_t s) {
  const int nt = 256;
  hipLaunchKernelGGL(to_f64_k<T>, dim3((n + nt - 1) / nt), dim3(nt), 0, s, (const T*)src, dst, n);
}

// save_fields.f90:129-133: p = Pi - |u|^2 / 2 (rotational form), in place
template <typename T>
__global__ void pressure_phys_k(long npts, const T* __restrict__ u, T* __restrict__ p) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npts) return;
  const double a = u[i], b = u[npts + i], c = u[2 * npts + i];
  p[i] = (T)((double)p[i] - 0.5 * ((a * a + b * b) + c * c));
}
template <typename T>
void launch_pressure_phys(long npts, const void* u, void* p, hipStream_t s) {
  const int nt = 256;
  hipLaunchKernelGGL(pressure_phys_k<T>, dim3((npts + nt - 1) / nt), dim3(nt), 0, s, npts, (const T*)u, (T*)p);
}
template void launch_pressure_phys<float>(long, const void*, void*, hipStream_t);
template void launch_pressure_phys<double>(long, const void*, void*, hipStream_t);

// integrals.f90:522-552 compute_energies_k: per-block compensated FP64 sums
// of w |uk|^2 over the 3 spectral fields (S layout, w = 1/2 on the kx = 0 and
// kx = nx/2 planes), one partial per block for the deterministic row reducer
template <typename T>
__global__ __launch_bounds__(256) void energy_k_k(GridGeom g, const typename V2<T>::t* __restrict__ uk,
                                                  double* __restrict__ partial) {
  __shared__ double red[256 / 64];
  const long n = 3 * g.ncplx, plane = (long)g.ny * g.nz;
  double s = 0.0, c = 0.0;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
    const long ix = (e % g.ncplx) / plane;
    const double w = (ix == 0 || ix == g.nx / 2) ? 0.5 : 1.0;
    const double re = (double)uk[e].x, im = (double)uk[e].y;
    const double v = w * (re * re + im * im);
    const double t = s + v;
    c += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
    s = t;
  }
  double v = s + c;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}
int energy_k_blocks() { return 512; }
template <typename T>
void launch_energy_k(const GridGeom& g, const void* uk, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(energy_k_k<T>, dim3(energy_k_blocks()), dim3(256), 0, s, g,
                     (const typename V2<T>::t*)uk, partial);
}
template void launch_energy_k<float>(const GridGeom&, const void*, double*, hipStream_t);
template void launch_energy_k<double>(const GridGeom&, const void*, double*, hipStream_t);

// vis.f90:9-66 cal_vis: exp(-dt nu k^2), 0 on the dealiased modes, in the
// spectral layout (ix, iy, iz), iz fastest
__global__ void cal_vis_k(GridGeom g, const double* kx, const double* ky, const double* kz, const double* kxt2,
                          const double* kyt2, const double* kzt2, double dt, double nu, int dealias, double* vis) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.ncplx) return;
  const int iz = (int)(e % g.nz), iy = (int)((e / g.nz) % g.ny), ix = (int)(e / ((long)g.nz * g.ny));
  const double k2 = (kx[ix] * kx[ix] + ky[iy] * ky[iy]) + kz[iz] * kz[iz];
  const bool alias = ((kxt2[ix] + kyt2[iy]) + kzt2[iz]) >= 1.0;
  vis[e] = (alias && dealias) ? 0.0 : exp(-dt * nu * k2);
}
void launch_cal_vis(const GridGeom& g, const double* kx, const double* ky, const double* kz, const double* kxt2,
                    const double* kyt2, const double* kzt2, double dt, double nu, int dealias, double* vis,
                    hipStream_t s) {
  const int nt = 256;
  hipLaunchKernelGGL(cal_vis_k, dim3((g.ncplx + nt - 1) / nt), dim3(nt), 0, s, g, kx, ky, kz, kxt2, kyt2, kzt2, dt,
                     nu, dealias, vis);
}

// |w| of 3 physical fields (save_fields.f90:156-160 vorabs)
template <typename T>
__global__ void magnitude_k(long npts, const T* __restrict__ w, T* __restrict__ out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  const double a = w[p], b = w[npts + p], c = w[2 * npts + p];
  out[p] = (T)sqrt(a * a + b * b + c * c);
}
template <typename T>
void launch_magnitude(long npts, const void* w, void* out, hipStream_t s) {
  const int nt = 256;
  hipLaunchKernelGGL(magnitude_k<T>, dim3((npts + nt - 1) / nt), dim3(nt), 0, s, npts, (const T*)w, (T*)out);
}
template void launch_magnitude<float>(long, const void*, void*, hipStream_t);
template void launch_magnitude<double>(long, const void*, void*, hipStream_t);

__global__ void zero_k(uint32_t* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}
void launch_zero(void* p, size_t bytes, hipStream_t s) {
  // a fill kernel rather than hipMemsetAsync, so the call may sit inside a captured graph
  const size_t n = bytes / 4;
  const int nt = 256;
  if (n) hipLaunchKernelGGL(zero_k, dim3((n + nt - 1) / nt), dim3(nt), 0, s, (uint32_t*)p, n);
}

#define INST(T)                                                                                              \
  template void launch_sphere<T>(const GridGeom&, const SphereArgs&, void*, void*, uint8_t*, hipStream_t);   \
  template void launch_cavity<T>(const GridGeom&, const CavityArgs&, void*, void*, uint8_t*, hipStream_t);   \
  template void launch_insect<T>(const GridGeom&, const InsectArgs&, const InsectArgs*, void*, void*, uint8_t*, hipStream_t, \
                                 const Batch&);                                                                           \
  template void launch_vortex_ring<T>(const GridGeom&, double, double, double, double, double, double, void*, hipStream_t); \
  template void launch_biot_savart<T>(const GridGeom&, const double*, const double*, const double*, const void*, void*, hipStream_t); \
  template void launch_divergence<T>(const GridGeom&, const double*, const double*, const double*, const void*, void*, hipStream_t); \
  template void launch_set_mode0<T>(const GridGeom&, void*, int, double, double, hipStream_t);               \
  template int diag_blocks<T>(const GridGeom&);                                                              \
  template void launch_diag<T>(const GridGeom&, const DiagArgs&, hipStream_t);                               \
  template void launch_from_f64<T>(const double*, void*, long, hipStream_t);                                 \
  template void launch_to_f64<T>(const void*, double*, long, hipStream_t);
INST(float)
INST(double)

}  // namespace flusi
