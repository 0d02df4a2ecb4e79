// FP64 CPU restatement of FLUSI's tethered-insect RK4 path, OpenMP-parallel
// with its own FFT -- TEST INFRASTRUCTURE ONLY.
//
// This program is the CPU baseline the benchmark times on the host cores, and
// a second checker beside oracle/flusi_oracle.py, which it must match (see
// tests/test_cpu_restatement.py).  Nothing in the product path runs it.
//
// Scope: the benchmark workload's configuration class -- a tethered insect
// with the ellipsoid_plain body and TwoEllipses wings driven by from_file
// Fourier kinematics, penalised Navier-Stokes in a periodic box, RK4 with
// explicit diffusion, 2/3 dealiasing, fixed dt, free mean flow, no sponge.
// Each routine follows the reference routine cited beside it (paths relative
// to the reference's src/); the arrays use the reference's single-rank
// layouts: physical fields (ix,iy,iz) with ix fastest, spectral half fields
// (iz,iy,ix) with iz fastest, i.e. [kx][ky][kz] in C order.
//
//   flusi_cpu <ini> <nsteps> [threads] [uk_in|-] [uk_out|-]
//
// Runs nsteps RK4 steps from the ini's quiescent start (or from uk_in, raw
// complex128 [3][nx/2+1][ny][nz]) and prints one JSON line with the set-up
// and stepping times; uk_out receives the final state in the same format.
#include <omp.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

const double PI = 3.14159265358979323846264338327950288;

struct cd {
  double re, im;
};
inline cd operator+(cd a, cd b) { return {a.re + b.re, a.im + b.im}; }
inline cd operator-(cd a, cd b) { return {a.re - b.re, a.im - b.im}; }
inline cd operator*(cd a, cd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline cd operator*(double s, cd a) { return {s * a.re, s * a.im}; }
inline cd conj(cd a) { return {a.re, -a.im}; }
inline cd expi(double t) { return {std::cos(t), std::sin(t)}; }

// ------------------------------------------------------------------- ini
// ini_files_parser.f90:249-319, 825-891: [section] key=value; lines, the
// value ending at ';', comments after ';' or '#', "<a>*dx" scaled by dx.
struct Ini {
  std::map<std::string, std::string> kv;
  explicit Ini(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::string line, sec;
    while (std::getline(f, line)) {
      size_t a = line.find_first_not_of(" \t\r");
      if (a == std::string::npos) continue;
      line = line.substr(a);
      if (line[0] == ';' || line[0] == '#' || line[0] == '!') continue;
      if (line[0] == '[') {
        sec = line.substr(1, line.find(']') - 1);
        continue;
      }
      size_t eq = line.find('=');
      if (eq == std::string::npos) continue;
      std::string key = line.substr(0, eq), val = line.substr(eq + 1);
      size_t sc = val.find(';');
      if (sc != std::string::npos) val = val.substr(0, sc);
      while (!key.empty() && (key.back() == ' ' || key.back() == '\t')) key.pop_back();
      kv[sec + "/" + key] = val;
    }
  }
  bool has(const std::string& s, const std::string& k) const { return kv.count(s + "/" + k) > 0; }
  std::string str(const std::string& s, const std::string& k, const std::string& d) const {
    auto it = kv.find(s + "/" + k);
    if (it == kv.end()) return d;
    std::string v = it->second;
    size_t a = v.find_first_not_of(" \t"), b = v.find_last_not_of(" \t");
    return a == std::string::npos ? d : v.substr(a, b - a + 1);
  }
  double dbl(const std::string& s, const std::string& k, double d, double dx = 0.0) const {
    std::string v = str(s, k, "");
    if (v.empty()) return d;
    size_t star = v.find("*dx");
    if (star != std::string::npos) return std::stod(v.substr(0, star)) * dx;
    return std::stod(v);
  }
  std::vector<double> vec(const std::string& s, const std::string& k, std::vector<double> d) const {
    std::string v = str(s, k, "");
    if (v.empty()) return d;
    std::istringstream in(v);
    std::vector<double> r;
    double x;
    while (in >> x) r.push_back(x);
    return r;
  }
};

// ------------------------------------------------------------------- FFT
// Mixed-radix (2, 3, 4, 5) Stockham transform of one line: stage s of radix
// r with sub-transform length ns reads in[j + q n/r], multiplies input q by
// w^(q (j mod ns)) with w = exp(sign 2 pi i / (ns r)), applies the radix-r DFT
// and writes out[(j / ns) ns r + j mod ns + q ns].  Unnormalised.
struct Line {
  int n = 0;
  std::vector<int> radix;
  std::vector<std::vector<cd>> twf, twi;  // per stage, [k * r + q]
  void init(int n_) {
    n = n_;
    radix.clear();
    int m = n;
    while (m % 4 == 0) { radix.push_back(4); m /= 4; }
    while (m % 2 == 0) { radix.push_back(2); m /= 2; }
    while (m % 3 == 0) { radix.push_back(3); m /= 3; }
    while (m % 5 == 0) { radix.push_back(5); m /= 5; }
    if (m != 1) throw std::runtime_error("FFT length must be 2^a 3^b 5^c");
    twf.clear();
    twi.clear();
    int ns = 1;
    for (int r : radix) {
      std::vector<cd> f(ns * r), b(ns * r);
      for (int k = 0; k < ns; ++k)
        for (int q = 0; q < r; ++q) {
          const double t = 2.0 * PI * (double)(q * k) / (double)(ns * r);
          f[k * r + q] = expi(-t);
          b[k * r + q] = expi(t);
        }
      twf.push_back(f);
      twi.push_back(b);
      ns *= r;
    }
  }
  static void dft(int r, int sign, cd* x) {
    if (r == 2) {
      const cd a = x[0], b = x[1];
      x[0] = a + b;
      x[1] = a - b;
    } else if (r == 4) {
      const cd s02 = x[0] + x[2], d02 = x[0] - x[2], s13 = x[1] + x[3];
      const cd t = x[1] - x[3];
      const cd d13 = sign < 0 ? cd{t.im, -t.re} : cd{-t.im, t.re};  // sign * i * t
      x[0] = s02 + s13;
      x[2] = s02 - s13;
      x[1] = d02 + d13;
      x[3] = d02 - d13;
    } else {
      cd y[5];
      for (int p = 0; p < r; ++p) {
        cd acc{0.0, 0.0};
        for (int q = 0; q < r; ++q) acc = acc + x[q] * expi(sign * 2.0 * PI * (double)((p * q) % r) / (double)r);
        y[p] = acc;
      }
      for (int p = 0; p < r; ++p) x[p] = y[p];
    }
  }
  // in-place on a (work: n more elements)
  void run(cd* a, cd* work, int sign) const {
    cd* in = a;
    cd* out = work;
    int ns = 1;
    for (size_t s = 0; s < radix.size(); ++s) {
      const int r = radix[s], nr = n / r;
      const cd* tw = sign < 0 ? twf[s].data() : twi[s].data();
      for (int j = 0; j < nr; ++j) {
        const int k = j % ns;
        cd x[5];
        for (int q = 0; q < r; ++q) x[q] = in[j + q * nr] * tw[k * r + q];
        dft(r, sign, x);
        const int base = (j / ns) * ns * r + k;
        for (int q = 0; q < r; ++q) out[base + q * ns] = x[q];
      }
      std::swap(in, out);
      ns *= r;
    }
    if (in != a) std::memcpy(a, in, sizeof(cd) * n);
  }
};

// cof_p3dfft.f90:33-83 fft (R2C times 1/N) and ifft (C2R, unnormalised, the
// imaginary parts of the kx = 0 and kx = nx/2 entries ignored), through
// length-nx/2 complex transforms of the even/odd split.
struct FFT3 {
  int nx, ny, nz, nxh;
  Line lx, ly, lz;
  std::vector<cd> twr;   // exp(-2 pi i k / nx), k <= nx/2
  std::vector<cd> mid;   // [z][y][kx]
  static const int B = 8;  // lines gathered together in the strided passes
  void init(int nx_, int ny_, int nz_) {
    nx = nx_; ny = ny_; nz = nz_; nxh = nx / 2 + 1;
    if (nx % 2) throw std::runtime_error("nx must be even");
    lx.init(nx / 2); ly.init(ny); lz.init(nz);
    twr.resize(nxh);
    for (int k = 0; k < nxh; ++k) twr[k] = expi(-2.0 * PI * (double)k / (double)nx);
    mid.assign((size_t)nz * ny * nxh, cd{0.0, 0.0});
  }
  // real [z][y][x] -> spectral [kx][ky][kz], times 1/N
  void forward(const double* f, cd* out) {
    const int m = nx / 2;
#pragma omp parallel
    {
      std::vector<cd> z(m), w(m);
#pragma omp for schedule(static)
      for (long zy = 0; zy < (long)nz * ny; ++zy) {
        const double* x = f + zy * nx;
        for (int i = 0; i < m; ++i) z[i] = {x[2 * i], x[2 * i + 1]};
        lx.run(z.data(), w.data(), -1);
        cd* o = mid.data() + zy * nxh;
        for (int k = 0; k <= m; ++k) {
          const cd zk = z[k % m], zm = conj(z[(m - k) % m]);
          const cd e = 0.5 * (zk + zm), d = zk - zm;
          const cd od = {0.5 * d.im, -0.5 * d.re};  // (zk - zm) / (2i)
          o[k] = e + twr[k] * od;
        }
      }
    }
    strided_y(-1);
    const double s = 1.0 / ((double)nx * ny * nz);
#pragma omp parallel
    {
      std::vector<cd> buf((size_t)B * nz), w(nz);
#pragma omp for schedule(static) collapse(2)
      for (int ky = 0; ky < ny; ++ky)
        for (int kb = 0; kb < nxh; kb += B) {
          const int nb = std::min(B, nxh - kb);
          for (int iz = 0; iz < nz; ++iz) {
            const cd* src = mid.data() + ((size_t)iz * ny + ky) * nxh + kb;
            for (int b = 0; b < nb; ++b) buf[(size_t)b * nz + iz] = src[b];
          }
          for (int b = 0; b < nb; ++b) {
            cd* l = buf.data() + (size_t)b * nz;
            lz.run(l, w.data(), -1);
            cd* o = out + ((size_t)(kb + b) * ny + ky) * nz;
            for (int iz = 0; iz < nz; ++iz) o[iz] = s * l[iz];
          }
        }
    }
  }
  // spectral [kx][ky][kz] -> real [z][y][x], unnormalised
  void inverse(const cd* in, double* f) {
#pragma omp parallel
    {
      std::vector<cd> buf((size_t)B * nz), w(nz);
#pragma omp for schedule(static) collapse(2)
      for (int ky = 0; ky < ny; ++ky)
        for (int kb = 0; kb < nxh; kb += B) {
          const int nb = std::min(B, nxh - kb);
          for (int b = 0; b < nb; ++b) {
            cd* l = buf.data() + (size_t)b * nz;
            std::memcpy(l, in + ((size_t)(kb + b) * ny + ky) * nz, sizeof(cd) * nz);
            lz.run(l, w.data(), +1);
          }
          for (int iz = 0; iz < nz; ++iz) {
            cd* dst = mid.data() + ((size_t)iz * ny + ky) * nxh + kb;
            for (int b = 0; b < nb; ++b) dst[b] = buf[(size_t)b * nz + iz];
          }
        }
    }
    strided_y(+1);
    const int m = nx / 2;
#pragma omp parallel
    {
      std::vector<cd> z(m), w(m);
#pragma omp for schedule(static)
      for (long zy = 0; zy < (long)nz * ny; ++zy) {
        const cd* X = mid.data() + zy * nxh;
        for (int k = 0; k < m; ++k) {
          cd a = X[k], b = conj(X[m - k]);
          if (k == 0) { a.im = 0.0; b = {X[m].re, 0.0}; }
          const cd e = a + b, d = conj(twr[k]) * (a - b);
          z[k] = {e.re - d.im, e.im + d.re};  // e + i d
        }
        lx.run(z.data(), w.data(), +1);
        double* x = f + zy * nx;
        for (int i = 0; i < m; ++i) { x[2 * i] = z[i].re; x[2 * i + 1] = z[i].im; }
      }
    }
  }
  void strided_y(int sign) {
#pragma omp parallel
    {
      std::vector<cd> buf((size_t)B * ny), w(ny);
#pragma omp for schedule(static) collapse(2)
      for (int iz = 0; iz < nz; ++iz)
        for (int kb = 0; kb < nxh; kb += B) {
          const int nb = std::min(B, nxh - kb);
          cd* base = mid.data() + (size_t)iz * ny * nxh + kb;
          for (int iy = 0; iy < ny; ++iy)
            for (int b = 0; b < nb; ++b) buf[(size_t)b * ny + iy] = base[(size_t)iy * nxh + b];
          for (int b = 0; b < nb; ++b) ly.run(buf.data() + (size_t)b * ny, w.data(), sign);
          for (int iy = 0; iy < ny; ++iy)
            for (int b = 0; b < nb; ++b) base[(size_t)iy * nxh + b] = buf[(size_t)b * ny + iy];
        }
    }
  }
};

// ------------------------------------------------------------- geometry
struct M3 {
  double a[3][3];
};
M3 mul(const M3& x, const M3& y) {
  M3 r{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) r.a[i][j] += x.a[i][k] * y.a[k][j];
  return r;
}
M3 tr(const M3& x) {
  M3 r{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = x.a[j][i];
  return r;
}
void mv(const M3& m, const double* v, double* o) {
  for (int i = 0; i < 3; ++i) o[i] = m.a[i][0] * v[0] + m.a[i][1] * v[1] + m.a[i][2] * v[2];
}
// rotation_matrices.f90:4-37 (passive)
M3 Rx(double t) { const double c = std::cos(t), s = std::sin(t); return M3{{{1, 0, 0}, {0, c, s}, {0, -s, c}}}; }
M3 Ry(double t) { const double c = std::cos(t), s = std::sin(t); return M3{{{c, 0, -s}, {0, 1, 0}, {s, 0, c}}}; }
M3 Rz(double t) { const double c = std::cos(t), s = std::sin(t); return M3{{{c, s, 0}, {-s, c, 0}, {0, 0, 1}}}; }

// module_insects.f90:408-422 steps (cosine smoothstep)
inline double steps(double x, double t, double h) {
  if (x <= t - h) return 1.0;
  if (x < t + h) return 0.5 * (1.0 + std::cos((x - t + h) * PI / (2.0 * h)));
  return 0.0;
}

// wings_motion.f90:98-171 from_file Fourier kinematics (degree, flusi
// convention); module_helpers.f90:30-55 Fseriesval
struct Kine {
  double a0[3] = {0, 0, 0};
  std::vector<double> ai[3], bi[3];
  void load(const std::string& path) {
    Ini k(path);
    if (k.str("kinematics", "type", "") != "fourier") throw std::runtime_error("kinematics: type=fourier only");
    if (k.str("kinematics", "units", "degree") != "degree") throw std::runtime_error("kinematics: degree only");
    if (k.str("kinematics", "convention", "flusi") != "flusi") throw std::runtime_error("kinematics: flusi only");
    const char* nm[3] = {"phi", "alpha", "theta"};
    for (int a = 0; a < 3; ++a) {
      const int n = (int)k.dbl("kinematics", std::string("nfft_") + nm[a], 0.0);
      a0[a] = k.dbl("kinematics", std::string("a0_") + nm[a], 0.0);
      ai[a] = k.vec("kinematics", std::string("ai_") + nm[a], std::vector<double>(n, 0.0));
      bi[a] = k.vec("kinematics", std::string("bi_") + nm[a], std::vector<double>(n, 0.0));
    }
  }
  // angle a (radians) and its rate at time t
  void eval(double t, double* ang, double* rate) const {
    const double f = 2.0 * PI;
    for (int a = 0; a < 3; ++a) {
      double u = 0.5 * a0[a], ud = 0.0;
      for (size_t i = 1; i <= ai[a].size(); ++i) {
        const double s = std::sin(f * (double)i * t), c = std::cos(f * (double)i * t);
        u = u + ai[a][i - 1] * c + bi[a][i - 1] * s;
        ud = ud + f * (double)i * (-ai[a][i - 1] * s + bi[a][i - 1] * c);
      }
      ang[a] = u * PI / 180.0;
      rate[a] = ud * PI / 180.0;
    }
  }
};

struct Sim {
  // parameters
  int nx, ny, nz, nxh;
  double xl, yl, zl, dx, dy, dz, nu, eps, dt;
  bool dealias_on;
  // insect (module_insects.f90:89-227, tethered)
  double b_top, b_bot, L_span, thick, b_body, L_body, x0[3], ypr[3], eta0, piv_l[3], piv_r[3];
  Kine kin_r, kin_l;
  bool body_drawn = false, cleaned = false;
  // fields
  FFT3 F;
  std::vector<double> kx, ky, kz;
  std::vector<unsigned char> alias;  // [kx][ky][kz]
  std::vector<double> mask, us;      // [z][y][x], us 3 fields
  std::vector<unsigned char> color;
  std::vector<cd> uk, u0, acc, nlk, wk;  // 3 spectral fields each
  std::vector<double> u, w;              // 3 physical fields each
  size_t npts, nspec;

  explicit Sim(const std::string& path) {
    Ini ini(path);
    nx = (int)ini.dbl("Resolution", "nx", 0); ny = (int)ini.dbl("Resolution", "ny", 0); nz = (int)ini.dbl("Resolution", "nz", 0);
    xl = ini.dbl("Geometry", "xl", 2 * PI); yl = ini.dbl("Geometry", "yl", 2 * PI); zl = ini.dbl("Geometry", "zl", 2 * PI);
    dx = xl / nx; dy = yl / ny; dz = zl / nz;
    nu = ini.dbl("ReynoldsNumber", "nu", 0.0);
    eps = ini.dbl("Penalization", "eps", 1e-2);
    dt = ini.dbl("Time", "dt_fixed", 0.0);
    dealias_on = (int)ini.dbl("Dealiasing", "iDealias", 1) == 1;
    auto need = [&](const std::string& s, const std::string& k, const std::string& v) {
      if (ini.str(s, k, "") != v) throw std::runtime_error("out of scope: " + s + "/" + k + " must be " + v);
    };
    need("Time", "iTimeMethodFluid", "RK4");
    need("Penalization", "iMask", "Insect");
    need("Insects", "BodyType", "ellipsoid_plain");
    need("Insects", "WingShape", "TwoEllipses");
    need("Insects", "BodyMotion", "tethered");
    for (const char* c : {"iMeanFlow_x", "iMeanFlow_y", "iMeanFlow_z"}) need("MeanFlow", c, "free");
    if (dt <= 0.0) throw std::runtime_error("out of scope: dt_fixed must be set");
    if (ini.str("Penalization", "periodic", "no") == "yes") throw std::runtime_error("out of scope: periodic insect");
    if (ini.str("Sponge", "iVorticitySponge", "no") == "yes") throw std::runtime_error("out of scope: sponge");
    b_top = ini.dbl("Insects", "b_top", 0.0); b_bot = ini.dbl("Insects", "b_bot", 0.0);
    L_span = ini.dbl("Insects", "L_span", 0.0);
    thick = ini.dbl("Insects", "WingThickness", 4.0 * dx, dx);
    b_body = ini.dbl("Insects", "b_body", 0.1); L_body = ini.dbl("Insects", "L_body", 1.0);
    auto v3 = [&](const std::string& k, std::vector<double> d, double* o) {
      std::vector<double> v = ini.vec("Insects", k, d);
      for (int i = 0; i < 3; ++i) o[i] = v.at(i);
    };
    v3("x0", {0.5 * xl, 0.5 * yl, 0.5 * zl}, x0);
    v3("yawpitchroll_0", {0, 0, 0}, ypr);
    for (double& a : ypr) a *= PI / 180.0;
    eta0 = ini.dbl("Insects", "eta0", 0.0) * PI / 180.0;
    v3("x_pivot_l", {0.0, b_body, 0.0}, piv_l);
    v3("x_pivot_r", {0.0, -b_body, 0.0}, piv_r);
    const std::string dir = path.find('/') == std::string::npos ? "." : path.substr(0, path.rfind('/'));
    for (auto [side, kn] : {std::pair<const char*, Kine*>{"FlappingMotion_right", &kin_r}, {"FlappingMotion_left", &kin_l}}) {
      const std::string pr = ini.str("Insects", side, "");
      const size_t p = pr.find("from_file::");
      if (p == std::string::npos) throw std::runtime_error("out of scope: from_file kinematics only");
      std::string f = pr.substr(p + 11);
      if (f[0] != '/') f = dir + "/" + f;
      kn->load(f);
    }
    nxh = nx / 2 + 1;
    npts = (size_t)nx * ny * nz;
    nspec = (size_t)nxh * ny * nz;
    F.init(nx, ny, nz);
    // cof_p3dfft.f90:661-680 wavenumbers; vis.f90:23-53 dealias predicate
    kx.resize(nxh); ky.resize(ny); kz.resize(nz);
    const double sx = 2 * PI / xl, sy = 2 * PI / yl, sz = 2 * PI / zl;
    for (int i = 0; i < nxh; ++i) kx[i] = sx * i;
    for (int i = 0; i < ny; ++i) ky[i] = sy * (double)(((i + ny / 2) % ny) - ny / 2);
    for (int i = 0; i < nz; ++i) kz[i] = sz * (double)(((i + nz / 2) % nz) - nz / 2);
    const double kxt = (2.0 / 3.0) * (double)(nx / 2 - 1), kyt = (2.0 / 3.0) * (double)(ny / 2 - 1),
                 kzt = (2.0 / 3.0) * (double)(nz / 2 - 1);
    alias.resize(nspec);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nxh; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
          const double a = (kx[i] / sx) / kxt, b = (ky[j] / sy) / kyt, c = (kz[k] / sz) / kzt;
          alias[((size_t)i * ny + j) * nz + k] = ((a * a + b * b) + c * c) >= 1.0;
        }
    auto zc = [&](std::vector<cd>& v) {
      v.resize(3 * nspec);
#pragma omp parallel for schedule(static)
      for (long i = 0; i < (long)v.size(); ++i) v[i] = {0.0, 0.0};
    };
    auto zr = [&](std::vector<double>& v, size_t n) {
      v.resize(n);
#pragma omp parallel for schedule(static)
      for (long i = 0; i < (long)n; ++i) v[i] = 0.0;
    };
    zc(uk); zc(u0); zc(acc); zc(nlk); zc(wk);
    zr(u, 3 * npts); zr(w, 3 * npts); zr(mask, npts); zr(us, 3 * npts);
    color.assign(npts, 0);
  }

  // ---- insect (module_insects.f90:281-399, body_geometry.f90:511-595 with
  // b_body and L_body from the file, wings_geometry.f90:5-110, 621-712)
  struct Motion {
    M3 Mb, Mwl, Mwr;
    double rl[3], rr[3];  // rot_rel_l_w, rot_rel_r_w
    double vb[3], rb[3];
  };
  Motion motion(double t) const {
    Motion m{};
    const double psi = ypr[2], beta = ypr[1], gamma = ypr[0];
    m.Mb = mul(Rx(psi), mul(Ry(beta), Rz(gamma)));
    for (int i = 0; i < 3; ++i) { m.vb[i] = 0.0; m.rb[i] = 0.0; }  // tethered: no body velocity
    double al[3], dl[3], ar[3], dr[3];
    kin_l.eval(t, al, dl);
    kin_r.eval(t, ar, dr);
    // module_insects.f90:650-750 wing_angular_velocities; angles (phi, alpha, theta)
    const M3 Msl = Ry(eta0), Msr = mul(Rx(PI), Ry(eta0));
    const M3 M1l = Ry(al[1]), M2l = Rz(al[2]), M3l = Rx(al[0]);
    const M3 M1r = Ry(-ar[1]), M2r = Rz(ar[2]), M3r = Rx(-ar[0]);
    m.Mwl = mul(M1l, mul(M2l, mul(M3l, Msl)));
    m.Mwr = mul(M1r, mul(M2r, mul(M3r, Msr)));
    auto rel = [](const M3& Mw, const M3& Ms, const M3& A, const M3& T, const M3& P, const double* ra, const double* rt,
                  const double* rp, double* out) {
      double t1[3], t2[3], t3[3], t4[3], t5[3];
      mv(tr(A), ra, t1);
      for (int i = 0; i < 3; ++i) t1[i] += rt[i];
      mv(tr(T), t1, t2);
      for (int i = 0; i < 3; ++i) t2[i] += rp[i];
      mv(tr(P), t2, t3);
      mv(tr(Ms), t3, t4);
      mv(Mw, t4, t5);
      for (int i = 0; i < 3; ++i) out[i] = t5[i];
    };
    const double ral[3] = {0, dl[1], 0}, rtl[3] = {0, 0, dl[2]}, rpl[3] = {dl[0], 0, 0};
    const double rar[3] = {0, -dr[1], 0}, rtr[3] = {0, 0, dr[2]}, rpr[3] = {-dr[0], 0, 0};
    rel(m.Mwl, Msl, M1l, M2l, M3l, ral, rtl, rpl, m.rl);
    rel(m.Mwr, Msr, M1r, M2r, M3r, rar, rtr, rpr, m.rr);
    return m;
  }
  void create_mask(double t) {
    const Motion m = motion(t);
    const double smooth = 1.0 * std::max(dx, std::max(dy, dz)), safety = 3.5 * smooth;
    const M3 Mbi = tr(m.Mb);
    const bool first = !cleaned;
    cleaned = true;
    const bool draw_body = !body_drawn;
    body_drawn = true;
#pragma omp parallel for schedule(static) collapse(2)
    for (int iz = 0; iz < nz; ++iz)
      for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
          const size_t p = ((size_t)iz * ny + iy) * nx + ix;
          // delete_old_mask (module_insects.f90:844-883): a resting body keeps
          // its mask; wing points and all of us are cleared
          if (first) { mask[p] = 0.0; color[p] = 0; }
          else if (color[p] == 2 || color[p] == 3) { mask[p] = 0.0; color[p] = 0; }
          us[p] = 0.0; us[npts + p] = 0.0; us[2 * npts + p] = 0.0;
          const double xg[3] = {ix * dx - x0[0], iy * dy - x0[1], iz * dz - x0[2]};
          double xb[3];
          mv(m.Mb, xg, xb);
          if (draw_body) {
            const double a = L_body / 2.0, b = b_body;
            if (std::fabs(xb[1]) <= b + safety && std::fabs(xb[2]) <= b + safety && std::fabs(xb[0]) < L_body / 2 + safety &&
                (xb[0] / a) * (xb[0] / a) <= 1.0) {
              const double R = std::sqrt(xb[1] * xb[1] + xb[2] * xb[2]);
              const double R0 = std::sqrt(b * b * (1.0 - (xb[0] / a) * (xb[0] / a)));
              if (R < R0 + safety) {
                mask[p] = std::max(steps(R, R0, smooth), mask[p]);
                color[p] = 1;
              }
            }
          }
          bool wing = false;
          for (int side = 0; side < 2; ++side) {  // right wing, then left
            const M3& Mw = side == 0 ? m.Mwr : m.Mwl;
            const double* pv = side == 0 ? piv_r : piv_l;
            const double* rot = side == 0 ? m.rr : m.rl;
            const double d[3] = {xb[0] - pv[0], xb[1] - pv[1], xb[2] - pv[2]};
            double xw[3];
            mv(Mw, d, xw);
            if (!(xw[1] >= -safety && xw[1] <= L_span + safety)) continue;
            if (!(std::fabs(xw[2]) <= 0.5 * thick + safety)) continue;
            const double ah = 0.5 * L_span, q = 1.0 - ((xw[1] - ah) * (xw[1] - ah)) / (ah * ah);
            const double x_top = q >= 0.0 ? std::sqrt((b_top * b_top) * q) : 0.0;
            const double x_bot = q >= 0.0 ? -std::sqrt((b_bot * b_bot) * q) : 0.0;
            if (!(xw[0] > x_bot - safety && xw[0] < x_top + safety)) continue;
            const double yt = xw[1] < 0.0 ? steps(-xw[1], 0.0, smooth) : steps(xw[1], L_span, smooth);
            const double zt = steps(std::fabs(xw[2]), 0.5 * thick, smooth);
            const double xt = xw[0] < 0.0 ? steps(-xw[0], -x_bot, smooth) : steps(xw[0], x_top, smooth);
            const double mt = zt * yt * xt;
            if (mask[p] < mt && mt > 0.0) {
              mask[p] = mt;
              color[p] = side == 0 ? 3 : 2;
              const double v[3] = {rot[1] * xw[2] - rot[2] * xw[1], rot[2] * xw[0] - rot[0] * xw[2], rot[0] * xw[1] - rot[1] * xw[0]};
              double o[3];
              mv(tr(Mw), v, o);
              us[p] = o[0]; us[npts + p] = o[1]; us[2 * npts + p] = o[2];
              wing = true;
            }
          }
          // stage II: wing velocities to the global frame (wings_geometry.f90:71-108)
          if (wing || ((color[p] == 2 || color[p] == 3) && mask[p] > 0.0)) {
            const double v[3] = {m.vb[0] + m.rb[1] * xb[2] - m.rb[2] * xb[1], m.vb[1] + m.rb[2] * xb[0] - m.rb[0] * xb[2],
                                 m.vb[2] + m.rb[0] * xb[1] - m.rb[1] * xb[0]};
            const double s[3] = {us[p] + v[0], us[npts + p] + v[1], us[2 * npts + p] + v[2]};
            double o[3];
            mv(Mbi, s, o);
            us[p] = o[0]; us[npts + p] = o[1]; us[2 * npts + p] = o[2];
          }
        }
  }

  // ---- rhs.f90:2-115 cal_nlk, :147-330 cal_nlk_fsi, :441-468 explicit
  // diffusion, :575-615 add_grad_pressure; basic_operators.f90:185-215 curl
  void cal_nlk(double t, const std::vector<cd>& v, std::vector<cd>& out) {
    create_mask(t);
    for (int c = 0; c < 3; ++c) F.inverse(v.data() + c * nspec, u.data() + c * npts);
#pragma omp parallel for schedule(static) collapse(2)
    for (int i = 0; i < nxh; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
          const size_t s = ((size_t)i * ny + j) * nz + k;
          const cd a = v[s], b = v[nspec + s], c = v[2 * nspec + s];
          auto ik = [](cd z) { return cd{-z.im, z.re}; };  // i z
          wk[s] = ik(ky[j] * c - kz[k] * b);
          wk[nspec + s] = ik(kz[k] * a - kx[i] * c);
          wk[2 * nspec + s] = ik(kx[i] * b - ky[j] * a);
        }
    for (int c = 0; c < 3; ++c) F.inverse(wk.data() + c * nspec, w.data() + c * npts);
    const double eps_inv = 1.0 / eps;
#pragma omp parallel for schedule(static)
    for (long p = 0; p < (long)npts; ++p) {
      const double chi = mask[p] * eps_inv;
      const double u0_ = u[p], u1 = u[npts + p], u2 = u[2 * npts + p];
      const double w0 = w[p], w1 = w[npts + p], w2 = w[2 * npts + p];
      const double p0 = chi * (u0_ - us[p]), p1 = chi * (u1 - us[npts + p]), p2 = chi * (u2 - us[2 * npts + p]);
      w[p] = u1 * w2 - u2 * w1 - p0;
      w[npts + p] = u2 * w0 - u0_ * w2 - p1;
      w[2 * npts + p] = u0_ * w1 - u1 * w0 - p2;
    }
    for (int c = 0; c < 3; ++c) F.forward(w.data() + c * npts, out.data() + c * nspec);
#pragma omp parallel for schedule(static) collapse(2)
    for (int i = 0; i < nxh; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
          const size_t s = ((size_t)i * ny + j) * nz + k;
          const double k2 = (kx[i] * kx[i] + ky[j] * ky[j]) + kz[k] * kz[k];
          cd n[3] = {out[s], out[nspec + s], out[2 * nspec + s]};
          for (int c = 0; c < 3; ++c) n[c] = n[c] - (nu * k2) * v[c * nspec + s];
          if (k2 != 0.0) {
            const cd dq = (kx[i] * n[0] + ky[j] * n[1]) + kz[k] * n[2];
            const cd q = {dq.re / k2, dq.im / k2};
            n[0] = n[0] - kx[i] * q;
            n[1] = n[1] - ky[j] * q;
            n[2] = n[2] - kz[k] * q;
          }
          for (int c = 0; c < 3; ++c) out[c * nspec + s] = n[c];
        }
  }
  // basic_operators.f90:1032-1066 dealias3
  void dealias(std::vector<cd>& f) {
    if (!dealias_on) return;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < (long)nspec; ++s)
      if (alias[s])
        for (int c = 0; c < 3; ++c) f[c * nspec + s] = {0.0, 0.0};
  }
  // fluid_time_step.f90:602-671 RK4 with explicit diffusion
  void rk4(double t) {
    auto axpy = [&](std::vector<cd>& dst, const std::vector<cd>& a, double s, const std::vector<cd>& b) {
#pragma omp parallel for schedule(static)
      for (long i = 0; i < (long)dst.size(); ++i) dst[i] = a[i] + s * b[i];
    };
    u0 = uk;
    cal_nlk(t, uk, nlk);
    dealias(nlk);
    acc = nlk;
    axpy(uk, u0, 0.5 * dt, nlk);
    cal_nlk(t + 0.5 * dt, uk, nlk);
    dealias(nlk);
    axpy(acc, acc, 2.0, nlk);
    axpy(uk, u0, 0.5 * dt, nlk);
    cal_nlk(t + 0.5 * dt, uk, nlk);
    dealias(nlk);
    axpy(acc, acc, 2.0, nlk);
    axpy(uk, u0, dt, nlk);
    cal_nlk(t + dt, uk, nlk);
    dealias(nlk);
    axpy(acc, acc, 1.0, nlk);
    axpy(uk, u0, dt / 6.0, acc);
    dealias(uk);
  }
};

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s <ini> <nsteps> [threads] [uk_in|-] [uk_out|-]\n", argv[0]);
    return 2;
  }
  try {
    const int nsteps = std::atoi(argv[2]);
    const int threads = argc > 3 ? std::atoi(argv[3]) : omp_get_max_threads();
    omp_set_num_threads(threads);
    const auto t0 = std::chrono::steady_clock::now();
    Sim s(argv[1]);
    double time = 0.0;
    // init_fields_fsi.f90:716-737 quiescent start, mask drawn at t = 0
    if (argc > 4 && std::string(argv[4]) != "-") {
      std::ifstream f(argv[4], std::ios::binary);
      f.read(reinterpret_cast<char*>(s.uk.data()), sizeof(cd) * s.uk.size());
      if (!f) throw std::runtime_error("cannot read the initial state");
    }
    s.create_mask(time);
    const auto t1 = std::chrono::steady_clock::now();
    for (int n = 0; n < nsteps; ++n) {
      s.rk4(time);
      time += s.dt;
    }
    const auto t2 = std::chrono::steady_clock::now();
    if (argc > 5 && std::string(argv[5]) != "-") {
      std::ofstream f(argv[5], std::ios::binary);
      f.write(reinterpret_cast<const char*>(s.uk.data()), sizeof(cd) * s.uk.size());
    }
    const double setup = std::chrono::duration<double>(t1 - t0).count();
    const double run = std::chrono::duration<double>(t2 - t1).count();
    std::printf("{\"grid\": [%d, %d, %d], \"steps\": %d, \"threads\": %d, \"setup_s\": %.6f, \"run_s\": %.6f, "
                "\"s_per_step\": %.6f}\n",
                s.nx, s.ny, s.nz, nsteps, threads, setup, run, nsteps ? run / nsteps : 0.0);
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "flusi_cpu: %s\n", e.what());
    return 1;
  }
}
