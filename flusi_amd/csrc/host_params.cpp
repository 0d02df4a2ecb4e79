// .ini reader, parameter defaults and tethered-insect kinematics.
#include "host_params.h"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace flusi {

// ------------------------------------------------------------------ .ini
IniFile::IniFile(const std::string& p) : path(p) {
  std::ifstream f(p);
  if (!f) throw FlusiError(1, "ERROR! file: " + p + " not found");
  std::string line;
  while (std::getline(f, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    size_t s = line.find_first_not_of(" \t");
    line = (s == std::string::npos) ? std::string() : line.substr(s);
    size_t sc = line.find(';');
    if (sc != std::string::npos) line = line.substr(0, sc + 1);
    if (!line.empty() && (line[0] == ';' || line[0] == '!' || line[0] == '#' || line[0] == '%')) line.clear();
    lines_.push_back(line);
  }
}

std::string IniFile::get(const std::string& section, const std::string& key) const {
  bool found = false;
  const std::string head = "[" + section + "]";
  const std::string kk = key + "=";
  for (const auto& l : lines_) {
    if (l.compare(0, head.size(), head) == 0) found = true;
    else if (!l.empty() && l[0] == '[') found = false;
    if (found && l.compare(0, kk.size(), kk) == 0) {
      const size_t i1 = l.find('=') + 1;
      const size_t sc = l.find(';');
      std::string v;
      if (sc != std::string::npos) v = l.substr(i1, sc - i1);
      else {
        v = l.substr(i1);
        size_t sp = v.find(' ');
        if (sp != std::string::npos) v = v.substr(0, sp);
      }
      size_t a = v.find_first_not_of(" \t");
      if (a == std::string::npos) return std::string();
      v = v.substr(a);
      size_t b = v.find_last_not_of(" \t");
      return v.substr(0, b + 1);
    }
  }
  return std::string();
}

static std::vector<std::string> tokens(const std::string& v) {
  std::vector<std::string> out;
  std::string cur;
  for (char c : v) {
    if (c == ' ' || c == '\t' || c == ',') { if (!cur.empty()) out.push_back(cur); cur.clear(); }
    else cur.push_back(c);
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

static double fortran_real(const std::string& tok) {
  std::string t = tok;
  for (auto& c : t) if (c == 'd' || c == 'D') c = 'e';
  char* end = nullptr;
  const double v = std::strtod(t.c_str(), &end);
  if (end == t.c_str()) throw FlusiError(2, "cannot read a real from '" + tok + "'");
  return v;
}

double IniFile::dbl(const std::string& s, const std::string& k, double def) const {
  const std::string v = get(s, k);
  if (v.empty()) return def;
  const size_t p = v.find("*dx");
  if (p == std::string::npos) {
    auto t = tokens(v);
    return fortran_real(t.empty() ? v : t[0]);
  }
  if (dx <= 0.0) throw FlusiError(3, "relative value read before dx is set");
  auto t = tokens(v.substr(0, p));
  return fortran_real(t.empty() ? v.substr(0, p) : t[0]) * dx;
}

int IniFile::integer(const std::string& s, const std::string& k, int def) const {
  const std::string v = get(s, k);
  if (v.empty()) return def;
  auto t = tokens(v);
  return std::atoi((t.empty() ? v : t[0]).c_str());
}

std::string IniFile::str(const std::string& s, const std::string& k, const std::string& def) const {
  const std::string v = get(s, k);
  return v.empty() ? def : v;
}

std::vector<double> IniFile::vct(const std::string& s, const std::string& k, const std::vector<double>& def) const {
  const std::string v = get(s, k);
  if (v.empty()) return def;
  auto t = tokens(v);
  std::vector<double> out;
  for (size_t i = 0; i < t.size() && out.size() < def.size(); ++i) out.push_back(fortran_real(t[i]));
  while (out.size() < def.size()) out.push_back(def[out.size()]);
  return out;
}

// params.f90:51-372 (keys and defaults of the fluid path)
Params load_params(const IniFile& ini_c) {
  IniFile& ini = const_cast<IniFile&>(ini_c);
  Params p;
  p.nx = ini.integer("Resolution", "nx", 4);
  p.ny = ini.integer("Resolution", "ny", 4);
  p.nz = ini.integer("Resolution", "nz", 4);
  p.xl = ini.dbl("Geometry", "xl", 6.283185307179586);
  p.yl = ini.dbl("Geometry", "yl", 6.283185307179586);
  p.zl = ini.dbl("Geometry", "zl", 6.283185307179586);
  ini.dx = p.xl / p.nx;
  p.length = ini.dbl("Geometry", "Size", 0.0);
  p.nt = ini.integer("Time", "nt", 9999999);
  p.method = ini.str("Time", "iTimeMethodFluid", "AB2");
  p.intelligent_dt = ini.str("Time", "intelligent_dt", "no");
  p.tmax = ini.dbl("Time", "Tmax", 1e9);
  p.cfl = ini.dbl("Time", "CFL", 0.1);
  p.cfl_eta = ini.dbl("Time", "CFL_eta", 0.99);
  p.dt_max = ini.dbl("Time", "dt_max", 0.0);
  p.dt_fixed = ini.dbl("Time", "dt_fixed", 0.0);
  p.checknan_every = ini.integer("Time", "checknan_every", 0);
  p.nu = ini.dbl("ReynoldsNumber", "nu", 1e-2);
  p.inicond = ini.str("InitialCondition", "inicond", "none");
  p.omega1 = ini.dbl("InitialCondition", "omega1", 0.0);
  p.core_a = ini.dbl("InitialCondition", "core_a", 0.0);
  p.iDealias = ini.integer("Dealiasing", "iDealias", 1);
  p.iPenalization = ini.integer("Penalization", "iPenalization", 0);
  p.iMoving = ini.integer("Penalization", "iMoving", 0);
  p.iMask = ini.str("Penalization", "iMask", "none");
  p.eps = ini.dbl("Penalization", "eps", 1e-2);
  p.periodic = ini.str("Penalization", "periodic", "no") == "yes";
  p.tsave = ini.dbl("Saving", "tsave", 9e9);
  p.tintegral = ini.dbl("Saving", "tintegral", 0.01);
  p.tsave_first = ini.dbl("Saving", "tsave_first", 0.0);
  p.itdrag = ini.integer("Saving", "itdrag", 99999);
  p.iDoBackup = ini.integer("Saving", "iDoBackup", 1);
  p.backup_type = ini.str("Saving", "backup_type", "one-file-backup");
  p.iSaveVelocity = ini.integer("Saving", "iSaveVelocity", 0);
  p.iSavePress = ini.integer("Saving", "iSavePress", 0);
  p.iSaveVorticity = ini.integer("Saving", "iSaveVorticity", 0);
  p.iSaveMagVorticity = ini.integer("Saving", "iSaveMagVorticity", 0);
  p.iSaveMask = ini.integer("Saving", "iSaveMask", 0);
  p.iSaveSolidVelocity = ini.integer("Saving", "iSaveSolidVelocity", 0);
  p.truntime = ini.dbl("Saving", "truntime", 1.0);
  p.wtimemax = ini.dbl("Saving", "wtimemax", 8760.0);
  p.tsave_period = ini.dbl("Saving", "tsave_period", 1.0);
  p.save_only_one_period = ini.str("Saving", "save_only_one_period", "no");
  p.naming = ini.str("Saving", "naming", "time");
  p.field_precision = ini.str("Saving", "field_precision", "single");
  p.striding = ini.integer("Saving", "striding", 1);
  if (p.striding < 1) throw FlusiError(8881, "Striding value is bad, exit!");
  p.iCavity = ini.str("Penalization", "iCavity", "no");
  p.cavity_size = ini.integer("Penalization", "cavity_size", 0);
  p.compute_forces = ini.integer("Penalization", "compute_forces", 1);
  p.iVorticitySponge = ini.str("Sponge", "iVorticitySponge", "no");
  p.iSpongeType = ini.str("Sponge", "iSpongeType", "top_cover");
  p.eps_sponge = ini.dbl("Sponge", "eps_sponge", 0.0);
  p.sponge_thickness = ini.integer("Sponge", "sponge_thickness", 0);
  p.x0 = ini.dbl("Geometry", "x0", p.xl / 2.0);
  p.y0 = ini.dbl("Geometry", "y0", p.yl / 2.0);
  p.z0 = ini.dbl("Geometry", "z0", p.zl / 2.0);
  const char* ax[3] = {"iMeanFlow_x", "iMeanFlow_y", "iMeanFlow_z"};
  for (int i = 0; i < 3; ++i) p.iMeanFlow[i] = ini.str("MeanFlow", ax[i], "free");
  if (ini.str("MeanFlow", "iMeanFlow", "unused") == "1") p.iMeanFlow = {{"fixed", "fixed", "fixed"}};
  p.m_fluid = ini.dbl("MeanFlow", "m_fluid", 999e9);
  p.umean = {ini.dbl("MeanFlow", "ux", 1.0), ini.dbl("MeanFlow", "uy", 1.0), ini.dbl("MeanFlow", "uz", 1.0)};
  p.umean_freq = ini.dbl("MeanFlow", "umean_freq", 0.0);
  p.startup_conditioner = ini.str("MeanFlow", "iMeanFlowStartupConditioner", "no");
  p.tau_meanflow = ini.dbl("MeanFlow", "tau_meanflow", 0.0);
  p.T_release_meanflow = ini.dbl("MeanFlow", "T_release_meanflow", 0.0);
  return p;
}

// ------------------------------------------------------------ 3x3 algebra
Mat3 mat_mul(const Mat3& a, const Mat3& b) {
  Mat3 c{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) c[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
  return c;
}
Mat3 transpose(const Mat3& a) { return {a[0], a[3], a[6], a[1], a[4], a[7], a[2], a[5], a[8]}; }
Vec3 mat_vec(const Mat3& a, const Vec3& v) {
  return {a[0] * v[0] + a[1] * v[1] + a[2] * v[2], a[3] * v[0] + a[4] * v[1] + a[5] * v[2],
          a[6] * v[0] + a[7] * v[1] + a[8] * v[2]};
}
static Vec3 vadd(const Vec3& a, const Vec3& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }

// rotation_matrices.f90:4-37 (passive rotations)
static Mat3 Rx(double a) { double c = std::cos(a), s = std::sin(a); return {1, 0, 0, 0, c, s, 0, -s, c}; }
static Mat3 Ry(double a) { double c = std::cos(a), s = std::sin(a); return {c, 0, -s, 0, 1, 0, s, 0, c}; }
static Mat3 Rz(double a) { double c = std::cos(a), s = std::sin(a); return {c, s, 0, -s, c, 0, 0, 0, 1}; }
static double deg2rad(double d) { return d * PI / 180.0; }  // vars.f90:255-260

// module_helpers.f90:30-55
static void fseries(double t, double a0, const std::vector<double>& ai, const std::vector<double>& bi,
                    double& u, double& udt) {
  const double f = 2.0 * PI;
  u = 0.5 * a0;
  udt = 0.0;
  for (size_t i = 1; i <= ai.size(); ++i) {
    const double s = std::sin(f * double(i) * t), c = std::cos(f * double(i) * t);
    u = u + ai[i - 1] * c + bi[i - 1] * s;
    udt = udt + f * double(i) * (-ai[i - 1] * s + bi[i - 1] * c);
  }
}

// Fortran default-real literals are single precision
static std::vector<double> f32(std::initializer_list<double> v) {
  std::vector<double> o;
  for (double x : v) o.push_back((double)(float)x);
  return o;
}

std::array<double, 6> WingKinematics::eval(double t, const std::string& base_dir) {
  std::array<double, 6> r{};
  if (protocol == "from_file") {
    if (!initialized) {
      if (infile.find(".ini") == std::string::npos) throw FlusiError(2030, "old kinematics file format: " + infile);
      std::string path = infile;
      if (!path.empty() && path[0] != '/') path = base_dir + "/" + path;
      IniFile kf(path);
      const std::string type = kf.str("kinematics", "type", "none");
      if (type != "Fourier" && type != "fourier" && type != "FOURIER")
        throw FlusiError(77771, "kinematics file: only type=fourier is supported");
      units = kf.str("kinematics", "units", "degree");
      convention = kf.str("kinematics", "convention", "flusi");
      const char* nm[3] = {"phi", "alpha", "theta"};
      for (int a = 0; a < 3; ++a) {
        const int n = kf.integer("kinematics", std::string("nfft_") + nm[a], 0);
        a0[a] = kf.dbl("kinematics", std::string("a0_") + nm[a], 0.0);
        ai[a] = kf.vct("kinematics", std::string("ai_") + nm[a], std::vector<double>(n, 0.0));
        bi[a] = kf.vct("kinematics", std::string("bi_") + nm[a], std::vector<double>(n, 0.0));
      }
      initialized = true;
    }
    for (int a = 0; a < 3; ++a) fseries(t, a0[a], ai[a], bi[a], r[a], r[3 + a]);
    // r = phi, alpha, theta, phi_dt, alpha_dt, theta_dt
    if (units == "degree" || units == "DEGREE" || units == "Degree") {
      for (auto& v : r) v = deg2rad(v);
    } else if (!(units == "radiant" || units == "RADIANT" || units == "Radiant")) {
      throw FlusiError(1718, "kinematics: units must be degree or radiant");
    }
    if (convention == "SISC" || convention == "sisc" || convention == "siam-sisc") { r[2] = -r[2]; r[5] = -r[5]; }
    else if (!(convention == "FLUSI" || convention == "flusi" || convention == "Flusi" || convention == "FluSI"))
      throw FlusiError(1719, "kinematics: convention must be flusi or sisc");
    return r;
  }
  if (protocol == "Drosophila_hovering_fry") {
    static const std::vector<double> aip = f32({71.1061858, 2.1685448, -0.1986978, 0.6095268, -0.0311298,
                                                -0.1255648, -0.0867778, 0.0543518, 0.0, 0.0});
    static const std::vector<double> bip = f32({5.4547058, -3.5461688, 0.6260698, 0.1573728, -0.0360498,
                                                -0.0205348, -0.0083818, -0.0076848, 0.0, 0.0});
    static const std::vector<double> aia = f32({3.3288788, 0.6303878, -10.9780518, 2.1123398, -3.2301198,
                                                -1.4473158, 0.6141758, -0.3071608, 0.1458498, 0.0848308});
    static const std::vector<double> bia = f32({67.5430838, 0.6566888, 9.9226018, 3.9183988, -2.6882828,
                                                0.6433518, -0.8792398, -0.4817838, 0.0300078, -0.1015118});
    static const std::vector<double> ait = f32({-3.9750378, -8.2808998, 0.0611208, 0.3906598, -0.4488778,
                                                0.120087, 0.0717048, -0.0699578, 0.0, 0.0});
    static const std::vector<double> bit = f32({-2.2839398, -3.5213068, 1.9296668, -1.0832488, -0.3011748,
                                                0.1786648, -0.1228608, 0.0004808, 0.0, 0.0});
    fseries(t, (double)(float)25.4649398, aip, bip, r[0], r[3]);
    fseries(t, (double)(float)-0.3056968, aia, bia, r[1], r[4]);
    fseries(t, (double)(float)-17.8244658, ait, bit, r[2], r[5]);
    for (auto& v : r) v = deg2rad(v);
    return r;
  }
  if (protocol == "flapper_dickinson") {
    static const double bi_alpha[29] = {48.23094285611071, 0.0, 10.224154661301371, 0.0, 2.1623763046726396,
        0.0, 0.05049394424178093, 0.0, -0.17550942623071494, 0.0, -0.06634193748204852, 0.0,
        -0.008925020495896451, 0.0, 0.0011292567942149407, 0.0, 6.471071566666472e-4, 0.0,
        1.0018757795834964e-4, 0.0, 3.0105550216312524e-6, 0.0, -1.237567150768195e-6, 0.0,
        -1.988004402010933e-7, 0.0, -1.10165545174181e-8, 0.0, 2.4135650975460306e-10};
    static const double ai_phi[31] = {63.24528806534019, 0.0, 5.753991800610726, 0.0, 1.3887974015525626,
        0.0, 0.3889856512386744, 0.0, 0.10577402496901325, 0.0, 0.026061339604144987, 0.0,
        0.005623376646981709, 0.0, 0.001042285996467963, 0.0, 1.639611509380189e-4, 0.0,
        2.1716252827442023e-5, 0.0, 2.408190194815521e-6, 0.0, 2.2268710288534648e-7, 0.0,
        1.7118916093759426e-8, 0.0, 1.0914870312823793e-9, 0.0, 5.76135101855556e-11, 0.0,
        2.513944479978149e-12};
    const double tadv = (double)0.08f;
    const double f = 2.0 * PI;
    double al = 0.0, al_dt = 0.0;
    for (int i = 1; i <= 29; ++i) {
      const double s = std::sin(f * double(i) * (t + tadv)), c = std::cos(f * double(i) * (t + tadv));
      al = al + bi_alpha[i - 1] * s;
      al_dt = al_dt + f * double(i) * bi_alpha[i - 1] * c;
    }
    const double alphacdeg = 90.0 - 40.0;
    al = alphacdeg / 40.0 * al;
    al_dt = alphacdeg / 40.0 * al_dt;
    double ph = 0.0, ph_dt = 0.0;
    for (int i = 1; i <= 31; ++i) {
      const double s = std::sin(f * double(i) * t), c = std::cos(f * double(i) * t);
      ph = ph + ai_phi[i - 1] * c;
      ph_dt = ph_dt + f * double(i) * (-ai_phi[i - 1] * s);
    }
    const double phicdeg = 180.0;
    ph = phicdeg / 180.0 * ph;
    ph_dt = phicdeg / 180.0 * ph_dt;
    return {deg2rad(ph), deg2rad(al), 0.0, deg2rad(ph_dt), deg2rad(al_dt), 0.0};
  }
  if (protocol == "none") return r;
  throw FlusiError(10624, "flapping motion protocol not supported: " + protocol);
}

int WingKinematics::ncontrols(const std::string& base_dir) {
  if (protocol != "from_file") return 0;
  if (!initialized) (void)eval(0.0, base_dir);
  const size_t n = ai[0].size();
  for (int a = 0; a < 3; ++a)
    if (ai[a].size() != n || bi[a].size() != n)
      throw FlusiError(900020, "kinematics controls need one nfft for phi, alpha and theta");
  return (int)(6 * n);
}
double& WingKinematics::control(int p, const std::string& base_dir) {
  const int nc = ncontrols(base_dir);
  if (p < 0 || p >= nc) throw FlusiError(900021, "kinematics control index out of range");
  const int n = nc / 6, a = p / (2 * n), r = p % (2 * n);
  return r < n ? ai[a][r] : bi[a][r - n];
}
int Insect::ncontrols() {
  const int r = kine_r.ncontrols(base_dir), l = kine_l.ncontrols(base_dir);
  if (r && l && r != l) throw FlusiError(900020, "kinematics controls: the wings' series differ in length");
  return std::max(r, l);
}
double Insect::get_control(int p) {
  if (kine_r.ncontrols(base_dir)) return kine_r.control(p, base_dir);
  return kine_l.control(p, base_dir);
}
void Insect::set_control(int p, double v) {
  if (kine_r.ncontrols(base_dir)) kine_r.control(p, base_dir) = v;
  if (kine_l.ncontrols(base_dir)) kine_l.control(p, base_dir) = v;
}

// insect_init_clean.f90:1-250 (tethered insect)
void Insect::init(const IniFile& ini_c, const Params& p, const std::string& bdir) {
  IniFile& ini = const_cast<IniFile&>(ini_c);
  base_dir = bdir;
  const double dxr = p.dx();
  WingShape = ini.str("Insects", "WingShape", "none");
  b_top = ini.dbl("Insects", "b_top", 0.0);
  b_bot = ini.dbl("Insects", "b_bot", 0.0);
  L_span = ini.dbl("Insects", "L_span", 0.0);
  auto mk_kine = [&](const std::string& prot, const std::string& infile_old) {
    WingKinematics k;
    const size_t pos = prot.find("from_file::");
    if (pos != std::string::npos) { k.protocol = "from_file"; k.infile = prot.substr(pos + 11); }
    else if (prot == "from_file") { k.protocol = "from_file"; k.infile = infile_old; }
    else k.protocol = prot;
    return k;
  };
  const std::string infile = ini.str("Insects", "infile", "none.in");
  kine_r = mk_kine(ini.str("Insects", "FlappingMotion_right", "none"), infile);
  kine_l = mk_kine(ini.str("Insects", "FlappingMotion_left", "none"), infile);
  BodyType = ini.str("Insects", "BodyType", "ellipsoid");
  HasDetails = ini.str("Insects", "HasDetails", "all");
  BodyMotion = ini.str("Insects", "BodyMotion", "yes");
  LeftWing = ini.str("Insects", "LeftWing", "yes");
  RightWing = ini.str("Insects", "RightWing", "yes");
  b_body = ini.dbl("Insects", "b_body", 0.1);
  L_body = ini.dbl("Insects", "L_body", 1.0);
  R_head = ini.dbl("Insects", "R_head", 0.1);
  R_eye = ini.dbl("Insects", "R_eye", 0.0);  // insect_init_clean.f90:142 default 0.d1 = 0
  WingThickness = ini.dbl("Insects", "WingThickness", 4.0 * dxr);
  auto v3 = [](const std::vector<double>& v) { return Vec3{v[0], v[1], v[2]}; };
  x0 = v3(ini.vct("Insects", "x0", {0.5 * p.xl, 0.5 * p.yl, 0.5 * p.zl}));
  yawpitchroll_0 = v3(ini.vct("Insects", "yawpitchroll_0", {0.0, 0.0, 0.0}));
  for (auto& v : yawpitchroll_0) v *= PI / 180.0;
  eta0 = ini.dbl("Insects", "eta0", 0.0) * (PI / 180.0);
  Jxx = ini.dbl("Insects", "Jxx", 0.0);
  Jyy = ini.dbl("Insects", "Jyy", 0.0);
  Jzz = ini.dbl("Insects", "Jzz", 0.0);
  Jxy = ini.dbl("Insects", "Jxy", 0.0);
  smooth = 1.0 * dxr;
  safety = 3.5 * smooth;
  x_head = v3(ini.vct("Insects", "x_head", {0.5 * L_body, 0.0, 0.0}));
  const double se = std::sin(45.0 * PI / 180.0) * R_head * 0.8;
  x_eye_r = v3(ini.vct("Insects", "x_eye_r", {x_head[0] + se, x_head[1] + se, x_head[2] + se}));
  x_eye_l = v3(ini.vct("Insects", "x_eye_l", {x_head[0] + se, x_head[1] - se, x_head[2] + se}));
  x_pivot_l_b = v3(ini.vct("Insects", "x_pivot_l", {0.0, +b_body, 0.0}));
  x_pivot_r_b = v3(ini.vct("Insects", "x_pivot_r", {0.0, -b_body, 0.0}));
  if (BodyMotion != "tethered") throw FlusiError(10625, "BodyMotion not supported: " + BodyMotion);
  if (BodyType != "jerry" && BodyType != "Jerry" && BodyType != "ellipsoid_plain" && BodyType != "nobody" &&
      BodyType != "bumblebee")
    throw FlusiError(10623, "Insect%BodyType unknown..." + BodyType);
  if (WingShape == "bumblebee") fwing.setup(WingShape, WingThickness);
  else if (WingShape != "TwoEllipses" && WingShape != "rectangular")
    throw FlusiError(10626, "WingShape not supported: " + WingShape);
  update(0.0);
}

InsectMotion Insect::motion(double t) {
  InsectMotion s{};
  // body_motion.f90:147-160 tethered
  s.psi = yawpitchroll_0[2]; s.beta = yawpitchroll_0[1]; s.gamma = yawpitchroll_0[0];
  s.psi_dt = s.beta_dt = s.gamma_dt = 0.0;
  s.xc = x0; s.vc = {0, 0, 0};
  auto r = kine_r.eval(t, base_dir);
  s.phi_r = r[0]; s.alpha_r = r[1]; s.theta_r = r[2]; s.phi_dt_r = r[3]; s.alpha_dt_r = r[4]; s.theta_dt_r = r[5];
  auto l = kine_l.eval(t, base_dir);
  s.phi_l = l[0]; s.alpha_l = l[1]; s.theta_l = l[2]; s.phi_dt_l = l[3]; s.alpha_dt_l = l[4]; s.theta_dt_l = l[5];
  s.eta = eta0;  // stroke_plane.f90:7-37
  // module_insects.f90:889-966 rotation matrices
  s.Mb = mat_mul(Rx(s.psi), mat_mul(Ry(s.beta), Rz(s.gamma)));
  const Mat3 Ms_l = Ry(s.eta), Ms_r = mat_mul(Rx(PI), Ry(s.eta));
  const Mat3 M1l = Ry(s.alpha_l), M2l = Rz(s.theta_l), M3l = Rx(s.phi_l);
  const Mat3 M1r = Ry(-s.alpha_r), M2r = Rz(s.theta_r), M3r = Rx(-s.phi_r);
  s.Mwl = mat_mul(M1l, mat_mul(M2l, mat_mul(M3l, Ms_l)));
  s.Mwr = mat_mul(M1r, mat_mul(M2r, mat_mul(M3r, Ms_r)));
  // module_insects.f90:604-636 body_angular_velocity
  const double cb = std::cos(s.beta), sb = std::sin(s.beta), cg = std::cos(s.gamma), sg = std::sin(s.gamma);
  const double cp = std::cos(s.psi), sp = std::sin(s.psi);
  s.rot_body_g = {s.psi_dt * cb * cg - s.beta_dt * sg, s.beta_dt * cg + s.psi_dt * cb * sg, s.gamma_dt - s.psi_dt * sb};
  s.rot_body_b = {s.psi_dt - s.gamma_dt * sb, s.beta_dt * cp + s.gamma_dt * cb * sp, s.gamma_dt * cb * cp - s.beta_dt * sp};
  // module_insects.f90:650-750 wing_angular_velocities
  const Vec3 ra_l{0, s.alpha_dt_l, 0}, rt_l{0, 0, s.theta_dt_l}, rp_l{s.phi_dt_l, 0, 0};
  const Vec3 ra_r{0, -s.alpha_dt_r, 0}, rt_r{0, 0, s.theta_dt_r}, rp_r{-s.phi_dt_r, 0, 0};
  s.rot_rel_l_w = mat_vec(s.Mwl, mat_vec(transpose(Ms_l), mat_vec(transpose(M3l),
                  vadd(rp_l, mat_vec(transpose(M2l), vadd(rt_l, mat_vec(transpose(M1l), ra_l)))))));
  s.rot_rel_r_w = mat_vec(s.Mwr, mat_vec(transpose(Ms_r), mat_vec(transpose(M3r),
                  vadd(rp_r, mat_vec(transpose(M2r), vadd(rt_r, mat_vec(transpose(M1r), ra_r)))))));
  const Vec3 rl_b = mat_vec(transpose(s.Mwl), s.rot_rel_l_w), rr_b = mat_vec(transpose(s.Mwr), s.rot_rel_r_w);
  s.rot_rel_l_g = mat_vec(transpose(s.Mb), rl_b);
  s.rot_rel_r_g = mat_vec(transpose(s.Mb), rr_b);
  return s;
}

void Insect::update(double t) {
  InsectMotion s = motion(t);
  // module_insects.f90:752-799 wing_angular_accel (forward difference, dt = 1e-8)
  const double dt = 1.0e-8;
  InsectMotion s2 = motion(t + dt);
  Vec3 dl, dr;
  for (int i = 0; i < 3; ++i) {
    dl[i] = (s2.rot_rel_l_g[i] - s.rot_rel_l_g[i]) / dt;
    dr[i] = (s2.rot_rel_r_g[i] - s.rot_rel_r_g[i]) / dt;
  }
  s.rot_dt_l_w = mat_vec(s.Mwl, mat_vec(s.Mb, dl));
  s.rot_dt_r_w = mat_vec(s.Mwr, mat_vec(s.Mb, dr));
  s.x_pivot_l_g = mat_vec(transpose(s.Mb), x_pivot_l_b);
  s.x_pivot_r_g = mat_vec(transpose(s.Mb), x_pivot_r_b);
  st = s;
  time = t;
}

// module_insects.f90:1139-1154 write_kinematics_file
std::vector<double> Insect::kinematics_row(double t) const {
  const InsectMotion& s = st;
  return {t, s.xc[0], s.xc[1], s.xc[2], s.psi, s.beta, s.gamma, s.eta, s.alpha_l, s.phi_l, s.theta_l,
          s.alpha_r, s.phi_r, s.theta_r, s.rot_rel_l_w[0], s.rot_rel_l_w[1], s.rot_rel_l_w[2],
          s.rot_rel_r_w[0], s.rot_rel_r_w[1], s.rot_rel_r_w[2], s.rot_dt_l_w[0], s.rot_dt_l_w[1],
          s.rot_dt_l_w[2], s.rot_dt_r_w[0], s.rot_dt_r_w[1], s.rot_dt_r_w[2]};
}


// body_geometry.f90:277-345.  The coordinates are Fortran default-real
// literals and the radii products of default reals: single precision,
// widened to double.
std::vector<Capsule> bumblebee_capsules() {
  auto f = [](float v) { return (double)v; };
  const float xl1[5] = {-0.74f, -0.63f, -0.4f, -0.1f, 0.1f}, yl1[5] = {0.32f, 0.32f, 0.31f, 0.3f, 0.12f},
              zl1[5] = {-0.35f, -0.37f, -0.2f, -0.1f, -0.16f};
  const float xl2[5] = {-0.24f, -0.15f, 0.02f, 0.17f, 0.19f}, yl2[5] = {0.33f, 0.33f, 0.32f, 0.3f, 0.15f},
              zl2[5] = {-0.29f, -0.28f, -0.2f, -0.15f, -0.19f};
  const float xl3[5] = {0.28f, 0.35f, 0.45f, 0.4f, 0.35f}, yl3[5] = {0.31f, 0.30f, 0.28f, 0.2f, 0.15f},
              zl3[5] = {-0.3f, -0.28f, -0.25f, -0.18f, -0.18f};
  const float r12[4] = {0.015f, 0.03f, 0.04f, 0.03f}, r3[4] = {0.015f, 0.02f, 0.03f, 0.02f};
  volatile float k13 = 1.3f;  // keep the single-precision products
  std::vector<Capsule> out;
  for (int j = 0; j < 4; ++j) {
    const float* xs[3] = {xl1, xl2, xl3};
    const float* ys[3] = {yl1, yl2, yl3};
    const float* zs[3] = {zl1, zl2, zl3};
    const float rr[3] = {r12[j] * k13, r12[j] * k13, r3[j] * k13};
    for (int side = 0; side < 2; ++side) {
      const double sg = side ? -1.0 : 1.0;
      for (int l = 0; l < 3; ++l)
        out.push_back(Capsule{Vec3{f(xs[l][j]), sg * f(ys[l][j]), f(zs[l][j])},
                              Vec3{f(xs[l][j + 1]), sg * f(ys[l][j + 1]), f(zs[l][j + 1])}, (double)rr[l]});
    }
  }
  const float ran = 0.015f * k13, rf = 0.017f * k13;
  out.push_back(Capsule{Vec3{f(0.63f), f(0.05f), f(-0.03f)}, Vec3{f(0.8f), f(0.27f), f(0.1f)}, (double)ran});
  out.push_back(Capsule{Vec3{f(0.63f), -f(0.05f), f(-0.03f)}, Vec3{f(0.8f), -f(0.27f), f(0.1f)}, (double)ran});
  out.push_back(Capsule{Vec3{f(0.43f), f(0.0f), f(-0.28f)}, Vec3{f(0.6f), f(0.0f), f(-0.23f)}, (double)rf});
  return out;
}

void FourierWing::setup(const std::string& shape, double thickness) {
  if (shape != "bumblebee") throw FlusiError(554329, "Fourier wing shape not supported: " + shape);
  // wings_geometry.f90:1115-1144
  static const double ai[25] = {-0.0128037920989526, -0.106777418654552, 0.0380851289321982,
      0.0330548081081197, -0.0178496286355627, -0.00328588543359649, 0.0108246924336137,
      -0.00489302388329943, -0.00708808441172961, 0.00518244772891516, 0.00445979960844562,
      -0.000108072056165527, 0.00204437845603716, 0.00147176382618797, -0.00229559463098105,
      -0.000514633972391526, 0.00134150515430486, -0.000149860228261824, 9.01456938813568e-05,
      0.00150639712261487, 0.000914624010720407, -0.000737650894315551, -0.000843849923321745,
      -0.000354971670482499, -0.000382956472432449};
  static const double bi[25] = {-0.0158061138788171, 0.0308243584184200, -0.00903330410923372,
      -0.0185758334697500, -0.000924452934252486, -0.00242101213359519, -0.00204549530064489,
      0.00291468131401423, -0.000140755032337495, -0.00135036427128534, 0.00141285439042451,
      -0.000334215276598231, -0.00161521722061879, -0.000164055684312904, -0.000256278551727569,
      -0.000740258481681094, 0.000847498161852221, 0.00157442110960973, -0.000559835622451578,
      -0.000617498559228280, 0.00115413452523474, 0.000322564770099778, -0.000917375185844477,
      4.44819399488798e-05, 0.000710028654602170};
  const double a0 = 0.594557593733011;
  xc = -0.1;
  yc = 0.501549263807117;
  dphi = (2.0 * PI) / double(N_RADIUS - 1);
  table.assign(N_RADIUS, 0.0);
  for (int j = 0; j < N_RADIUS; ++j) {
    const double th = double(j) * dphi;
    double R = a0 / 2.0;
    for (int i = 1; i <= 25; ++i)
      R = R + ai[i - 1] * std::cos(2.0 * PI * double(i) * th) + bi[i - 1] * std::sin(2.0 * PI * double(i) * th);
    table[j] = R;
  }
  double xmin = 999.0e9, ymin = 999.0e9, xmax = -999.0e9, ymax = -999.0e9, theta = 0.0;
  while (theta < 1.0) {
    const double r = radius(theta), tp = 2.0 * PI * theta - PI;
    const double x = xc + r * std::cos(tp), y = yc + r * std::sin(tp);
    xmin = std::min(xmin, x); ymin = std::min(ymin, y);
    xmax = std::max(xmax, x); ymax = std::max(ymax, y);
    theta = theta + 1.0e-3;
  }
  bbox[0] = xmin; bbox[1] = xmax; bbox[2] = ymin; bbox[3] = ymax;
  bbox[4] = -thickness / 2.0; bbox[5] = thickness / 2.0;
}

double FourierWing::radius(double theta) const {
  const long j = (long)std::floor(theta / dphi);
  return table[j] + ((theta - double(j) * dphi) / dphi) * (table[j + 1] - table[j]);
}

}  // namespace flusi
