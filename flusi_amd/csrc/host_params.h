// Host-side configuration: FLUSI's .ini format and the parameters of the
// fluid hot path, plus the tethered-insect kinematics evaluated once per mask
// update on the host.
#pragma once
#include <array>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

namespace flusi {

constexpr double PI = 3.14159265358979323846264338327950288;

struct FlusiError : std::runtime_error {
  int code;
  FlusiError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// ini_files_parser.f90: [Section] headers, key=value; with the value ending
// at the first ';', comment lines starting with ; ! # %, N*dx grid units.
class IniFile {
 public:
  explicit IniFile(const std::string& path);
  std::string get(const std::string& section, const std::string& key) const;
  double dbl(const std::string& s, const std::string& k, double def) const;
  int integer(const std::string& s, const std::string& k, int def) const;
  std::string str(const std::string& s, const std::string& k, const std::string& def) const;
  std::vector<double> vct(const std::string& s, const std::string& k, const std::vector<double>& def) const;
  double dx = -1.0;
  std::string path;
 private:
  std::vector<std::string> lines_;
};

using Vec3 = std::array<double, 3>;
using Mat3 = std::array<double, 9>;  // row-major

struct Params {
  int nx = 4, ny = 4, nz = 4;
  double xl = 2 * PI, yl = 2 * PI, zl = 2 * PI;
  double length = 0.0;
  int nt = 9999999;
  std::string method = "AB2";
  std::string intelligent_dt = "no";
  double tmax = 1e9, cfl = 0.1, cfl_eta = 0.99, dt_max = 0.0, dt_fixed = 0.0;
  // build extension: check uk for NaN after every N-th step (0: never), the
  // reference's checknan_cmplx (basic_operators.f90:925-950, abort 202)
  int checknan_every = 0;
  double nu = 1e-2;
  std::string inicond = "none";
  double omega1 = 0.0, core_a = 0.0;
  int iDealias = 1;
  int iPenalization = 0, iMoving = 0;
  std::string iMask = "none";
  double eps = 1e-2;
  bool periodic = false;
  double tsave = 9e9, tintegral = 0.01, tsave_first = 0.0;
  int itdrag = 99999;
  // [Saving] field snapshots and runtime backups (params.f90:166-187, 305)
  int iDoBackup = 1;
  std::string backup_type = "one-file-backup";
  int iSaveVelocity = 0, iSavePress = 0, iSaveVorticity = 0, iSaveMagVorticity = 0, iSaveMask = 0,
      iSaveSolidVelocity = 0;
  double truntime = 1.0, wtimemax = 8760.0, tsave_period = 1.0;
  std::string save_only_one_period = "no", naming = "time", field_precision = "single";
  int striding = 1;
  std::string iCavity = "no";
  int cavity_size = 0;
  int compute_forces = 1;
  std::string iVorticitySponge = "no", iSpongeType = "top_cover";
  double eps_sponge = 0.0;
  int sponge_thickness = 0;
  double x0 = 0, y0 = 0, z0 = 0;
  std::array<std::string, 3> iMeanFlow{{"free", "free", "free"}};
  double m_fluid = 999e9;
  Vec3 umean{{1.0, 1.0, 1.0}};
  double umean_freq = 0.0;
  std::string startup_conditioner = "no";
  double tau_meanflow = 0.0, T_release_meanflow = 0.0;
  double dx() const { return xl / nx; }
  double dy() const { return yl / ny; }
  double dz() const { return zl / nz; }
};

Params load_params(const IniFile& ini);

// wings_motion.f90:69-305 FlappingMotion (from_file Fourier, and the
// Drosophila_hovering_fry / flapper_dickinson protocols)
struct WingKinematics {
  std::string protocol, infile;
  bool initialized = false;
  std::string units = "degree", convention = "flusi";
  double a0[3] = {0, 0, 0};             // phi, alpha, theta
  std::vector<double> ai[3], bi[3];
  // returns phi, alpha, theta, phi_dt, alpha_dt, theta_dt (radians)
  std::array<double, 6> eval(double t, const std::string& base_dir);
  // Fourier coefficients of a from_file kinematics as controls: for angle a
  // (phi, alpha, theta) and n = nfft, control a*2n + i is ai_a(i) and
  // a*2n + n + i is bi_a(i), in the file's units; all angles need one nfft
  int ncontrols(const std::string& base_dir);
  double& control(int p, const std::string& base_dir);
};

struct InsectMotion {
  double psi, beta, gamma, psi_dt, beta_dt, gamma_dt, eta;
  Vec3 xc, vc;
  double phi_r, alpha_r, theta_r, phi_dt_r, alpha_dt_r, theta_dt_r;
  double phi_l, alpha_l, theta_l, phi_dt_l, alpha_dt_l, theta_dt_l;
  Mat3 Mb, Mwl, Mwr;
  Vec3 rot_body_b, rot_body_g, rot_rel_l_w, rot_rel_r_w, rot_rel_l_g, rot_rel_r_g;
  Vec3 rot_dt_l_w, rot_dt_r_w, x_pivot_l_g, x_pivot_r_g;
};

// module_insects.f90 type diptera, restricted to a tethered insect
// Capsule of the bumblebee's legs, antennae and proboscis: segment a-b in
// the body frame and radius (body_geometry.f90:277-345)
struct Capsule { Vec3 a, b; double r; };
std::vector<Capsule> bumblebee_capsules();

// Wing outline R(theta) as a Fourier series in the normalised angle, sampled
// into a table and interpolated linearly (wings_geometry.f90:819-858,
// :1115-1144, :1438-1511)
struct FourierWing {
  static constexpr int N_RADIUS = 25000;
  double xc = 0, yc = 0, dphi = 0;
  double bbox[6] = {0, 0, 0, 0, 0, 0};
  std::vector<double> table;
  void setup(const std::string& shape, double thickness);
  double radius(double theta) const;
};

struct Insect {
  std::string WingShape, BodyType, BodyMotion, LeftWing, RightWing, HasDetails;
  FourierWing fwing;
  double b_top = 0, b_bot = 0, L_span = 0, b_body = 0.1, L_body = 1.0, R_head = 0.1, R_eye = 0.0;
  double WingThickness = 0, eta0 = 0, Jxx = 0, Jyy = 0, Jzz = 0, Jxy = 0, smooth = 0, safety = 0;
  Vec3 x0, yawpitchroll_0, x_head, x_eye_r, x_eye_l, x_pivot_l_b, x_pivot_r_b;
  WingKinematics kine_l, kine_r;
  std::string base_dir;
  bool body_already_drawn = false, cleaned_already_once = false;
  InsectMotion st;
  double time = 0.0;
  void init(const IniFile& ini, const Params& p, const std::string& base_dir);
  void update(double t);                       // Update_Insect (module_insects.f90:281-333)
  // the wings' from_file Fourier coefficients (shared by both wings when both
  // read the same series): count, value, and setting it on every such wing
  int ncontrols();
  double get_control(int p);
  void set_control(int p, double v);
  std::vector<double> kinematics_row(double t) const;  // write_kinematics_file columns
 private:
  InsectMotion motion(double t);
};

Mat3 mat_mul(const Mat3& a, const Mat3& b);
Mat3 transpose(const Mat3& a);
Vec3 mat_vec(const Mat3& a, const Vec3& v);

}  // namespace flusi
