// Solver context: device state, the RHS pass pipeline, the AB2/RK2/RK4
// steppers, create_mask, write_integrals and the C ABI of include/flusi_hip.h.
#include "../../include/flusi_hip.h"
#include "adjoint.h"
#include "fft_core.h"
#include "fields.h"
#include "flusi_kernels.h"
#include "h5bind.h"
#include "host_params.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <memory>
#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace flusi {

#define HIPCHECK(x)                                                                        \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) throw FlusiError(900001, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

static void kcheck(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw FlusiError(900002, std::string(what) + ": " + hipGetErrorString(e));
}

// Scratch of the allocation-free entry point flusi_energy_k_noalloc: the
// Parseval partial sums, the row reducer's middle rows and the result.  Each
// GPU holds its own copy until the process ends.  In host code the name
// `scratch` is a stand-in, not a device address: the address comes from
// hipGetSymbolAddress, once per device, and only that address is passed on.
// Every call on a GPU shares its copy (calls that overlap on one GPU race);
// each call writes every element it reads before reading it.
constexpr size_t kNoAllocScratchDoubles = 4096;  // 32 KiB
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];
constexpr int kMaxDevices = 64;
static double* scratch_dev[kMaxDevices] = {};
static std::mutex scratch_mu;

// the device address of `scratch` on the current device
static double* scratch_address() {
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    throw FlusiError(900020, "device " + std::to_string(dev) + " beyond the " + std::to_string(kMaxDevices) +
                                 " devices the allocation-free scratch serves");
  std::lock_guard<std::mutex> g(scratch_mu);
  if (!scratch_dev[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      throw FlusiError(900021, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    scratch_dev[dev] = static_cast<double*>(ptr);
  }
  return scratch_dev[dev];
}

// fluid_time_step.f90:1051-1060 truncate: write(str,'(es7.1)') a; read(str,*) a
static double truncate_dt(double a) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.1E", a);
  return std::strtod(buf, nullptr);
}

// vars.f90:429-462 time_for_output
static bool time_for_output(double time, double dt, int it, double tfrequ, int ifrequ, double tmax, double tfirst,
                            bool intelligent) {
  if (time < tfirst) return false;
  if (intelligent) {
    const double t1 = std::ceil(time / tfrequ) * tfrequ, t2 = std::floor(time / tfrequ) * tfrequ;
    return std::fabs(time - t1) <= 1e-6 || std::fabs(time - t2) <= 1e-6 || (it % ifrequ) == 0 ||
           std::fabs(time - tmax) <= 1e-6;
  }
  const double md = time - std::floor(time / tfrequ) * tfrequ;
  return (md < dt) || (it % ifrequ) == 0 || (time == tmax);
}

// vars.f90:408-422
static double startup_conditioner(double time, double trel, double tau) {
  const double t = time - trel;
  if (time <= trel) return 0.0;
  if (time > trel && time < trel + tau) return (t * t * t) / (-0.5 * tau * tau * tau) + 3.0 * (t * t) / (tau * tau);
  return 1.0;
}

static int fortran_nint(double x) { return x >= 0 ? (int)std::floor(x + 0.5) : -(int)std::floor(-x + 0.5); }

static bool file_exists(const std::string& f) {
  FILE* h = std::fopen(f.c_str(), "rb");
  if (h) std::fclose(h);
  return h != nullptr;
}

static const char* kTimerNames[FLUSI_TIMER_SLOTS] = {"ZINV", "YINV", "XPASS", "YFWD", "ZSTEP", "MASK", "DIAG", "OTHER"};
enum { TM_ZINV, TM_YINV, TM_X, TM_YFWD, TM_ZSTEP, TM_MASK, TM_DIAG, TM_OTHER };

class SolverBase {
 public:
  virtual ~SolverBase() = default;
  std::string last_error;
  Params p;
  std::unique_ptr<IniFile> ini;
  std::string base_dir, out_dir;
  int nx = 0, ny = 0, nz = 0, nxh = 0;
  long npts = 0, nc = 0;
  double time = 0.0, dt0 = 0.0, dt1 = 0.0;
  int it = 0, n0 = 0, n1 = 1;
  double vis_dt = -1.0;  // dt of the stored integrating factor, < 0: factor is zero
  bool fields_initialized = false;
  Insect insect;
  bool is_insect = false;
  Vec3 last_power4_xyz{0, 0, 0};
  double last_power_total = 0.0;

  virtual void init_fields() = 0;
  virtual void set_uk(const double* uk) = 0;
  virtual void get_uk(double* uk) = 0;
  virtual void get_u(double* u, bool vort) = 0;
  virtual void get_pressure(double time, double* p) = 0;
  virtual void cal_nlk(double time, double* nlk, double* u) = 0;
  virtual void fluid_timestep() = 0;
  virtual void create_mask(double time) = 0;
  virtual void get_mask(double* mask, double* us, int16_t* color) = 0;
  virtual void write_integrals(double time, double* out) = 0;
  virtual void cal_drag(double time, const double* u, double* out) = 0;
  virtual double energy_k() = 0;
  virtual double energy_k_noalloc() = 0;
  virtual void cal_vis(double dt, double* vis) = 0;
  virtual void fft1(const double* f, double* fk) = 0;
  virtual void ifft1(const double* fk, double* f) = 0;
  virtual void timers(int enable, double* ms) = 0;
  virtual void autotune_v(int reps, int* variant, double* ms) = 0;
  virtual int adjoint_ncontrols() = 0;
  virtual void adjoint_get_controls(double* theta) = 0;
  virtual double adjoint_forward(const double* theta, int nsteps, int every) = 0;
  virtual void adjoint_backward(double* grad) = 0;
  virtual void save_fields() = 0;
  virtual void dump_runtime_backup() = 0;
  virtual void read_runtime_backup(const std::string& file) = 0;

  // runtime backups: the file written next (runtime_backup0/1, time_step.f90:15)
  int nbackup = 0;
  bool restarted = false;  // the fields came from a runtime backup

  // output files: relative names live in the output directory
  std::string out_path(const std::string& name) const {
    if (name.empty() || name[0] == '/' || out_dir.empty()) return name;
    return out_dir + "/" + name;
  }
  bool saves_fields() const {
    return p.iSaveVelocity == 1 || p.iSavePress == 1 || p.iSaveVorticity == 1 || p.iSaveMagVorticity == 1 ||
           (p.iSaveMask == 1 && p.iPenalization == 1) ||
           (p.iSaveSolidVelocity == 1 && p.iPenalization == 1 && p.iMoving == 1);
  }

  // wall clock of the time-stepping loop's start (time_step.f90:53, t4)
  std::chrono::steady_clock::time_point wtime_start{};
  bool wtime_set = false;
  virtual void check_nan() = 0;

  // time_step.f90:57-167 loop body
  void step() {
    if (!wtime_set) { wtime_start = std::chrono::steady_clock::now(); wtime_set = true; }
    dt0 = dt1;
    fluid_timestep();
    time = time + dt1;
    it += 1;
    if (p.checknan_every > 0 && it % p.checknan_every == 0) check_nan();
    // time_step.f90:163, 189-201 save_time_stepping_info: time, wall time
    // since the loop started, it, dt, number of MPI ranks (one)
    if (!out_dir.empty()) {
      const double wt = std::chrono::duration<double>(std::chrono::steady_clock::now() - wtime_start).count();
      char line[160];
      std::snprintf(line, sizeof(line), "%15.8E %15.8E %9d %15.8E %6d\n", time, wt, it, dt1, 1);
      append_line("timesteps_info.t", line);
    }
    if (time_for_output(time, dt1, it, p.tintegral, p.itdrag, p.tmax, 0.0, p.intelligent_dt == "yes"))
      write_integrals(time, nullptr);
    // time_step.f90:97-105 field snapshots (written only with an output directory)
    if (!out_dir.empty() && saves_fields() &&
        time_for_output(time, dt1, it, p.tsave, 99999999, p.tmax, p.tsave_first, p.intelligent_dt == "yes"))
      save_fields();
  }

  // time_step.f90:1-180: the loop, with the wall-clock backups every truntime
  // hours, the wtimemax stop and the final backup
  void run(int nsteps) {
    const auto t0 = std::chrono::steady_clock::now();
    auto hours = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / 3600.0; };
    double truntimenext = p.truntime;
    const bool backups = p.iDoBackup == 1 && !out_dir.empty();
    // init_fields.f90:108-111 and time_step.f90:47-50: initial snapshot and
    // integrals, unless resuming a backup
    if (!restarted) {
      if (!out_dir.empty() && saves_fields() && time >= p.tsave_first) save_fields();
      write_integrals(time, nullptr);
    }
    int n = 0;
    wtime_start = std::chrono::steady_clock::now();
    wtime_set = true;
    while (time < p.tmax && it <= p.nt && (nsteps <= 0 || n < nsteps)) {
      step();
      ++n;
      if (backups && truntimenext < hours()) { dump_runtime_backup(); truntimenext += p.truntime; }
      if (p.wtimemax < hours()) break;
    }
    if (backups) dump_runtime_backup();
  }

  void append_line(const std::string& name, const char* line) {
    if (out_dir.empty()) return;
    const std::string path = out_dir + "/" + name;
    FILE* f = std::fopen(path.c_str(), "a");
    if (!f) throw FlusiError(900010, "cannot open " + path);
    std::fputs(line, f);
    std::fclose(f);
  }
  void append(const std::string& name, const std::vector<double>& row, const char* header) {
    if (out_dir.empty()) return;
    const std::string path = out_dir + "/" + name;
    FILE* f = std::fopen(path.c_str(), "a");
    if (!f) throw FlusiError(900010, "cannot open " + path);
    if (std::ftell(f) == 0 && header) std::fprintf(f, "%s\n", header);
    for (double v : row) std::fprintf(f, "%15.8E ", v);
    std::fprintf(f, "\n");
    std::fclose(f);
  }
};

template <typename T>
class Solver : public SolverBase {
 public:
  using C = typename V2<T>::t;
  static constexpr bool FP64 = sizeof(T) == 8;
  hipStream_t stream = nullptr;
  // device memory
  void* uk = nullptr;
  void* reg[2] = {nullptr, nullptr};
  void* Q = nullptr;
  void* P = nullptr;
  long prow = 0, pc = 0;    // P-layout row stride (kx) and field stride
  void* mask = nullptr;
  void* us = nullptr;
  uint8_t* color = nullptr;
  // per x line: may the mask be nonzero (the x pass reads only those lines'
  // mask); valid after an unbatched, uncaptured insect mask update, which
  // recomputes the lines it redrew
  uint8_t* mflag = nullptr;
  void* mzero = nullptr;
  bool mflag_valid = false;
  bool mflag_on = true;     // FLUSI_MASK_FLAGS=0: every line's mask is read
  void* phys = nullptr;     // 7 real fields for diagnostics (u, w, div)
  double* tab = nullptr;    // kx ky kz kxt2 kyt2 kzt2
  double *kx_d, *ky_d, *kz_d, *kxt2_d, *kyt2_d, *kzt2_d;
  void* tw_z = nullptr; void* tw_y = nullptr; void* tw_M = nullptr; void* tw_N = nullptr;
  double* xpart = nullptr;  // x-pass partial rows (max|u|^2, penalization sums)
  double* caps_d = nullptr;   // bumblebee capsule records
  // insect mask arguments live in device memory, one record per batch member:
  // eager calls cycle through INS_SLOTS slots, each filled from a pinned host
  // copy whose event says when it may be rewritten; a captured step reads
  // slot s of ins_cap for its s-th mask update, which the host fills before
  // every replay
  static constexpr int INS_SLOTS = 8;
  InsectArgs* ins_d = nullptr;
  InsectArgs* ins_h = nullptr;
  int ins_per_slot = 0;
  hipEvent_t ins_ev[INS_SLOTS] = {};
  int ins_next = 0;
  InsectArgs* ins_cap = nullptr;
  bool capture = false;        // recording a graph: fixed launch shapes, no host reads
  int stage_slot = 0;
  // batches: a leader launches every pass once for all members of `batch`
  // (itself first), whose buffers sit in slabs bat.mstride bytes apart
  std::vector<Solver<T>*> batch;
  Batch bat{1, 0};             // the launch batch (nmem > 1 only while a batched step runs)
  bool ens_step = false;       // an ensemble step is being launched (it records the outputs)
  char* slab = nullptr;        // setup() places its buffers here when set
  size_t slab_cap = 0, slab_used = 0;
  bool in_setup = false;
  std::vector<void*> owned;    // device allocations freed by this solver
  bool zwave = false;          // z-forward step pass with in-wave line transforms
  int zeu = 0;                 // block z step, nz >= 256: epilogue positions scheduled together (0 = all)
  double* fw_d = nullptr;     // Fourier wing radius table
  bool xwave = false;       // right-hand-side x pass with one wavefront per line
  double* zpart = nullptr; int nzblocks = 0;
  double* dpart = nullptr; int ndblocks = 0;
  double* noalloc_scratch = nullptr;  // device address of `scratch` on this context's GPU
  double* red = nullptr;    // small results
  double* rmid = nullptr;   // reduction scratch
  int* kinds = nullptr;     // reduction kinds: [0..3] x pass, [4] energy, [8..8+DIAG_NV) diag
  double* stage = nullptr; size_t stage_n = 0;
  std::vector<double> kx_h, ky_h, kz_h, kxt2_h, kyt2_h, kzt2_h;
  double ksy = 0.0, ksz = 0.0;  // 2 pi / yl, 2 pi / zl
  GridGeom geom{};
  bool timing = false;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev;
  double tm_ms[FLUSI_TIMER_SLOTS] = {0};

  ~Solver() override {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& e : ev) { (void)hipEventDestroy(e.second.first); (void)hipEventDestroy(e.second.second); }
    for (void* q : owned) (void)hipFree(q);
    for (hipEvent_t e : ins_ev) if (e) (void)hipEventDestroy(e);
    if (ins_h) (void)hipHostFree(ins_h);
    if (stream) (void)hipStreamDestroy(stream);
  }

  // device memory: from the slab while setup() runs with one, else its own
  // allocation; slab_used counts setup()'s bytes either way
  void* dalloc(size_t bytes) {
    bytes = (std::max<size_t>(bytes, 256) + 255) / 256 * 256;
    if (in_setup) {
      slab_used += bytes;
      if (slab) {
        if (slab_used > slab_cap) throw FlusiError(900013, "batch slab too small");
        return slab + (slab_used - bytes);
      }
    }
    void* q = nullptr;
    HIPCHECK(hipMalloc(&q, bytes));
    owned.push_back(q);
    return q;
  }
  template <typename P> void dfree(P*& q) {
    if (!q) return;
    for (auto& o : owned)
      if (o == (void*)q) { (void)hipFree(o); o = owned.back(); owned.pop_back(); break; }
    q = nullptr;
  }

  bool fft_ok = false;
  void need_fft() const {
    if (!fft_ok)
      throw FlusiError(900003, "grid " + std::to_string(nx) + "x" + std::to_string(ny) + "x" + std::to_string(nz) +
                                   " has no FFT plan (nx/2 in 8..256 and ny, nz in 16..512, each 2^a or 3*2^a); "
                                   "only mask generation is available");
  }

  void setup() {
    in_setup = true;
    slab_used = 0;
    setup_buffers();
    in_setup = false;
  }
  void setup_buffers() {
    nx = p.nx; ny = p.ny; nz = p.nz; nxh = nx / 2 + 1;
    fft_ok = nx % 2 == 0 && x_length_supported(nx / 2) && fft_length_supported(ny) && fft_length_supported(nz);
    npts = (long)nx * ny * nz;
    nc = (long)nxh * ny * nz;
    geom = GridGeom{nx, ny, nz, nxh, npts, nc, p.dx(), p.dy(), p.dz(), p.xl, p.yl, p.zl, p.periodic ? 1 : 0};
    HIPCHECK(hipStreamCreate(&stream));
    noalloc_scratch = scratch_address();  // once per device, with this context's device current
    const size_t cs = sizeof(C) * nc;
    mask = dalloc(sizeof(T) * npts);
    us = dalloc(3 * sizeof(T) * npts);
    color = (uint8_t*)dalloc((npts + 3) / 4 * 4);
    launch_zero(mask, sizeof(T) * npts, stream);
    launch_zero(us, 3 * sizeof(T) * npts, stream);
    launch_zero(color, (npts + 3) / 4 * 4, stream);
    if (fft_ok) {
      uk = dalloc(3 * cs);
      reg[0] = dalloc(3 * cs);
      reg[1] = dalloc(3 * cs);
      Q = dalloc(6 * cs);
      // P rows padded to whole 128-byte lines: the y passes' 16-kx (8 for
      // FP64) segments and the x pass's lines then start on line boundaries
      const long seg = 128 / (long)sizeof(C);
      prow = (nxh + seg - 1) / seg * seg;
      pc = prow * ny * nz;
      P = dalloc(6 * sizeof(C) * pc);
      launch_zero(uk, 3 * cs, stream);
      launch_zero(reg[0], 3 * cs, stream);
      launch_zero(reg[1], 3 * cs, stream);
    }
    kcheck("zero state");
    // wavenumbers (cof_p3dfft.f90:661-680) and dealias terms (vis.f90:23-53)
    const double sx = 2.0 * PI / p.xl, sy = 2.0 * PI / p.yl, sz = 2.0 * PI / p.zl;
    ksy = sy; ksz = sz;
    const double kxt = (2.0 / 3.0) * double(nx / 2 - 1), kyt = (2.0 / 3.0) * double(ny / 2 - 1),
                 kzt = (2.0 / 3.0) * double(nz / 2 - 1);
    kx_h.resize(nxh); ky_h.resize(ny); kz_h.resize(nz);
    kxt2_h.resize(nxh); kyt2_h.resize(ny); kzt2_h.resize(nz);
    auto wrap = [](int i, int n) { return ((i + n / 2) % n) - n / 2; };
    for (int i = 0; i < nxh; ++i) { kx_h[i] = sx * double(i); double t = (kx_h[i] / sx) / kxt; kxt2_h[i] = t * t; }
    for (int i = 0; i < ny; ++i) { ky_h[i] = sy * double(wrap(i, ny)); double t = (ky_h[i] / sy) / kyt; kyt2_h[i] = t * t; }
    for (int i = 0; i < nz; ++i) { kz_h[i] = sz * double(wrap(i, nz)); double t = (kz_h[i] / sz) / kzt; kzt2_h[i] = t * t; }
    const size_t ntab = 2 * ((size_t)nxh + ny + nz);
    tab = (double*)dalloc(ntab * sizeof(double));
    std::vector<double> h;
    h.insert(h.end(), kx_h.begin(), kx_h.end());
    h.insert(h.end(), ky_h.begin(), ky_h.end());
    h.insert(h.end(), kz_h.begin(), kz_h.end());
    h.insert(h.end(), kxt2_h.begin(), kxt2_h.end());
    h.insert(h.end(), kyt2_h.begin(), kyt2_h.end());
    h.insert(h.end(), kzt2_h.begin(), kzt2_h.end());
    HIPCHECK(hipMemcpy(tab, h.data(), ntab * sizeof(double), hipMemcpyHostToDevice));
    kx_d = tab; ky_d = kx_d + nxh; kz_d = ky_d + ny; kxt2_d = kz_d + nz; kyt2_d = kxt2_d + nxh; kzt2_d = kyt2_d + ny;
    if (fft_ok) {
      tw_z = make_tw(nz, nz); tw_y = make_tw(ny, ny); tw_M = make_tw(nx / 2, nx / 2); tw_N = make_tw(nxh, nx);
      if (const char* e = getenv("FLUSI_TILE_WAVE")) twave = atoi(e) != 0;
      if (const char* e = getenv("FLUSI_ZWAVE")) zwave = atoi(e) != 0;
      if (const char* e = getenv("FLUSI_ZSTEP_EU")) zeu = atoi(e);
      if (const char* e = getenv("FLUSI_MASK_FLAGS")) mflag_on = atoi(e) != 0;
      if (const char* e = getenv("FLUSI_TILE_PREFETCH")) tpf = atoi(e) != 0;
      if (const char* e = getenv("FLUSI_XWAVE_WPE")) xwpe = atoi(e);
      xwave = x_rhs_wave_supported(nx);
      if (const char* e = getenv("FLUSI_XWAVE")) xwave = xwave && atoi(e) != 0;
      // partial rows for either x-pass form, so the variant can change later
      const int nxb = std::max(x_rhs_wave_supported(nx) ? x_rhs_wave_blocks(nx, ny * nz) : 0,
                               x_pass_blocks(nx, ny * nz, FP64));
      xpart = (double*)dalloc(sizeof(double) * 4 * nxb);
      // room for either z-step form's partials (the variant can change later)
      nzblocks = std::max(zstep_blocks(nz, ny, nxh, FP64, false), zstep_blocks(nz, ny, nxh, FP64, true));
      zpart = (double*)dalloc(sizeof(double) * nzblocks);
      ndblocks = diag_blocks<T>(geom);
      dpart = (double*)dalloc(sizeof(double) * DIAG_NV * ndblocks);
    }
    red = (double*)dalloc(sizeof(double) * 128);
    rmid = (double*)dalloc(sizeof(double) * reduce_rows_scratch(DIAG_NV));
    std::vector<int> k(8 + DIAG_NV, 0);
    k[0] = 1;  // umax^2: max
    for (int v = 0; v < DIAG_NV; ++v) {
      if (v >= DG_URES && v < DG_URES + 4) k[8 + v] = 2;
      if (v == DG_DIVMAX || v == DG_DIVMAXF) k[8 + v] = 1;
    }
    kinds = (int*)dalloc(sizeof(int) * k.size());
    HIPCHECK(hipMemcpy(kinds, k.data(), sizeof(int) * k.size(), hipMemcpyHostToDevice));
    HIPCHECK(hipStreamSynchronize(stream));
  }

  void* make_tw(int n, int period) {
    std::vector<C> t(n);
    for (int m = 0; m < n; ++m) {
      const double a = -2.0 * PI * double(m) / double(period);
      t[m].x = (T)std::cos(a);
      t[m].y = (T)std::sin(a);
    }
    void* d = dalloc(sizeof(C) * n);
    HIPCHECK(hipMemcpy(d, t.data(), sizeof(C) * n, hipMemcpyHostToDevice));
    return d;
  }

  // ------------------------------------------------------------ timing
  template <typename F> void timed(int slot, F&& f) {
    if (!timing) { f(); return; }
    hipEvent_t a, b;
    HIPCHECK(hipEventCreate(&a)); HIPCHECK(hipEventCreate(&b));
    HIPCHECK(hipEventRecord(a, stream));
    f();
    HIPCHECK(hipEventRecord(b, stream));
    ev.push_back({slot, {a, b}});
  }
  void timers(int enable, double* ms) override {
    HIPCHECK(hipStreamSynchronize(stream));
    for (auto& e : ev) {
      float t = 0.f;
      (void)hipEventElapsedTime(&t, e.second.first, e.second.second);
      tm_ms[e.first] += t;
      (void)hipEventDestroy(e.second.first); (void)hipEventDestroy(e.second.second);
    }
    ev.clear();
    if (ms) for (int i = 0; i < FLUSI_TIMER_SLOTS; ++i) ms[i] = tm_ms[i];
    if (enable >= 0) { timing = enable != 0; for (double& v : tm_ms) v = 0.0; }
  }

  // ------------------------------------------------------------ passes
  C* cf(void* base, int f) { return reinterpret_cast<C*>(base) + (long)f * nc; }
  // kernel variants (chosen by autotune(), or by the FLUSI_* environment overrides)
  bool twave = true;        // y/z passes with the line transform inside one wavefront
  bool tpf = true;          // block y/z passes: prefetch the next job's operand
  // one-wave-per-line x pass: waves-per-SIMD register budget (3 or 0); FP64
  // keeps the unconstrained allocation (at 3 it would spill ~480 B per lane)
  int xwpe = FP64 ? 0 : 3;
  void tile(const TilePassArgs& a0, int L, int dir, bool contig_in) {
    TilePassArgs a = a0;
    a.pf = tpf ? 1 : 0;
    a.bat = bat;
    if (twave && tile_wave_supported(L)) launch_tile_wave<T>(a, L, dir, contig_in, stream);
    else launch_tile_pass<T>(a, L, dir, contig_in, stream);
  }
  C* cfp(int f) { return reinterpret_cast<C*>(P) + (long)f * pc; }  // field f of P

  void zinv(const TileJob* jobs, int nj) {  // S -> Q
    TilePassArgs a{};
    for (int j = 0; j < nj; ++j) a.jobs[j] = jobs[j];
    a.njobs = nj; a.nb = ny; a.nc = nxh;
    a.sin = Strides{1, nz, (long)ny * nz};
    a.sout = Strides{ny, 1, (long)nz * ny};
    a.kscale = ksz; a.scale = 1.0; a.tw = tw_z;
    timed(TM_ZINV, [&] { tile(a, nz, +1, true); });
    kcheck("zinv");
  }
  void yinv(const TileJob* jobs, int nj) {  // Q -> P
    TilePassArgs a{};
    for (int j = 0; j < nj; ++j) a.jobs[j] = jobs[j];
    a.njobs = nj; a.nb = nxh; a.nc = nz;
    // the last tile column stores whole 128-byte segments (zeros beyond nxh):
    // a store of the lone Nyquist value alone is a partial-line write that the
    // memory side turns into a read-modify-write (YINV at 512^3: 5.41 -> 4.70 ms
    // per RK4 step)
    a.nb_store = (int)std::min<long>(prow, (nxh + 15) / 16 * 16);
    a.sin = Strides{1, (long)nz * ny, ny};
    a.sout = Strides{prow, 1, (long)ny * prow};
    a.kscale = ksy; a.scale = 1.0; a.tw = tw_y;
    timed(TM_YINV, [&] { tile(a, ny, +1, true); });
    kcheck("yinv");
  }
  void yfwd(int nf) {  // P -> Q
    TilePassArgs a{};
    for (int j = 0; j < nf; ++j) a.jobs[j] = TileJob{cfp(j), nullptr, cf(Q, j), OP_COPY, nullptr, 0};
    a.njobs = nf; a.nb = nxh; a.nc = nz;
    a.sin = Strides{prow, 1, (long)ny * prow};
    a.sout = Strides{1, (long)nz * ny, ny};
    a.kscale = ksy; a.scale = 1.0; a.tw = tw_y;
    timed(TM_YFWD, [&] { tile(a, ny, -1, false); });
    kcheck("yfwd");
  }
  void zfwd_plain(int nf, void* dst) {  // Q -> S with 1/N (cof_p3dfft.f90:52-56)
    TilePassArgs a{};
    for (int j = 0; j < nf; ++j) a.jobs[j] = TileJob{cf(Q, j), nullptr, cf(dst, j), OP_COPY, nullptr, 0};
    a.njobs = nf; a.nb = ny; a.nc = nxh;
    a.sin = Strides{ny, 1, (long)nz * ny};
    a.sout = Strides{1, nz, (long)ny * nz};
    a.kscale = ksz; a.scale = 1.0 / (double(nx) * double(ny) * double(nz)); a.tw = tw_z;
    timed(TM_ZSTEP, [&] { tile(a, nz, -1, false); });
    kcheck("zfwd");
  }
  XPassArgs xargs() {
    XPassArgs a{};
    a.nlines = ny * nz; a.nx = nx; a.ny = ny; a.nz = nz; a.nxh = nxh; a.prow = prow;
    a.kx = kx_d; a.kx_scale = 2.0 * PI / p.xl; a.twM = tw_M; a.twN = tw_N;
    a.npts = npts; a.wpe = xwpe;
    a.bat = bat;
    return a;
  }
  // the five fields of the z-inverse pass and the six of the y-inverse pass:
  // Q: 0 U_x, 1 U_y, 2 U_z, 3 D_x = F_z^-1(i kz u_x), 4 D_y = F_z^-1(i kz u_y)
  // P: 0 V_x, 1 V_y, 2 V_z, 3 w_x, 4 dz u_x, 5 dy u_x
  void inv_uw(const void* u_in) {
    const C* u = reinterpret_cast<const C*>(u_in);
    // each operand is read once: u_x gives U_x and D_x, u_y gives U_y and D_y
    TileJob zj[3] = {{u, nullptr, cf(Q, 0), OP_COPY, cf(Q, 3), OP_DERIV},
                     {u + nc, nullptr, cf(Q, 1), OP_COPY, cf(Q, 4), OP_DERIV},
                     {u + 2 * nc, nullptr, cf(Q, 2), OP_COPY, nullptr, 0}};
    zinv(zj, 3);
    // (U_z, D_y) give V_z and w_x; U_x gives V_x and dy u_x.  The two-operand
    // job goes first: both its reads start with the kernel, the later jobs'
    // reads are prefetched behind the transforms
    TileJob yj[4] = {{cf(Q, 2), cf(Q, 4), cfp(2), OP_COPY, cfp(3), OP_DERIV_MINUS},
                     {cf(Q, 0), nullptr, cfp(0), OP_COPY, cfp(5), OP_DERIV},
                     {cf(Q, 1), nullptr, cfp(1), OP_COPY, nullptr, 0},
                     {cf(Q, 3), nullptr, cfp(4), OP_COPY, nullptr, 0}};
    yinv(yj, 4);
  }

  int sponge_code() const {
    if (p.iVorticitySponge != "yes") return 0;
    if (p.iSpongeType == "cavity") return 1;
    if (p.iSpongeType == "outlet_x") return 2;
    throw FlusiError(3335, "iSpongeType not supported: " + p.iSpongeType);
  }

  // rhs.f90:2-115 up to the forward transform: mask, inverse passes, x pass,
  // and the reduction of max|u|^2 and sum chi (u - us) into red[0..3]
  void pre_rhs(const void* u_in, double t, bool with_mask = true) {
    if (with_mask && p.iMoving == 1 && p.iPenalization == 1) create_mask(t);
    if (p.eps < 1.0e-11) throw FlusiError(77363800, "Value of eps is very small, maybe even zero.");
    inv_uw(u_in);
    XPassArgs a = xargs();
    a.in = P; a.out = P; a.fstride_in = pc; a.fstride_out = pc; a.nf = 6;
    a.mask = mask; a.us = us; a.eps_inv = 1.0 / p.eps;
    a.mflag = mflag_valid ? mflag : nullptr; a.mzero = mzero;
    a.sponge = sponge_code();
    a.sponge_thickness = p.sponge_thickness;
    a.sponge_eps_inv = a.sponge ? 1.0 / p.eps_sponge : 0.0;
    a.sponge_mask_thr = (a.sponge == 2) ? (double)1.0e-12f : 1.0e-12;
    a.partial = xpart;
    a.want_stats = (p.dt_fixed <= 0.0 || p.iMeanFlow[0] == "dynamic" || p.iMeanFlow[1] == "dynamic" ||
                    p.iMeanFlow[2] == "dynamic") ? 1 : 0;
    timed(TM_X, [&] {
      if (xwave) launch_x_rhs_wave<T>(a, stream);
      else launch_x_pass<T>(a, XM_RHS, stream);
    });
    kcheck("x pass");
    if (a.want_stats) {
      const int nxblocks = xwave ? x_rhs_wave_blocks(nx, ny * nz) : x_pass_blocks(nx, ny * nz, FP64);
      for (int m = 0; m < batch_n(bat); ++m) {
        const long o = m * bat.mstride;
        launch_reduce_rows(mshift(xpart, o), nxblocks, 4, mshift(kinds, o), mshift(rmid, o), mshift(red, o), stream);
      }
    }
    kcheck("reduce x");
  }

  ZStepArgs zargs(int mode) {
    ZStepArgs z{};
    z.nq = Q; z.fs = nc; z.nx = nx; z.ny = ny; z.nz = nz; z.nxh = nxh;
    z.kx = kx_d; z.ky = ky_d; z.kz = kz_d; z.kxt2 = kxt2_d; z.kyt2 = kyt2_d; z.kzt2 = kzt2_d;
    z.tw = tw_z; z.scale = 1.0 / (double(nx) * double(ny) * double(nz));
    z.mode = mode; z.sponge = sponge_code() ? 1 : 0; z.dealias_on = p.iDealias == 1; z.nu = p.nu;
    z.explicit_diffusion = (p.method == "RK4") ? 1 : 0;
    z.vis_dt_new = -1.0; z.vis_dt_old = -1.0;
    z.pen_sum = red + 1;
    double sc = p.dx() * p.dy() * p.dz();
    if (p.startup_conditioner == "yes") sc *= startup_conditioner(time, p.T_release_meanflow, p.tau_meanflow);
    z.dynamic_scale = sc / p.m_fluid;
    for (int c = 0; c < 3; ++c) z.meanflow_dynamic[c] = p.iMeanFlow[c] == "dynamic";
    z.wave = zwave ? 1 : 0;
    z.eu = zeu;
    z.bat = bat;
    return z;
  }
  void post_rhs(ZStepArgs& z) {
    yfwd(z.sponge ? 6 : 3);
    const bool fin = z.mode == SM_EULER || z.mode == SM_AB2 || z.mode == SM_RK2_2 || z.mode == SM_RK4_4;
    if (fin) {
      // set_mean_flow (fluid_time_step.f90:1066-1090) with the time of the step start
      for (int c = 0; c < 3; ++c) {
        z.set_mean[c] = 0;
        if (p.iMeanFlow[c] == "fixed") { z.set_mean[c] = 1; z.mean_value[c] = p.umean[c]; }
        else if (p.iMeanFlow[c] == "sinusoidal") {
          z.set_mean[c] = 1; z.mean_value[c] = p.umean[c] * std::sin(2.0 * PI * p.umean_freq * time);
        }
      }
      z.energy_partial = zpart;
    }
    timed(TM_ZSTEP, [&] { launch_zstep<T>(z, stream); });
    kcheck("zstep");
  }

  double host_umax() {
    double v = 0.0;
    HIPCHECK(hipMemcpyAsync(&v, red, sizeof(double), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    return std::sqrt(v);
  }

  // fluid_time_step.f90:917-1042 adjust_dt
  double adjust_dt(double t) {
    if (p.dt_fixed > 0.0) return p.dt_fixed;
    const double umax = host_umax();
    if (std::isnan(umax)) throw FlusiError(100011, "Evolved field contains a NAN: aborting run.");
    if (umax >= 1.0e3) throw FlusiError(100012, "Umax is very big, surely this is an error");
    const double h = std::min(p.dx(), std::min(p.dy(), p.dz()));
    double d = (umax >= 1.0e-8) ? h * p.cfl / umax : 1.0e-3;
    d = truncate_dt(d);
    if (p.dt_max > 0.0) d = std::min(d, p.dt_max);
    if (p.iPenalization > 0) d = std::min(p.cfl_eta * p.eps, d);
    if (p.method == "RK4") d = std::min(d, 0.5 * h * h / p.nu);
    if (p.intelligent_dt == "yes") {
      if (t >= p.tsave_first) {
        d = std::min(d, 0.98 * p.tsave);
        const double tt = std::ceil(t / p.tsave) * p.tsave;
        if (t + d > tt && std::fabs(t - tt) >= 1.0e-6) d = tt - t;
      }
      d = std::min(d, 0.98 * p.tintegral);
      double tt = std::ceil(t / p.tintegral) * p.tintegral;
      if (t + d > tt && std::fabs(t - tt) >= 1.0e-6) d = tt - t;
      tt = p.tmax;
      if (t + d > tt && std::fabs(t - tt) >= 1.0e-6) d = p.tmax - t;
    }
    append("dt.t", {t, d, d * umax / h, d * p.nu / (h * h), d / p.eps},
           "%          time              dt             CFL         viscous    penalization");
    return d;
  }

  // force: record also inside an ensemble step (whose bookkeeping calls this
  // for every member after the step)
  void record_ekin(bool force = false) {
    if (out_dir.empty() || (ens_step && !force)) return;
    launch_reduce_rows(zpart, zstep_blocks(nz, ny, nxh, FP64, zwave), 1, kinds + 4, rmid, red + 8, stream);
    double e = 0.0;
    HIPCHECK(hipMemcpyAsync(&e, red + 8, sizeof(double), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    append("ekin.t", {time + dt1, e * p.xl * p.yl * p.zl}, "%          time            ekin");
  }

  // basic_operators.f90:925-950 checknan_cmplx on uk, from the Parseval sum
  // the last z step of a step accumulates over every entry of the new uk (a
  // NaN or Inf anywhere makes it non-finite)
  void check_nan() override {
    need_fft();
    launch_reduce_rows(zpart, zstep_blocks(nz, ny, nxh, FP64, zwave), 1, kinds + 4, rmid, red + 10, stream);
    kcheck("checknan");
    double e = 0.0;
    HIPCHECK(hipMemcpyAsync(&e, red + 10, sizeof(double), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    if (!std::isfinite(e)) throw FlusiError(202, "NaN in uk after time step " + std::to_string(it));
  }

  // ------------------------------------------------------------ autotuning
  // The right-hand side has interchangeable kernel variants: in-wave or block
  // y/z passes, operand prefetch in the block passes, the one-wave-per-line or
  // the block x pass with its register budget, and the in-wave or block z
  // step.  autotune() runs `reps` right-hand sides of the current state (the
  // whole batch when one is active) for every variant the grid allows and
  // keeps the fastest; results are cached per grid, precision, sponge and
  // batch size for the life of the process.
  struct Variant { int twave, tpf, xwave, xwpe, zwave, zeu; };
  Variant variant() const { return Variant{twave, tpf, xwave, xwpe, zwave, zeu}; }
  void apply(const Variant& v) {
    twave = v.twave != 0; tpf = v.tpf != 0; xwave = v.xwave != 0; xwpe = v.xwpe; zwave = v.zwave != 0; zeu = v.zeu;
  }
  double time_rhs(int reps) {
    auto rhs = [&] {
      pre_rhs(uk, time, false);
      ZStepArgs z = zargs(SM_RHS_ONLY);
      z.rhs_out = P; z.udiff = uk;
      post_rhs(z);
    };
    rhs();  // first launches of a variant set its LDS limits
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0)); HIPCHECK(hipEventCreate(&e1));
    HIPCHECK(hipEventRecord(e0, stream));
    for (int r = 0; r < reps; ++r) rhs();
    HIPCHECK(hipEventRecord(e1, stream));
    HIPCHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
    return ms / std::max(reps, 1);
  }
  std::vector<Variant> candidates() const {
    const bool tw_any = tile_wave_supported(ny) || tile_wave_supported(nz);
    const bool tw_all = tile_wave_supported(ny) && tile_wave_supported(nz);
    const bool xw = x_rhs_wave_supported(nx), zw = zstep_wave_supported(nz);
    std::vector<Variant> c;
    for (int tw : {1, 0}) {
      if (tw && !tw_any) continue;
      for (int pf : {1, 0}) {
        if (!pf && tw && tw_all) continue;  // prefetch only acts in block passes
        for (int xv : {1, 0}) {
          if (xv && !xw) continue;
          for (int wpe : {3, 0}) {
            if (!xv && wpe == 0) continue;  // register budget only acts in the wave x pass
            for (int zv : {0, 1}) {
              if (zv && !zw) continue;
              for (int eu : {0, 2, 4}) {
                if (eu && (zv || nz < 256)) continue;  // grouping acts in the block z step at nz >= 256
                c.push_back(Variant{tw, pf, xv, wpe, zv, eu});
              }
            }
          }
        }
      }
    }
    return c;
  }
  static std::map<std::string, std::pair<Variant, double>>& tune_cache() {
    static std::map<std::string, std::pair<Variant, double>> m;
    return m;
  }
  static std::mutex& tune_mutex() { static std::mutex mu; return mu; }
  std::string tune_key() const {
    return std::to_string(nx) + "x" + std::to_string(ny) + "x" + std::to_string(nz) + (FP64 ? ":f64" : ":f32") +
           ":sp" + std::to_string(sponge_code()) + ":m" + std::to_string(batch_n(bat));
  }
  static constexpr int kVariantInts = 6;
  // v: the chosen variant (6 ints: twave, tpf, xwave, xwpe, zwave, zeu); ms: its
  // time per right-hand side; force: tune again even when cached
  void autotune(int reps, int* v, double* ms, bool force) {
    need_fft();
    if (!fields_initialized) throw FlusiError(900004, "fields not initialised (call flusi_init_fields)");
    const std::string key = tune_key();
    std::pair<Variant, double> best{variant(), 1e300};
    bool hit = false;
    {
      std::lock_guard<std::mutex> g(tune_mutex());
      auto it = tune_cache().find(key);
      if (it != tune_cache().end() && !force) { best = it->second; hit = true; }
    }
    if (!hit) {
      const Variant keep = variant();
      for (const Variant& c : candidates()) {
        apply(c);
        const double t = time_rhs(reps);
        if (t < best.second) best = {c, t};
      }
      apply(keep);
      std::lock_guard<std::mutex> g(tune_mutex());
      tune_cache()[key] = best;
    }
    apply(best.first);
    if (v) {
      const Variant& b = best.first;
      const int o[6] = {b.twave, b.tpf, b.xwave, b.xwpe, b.zwave, b.zeu};
      for (int i = 0; i < kVariantInts; ++i) v[i] = o[i];
    }
    if (ms) *ms = best.second;
  }
  void autotune_v(int reps, int* v, double* ms) override { autotune(reps, v, ms, false); }

  // ------------------------------------------------------------ steppers
  void fluid_timestep() override {
    need_fft();
    if (!fields_initialized) throw FlusiError(900004, "fields not initialised (call flusi_init_fields)");
    const std::string& m = p.method;
    if (m == "AB2") {
      if (it == 0) {  // fluid_time_step.f90:675-748 euler_startup
        pre_rhs(uk, time);
        dt1 = adjust_dt(time);
        if (dt1 != dt0) vis_dt = dt1;
        ZStepArgs z = zargs(SM_EULER);
        z.u = uk; z.r0 = reg[n0]; z.dt = dt1; z.vis_dt_new = vis_dt;
        post_rhs(z);
      } else {        // :752-832 adamsbashforth
        pre_rhs(uk, time);
        double d = adjust_dt(time);
        if (d > p.tmax - time) d = p.tmax - time;
        dt1 = d;
        ZStepArgs z = zargs(SM_AB2);
        z.b10 = dt1 / dt0 * (0.5 * dt1 + dt0);
        z.b11 = -0.5 * dt1 * dt1 / dt0;
        if (dt1 != dt0) vis_dt = dt1;
        z.u = uk; z.r0 = reg[n0]; z.r1 = reg[n1]; z.dt = dt1; z.vis_dt_new = vis_dt;
        post_rhs(z);
      }
    } else if (m == "RK2") {  // :519-598
      pre_rhs(uk, time);
      dt1 = adjust_dt(time);
      const double old = vis_dt;
      if (dt1 != dt0) vis_dt = dt1;
      ZStepArgs z = zargs(SM_RK2_1);
      z.u = uk; z.r0 = reg[0]; z.dt = dt1; z.vis_dt_old = old; z.vis_dt_new = vis_dt;
      post_rhs(z);
      pre_rhs(uk, time);
      ZStepArgs z2 = zargs(SM_RK2_2);
      z2.u = uk; z2.r0 = reg[0]; z2.dt = dt1;
      post_rhs(z2);
    } else if (m == "RK4") {  // :602-671 (explicit diffusion, accumulator register)
      pre_rhs(uk, time);
      dt1 = adjust_dt(time);
      const int modes[4] = {SM_RK4_1, SM_RK4_2, SM_RK4_3, SM_RK4_4};
      const double ts[4] = {time, time + 0.5 * dt1, time + 0.5 * dt1, time + dt1};
      for (int s = 0; s < 4; ++s) {
        if (s > 0) pre_rhs(reg[0], ts[s]);
        ZStepArgs z = zargs(modes[s]);
        z.u = uk; z.udiff = (s == 0) ? uk : reg[0]; z.ustage = reg[0]; z.acc = reg[1]; z.dt = dt1;
        post_rhs(z);
      }
    } else {
      throw FlusiError(10001, "Error! iTimeMethodFluid unknown. Abort.");
    }
    record_ekin();
    std::swap(n0, n1);  // time_step.f90:72
  }

  void cal_nlk(double t, double* nlk, double* u) override {
    need_fft();
    if (!fields_initialized) throw FlusiError(900004, "fields not initialised (call flusi_init_fields)");
    pre_rhs(uk, t);
    ZStepArgs z = zargs(SM_RHS_ONLY);
    z.rhs_out = P; z.udiff = uk;
    post_rhs(z);
    if (nlk) copy_spectral_out(P, nlk);
    if (u) get_u(u, false);
  }

  // ------------------------------------------------------------ adjoint
  // Gradient of the mean lift J = (1/K) sum_{n=1..K} F_z(t_n) over K RK4 steps
  // (F_z = dV/eta sum chi (u_z - us_z), the force of integrals.f90:87-100 /
  // drag.f90:17-454 along z) with respect to the wing kinematics' Fourier
  // coefficients.  The forward run keeps the state every `every` steps; the
  // backward sweep recomputes one segment at a time, then each step's stage
  // states, and applies the transposed stages:
  //   u_{n+1} = u_n + dt/6 (k1 + 2 k2 + 2 k3 + k4),  k_s = N(u_s, t_s),
  //   kb4 = dt/6 ub;              W4 = N_u(u_4)^T kb4;  un = ub + W4
  //   kb3 = dt/3 ub + dt W4;      W3 = N_u(u_3)^T kb3;  un += W3
  //   kb2 = dt/3 ub + dt/2 W3;    W2 = N_u(u_2)^T kb2;  un += W2
  //   kb1 = dt/6 ub + dt/2 W2;    W1 = N_u(u_n)^T kb1;  un += W1
  // with theta-gradient terms <phi_s, dg/dtheta> at every stage and the
  // objective's own terms at every t_n.  dchi/dtheta and dus/dtheta come from
  // central differences of the mask drawing (FP64 scratch fields).
  struct AdjointState {
    int K = 0, every = 0;
    std::vector<double> theta, tn;  // controls and the step times t_0..t_K
    void* ck = nullptr;             // states at steps 0, every, 2 every, ...
    void* seg = nullptr;            // states of one segment
    void* st[3] = {nullptr, nullptr, nullptr};  // stage states u_2, u_3, u_4
    void *ub = nullptr, *un = nullptr, *kb = nullptr, *wv = nullptr, *mw = nullptr, *A = nullptr, *B = nullptr;
    void* ph = nullptr;             // physical: phi (3), a_w (3)
    double *mP = nullptr, *usP = nullptr, *mM = nullptr, *usM = nullptr;
    uint8_t* colS = nullptr;
    GridBox box{};                  // where the last pair of perturbed drawings can differ
    char* slab = nullptr;           // batched drawings: per drawing an FP64 mask, its us, a colour field
    long slab_stride = 0;           // bytes between two drawings' records
    int slab_n = 0;                 // drawings the slab holds
    double* part = nullptr;         // reduction partials (blocks x controls)
    double* gstage = nullptr;       // one call's per-control sums
    double* gacc = nullptr;         // the gradient, accumulated on the device
    int* kzero = nullptr;           // reduction kinds: all sums
    double J = 0.0;
    int nck = 0;
    bool have_forward = false;
  } adj;
  double mask_h = 1.0e-6;  // control step of the mask differences (the controls' units)
  static constexpr int kMaxControls = 64;
  // Adjoint runs redraw the whole insect at every mask update, so that the
  // mask is a function of time and controls only.  (The reference keeps the
  // body from its first drawing and clears only the previous wings' points;
  // where wing and body smoothing layers overlap that history makes the mask
  // a non-differentiable function of the kinematics.)
  bool mask_fresh = false;

  SpecGeom sgeom() const {
    return SpecGeom{nx, ny, nz, nxh, nc, kx_d, ky_d, kz_d, kxt2_d, kyt2_d, kzt2_d, p.iDealias == 1 ? 1 : 0};
  }
  size_t state_bytes() const { return 3 * sizeof(C) * nc; }
  void* state_at(void* base, int i) { return (char*)base + (size_t)i * state_bytes(); }
  void copy_state(void* dst, const void* src) {
    HIPCHECK(hipMemcpyAsync(dst, src, state_bytes(), hipMemcpyDeviceToDevice, stream));
  }
  double reduce_partials(double* part, int nb) {
    launch_reduce_rows(part, nb, 1, kinds + 4, rmid, red + 100, stream);  // red[16..77]: diagnostics
    double v = 0.0;
    HIPCHECK(hipMemcpyAsync(&v, red + 100, sizeof(double), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    return v;
  }
  void adjoint_check() {
    need_fft();
    if (!(p.iMask == "Insect" || p.iMask == "insect") || p.iMoving != 1 || p.iPenalization != 1)
      throw FlusiError(900023, "adjoint: a moving insect mask is required");
    if (p.method != "RK4") throw FlusiError(900023, "adjoint: RK4 only");
    if (p.dt_fixed <= 0.0) throw FlusiError(900023, "adjoint: dt_fixed > 0 required");
    if (sponge_code()) throw FlusiError(900023, "adjoint: no vorticity sponge");
    for (int c = 0; c < 3; ++c)
      if (p.iMeanFlow[c] != "free") throw FlusiError(900023, "adjoint: iMeanFlow must be free");
    if (p.startup_conditioner == "yes") throw FlusiError(900023, "adjoint: no startup conditioner");
  }
  int adjoint_ncontrols() override {
    if (!fields_initialized) init_fields();
    adjoint_check();
    return insect.ncontrols();
  }
  void adjoint_get_controls(double* th) override {
    const int n = adjoint_ncontrols();
    for (int i = 0; i < n; ++i) th[i] = insect.get_control(i);
  }
  void adjoint_alloc(int K, int every) {
    AdjointState& a = adj;
    const int nck = K / every + 1;
    if (!a.ub) {
      for (void** q : {&a.ub, &a.un, &a.kb, &a.wv, &a.mw, &a.A, &a.B}) *q = dalloc(state_bytes());
      for (int s = 0; s < 3; ++s) a.st[s] = dalloc(state_bytes());
      a.ph = dalloc(6 * sizeof(T) * npts);
      a.mP = (double*)dalloc(sizeof(double) * npts);
      a.mM = (double*)dalloc(sizeof(double) * npts);
      a.usP = (double*)dalloc(3 * sizeof(double) * npts);
      a.usM = (double*)dalloc(3 * sizeof(double) * npts);
      a.colS = (uint8_t*)dalloc(npts);
    }
    const int ncon = (int)a.theta.size();
    if (!a.part || a.gstage == nullptr || ncon > kMaxControls) {
      if (ncon > kMaxControls) throw FlusiError(900026, "adjoint: at most 64 controls");
      dfree(a.part); dfree(a.gstage); dfree(a.gacc); dfree(a.kzero);
      a.part = (double*)dalloc(sizeof(double) * adjoint_reduce_blocks(npts) * kMaxControls);
      a.gstage = (double*)dalloc(sizeof(double) * kMaxControls);
      a.gacc = (double*)dalloc(sizeof(double) * kMaxControls);
      a.kzero = (int*)dalloc(sizeof(int) * kMaxControls);
      launch_zero(a.kzero, sizeof(int) * kMaxControls, stream);
    }
    if (a.nck < nck || a.every != every) {
      dfree(a.ck); dfree(a.seg);
      a.ck = dalloc((size_t)nck * state_bytes());
      a.seg = dalloc((size_t)(every + 1) * state_bytes());
      a.nck = nck;
    }
    a.K = K; a.every = every;
  }
  // z-force of the current state (its physical velocity in pf(0..2)) with the current mask
  double lift_now() {
    launch_lift<T>(npts, pf(0), us, mask, adj.part, stream);
    return reduce_partials(adj.part, adjoint_reduce_blocks(npts)) * p.dx() * p.dy() * p.dz() / p.eps;
  }
  void step_plain() {  // time_step.f90:57-167 without the output
    dt0 = dt1;
    fluid_timestep();
    time = time + dt1;
    it += 1;
  }
  // J for the controls th over K steps from the initial condition; keeps the
  // checkpoints for adjoint_backward
  double adjoint_forward(const double* th, int K, int every) override {
    if (K < 1 || every < 1) throw FlusiError(900024, "adjoint: nsteps and the checkpoint interval must be >= 1");
    mask_fresh = true;
    init_fields();
    adjoint_check();
    const int ncon = insect.ncontrols();
    if (ncon == 0) throw FlusiError(900023, "adjoint: the wing kinematics must be from_file Fourier series");
    adj.theta.assign(th, th + ncon);
    for (int i = 0; i < ncon; ++i) insect.set_control(i, th[i]);
    create_mask(time);  // the initial mask with these controls
    adjoint_alloc(K, every);
    ensure_phys();
    adj.tn.assign(K + 1, 0.0);
    adj.tn[0] = time;
    copy_state(state_at(adj.ck, 0), uk);
    double J = 0.0;
    for (int n = 1; n <= K; ++n) {
      step_plain();
      adj.tn[n] = time;
      ifft_fields(uk, 3, pf(0));
      J += lift_now() / K;  // the last stage left the mask at t_n
      if (n % every == 0) copy_state(state_at(adj.ck, n / every), uk);
    }
    adj.J = J;
    adj.have_forward = true;
    return J;
  }
  // central-difference drawings of the insect at t with control i moved by
  // +-h.  The controls are wing kinematics coefficients, so the two drawings
  // differ only inside the wing boxes of the two perturbed states: both are
  // drawn over the union of those boxes (all parts are evaluated there), and
  // the theta terms sum over it (adj.box).  Periodic insects, whose parts may
  // wrap, use the whole grid.
  void perturbed_masks(double t, int i) {
    const double v = insect.get_control(i);
    InsectArgs ap[2];
    for (int k = 0; k < 2; ++k) {
      insect.set_control(i, v + (k == 0 ? 1 : -1) * mask_h);
      insect.update(t);
      ap[k] = insect_args(true, true);
    }
    insect.set_control(i, v);
    const int nn[3] = {nx, ny, nz};
    GridBox& b = adj.box;
    b.nx = nx; b.ny = ny;
    if (ap[0].use_box) {
      for (int d = 0; d < 3; ++d) { b.lo[d] = nn[d]; b.hi[d] = -1; }
      for (int k = 0; k < 2; ++k)
        for (int w = 0; w < 2; ++w) {
          if (!ap[k].draw_wing[w]) continue;
          const int* lo = ap[k].box_lo[1 + w];
          const int* hi = ap[k].box_hi[1 + w];
          bool inside = true;
          for (int d = 0; d < 3; ++d) inside = inside && lo[d] <= hi[d];
          if (!inside) continue;
          for (int d = 0; d < 3; ++d) { b.lo[d] = std::min(b.lo[d], lo[d]); b.hi[d] = std::max(b.hi[d], hi[d]); }
        }
    } else {
      for (int d = 0; d < 3; ++d) { b.lo[d] = 0; b.hi[d] = nn[d] - 1; }
    }
    const bool empty = b.lo[0] > b.hi[0] || b.lo[1] > b.hi[1] || b.lo[2] > b.hi[2];
    for (int k = 0; k < 2; ++k) {
      if (empty) break;  // no wing in the grid: both drawings agree everywhere
      InsectArgs& a = ap[k];
      for (int d = 0; d < 3; ++d) { a.reg_lo[d] = b.lo[d]; a.reg_hi[d] = b.hi[d]; }
      a.clip = 0;
      if (!ins_h) {  // the eager argument ring
        HIPCHECK(hipStreamSynchronize(stream));
        ins_per_slot = std::max(ins_per_slot, 1);
        if (!ins_d) ins_d = (InsectArgs*)dalloc(sizeof(InsectArgs) * INS_SLOTS * ins_per_slot);
        HIPCHECK(hipHostMalloc((void**)&ins_h, sizeof(InsectArgs) * INS_SLOTS * ins_per_slot, hipHostMallocDefault));
      }
      const int slot = ins_next;
      ins_next = (ins_next + 1) % INS_SLOTS;
      if (ins_ev[slot]) HIPCHECK(hipEventSynchronize(ins_ev[slot]));
      else HIPCHECK(hipEventCreateWithFlags(&ins_ev[slot], hipEventDisableTiming));
      InsectArgs* h = ins_h + (size_t)slot * ins_per_slot;
      *h = a;
      InsectArgs* d = ins_d + (size_t)slot * ins_per_slot;
      HIPCHECK(hipMemcpyAsync(d, h, sizeof(InsectArgs), hipMemcpyHostToDevice, stream));
      HIPCHECK(hipEventRecord(ins_ev[slot], stream));
      launch_insect<double>(geom, a, d, k == 0 ? adj.mP : adj.mM, k == 0 ? adj.usP : adj.usM, adj.colS, stream,
                            Batch{1, 0});
    }
    if (empty)
      for (int d = 0; d < 3; ++d) { b.lo[d] = 0; b.hi[d] = -1; }
    kcheck("perturbed masks");
  }
  // n insect argument records into the next slot of the eager ring (grown
  // to n records per slot when needed); returns their device copy
  const InsectArgs* upload_insect_args(const InsectArgs* recs, int n) {
    if (!ins_h || ins_per_slot < n) {
      HIPCHECK(hipStreamSynchronize(stream));
      dfree(ins_d);
      if (ins_h) HIPCHECK(hipHostFree(ins_h));
      ins_per_slot = std::max(ins_per_slot, n);
      ins_d = (InsectArgs*)dalloc(sizeof(InsectArgs) * INS_SLOTS * ins_per_slot);
      HIPCHECK(hipHostMalloc((void**)&ins_h, sizeof(InsectArgs) * INS_SLOTS * ins_per_slot, hipHostMallocDefault));
    }
    const int slot = ins_next;
    ins_next = (ins_next + 1) % INS_SLOTS;
    if (ins_ev[slot]) HIPCHECK(hipEventSynchronize(ins_ev[slot]));
    else HIPCHECK(hipEventCreateWithFlags(&ins_ev[slot], hipEventDisableTiming));
    InsectArgs* h = ins_h + (size_t)slot * ins_per_slot;
    std::copy(recs, recs + n, h);
    InsectArgs* d = ins_d + (size_t)slot * ins_per_slot;
    HIPCHECK(hipMemcpyAsync(d, h, sizeof(InsectArgs) * n, hipMemcpyHostToDevice, stream));
    HIPCHECK(hipEventRecord(ins_ev[slot], stream));
    return d;
  }

  // All 2 ncon perturbed drawings of one theta_terms call in one batched
  // launch (member 2i: control i at +h, 2i+1: at -h), over the union of their
  // wing boxes; each drawing has its own record in adj.slab.  Used while the
  // slab fits kSlabMax bytes.
  static constexpr size_t kSlabMax = size_t(4) << 30;
  bool batched_drawings(int ncon) {
    const long stride = ((33L * npts + 255) / 256) * 256;  // mask 8 B, us 24 B, colour 1 B per point
    if ((size_t)stride * 2 * ncon > kSlabMax || (long)nz * 2 * ncon > 65535) return false;
    if (adj.slab_n < 2 * ncon || adj.slab_stride != stride) {
      HIPCHECK(hipStreamSynchronize(stream));
      dfree(adj.slab);
      adj.slab = (char*)dalloc((size_t)stride * 2 * ncon);
      adj.slab_stride = stride;
      adj.slab_n = 2 * ncon;
    }
    return true;
  }
  void perturbed_masks_all(double t, int ncon) {
    std::vector<InsectArgs> recs(2 * ncon);
    for (int i = 0; i < ncon; ++i) {
      const double v = insect.get_control(i);
      for (int k = 0; k < 2; ++k) {
        insect.set_control(i, v + (k == 0 ? 1 : -1) * mask_h);
        insect.update(t);
        recs[2 * i + k] = insect_args(true, true);
      }
      insect.set_control(i, v);
    }
    const int nn[3] = {nx, ny, nz};
    GridBox& b = adj.box;
    b.nx = nx; b.ny = ny;
    if (recs[0].use_box) {
      for (int d = 0; d < 3; ++d) { b.lo[d] = nn[d]; b.hi[d] = -1; }
      for (const InsectArgs& a : recs)
        for (int w = 0; w < 2; ++w) {
          if (!a.draw_wing[w]) continue;
          bool inside = true;
          for (int d = 0; d < 3; ++d) inside = inside && a.box_lo[1 + w][d] <= a.box_hi[1 + w][d];
          if (!inside) continue;
          for (int d = 0; d < 3; ++d) {
            b.lo[d] = std::min(b.lo[d], a.box_lo[1 + w][d]);
            b.hi[d] = std::max(b.hi[d], a.box_hi[1 + w][d]);
          }
        }
    } else {
      for (int d = 0; d < 3; ++d) { b.lo[d] = 0; b.hi[d] = nn[d] - 1; }
    }
    if (b.lo[0] > b.hi[0] || b.lo[1] > b.hi[1] || b.lo[2] > b.hi[2]) {
      for (int d = 0; d < 3; ++d) { b.lo[d] = 0; b.hi[d] = -1; }
      return;  // no wing in the grid: every pair of drawings agrees everywhere
    }
    for (InsectArgs& a : recs) {
      for (int d = 0; d < 3; ++d) { a.reg_lo[d] = b.lo[d]; a.reg_hi[d] = b.hi[d]; }
      a.clip = 0;
    }
    const InsectArgs* dev = upload_insect_args(recs.data(), 2 * ncon);
    launch_insect<double>(geom, recs[0], dev, adj.slab, adj.slab + 8 * npts, (uint8_t*)(adj.slab + 32 * npts), stream,
                          Batch{2 * ncon, adj.slab_stride});
    kcheck("perturbed masks (batched)");
  }

  // gacc_i += scale * sum f . (dchi_i (u - us) - chi dus_i) / (2h) at time t,
  // with f = phi or e_z (phi null), u in pf(0..2), mask and us at t; all on
  // the device
  void theta_terms(double t, const void* phi, double scale) {
    const int ncon = (int)adj.theta.size();
    const int nb = adjoint_reduce_blocks(npts);
    if (batched_drawings(ncon)) {
      perturbed_masks_all(t, ncon);
      launch_theta_batch<T>(npts, adj.box, phi, pf(0), us, mask, adj.slab, adj.slab_stride, ncon, adj.part, stream);
    } else {
      for (int i = 0; i < ncon; ++i) {
        perturbed_masks(t, i);
        launch_theta<T>(npts, adj.box, phi, pf(0), us, mask, adj.mP, adj.usP, adj.mM, adj.usM, adj.part + i, ncon,
                        stream);
      }
    }
    launch_reduce_rows(adj.part, nb, ncon, adj.kzero, rmid, adj.gstage, stream);
    launch_axpby<double>(ncon, 1.0, adj.gacc, scale / (2.0 * mask_h), adj.gstage, adj.gacc, stream);
    kcheck("theta terms");
    insect.update(t);
  }
  // W = N_u(u_s)^T lam at time t_s; grad += <phi, dg/dtheta>
  void adjoint_rhs(const void* u_s, double t_s, const void* lam, void* W) {
    const SpecGeom g = sgeom();
    const double N = double(nx) * double(ny) * double(nz);
    create_mask(t_s);
    inv_uw_phys(u_s);  // pf(0..2) u, pf(3..5) w
    launch_adj_pre<T>(g, p.nu, 1.0 / N, lam, W, adj.mw, stream);
    void* phi = adj.ph;
    void* aw = (char*)adj.ph + 3 * sizeof(T) * npts;
    ifft_fields(adj.mw, 3, phi);
    // g = u x w - chi/eta (u - us): dg/dtheta = -(1/eta)(dchi (u - us) - chi dus)
    theta_terms(t_s, phi, -1.0 / p.eps);
    launch_adj_point<T>(npts, 1.0 / p.eps, phi, pf(0), pf(3), mask, aw, stream);
    fft_fields(phi, 3, adj.A);
    fft_fields(aw, 3, adj.B);
    launch_adj_post<T>(g, N, adj.A, adj.B, W, stream);
    kcheck("adjoint rhs");
  }
  // ub += dJ/du at step n (its state u_n), grad += dJ/dtheta there
  void objective_terms(int n, const void* u_n) {
    const double dV = p.dx() * p.dy() * p.dz();
    const double N = double(nx) * double(ny) * double(nz);
    const double s = dV / (p.eps * adj.K);
    create_mask(adj.tn[n]);
    ifft_fields(u_n, 3, pf(0));
    theta_terms(adj.tn[n], nullptr, s);
    void* src = adj.ph;
    launch_ez_mask<T>(npts, s, mask, src, stream);
    fft_fields(src, 3, adj.A);
    launch_adj_post<T>(sgeom(), N, adj.A, nullptr, adj.ub, stream);
    kcheck("objective terms");
  }
  // stage states u_2, u_3, u_4 of the step from u_n at t_n
  void stage_states(const void* u_n, double t_n) {
    copy_state(uk, u_n);
    const double dt = p.dt_fixed;
    const int modes[3] = {SM_RK4_1, SM_RK4_2, SM_RK4_3};
    const double ts[3] = {t_n, t_n + 0.5 * dt, t_n + 0.5 * dt};
    for (int s = 0; s < 3; ++s) {
      pre_rhs(s == 0 ? uk : reg[0], ts[s]);
      ZStepArgs z = zargs(modes[s]);
      z.u = uk; z.udiff = (s == 0) ? uk : reg[0]; z.ustage = reg[0]; z.acc = reg[1]; z.dt = dt;
      post_rhs(z);
      copy_state(adj.st[s], reg[0]);
    }
  }
  void adjoint_backward(double* grad_out) override {
    if (!adj.have_forward) throw FlusiError(900025, "adjoint: run adjoint_forward first");
    AdjointState& a = adj;
    const int K = a.K, E = a.every, ncon = (int)a.theta.size();
    const double dt = p.dt_fixed;
    const size_t ns = 3 * nc * 2;  // reals of a state
    launch_zero(a.ub, state_bytes(), stream);
    launch_zero(a.gacc, sizeof(double) * kMaxControls, stream);
    const int nseg = (K + E - 1) / E;
    for (int sg = nseg - 1; sg >= 0; --sg) {
      const int n0 = sg * E, n1 = std::min(K, n0 + E);
      // the segment's states u_{n0} .. u_{n1}
      copy_state(uk, state_at(a.ck, sg));
      time = a.tn[n0]; it = n0; dt1 = dt;
      copy_state(state_at(a.seg, 0), uk);
      for (int n = n0; n < n1; ++n) {
        step_plain();
        copy_state(state_at(a.seg, n + 1 - n0), uk);
      }
      for (int n = n1 - 1; n >= n0; --n) {
        objective_terms(n + 1, state_at(a.seg, n + 1 - n0));
        const void* u_n = state_at(a.seg, n - n0);
        stage_states(u_n, a.tn[n]);
        const double t = a.tn[n];
        // stage 4 (u_4 at t + dt)
        launch_axpby<T>(ns, dt / 6.0, a.ub, 0.0, nullptr, a.kb, stream);
        adjoint_rhs(a.st[2], t + dt, a.kb, a.wv);
        launch_axpby<T>(ns, 1.0, a.ub, 1.0, a.wv, a.un, stream);
        // stage 3 (u_3 at t + dt/2)
        launch_axpby<T>(ns, dt / 3.0, a.ub, dt, a.wv, a.kb, stream);
        adjoint_rhs(a.st[1], t + 0.5 * dt, a.kb, a.wv);
        launch_axpby<T>(ns, 1.0, a.un, 1.0, a.wv, a.un, stream);
        // stage 2 (u_2 at t + dt/2)
        launch_axpby<T>(ns, dt / 3.0, a.ub, 0.5 * dt, a.wv, a.kb, stream);
        adjoint_rhs(a.st[0], t + 0.5 * dt, a.kb, a.wv);
        launch_axpby<T>(ns, 1.0, a.un, 1.0, a.wv, a.un, stream);
        // stage 1 (u_n at t)
        launch_axpby<T>(ns, dt / 6.0, a.ub, 0.5 * dt, a.wv, a.kb, stream);
        adjoint_rhs(u_n, t, a.kb, a.wv);
        launch_axpby<T>(ns, 1.0, a.un, 1.0, a.wv, a.ub, stream);
      }
    }
    HIPCHECK(hipMemcpyAsync(grad_out, a.gacc, sizeof(double) * ncon, hipMemcpyDeviceToHost, stream));
    // the solver is left at the initial state
    copy_state(uk, state_at(a.ck, 0));
    time = a.tn[0]; it = 0;
    create_mask(time);
    HIPCHECK(hipStreamSynchronize(stream));
  }

  // ------------------------------------------------------------ masks
  void create_mask(double t) override {
    if (p.iMask == "Insect" && (p.iMoving != 1 || p.iPenalization != 1))
      throw FlusiError(4453, "insects.f90::DrawInsect: the parameters iMoving or iPenalization are wrong.");
    const int M = batch_n(bat);
    auto member = [&](int m) -> Solver<T>& { return M > 1 ? *batch[m] : *this; };
    // the mask changes below; the line flags are brought up to date only by
    // the unbatched, uncaptured insect update
    const bool flags_were_valid = mflag_valid;
    mflag_valid = false;
    timed(TM_MASK, [&] {
      if (p.iPenalization == 1) {
        if (p.iMask == "sphere" || p.iMask == "Sphere") {
          for (int m = 0; m < M; ++m) {  // each member's own sphere
            const Params& q = member(m).p;
            SphereArgs a{};
            a.x0 = q.x0; a.y0 = q.y0; a.z0 = q.z0; a.radius = 0.5 * q.length;
            const double h = std::max(q.dx(), std::max(q.dy(), q.dz()));
            a.smoothing = 1.5 * h; a.safety = 2.0 * 1.5 * h; a.color_thr = (double)1.0e-12f;
            const long o = m * bat.mstride;
            launch_sphere<T>(geom, a, mshift((char*)mask, o), mshift((char*)us, o), mshift(color, o), stream);
          }
        } else if (p.iMask == "Insect" || p.iMask == "insect") {
          // one argument record per member, all over the same region (the
          // whole grid when batched or captured)
          const bool full = M > 1 || capture;
          std::vector<InsectArgs> recs(M);
          for (int m = 0; m < M; ++m) {
            member(m).insect.update(t);
            recs[m] = member(m).insect_args(full);
          }
          const InsectArgs* dev;
          if (capture) {
            if (stage_slot >= INS_SLOTS) throw FlusiError(900012, "too many mask updates in a captured step");
            if (!ins_cap) ins_cap = (InsectArgs*)dalloc(sizeof(InsectArgs) * INS_SLOTS * M);
            dev = ins_cap + (size_t)stage_slot++ * M;
          } else {
            if (ins_per_slot < M) {  // (re)size the ring for M records per slot
              HIPCHECK(hipStreamSynchronize(stream));
              dfree(ins_d);
              if (ins_h) HIPCHECK(hipHostFree(ins_h));
              ins_h = nullptr;
              ins_per_slot = M;
              ins_d = (InsectArgs*)dalloc(sizeof(InsectArgs) * INS_SLOTS * M);
              HIPCHECK(hipHostMalloc((void**)&ins_h, sizeof(InsectArgs) * INS_SLOTS * M, hipHostMallocDefault));
            }
            const int slot = ins_next;
            ins_next = (ins_next + 1) % INS_SLOTS;
            if (ins_ev[slot]) HIPCHECK(hipEventSynchronize(ins_ev[slot]));
            else HIPCHECK(hipEventCreateWithFlags(&ins_ev[slot], hipEventDisableTiming));
            InsectArgs* h = ins_h + (size_t)slot * ins_per_slot;
            std::copy(recs.begin(), recs.end(), h);
            InsectArgs* d = ins_d + (size_t)slot * ins_per_slot;
            HIPCHECK(hipMemcpyAsync(d, h, sizeof(InsectArgs) * M, hipMemcpyHostToDevice, stream));
            HIPCHECK(hipEventRecord(ins_ev[slot], stream));
            dev = d;
          }
          launch_insect<T>(geom, recs[0], dev, mask, us, color, stream, bat);
          if (M == 1 && !capture && mflag_on) update_mask_flags(recs[0], flags_were_valid);
        } else if (p.iMask == "none" || p.iMask == "empty" || p.iMask == "no") {
        } else {
          throw FlusiError(3333, "create_mask_fsi(): unkown mask function iMask");
        }
      }
      if (p.iCavity != "no" && p.iPenalization == 1) {
        mflag_valid = false;
        CavityArgs c{};
        c.size = p.cavity_size;
        if (p.iCavity == "yes") { c.ux = c.uy = c.uz = 0.0; }
        else if (p.iCavity == "freestream") { c.ux = p.umean[0]; c.uy = p.umean[1]; c.uz = p.umean[2]; }
        else throw FlusiError(3334, "iCavity not supported: " + p.iCavity);
        for (int m = 0; m < M; ++m) {
          const long o = m * bat.mstride;
          launch_cavity<T>(geom, c, mshift((char*)mask, o), mshift((char*)us, o), mshift(color, o), stream);
        }
      }
    });
    kcheck("create_mask");
  }

  // Line flags after an insect update: insect_k writes only inside its region
  // box, so lines outside it keep their flags; the lines crossing the box
  // (or every line, when the flags were not valid before) are recomputed
  // from the mask.
  void update_mask_flags(const InsectArgs& r, bool were_valid) {
    if (!mflag) {
      mflag = (uint8_t*)dalloc((size_t)ny * nz);
      mzero = dalloc(sizeof(T) * nx);
      launch_zero(mzero, sizeof(T) * nx, stream);
    }
    const bool empty = r.reg_lo[0] > r.reg_hi[0] || r.reg_lo[1] > r.reg_hi[1] || r.reg_lo[2] > r.reg_hi[2];
    if (!were_valid) launch_mask_flags<T>(mask, mflag, nx, ny, 0, ny - 1, 0, nz - 1, stream);
    else if (!empty) launch_mask_flags<T>(mask, mflag, nx, ny, r.reg_lo[1], r.reg_hi[1], r.reg_lo[2], r.reg_hi[2], stream);
    mflag_valid = true;
  }

  // full: visit the whole grid (a batched or captured launch has one fixed
  // region for every member and stage).  scratch: a complete drawing (body
  // and wings, everything reset) into other buffers, leaving the record of
  // what the solver's own mask holds untouched.
  InsectArgs insect_args(bool full = false, bool scratch = false) {
    Insect& I = insect;
    InsectArgs a{};
    const double h = std::max(p.dx(), std::max(p.dy(), p.dz()));
    I.smooth = 1.0 * h;
    I.safety = 3.5 * I.smooth;
    // delete_old_mask: a resting body keeps its mask after the first call
    a.full_reset = I.cleaned_already_once ? 0 : 1;
    if (!scratch) I.cleaned_already_once = true;
    a.draw_body = 0;
    if (scratch || (mask_fresh && I.body_already_drawn)) {
      if (!I.body_already_drawn) throw FlusiError(900022, "scratch insect drawing before the first mask");
      a.full_reset = 1;
      a.draw_body = (I.BodyType != "nobody") ? 1 : 0;
    } else if (!I.body_already_drawn) {
      a.draw_body = (I.BodyType != "nobody") ? 1 : 0;
      I.body_already_drawn = true;
      if (I.BodyType == "jerry" || I.BodyType == "Jerry") {
        // body_geometry.f90:535-545 hard-coded Jerry geometry
        I.R_head = 0.125; I.R_eye = 0.0625;
        I.x_pivot_r_b = {0.05, -0.2165, 0.0};
        I.x_pivot_l_b = {0.05, +0.2165, 0.0};
        I.b_body = 0.1; I.L_body = 1.0;
        I.x_head = {0.5 * I.L_body, 0.0, 0.0};
        const double se = std::sin(45.0 * PI / 180.0) * I.R_head * 0.8;
        I.x_eye_r = {I.x_head[0] + se, I.x_head[1] + se, I.x_head[2] + se};
        I.x_eye_l = {I.x_head[0] + se, I.x_head[1] - se, I.x_head[2] + se};
      }
    }
    const InsectMotion& s = I.st;
    a.body_type = (I.BodyType == "jerry" || I.BodyType == "Jerry") ? 1 : (I.BodyType == "bumblebee" ? 3 : 2);
    for (int i = 0; i < 3; ++i) a.xc[i] = s.xc[i];
    for (int i = 0; i < 9; ++i) a.Mb[i] = s.Mb[i];
    a.b_body = I.b_body; a.L_body = I.L_body; a.smooth = I.smooth; a.safety = I.safety;
    a.nspheres = 0;
    if (a.draw_body && a.body_type == 1) {
      const Vec3 parts[3] = {I.x_head, I.x_eye_l, I.x_eye_r};
      const double rr[3] = {I.R_head, I.R_eye, I.R_eye};
      const double box[3] = {p.xl, p.yl, p.zl};
      const double ddx[3] = {p.dx(), p.dy(), p.dz()};
      const int nn[3] = {nx, ny, nz};
      const double mind = std::min(ddx[0], std::min(ddx[1], ddx[2]));
      for (int k = 0; k < 3; ++k) {
        Vec3 c = mat_vec(transpose(s.Mb), parts[k]);
        for (int i = 0; i < 3; ++i) c[i] += s.xc[i];
        for (int i = 0; i < 3; ++i) if (c[i] < 0.0) c[i] += box[i];
        for (int i = 0; i < 3; ++i) if (c[i] >= box[i]) c[i] -= box[i];
        const int ns = fortran_nint((rr[k] + I.safety) / mind);
        for (int i = 0; i < 3; ++i) {
          a.sph_c[k][i] = c[i];
          const int ci = fortran_nint(c[i] / ddx[i]);
          a.sph_lo[k][i] = std::max(ci - ns, 0);
          a.sph_hi[k][i] = std::min(ci + ns, nn[i] - 1);
        }
        a.sph_r[k] = rr[k];
      }
      a.nspheres = 3;
    }
    if (a.draw_body && a.body_type == 3 && !((scratch || mask_fresh) && caps_d) &&
        (I.HasDetails == "all" || I.HasDetails == "legs" || I.HasDetails == "antennae_proboscis")) {
      // legs, antennae, proboscis (body_geometry.f90:276-345): capsules in the
      // global frame with their end spheres' wrapped centres
      const std::vector<Capsule> cl = bumblebee_capsules();
      std::vector<double> rec(cl.size() * CAPSULE_W, 0.0);
      const double box[3] = {p.xl, p.yl, p.zl};
      const Mat3 MbT = transpose(s.Mb);
      for (size_t k = 0; k < cl.size(); ++k) {
        double* r = rec.data() + k * CAPSULE_W;
        Vec3 A = mat_vec(MbT, cl[k].a), B = mat_vec(MbT, cl[k].b);
        for (int i = 0; i < 3; ++i) { A[i] += s.xc[i]; B[i] += s.xc[i]; }
        for (int i = 0; i < 3; ++i) { r[CAP_A + i] = A[i]; r[CAP_B + i] = B[i]; }
        r[CAP_R] = cl[k].r;
        for (int e = 0; e < 2; ++e) {  // drawsphere's wrap of the centre (body_geometry.f90:1674-1681)
          Vec3 c = e ? B : A;
          for (int i = 0; i < 3; ++i) if (c[i] < 0.0) c[i] += box[i];
          for (int i = 0; i < 3; ++i) if (c[i] >= box[i]) c[i] -= box[i];
          for (int i = 0; i < 3; ++i) r[(e ? CAP_SB : CAP_SA) + i] = c[i];
        }
      }
      if (!caps_d) caps_d = (double*)dalloc(sizeof(double) * rec.size());
      HIPCHECK(hipMemcpyAsync(caps_d, rec.data(), sizeof(double) * rec.size(), hipMemcpyHostToDevice, stream));
      HIPCHECK(hipStreamSynchronize(stream));
      a.ncaps = (int)cl.size();
      a.caps = caps_d;
    } else if (a.draw_body && a.body_type == 3 && caps_d) {  // a tethered body's capsules do not move
      a.ncaps = (int)bumblebee_capsules().size();
      a.caps = caps_d;
    }
    if (I.WingShape == "bumblebee") {
      if (!fw_d) {
        fw_d = (double*)dalloc(sizeof(double) * I.fwing.table.size());
        HIPCHECK(hipMemcpy(fw_d, I.fwing.table.data(), sizeof(double) * I.fwing.table.size(), hipMemcpyHostToDevice));
      }
      a.fw_table = fw_d;
      a.fw_xc = I.fwing.xc; a.fw_yc = I.fwing.yc; a.fw_dphi = I.fwing.dphi;
      for (int i = 0; i < 6; ++i) a.fw_bbox[i] = I.fwing.bbox[i];
    }
    a.wing_shape = (I.WingShape == "TwoEllipses") ? 1 : (I.WingShape == "bumblebee" ? 3 : 2);
    a.smooth_mode = mask_fresh ? 1 : 0;
    a.L_span = I.L_span; a.b_top = I.b_top; a.b_bot = I.b_bot; a.thickness = I.WingThickness;
    a.draw_wing[0] = I.RightWing == "yes"; a.draw_wing[1] = I.LeftWing == "yes";
    for (int i = 0; i < 9; ++i) { a.Mw[0][i] = s.Mwr[i]; a.Mw[1][i] = s.Mwl[i]; }
    for (int i = 0; i < 3; ++i) {
      a.piv[0][i] = I.x_pivot_r_b[i]; a.piv[1][i] = I.x_pivot_l_b[i];
      a.rot[0][i] = s.rot_rel_r_w[i]; a.rot[1][i] = s.rot_rel_l_w[i];
    }
    a.color_w[0] = 3; a.color_w[1] = 2;
    const Vec3 vb = mat_vec(s.Mb, s.vc);
    for (int i = 0; i < 3; ++i) { a.vb[i] = vb[i]; a.rot_body_b[i] = s.rot_body_b[i]; }
    // index boxes enclosing every point a part's geometric tests can accept
    // (corners of its local bounding box mapped to the global frame, one cell
    // of margin); unused with periodic insects, whose parts may wrap
    a.use_box = p.periodic ? 0 : 1;
    const Mat3 MbT = transpose(s.Mb);
    auto set_box = [&](int k, const Vec3& lo, const Vec3& hi, const Mat3* Mw, const Vec3& piv) {
      double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
      for (int c = 0; c < 8; ++c) {
        Vec3 v{(c & 1) ? hi[0] : lo[0], (c & 2) ? hi[1] : lo[1], (c & 4) ? hi[2] : lo[2]};
        if (Mw) { v = mat_vec(transpose(*Mw), v); for (int i = 0; i < 3; ++i) v[i] += piv[i]; }
        v = mat_vec(MbT, v);
        for (int i = 0; i < 3; ++i) { mn[i] = std::min(mn[i], v[i] + s.xc[i]); mx[i] = std::max(mx[i], v[i] + s.xc[i]); }
      }
      const double d[3] = {p.dx(), p.dy(), p.dz()};
      const int nn[3] = {nx, ny, nz};
      for (int i = 0; i < 3; ++i) {
        a.box_lo[k][i] = std::max(0, (int)std::floor(mn[i] / d[i]) - 1);
        a.box_hi[k][i] = std::min(nn[i] - 1, (int)std::ceil(mx[i] / d[i]) + 1);
      }
    };
    const double sf = I.safety;
    if (a.body_type == 3)  // bumblebee body, head and appendages lie well inside this body-frame box
      set_box(0, {-1.5 - sf, -0.8 - sf, -1.0 - sf}, {1.5 + sf, 0.8 + sf, 0.8 + sf}, nullptr, {0, 0, 0});
    else
      set_box(0, {-I.L_body / 2 - sf, -I.b_body - sf, -I.b_body - sf}, {I.L_body / 2 + sf, I.b_body + sf, I.b_body + sf},
              nullptr, {0, 0, 0});
    Vec3 wlo{-I.b_bot - sf, -sf, -0.5 * I.WingThickness - sf}, whi{I.b_top + sf, I.L_span + sf, 0.5 * I.WingThickness + sf};
    if (a.wing_shape == 3) {
      const double* bb = I.fwing.bbox;
      wlo = Vec3{bb[0] - sf, bb[2] - sf, bb[4] - sf};
      whi = Vec3{bb[1] + sf, bb[3] + sf, bb[5] + sf};
    }
    set_box(1, wlo, whi, &s.Mwr, I.x_pivot_r_b);
    set_box(2, wlo, whi, &s.Mwl, I.x_pivot_l_b);

    // region visited: the parts drawn now and the wings of the previous call
    // (the only places where wing colours can be left); everything on the
    // first call and for periodic insects
    // (a batched or captured launch visits the whole grid and skips the
    // points outside that region, a.clip)
    const int nn[3] = {nx, ny, nz};
    a.clip = 0;
    if (!a.use_box || a.full_reset || scratch) {
      for (int i = 0; i < 3; ++i) { a.reg_lo[i] = 0; a.reg_hi[i] = nn[i] - 1; }
    } else {
      for (int i = 0; i < 3; ++i) { a.reg_lo[i] = nn[i]; a.reg_hi[i] = -1; }
      auto include = [&](const int* lo, const int* hi) {
        for (int i = 0; i < 3; ++i) if (lo[i] > hi[i]) return;  // part outside the grid
        for (int i = 0; i < 3; ++i) { a.reg_lo[i] = std::min(a.reg_lo[i], lo[i]); a.reg_hi[i] = std::max(a.reg_hi[i], hi[i]); }
      };
      if (a.draw_body) {
        include(a.box_lo[0], a.box_hi[0]);
        for (int k = 0; k < a.nspheres; ++k) include(a.sph_lo[k], a.sph_hi[k]);
      }
      for (int w = 0; w < 2; ++w) {
        if (a.draw_wing[w]) include(a.box_lo[1 + w], a.box_hi[1 + w]);
        if (prev_wing_valid[w]) include(prev_wing_lo[w], prev_wing_hi[w]);
      }
      if (full) {
        a.clip = 1;
        for (int i = 0; i < 3; ++i) {
          a.clip_lo[i] = a.reg_lo[i]; a.clip_hi[i] = a.reg_hi[i];  // empty when lo > hi: every point returns
          a.reg_lo[i] = 0; a.reg_hi[i] = nn[i] - 1;
        }
      }
    }
    if (!scratch)
      for (int w = 0; w < 2; ++w) {
        prev_wing_valid[w] = a.draw_wing[w] != 0;
        for (int i = 0; i < 3; ++i) { prev_wing_lo[w][i] = a.box_lo[1 + w][i]; prev_wing_hi[w][i] = a.box_hi[1 + w][i]; }
      }
    return a;
  }
  // index boxes of the wings drawn by the previous insect mask update
  bool prev_wing_valid[2] = {false, false};
  int prev_wing_lo[2][3] = {}, prev_wing_hi[2][3] = {};

  void get_mask(double* m, double* u_s, int16_t* col) override {
    HIPCHECK(hipStreamSynchronize(stream));
    if (m) copy_real_out(mask, m, npts);
    if (u_s) copy_real_out(us, u_s, 3 * npts);
    if (col) {
      std::vector<uint8_t> c(npts);
      HIPCHECK(hipMemcpy(c.data(), color, npts, hipMemcpyDeviceToHost));
      for (long i = 0; i < npts; ++i) col[i] = c[i];
    }
  }

  // ------------------------------------------------------------ host copies
  double* staging(size_t n) {
    if (n > stage_n) {
      dfree(stage);
      stage = (double*)dalloc(n * sizeof(double));
      stage_n = n;
    }
    return stage;
  }
  void copy_real_out(const void* src, double* dst, long n) {
    if (FP64) { HIPCHECK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, stream)); }
    else {
      double* s = staging(n);
      launch_to_f64<T>(src, s, n, stream);
      HIPCHECK(hipMemcpyAsync(dst, s, n * sizeof(double), hipMemcpyDeviceToHost, stream));
    }
    HIPCHECK(hipStreamSynchronize(stream));
  }
  void copy_real_in(const double* src, void* dst, long n) {
    if (FP64) { HIPCHECK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, stream)); }
    else {
      double* s = staging(n);
      HIPCHECK(hipMemcpyAsync(s, src, n * sizeof(double), hipMemcpyHostToDevice, stream));
      launch_from_f64<T>(s, dst, n, stream);
    }
    HIPCHECK(hipStreamSynchronize(stream));
  }
  void copy_spectral_out(const void* src, double* dst) { copy_real_out(src, dst, 6 * nc); }

  void set_uk(const double* h) override {
    need_fft();
 copy_real_in(h, uk, 6 * nc); fields_initialized = true; }
  void get_uk(double* h) override {
    need_fft();
 copy_real_out(uk, h, 6 * nc); }

  void ensure_phys() { if (!phys) phys = dalloc(7 * sizeof(T) * npts); }
  T* pf(int f) { return reinterpret_cast<T*>(phys) + (long)f * npts; }

  void inv_uw_phys(const void* u_in) {
    inv_uw(u_in);
    ensure_phys();
    XPassArgs a = xargs();
    a.in = P; a.out = phys; a.fstride_in = pc; a.fstride_out = npts; a.nf = 6;
    timed(TM_X, [&] { launch_x_pass<T>(a, XM_INV_UW, stream); });
    kcheck("x inverse uw");
  }
  // 1 or 3 spectral S-layout fields -> physical
  void ifft_fields(const void* src, int nf, void* dst) {
    const C* s = reinterpret_cast<const C*>(src);
    TileJob zj[3], yj[3];
    for (int f = 0; f < nf; ++f) {
      zj[f] = TileJob{s + (long)f * nc, nullptr, cf(Q, f), OP_COPY, nullptr, 0};
      yj[f] = TileJob{cf(Q, f), nullptr, cfp(f), OP_COPY, nullptr, 0};
    }
    zinv(zj, nf);
    yinv(yj, nf);
    XPassArgs a = xargs();
    a.in = P; a.out = dst; a.fstride_in = pc; a.fstride_out = npts; a.nf = nf;
    timed(TM_X, [&] { launch_x_pass<T>(a, XM_INV, stream); });
    kcheck("x inverse");
  }
  // 1 or 3 physical fields -> S-layout spectral (times 1/N)
  void fft_fields(const void* src, int nf, void* dst) {
    XPassArgs a = xargs();
    a.in = src; a.out = P; a.fstride_in = npts; a.fstride_out = pc; a.nf = nf;
    timed(TM_X, [&] { launch_x_pass<T>(a, XM_FWD, stream); });
    kcheck("x forward");
    yfwd(nf);
    zfwd_plain(nf, dst);
  }

  // save_fields.f90:98-133: the mask at t, the unprojected right-hand side,
  // pressure (rhs.f90:342-379), ifft, minus |u|^2/2
  void get_pressure(double t, double* out) override {
    need_fft();
    if (!fields_initialized) throw FlusiError(900004, "fields not initialised (call flusi_init_fields)");
    pre_rhs(uk, t);
    ZStepArgs z = zargs(SM_PRESSURE);
    z.rhs_out = P; z.udiff = uk;
    post_rhs(z);
    ensure_phys();
    ifft_fields(P, 1, pf(6));
    ifft_fields(uk, 3, pf(0));
    launch_pressure_phys<T>(npts, pf(0), pf(6), stream);
    kcheck("pressure");
    copy_real_out(pf(6), out, npts);
  }

  // ------------------------------------------------------------ HDF5 outputs
  // flusi_hdf5_interface.f90:21-81 save_field_hdf5: one field per file
  // "<name>.h5", dataset = the name up to its first '_', with the
  // attributes time, viscosity, epsi, domain_size, origin and nxyz
  // flusi_hdf5_interface.f90:21-81 save_field_hdf5; with striding > 1 the
  // field at ix, iy, iz = 0, s, 2s, ... (every index below n) and nxyz =
  // (nx, ny, nz) / s (integer division), as save_field_hdf5_strided (:95-169)
  void save_field_h5(const std::string& name, const double* f) {
    const std::string file = out_path(name + ".h5");
    const std::string dset = name.substr(0, name.find('_'));
    const int st = p.striding;
    if (st < 1) throw FlusiError(8881, "Striding value is bad, exit!");
    std::vector<double> red;
    int mx = nx, my = ny, mz = nz;
    if (st > 1) {
      mx = (nx - 1) / st + 1; my = (ny - 1) / st + 1; mz = (nz - 1) / st + 1;
      red.resize((size_t)mx * my * mz);
      for (int iz = 0; iz < mz; ++iz)
        for (int iy = 0; iy < my; ++iy)
          for (int ix = 0; ix < mx; ++ix)
            red[((size_t)iz * my + iy) * mx + ix] = f[((size_t)(st * iz) * ny + st * iy) * nx + st * ix];
      f = red.data();
    }
    h5_write(file, dset, mx, my, mz, f, p.field_precision != "double", true,
             {{"time", {time}, 0}, {"viscosity", {p.nu}, 0}, {"epsi", {p.eps}, 0},
              {"domain_size", {p.xl, p.yl, p.zl}, 0}, {"origin", {0.0, 0.0, 0.0}, 0},
              {"nxyz", {(double)(nx / st), (double)(ny / st), (double)(nz / st)}, 1}});
  }

  // save_fields.f90:46-237 save_fields_fsi at the current time
  void save_fields() override {
    need_fft();
    if (!fields_initialized) throw FlusiError(900004, "fields not initialised (call flusi_init_fields)");
    if (p.striding < 1) throw FlusiError(8881, "Striding value is bad, exit!");
    char name[32];
    if (p.naming == "timestep") std::snprintf(name, sizeof(name), "%06d", it);
    else if (p.save_only_one_period == "yes")  // save_fields.f90:86 (t minus the period count)
      std::snprintf(name, sizeof(name), "%06d", fortran_nint((time - std::floor(time / p.tsave_period)) * 1000.0));
    else std::snprintf(name, sizeof(name), "%06d", fortran_nint(time * 1000.0));
    const std::string tag(name);
    if (p.iMoving == 1) create_mask(time);
    std::vector<double> a(3 * (size_t)npts), b;
    if (p.iSaveVelocity == 1) {
      get_u(a.data(), false);
      save_field_h5("ux_" + tag, a.data());
      save_field_h5("uy_" + tag, a.data() + npts);
      save_field_h5("uz_" + tag, a.data() + 2 * npts);
    }
    if (p.iSavePress == 1) {
      // rhs.f90:342-379 and save_fields.f90:127-132: the pressure of the
      // unprojected right-hand side at t, minus |u|^2/2
      b.resize(npts);
      get_pressure(time, b.data());
      save_field_h5("p_" + tag, b.data());
    }
    if (p.iSaveVorticity == 1 || p.iSaveMagVorticity == 1) {
      get_u(a.data(), true);
      if (p.iSaveVorticity == 1) {
        save_field_h5("vorx_" + tag, a.data());
        save_field_h5("vory_" + tag, a.data() + npts);
        save_field_h5("vorz_" + tag, a.data() + 2 * npts);
      }
      if (p.iSaveMagVorticity == 1) {
        b.resize(npts);
        launch_magnitude<T>(npts, pf(3), pf(6), stream);  // get_u left the vorticity in pf(3..5)
        kcheck("vorabs");
        copy_real_out(pf(6), b.data(), npts);
        save_field_h5("vorabs_" + tag, b.data());
      }
    }
    const bool smask = p.iSaveMask == 1 && p.iPenalization == 1;
    const bool sus = p.iSaveSolidVelocity == 1 && p.iPenalization == 1 && p.iMoving == 1;
    if (smask || sus) {
      b.resize(npts);
      std::vector<int16_t> col(npts);
      get_mask(b.data(), a.data(), col.data());
      if (smask) save_field_h5("mask_" + tag, b.data());
      if (sus) {
        save_field_h5("usx_" + tag, a.data());
        save_field_h5("usy_" + tag, a.data() + npts);
        save_field_h5("usz_" + tag, a.data() + 2 * npts);
      }
    }
  }

  // runtime_backuping.f90:3-193 (fluid part): u and the registers nlk(0),
  // nlk(1) in physical space, each dataset with the attribute
  // bckp = (time, dt1, dt0, n1, it, nx, ny, nz); then nbackup = 1 - nbackup
  void dump_runtime_backup() override {
    need_fft();
    if (!fields_initialized) throw FlusiError(900004, "fields not initialised (call flusi_init_fields)");
    ensure_phys();
    const bool one = p.backup_type == "one-file-backup";
    const std::string base = out_path("runtime_backup" + std::to_string(nbackup));
    const std::vector<double> bckp = {time, dt1, dt0, (double)n1, (double)it, (double)nx, (double)ny, (double)nz};
    std::vector<double> f(npts);
    bool first = true;
    const void* src[3] = {uk, reg[0], reg[1]};
    const char* names[3][3] = {{"ux", "uy", "uz"}, {"nlkx0", "nlky0", "nlkz0"}, {"nlkx1", "nlky1", "nlkz1"}};
    for (int g = 0; g < 3; ++g)
      for (int c = 0; c < 3; ++c) {
        ifft_fields(cf(const_cast<void*>(src[g]), c), 1, pf(0));
        copy_real_out(pf(0), f.data(), npts);
        const std::string file = one ? base + ".h5" : base + "_" + names[g][c] + ".h5";
        h5_write(file, names[g][c], nx, ny, nz, f.data(), p.field_precision != "double", one ? first : true,
                 {{"bckp", bckp, 0}});
        first = false;
      }
    nbackup = 1 - nbackup;
  }

  // runtime_backuping.f90:198-361 (fluid part): the state, time, dt0, dt1,
  // n1 and it from a backup; the integrating factor is that of dt1
  void read_runtime_backup(const std::string& given) override {
    need_fft();
    ensure_phys();
    const bool one = p.backup_type == "one-file-backup";
    std::string base = out_path(given);
    if (base.size() > 3 && base.compare(base.size() - 3, 3, ".h5") == 0) base.resize(base.size() - 3);
    auto file_of = [&](const char* dset) { return one ? base + ".h5" : base + "_" + dset + ".h5"; };
    if (!file_exists(file_of("ux"))) throw FlusiError(81, "backup file not found: " + file_of("ux"));
    double b[8];
    h5_attr(file_of("ux"), "ux", "bckp", b, 8);
    if ((int)b[5] != nx || (int)b[6] != ny || (int)b[7] != nz)
      throw FlusiError(77776, "ERROR! read_runtime_backup(): Resolution mismatch");
    std::vector<double> f(npts);
    void* dst[3] = {uk, reg[0], reg[1]};
    const char* names[3][3] = {{"ux", "uy", "uz"}, {"nlkx0", "nlky0", "nlkz0"}, {"nlkx1", "nlky1", "nlkz1"}};
    for (int g = 0; g < 3; ++g)
      for (int c = 0; c < 3; ++c) {
        h5_read(file_of(names[g][c]), names[g][c], nx, ny, nz, f.data());
        for (long i = 0; i < npts; ++i)
          if (std::isnan(f[i])) throw FlusiError(100011, "NaN in recently read backup file!! (" + std::string(names[g][c]) + ")");
        copy_real_in(f.data(), pf(0), npts);
        fft_fields(pf(0), 1, cf(dst[g], c));
      }
    time = b[0]; dt1 = b[1]; dt0 = b[2]; n1 = (int)b[3]; it = (int)b[4];
    n0 = 1 - n1;
    vis_dt = dt1;  // runtime_backuping.f90:348 cal_vis(dt1, explin)
    HIPCHECK(hipStreamSynchronize(stream));
  }

  // params.f90:199-233 inicond = "backup": the newer of runtime_backup0/1.h5
  std::string newest_backup() {
    double best = -9.0e9;
    std::string pick;
    const bool one = p.backup_type == "one-file-backup";
    for (int k = 0; k < 2; ++k) {
      const std::string nm = "runtime_backup" + std::to_string(k) + ".h5";
      const std::string probe = one ? out_path(nm) : out_path("runtime_backup" + std::to_string(k) + "_ux.h5");
      if (!file_exists(probe)) continue;
      double b[8];
      h5_attr(probe, "ux", "bckp", b, 8);
      if (b[0] > best) { best = b[0]; pick = nm; }
    }
    if (best < -1.0) throw FlusiError(81, "no backup files found...");
    return pick;
  }

  void get_u(double* u, bool vort) override {
    need_fft();
    inv_uw_phys(uk);
    copy_real_out(vort ? (void*)pf(3) : (void*)pf(0), u, 3 * npts);
  }

  void fft1(const double* f, double* fk) override {
    need_fft();
    ensure_phys();
    copy_real_in(f, pf(0), npts);
    fft_fields(pf(0), 1, reg_scratch());
    copy_real_out(reg_scratch(), fk, 2 * nc);
  }
  void ifft1(const double* fk, double* f) override {
    need_fft();
    ensure_phys();
    copy_real_in(fk, reg_scratch(), 2 * nc);
    ifft_fields(reg_scratch(), 1, pf(0));
    copy_real_out(pf(0), f, npts);
  }
  void* reg_scratch() { return cfp(3); }  // S-layout scratch (3 fields) outside the RHS pipeline

  // ------------------------------------------------------------ init
  void init_fields() override {
    // init_fields_fsi.f90:33-48
    if (!fft_ok && p.inicond != "quiescent") need_fft();
    time = 0.0; dt1 = p.tsave; dt0 = p.tsave; it = 0; vis_dt = -1.0;
    restarted = false;
    const size_t cs = sizeof(C) * nc;
    if (fft_ok) {
      launch_zero(uk, 3 * cs, stream);
      launch_zero(reg[0], 3 * cs, stream);
      launch_zero(reg[1], 3 * cs, stream);
    }
    if (p.inicond == "VortexRing") {
      // init_fields_fsi.f90:381-435 (core radius 0.82*Size unless core_a is given)
      const double r00 = p.length;
      const double a = p.core_a > 0.0 ? p.core_a : 0.82 * r00;
      ensure_phys();
      launch_vortex_ring<T>(geom, p.x0, p.y0, p.z0, r00, a, p.omega1, phys, stream);
      kcheck("vortex ring");
      fft_fields(phys, 3, reg[1]);
      launch_biot_savart<T>(geom, kx_d, ky_d, kz_d, reg[1], uk, stream);
      kcheck("biot-savart");
      launch_zero(reg[1], 3 * cs, stream);
      set_mean_flow(0.0);
    } else if (p.inicond == "MeanFlow") {
      need_fft();
      for (int c = 0; c < 3; ++c) launch_set_mode0<T>(geom, uk, c, p.umean[c], 0.0, stream);
    } else if (p.inicond == "quiescent") {
    } else if (p.inicond == "backup" || p.inicond.rfind("backup::", 0) == 0) {
      // init_fields_fsi.f90:740-747 (and params.f90:199-233 for "backup")
      need_fft();
      read_runtime_backup(p.inicond == "backup" ? newest_backup() : p.inicond.substr(8));
      restarted = true;
    } else {
      throw FlusiError(55523, "??? ERROR: Invalid initial condition" + p.inicond);
    }
    if (!restarted) { n1 = 1; n0 = 0; }
    if (p.iMask == "Insect" && p.iPenalization == 1) {
      is_insect = true;
      insect = Insect();
      insect.init(*ini, p, base_dir);
    }
    if (p.iPenalization == 1) create_mask(time);
    HIPCHECK(hipStreamSynchronize(stream));
    fields_initialized = true;
  }
  void set_mean_flow(double t) {
    for (int c = 0; c < 3; ++c) {
      if (p.iMeanFlow[c] == "fixed") launch_set_mode0<T>(geom, uk, c, p.umean[c], 0.0, stream);
      else if (p.iMeanFlow[c] == "sinusoidal")
        launch_set_mode0<T>(geom, uk, c, p.umean[c] * std::sin(2.0 * PI * p.umean_freq * t), 0.0, stream);
    }
  }

  // ------------------------------------------------------------ diagnostics
  // integrals.f90:522-552 compute_energies_k of the current uk, on the device
  double* epart = nullptr;
  double energy_k() override {
    need_fft();
    if (!epart) epart = (double*)dalloc(sizeof(double) * energy_k_blocks());
    launch_energy_k<T>(geom, uk, epart, stream);
    launch_reduce_rows(epart, energy_k_blocks(), 1, kinds + 4, rmid, red + 9, stream);
    kcheck("energy_k");
    double e = 0.0;
    HIPCHECK(hipMemcpyAsync(&e, red + 9, sizeof(double), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    return e * p.xl * p.yl * p.zl;
  }

  // energy_k() without device allocation: the partial sums, the reducer's
  // middle rows and the result in the file-scope scratch (energy_k() keeps
  // its partials in memory it allocates on its first call); the same kernels
  // in the same order, so the result is the same to the bit
  double energy_k_noalloc() override {
    need_fft();
    const size_t nb = (size_t)energy_k_blocks(), nmid = reduce_rows_scratch(1);
    if (nb + nmid + 1 > kNoAllocScratchDoubles)
      throw FlusiError(900022, "energy_k_noalloc needs " + std::to_string(nb + nmid + 1) +
                                   " doubles of scratch, the maximum is " + std::to_string(kNoAllocScratchDoubles));
    double* part = noalloc_scratch;
    double* mid = part + nb;
    double* out = mid + nmid;
    launch_energy_k<T>(geom, uk, part, stream);
    launch_reduce_rows(part, (int)nb, 1, kinds + 4, mid, out, stream);
    kcheck("energy_k_noalloc");
    double e = 0.0;
    HIPCHECK(hipMemcpyAsync(&e, out, sizeof(double), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    return e * p.xl * p.yl * p.zl;
  }

  // vis.f90:9-66 cal_vis into a host table (the boundary copy; the steppers
  // evaluate the same factors inside the z step)
  double* visd = nullptr;
  void cal_vis(double dt, double* vis) override {
    if (!visd) visd = (double*)dalloc(sizeof(double) * nc);
    double* d = visd;
    launch_cal_vis(geom, kx_d, ky_d, kz_d, kxt2_d, kyt2_d, kzt2_d, dt, p.nu, p.iDealias == 1 ? 1 : 0, d, stream);
    kcheck("cal_vis");
    HIPCHECK(hipMemcpyAsync(vis, d, sizeof(double) * nc, hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
  }

  // drag.f90:17-454 and integrals.f90:48-238 reductions over the physical
  // velocity u (and, with w, vorticity and divergence) with the current mask,
  // us and colour: the DIAG_NV sums of the diagnostics kernel
  std::vector<double> diag_sums(const void* u, const void* w, const void* div) {
    DiagArgs d{};
    d.u = u; d.w = w; d.div = div; d.want_w = w ? 1 : 0;
    d.mask = mask; d.us = us; d.color = color;
    d.eps_inv = 1.0 / p.eps;
    // moments about (x0, y0, z0), which body_motion.f90:250-255 sets to the
    // insect's body centre at every update
    d.x0[0] = p.x0; d.x0[1] = p.y0; d.x0[2] = p.z0;
    if (is_insect) for (int i = 0; i < 3; ++i) d.x0[i] = insect.st.xc[i];
    d.insect = is_insect ? 1 : 0;
    if (is_insect) for (int i = 0; i < 3; ++i) { d.piv_l[i] = insect.st.x_pivot_l_g[i]; d.piv_r[i] = insect.st.x_pivot_r_g[i]; }
    if (!dpart) {
      ndblocks = diag_blocks<T>(geom);
      dpart = (double*)dalloc(sizeof(double) * DIAG_NV * ndblocks);
    }
    d.partial = dpart;
    timed(TM_DIAG, [&] { launch_diag<T>(geom, d, stream); });
    kcheck("diag");
    launch_reduce_rows(dpart, ndblocks, DIAG_NV, kinds + 8, rmid, red + 16, stream);
    kcheck("reduce diag");
    std::vector<double> r(DIAG_NV);
    HIPCHECK(hipMemcpyAsync(r.data(), red + 16, sizeof(double) * DIAG_NV, hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    return r;
  }

  // drag.f90:336-453: forces, moments, penalisation power, residual velocity
  // and (insects) aerodynamic and inertial power from the sums r into the
  // packed integrals o (slots 0..9 and 32..60); with `forces` (cal_drag
  // runs) the penalisation power is updated and the forces.t,
  // forces_partN.t and u_residual.t rows are written, otherwise slots 6..9
  // keep the last cal_drag's power
  void drag_assemble(double t, const std::vector<double>& r, double* o, bool forces) {
    const double dv = p.dx() * p.dy() * p.dz();
    double F[3] = {0, 0, 0}, Tq[3] = {0, 0, 0};
    for (int c = 1; c < 4; ++c)
      for (int i = 0; i < 3; ++i) { F[i] += r[DG_FORCE + 3 * c + i] * dv; Tq[i] += r[DG_TORQUE0 + 3 * c + i] * dv; }
    if (forces) {
      double pw[3];
      for (int i = 0; i < 3; ++i) pw[i] = r[DG_POWER + i] * dv;
      last_power_total = pw[0] + pw[1] + pw[2];
      last_power4_xyz = {pw[0], pw[1], pw[2]};
    }
    for (int i = 0; i < 3; ++i) { o[i] = F[i]; o[3 + i] = Tq[i]; }
    o[6] = last_power_total; o[7] = last_power4_xyz[0]; o[8] = last_power4_xyz[1]; o[9] = last_power4_xyz[2];
    for (int c = 1; c < 4; ++c)
      for (int i = 0; i < 3; ++i) { o[32 + 3 * (c - 1) + i] = r[DG_FORCE + 3 * c + i] * dv; o[44 + 3 * (c - 1) + i] = r[DG_TORQUE + 3 * c + i] * dv; }
    for (int c = 0; c < 6; ++c) o[53 + c] = std::sqrt(c < 4 ? r[DG_URES + c] : 9.0e9);
    if (is_insect) {
      // module_insects.f90:469-586 aero_power / inert_power
      const InsectMotion& s = insect.st;
      double ap_l = 0, ap_r = 0;
      for (int i = 0; i < 3; ++i) { ap_l -= o[44 + 3 + i] * s.rot_rel_l_g[i]; ap_r -= o[44 + 6 + i] * s.rot_rel_r_g[i]; }
      auto ipow = [&](const Vec3& rw, const Vec3& rdt) {
        const Insect& I = insect;
        const double a0 = I.Jxx * rdt[0] + I.Jxy * rdt[1], a1 = I.Jxy * rdt[0] + I.Jyy * rdt[1], a2 = I.Jzz * rdt[2];
        const double b0 = I.Jxx * rw[0] + I.Jxy * rw[1], b1 = I.Jxy * rw[0] + I.Jyy * rw[1], b2 = I.Jzz * rw[2];
        return rw[0] * (a0 + rw[1] * b2 - rw[2] * b1) + rw[1] * (a1 + rw[2] * b0 - rw[0] * b2) + rw[2] * (a2 + rw[0] * b1 - rw[1] * b0);
      };
      const double ip_l = ipow(s.rot_rel_l_w, s.rot_dt_l_w), ip_r = ipow(s.rot_rel_r_w, s.rot_dt_r_w);
      o[59] = 0.0 + ap_l + ap_r;
      o[60] = ip_r + ip_l;
      if (!forces) return;
      append("forces.t", {t, F[0], F[1], F[2], 0, 0, 0, Tq[0], Tq[1], Tq[2], 0, 0, 0, o[59], o[60]}, forces_header(true));
      const double apc[3] = {0.0, ap_l, ap_r}, ipc[3] = {0.0, ip_l, ip_r};
      for (int c = 1; c <= 3; ++c) {
        std::vector<double> row = {t};
        for (int i = 0; i < 3; ++i) row.push_back(o[32 + 3 * (c - 1) + i]);
        row.insert(row.end(), {0, 0, 0});
        for (int i = 0; i < 3; ++i) row.push_back(o[44 + 3 * (c - 1) + i]);
        row.insert(row.end(), {0, 0, 0, apc[c - 1], ipc[c - 1]});
        append("forces_part" + std::to_string(c) + ".t", row, forces_header(true));
      }
    } else {
      if (!forces) return;
      append("forces.t", {t, F[0], F[1], F[2], 0, 0, 0, Tq[0], Tq[1], Tq[2], 0, 0, 0}, forces_header(false));
    }
    append("u_residual.t", {t, o[53], o[54], o[55], o[56], o[57], o[58]}, "%          time     u_residual(0:5)");
  }

  // drag.f90:17-454 cal_drag(time, u, Insect) as its own entry point: the
  // given physical velocity (host, reference layout, float64; null: the
  // velocity of the current uk) with the mask, us and colour as they are (the
  // caller redraws them with create_mask, as integrals.f90:87-100 does)
  void cal_drag(double t, const double* u_host, double* out) override {
    if (!fields_initialized) throw FlusiError(900004, "fields not initialised (call flusi_init_fields)");
    double o[FLUSI_INTEGRALS_NV] = {0};
    if (p.iPenalization == 1) {  // drag.f90:53: nothing to do without penalisation
      ensure_phys();
      if (u_host) copy_real_in(u_host, pf(0), 3 * npts);
      else { need_fft(); ifft_fields(uk, 3, pf(0)); }
      drag_assemble(t, diag_sums(pf(0), nullptr, nullptr), o, true);
    }
    if (out) std::memcpy(out, o, sizeof(o));
  }

  // integrals.f90:48-238 write_integrals_fsi
  void write_integrals(double t, double* out) override {
    need_fft();
    if (!fields_initialized) throw FlusiError(900004, "fields not initialised (call flusi_init_fields)");
    const bool ins = is_insect;
    if (ins) append("kinematics.t", insect.kinematics_row(t), kin_header());
    inv_uw_phys(uk);
    const bool forces = p.compute_forces == 1 && p.iPenalization == 1;
    if (p.compute_forces == 1 && p.iMoving == 1) create_mask(t);
    // divergence (basic_operators.f90:707-734) -> physical
    void* dk = reg_scratch();
    launch_divergence<T>(geom, kx_d, ky_d, kz_d, uk, dk, stream);
    kcheck("divergence");
    ifft_fields(dk, 1, pf(6));
    const std::vector<double> r = diag_sums(pf(0), pf(3), pf(6));
    std::vector<C> m0(3);
    for (int c = 0; c < 3; ++c)
      HIPCHECK(hipMemcpyAsync(&m0[c], cf(uk, c), sizeof(C), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    const double dv = p.dx() * p.dy() * p.dz();
    double o[FLUSI_INTEGRALS_NV] = {0};
    drag_assemble(t, r, o, forces);
    const double ekf[3] = {0.5 * dv * r[DG_EKF], 0.5 * dv * r[DG_EKF + 1], 0.5 * dv * r[DG_EKF + 2]};
    const double ek[3] = {0.5 * dv * r[DG_EK], 0.5 * dv * r[DG_EK + 1], 0.5 * dv * r[DG_EK + 2]};
    const double dsf[3] = {p.nu * dv * r[DG_DISSF], p.nu * dv * r[DG_DISSF + 1], p.nu * dv * r[DG_DISSF + 2]};
    const double ds[3] = {p.nu * dv * r[DG_DISS], p.nu * dv * r[DG_DISS + 1], p.nu * dv * r[DG_DISS + 2]};
    auto put4 = [&](int off, const double* v3, double tot) { o[off] = tot; o[off + 1] = v3[0]; o[off + 2] = v3[1]; o[off + 3] = v3[2]; };
    put4(10, ekf, 0.5 * dv * (r[DG_EKF] + r[DG_EKF + 1] + r[DG_EKF + 2]));
    put4(14, dsf, p.nu * dv * (r[DG_DISSF] + r[DG_DISSF + 1] + r[DG_DISSF + 2]));
    put4(18, ek, 0.5 * dv * (r[DG_EK] + r[DG_EK + 1] + r[DG_EK + 2]));
    put4(22, ds, p.nu * dv * (r[DG_DISS] + r[DG_DISS + 1] + r[DG_DISS + 2]));
    o[26] = r[DG_DIVMAX]; o[27] = r[DG_DIVMAXF];
    o[28] = r[DG_MASKVOL] * dv;
    for (int i = 0; i < 3; ++i) o[29 + i] = r[DG_MASKUS + i] * dv;
    append("divu.t", {t, o[26], o[27]}, "%          time         max_div   max_div_fluid");
    std::vector<double> erow = {t};
    for (int i = 10; i < 26; ++i) erow.push_back(o[i]);
    for (int i = 6; i < 10; ++i) erow.push_back(o[i]);
    append("energy.t", erow, "%          time         E_kin_f       E_kin_x_f       E_kin_y_f       E_kin_z_f          diss_f        diss_x_f        diss_y_f        diss_z_f       E_kin_tot     E_kin_tot_x     E_kin_tot_y     E_kin_tot_z        diss_tot      diss_tot_x      diss_tot_y      diss_tot_z      flux_penal");
    append("meanflow.t", {t, (double)m0[0].x, (double)m0[1].x, (double)m0[2].x}, "%          time          uxmean          uymean          uzmean");
    append("mask_volume.t", {t, o[28], o[29], o[30], o[31]}, "%          time     mask_volume      usx_integral    usy_integral    usz_integral");
    if (out) std::memcpy(out, o, sizeof(o));
  }
  static const char* kin_header() {
    return "%          time       xc_body_g         yc_body         zc_body             psi            beta           gamma      eta_stroke         alpha_l           phi_l         theta_l         alpha_r           phi_r         theta_r         rot_l_x         rot_l_y         rot_l_z         rot_r_x         rot_r_y         rot_r_z      rot_dt_l_x      rot_dt_l_y      rot_dt_l_z      rot_dt_r_x      rot_dt_r_y      rot_dt_r_z";
  }
  static const char* forces_header(bool ins) {
    return ins ? "%          time          Forcex          Forcey          Forcez     Forcex_unst     Forcey_unst     Forcez_unst         Momentx         Momenty         Momentz    Momentx_unst    Momenty_unst    Momentz_unst        Powerx     Inertial_power"
               : "%          time          Forcex          Forcey          Forcez     Forcex_unst     Forcey_unst     Forcez_unst         Momentx         Momenty         Momentz    Momentx_unst    Momenty_unst    Momentz_unst";
  }
};

}  // namespace flusi

// =========================================================================== C ABI
struct flusi_ctx {
  std::unique_ptr<flusi::SolverBase> s;
  std::string err;
  bool borrowed = false;  // a member of an ensemble, freed with it
};

namespace flusi {

// ------------------------------------------------------------ ensembles
// Many independent simulations advanced together.  Members whose grids and
// kernel-relevant parameters agree form a batch: each member's device state
// lives in one slab, the slabs of a batch sit a fixed stride apart, and every
// pass of a step is one launch over the whole batch (one grid dimension per
// member).  One step of all batches can be captured into a hipGraph (one
// branch per batch, forked from an origin stream) and replayed; before each
// replay the host writes only the moving insects' per-stage mask arguments.

// parameters that enter the batched kernels' arguments (the mask geometry
// does not: spheres are drawn per member, insects get one record per member)
static std::string batch_key(const Params& p, int precision) {
  char b[1024];
  std::snprintf(b, sizeof(b),
                "%d|%d %d %d|%.17g %.17g %.17g|%s|%.17g|%.17g|%d|%d %d|%s|%.17g|%s %d|%s %s %.17g %d|%s %s %s|%.17g|"
                "%.17g %.17g %.17g|%.17g|%s %.17g %.17g|%d",
                precision, p.nx, p.ny, p.nz, p.xl, p.yl, p.zl, p.method.c_str(), p.dt_fixed, p.nu, p.iDealias,
                p.iPenalization, p.iMoving, p.iMask.c_str(), p.eps, p.iCavity.c_str(), p.cavity_size,
                p.iVorticitySponge.c_str(), p.iSpongeType.c_str(), p.eps_sponge, p.sponge_thickness,
                p.iMeanFlow[0].c_str(), p.iMeanFlow[1].c_str(), p.iMeanFlow[2].c_str(), p.m_fluid, p.umean[0],
                p.umean[1], p.umean[2], p.umean_freq, p.startup_conditioner.c_str(), p.tau_meanflow,
                p.T_release_meanflow, p.periodic ? 1 : 0);
  return b;
}

class EnsembleBase {
 public:
  virtual ~EnsembleBase() = default;
  std::string err;
  std::vector<flusi_ctx*> ctxs;
  virtual int nbatches() const = 0;
  virtual void batch_info(int b, int* nmem, int* nx, int* ny, int* nz, int* lead) const = 0;
  virtual void init_fields() = 0;
  virtual void autotune(int reps, int* variants, double* ms) = 0;
  virtual void advance(int nsteps, int mode) = 0;
};

template <typename T>
class Ensemble : public EnsembleBase {
  struct Group {
    std::vector<int> idx;   // member indices, leader first
    char* slab = nullptr;
    size_t stride = 0;
    int nstage = 0;         // insect mask updates per captured step
    hipEvent_t join = nullptr;
    Solver<T>* lead = nullptr;
  };
  std::vector<Group> groups;
  hipStream_t origin = nullptr;
  hipEvent_t fork = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  static constexpr int RING = 4;  // pinned buffers of per-replay mask arguments
  InsectArgs* ring_h[RING] = {};
  hipEvent_t ring_ev[RING] = {};
  size_t ring_n = 0;
  int ring_next = 0;

  Solver<T>& S(int i) { return *static_cast<Solver<T>*>(ctxs[i]->s.get()); }

 public:
  // separate from the constructor so that a failure part-way is cleaned up
  // by the destructor
  void build(const std::vector<std::string>& paths, int precision) {
    HIPCHECK(hipStreamCreate(&origin));
    HIPCHECK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    const int n = (int)paths.size();
    std::vector<std::unique_ptr<IniFile>> inis(n);
    std::vector<Params> ps(n);
    std::vector<std::string> keys(n);
    for (int i = 0; i < n; ++i) {
      inis[i].reset(new IniFile(paths[i]));
      ps[i] = load_params(*inis[i]);
      keys[i] = batch_key(ps[i], precision);
    }
    for (int i = 0; i < n; ++i) {
      bool placed = false;
      for (Group& g : groups)
        if (keys[g.idx[0]] == keys[i]) { g.idx.push_back(i); placed = true; break; }
      if (!placed) { groups.emplace_back(); groups.back().idx.push_back(i); }
    }
    ctxs.assign(n, nullptr);
    for (Group& g : groups) {
      {  // one member's setup() bytes, from a throw-away solver
        Solver<T> probe;
        probe.p = ps[g.idx[0]];
        probe.setup();
        probe.need_fft();
        g.stride = (probe.slab_used + 65535) / 65536 * 65536;
      }
      HIPCHECK(hipMalloc((void**)&g.slab, g.stride * g.idx.size()));
      HIPCHECK(hipEventCreateWithFlags(&g.join, hipEventDisableTiming));
      for (size_t k = 0; k < g.idx.size(); ++k) {
        const int i = g.idx[k];
        auto c = new flusi_ctx();
        c->borrowed = true;
        ctxs[i] = c;
        auto s = new Solver<T>();
        c->s.reset(s);
        s->ini = std::move(inis[i]);
        s->p = ps[i];
        const size_t sl = paths[i].find_last_of('/');
        s->base_dir = (sl == std::string::npos) ? "." : paths[i].substr(0, sl);
        s->slab = g.slab + k * g.stride;
        s->slab_cap = g.stride;
        s->setup();
      }
      // every device pointer of a member must sit k * stride after the leader's
      Solver<T>& L = S(g.idx[0]);
      for (size_t k = 1; k < g.idx.size(); ++k) {
        Solver<T>& m = S(g.idx[k]);
        const void* a[] = {L.uk, L.reg[0], L.reg[1], L.Q, L.P, L.mask, L.us, L.color, L.tab, L.tw_z, L.tw_y, L.tw_M,
                           L.tw_N, L.xpart, L.zpart, L.dpart, L.red, L.rmid, L.kinds};
        const void* b[] = {m.uk, m.reg[0], m.reg[1], m.Q, m.P, m.mask, m.us, m.color, m.tab, m.tw_z, m.tw_y, m.tw_M,
                           m.tw_N, m.xpart, m.zpart, m.dpart, m.red, m.rmid, m.kinds};
        for (size_t j = 0; j < sizeof(a) / sizeof(a[0]); ++j)
          if ((const char*)b[j] - (const char*)a[j] != (long)(k * g.stride))
            throw FlusiError(900014, "ensemble: member buffers do not follow the batch layout");
      }
      g.lead = &L;
      for (int i : g.idx) L.batch.push_back(&S(i));
    }
  }
  ~Ensemble() override {
    if (origin) (void)hipStreamSynchronize(origin);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (flusi_ctx* c : ctxs) delete c;  // solvers first: their buffers live in the slabs
    for (Group& g : groups) {
      if (g.slab) (void)hipFree(g.slab);
      if (g.join) (void)hipEventDestroy(g.join);
    }
    for (int r = 0; r < RING; ++r) {
      if (ring_ev[r]) (void)hipEventDestroy(ring_ev[r]);
      if (ring_h[r]) (void)hipHostFree(ring_h[r]);
    }
    if (fork) (void)hipEventDestroy(fork);
    if (origin) (void)hipStreamDestroy(origin);
  }

  int nbatches() const override { return (int)groups.size(); }
  void batch_info(int b, int* nmem, int* nx, int* ny, int* nz, int* lead) const override {
    if (b < 0 || b >= (int)groups.size()) throw FlusiError(900015, "ensemble: batch index out of range");
    const Group& g = groups[b];
    if (nmem) *nmem = (int)g.idx.size();
    if (lead) *lead = g.idx[0];
    if (nx) *nx = g.lead->nx;
    if (ny) *ny = g.lead->ny;
    if (nz) *nz = g.lead->nz;
  }

  void init_fields() override {
    for (flusi_ctx* c : ctxs) c->s->init_fields();
  }

  struct BatchScope {  // the leader launches for its whole batch while alive
    Solver<T>& L;
    BatchScope(Solver<T>& l, size_t n, size_t stride) : L(l) {
      L.bat = Batch{(int)n, (long)stride};
      L.ens_step = true;
    }
    ~BatchScope() { L.bat = Batch{1, 0}; L.ens_step = false; }
  };

  void autotune(int reps, int* variants, double* ms) override {
    for (size_t b = 0; b < groups.size(); ++b) {
      Group& g = groups[b];
      BatchScope bs(*g.lead, g.idx.size(), g.stride);
      g.lead->autotune(reps, variants ? variants + Solver<T>::kVariantInts * b : nullptr, ms ? ms + b : nullptr, false);
      for (int i : g.idx) S(i).apply(g.lead->variant());
    }
    if (exec) { (void)hipGraphExecDestroy(exec); exec = nullptr; (void)hipGraphDestroy(graph); graph = nullptr; }
  }

  // host bookkeeping of one step for every member (time_step.f90:57-167),
  // and their output at output times
  void after_step(bool synced) {
    for (Group& g : groups) {
      Solver<T>& L = *g.lead;
      for (int i : g.idx) {
        Solver<T>& m = S(i);
        if (&m != &L) { m.dt0 = L.dt0; m.dt1 = L.dt1; m.vis_dt = L.vis_dt; m.n0 = L.n0; m.n1 = L.n1; }
      }
    }
    for (flusi_ctx* c : ctxs) {
      Solver<T>& m = *static_cast<Solver<T>*>(c->s.get());
      if (!m.out_dir.empty()) {
        if (!synced) { HIPCHECK(hipDeviceSynchronize()); synced = true; }
        m.record_ekin(true);
      }
      m.time = m.time + m.dt1;
      m.it += 1;
      if (!m.out_dir.empty() &&
          time_for_output(m.time, m.dt1, m.it, m.p.tintegral, m.p.itdrag, m.p.tmax, 0.0, m.p.intelligent_dt == "yes")) {
        if (!synced) { HIPCHECK(hipDeviceSynchronize()); synced = true; }
        m.write_integrals(m.time, nullptr);
      }
    }
  }

  void batched_step() {
    for (Group& g : groups) {
      Solver<T>& L = *g.lead;
      if (L.p.dt_fixed <= 0.0 && g.idx.size() > 1)
        throw FlusiError(900016, "ensemble: batched steps need dt_fixed > 0 (one time step for the whole batch)");
      BatchScope bs(L, g.idx.size(), g.stride);
      L.dt0 = L.dt1;
      L.fluid_timestep();
    }
    after_step(false);
  }

  void check_capturable(const Solver<T>& L) const {
    const Params& p = L.p;
    if (p.method != "RK2" && p.method != "RK4")
      throw FlusiError(900017, "ensemble graph: RK2 or RK4 only (AB2 alternates its registers)");
    if (p.dt_fixed <= 0.0) throw FlusiError(900017, "ensemble graph: dt_fixed > 0 required");
    if (p.startup_conditioner == "yes") throw FlusiError(900017, "ensemble graph: startup_conditioner depends on time");
    for (int c = 0; c < 3; ++c)
      if (p.iMeanFlow[c] == "sinusoidal") throw FlusiError(900017, "ensemble graph: sinusoidal mean flow depends on time");
  }
  int mask_updates(const Solver<T>& L) const {
    if (!(L.p.iMoving == 1 && L.p.iPenalization == 1 && (L.p.iMask == "Insect" || L.p.iMask == "insect"))) return 0;
    return L.p.method == "RK4" ? 4 : 2;
  }

  void capture() {
    for (Group& g : groups) {
      check_capturable(*g.lead);
      g.nstage = mask_updates(*g.lead);
      Solver<T>& L = *g.lead;
      L.timers(0, nullptr);  // no event timing inside the graph
      if (g.nstage && !L.ins_cap)
        L.ins_cap = (InsectArgs*)L.dalloc(sizeof(InsectArgs) * Solver<T>::INS_SLOTS * g.idx.size());
    }
    HIPCHECK(hipDeviceSynchronize());
    HIPCHECK(hipStreamBeginCapture(origin, hipStreamCaptureModeThreadLocal));
    try {
      HIPCHECK(hipEventRecord(fork, origin));
      for (Group& g : groups) {
        Solver<T>& L = *g.lead;
        HIPCHECK(hipStreamWaitEvent(L.stream, fork, 0));
        BatchScope bs(L, g.idx.size(), g.stride);
        // the host state is restored afterwards: recording runs no step
        const double d0 = L.dt0, d1 = L.dt1, vd = L.vis_dt;
        const int n0 = L.n0, n1 = L.n1;
        std::vector<Insect> saved;
        for (int i : g.idx) saved.push_back(S(i).insect);
        L.capture = true;
        L.stage_slot = 0;
        L.dt0 = L.dt1;
        try {
          L.fluid_timestep();
        } catch (...) {
          L.capture = false;
          throw;
        }
        L.capture = false;
        if (L.stage_slot != g.nstage) throw FlusiError(900018, "ensemble graph: unexpected number of mask updates");
        L.dt0 = d0; L.dt1 = d1; L.vis_dt = vd; L.n0 = n0; L.n1 = n1;
        for (size_t k = 0; k < g.idx.size(); ++k) S(g.idx[k]).insect = saved[k];
        HIPCHECK(hipEventRecord(g.join, L.stream));
        HIPCHECK(hipStreamWaitEvent(origin, g.join, 0));
      }
    } catch (...) {
      hipGraph_t gdrop = nullptr;
      (void)hipStreamEndCapture(origin, &gdrop);
      if (gdrop) (void)hipGraphDestroy(gdrop);
      throw;
    }
    HIPCHECK(hipStreamEndCapture(origin, &graph));
    HIPCHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    size_t need = 0;
    for (Group& g : groups) need += (size_t)g.nstage * g.idx.size();
    if (need > ring_n) {
      for (int r = 0; r < RING; ++r) {
        if (ring_h[r]) { if (ring_ev[r]) HIPCHECK(hipEventSynchronize(ring_ev[r])); HIPCHECK(hipHostFree(ring_h[r])); }
        HIPCHECK(hipHostMalloc((void**)&ring_h[r], sizeof(InsectArgs) * need, hipHostMallocDefault));
      }
      ring_n = need;
    }
  }

  void graph_step() {
    const int r = ring_next;
    ring_next = (ring_next + 1) % RING;
    if (ring_ev[r]) HIPCHECK(hipEventSynchronize(ring_ev[r]));
    else HIPCHECK(hipEventCreateWithFlags(&ring_ev[r], hipEventDisableTiming));
    size_t off = 0;
    for (Group& g : groups) {
      if (!g.nstage) continue;
      Solver<T>& L = *g.lead;
      const size_t M = g.idx.size();
      const double t = L.time, dt = L.p.dt_fixed;
      // the stage times at which fluid_timestep redraws the mask
      const double ts[4] = {t, g.nstage == 4 ? t + 0.5 * dt : t, t + 0.5 * dt, t + dt};
      InsectArgs* h = ring_h[r] + off;
      for (int s = 0; s < g.nstage; ++s)
        for (size_t k = 0; k < M; ++k) {
          Solver<T>& m = S(g.idx[k]);
          m.insect.update(ts[s]);
          h[s * M + k] = m.insect_args(true);
        }
      HIPCHECK(hipMemcpyAsync(L.ins_cap, h, sizeof(InsectArgs) * g.nstage * M, hipMemcpyHostToDevice, origin));
      off += g.nstage * M;
    }
    HIPCHECK(hipEventRecord(ring_ev[r], origin));
    HIPCHECK(hipGraphLaunch(exec, origin));
    for (Group& g : groups) g.lead->dt0 = g.lead->dt1;
    after_step(false);
  }

  // mode 0: every member steps on its own; 1: batched launches; 2: batched
  // launches replayed from a captured graph
  void advance(int nsteps, int mode) override {
    for (flusi_ctx* c : ctxs)
      if (!c->s->fields_initialized) throw FlusiError(900004, "fields not initialised (call flusi_ensemble_init_fields)");
    if (mode == 0) {
      for (int k = 0; k < nsteps; ++k)
        for (flusi_ctx* c : ctxs) c->s->step();
      HIPCHECK(hipDeviceSynchronize());
      return;
    }
    if (mode != 1 && mode != 2) throw FlusiError(900019, "ensemble: mode must be 0, 1 or 2");
    int k = 0;
    if (mode == 2) {
      // the captured step assumes an established integrating factor
      for (; k < nsteps; ++k) {
        bool fresh = false;
        for (flusi_ctx* c : ctxs) fresh = fresh || c->s->it == 0;
        if (!fresh) break;
        batched_step();
      }
      if (k < nsteps && !exec) capture();
      for (; k < nsteps; ++k) graph_step();
    } else {
      for (; k < nsteps; ++k) batched_step();
    }
    HIPCHECK(hipStreamSynchronize(origin));
    for (Group& g : groups) HIPCHECK(hipStreamSynchronize(g.lead->stream));
  }
};

}  // namespace flusi

using flusi::FlusiError;

struct flusi_ensemble {
  std::unique_ptr<flusi::EnsembleBase> e;
  std::string err;
};

template <typename F> static int eguarded(flusi_ensemble* c, F&& f) {
  if (!c || !c->e) return 900000;
  try {
    f();
    return 0;
  } catch (const FlusiError& x) {
    c->err = x.what();
    return x.code;
  } catch (const std::exception& x) {
    c->err = x.what();
    return 900099;
  }
}

template <typename F> static int guarded(flusi_ctx* c, F&& f) {
  if (!c || !c->s) return 900000;
  try {
    f();
    return 0;
  } catch (const FlusiError& e) {
    c->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    c->err = e.what();
    return 900099;
  }
}

extern "C" {

int flusi_init(const char* ini_path, int precision, flusi_ctx** out) {
  if (!out) return 900000;
  *out = nullptr;
  auto ctx = new flusi_ctx();
  try {
    if (precision == FLUSI_FP64) ctx->s.reset(new flusi::Solver<double>());
    else if (precision == FLUSI_FP32_MIXED) ctx->s.reset(new flusi::Solver<float>());
    else throw FlusiError(900005, "precision must be FLUSI_FP64 or FLUSI_FP32_MIXED");
    flusi::SolverBase* s = ctx->s.get();
    s->ini.reset(new flusi::IniFile(ini_path));
    s->p = flusi::load_params(*s->ini);
    std::string path(ini_path);
    const size_t sl = path.find_last_of('/');
    s->base_dir = (sl == std::string::npos) ? "." : path.substr(0, sl);
    if (precision == FLUSI_FP64) static_cast<flusi::Solver<double>*>(s)->setup();
    else static_cast<flusi::Solver<float>*>(s)->setup();
  } catch (const FlusiError& e) {
    ctx->err = e.what();
    const int code = e.code;
    *out = ctx;  // keep the context so flusi_last_error can report
    ctx->s.reset();
    return code;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    *out = ctx;
    ctx->s.reset();
    return 900099;
  }
  *out = ctx;
  return 0;
}

void flusi_free(flusi_ctx* ctx) {
  if (ctx && !ctx->borrowed) delete ctx;  // ensemble members go with their ensemble
}

const char* flusi_last_error(const flusi_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int flusi_get_dims(const flusi_ctx* ctx, int* nx, int* ny, int* nz) {
  if (!ctx || !ctx->s) return 900000;
  *nx = ctx->s->nx; *ny = ctx->s->ny; *nz = ctx->s->nz;
  return 0;
}

int flusi_set_output_dir(flusi_ctx* ctx, const char* dir) {
  return guarded(ctx, [&] { ctx->s->out_dir = dir ? dir : ""; });
}

int flusi_init_fields(flusi_ctx* ctx) { return guarded(ctx, [&] { ctx->s->init_fields(); }); }
int flusi_set_uk(flusi_ctx* ctx, const double* uk) { return guarded(ctx, [&] { ctx->s->set_uk(uk); }); }
int flusi_get_uk(flusi_ctx* ctx, double* uk) { return guarded(ctx, [&] { ctx->s->get_uk(uk); }); }
int flusi_get_u(flusi_ctx* ctx, double* u) { return guarded(ctx, [&] { ctx->s->get_u(u, false); }); }
int flusi_get_vorticity(flusi_ctx* ctx, double* w) { return guarded(ctx, [&] { ctx->s->get_u(w, true); }); }
int flusi_get_pressure(flusi_ctx* ctx, double time, double* p) {
  return guarded(ctx, [&] { ctx->s->get_pressure(time, p); });
}
int flusi_cal_nlk(flusi_ctx* ctx, double time, int it, double* nlk, double* u) {
  (void)it;
  return guarded(ctx, [&] { ctx->s->cal_nlk(time, nlk, u); });
}
int flusi_fluid_timestep(flusi_ctx* ctx, double* time, int* it, double* dt0, double* dt1) {
  return guarded(ctx, [&] {
    flusi::SolverBase* s = ctx->s.get();
    if (time) s->time = *time;
    if (it) s->it = *it;
    if (dt0) s->dt0 = *dt0;
    s->fluid_timestep();
    if (dt1) *dt1 = s->dt1;
  });
}
int flusi_step(flusi_ctx* ctx) { return guarded(ctx, [&] { ctx->s->step(); }); }
int flusi_run(flusi_ctx* ctx, int nsteps) {
  return guarded(ctx, [&] {
    flusi::SolverBase* s = ctx->s.get();
    if (!s->fields_initialized) s->init_fields();
    s->run(nsteps);
  });
}
int flusi_save_fields(flusi_ctx* ctx) { return guarded(ctx, [&] { ctx->s->save_fields(); }); }
int flusi_dump_runtime_backup(flusi_ctx* ctx) { return guarded(ctx, [&] { ctx->s->dump_runtime_backup(); }); }
int flusi_read_runtime_backup(flusi_ctx* ctx, const char* file) {
  return guarded(ctx, [&] {
    if (!file) throw FlusiError(900000, "no backup file given");
    flusi::SolverBase* s = ctx->s.get();
    if (!s->fields_initialized) throw FlusiError(900004, "fields not initialised (call flusi_init_fields)");
    s->read_runtime_backup(file);
    s->restarted = true;
    if (s->p.iPenalization == 1) s->create_mask(s->time);  // init_fields.f90:98-101
  });
}
int flusi_advance(flusi_ctx* ctx, int nsteps) {
  return guarded(ctx, [&] { for (int i = 0; i < nsteps; ++i) ctx->s->step(); });
}
int flusi_get_time(const flusi_ctx* ctx, double* time, int* it, double* dt0, double* dt1) {
  if (!ctx || !ctx->s) return 900000;
  if (time) *time = ctx->s->time;
  if (it) *it = ctx->s->it;
  if (dt0) *dt0 = ctx->s->dt0;
  if (dt1) *dt1 = ctx->s->dt1;
  return 0;
}
int flusi_create_mask(flusi_ctx* ctx, double time) { return guarded(ctx, [&] { ctx->s->create_mask(time); }); }
int flusi_get_mask(flusi_ctx* ctx, double* mask, double* us, int16_t* color) {
  return guarded(ctx, [&] { ctx->s->get_mask(mask, us, color); });
}
int flusi_write_integrals(flusi_ctx* ctx, double time, double* out) {
  return guarded(ctx, [&] { ctx->s->write_integrals(time, out); });
}
int flusi_cal_drag(flusi_ctx* ctx, double time, const double* u, double* out) {
  return guarded(ctx, [&] { ctx->s->cal_drag(time, u, out); });
}
int flusi_energy_k(flusi_ctx* ctx, double* E) { return guarded(ctx, [&] { *E = ctx->s->energy_k(); }); }
int flusi_energy_k_noalloc(flusi_ctx* ctx, double* E) {
  return guarded(ctx, [&] { *E = ctx->s->energy_k_noalloc(); });
}
int flusi_cal_vis(flusi_ctx* ctx, double dt, double* vis) { return guarded(ctx, [&] { ctx->s->cal_vis(dt, vis); }); }
int flusi_fft(flusi_ctx* ctx, const double* f, double* fk) { return guarded(ctx, [&] { ctx->s->fft1(f, fk); }); }
int flusi_ifft(flusi_ctx* ctx, const double* fk, double* f) { return guarded(ctx, [&] { ctx->s->ifft1(fk, f); }); }
int flusi_autotune(flusi_ctx* ctx, int reps, int* variant, double* ms) {
  return guarded(ctx, [&] { ctx->s->autotune_v(reps, variant, ms); });
}
int flusi_adjoint_ncontrols(flusi_ctx* ctx, int* n) {
  return guarded(ctx, [&] { *n = ctx->s->adjoint_ncontrols(); });
}
int flusi_adjoint_get_controls(flusi_ctx* ctx, double* theta) {
  return guarded(ctx, [&] { ctx->s->adjoint_get_controls(theta); });
}
int flusi_adjoint_forward(flusi_ctx* ctx, const double* theta, int nsteps, int every, double* J) {
  return guarded(ctx, [&] {
    const double j = ctx->s->adjoint_forward(theta, nsteps, every);
    if (J) *J = j;
  });
}
int flusi_adjoint_backward(flusi_ctx* ctx, double* grad) {
  return guarded(ctx, [&] { ctx->s->adjoint_backward(grad); });
}
int flusi_timers(flusi_ctx* ctx, int enable, double* ms, const char** names) {
  return guarded(ctx, [&] {
    ctx->s->timers(enable, ms);
    if (names) for (int i = 0; i < FLUSI_TIMER_SLOTS; ++i) names[i] = flusi::kTimerNames[i];
  });
}

// ensembles
int flusi_ensemble_create(int n, const char* const* ini_paths, int precision, flusi_ensemble** out) {
  if (!out || n < 1 || !ini_paths) return 900000;
  *out = nullptr;
  auto e = new flusi_ensemble();
  try {
    std::vector<std::string> paths(ini_paths, ini_paths + n);
    if (precision == FLUSI_FP64) {
      auto x = new flusi::Ensemble<double>();
      e->e.reset(x);
      x->build(paths, precision);
    } else if (precision == FLUSI_FP32_MIXED) {
      auto x = new flusi::Ensemble<float>();
      e->e.reset(x);
      x->build(paths, precision);
    } else {
      throw FlusiError(900005, "precision must be FLUSI_FP64 or FLUSI_FP32_MIXED");
    }
  } catch (const FlusiError& x) {
    e->err = x.what();
    e->e.reset();
    *out = e;
    return x.code;
  } catch (const std::exception& x) {
    e->err = x.what();
    e->e.reset();
    *out = e;
    return 900099;
  }
  *out = e;
  return 0;
}
void flusi_ensemble_free(flusi_ensemble* e) { delete e; }
const char* flusi_ensemble_last_error(const flusi_ensemble* e) { return e ? e->err.c_str() : "null ensemble"; }
int flusi_ensemble_size(const flusi_ensemble* e, int* nmembers, int* nbatches) {
  if (!e || !e->e) return 900000;
  if (nmembers) *nmembers = (int)e->e->ctxs.size();
  if (nbatches) *nbatches = e->e->nbatches();
  return 0;
}
int flusi_ensemble_batch(flusi_ensemble* e, int b, int* nmem, int* nx, int* ny, int* nz, int* lead) {
  return eguarded(e, [&] { e->e->batch_info(b, nmem, nx, ny, nz, lead); });
}
int flusi_ensemble_member(flusi_ensemble* e, int i, flusi_ctx** member) {
  return eguarded(e, [&] {
    if (i < 0 || i >= (int)e->e->ctxs.size()) throw FlusiError(900015, "ensemble: member index out of range");
    *member = e->e->ctxs[i];
  });
}
int flusi_ensemble_init_fields(flusi_ensemble* e) { return eguarded(e, [&] { e->e->init_fields(); }); }
int flusi_ensemble_autotune(flusi_ensemble* e, int reps, int* variants, double* ms) {
  return eguarded(e, [&] { e->e->autotune(reps, variants, ms); });
}
int flusi_ensemble_advance(flusi_ensemble* e, int nsteps, int mode) {
  return eguarded(e, [&] { e->e->advance(nsteps, mode); });
}

}  // extern "C"
