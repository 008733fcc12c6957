// Host runtime and C ABI of libkmm (see include/kmm.h).
//
// Owns the device-resident state of one channel, the per-pencil tau and
// Green's-function tables, the immersed-boundary preprocessing, restart I/O
// and the launch sequence of one time step:
//   k_pencil (y)  ->  k_zinv  ->  k_ibdrag  ->  k_xplane  ->  k_zfwd
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/kmm.h"
#include "kmm_kernels.h"

using kmm::PencilArgs;
using kmm::PlaneArgs;

static thread_local std::string g_last_error;

#define KMM_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      g_last_error = std::string(#call) + ": " + hipGetErrorString(e_);                 \
      return KMM_E_HIP;                                                                 \
    }                                                                                   \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
};

struct kmm_ctx {
  kmm_config cfg;
  int device = 0;
  hipStream_t stream = nullptr;
  int nx, ny, nz, nyp, nxh, nzh, mx, mz, P;
  double alpha, beta, dyde, dPdx, gH, gS;
  std::vector<double> ycoord, seght;
  size_t S;  // complex elements of one spectral field
  // state
  double2 *V, *OMY, *FN, *FNM1, *GN, *GNM1, *TH, *SCN, *SCNM1;
  double *u0, *w0, *h1n, *h1nm1, *h3n, *h3nm1;
  void* NL[4];
  void* T[9];
  void *Z, *ZF;
  // tables
  double2 *tabH, *tabP, *tabS, *G1, *G2;
  double4* gk;
  void *twM, *twZ, *twX;   // transfer-precision twiddles
  size_t vsz = 16;         // bytes per complex transfer element (16 FP64, 8 mixed)
  double *smz, *smy, *seght_d;
  unsigned long long* cfl_d;
  int cfl_cap = 0;
  // immersed boundary
  int nsrc = 0;
  int4* src = nullptr;
  double* srcp = nullptr;   // [nsrc][6] per-type constants of the IB cells
  double *fint = nullptr, *drag = nullptr;
  int *trow = nullptr, *tx = nullptr, *tptr = nullptr, *ent_src = nullptr, *fixrow = nullptr, *fixx = nullptr;
  double* ent_w = nullptr;
  std::vector<int4> src_h;
  std::vector<double> fint_full[3];  // host copies (nyp,mz,mx) Fortran order
  // bookkeeping
  int it = 0;
  bool nl_valid = false;  // NL buffers hold the nonlinear terms of the state
  double last_cfl = 0.0;
  int last_cfl_it = 0;
  double t_pencil = 0.0, t_plane = 0.0;
  hipEvent_t ev[3];
  // per-kernel timing (pencil, zinv, ibdrag, xplane, zfwd) over a kmm_step call
  int timing = 0;
  std::vector<hipEvent_t> tev;
  double ktime[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int ksamples = 0;
  std::vector<void*> allocs;
  // Lagrangian tracers
  int npart = 0;
  int part_cap = 0;
  double* part = nullptr;    // [7][part_cap]: x, y, z, u, v, w, swirl
  double* scsrc = nullptr;   // [nyp][mz][mx] Gaussian-blob scalar source (scl_flag 2)
  int tune_pencil = 0, tune_xplane = 0;  // occupancy variants (kmm_autotune)
  double* stats_d = nullptr;             // [8][nyp] statistics sums (kmm_statistics)
  double* swirl_tmp = nullptr;           // [nyp][mz][mx] swirl for statistics without tracers
  std::vector<int32_t> imatrix_h;        // geometry as set (the autotuner's twin needs it)
  double* cap = nullptr;     // [4][nyp][mz][mx] physical u, v, w (and swirl) of the last evaluation
  bool part_swirl = false;   // swirl strength at the particles (solvers.f90:3479-3481)
  double* ycoord_d = nullptr;
  int* perr = nullptr;
  int part_state = 0;        // 0 none, 1 initialise at the next evaluation, 2 tracking
  int substep = 1;
  // inertial particles (particle_flag > 0): FP64 fields on the 3/2 grid
  double* mrf = nullptr;                 // [12][P] vorticity, velocity gradient
  double* lap[2] = {nullptr, nullptr};   // [3][P] Laplacians, current / previous evaluation
  int lap_cur = 0;
  double* uold = nullptr;                // [3][P] u, v, w of the previous evaluation
  void* lapT = nullptr;                  // [3] Laplacians in physical y (transfer precision)
  void* lapZ = nullptr;                  // [3] after the z inverse
  bool cap_valid = false;
  double* dpdx_d = nullptr;  // mean pressure gradient (device; updated under constant mass flux)
  // captured steady steps (kmm_set_graph)
  int graph_steps = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  long long graph_sig[4] = {0, 0, 0, 0};
  int graph_slot = 0;
  double ubulk = 0.0;        // bulk-velocity target of the constant-mass-flux correction
  // FENE-P conformation tensor: independent components 11 12 13 22 23 33
  double2 *C[6] = {}, *CN[6] = {}, *CNM1[6] = {}, *STRN[6] = {}, *STRNM1[6] = {}, *TF[3] = {};
  double2* tabC = nullptr;
  void* NLP[12] = {};   // plane outputs c_n, str_n (physical y, transfer precision)
  void* TP[12] = {};    // transfer C, D C
  void* TPD[3] = {};    // transfer D u, D v, D w
  void *ZP = nullptr, *ZFP = nullptr;
  struct PolyFlags { bool init = false, solve = false, evaluate = false, force = false, copy = false; };
  PolyFlags pprev;      // flags of the evaluation whose results sit in NLP
  // learned closure: physical captures of the last evaluations (FP32)
  bool poly_capture = false;
  float* hist[3] = {nullptr, nullptr, nullptr};  // [6][Pphys] stress, slot it % 3
  float* capg = nullptr;                          // [9][Pphys] du_i/dx_k of the last evaluation
  int hist_n = 0, hist_it = -1000;                // consecutive captured evaluations, last one
  kmm_mlp* closure = nullptr;                     // network replacing the conformation transport
  bool nlp_closure = false;                       // NLP holds only a closure stress (no c_n)
  // audits of the closure: windows of full FENE-P transport every audit_every
  // steps, restarted from C rebuilt from the closure's stress
  bool closure_active = false;                    // false after a failed audit (FENE-P fallback)
  int closure_it0 = 0;                            // first step of the current closure period
  int audit_every = 0, audit_window = 0, audit_w = -1;  // audit_w: position in a running window
  double audit_gate = 0.0, audit_last = -1.0;
  int audits = 0, fallback_it = -1;
  float* audit_pred = nullptr;                    // [6][Pphys] network prediction in a window
  float* conf_phys = nullptr;                     // [6][Pphys] rebuilt C on the 3/2 grid
  void* conf_nlp[6] = {};                         // its x-z transforms (physical y)
  double* audit_part = nullptr;                   // residual partial sums
  double* audit_acc = nullptr;                    // [2] residual sums of the running window
  // retraining at the end of each passed audit window on pairs drawn from the
  // window's FENE-P evaluations (kmm_set_closure_retrain)
  int retrain_steps = 0, retrain_batch = 0;
  float retrain_lr = 0.0f;
  unsigned long long retrain_seed = 0;
  float *rX = nullptr, *rT = nullptr;
  // per-routine slab solves (kmm_solve/kmm_penta): one kx slab and its tables,
  // allocated at the first call and kept until kmm_destroy
  void* slab = nullptr;
  // per-routine entries (kmm_phirhs ...): 9 FP64 spectral fields and four
  // mean profiles, allocated at the first call and kept until kmm_destroy
  double2* rt = nullptr;
  // batched ensemble launches (kmm_ensemble_step): the argument blocks of
  // the steps of one chunk for every channel of the group this context leads
  void* batch_args = nullptr;
  size_t batch_cap = 0;
  std::vector<char> batch_host;
  long long batched_steps = 0;  // steps this channel ran in shared launches
};

static void drop_graph(kmm_ctx* c);
int kmm_mlp_train_device(kmm_mlp* n, int npts, const float* dX, const float* dT, int steps, int batch, float lr,
                         hipStream_t s);  // kmm_mlp_api.cpp
int kmm_mlp_activation(const kmm_mlp* m);  // kmm_mlp_api.cpp
int kmm_mlp_last_loss(kmm_mlp* m, float* loss, hipStream_t s);  // kmm_mlp_api.cpp

static int dalloc(kmm_ctx* c, void** p, size_t bytes) {
  KMM_HIP(hipMalloc(p, bytes > 0 ? bytes : 16));
  KMM_HIP(hipMemsetAsync(*p, 0, bytes > 0 ? bytes : 16, c->stream));
  c->allocs.push_back(*p);
  return KMM_OK;
}
#define DALLOC(ptr, bytes)                                                  \
  do {                                                                      \
    int r_ = dalloc(c, (void**)&(ptr), (bytes));                            \
    if (r_) return r_;                                                      \
  } while (0)

extern "C" const char* kmm_status_string(int s) {
  switch (s) {
    case KMM_OK: return "ok";
    case KMM_E_CFL: return g_last_error.empty() ? "CFL failure" : g_last_error.c_str();
    case KMM_E_PARTICLE: return "particle out of domain";
    case KMM_E_HIP: return g_last_error.empty() ? "HIP error" : g_last_error.c_str();
    case KMM_E_CONFIG: return g_last_error.empty() ? "configuration error" : g_last_error.c_str();
    case KMM_E_IO: return g_last_error.empty() ? "I/O error" : g_last_error.c_str();
    case KMM_E_ARG: return g_last_error.empty() ? "bad argument" : g_last_error.c_str();
    default: return "unknown";
  }
}

static int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

// shared with the other translation units of the library
const char* kmm_last_error_set(const std::string& msg) {
  g_last_error = msg;
  return g_last_error.c_str();
}

// ----------------------------------------------------------- dns.config --
static bool parse_num(const std::string& line, int nvals, double* out) {
  std::string s = line;
  for (auto& ch : s)
    if (ch == 'd' || ch == 'D') ch = 'e';
  std::istringstream is(s);
  for (int i = 0; i < nvals; ++i) {
    std::string tok;
    if (!(is >> tok)) return false;
    // list-directed input also accepts a trailing comma separator
    if (!tok.empty() && tok.back() == ',') tok.pop_back();
    char* end = nullptr;
    out[i] = std::strtod(tok.c_str(), &end);
    if (end == tok.c_str()) return false;
  }
  return true;
}

extern "C" int kmm_config_read(const char* path, kmm_config* cfg) {
  if (!path || !cfg) return fail(KMM_E_ARG, "null argument");
  std::ifstream f(path);
  if (!f) return fail(KMM_E_IO, std::string("cannot open ") + path);
  std::vector<std::string> lines;
  std::string ln;
  while (std::getline(f, ln)) lines.push_back(ln);
  if (lines.size() < 75) return fail(KMM_E_IO, "dns.config has fewer than 75 lines");
  auto get = [&](int lineno, double* v, int n = 1) -> bool { return parse_num(lines[lineno - 1], n, v); };
  double v[3];
  std::memset(cfg, 0, sizeof(*cfg));
#define RD(line, dst, T)                                                        \
  do {                                                                          \
    if (!get(line, v)) return fail(KMM_E_IO, "dns.config line " #line);          \
    cfg->dst = (T)v[0];                                                         \
  } while (0)
  RD(4, nsteps, int); RD(5, iprnfrq, int); RD(6, dt, double); RD(7, gain, double); RD(8, ugain, double);
  RD(9, theta, double); RD(10, ny, int); RD(11, nz, int); RD(12, nx, int); RD(13, bftail, int);
  RD(14, yl, double); RD(18, xstart, double);
  RD(15, zl, double); RD(16, xl, double); RD(17, re, double); RD(19, uinf, double); RD(20, kwall, int);
  RD(21, perturbtime, int); RD(22, r_tau, double); RD(23, gravity, double); RD(29, irstrt, int);
  RD(30, crstrt, int);
  RD(34, flow_select, int); RD(38, scl_flag, int); RD(63, diff, double); RD(64, deltaT, double);
  RD(66, src_start, int); RD(67, src_stop, int); RD(74, cadence, int);
  {
    double sg[3];
    if (!get(65, sg, 3)) return fail(KMM_E_IO, "dns.config line 65");
    cfg->sigmax = sg[0]; cfg->sigmay = sg[1]; cfg->sigmaz = sg[2];
  }
  RD(33, particle_flag, int); RD(46, ratio, double); RD(47, ap, double); RD(48, c_mu, double);
  RD(49, cd_switch, int);
  RD(35, itarget, int); RD(36, ipeter, int); RD(37, ipolyflag, int);
  RD(55, zlmax, double); RD(56, tpoly, double); RD(57, alpha_poly, double); RD(58, diffpoly, double);
  RD(59, c11z, double); RD(60, c22z, double); RD(61, c33z, double); RD(62, qbeta, double);
#undef RD
  if (cfg->src_stop < cfg->src_start) cfg->src_stop = cfg->nsteps;  // main_dns.f90:1071-1075
  cfg->scalar = 0;
  cfg->polymer = 0;
  cfg->particle_swirl = 1;  // part_track's caller always interpolates swirl_3d
  return KMM_OK;
}

// ----------------------------------------------------------- layout ----
// reference (nyp,nz,nxh) column-major <-> device [y][kx][kz]
static void ref_to_dev(const kmm_ctx* c, const double* ref, std::vector<double2>& dev) {
  dev.resize(c->S);
  for (int k = 0; k < c->nxh; ++k)
    for (int j = 0; j < c->nz; ++j)
      for (int i = 0; i < c->nyp; ++i) {
        const size_t r = (size_t)i + (size_t)c->nyp * ((size_t)j + (size_t)c->nz * k);
        dev[((size_t)i * c->nxh + k) * c->nz + j] = make_double2(ref[2 * r], ref[2 * r + 1]);
      }
}
static void dev_to_ref(const kmm_ctx* c, const std::vector<double2>& dev, double* ref) {
  for (int k = 0; k < c->nxh; ++k)
    for (int j = 0; j < c->nz; ++j)
      for (int i = 0; i < c->nyp; ++i) {
        const size_t r = (size_t)i + (size_t)c->nyp * ((size_t)j + (size_t)c->nz * k);
        const double2 v = dev[((size_t)i * c->nxh + k) * c->nz + j];
        ref[2 * r] = v.x;
        ref[2 * r + 1] = v.y;
      }
}
static int upload_field(kmm_ctx* c, double2* d, const double* ref) {
  std::vector<double2> h;
  ref_to_dev(c, ref, h);
  KMM_HIP(hipMemcpyAsync(d, h.data(), c->S * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  KMM_HIP(hipStreamSynchronize(c->stream));
  return KMM_OK;
}
static int download_field(kmm_ctx* c, const double2* d, double* ref) {
  std::vector<double2> h(c->S);
  KMM_HIP(hipMemcpyAsync(h.data(), d, c->S * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
  KMM_HIP(hipStreamSynchronize(c->stream));
  dev_to_ref(c, h, ref);
  return KMM_OK;
}

// fields stored pencil-major on the device, [kx][kz][y]: the state of the
// step (v, omega_y, theta and their nonlinear-term histories, the FENE-P
// conformation and its histories), read and written only by the pencil
// kernels; the reference's own order of a spectral field, so they move as
// plain copies
static bool pencil_major(const kmm_ctx* c, int f) {
  switch (f) {
    case KMM_V: case KMM_FN: case KMM_FNM1: case KMM_OMY: case KMM_GN: case KMM_GNM1:
    case KMM_SCALAR: case KMM_SCN: case KMM_SCNM1:
      return true;
    default:
      return c->cfg.polymer && f >= KMM_C11 && f <= KMM_STR33NM1;
  }
}
static int upload_pm(kmm_ctx* c, double2* d, const double* ref) {
  KMM_HIP(hipMemcpyAsync(d, ref, c->S * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  KMM_HIP(hipStreamSynchronize(c->stream));
  return KMM_OK;
}
static int download_pm(kmm_ctx* c, const double2* d, double* ref) {
  KMM_HIP(hipMemcpyAsync(ref, d, c->S * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
  KMM_HIP(hipStreamSynchronize(c->stream));
  return KMM_OK;
}

static std::vector<double2> twiddles(int n) {
  std::vector<double2> t(n);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int k = 0; k < n; ++k) {
    const long double a = -2.0L * pi * (long double)k / (long double)n;
    t[k] = make_double2((double)cosl(a), (double)sinl(a));
  }
  return t;
}

static std::vector<float2> twiddles_f(int n) {
  std::vector<double2> d = twiddles(n);
  std::vector<float2> f(n);
  for (int k = 0; k < n; ++k) f[k] = make_float2((float)d[k].x, (float)d[k].y);
  return f;
}

static bool fft_len_ok(int n) {
  int m = n;
  while (m % 2 == 0) m /= 2;
  while (m % 3 == 0) m /= 3;
  return m == 1;
}

// -------------------------------------------------------------- create --
static PencilArgs pencil_args(kmm_ctx* c) {
  PencilArgs a;
  std::memset(&a, 0, sizeof(a));
  a.nxh = c->nxh; a.nz = c->nz; a.nzh = c->nzh; a.P = c->P; a.ny = c->ny;
  a.tune = c->tune_pencil;
  a.alpha = c->alpha; a.beta = c->beta; a.dyde = c->dyde; a.re = c->cfg.re; a.dt = c->cfg.dt;
  a.theta = c->cfg.theta; a.diff = c->cfg.diff; a.gH = c->gH; a.gS = c->gS;
  // setbcs (main_dns.f90:1240-1265): no-slip for channel / Couette
  a.atu = 1.0; a.btu = 0.0; a.gtu = 0.0; a.abu = 1.0; a.bbu = 0.0; a.gbu = 0.0;
  a.atw = 1.0; a.btw = 0.0; a.gtw = 0.0; a.abw = 1.0; a.bbw = 0.0; a.gbw = 0.0;
  if (c->cfg.flow_select == 2 || c->cfg.flow_select == 12 || c->cfg.flow_select == 3) a.gtu = c->cfg.uinf;
  if (c->cfg.flow_select == 3) a.gtw = c->cfg.uinf;
  a.scalar = c->cfg.scalar;
  a.mixed = c->cfg.precision == 1;
  a.smooth = 1;
  a.first_nl = 1;
  a.smz = c->smz; a.smy = c->smy; a.twM = c->twM;
  a.tabH = c->tabH; a.tabP = c->tabP; a.tabS = c->tabS; a.G1 = c->G1; a.G2 = c->G2; a.gk = c->gk;
  a.NL1 = c->NL[0]; a.NL2 = c->NL[1]; a.NL3 = c->NL[2]; a.NLS = c->NL[3];
  a.V = c->V; a.OMY = c->OMY; a.FN = c->FN; a.FNM1 = c->FNM1; a.GN = c->GN; a.GNM1 = c->GNM1;
  a.TH = c->TH; a.SCN = c->SCN; a.SCNM1 = c->SCNM1;
  a.u0 = c->u0; a.w0 = c->w0; a.h1n = c->h1n; a.h1nm1 = c->h1nm1; a.h3n = c->h3n; a.h3nm1 = c->h3nm1;
  for (int t = 0; t < 9; ++t) a.T[t] = c->T[t];
  a.cmf = (c->cfg.cmf && c->cfg.flow_select == 1) ? 1 : 0;
  a.ubulk = c->ubulk;
  a.dpdx = c->dpdx_d;
  return a;
}

static PlaneArgs plane_args(kmm_ctx* c, int it, unsigned long long* cfl) {
  PlaneArgs a;
  std::memset(&a, 0, sizeof(a));
  a.nxh = c->nxh; a.nz = c->nz; a.nzh = c->nzh; a.nyp = c->nyp; a.mx = c->mx; a.mz = c->mz; a.P = c->P;
  a.nfin = c->cfg.scalar ? 9 : 6;
  a.nfout = c->cfg.scalar ? 4 : 3;
  a.tune = c->tune_xplane;
  a.mixed = c->cfg.precision == 1;
  a.dt = c->cfg.dt; a.dPdx = c->dPdx; a.dpdx_d = c->dpdx_d;
  a.delxm = c->cfg.xl / c->mx; a.delzm = c->cfg.zl / c->mz;
  a.src_uniform = (c->cfg.scalar && c->cfg.scl_flag == 6) ? c->cfg.deltaT : 0.0;
  a.flow_select = c->cfg.flow_select;
  a.seght = c->seght_d; a.twZ = c->twZ; a.twX = c->twX;
  for (int f = 0; f < 9; ++f) a.in[f] = c->T[f];
  a.Z = c->Z; a.ZF = c->ZF;
  for (int o = 0; o < 4; ++o) a.out[o] = c->NL[o];
  a.cfl = cfl;
  a.nsrc = c->nsrc; a.src = c->src; a.srcp = c->srcp; a.fint = c->fint; a.drag = c->drag;
  a.gain = c->cfg.gain; a.ugain = c->cfg.ugain; a.it = it; a.perturbtime = c->cfg.perturbtime;
  a.trow = c->trow; a.tx = c->tx; a.tptr = c->tptr; a.ent_src = c->ent_src; a.ent_w = c->ent_w;
  a.fixrow = c->fixrow; a.fixx = c->fixx;
  return a;
}

extern "C" int kmm_create(const kmm_config* cfg, int device, kmm_ctx** out) {
  if (!cfg || !out) return fail(KMM_E_ARG, "null argument");
  *out = nullptr;
  const int nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
  if (nx <= 0 || ny <= 0 || nz <= 0 || nx % 2 || ny % 2 || nz % 2)
    return fail(KMM_E_CONFIG, "grid sizes must be positive and even");
  if (!(ny == 32 || ny == 48 || ny == 64 || ny == 80 || ny == 96 || ny == 128 || ny == 192))
    return fail(KMM_E_CONFIG, "ny must be one of 32, 48, 64, 80, 96, 128, 192");
  if ((nx / 2) % 8 != 0) return fail(KMM_E_CONFIG, "nx/2 must be a multiple of 8");
  if (!fft_len_ok(3 * nx / 2) || !fft_len_ok(3 * nz / 2)) return fail(KMM_E_CONFIG, "3/2 grid not 2^a 3^b");
  if (cfg->theta > 0.99) return fail(KMM_E_CONFIG, "THETA > 0.99 is not supported");
  if (cfg->precision != 0 && cfg->precision != 1) return fail(KMM_E_CONFIG, "precision must be 0 (FP64) or 1 (mixed)");
  if (!(cfg->flow_select == 1 || cfg->flow_select == 2 || cfg->flow_select == 3 || cfg->flow_select == 11 ||
        cfg->flow_select == 12))
    return fail(KMM_E_CONFIG, "flow_select must be a wall-bounded case with nt=nb=1 (1,2,3,11,12)");
  if (cfg->scalar && !(cfg->scl_flag == 0 || cfg->scl_flag == 1 || cfg->scl_flag == 2 || cfg->scl_flag == 3 ||
                      cfg->scl_flag == 6))
    return fail(KMM_E_CONFIG, "scl_flag must be 0, 1, 2 (blobs at the particles), 3 or 6 (uniform source)");
  if (cfg->scalar && cfg->scl_flag == 2 && !(cfg->sigmax > 0.0 && cfg->sigmay > 0.0 && cfg->sigmaz > 0.0))
    return fail(KMM_E_CONFIG, "scl_flag 2 needs positive sigmax, sigmay, sigmaz");
  if (cfg->polymer) {
    if (cfg->itarget != 0) return fail(KMM_E_CONFIG, "polymer: only itarget 0 (uniform beta) is supported");
    if (!(cfg->tpoly > 0.0) || !(cfg->diffpoly > 0.0) || !(cfg->qbeta > 0.0 && cfg->qbeta <= 1.0))
      return fail(KMM_E_CONFIG, "polymer: tpoly, diffpoly must be positive and 0 < qbeta <= 1");
    if (cfg->ipeter != 0 && !(cfg->zlmax * cfg->zlmax > 3.0))
      return fail(KMM_E_CONFIG, "polymer: zlmax^2 must exceed 3 for the Peterlin function");
  }
  if (cfg->cmf && cfg->flow_select != 1) return fail(KMM_E_CONFIG, "constant mass flux needs flow_select 1 (channel)");
  kmm_ctx* c = new kmm_ctx();
  c->cfg = *cfg;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) { delete c; return fail(KMM_E_HIP, "hipSetDevice failed"); }
  if (hipStreamCreate(&c->stream) != hipSuccess) { delete c; return fail(KMM_E_HIP, "hipStreamCreate failed"); }
  for (auto& e : c->ev) hipEventCreate(&e);
  c->nx = nx; c->ny = ny; c->nz = nz; c->nyp = ny + 1; c->nxh = nx / 2; c->nzh = nz / 2;
  c->mx = 3 * nx / 2; c->mz = 3 * nz / 2; c->P = c->nxh * nz;
  c->S = (size_t)c->nyp * c->P;
  const double pi = 2.0 * std::acos(0.0);
  c->alpha = 2.0 * pi / cfg->xl;
  c->beta = 2.0 * pi / cfg->zl;
  c->dyde = 2.0 / cfg->yl;
  c->dPdx = -2.0 / cfg->yl * std::pow(2.0 * cfg->r_tau / (cfg->yl * cfg->re), 2);
  c->gH = cfg->re / (cfg->dt * cfg->theta);
  c->gS = cfg->re * cfg->diff / (cfg->dt * cfg->theta);
  c->ycoord.resize(c->nyp);
  c->seght.resize(c->nyp);
  c->ycoord[0] = cfg->yl / 2.0;
  for (int k = 1; k < c->nyp; ++k) {
    c->ycoord[k] = std::cos((double)k * pi / (double)ny) * (cfg->yl / 2.0);
    c->seght[k] = std::fabs(c->ycoord[k] - c->ycoord[k - 1]);
  }
  c->seght[0] = c->seght[c->nyp - 1];
  int r = 0;
  {
    const size_t fb = c->S * sizeof(double2);
    double2** st[] = {&c->V, &c->OMY, &c->FN, &c->FNM1, &c->GN, &c->GNM1, &c->TH, &c->SCN, &c->SCNM1};
    for (auto p : st) if ((r = dalloc(c, (void**)p, fb))) goto fail_;
    for (auto p : {&c->u0, &c->w0, &c->h1n, &c->h1nm1, &c->h3n, &c->h3nm1})
      if ((r = dalloc(c, (void**)p, c->nyp * sizeof(double)))) goto fail_;
    c->vsz = (cfg->precision == 1) ? sizeof(float2) : sizeof(double2);
    const size_t tbv = c->S * c->vsz;
    for (int i = 0; i < 4; ++i) if ((r = dalloc(c, (void**)&c->NL[i], tbv))) goto fail_;
    for (int i = 0; i < 9; ++i) if ((r = dalloc(c, (void**)&c->T[i], tbv))) goto fail_;
    const size_t zb = (size_t)c->nyp * c->mz * c->nxh * c->vsz;
    if ((r = dalloc(c, (void**)&c->Z, 9 * zb))) goto fail_;
    if ((r = dalloc(c, (void**)&c->ZF, 4 * zb))) goto fail_;
    const size_t tb = (size_t)c->P * kmm::tab_width(ny) * sizeof(double2);
    for (auto p : {&c->tabH, &c->tabP, &c->tabS}) if ((r = dalloc(c, (void**)p, tb))) goto fail_;
    const size_t gb = (size_t)c->P * (ny / 2 + 1) * sizeof(double2);
    for (auto p : {&c->G1, &c->G2}) if ((r = dalloc(c, (void**)p, gb))) goto fail_;
    if ((r = dalloc(c, (void**)&c->gk, (size_t)c->P * sizeof(double4)))) goto fail_;
    if (cfg->polymer) {
      for (int q = 0; q < 6; ++q)
        for (double2** b : {&c->C[q], &c->CN[q], &c->CNM1[q], &c->STRN[q], &c->STRNM1[q]})
          if ((r = dalloc(c, (void**)b, fb))) goto fail_;
      for (int q = 0; q < 3; ++q) if ((r = dalloc(c, (void**)&c->TF[q], fb))) goto fail_;
      for (int q = 0; q < 12; ++q) {
        if ((r = dalloc(c, &c->NLP[q], tbv))) goto fail_;
        if ((r = dalloc(c, &c->TP[q], tbv))) goto fail_;
      }
      for (int q = 0; q < 3; ++q) if ((r = dalloc(c, &c->TPD[q], tbv))) goto fail_;
      if ((r = dalloc(c, &c->ZP, 24 * zb))) goto fail_;
      if ((r = dalloc(c, &c->ZFP, 12 * zb))) goto fail_;
      if ((r = dalloc(c, (void**)&c->tabC, tb))) goto fail_;
    }
    if (cfg->scalar && cfg->scl_flag == 2)
      if ((r = dalloc(c, (void**)&c->scsrc, (size_t)c->nyp * c->mz * c->mx * sizeof(double)))) goto fail_;
    if ((r = dalloc(c, (void**)&c->cfl_d, 4096 * sizeof(unsigned long long)))) goto fail_;
    c->cfl_cap = 4096;
    // twiddles and smoothing
    auto up = [&](void** d, const void* h, size_t b) -> int {
      int rr = dalloc(c, d, b);
      if (rr) return rr;
      KMM_HIP(hipMemcpyAsync(*d, h, b, hipMemcpyHostToDevice, c->stream));
      return KMM_OK;
    };
    auto tM = twiddles(2 * ny), tZ = twiddles(c->mz), tX = twiddles(c->mx);
    if (cfg->precision == 1) {
      auto fM = twiddles_f(2 * ny), fZ = twiddles_f(c->mz), fX = twiddles_f(c->mx);
      if ((r = up((void**)&c->twM, fM.data(), fM.size() * sizeof(float2)))) goto fail_;
      if ((r = up((void**)&c->twZ, fZ.data(), fZ.size() * sizeof(float2)))) goto fail_;
      if ((r = up((void**)&c->twX, fX.data(), fX.size() * sizeof(float2)))) goto fail_;
    } else {
      if ((r = up((void**)&c->twM, tM.data(), tM.size() * sizeof(double2)))) goto fail_;
      if ((r = up((void**)&c->twZ, tZ.data(), tZ.size() * sizeof(double2)))) goto fail_;
      if ((r = up((void**)&c->twX, tX.data(), tX.size() * sizeof(double2)))) goto fail_;
    }
    std::vector<double> smz(nz), smy(c->nyp);
    for (int j = 0; j < nz; ++j) {  // main_dns.f90:712-717
      const int jj = (j < c->nzh) ? j : nz - j;
      smz[j] = std::exp(-3.0 * std::pow((double)jj / (double)c->nzh, 10));
    }
    for (int i = 0; i < c->nyp; ++i) smy[i] = std::exp(-3.0 * std::pow((double)i / (double)ny, 10));
    if ((r = up((void**)&c->smz, smz.data(), nz * sizeof(double)))) goto fail_;
    if ((r = up((void**)&c->smy, smy.data(), c->nyp * sizeof(double)))) goto fail_;
    if ((r = up((void**)&c->seght_d, c->seght.data(), c->nyp * sizeof(double)))) goto fail_;
    if ((r = up((void**)&c->ycoord_d, c->ycoord.data(), c->nyp * sizeof(double)))) goto fail_;
    if ((r = up((void**)&c->dpdx_d, &c->dPdx, sizeof(double)))) goto fail_;
    c->ubulk = cfg->ubulk;
    if ((r = dalloc(c, (void**)&c->perr, sizeof(int)))) goto fail_;
    for (int d = 0; d < 3; ++d) c->fint_full[d].assign((size_t)c->nyp * c->mz * c->mx, 0.0);
    // tau tables and Green's functions (gfcn, main_dns.f90:151-161)
    PencilArgs a = pencil_args(c);
    hipError_t e;
    if ((e = kmm::launch_tau_tables(ny, a, c->gH, c->tabH, c->stream)) != hipSuccess) { r = fail(KMM_E_HIP, hipGetErrorString(e)); goto fail_; }
    if ((e = kmm::launch_tau_tables(ny, a, 0.0, c->tabP, c->stream)) != hipSuccess) { r = fail(KMM_E_HIP, hipGetErrorString(e)); goto fail_; }
    if ((e = kmm::launch_tau_tables(ny, a, c->gS, c->tabS, c->stream)) != hipSuccess) { r = fail(KMM_E_HIP, hipGetErrorString(e)); goto fail_; }
    if (cfg->polymer &&
        (e = kmm::launch_tau_tables(ny, a, cfg->re * cfg->diffpoly / (cfg->dt * cfg->theta), c->tabC, c->stream)) != hipSuccess) {
      r = fail(KMM_E_HIP, hipGetErrorString(e));
      goto fail_;
    }
    // p1b=1, p1t=-1, p2b=-1, p2t=0.5 (setbcs, main_dns.f90:1406-1409)
    if ((e = kmm::launch_green(ny, a, 1.0, -1.0, -1.0, 0.5, c->G1, c->G2, c->gk, c->stream)) != hipSuccess) { r = fail(KMM_E_HIP, hipGetErrorString(e)); goto fail_; }
    if (hipStreamSynchronize(c->stream) != hipSuccess) { r = fail(KMM_E_HIP, "setup kernels failed"); goto fail_; }
  }
  *out = c;
  return KMM_OK;
fail_:
  kmm_destroy(c);
  return r;
}

static int part_free(kmm_ctx* c);

extern "C" int kmm_destroy(kmm_ctx* c) {
  if (!c) return KMM_OK;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->gexec) hipGraphExecDestroy(c->gexec);
  if (c->graph) hipGraphDestroy(c->graph);
  for (void* p : c->allocs) hipFree(p);
  part_free(c);
  for (auto& e : c->ev) if (e) hipEventDestroy(e);
  for (auto& e : c->tev) hipEventDestroy(e);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return KMM_OK;
}

// ---------------------------------------------------------- geometry ----
// Solid cells (imatrix 1 or 4) carry Goldstein feedback; their force is
// spread with exp(-2 dsq) over a 5x5 stencil (main_dns.f90:1550-1558) on the
// planes i <= kmaxsurf (quirk: kmaxsurf is the largest solid x index,
// main_dns.f90:1526-1535).  The reference walks the (mz,mx) plane
// column-major and assigns the product at each point before spreading, so a
// contribution survives only at linear addresses not after its source
// (quirk Q6); z offsets are not wrapped (quirk Q5).
extern "C" int kmm_set_geometry(kmm_ctx* c, const int32_t* im) {
  if (c) drop_graph(c);  // the IB tables are reallocated
  if (!c || !im) return fail(KMM_E_ARG, "null argument");
  const int nyp = c->nyp, mz = c->mz, mx = c->mx;
  c->imatrix_h.assign(im, im + (size_t)nyp * mz * mx);
  auto at = [&](int i, int j, int k) { return im[(size_t)i + (size_t)nyp * ((size_t)j + (size_t)mz * k)]; };
  int kmax = 0;
  std::vector<int4> src;
  bool direct = false, moving = false;
  for (int k = 0; k < mx; ++k)
    for (int j = 0; j < mz; ++j)
      for (int i = 0; i < nyp; ++i) {
        const int t = at(i, j, k);
        if (!(t == 0 || t == 1 || t == 3 || t == 4 || t == 6 || t == 7 || t == 8 || t == 11))
          return fail(KMM_E_CONFIG, "imatrix types are 0, 1, 3, 4, 6, 7, 8, 11 (solvers.f90:2996-3043)");
        if (t == 1 || t == 3 || t == 4) kmax = std::max(kmax, k + 1);  // kmaxsurf, main_dns.f90:1526-1535
        if (t == 6 || t == 7) direct = true;
        if (t == 8) moving = true;
      }
  // forcepre (main_dns.f90:1490-1546): buffer gains, spanwise wall slope and
  // the suction profile, with its x spacing xl/(mx-1)
  const int bfhead = 2, bfwidth = c->cfg.bftail - bfhead, bfwh = bfwidth / 2;
  const double slope = bfwidth / 2.0;
  if (moving && !(slope > 0.0)) return fail(KMM_E_CONFIG, "imatrix 8 needs BFTAIL > 2 (slope length bfwidth/2)");
  if (direct && !(c->cfg.flow_select == 1 || c->cfg.flow_select == 2))
    return fail(KMM_E_CONFIG, "imatrix 6/7 need the initial profile of flow_select 1 or 2");
  std::vector<double> bfgain(1200, 0.0), bfugain(1200, 0.0);
  for (int i = 1; i <= bfwidth + 1 && i <= 1200; ++i) {
    const int ib = i - (bfwh + 1);
    if (bfwh > 0 && std::abs(ib) <= bfwh / 3) bfgain[i - 1] = 20.0 * std::exp(-60.0 * std::pow((double)ib / bfwh, 2));
    if (bfwh > 0) bfugain[i - 1] = 20.0 * std::exp(-6.0 * std::pow((double)ib / bfwh, 2));
  }
  const double delxf = c->cfg.xl / (double)(mx - 1), xsuction = delxf * (c->cfg.bftail - 1);
  auto vdes = [&](int i1) {  // indexed by the y index, as vcw3dp does (solvers.f90:3034)
    const double x = delxf * (i1 - 1);
    return x <= xsuction ? 0.0 : (-0.8604 * c->cfg.uinf) / std::sqrt(c->cfg.re * c->cfg.uinf * (x + c->cfg.xstart - xsuction));
  };
  // initial profile of init_flow (x and z independent): channel or Couette
  std::vector<double> initu(nyp, 0.0);
  if (direct) {
    const double yl = c->cfg.yl, ymin = c->ycoord[nyp - 1];
    for (int i = 0; i < nyp; ++i) {
      const double d = std::fabs(c->ycoord[i] - ymin);
      if (c->cfg.flow_select == 2) initu[i] = c->cfg.uinf * d / std::fabs(c->ycoord[0] - ymin);
      else if (c->cfg.flow_select == 1 && !(i + 1 <= c->cfg.kwall || nyp - (i + 1) <= c->cfg.kwall))
        initu[i] = c->cfg.re * c->dPdx * (0.5 * d * d - 0.5 * yl * d);
    }
  }
  // sources ordered by (y, linear in-plane address), with per-type constants
  std::vector<double> par;   // [6] per source
  for (int i = 0; i < nyp; ++i)
    for (int k = 0; k < mx; ++k)
      for (int j = 0; j < mz; ++j) {
        const int t = at(i, j, k);
        if (t == 0) continue;
        double q[6] = {0, 0, 0, 0, 0, 0};
        if (t == 6) {  // bfugain, bfgain at k - bfhead + 1, initu, initv, initw at (i, j, k)
          const int b = (k + 1) - bfhead + 1;
          if (b < 1 || b > 1200) return fail(KMM_E_CONFIG, "imatrix 6 cell outside the buffer gains (x index < 2)");
          q[0] = bfugain[b - 1]; q[1] = bfgain[b - 1]; q[2] = initu[i];
        } else if (t == 7) {  // vdes(i), Uinf, initw
          q[0] = vdes(i + 1); q[1] = c->cfg.uinf;
        } else if (t == 8) {  // wdes at x index k
          const int k1 = k + 1;
          q[0] = (k1 <= bfwidth * 3 + 1) ? (0.3 * c->cfg.uinf) / slope * (k1 - (bfwidth * 3 + 1 - slope))
                                         : 0.3 * c->cfg.uinf;
        }
        src.push_back(make_int4(i, j, k, t));
        par.insert(par.end(), q, q + 6);
      }
  struct Ent { int y, z, x, s; double w; };
  std::vector<Ent> ents;
  for (size_t s = 0; s < src.size(); ++s) {
    const int i = src[s].x, j = src[s].y, k = src[s].z;
    if (src[s].w == 6 || src[s].w == 7) {
      // buffer and suction cells add their force at the cell itself, after
      // its product is assigned, and spread nothing
      ents.push_back({i, j, k, (int)s, 1.0});
      continue;
    }
    if (i + 1 > kmax) continue;
    const long ls = (long)k * mz + j;
    for (int ii = -2; ii <= 2; ++ii) {
      int ip = k + ii;
      if (ip < 0) ip += mx;
      if (ip >= mx) ip -= mx;
      for (int jj = -2; jj <= 2; ++jj) {
        const long lt = (long)ip * mz + (j + jj);
        if (lt < 0 || lt >= (long)mz * mx)
          return fail(KMM_E_CONFIG, "IB spreading leaves the (mz,mx) plane (solid cell at the x=1/mx edge)");
        if (lt > ls) continue;
        const double w = std::exp(-2.0 * (double)(ii * ii + jj * jj));
        ents.push_back({i, (int)(lt % mz), (int)(lt / mz), (int)s, w});
      }
    }
  }
  std::stable_sort(ents.begin(), ents.end(), [](const Ent& a, const Ent& b) {
    if (a.y != b.y) return a.y < b.y;
    if (a.z != b.z) return a.z < b.z;
    return a.x < b.x;
  });
  std::vector<int> trow((size_t)nyp * mz + 1, 0), tx, tptr, esrc;
  std::vector<double> ew;
  {
    size_t q = 0;
    for (int row = 0; row < nyp * mz; ++row) {
      trow[row] = (int)tx.size();
      const int y = row / mz, z = row % mz;
      while (q < ents.size() && ents[q].y == y && ents[q].z == z) {
        const int x = ents[q].x;
        tx.push_back(x);
        tptr.push_back((int)esrc.size());
        while (q < ents.size() && ents[q].y == y && ents[q].z == z && ents[q].x == x) {
          esrc.push_back(ents[q].s);
          ew.push_back(ents[q].w);
          ++q;
        }
      }
    }
    trow[(size_t)nyp * mz] = (int)tx.size();
    tptr.push_back((int)esrc.size());
  }
  std::vector<int> fixrow((size_t)nyp * mz + 1, 0), fixx;
  for (int row = 0; row < nyp * mz; ++row) {
    fixrow[row] = (int)fixx.size();
    const int y = row / mz, z = row % mz;
    for (int k = 0; k < mx; ++k)
      if (at(y, z, k) != 0) fixx.push_back(k);
  }
  fixrow[(size_t)nyp * mz] = (int)fixx.size();
  auto up = [&](void** d, const void* h, size_t b) -> int {
    int rr = dalloc(c, d, b);
    if (rr) return rr;
    if (b) KMM_HIP(hipMemcpyAsync(*d, h, b, hipMemcpyHostToDevice, c->stream));
    return KMM_OK;
  };
  int r;
  c->nsrc = (int)src.size();
  c->src_h = src;
  if ((r = up((void**)&c->src, src.data(), src.size() * sizeof(int4)))) return r;
  if ((r = up((void**)&c->srcp, par.data(), par.size() * sizeof(double)))) return r;
  if ((r = dalloc(c, (void**)&c->fint, 3 * src.size() * sizeof(double)))) return r;
  if ((r = dalloc(c, (void**)&c->drag, 3 * src.size() * sizeof(double)))) return r;
  if ((r = up((void**)&c->trow, trow.data(), trow.size() * sizeof(int)))) return r;
  if ((r = up((void**)&c->tx, tx.data(), tx.size() * sizeof(int)))) return r;
  if ((r = up((void**)&c->tptr, tptr.data(), tptr.size() * sizeof(int)))) return r;
  if ((r = up((void**)&c->ent_src, esrc.data(), esrc.size() * sizeof(int)))) return r;
  if ((r = up((void**)&c->ent_w, ew.data(), ew.size() * sizeof(double)))) return r;
  if ((r = up((void**)&c->fixrow, fixrow.data(), fixrow.size() * sizeof(int)))) return r;
  if ((r = up((void**)&c->fixx, fixx.data(), fixx.size() * sizeof(int)))) return r;
  // load integrals of the solid cells from the host copies
  std::vector<double> fi(3 * src.size());
  for (int d = 0; d < 3; ++d)
    for (size_t s = 0; s < src.size(); ++s) {
      const int4 q = src[s];
      fi[d * src.size() + s] = c->fint_full[d][(size_t)q.x + (size_t)nyp * ((size_t)q.y + (size_t)mz * q.z)];
    }
  if (!fi.empty()) KMM_HIP(hipMemcpyAsync(c->fint, fi.data(), fi.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  KMM_HIP(hipStreamSynchronize(c->stream));
  return KMM_OK;
}

static int sync_fint_to_host(kmm_ctx* c) {
  if (c->nsrc == 0) return KMM_OK;
  std::vector<double> fi(3 * (size_t)c->nsrc);
  KMM_HIP(hipMemcpyAsync(fi.data(), c->fint, fi.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  KMM_HIP(hipStreamSynchronize(c->stream));
  for (int d = 0; d < 3; ++d)
    for (int s = 0; s < c->nsrc; ++s) {
      const int4 q = c->src_h[s];
      c->fint_full[d][(size_t)q.x + (size_t)c->nyp * ((size_t)q.y + (size_t)c->mz * q.z)] = fi[(size_t)d * c->nsrc + s];
    }
  return KMM_OK;
}
static int sync_fint_to_dev(kmm_ctx* c) {
  if (c->nsrc == 0) return KMM_OK;
  std::vector<double> fi(3 * (size_t)c->nsrc);
  for (int d = 0; d < 3; ++d)
    for (int s = 0; s < c->nsrc; ++s) {
      const int4 q = c->src_h[s];
      fi[(size_t)d * c->nsrc + s] = c->fint_full[d][(size_t)q.x + (size_t)c->nyp * ((size_t)q.y + (size_t)c->mz * q.z)];
    }
  KMM_HIP(hipMemcpyAsync(c->fint, fi.data(), fi.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  KMM_HIP(hipStreamSynchronize(c->stream));
  return KMM_OK;
}

// ------------------------------------------------------------ fields ----
static double2* spectral_ptr(kmm_ctx* c, int f) {
  switch (f) {
    case KMM_V: return c->V;
    case KMM_FN: return c->FN;
    case KMM_FNM1: return c->FNM1;
    case KMM_OMY: return c->OMY;
    case KMM_GN: return c->GN;
    case KMM_GNM1: return c->GNM1;
    case KMM_SCALAR: return c->TH;
    case KMM_SCN: return c->SCN;
    case KMM_SCNM1: return c->SCNM1;
    default: break;
  }
  if (c->cfg.polymer && f >= KMM_C11 && f <= KMM_STR33NM1) {
    const int g = (f - KMM_C11) / 6, q = (f - KMM_C11) % 6;
    double2** tab[5] = {c->C, c->CN, c->CNM1, c->STRN, c->STRNM1};
    return tab[g][q];
  }
  switch (f) {
    default: return nullptr;
  }
}
static double* mean_ptr(kmm_ctx* c, int f) {
  switch (f) {
    case KMM_U0: return c->u0;
    case KMM_H1N: return c->h1n;
    case KMM_H1NM1: return c->h1nm1;
    case KMM_W0: return c->w0;
    case KMM_H3N: return c->h3n;
    case KMM_H3NM1: return c->h3nm1;
    default: return nullptr;
  }
}

static int materialize(kmm_ctx* c);

// fields the next step rebuilds from the NL buffers while they are valid
static bool nl_derived(int f) {
  return f == KMM_FN || f == KMM_GN || f == KMM_SCN || f == KMM_H1N || f == KMM_H3N ||
         (f >= KMM_C11N && f <= KMM_STR33NM1);
}

extern "C" int kmm_put_field(kmm_ctx* c, int f, const double* h) {
  if (!c || !h) return fail(KMM_E_ARG, "null argument");
  KMM_HIP(hipSetDevice(c->device));
  if (c->nl_valid && nl_derived(f)) {
    int r = materialize(c);
    if (r) return r;
    c->nl_valid = false;
  }
  if (double2* d = spectral_ptr(c, f)) return pencil_major(c, f) ? upload_pm(c, d, h) : upload_field(c, d, h);
  if (double* m = mean_ptr(c, f)) {
    KMM_HIP(hipMemcpyAsync(m, h, c->nyp * sizeof(double), hipMemcpyHostToDevice, c->stream));
    KMM_HIP(hipStreamSynchronize(c->stream));
    return KMM_OK;
  }
  if (f >= KMM_FXINTG && f <= KMM_FZINTG) {
    std::memcpy(c->fint_full[f - KMM_FXINTG].data(), h, c->fint_full[0].size() * sizeof(double));
    return sync_fint_to_dev(c);
  }
  return fail(KMM_E_ARG, "unknown field id");
}

extern "C" int kmm_get_field(kmm_ctx* c, int f, double* h) {
  if (!c || !h) return fail(KMM_E_ARG, "null argument");
  KMM_HIP(hipSetDevice(c->device));
  if (c->nl_valid && nl_derived(f)) {
    int r = materialize(c);
    if (r) return r;
  }
  if (double2* d = spectral_ptr(c, f)) return pencil_major(c, f) ? download_pm(c, d, h) : download_field(c, d, h);
  if (double* m = mean_ptr(c, f)) {
    KMM_HIP(hipMemcpyAsync(h, m, c->nyp * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    KMM_HIP(hipStreamSynchronize(c->stream));
    return KMM_OK;
  }
  if (f >= KMM_FXINTG && f <= KMM_FZINTG) {
    int r = sync_fint_to_host(c);
    if (r) return r;
    std::memcpy(h, c->fint_full[f - KMM_FXINTG].data(), c->fint_full[0].size() * sizeof(double));
    return KMM_OK;
  }
  return fail(KMM_E_ARG, "unknown field id");
}

// -------------------------------------------------------------- step ----
// Time-line of the POLYMER branches (main_dns.f90:206, 240, 637, 674, 682):
// init C at the start of step it; polyNL + implicit solve; polymer terms in
// vcw3dp; polymer force; the n -> n-1 copy before smoothing, which sits inside
// the force branch so that its it == src_start-1 arm never fires.
static kmm_ctx::PolyFlags poly_flags(const kmm_ctx* c, int it) {
  kmm_ctx::PolyFlags f;
  if (!c->cfg.polymer) return f;
  const kmm_config& g = c->cfg;
  f.init = (it == g.src_start - 1) || (it == g.irstrt && g.crstrt == 0);
  f.solve = it >= g.src_start;
  f.evaluate = it >= g.src_start - 1;
  f.force = it >= g.src_start;
  f.copy = f.force && f.init;
  return f;
}

static kmm::PolyArgs poly_args(kmm_ctx* c) {
  kmm::PolyArgs a;
  std::memset(&a, 0, sizeof(a));
  const kmm_config& g = c->cfg;
  a.nxh = c->nxh; a.nz = c->nz; a.nzh = c->nzh; a.P = c->P; a.ny = c->ny;
  a.alpha = c->alpha; a.beta = c->beta; a.dyde = c->dyde; a.re = g.re; a.dt = g.dt; a.theta = g.theta;
  a.diffpoly = g.diffpoly; a.tpoly = g.tpoly;
  a.zbeta1 = (1.0 - g.qbeta) / (g.re * g.qbeta * g.tpoly);   // solvers.f90:2971
  a.gC = g.re * g.diffpoly / (g.dt * g.theta);
  a.c0[0] = g.c11z; a.c0[1] = g.c22z; a.c0[2] = g.c33z;
  a.smz = c->smz; a.smy = c->smy; a.twM = c->twM; a.tabC = c->tabC;
  for (int q = 0; q < 12; ++q) { a.NLP[q] = c->NLP[q]; a.TP[q] = c->TP[q]; }
  for (int q = 0; q < 6; ++q) {
    a.C[q] = c->C[q]; a.CN[q] = c->CN[q]; a.CNM1[q] = c->CNM1[q]; a.STRN[q] = c->STRN[q]; a.STRNM1[q] = c->STRNM1[q];
  }
  for (int q = 0; q < 3; ++q) a.TF[q] = c->TF[q];
  return a;
}

// From this ny up, a FENE-P step runs its polymer pencil work as two passes
// (part B, then C init / solve / evaluate) instead of one fused pass.
#ifndef KMM_POLY_SPLIT_MIN_NY
#define KMM_POLY_SPLIT_MIN_NY 192
#endif

static int run_poly_pencil(kmm_ctx* c, int mode, const kmm::PolyArgs& a) {
  hipError_t e = kmm::launch_poly_pencil(c->ny, mode, a, c->cfg.precision == 1, c->stream);
  if (e != hipSuccess) return fail(KMM_E_HIP, std::string("polymer pencil: ") + hipGetErrorString(e));
  return KMM_OK;
}

static inline int hslot(int it) { return ((it % 3) + 3) % 3; }

static void hist_note(kmm_ctx* c, int it) {
  c->hist_n = (it == c->hist_it + 1) ? c->hist_n + 1 : 1;
  c->hist_it = it;
}

// learned closure at step it: velocity gradient on the 3/2 grid, the network
// on every point (inputs: gradient, wall distance, stress of it-1 and it-2),
// forward transforms of the predicted stress into NLP[6..11]
static int run_closure_plane(kmm_ctx* c, int it, bool swirl, double* mrf) {
  kmm::PolyPlaneArgs sa;
  std::memset(&sa, 0, sizeof(sa));
  sa.nxh = c->nxh; sa.nz = c->nz; sa.nzh = c->nzh; sa.nyp = c->nyp; sa.mx = c->mx; sa.mz = c->mz;
  sa.mixed = c->cfg.precision == 1;
  sa.alpha = c->alpha; sa.beta = c->beta;
  sa.twZ = c->twZ; sa.twX = c->twX;
  sa.U = c->Z;
  sa.in[0] = c->T[1];
  sa.in[1] = c->T[2];
  sa.gz = 0;
  sa.ZP = c->ZP;
  sa.ZFP = c->ZFP;
  sa.capg = c->capg;
  sa.swirl = swirl ? c->cap + (size_t)3 * c->nyp * c->mz * c->mx : nullptr;
  sa.mrf = mrf;
  hipError_t e = kmm::launch_swirl_plane(sa, true, c->stream);
  if (e != hipSuccess) return fail(KMM_E_HIP, std::string("closure gradient: ") + hipGetErrorString(e));
  const void* wb;
  const float *par, *nrm;
  int dev;
  kmm_mlp_view(c->closure, &wb, &par, &nrm, &dev);
  kmm::MlpFieldIO io;
  io.P = (size_t)c->nyp * c->mz * c->mx;
  io.plane_pts = c->mz * c->mx;
  io.half = 0.5 * c->cfg.yl;
  io.ycoord = c->ycoord_d;
  io.g = c->capg;
  io.h1 = c->hist[hslot(it - 1)];
  io.h2 = c->hist[hslot(it - 2)];
  io.out = c->hist[hslot(it)];
  if ((e = kmm::launch_mlp_field(io, wb, par, nrm, kmm_mlp_activation(c->closure), c->stream)) != hipSuccess)
    return fail(KMM_E_HIP, std::string("closure network: ") + hipGetErrorString(e));
  sa.caps = c->hist[hslot(it)];
  for (int q = 0; q < 12; ++q) sa.out[q] = c->NLP[q];
  if ((e = kmm::launch_closure_fwd(sa, c->stream)) != hipSuccess)
    return fail(KMM_E_HIP, std::string("closure transforms: ") + hipGetErrorString(e));
  hist_note(c, it);
  return KMM_OK;
}

// buffers of the conformation rebuild and the audit residual (allocated once)
static int audit_buffers(kmm_ctx* c) {
  if (c->conf_phys) return KMM_OK;
  const size_t P = (size_t)c->nyp * c->mz * c->mx;
  const size_t tbv = c->S * c->vsz;
  KMM_HIP(hipMalloc((void**)&c->conf_phys, 6 * P * sizeof(float)));
  c->allocs.push_back(c->conf_phys);
  KMM_HIP(hipMalloc((void**)&c->audit_pred, 6 * P * sizeof(float)));
  c->allocs.push_back(c->audit_pred);
  for (int q = 0; q < 6; ++q) {
    KMM_HIP(hipMalloc(&c->conf_nlp[q], tbv));
    c->allocs.push_back(c->conf_nlp[q]);
  }
  KMM_HIP(hipMalloc((void**)&c->audit_part, 2 * (size_t)kmm::audit_part_count() * sizeof(double)));
  c->allocs.push_back(c->audit_part);
  KMM_HIP(hipMalloc((void**)&c->audit_acc, 2 * sizeof(double)));
  c->allocs.push_back(c->audit_acc);
  return KMM_OK;
}

// C (spectral) rebuilt from a physical stress field on the 3/2 grid: the
// inverse of the stress map pointwise, forward transforms in x and z, and the
// polymer pencil's forward Chebyshev transform (mode 3) into C.  With prev
// given, the stress is 2 str - prev (one step of linear extrapolation).  The
// pending plane outputs in NLP are left untouched.
static int rebuild_conf(kmm_ctx* c, const float* str, const float* prev = nullptr) {
  int r = audit_buffers(c);
  if (r) return r;
  const size_t P = (size_t)c->nyp * c->mz * c->mx;
  const double L2 = c->cfg.zlmax * c->cfg.zlmax;
  hipError_t e = kmm::launch_conf_from_stress(str, prev, c->conf_phys, P, L2, c->cfg.ipeter, c->stream);
  if (e != hipSuccess) return fail(KMM_E_HIP, std::string("conformation rebuild: ") + hipGetErrorString(e));
  kmm::PolyPlaneArgs sa;
  std::memset(&sa, 0, sizeof(sa));
  sa.nxh = c->nxh; sa.nz = c->nz; sa.nzh = c->nzh; sa.nyp = c->nyp; sa.mx = c->mx; sa.mz = c->mz;
  sa.mixed = c->cfg.precision == 1;
  sa.alpha = c->alpha; sa.beta = c->beta;
  sa.twZ = c->twZ; sa.twX = c->twX;
  sa.ZFP = c->ZFP;
  sa.caps = c->conf_phys;
  for (int q = 0; q < 6; ++q) sa.out[6 + q] = c->conf_nlp[q];
  if ((e = kmm::launch_closure_fwd(sa, c->stream)) != hipSuccess)
    return fail(KMM_E_HIP, std::string("conformation rebuild transforms: ") + hipGetErrorString(e));
  kmm::PolyArgs pa = poly_args(c);
  for (int q = 0; q < 6; ++q) pa.NLP[q] = c->conf_nlp[q];
  if ((e = kmm::launch_poly_pencil(c->ny, 3, pa, c->cfg.precision == 1, c->stream)) != hipSuccess)
    return fail(KMM_E_HIP, std::string("conformation rebuild pencil: ") + hipGetErrorString(e));
  return KMM_OK;
}

// network prediction of step it's stress from the FENE-P history of the
// window, residual sums against the FENE-P stress just captured
static int audit_compare(kmm_ctx* c, int it) {
  const void* wb;
  const float *par, *nrm;
  int dev;
  kmm_mlp_view(c->closure, &wb, &par, &nrm, &dev);
  kmm::MlpFieldIO io;
  io.P = (size_t)c->nyp * c->mz * c->mx;
  io.plane_pts = c->mz * c->mx;
  io.half = 0.5 * c->cfg.yl;
  io.ycoord = c->ycoord_d;
  io.g = c->capg;
  io.h1 = c->hist[hslot(it - 1)];
  io.h2 = c->hist[hslot(it - 2)];
  io.out = c->audit_pred;
  hipError_t e = kmm::launch_mlp_field(io, wb, par, nrm, kmm_mlp_activation(c->closure), c->stream);
  if (e == hipSuccess)
    e = kmm::launch_audit_residual(c->audit_pred, c->hist[hslot(it)], 6 * io.P, c->audit_part, c->audit_acc,
                                   c->stream);
  if (e != hipSuccess) return fail(KMM_E_HIP, std::string("closure audit: ") + hipGetErrorString(e));
  return KMM_OK;
}

// first step after step `it` that opens an audit window: closure_it0 +
// m * audit_every, m >= 1
static int next_audit_from(const kmm_ctx* c, int it) {
  const int e = c->audit_every;
  const int d = it + 1 - c->closure_it0;
  return c->closure_it0 + (d <= e ? e : ((d + e - 1) / e) * e);
}

static int run_poly_plane(kmm_ctx* c, int it) {
  kmm::PolyPlaneArgs a;
  std::memset(&a, 0, sizeof(a));
  a.nxh = c->nxh; a.nz = c->nz; a.nzh = c->nzh; a.nyp = c->nyp; a.mx = c->mx; a.mz = c->mz;
  a.mixed = c->cfg.precision == 1;
  a.alpha = c->alpha; a.beta = c->beta;
  a.ipeter = c->cfg.ipeter;
  a.L2 = c->cfg.zlmax * c->cfg.zlmax;
  a.twZ = c->twZ; a.twX = c->twX;
  a.U = c->Z;
  // z-stage inputs: dz u(3), dy u(3), C(6), dz C(6), dy C(6)
  for (int q = 0; q < 3; ++q) { a.in[q] = c->T[q]; a.in[3 + q] = c->TPD[q]; }
  for (int q = 0; q < 6; ++q) { a.in[6 + q] = c->TP[q]; a.in[12 + q] = c->TP[q]; a.in[18 + q] = c->TP[6 + q]; }
  a.ZP = c->ZP; a.ZFP = c->ZFP;
  for (int q = 0; q < 12; ++q) a.out[q] = c->NLP[q];
  const bool cap = c->poly_capture || c->audit_w >= 0;  // audit windows capture for the comparison
  if (cap) {
    a.capg = c->capg;
    a.caps = c->hist[hslot(it)];
  }
  hipError_t e = kmm::launch_poly_plane(a, c->stream);
  if (e != hipSuccess) return fail(KMM_E_HIP, std::string("polymer plane: ") + hipGetErrorString(e));
  if (cap) hist_note(c, it);
  return KMM_OK;
}

static kmm::TracerArgs tracer_args(kmm_ctx* c) {
  kmm::TracerArgs t;
  std::memset(&t, 0, sizeof(t));
  t.mx = c->mx; t.mz = c->mz; t.ny = c->ny;
  t.xl = c->cfg.xl; t.yl = c->cfg.yl; t.zl = c->cfg.zl;
  t.delxm = c->cfg.xl / c->mx; t.delzm = c->cfg.zl / c->mz;
  t.pi = 2.0 * std::acos(0.0);
  t.dt = c->cfg.dt / (double)c->substep;
  t.ap = c->cfg.ap;
  t.ycoord = c->ycoord_d;
  t.err = c->perr;
  return t;
}

// One evaluation of vcw3dp's plane stage.  With track set and particles
// present, the physical velocity is kept and the tracers are initialised or
// advanced (part_track, solvers.f90:3463-3466).
// The scl_flag 2 source is computed for src_start <= it <= src_stop; outside
// that window the reference adds its uninitialised local array, here nothing.
static bool blob_active(const kmm_ctx* c, int it) {
  return c->scsrc && it >= c->cfg.src_start && it <= c->cfg.src_stop;
}

// inertial (Maxey-Riley) particles are tracked (particle_flag > 0)
static bool inertial(const kmm_ctx* c) { return c->cfg.particle_flag > 0 && c->npart > 0; }

// Laplacians of u, v, w of this evaluation on the 3/2 grid into lap[lap_cur]
static int mr_laplacians(kmm_ctx* c) {
  kmm::MrLapArgs la;
  std::memset(&la, 0, sizeof(la));
  la.P = (int)c->P; la.nz = c->nz; la.nzh = c->nzh; la.nxh = c->nxh; la.nyp = c->nyp; la.mx = c->mx; la.mz = c->mz;
  la.mixed = c->cfg.precision == 1;
  la.alpha = c->alpha; la.beta = c->beta; la.dyde = c->dyde;
  la.twM = c->twM; la.twZ = c->twZ; la.twX = c->twX;
  la.V = c->V; la.OMY = c->OMY; la.u0 = c->u0; la.w0 = c->w0;
  const size_t tb = (size_t)c->nyp * c->nxh * c->nz * c->vsz;
  for (int f = 0; f < 3; ++f) la.LT[f] = static_cast<char*>(c->lapT) + f * tb;
  la.LZ = c->lapZ;
  la.out = c->lap[c->lap_cur];
  hipError_t e = kmm::launch_mr_lap_pencil(c->ny, la, c->stream);
  if (e == hipSuccess) e = kmm::launch_mr_lap_plane(la, c->stream);
  if (e != hipSuccess) return fail(KMM_E_HIP, std::string("particle Laplacians: ") + hipGetErrorString(e));
  return KMM_OK;
}

// u, v, w and the Laplacians of this evaluation become the old ones of the
// next (solvers.f90:3468-3474)
static int mr_keep_old(kmm_ctx* c) {
  const size_t P = (size_t)c->nyp * c->mz * c->mx;
  KMM_HIP(hipMemcpyAsync(c->uold, c->cap, 3 * P * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  c->lap_cur = 1 - c->lap_cur;
  return KMM_OK;
}

static int run_plane(kmm_ctx* c, int it, unsigned long long* cfl, hipEvent_t* ev = nullptr, bool track = false,
                     bool poly = false, bool closure = false) {
  PlaneArgs pa = plane_args(c, it, cfl);
  const bool trk = track && c->npart > 0;
  if (trk) pa.cap = c->cap;
  if (blob_active(c, it)) {
    // solvers.f90:2078-2118 at the particle positions before this step's update
    kmm::BlobArgs b;
    b.n = c->npart; b.mx = c->mx; b.mz = c->mz; b.nyp = c->nyp;
    b.delxm = c->cfg.xl / c->mx; b.delzm = c->cfg.zl / c->mz;
    b.sigmax = c->cfg.sigmax; b.sigmay = c->cfg.sigmay; b.sigmaz = c->cfg.sigmaz;
    b.deltaT = c->cfg.deltaT;
    b.ycoord = c->ycoord_d;
    b.part = c->part;
    b.src = c->scsrc;
    hipError_t eb = kmm::launch_blob_source(b, c->stream);
    if (eb != hipSuccess) return fail(KMM_E_HIP, std::string("scalar source: ") + hipGetErrorString(eb));
    pa.scsrc = c->scsrc;
  }
  hipError_t e = ev ? kmm::launch_plane_timed(pa, c->stream, ev) : kmm::launch_plane(pa, c->stream);
  if (e != hipSuccess) return fail(KMM_E_HIP, std::string("plane stage: ") + hipGetErrorString(e));
  if (poly) {
    int r = run_poly_plane(c, it);
    if (r) return r;
  }
  const bool mr = trk && inertial(c);
  if (closure) {
    int r = run_closure_plane(c, it, trk && c->part_swirl, mr ? c->mrf : nullptr);
    if (r) return r;
  } else if (trk && (c->part_swirl || mr)) {
    // swirl_3d (solvers.f90:3147-3150 via calcswirl) for swirl_part
    kmm::PolyPlaneArgs sa;
    std::memset(&sa, 0, sizeof(sa));
    sa.nxh = c->nxh; sa.nz = c->nz; sa.nzh = c->nzh; sa.nyp = c->nyp; sa.mx = c->mx; sa.mz = c->mz;
    sa.mixed = c->cfg.precision == 1;
    sa.alpha = c->alpha; sa.beta = c->beta;
    sa.twZ = c->twZ; sa.twX = c->twX;
    sa.U = c->Z;
    sa.in[0] = c->T[1];
    sa.in[1] = c->T[2];
    sa.gz = poly ? 1 : 0;   // the polymer stage left dz u, dz v, dz w in ZP rows 0..2
    sa.ZP = c->ZP;
    sa.swirl = c->part_swirl ? c->cap + (size_t)3 * c->nyp * c->mz * c->mx : nullptr;
    sa.mrf = mr ? c->mrf : nullptr;  // vorticity and gradient for inertial particles
    if ((e = kmm::launch_swirl_plane(sa, !poly, c->stream)) != hipSuccess)
      return fail(KMM_E_HIP, std::string("swirl: ") + hipGetErrorString(e));
  }
  if (ev) hipEventRecord(ev[4], c->stream);
  if (mr && c->part_state == 2) {
    // inertial particles: Laplacians of this evaluation, then RK4 sub-steps
    // (part_track.f90:131-150)
    const size_t P = (size_t)c->nyp * c->mz * c->mx;
    int r = mr_laplacians(c);
    if (r) return r;
    kmm::MrArgs m;
    std::memset(&m, 0, sizeof(m));
    m.t = tracer_args(c);
    m.t.n = c->npart;
    m.t.fstride = (long long)P;
    m.t.sy = (long long)c->mz * c->mx; m.t.sz = c->mx; m.t.sx = 1;
    m.t.part = c->part;
    m.cap = c->cap;
    m.mrf = c->mrf;
    m.lap = c->lap[c->lap_cur];
    m.lold = c->lap[1 - c->lap_cur];
    m.uold = c->uold;
    m.substep = c->substep;
    m.cd_switch = c->cfg.cd_switch;
    m.swirl = c->part_swirl ? 1 : 0;
    m.re = c->cfg.re; m.ratio = c->cfg.ratio; m.c_mu = c->cfg.c_mu; m.gravity = c->cfg.gravity;
    m.dtfull = c->cfg.dt;
    if ((e = kmm::launch_inertial(m, c->stream)) != hipSuccess)
      return fail(KMM_E_HIP, std::string("inertial particles: ") + hipGetErrorString(e));
    if ((r = mr_keep_old(c))) return r;
    c->cap_valid = true;
  } else if (trk) {
    kmm::TracerArgs t = tracer_args(c);
    t.n = c->npart; t.nf = c->part_swirl ? 4 : 3;
    t.fld = c->cap;
    t.fstride = (long long)c->nyp * c->mz * c->mx;
    t.sy = (long long)c->mz * c->mx; t.sz = c->mx; t.sx = 1;
    t.update = (c->part_state == 2) ? 1 : 0;
    t.part = c->part;
    if ((e = kmm::launch_tracer(t, c->stream)) != hipSuccess)
      return fail(KMM_E_HIP, std::string("tracers: ") + hipGetErrorString(e));
    if (mr) {  // init_part of inertial particles: keep this evaluation's fields as the old ones
      int r = mr_laplacians(c);
      if (r) return r;
      if ((r = mr_keep_old(c))) return r;
    }
    c->part_state = 2;
    c->cap_valid = true;
  }
  if (ev) hipEventRecord(ev[5], c->stream);
  return KMM_OK;
}

static int check_particles(kmm_ctx* c) {
  if (c->npart == 0) return KMM_OK;
  int h = 0;
  KMM_HIP(hipMemcpyAsync(&h, c->perr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  KMM_HIP(hipStreamSynchronize(c->stream));
  if (h & 1) return fail(KMM_E_PARTICLE, "a particle left the grid (cell index out of range or not finite)");
  if (h & 2) return fail(KMM_E_CFL, "particle cfl failure: max |u_f - u_p| dt / a >= 10 (part_track.f90:216-221)");
  return KMM_OK;
}

// NL buffers -> FN, GN, SCN, h1n, h3n (smoothed), without touching the
// rest of the state.
static int materialize(kmm_ctx* c) {
  PencilArgs a = pencil_args(c);
  a.smooth = 1;
  if (c->cfg.polymer && c->pprev.evaluate) {
    kmm::PolyArgs pa = poly_args(c);
    pa.partB = 1; pa.force = c->pprev.force; pa.copy = c->pprev.copy; pa.smooth = 1;
    pa.closure = c->nlp_closure ? 1 : 0;
    int r = run_poly_pencil(c, 2, pa);
    if (r) return r;
    if (c->pprev.force) for (int q = 0; q < 3; ++q) a.TF[q] = c->TF[q];
  }
  hipError_t e = kmm::launch_pencil_finish(c->ny, a, c->stream);
  if (e != hipSuccess) return fail(KMM_E_HIP, std::string("pencil finish: ") + hipGetErrorString(e));
  KMM_HIP(hipStreamSynchronize(c->stream));
  return KMM_OK;
}

static int check_cfl(kmm_ctx* c, int n, int it0) {
  std::vector<unsigned long long> h(n);
  KMM_HIP(hipMemcpyAsync(h.data(), c->cfl_d, n * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
  KMM_HIP(hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; ++i) {
    double v;
    std::memcpy(&v, &h[i], sizeof(double));
    c->last_cfl = v;
    c->last_cfl_it = it0 + i;
    if (v > 1.0) {
      char msg[128];
      std::snprintf(msg, sizeof msg, "CFL failure at time step %d: %g", it0 + i, v);
      g_last_error = msg;
      return KMM_E_CFL;
    }
  }
  return KMM_OK;
}

static double host_bulk(const std::vector<double>& a) {
  double s = 0.0;
  for (size_t n = 0; n < a.size(); n += 2) s += a[n] / (1.0 - (double)n * (double)n);
  return s;
}

extern "C" int kmm_startup(kmm_ctx* c) {
  if (!c) return fail(KMM_E_ARG, "null argument");
  KMM_HIP(hipSetDevice(c->device));
  if (c->cfg.cmf && c->cfg.ubulk <= 0.0) {  // target = bulk velocity of the initial mean flow
    std::vector<double> u0(c->nyp);
    KMM_HIP(hipMemcpyAsync(u0.data(), c->u0, c->nyp * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    KMM_HIP(hipStreamSynchronize(c->stream));
    c->ubulk = host_bulk(u0);
  }
  PencilArgs a = pencil_args(c);
  // initial() evaluates vcw3dp at it = 0: no C initialisation or solve there
  const kmm_ctx::PolyFlags pf = poly_flags(c, 0);
  int r = 0;
  if (pf.evaluate) {
    kmm::PolyArgs pa = poly_args(c);
    pa.evaluate = 1;
    if ((r = run_poly_pencil(c, 1, pa))) return r;
  }
  if (pf.evaluate)
    for (int q = 0; q < 3; ++q) a.TPD[q] = c->TPD[q];
  hipError_t e = kmm::launch_pencil_startup(c->ny, a, c->stream);
  if (e != hipSuccess) return fail(KMM_E_HIP, hipGetErrorString(e));
  for (int q = 0; q < 3; ++q) a.TPD[q] = nullptr;
  KMM_HIP(hipMemsetAsync(c->cfl_d, 0, sizeof(unsigned long long), c->stream));
  c->it = 0;
  r = run_plane(c, 0, c->cfl_d, nullptr, true, pf.evaluate);
  if (r) return r;
  a.smooth = 0;  // initial() applies no smoothing (main_dns.f90:2155-2206)
  if (c->cfg.polymer) {
    if (pf.evaluate) {
      kmm::PolyArgs pa = poly_args(c);
      pa.partB = 1; pa.force = pf.force; pa.smooth = 0;
      if ((r = run_poly_pencil(c, 2, pa))) return r;
      if (pf.force) for (int q = 0; q < 3; ++q) a.TF[q] = c->TF[q];
    } else {
      for (int q = 0; q < 6; ++q) {  // no polymer terms before release
        KMM_HIP(hipMemsetAsync(c->CN[q], 0, c->S * sizeof(double2), c->stream));
        KMM_HIP(hipMemsetAsync(c->STRN[q], 0, c->S * sizeof(double2), c->stream));
      }
    }
  }
  e = kmm::launch_pencil_finish(c->ny, a, c->stream);
  if (e != hipSuccess) return fail(KMM_E_HIP, hipGetErrorString(e));
  if (c->cfg.polymer) {  // initial(): n-1 polymer terms = n terms (main_dns.f90:2200-2225)
    for (int q = 0; q < 6; ++q) {
      KMM_HIP(hipMemcpyAsync(c->CNM1[q], c->CN[q], c->S * sizeof(double2), hipMemcpyDeviceToDevice, c->stream));
      KMM_HIP(hipMemcpyAsync(c->STRNM1[q], c->STRN[q], c->S * sizeof(double2), hipMemcpyDeviceToDevice, c->stream));
    }
  }
  c->pprev = pf;
  const size_t fb = c->S * sizeof(double2);
  KMM_HIP(hipMemcpyAsync(c->FNM1, c->FN, fb, hipMemcpyDeviceToDevice, c->stream));
  KMM_HIP(hipMemcpyAsync(c->GNM1, c->GN, fb, hipMemcpyDeviceToDevice, c->stream));
  if (c->cfg.scalar) KMM_HIP(hipMemcpyAsync(c->SCNM1, c->SCN, fb, hipMemcpyDeviceToDevice, c->stream));
  KMM_HIP(hipMemcpyAsync(c->h1nm1, c->h1n, c->nyp * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  KMM_HIP(hipMemcpyAsync(c->h3nm1, c->h3n, c->nyp * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  c->nl_valid = false;
  r = check_cfl(c, 1, 0);
  return r ? r : check_particles(c);
}

// Adam steps of the closure on training pairs drawn from the captures of
// the window's last evaluations (FENE-P stresses at it, it-1; it-2 from the
// window's opening step), queued on the channel's stream.
static int retrain_closure(kmm_ctx* c) {
  const int n = c->retrain_steps * c->retrain_batch;
  kmm::MlpFieldIO io;
  io.P = (size_t)c->nyp * c->mz * c->mx;
  io.plane_pts = c->mz * c->mx;
  io.half = 0.5 * c->cfg.yl;
  io.ycoord = c->ycoord_d;
  io.g = c->capg;
  io.h1 = c->hist[hslot(c->it - 1)];
  io.h2 = c->hist[hslot(c->it - 2)];
  io.out = nullptr;
  hipError_t e = kmm::launch_mlp_sample(io, c->hist[hslot(c->it)], n, c->retrain_seed + (unsigned long long)c->it,
                                        c->rX, c->rT, c->stream);
  if (e != hipSuccess) return fail(KMM_E_HIP, std::string("closure retraining samples: ") + hipGetErrorString(e));
  return kmm_mlp_train_device(c->closure, n, c->rX, c->rT, c->retrain_steps, c->retrain_batch, c->retrain_lr,
                              c->stream);
}

// Launch one time step on the library stream (no synchronisation).  s is the
// step's slot in the CFL buffer; ev (or nullptr) receives per-kernel events;
// stage_ev records the first-step stage events of the untimed path.
static int enqueue_step(kmm_ctx* c, int s, hipEvent_t* ev, bool stage_ev) {
  c->it += 1;
  PencilArgs a = pencil_args(c);
  a.first_nl = c->nl_valid ? 1 : 0;
  // timing 2: events around the pencil step kernel only (ev[1], ev[2])
  const bool all_ev = ev && c->timing != 2;
  if (all_ev) hipEventRecord(ev[0], c->stream);
  else if (!ev && stage_ev) hipEventRecord(c->ev[0], c->stream);
  const bool att = c->closure != nullptr && c->closure_active;
  if (att && c->audit_every > 0 && c->audit_w < 0 && c->it == next_audit_from(c, c->it - 1)) c->audit_w = 0;
  const bool clo = att && c->audit_w < 0;
  kmm_ctx::PolyFlags pf = poly_flags(c, c->it);
  if (clo) {  // learned closure: no conformation transport; the stress comes from the network
    pf = kmm_ctx::PolyFlags();
    pf.evaluate = true;
    pf.force = true;
  } else if (att && c->audit_w == 0) {
    // audit window opens: C at this step (no solve) is rebuilt from the
    // closure's last two stresses extrapolated to this step -- holding the
    // last one would freeze the stress for a step, an error of a whole
    // step's increment; the next step starts the transport history afresh
    // (the n -> n-1 copy of an initial step)
    int r = rebuild_conf(c, c->hist[hslot(c->it - 1)], c->hist[hslot(c->it - 2)]);
    if (r) return r;
    KMM_HIP(hipMemsetAsync(c->audit_acc, 0, 2 * sizeof(double), c->stream));
    pf = kmm_ctx::PolyFlags();
    pf.evaluate = true;
    pf.force = true;
    pf.copy = true;
  }
  if (c->cfg.polymer) {
    kmm::PolyArgs pa = poly_args(c);
    pa.partB = c->nl_valid && c->pprev.evaluate;
    pa.force = pa.partB && c->pprev.force;
    pa.copy = pa.partB && c->pprev.copy;
    pa.closure = c->nlp_closure ? 1 : 0;
    pa.smooth = 1;
    if (!clo) { pa.init = pf.init; pa.solve = pf.solve; pa.evaluate = pf.evaluate; }
    // a closure step finishing a closure evaluation only forms the force
    const bool force_only = clo && pa.partB && pa.closure && pa.force && !pa.copy;
    int r;
    if (!clo && !pa.closure && c->ny >= KMM_POLY_SPLIT_MIN_NY) {
      // part B, then the step without it: each pass holds fewer registers
      if (pa.partB && (r = run_poly_pencil(c, 2, pa))) return r;
      kmm::PolyArgs pb = pa;
      pb.partB = 0;
      r = run_poly_pencil(c, 5, pb);
    } else {
      r = run_poly_pencil(c, force_only ? 4 : 0, pa);
    }
    if (r) return r;
    if (pa.force) for (int q = 0; q < 3; ++q) a.TF[q] = c->TF[q];
  }
  if (pf.evaluate && !clo) for (int q = 0; q < 3; ++q) a.TPD[q] = c->TPD[q];  // D u_i for the FENE-P plane
  if (ev) hipEventRecord(ev[1], c->stream);
  hipError_t e = kmm::launch_pencil_step(c->ny, a, c->stream);
  if (e != hipSuccess) return fail(KMM_E_HIP, std::string("pencil step: ") + hipGetErrorString(e));
  if (ev) hipEventRecord(ev[2], c->stream);
  else if (stage_ev) hipEventRecord(c->ev[1], c->stream);
  int r = run_plane(c, c->it, c->cfl_d + s, all_ev ? ev + 3 : nullptr, true, pf.evaluate && !clo, clo);
  if (r) return r;
  if (!ev && stage_ev) hipEventRecord(c->ev[2], c->stream);
  c->pprev = pf;
  c->nlp_closure = clo;
  c->nl_valid = true;
  if (att && c->audit_w >= 0) {
    if (c->audit_w >= 2 && (r = audit_compare(c, c->it))) return r;
    if (++c->audit_w == c->audit_window) {
      // the gate: relative residual of the network's one-step predictions
      double acc[2] = {0.0, 0.0};
      KMM_HIP(hipMemcpyAsync(acc, c->audit_acc, sizeof(acc), hipMemcpyDeviceToHost, c->stream));
      KMM_HIP(hipStreamSynchronize(c->stream));
      c->audit_last = acc[1] > 0.0 ? std::sqrt(acc[0] / acc[1]) : (acc[0] > 0.0 ? INFINITY : 0.0);
      // a non-finite state or prediction fails the gate
      if (!(std::isfinite(acc[0]) && std::isfinite(acc[1]))) c->audit_last = INFINITY;
      c->audits += 1;
      c->audit_w = -1;
      if (c->audit_last <= c->audit_gate) {
        c->closure_it0 = c->it + 1;  // the closure resumes from the FENE-P stress history
        if (c->retrain_steps > 0) {
          if ((r = retrain_closure(c))) return r;
          // a diverged retraining (non-finite loss) ends the closure: FENE-P carries on
          float loss = 0.0f;
          if ((r = kmm_mlp_last_loss(c->closure, &loss, c->stream))) return r;
          if (!std::isfinite(loss)) {
            c->closure_active = false;
            c->fallback_it = c->it + 1;
          }
        }
      } else {
        c->closure_active = false;   // fallback: the transport carries on
        c->fallback_it = c->it + 1;
      }
    }
  }
  return KMM_OK;
}

// Steps it+1 .. it+k launch the same kernels with the same arguments (apart
// from the IB activation step, constant here) -- the condition for replaying
// a captured graph.
static bool steady(const kmm_ctx* c, int k) {
  if (!c->nl_valid) return false;
  if (c->closure && c->closure_active) {
    if (c->audit_every > 0 && (c->audit_w >= 0 || next_audit_from(c, c->it) <= c->it + k)) return false;
    return c->nlp_closure && k % 3 == 0;  // history slots repeat every 3 steps
  }
  if (c->audit_w >= 0) return false;
  if (c->poly_capture) return false;                     // capture bookkeeping runs per launch
  if (c->npart > 0 && c->part_state != 2) return false;
  if (inertial(c)) return false;  // per-step copies of the old fields and buffer swaps
  if (c->nsrc > 0 && c->it < c->cfg.perturbtime + 1) return false;
  if (c->scsrc) {  // the source window must not open or close inside the graph
    const int a = c->it + 1, b = c->it + k, s0 = c->cfg.src_start, s1 = c->cfg.src_stop;
    if ((s0 > a && s0 <= b) || (s1 >= a && s1 < b)) return false;
  }
  if (c->cfg.polymer) {
    for (int j = 0; j <= k; ++j) {
      const kmm_ctx::PolyFlags f = poly_flags(c, c->it + j);
      const kmm_ctx::PolyFlags& g = c->pprev;
      if (f.init != g.init || f.solve != g.solve || f.evaluate != g.evaluate || f.force != g.force || f.copy != g.copy)
        return false;
    }
  }
  return true;
}

static void drop_graph(kmm_ctx* c) {
  if (c->gexec) hipGraphExecDestroy(c->gexec);
  if (c->graph) hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
}

extern "C" int kmm_set_graph(kmm_ctx* c, int steps) {
  if (!c || steps < 0 || steps > 4096) return fail(KMM_E_ARG, "steps per graph must be in 0..4096");
  KMM_HIP(hipSetDevice(c->device));
  drop_graph(c);
  c->graph_steps = steps;
  return KMM_OK;
}

// One chunk of kmm_step: launches up to `remaining` steps (a graph replay
// when the steps are steady) without waiting; step_chunk_finish waits and
// checks CFL and particles.
constexpr int kEagerChunk = 64;
struct StepChunk {
  int n = 0, it0 = 0;
  bool graph = false;
};

static int step_chunk_enqueue(kmm_ctx* c, int remaining, StepChunk& ch) {
  const int K = c->graph_steps;
  ch.graph = K > 0 && !c->timing && remaining >= K && steady(c, K);
  // eager steps run in chunks of at most kEagerChunk between CFL checks, so
  // that a failing step is reported at most that many steps late
  ch.n = ch.graph ? K : std::min(remaining, kEagerChunk);
  ch.it0 = c->it + 1;
  const int n = ch.n;
  if (c->timing) {
    const size_t need = (size_t)n * 9;
    while (c->tev.size() < need) {
      hipEvent_t e;
      KMM_HIP(hipEventCreate(&e));
      c->tev.push_back(e);
    }
  }
  KMM_HIP(hipMemsetAsync(c->cfl_d, 0, n * sizeof(unsigned long long), c->stream));
  if (ch.graph) {
    // kernel arguments and buffers of a steady step are the same every time;
    // the signature guards what could change between calls
    const long long sig[4] = {(long long)c->npart, (long long)(uintptr_t)c->part,
                              (long long)c->nsrc * 2 + (c->part_swirl ? 1 : 0), (long long)(uintptr_t)c->closure};
    if (c->gexec && (std::memcmp(sig, c->graph_sig, sizeof(sig)) != 0 ||
                     (c->closure && hslot(c->it + 1) != c->graph_slot)))
      drop_graph(c);
    if (!c->gexec) {
      KMM_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int r = 0;
      const int itsave = c->it;
      for (int s = 0; s < n && !r; ++s) r = enqueue_step(c, s, nullptr, false);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(c->stream, &g);
      if (r) {
        if (g) hipGraphDestroy(g);
        return r;
      }
      if (e != hipSuccess) return fail(KMM_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
      c->graph = g;
      KMM_HIP(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
      std::memcpy(c->graph_sig, sig, sizeof(sig));
      c->graph_slot = hslot(itsave + 1);
      c->it = itsave;  // the capture recorded the steps; the replay below runs them
      if (c->closure) c->hist_it = itsave;
    }
    KMM_HIP(hipGraphLaunch(c->gexec, c->stream));
    if (c->closure)
      for (int q = 1; q <= n; ++q) hist_note(c, c->it + q);
    c->it += n;
  } else {
    for (int s = 0; s < n; ++s) {
      hipEvent_t* ev = c->timing ? &c->tev[(size_t)s * 9] : nullptr;
      int r = enqueue_step(c, s, ev, s == 0);
      if (r) return r;
    }
  }
  return KMM_OK;
}

static int step_chunk_finish(kmm_ctx* c, const StepChunk& ch) {
  const int n = ch.n;
  int r = check_cfl(c, n, ch.it0);
  if (c->timing) {
    // intervals: poly pencil, pencil, zinv, ibdrag, xplane, zfwd, poly plane,
    // tracers (timing 2: the pencil only)
    static const int slot[8] = {5, 0, 1, 2, 3, 4, 6, 7};
    for (int s = 0; s < n; ++s) {
      hipEvent_t* ev = &c->tev[(size_t)s * 9];
      for (int k = 0; k < 8; ++k) {
        if (c->timing == 2 && k != 1) continue;
        float t = 0.0f;
        hipEventElapsedTime(&t, ev[k], ev[k + 1]);
        c->ktime[slot[k]] += t;
      }
    }
    c->ksamples += n;
  } else if (!ch.graph) {
    float t1 = 0, t2 = 0;
    hipEventElapsedTime(&t1, c->ev[0], c->ev[1]);
    hipEventElapsedTime(&t2, c->ev[1], c->ev[2]);
    c->t_pencil = t1;
    c->t_plane = t2;
  }
  if (r) return r;
  return check_particles(c);
}

static void reset_step_timers(kmm_ctx* c) {
  c->t_pencil = c->t_plane = 0.0;
  for (double& t : c->ktime) t = 0.0;
  c->ksamples = 0;
}

extern "C" int kmm_step(kmm_ctx* c, int nsteps) {
  if (!c || nsteps < 0) return fail(KMM_E_ARG, "bad argument");
  KMM_HIP(hipSetDevice(c->device));
  reset_step_timers(c);
  int done = 0;
  while (done < nsteps) {
    StepChunk ch;
    int r = step_chunk_enqueue(c, nsteps - done, ch);
    if (!r) r = step_chunk_finish(c, ch);
    if (r) return r;
    done += ch.n;
  }
  return KMM_OK;
}

// A channel whose steady step is just the pencil kernel and the plane stage
// (no polymer, particles, IB cells, scalar source window, closure, timing
// or per-step bookkeeping): such channels of one size can share launches.
static bool batchable(const kmm_ctx* c) {
  return !c->cfg.polymer && c->npart == 0 && c->nsrc == 0 && !c->scsrc && !c->closure && !c->poly_capture &&
         !c->timing && !inertial(c) && c->nl_valid;
}
static bool same_shape(const kmm_ctx* a, const kmm_ctx* b) {
  return a->nx == b->nx && a->ny == b->ny && a->nz == b->nz && a->cfg.precision == b->cfg.precision &&
         a->cfg.scalar == b->cfg.scalar && a->tune_pencil == b->tune_pencil && a->tune_xplane == b->tune_xplane;
}

// n steps of the channels g (same shape, all batchable) with one launch per
// stage and step on the first channel's stream, queued without waiting
// (batch_finish waits); each channel keeps its own CFL slots.  The argument
// blocks are staged in the first channel's host buffer, which lives until
// its next chunk.
struct BatchChunk {
  std::vector<kmm_ctx*> g;
  int n = 0;
  std::vector<int> it0;
};
static int batch_enqueue(BatchChunk& b) {
  kmm_ctx* c0 = b.g[0];
  const int M = (int)b.g.size(), n = b.n;
  const size_t pb = (size_t)n * M * sizeof(PencilArgs), lb = (size_t)n * M * sizeof(PlaneArgs);
  c0->batch_host.resize(pb + lb);
  PencilArgs* pa = reinterpret_cast<PencilArgs*>(c0->batch_host.data());
  PlaneArgs* la = reinterpret_cast<PlaneArgs*>(c0->batch_host.data() + pb);
  b.it0.assign(M, 0);
  for (int m = 0; m < M; ++m) {
    kmm_ctx* c = b.g[m];
    // the channel's earlier work on its own stream ends before the group's
    KMM_HIP(hipMemsetAsync(c->cfl_d, 0, n * sizeof(unsigned long long), c->stream));
    KMM_HIP(hipStreamSynchronize(c->stream));
    b.it0[m] = c->it + 1;
    for (int s = 0; s < n; ++s) {
      c->it += 1;
      pa[(size_t)s * M + m] = pencil_args(c);
      la[(size_t)s * M + m] = plane_args(c, c->it, c->cfl_d + s);
    }
  }
  if (c0->batch_cap < pb + lb) {
    int r = dalloc(c0, &c0->batch_args, pb + lb);
    if (r) return r;
    c0->batch_cap = pb + lb;
  }
  PencilArgs* dpa = static_cast<PencilArgs*>(c0->batch_args);
  PlaneArgs* dla = reinterpret_cast<PlaneArgs*>(static_cast<char*>(c0->batch_args) + pb);
  KMM_HIP(hipMemcpyAsync(c0->batch_args, c0->batch_host.data(), pb + lb, hipMemcpyHostToDevice, c0->stream));
  for (int s = 0; s < n; ++s) {
    hipError_t e = kmm::launch_pencil_step_batch(c0->ny, pa[(size_t)s * M], dpa + (size_t)s * M, M, c0->stream);
    if (e == hipSuccess) e = kmm::launch_plane_batch(la[(size_t)s * M], dla + (size_t)s * M, M, c0->stream);
    if (e != hipSuccess) return fail(KMM_E_HIP, std::string("batched ensemble step: ") + hipGetErrorString(e));
  }
  return KMM_OK;
}
static int batch_finish(const BatchChunk& b) {
  KMM_HIP(hipStreamSynchronize(b.g[0]->stream));
  for (size_t m = 0; m < b.g.size(); ++m) {
    b.g[m]->batched_steps += b.n;
    int r = check_cfl(b.g[m], b.n, b.it0[m]);
    if (r) return r;
  }
  return KMM_OK;
}

extern "C" int kmm_batched_steps(kmm_ctx* c, long long* n) {
  if (!c || !n) return fail(KMM_E_ARG, "null argument");
  *n = c->batched_steps;
  return KMM_OK;
}

extern "C" int kmm_ensemble_step_ex(kmm_ctx** ctxs, int n, int nsteps, int flags);
extern "C" int kmm_ensemble_step(kmm_ctx** ctxs, int n, int nsteps) {
  return kmm_ensemble_step_ex(ctxs, n, nsteps, 0);
}

extern "C" int kmm_ensemble_step_ex(kmm_ctx** ctxs, int n, int nsteps, int flags) {
  if (!ctxs || n <= 0 || nsteps < 0 || (flags & ~KMM_ENSEMBLE_NO_BATCH)) return fail(KMM_E_ARG, "bad argument");
  const bool allow_batch = !(flags & KMM_ENSEMBLE_NO_BATCH);
  for (int m = 0; m < n; ++m) {
    if (!ctxs[m]) return fail(KMM_E_ARG, "null member");
    if (m && ctxs[m]->device != ctxs[0]->device) return fail(KMM_E_ARG, "ensemble members must share a device");
    reset_step_timers(ctxs[m]);
  }
  KMM_HIP(hipSetDevice(ctxs[0]->device));
  std::vector<int> done(n, 0);
  std::vector<StepChunk> ch(n);
  for (;;) {
    bool any = false;
    // groups of batchable members of one shape, each with at least two
    // members: one chunk of their steps in shared launches
    std::vector<int> grouped(n, 0);
    std::vector<std::vector<int>> groups;
    for (int m = 0; m < n; ++m) {
      if (!allow_batch || grouped[m] || done[m] >= nsteps || !batchable(ctxs[m])) continue;
      std::vector<int> gi{m};
      for (int q = m + 1; q < n; ++q)
        if (!grouped[q] && done[q] == done[m] && batchable(ctxs[q]) && same_shape(ctxs[m], ctxs[q])) gi.push_back(q);
      if (gi.size() < 2) continue;
      for (int q : gi) grouped[q] = 1;
      groups.push_back(gi);
    }
    // the other members: one chunk queued on every member's stream
    for (int m = 0; m < n; ++m) {
      ch[m].n = 0;
      if (done[m] >= nsteps || grouped[m]) continue;
      any = true;
      int r = step_chunk_enqueue(ctxs[m], nsteps - done[m], ch[m]);
      if (r) return r;
    }
    std::vector<BatchChunk> bc(groups.size());
    for (size_t q = 0; q < groups.size(); ++q) {
      any = true;
      bc[q].n = std::min(nsteps - done[groups[q][0]], kEagerChunk);
      for (int m : groups[q]) bc[q].g.push_back(ctxs[m]);
      int r = batch_enqueue(bc[q]);
      if (r) return r;
    }
    if (!any) break;
    for (size_t q = 0; q < groups.size(); ++q) {
      int r = batch_finish(bc[q]);
      if (r) {
        char msg[64];
        std::snprintf(msg, sizeof msg, " (ensemble group led by member %d)", groups[q][0]);
        g_last_error += msg;
        return r;
      }
      for (int m : groups[q]) done[m] += bc[q].n;
    }
    for (int m = 0; m < n; ++m) {
      if (!ch[m].n) continue;
      int r = step_chunk_finish(ctxs[m], ch[m]);
      if (r) {
        char msg[64];
        std::snprintf(msg, sizeof msg, " (ensemble member %d)", m);
        g_last_error += msg;
        return r;
      }
      done[m] += ch[m].n;
    }
  }
  return KMM_OK;
}

extern "C" int kmm_set_timing(kmm_ctx* c, int on) {
  if (!c) return fail(KMM_E_ARG, "null argument");
  if (on < 0 || on > 2) return fail(KMM_E_ARG, "timing mode is 0, 1 or 2");
  c->timing = on;
  return KMM_OK;
}

extern "C" int kmm_kernel_times(kmm_ctx* c, double* ms, int* nsamples) {
  if (!c || !ms) return fail(KMM_E_ARG, "null argument");
  for (int k = 0; k < 8; ++k) ms[k] = c->ksamples ? c->ktime[k] / c->ksamples : 0.0;
  if (nsamples) *nsamples = c->ksamples;
  return KMM_OK;
}

extern "C" int kmm_last_cfl(kmm_ctx* c, double* cflmax, int* it) {
  if (!c) return fail(KMM_E_ARG, "null argument");
  if (cflmax) *cflmax = c->last_cfl;
  if (it) *it = c->last_cfl_it;
  return KMM_OK;
}

extern "C" int kmm_stage_times(kmm_ctx* c, double* p, double* q) {
  if (!c) return fail(KMM_E_ARG, "null argument");
  if (p) *p = c->t_pencil;
  if (q) *q = c->t_plane;
  return KMM_OK;
}

// ----------------------------------------------------------- restart ----
// Fortran unformatted sequential: 4-byte little-endian length before and
// after each record.
static bool write_record(FILE* f, const std::vector<std::pair<const void*, size_t>>& parts) {
  size_t tot = 0;
  for (auto& p : parts) tot += p.second;
  if (tot > 0x7fffffffu) return false;
  const uint32_t m = (uint32_t)tot;
  if (fwrite(&m, 4, 1, f) != 1) return false;
  for (auto& p : parts)
    if (p.second && fwrite(p.first, 1, p.second, f) != p.second) return false;
  return fwrite(&m, 4, 1, f) == 1;
}
static bool read_record(FILE* f, const std::vector<std::pair<void*, size_t>>& parts) {
  uint32_t m = 0, m2 = 0;
  if (fread(&m, 4, 1, f) != 1) return false;
  size_t tot = 0;
  for (auto& p : parts) tot += p.second;
  if (m != tot) return false;
  for (auto& p : parts)
    if (p.second && fread(p.first, 1, p.second, f) != p.second) return false;
  if (fread(&m2, 4, 1, f) != 1) return false;
  return m2 == m;
}

extern "C" int kmm_write_restart(kmm_ctx* c, const char* path) {
  if (!c || !path) return fail(KMM_E_ARG, "null argument");
  const size_t fs = c->S * 2, ms = c->nyp;
  std::vector<double> buf(6 * fs + 6 * ms);
  const int ids[6] = {KMM_V, KMM_FN, KMM_FNM1, KMM_OMY, KMM_GN, KMM_GNM1};
  for (int q = 0; q < 6; ++q) {
    int r = kmm_get_field(c, ids[q], buf.data() + q * fs);
    if (r) return r;
  }
  const int mids[6] = {KMM_U0, KMM_H1N, KMM_H1NM1, KMM_W0, KMM_H3N, KMM_H3NM1};
  for (int q = 0; q < 6; ++q) {
    int r = kmm_get_field(c, mids[q], buf.data() + 6 * fs + q * ms);
    if (r) return r;
  }
  int r = sync_fint_to_host(c);
  if (r) return r;
  FILE* f = fopen(path, "wb");
  if (!f) return fail(KMM_E_IO, std::string("cannot write ") + path);
  const size_t pb = c->fint_full[0].size() * sizeof(double);
  bool ok = write_record(f, {{buf.data(), buf.size() * sizeof(double)}}) &&
            write_record(f, {{c->fint_full[0].data(), pb}, {c->fint_full[1].data(), pb}, {c->fint_full[2].data(), pb}});
  fclose(f);
  return ok ? KMM_OK : fail(KMM_E_IO, "restart write failed");
}

extern "C" int kmm_read_restart(kmm_ctx* c, const char* path) {
  if (!c || !path) return fail(KMM_E_ARG, "null argument");
  // irstrt = 2 (main_dns.f90:1748-1809): the file holds the channel at half
  // the streamwise resolution (nxh/2 modes, mx/2 columns of the IB
  // integrals, the other sizes equal); it fills the low kx modes and the
  // first x columns, the rest starts at zero.  In the reference layouts
  // (nyp, nz, nxh) and (nyp, mz, mx), x slowest, that is a prefix of each array.
  const bool half = c->cfg.irstrt == 2;
  if (half && (c->nxh % 2 || c->mx % 2)) return fail(KMM_E_CONFIG, "irstrt = 2 needs nx/2 and 3 nx/2 even");
  const size_t fs = c->S * 2, ms = c->nyp, fsf = half ? fs / 2 : fs;
  const size_t psz = c->fint_full[0].size(), psf = half ? psz / 2 : psz;
  std::vector<double> rec(6 * fsf + 6 * ms);
  FILE* f = fopen(path, "rb");
  if (!f) return fail(KMM_E_IO, std::string("cannot open ") + path);
  std::vector<double> p0(psz, 0.0), p1(psz, 0.0), p2(psz, 0.0);
  const size_t pb = psf * sizeof(double);
  bool ok = read_record(f, {{rec.data(), rec.size() * sizeof(double)}}) &&
            read_record(f, {{p0.data(), pb}, {p1.data(), pb}, {p2.data(), pb}});
  fclose(f);
  if (!ok) return fail(KMM_E_IO, std::string("restart file ") + path + " has the wrong size or record markers");
  std::vector<double> buf(6 * fs + 6 * ms, 0.0);
  for (int q = 0; q < 6; ++q) std::copy(rec.begin() + q * fsf, rec.begin() + (q + 1) * fsf, buf.begin() + q * fs);
  std::copy(rec.begin() + 6 * fsf, rec.end(), buf.begin() + 6 * fs);
  c->nl_valid = false;
  const int ids[6] = {KMM_V, KMM_FN, KMM_FNM1, KMM_OMY, KMM_GN, KMM_GNM1};
  for (int q = 0; q < 6; ++q) {
    int r = kmm_put_field(c, ids[q], buf.data() + q * fs);
    if (r) return r;
  }
  const int mids[6] = {KMM_U0, KMM_H1N, KMM_H1NM1, KMM_W0, KMM_H3N, KMM_H3NM1};
  for (int q = 0; q < 6; ++q) {
    int r = kmm_put_field(c, mids[q], buf.data() + 6 * fs + q * ms);
    if (r) return r;
  }
  c->fint_full[0] = p0;
  c->fint_full[1] = p1;
  c->fint_full[2] = p2;
  c->it = 0;
  c->pprev = kmm_ctx::PolyFlags();
  if (c->cfg.polymer && c->cfg.crstrt == 0) {
    // initial() with crstrt = 0: no polymer state until it is released
    // (main_dns.f90:1887-1907)
    for (int q = 0; q < 6; ++q)
      for (double2* b : {c->C[q], c->CN[q], c->CNM1[q], c->STRN[q], c->STRNM1[q]})
        KMM_HIP(hipMemsetAsync(b, 0, c->S * sizeof(double2), c->stream));
  }
  return sync_fint_to_dev(c);
}

// c-restart records: SCALAR only: scalar | scn | scnm1.  POLYMER (the
// reference's SCALAR+POLYMER layout, main_dns.f90:782-793, 1910-1916):
// scalar, c11..c33 | scn, c11n.. | scnm1, c11nm1.. | str11n.. | str11nm1..
static const int kPolyFull[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};  // c11 c12 c13 c21 c22 c23 c31 c32 c33

extern "C" int kmm_write_crestart(kmm_ctx* c, const char* path) {
  if (!c || !path) return fail(KMM_E_ARG, "null argument");
  if (!c->cfg.scalar && !c->cfg.polymer) return fail(KMM_E_CONFIG, "no scalar or polymer in this configuration");
  const size_t fs = c->S * 2, nb = fs * sizeof(double);
  int r;
  std::vector<double> sc(fs, 0.0), scn(fs, 0.0), scm(fs, 0.0);
  if (c->cfg.scalar) {
    if ((r = kmm_get_field(c, KMM_SCALAR, sc.data())) || (r = kmm_get_field(c, KMM_SCN, scn.data())) ||
        (r = kmm_get_field(c, KMM_SCNM1, scm.data())))
      return r;
  }
  std::vector<std::vector<double>> pol;
  if (c->cfg.polymer) {
    pol.assign(30, std::vector<double>(fs));
    for (int q = 0; q < 30; ++q)
      if ((r = kmm_get_field(c, KMM_C11 + q, pol[q].data()))) return r;
  }
  FILE* f = fopen(path, "wb");
  if (!f) return fail(KMM_E_IO, std::string("cannot write ") + path);
  bool ok;
  if (!c->cfg.polymer) {
    ok = write_record(f, {{sc.data(), nb}}) && write_record(f, {{scn.data(), nb}}) && write_record(f, {{scm.data(), nb}});
  } else {
    std::vector<std::pair<const void*, size_t>> r1 = {{sc.data(), nb}}, r2 = {{scn.data(), nb}}, r3 = {{scm.data(), nb}}, r4,
                                                   r5;
    for (int q = 0; q < 9; ++q) r1.push_back({pol[kPolyFull[q]].data(), nb});
    for (int q = 0; q < 6; ++q) {
      r2.push_back({pol[6 + q].data(), nb});
      r3.push_back({pol[12 + q].data(), nb});
      r4.push_back({pol[18 + q].data(), nb});
      r5.push_back({pol[24 + q].data(), nb});
    }
    ok = write_record(f, r1) && write_record(f, r2) && write_record(f, r3) && write_record(f, r4) &&
         write_record(f, r5);
  }
  fclose(f);
  return ok ? KMM_OK : fail(KMM_E_IO, "c-restart write failed");
}

extern "C" int kmm_read_crestart(kmm_ctx* c, const char* path) {
  if (!c || !path) return fail(KMM_E_ARG, "null argument");
  if (!c->cfg.scalar && !c->cfg.polymer) return fail(KMM_E_CONFIG, "no scalar or polymer in this configuration");
  const size_t fs = c->S * 2, nb = fs * sizeof(double);
  std::vector<double> sc(fs), scn(fs), scm(fs);
  FILE* f = fopen(path, "rb");
  if (!f) return fail(KMM_E_IO, std::string("cannot open ") + path);
  bool ok;
  std::vector<std::vector<double>> pol;
  if (!c->cfg.polymer) {
    ok = read_record(f, {{sc.data(), nb}}) && read_record(f, {{scn.data(), nb}}) && read_record(f, {{scm.data(), nb}});
  } else {
    pol.assign(33, std::vector<double>(fs));  // 0..29 as KMM_C11.., 30..32 c21 c31 c32 (ignored)
    static const int full_dst[9] = {0, 1, 2, 30, 3, 4, 31, 32, 5};
    std::vector<std::pair<void*, size_t>> r1 = {{sc.data(), nb}}, r2 = {{scn.data(), nb}}, r3 = {{scm.data(), nb}}, r4, r5;
    for (int q = 0; q < 9; ++q) r1.push_back({pol[full_dst[q]].data(), nb});
    for (int q = 0; q < 6; ++q) {
      r2.push_back({pol[6 + q].data(), nb});
      r3.push_back({pol[12 + q].data(), nb});
      r4.push_back({pol[18 + q].data(), nb});
      r5.push_back({pol[24 + q].data(), nb});
    }
    ok = read_record(f, r1) && read_record(f, r2) && read_record(f, r3) && read_record(f, r4) && read_record(f, r5);
  }
  fclose(f);
  if (!ok) return fail(KMM_E_IO, std::string("c-restart file ") + path + " is malformed");
  int r;
  if (c->cfg.scalar) {
    if ((r = kmm_put_field(c, KMM_SCALAR, sc.data())) || (r = kmm_put_field(c, KMM_SCN, scn.data())) ||
        (r = kmm_put_field(c, KMM_SCNM1, scm.data())))
      return r;
  }
  if (c->cfg.polymer)
    for (int q = 0; q < 30; ++q)
      if ((r = kmm_put_field(c, KMM_C11 + q, pol[q].data()))) return r;
  return KMM_OK;
}

// ------------------------------------------------- per-routine entries --
extern "C" int kmm_cderiv(kmm_ctx* c, const double* f, double* df) {
  if (!c || !f || !df) return fail(KMM_E_ARG, "null argument");
  KMM_HIP(hipSetDevice(c->device));
  // the z-intermediate buffer holds at least two FP64 spectral fields
  double2* a = static_cast<double2*>(c->ZF);
  double2* b = a + c->S;
  int r = upload_field(c, a, f);
  if (r) return r;
  hipError_t e = kmm::launch_cderiv(c->ny, c->P, c->dyde, a, b, c->stream);
  if (e != hipSuccess) return fail(KMM_E_HIP, hipGetErrorString(e));
  return download_field(c, b, df);
}

// Helmholtz constants of one kx slab: wn = kz^2 + x + g per kz, with the
// Neumann zero-mode guard of solve/penta (wn <= 1e-20 -> 1, flagged).
// Returns whether any kz was flagged.
static bool slab_wn(const kmm_ctx* c, double x, double g, int ib, double* wn, int* kz) {
  bool anyk = false;
  for (int j = 0; j < c->nz; ++j) {
    const int js = (j < c->nzh) ? j : j - c->nz;
    wn[j] = (js * c->beta) * (js * c->beta) + x + g;
    kz[j] = 0;
    if (ib != 0 && wn[j] <= 1.0e-20) { wn[j] = 1.0; kz[j] = 1; anyk = true; }
  }
  return anyk;
}

// tau tables and the solve of one slab du [y][kz] (P = nz pencils) whose
// wn, flags and boundary values are already on the device
static hipError_t slab_launch(kmm_ctx* c, double2* du, double2* dtab, const double2* dbt, const double2* dbb,
                              const double* dwn, const int* dkz, bool anyk, double dyde, int ib, bool penta,
                              double at_, double btt, double ab_, double bbb) {
  hipError_t e = kmm::launch_tau_tables_wn(c->ny, c->nz, dyde, dwn, dtab, c->stream);
  PencilArgs a = pencil_args(c);
  a.P = c->nz;
  a.dyde = dyde;
  if (e != hipSuccess) return e;
  if (!penta) return kmm::launch_solve(c->ny, a, dtab, dwn, du, dbt, dbb, ib, anyk ? dkz : nullptr, c->stream);
  return kmm::launch_penta(c->ny, a, dtab, dwn, du, at_, btt, dbt, ab_, bbb, dbb, +1, anyk ? dkz : nullptr,
                           c->stream);
}

// one kx slab: u complex (nyp,nz) -> device [y][kz] with P = nz pencils
static int slab_solve(kmm_ctx* c, double* u, double x, double g, double dyde, const double* bt, const double* bb,
                      int ib, bool penta, double at_, double btt, double ab_, double bbb) {
  KMM_HIP(hipSetDevice(c->device));
  const int nz = c->nz, nyp = c->nyp;
  std::vector<double2> h((size_t)nyp * nz);
  for (int j = 0; j < nz; ++j)
    for (int i = 0; i < nyp; ++i) {
      const size_t r = (size_t)i + (size_t)nyp * j;
      h[(size_t)i * nz + j] = make_double2(u[2 * r], u[2 * r + 1]);
    }
  std::vector<double> wn(nz);
  std::vector<int> kz(nz, 0);
  const bool anyk = slab_wn(c, x, g, ib, wn.data(), kz.data());
  // slab layout: u [nyp][nz], 1/d table [nz][tab_width], bctop, bcbot [nz]
  // (double2), wn [nz] (double), zero-mode flags [nz] (int)
  const size_t nu = (size_t)nyp * nz, ntab = (size_t)nz * kmm::tab_width(c->ny);
  if (!c->slab) {
    int r = dalloc(c, &c->slab, (nu + ntab + 2 * (size_t)nz) * sizeof(double2) + (size_t)nz * (sizeof(double) + sizeof(int)));
    if (r) return r;
  }
  double2* du = static_cast<double2*>(c->slab);
  double2* dtab = du + nu;
  double2* dbt = dtab + ntab;
  double2* dbb = dbt + nz;
  double* dwn = reinterpret_cast<double*>(dbb + nz);
  int* dkz = reinterpret_cast<int*>(dwn + nz);
  KMM_HIP(hipMemcpyAsync(du, h.data(), h.size() * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  KMM_HIP(hipMemcpyAsync(dbt, bt, nz * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  KMM_HIP(hipMemcpyAsync(dbb, bb, nz * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  KMM_HIP(hipMemcpyAsync(dwn, wn.data(), nz * sizeof(double), hipMemcpyHostToDevice, c->stream));
  KMM_HIP(hipMemcpyAsync(dkz, kz.data(), nz * sizeof(int), hipMemcpyHostToDevice, c->stream));
  hipError_t e = slab_launch(c, du, dtab, dbt, dbb, dwn, dkz, anyk, dyde, ib, penta, at_, btt, ab_, bbb);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h.data(), du, h.size() * sizeof(double2), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(KMM_E_HIP, hipGetErrorString(e));
  for (int j = 0; j < nz; ++j)
    for (int i = 0; i < nyp; ++i) {
      const size_t r = (size_t)i + (size_t)nyp * j;
      u[2 * r] = h[(size_t)i * nz + j].x;
      u[2 * r + 1] = h[(size_t)i * nz + j].y;
    }
  return KMM_OK;
}

extern "C" int kmm_solve(kmm_ctx* c, double* u, double x, double g, double dyde, const double* bctop,
                         const double* bcbot, int ib) {
  if (!c || !u || !bctop || !bcbot) return fail(KMM_E_ARG, "null argument");
  return slab_solve(c, u, x, g, dyde, bctop, bcbot, ib, false, 0, 0, 0, 0);
}

extern "C" int kmm_penta(kmm_ctx* c, double* u, double x, double g, double dyde, int ib, double at, double bt,
                         const double* gtop, double ab, double bb, const double* gbot) {
  if (!c || !u || !gtop || !gbot) return fail(KMM_E_ARG, "null argument");
  return slab_solve(c, u, x, g, dyde, gtop, gbot, ib, true, at, bt, ab, bb);
}

// --------------------------------------------- per-routine field entries --
static int rt_buffers(kmm_ctx* c) {
  if (!c->rt) {
    int r = dalloc(c, (void**)&c->rt, 9 * c->S * sizeof(double2) + 4 * (size_t)c->nyp * sizeof(double));
    if (r) return r;
  }
  return KMM_OK;
}
static double* rt_mean(kmm_ctx* c, int i) { return reinterpret_cast<double*>(c->rt + 9 * c->S) + (size_t)i * c->nyp; }

static kmm::RoutineArgs routine_args(kmm_ctx* c, int op) {
  kmm::RoutineArgs a;
  std::memset(&a, 0, sizeof(a));
  a.op = op;
  a.nxh = c->nxh; a.nz = c->nz; a.nzh = c->nzh; a.P = c->P;
  a.alpha = c->alpha; a.beta = c->beta; a.dyde = c->dyde; a.re = c->cfg.re; a.dt = c->cfg.dt;
  a.theta = c->cfg.theta;
  return a;
}

// Uploads nin spectral fields (reference layout) into rt[0..nin), runs op
// with outputs rt[nin..nin+nout), downloads them into out[].  m0/m1 are
// optional mean profiles.
static int run_routine(kmm_ctx* c, int op, const double* const* in, int nin, double* const* out, int nout,
                       double par, const double* m0 = nullptr, const double* m1 = nullptr) {
  for (int i = 0; i < nin; ++i) if (!in[i]) return fail(KMM_E_ARG, "null argument");
  for (int o = 0; o < nout; ++o) if (!out[o]) return fail(KMM_E_ARG, "null argument");
  KMM_HIP(hipSetDevice(c->device));
  int r = rt_buffers(c);
  if (r) return r;
  kmm::RoutineArgs a = routine_args(c, op);
  a.par = par;
  for (int i = 0; i < nin; ++i) {
    if ((r = upload_field(c, c->rt + (size_t)i * c->S, in[i]))) return r;
    a.in[i] = c->rt + (size_t)i * c->S;
  }
  for (int o = 0; o < nout; ++o) a.out[o] = c->rt + (size_t)(nin + o) * c->S;
  if (m0) { a.m0 = rt_mean(c, 0); KMM_HIP(hipMemcpyAsync(a.m0, m0, c->nyp * sizeof(double), hipMemcpyHostToDevice, c->stream)); }
  if (m1) { a.m1 = rt_mean(c, 1); KMM_HIP(hipMemcpyAsync(a.m1, m1, c->nyp * sizeof(double), hipMemcpyHostToDevice, c->stream)); }
  hipError_t e = kmm::launch_routine(c->ny, a, c->stream);
  if (e != hipSuccess) return fail(KMM_E_HIP, hipGetErrorString(e));
  for (int o = 0; o < nout; ++o)
    if ((r = download_field(c, a.out[o], out[o]))) return r;
  return KMM_OK;
}

extern "C" int kmm_phirhs(kmm_ctx* c, double* v, const double* fn, const double* fnm1, double* bcbot, double* bctop,
                          double var) {
  if (!c || !v || !fn || !fnm1) return fail(KMM_E_ARG, "null argument");
  const double* in[3] = {v, fn, fnm1};
  double* out[1] = {v};
  int r = run_routine(c, kmm::RT_PHIRHS, in, 3, out, 1, var);
  if (r) return r;
  // phirhs zeroes the boundary values (solvers.f90:618-625)
  const size_t nb = 2 * (size_t)c->nz * c->nxh;
  if (bcbot) std::memset(bcbot, 0, nb * sizeof(double));
  if (bctop) std::memset(bctop, 0, nb * sizeof(double));
  return KMM_OK;
}

extern "C" int kmm_nonlin(kmm_ctx* c, double* fn, const double* omz, double* gn) {
  if (!c) return fail(KMM_E_ARG, "null argument");
  const double* in[3] = {fn, omz, gn};
  double* out[2] = {fn, gn};
  return run_routine(c, kmm::RT_NONLIN, in, 3, out, 2, 0.0);
}

extern "C" int kmm_veloc(kmm_ctx* c, double* u, double* v, double* w, const double* u0, const double* w0,
                         const double* omy) {
  if (!c || !u0 || !w0) return fail(KMM_E_ARG, "null argument");
  const double* in[2] = {v, omy};
  double* out[3] = {u, w, v};
  return run_routine(c, kmm::RT_VELOC, in, 2, out, 3, 0.0, u0, w0);
}

extern "C" int kmm_vort(kmm_ctx* c, const double* u, const double* v, const double* w, double* omx, double* omz) {
  if (!c) return fail(KMM_E_ARG, "null argument");
  const double* in[3] = {u, v, w};
  double* out[2] = {omx, omz};
  return run_routine(c, kmm::RT_VORT, in, 3, out, 2, 0.0);
}

extern "C" int kmm_gradscl(kmm_ctx* c, const double* scalar, double* sclx, double* scly, double* sclz) {
  if (!c) return fail(KMM_E_ARG, "null argument");
  const double* in[1] = {scalar};
  double* out[3] = {sclx, scly, sclz};
  return run_routine(c, kmm::RT_GRAD, in, 1, out, 3, 0.0);
}

extern "C" int kmm_derivscji(kmm_ctx* c, const double* const* cij, double* const* dc) {
  if (!c || !cij || !dc) return fail(KMM_E_ARG, "null argument");
  // cij: c11 c12 c13 c21 c22 c23 c31 c32 c33; dc in derivscji's dummy order
  // (derivatives.f90:148-151): dc111 dc112 dc113 dc211 dc212 dc213 dc311
  // dc312 dc313, dc121 ... dc323, dc131 ... dc333; dc_ijk = d c_ij / d x_k
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const double* in[1] = {cij[3 * i + j]};
      double* out[3];
      for (int k = 0; k < 3; ++k) out[k] = dc[9 * j + 3 * i + k];
      int r = run_routine(c, kmm::RT_GRAD, in, 1, out, 3, 0.0);
      if (r) return r;
    }
  return KMM_OK;
}

extern "C" int kmm_polyforce(kmm_ctx* c, const double* s11, const double* s12, const double* s13,
                             const double* s22, const double* s23, const double* s33, double* t1, double* t2,
                             double* t3) {
  if (!c) return fail(KMM_E_ARG, "null argument");
  const double* in[6] = {s11, s12, s13, s22, s23, s33};
  double* out[3] = {t1, t2, t3};
  return run_routine(c, kmm::RT_POLYFORCE, in, 6, out, 3, 0.0);
}

extern "C" int kmm_subforce(kmm_ctx* c, double* gn, double* fn, double* omz, const double* t1, const double* t2,
                            const double* t3) {
  if (!c) return fail(KMM_E_ARG, "null argument");
  const double* in[6] = {gn, fn, omz, t1, t2, t3};
  double* out[3] = {gn, fn, omz};
  return run_routine(c, kmm::RT_SUBFORCE, in, 6, out, 3, 0.0);
}

extern "C" int kmm_polyNL(kmm_ctx* c, double* const* cNL, const double* const* cn, const double* const* cnm1,
                          const double* const* strn, const double* const* strnm1) {
  if (!c || !cNL || !cn || !cnm1 || !strn || !strnm1) return fail(KMM_E_ARG, "null argument");
  if (!(c->cfg.tpoly > 0.0)) return fail(KMM_E_CONFIG, "polyNL needs tpoly > 0");
  for (int q = 0; q < 6; ++q) {  // components 11 12 13 22 23 33
    const double* in[4] = {cn[q], cnm1[q], strn[q], strnm1[q]};
    double* out[1] = {cNL[q]};
    int r = run_routine(c, kmm::RT_POLYNL, in, 4, out, 1, 1.0 / c->cfg.tpoly);
    if (r) return r;
  }
  return KMM_OK;
}

extern "C" int kmm_norm(kmm_ctx* c, double* a) {
  if (!c) return fail(KMM_E_ARG, "null argument");
  const double* in[1] = {a};
  double* out[1] = {a};
  // in place on the device copy: every element is read and written by one wave
  KMM_HIP(hipSetDevice(c->device));
  int r = rt_buffers(c);
  if (r) return r;
  if (!a) return fail(KMM_E_ARG, "null argument");
  kmm::RoutineArgs ar = routine_args(c, kmm::RT_NORM);
  if ((r = upload_field(c, c->rt, in[0]))) return r;
  ar.in[0] = c->rt;
  ar.out[0] = c->rt;
  hipError_t e = kmm::launch_routine(c->ny, ar, c->stream);
  if (e != hipSuccess) return fail(KMM_E_HIP, hipGetErrorString(e));
  return download_field(c, c->rt, out[0]);
}

extern "C" int kmm_uzero(kmm_ctx* c, double* u0, const double* h1, const double* h1l, int uflag) {
  if (!c || !u0 || !h1 || !h1l) return fail(KMM_E_ARG, "null argument");
  if (uflag != 1 && uflag != 3) return fail(KMM_E_ARG, "uflag must be 1 (u0) or 3 (w0)");
  KMM_HIP(hipSetDevice(c->device));
  int r = rt_buffers(c);
  if (r) return r;
  kmm::RoutineArgs a = routine_args(c, 0);
  const PencilArgs pa = pencil_args(c);
  // Robin rows of psolve1d: uflag 3 passes btu for w (solvers.f90:1173)
  const double rb[6] = {uflag == 1 ? pa.atu : pa.atw, pa.btu, uflag == 1 ? pa.gtu : pa.gtw,
                        uflag == 1 ? pa.abu : pa.abw, uflag == 1 ? pa.bbu : pa.bbw, uflag == 1 ? pa.gbu : pa.gbw};
  std::memcpy(a.rb, rb, sizeof(rb));
  const double* src[3] = {u0, h1, h1l};
  double** dst[3] = {&a.m0, &a.m1, &a.m2};
  for (int i = 0; i < 3; ++i) {
    *dst[i] = rt_mean(c, i);
    KMM_HIP(hipMemcpyAsync(*dst[i], src[i], c->nyp * sizeof(double), hipMemcpyHostToDevice, c->stream));
  }
  a.m3 = rt_mean(c, 3);
  hipError_t e = kmm::launch_uzero(c->ny, a, c->tabH, c->stream);
  if (e != hipSuccess) return fail(KMM_E_HIP, hipGetErrorString(e));
  KMM_HIP(hipMemcpyAsync(u0, a.m3, c->nyp * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  KMM_HIP(hipStreamSynchronize(c->stream));
  return KMM_OK;
}

// ---------------------------------------------- slab solve, no allocation --
// kmm_solve_noalloc runs kmm_solve's kernels with every device buffer carved
// from `scratch`, a fixed array each GPU holds until the process ends: two
// slab copies (reference column-major order and the kernels' [y][kz] order),
// the tau table and the per-kz constants.  Nothing is allocated on the device
// or the heap; the host stages only through stack arrays.  Every call fully
// writes what it reads back, so no part of the array needs clearing; calls
// that overlap on one GPU share the array and race.
namespace {
constexpr size_t kNaSlab = 131072;  // max nyp * nz complex elements of one slab
constexpr int kNaNz = 1024;         // max nz
constexpr int kNaTab = kmm::tab_width(192);  // widest 1/d table (ny = 192)
constexpr size_t kNaBytes = 2 * kNaSlab * sizeof(double2) + (size_t)kNaNz * (kNaTab + 2) * sizeof(double2) +
                            (size_t)kNaNz * (sizeof(double) + sizeof(int));
constexpr int kNaDevices = 64;
std::mutex g_na_mu;
void* g_na_ptr[kNaDevices] = {};    // device address of `scratch`, per device
}  // namespace
static __device__ __align__(256) unsigned char scratch[kNaBytes];

// out[c * R + r] = in[r * C + c]: between the reference's column-major slab
// (R = nz rows of nyp) and the kernels' [y][kz] order (R = nyp rows of nz)
__global__ void k_slab_transpose(const double2* __restrict__ in, double2* __restrict__ out, int R, int C) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= R * C) return;
  const int r = t / C, col = t - r * C;
  out[(size_t)col * R + r] = in[t];
}

extern "C" int kmm_solve_noalloc(kmm_ctx* c, double* u, double x, double g, double dyde, const double* bctop,
                                 const double* bcbot, int ib) {
  if (!c || !u || !bctop || !bcbot) return fail(KMM_E_ARG, "null argument");
  const int nz = c->nz, nyp = c->nyp;
  if (nz > kNaNz || (size_t)nyp * nz > kNaSlab)
    return fail(KMM_E_ARG, "kmm_solve_noalloc: slab larger than its fixed scratch (nz <= 1024, nyp*nz <= 131072)");
  if (c->device < 0 || c->device >= kNaDevices) return fail(KMM_E_ARG, "kmm_solve_noalloc: device index >= 64");
  KMM_HIP(hipSetDevice(c->device));
  unsigned char* base;
  {
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[c->device]) {
      void* p = nullptr;
      KMM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      g_na_ptr[c->device] = p;
    }
    base = static_cast<unsigned char*>(g_na_ptr[c->device]);
  }
  double2* dcol = reinterpret_cast<double2*>(base);           // reference order
  double2* du = dcol + kNaSlab;                               // [y][kz]
  double2* dtab = du + kNaSlab;
  double2* dbt = dtab + (size_t)kNaNz * kNaTab;
  double2* dbb = dbt + kNaNz;
  double* dwn = reinterpret_cast<double*>(dbb + kNaNz);
  int* dkz = reinterpret_cast<int*>(dwn + kNaNz);
  double wn[kNaNz];
  int kz[kNaNz];
  const bool anyk = slab_wn(c, x, g, ib, wn, kz);
  const size_t nu = (size_t)nyp * nz;
  const int nb = (int)((nu + 255) / 256);
  KMM_HIP(hipMemcpyAsync(dcol, u, nu * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  KMM_HIP(hipMemcpyAsync(dbt, bctop, nz * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  KMM_HIP(hipMemcpyAsync(dbb, bcbot, nz * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  KMM_HIP(hipMemcpyAsync(dwn, wn, nz * sizeof(double), hipMemcpyHostToDevice, c->stream));
  KMM_HIP(hipMemcpyAsync(dkz, kz, nz * sizeof(int), hipMemcpyHostToDevice, c->stream));
  k_slab_transpose<<<nb, 256, 0, c->stream>>>(dcol, du, nz, nyp);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = slab_launch(c, du, dtab, dbt, dbb, dwn, dkz, anyk, dyde, ib, false, 0, 0, 0, 0);
  if (e == hipSuccess) {
    k_slab_transpose<<<nb, 256, 0, c->stream>>>(du, dcol, nyp, nz);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(u, dcol, nu * sizeof(double2), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(KMM_E_HIP, hipGetErrorString(e));
  return KMM_OK;
}

// vcw3dp's body for both argument lists: NS (+ scalar) inputs, and with
// poly set the conformation C (six independent components), its y
// derivatives and the velocity's y derivatives for the FENE-P products.
struct Vcw3dpPoly {
  const double* c[6];    // C11 C12 C13 C22 C23 C33 (spectral)
  const double* dyc[6];  // their y derivatives (dc112, dc122, dc132, dc222, dc232, dc332)
  const double* dyu[3];  // du/dy, dv/dy, dw/dy (u12, u22, u32)
  double* cn[6];         // outputs: c_n, str_n, qp of the six components
  double* strn[6];
  double* qp[6];
};
static int vcw3dp_body(kmm_ctx* c, const double* u, const double* v, const double* w, const double* omx,
                       const double* omy, double* omz, double* fn, double* gn, const double* sclx,
                       const double* scly, const double* sclz, double* scn, int it, const Vcw3dpPoly* poly) {
  double* omz_out = omz;
  if (!c || !u || !v || !w || !omx || !omy || !omz || !gn || !fn) return fail(KMM_E_ARG, "null argument");
  const bool sc = c->cfg.scalar != 0;
  if (sc && (!sclx || !scly || !sclz || !scn)) return fail(KMM_E_ARG, "scalar arrays required");
  KMM_HIP(hipSetDevice(c->device));
  const double* ins[9] = {u, v, w, omx, omy, omz, sclx, scly, sclz};
  const int nin = sc ? 9 : 6;
  // spectral inputs -> NL buffers (scratch) -> Chebyshev-to-physical -> T
  auto to_transfer = [&](const double* src, void* dst) -> int {
    int r = upload_field(c, static_cast<double2*>(c->ZF), src);
    if (r) return r;
    hipError_t e = kmm::launch_cheb(c->ny, c->P, c->twM, c->ZF, dst, 1, c->cfg.precision, c->stream);
    return e == hipSuccess ? KMM_OK : fail(KMM_E_HIP, hipGetErrorString(e));
  };
  int r;
  for (int f = 0; f < nin; ++f)
    if ((r = to_transfer(ins[f], c->T[f]))) return r;
  // the FENE-P products are evaluated for it >= src_start - 1 (solvers.f90:2901)
  const bool pterms = poly && it >= c->cfg.src_start - 1;
  if (pterms) {
    for (int q = 0; q < 6; ++q) {
      if ((r = to_transfer(poly->c[q], c->TP[q]))) return r;
      if ((r = to_transfer(poly->dyc[q], c->TP[6 + q]))) return r;
    }
    for (int q = 0; q < 3; ++q)
      if ((r = to_transfer(poly->dyu[q], c->TPD[q]))) return r;
  }
  KMM_HIP(hipMemsetAsync(c->cfl_d, 0, sizeof(unsigned long long), c->stream));
  r = run_plane(c, it, c->cfl_d, nullptr, false, pterms, false);
  if (r) return r;
  auto from_nl = [&](const void* src, double* dst) -> int {
    hipError_t e = kmm::launch_cheb(c->ny, c->P, c->twM, src, c->ZF, 0, c->cfg.precision, c->stream);
    if (e != hipSuccess) return fail(KMM_E_HIP, hipGetErrorString(e));
    return download_field(c, static_cast<double2*>(c->ZF), dst);
  };
  double* outs[4] = {gn, fn, omz_out, scn};
  const int nout = sc ? 4 : 3;
  for (int o = 0; o < nout; ++o)
    if ((r = from_nl(c->NL[o], outs[o]))) return r;
  if (pterms) {
    // qp = (1 - beta)/(re beta tpoly) str with beta = qbeta (itarget 0)
    const double zb = (1.0 - c->cfg.qbeta) / (c->cfg.re * c->cfg.qbeta * c->cfg.tpoly);
    for (int q = 0; q < 6; ++q) {
      if ((r = from_nl(c->NLP[q], poly->cn[q]))) return r;
      if ((r = from_nl(c->NLP[6 + q], poly->strn[q]))) return r;
      for (size_t e = 0; e < 2 * c->S; ++e) poly->qp[q][e] = zb * poly->strn[q][e];
    }
  }
  c->nl_valid = false;
  return check_cfl(c, 1, it);
}

extern "C" int kmm_vcw3dp(kmm_ctx* c, const double* u, const double* v, const double* w, const double* omx,
                          const double* omy, double* omz, double* fn, double* gn, const double* const* uij,
                          const double* scalar, const double* sclx, const double* scly, const double* sclz,
                          double* scn, int it) {
  (void)uij;     // the velocity gradients feed calcswirl/part_track, which kmm_step runs itself
  (void)scalar;  // the scalar enters only through its gradient
  return vcw3dp_body(c, u, v, w, omx, omy, omz, fn, gn, sclx, scly, sclz, scn, it, nullptr);
}

extern "C" int kmm_vcw3dp_poly(kmm_ctx* c, const double* u, const double* v, const double* w, const double* omx,
                               const double* omy, double* omz, double* fn, double* gn, const double* const* uij,
                               const double* scalar, const double* sclx, const double* scly, const double* sclz,
                               double* scn, const double* const* cij, const double* const* dc, double* const* cn,
                               double* const* strn, double* const* qp, int it) {
  (void)scalar;
  if (!c || !uij || !cij || !dc || !cn || !strn || !qp) return fail(KMM_E_ARG, "null argument");
  if (!c->cfg.polymer) return fail(KMM_E_CONFIG, "the POLYMER argument list needs polymer = 1");
  if (c->cfg.itarget != 0) return fail(KMM_E_CONFIG, "only itarget 0 is built");
  Vcw3dpPoly pl;
  // independent components 11 12 13 22 23 33 of the nine (c21 = c12, ...)
  static const int ci[6] = {0, 1, 2, 4, 5, 8};
  for (int q = 0; q < 6; ++q) {
    const int i = ci[q] / 3, j = ci[q] % 3;   // C_(i+1)(j+1)
    pl.c[q] = cij[ci[q]];
    pl.dyc[q] = dc[9 * j + 3 * i + 1];       // dc_ij2 in derivscji's dummy order
    pl.cn[q] = cn[q];
    pl.strn[q] = strn[q];
    pl.qp[q] = qp[q];
    if (!pl.c[q] || !pl.dyc[q] || !pl.cn[q] || !pl.strn[q] || !pl.qp[q]) return fail(KMM_E_ARG, "null argument");
  }
  pl.dyu[0] = uij[1]; pl.dyu[1] = uij[4]; pl.dyu[2] = uij[7];   // u12, u22, u32
  if (!pl.dyu[0] || !pl.dyu[1] || !pl.dyu[2]) return fail(KMM_E_ARG, "u12, u22, u32 required");
  return vcw3dp_body(c, u, v, w, omx, omy, omz, fn, gn, sclx, scly, sclz, scn, it, &pl);
}

// --------------------------------------------------------- particles ----
static int part_free(kmm_ctx* c) {
  if (c->part) hipFree(c->part);
  if (c->cap) hipFree(c->cap);
  c->part = nullptr;
  c->cap = nullptr;
  c->part_cap = 0;
  void** mr[6] = {(void**)&c->mrf, (void**)&c->lap[0], (void**)&c->lap[1], (void**)&c->uold, &c->lapT, &c->lapZ};
  for (void** q : mr) {
    if (*q) hipFree(*q);
    *q = nullptr;
  }
  return KMM_OK;
}

// buffers of the inertial-particle fields (allocated once per context)
static int mr_buffers(kmm_ctx* c) {
  if (c->mrf) return KMM_OK;
  const size_t P = (size_t)c->nyp * c->mz * c->mx;
  KMM_HIP(hipMalloc((void**)&c->mrf, 12 * P * sizeof(double)));
  for (int b = 0; b < 2; ++b) KMM_HIP(hipMalloc((void**)&c->lap[b], 3 * P * sizeof(double)));
  KMM_HIP(hipMalloc((void**)&c->uold, 3 * P * sizeof(double)));
  KMM_HIP(hipMalloc(&c->lapT, 3 * (size_t)c->nyp * c->nxh * c->nz * c->vsz));
  KMM_HIP(hipMalloc(&c->lapZ, 3 * (size_t)c->nyp * c->mz * c->nxh * c->vsz));
  c->lap_cur = 0;
  return KMM_OK;
}

extern "C" int kmm_set_particles(kmm_ctx* c, int n, const double* x, const double* y, const double* z) {
  if (!c || n < 0 || (n > 0 && (!x || !y || !z))) return fail(KMM_E_ARG, "bad argument");
  KMM_HIP(hipSetDevice(c->device));
  KMM_HIP(hipStreamSynchronize(c->stream));
  drop_graph(c);  // particle buffers may be reallocated
  if (n == 0) {
    c->npart = 0;
    c->part_state = 0;
    c->cap_valid = false;
    return part_free(c);
  }
  if (c->cfg.particle_flag < 0) return fail(KMM_E_CONFIG, "particle_flag must be 0 (tracers) or > 0 (inertial)");
  // substep_check, part_track.f90:335-350
  const kmm_config& g = c->cfg;
  if (g.particle_flag > 0 && !(g.ap > 0.0))
    return fail(KMM_E_CONFIG, "inertial particles need a positive radius a");
  double tau = 1.0;
  if (g.cd_switch != 0) {
    const double pi_f = 1.0 / (g.ratio + 0.5);
    tau = (g.cd_switch == 1) ? g.re * (g.ap * g.ap) / (6.0 * pi_f) : 2.0 * g.re * (g.ap * g.ap) / (9.0 * pi_f);
  }
  if (!(tau > 0.0) || !std::isfinite(2.0 * g.dt / tau))
    return fail(KMM_E_CONFIG, "particle time scale is zero (set the particle radius or CD_switch 0)");
  c->substep = (int)std::ceil(std::max(1.0, 2.0 * g.dt / tau));
  if (n > c->part_cap) {
    if (c->part) { hipFree(c->part); c->part = nullptr; }
    KMM_HIP(hipMalloc((void**)&c->part, (size_t)7 * n * sizeof(double)));
    c->part_cap = n;
  }
  if (!c->cap) KMM_HIP(hipMalloc((void**)&c->cap, (size_t)4 * c->nyp * c->mz * c->mx * sizeof(double)));
  c->part_swirl = g.particle_swirl != 0;
  if (g.particle_flag > 0) {
    int r = mr_buffers(c);
    if (r) return r;
  }
  if ((c->part_swirl || g.particle_flag > 0) && !c->ZP) {  // dz v, dz w z lines (the polymer stage allocates ZP otherwise)
    const size_t zb = (size_t)c->nyp * c->mz * c->nxh * c->vsz;
    KMM_HIP(hipMalloc(&c->ZP, 2 * zb));
    c->allocs.push_back(c->ZP);
  }
  const size_t b = (size_t)n * sizeof(double);
  KMM_HIP(hipMemcpyAsync(c->part, x, b, hipMemcpyHostToDevice, c->stream));
  KMM_HIP(hipMemcpyAsync(c->part + n, y, b, hipMemcpyHostToDevice, c->stream));
  KMM_HIP(hipMemcpyAsync(c->part + 2 * (size_t)n, z, b, hipMemcpyHostToDevice, c->stream));
  KMM_HIP(hipMemsetAsync(c->part + 3 * (size_t)n, 0, 4 * b, c->stream));
  KMM_HIP(hipMemsetAsync(c->perr, 0, sizeof(int), c->stream));
  KMM_HIP(hipStreamSynchronize(c->stream));
  c->npart = n;
  c->part_state = 1;
  c->cap_valid = false;
  return KMM_OK;
}

extern "C" int kmm_read_particles(kmm_ctx* c, const char* path) {
  if (!c || !path) return fail(KMM_E_ARG, "null argument");
  std::ifstream f(path);
  if (!f) return fail(KMM_E_IO, std::string("cannot open ") + path);
  std::string ln;
  if (!std::getline(f, ln)) return fail(KMM_E_IO, std::string(path) + ": missing header line");
  std::vector<double> vals;
  while (std::getline(f, ln)) {
    for (auto& ch : ln) {
      if (ch == 'd' || ch == 'D') ch = 'e';
      if (ch == ',') ch = ' ';
    }
    std::istringstream is(ln);
    std::string tok;
    while (is >> tok) {
      char* end = nullptr;
      const double v = std::strtod(tok.c_str(), &end);
      if (end == tok.c_str() || *end != '\0') return fail(KMM_E_IO, std::string(path) + ": bad number '" + tok + "'");
      vals.push_back(v);
    }
  }
  if (vals.size() % 7 != 0) return fail(KMM_E_IO, std::string(path) + ": expected 7 values per particle");
  const int n = (int)(vals.size() / 7);
  std::vector<double> x(n), y(n), z(n);
  for (int j = 0; j < n; ++j) {
    x[j] = vals[7 * (size_t)j];
    y[j] = vals[7 * (size_t)j + 1];
    z[j] = vals[7 * (size_t)j + 2];
  }
  return kmm_set_particles(c, n, x.data(), y.data(), z.data());
}

extern "C" int kmm_particle_count(kmm_ctx* c, int* n) {
  if (!c || !n) return fail(KMM_E_ARG, "null argument");
  *n = c->npart;
  return KMM_OK;
}

extern "C" int kmm_get_particles(kmm_ctx* c, int n, double* x, double* y, double* z, double* u, double* v,
                                 double* w) {
  if (!c) return fail(KMM_E_ARG, "null argument");
  if (n != c->npart) return fail(KMM_E_ARG, "n differs from the number of particles set");
  if (n == 0) return KMM_OK;
  KMM_HIP(hipSetDevice(c->device));
  double* dst[6] = {x, y, z, u, v, w};
  for (int k = 0; k < 6; ++k)
    if (dst[k])
      KMM_HIP(hipMemcpyAsync(dst[k], c->part + (size_t)k * n, (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
  KMM_HIP(hipStreamSynchronize(c->stream));
  return KMM_OK;
}

extern "C" int kmm_get_particle_swirl(kmm_ctx* c, int n, double* swirl) {
  if (!c || !swirl) return fail(KMM_E_ARG, "null argument");
  if (n != c->npart) return fail(KMM_E_ARG, "n differs from the number of particles set");
  if (!c->part_swirl) return fail(KMM_E_CONFIG, "particle swirl is off (particle_swirl = 0)");
  if (n == 0) return KMM_OK;
  KMM_HIP(hipSetDevice(c->device));
  KMM_HIP(hipMemcpyAsync(swirl, c->part + (size_t)6 * n, (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                         c->stream));
  KMM_HIP(hipStreamSynchronize(c->stream));
  return KMM_OK;
}

// Fortran Ew.d edit descriptor with w = 14, d = 6: [-]0.ddddddE+ee
static std::string fortran_e14_6(double v) {
  char out[32];
  if (std::isnan(v)) { std::snprintf(out, sizeof out, "%14s", "NaN"); return out; }
  if (std::isinf(v)) { std::snprintf(out, sizeof out, "%14s", v > 0 ? "Infinity" : "-Infinity"); return out; }
  std::string digits = "000000";
  int ex = 0;
  const bool neg = std::signbit(v) && v != 0.0;
  if (v != 0.0) {
    char t[32];
    std::snprintf(t, sizeof t, "%.5E", std::fabs(v));  // d.dddddE+xx: same 6 significant digits
    digits = std::string(1, t[0]) + std::string(t + 2, 5);
    ex = std::atoi(std::strchr(t, 'E') + 1) + 1;
  }
  char e[8];
  if (std::abs(ex) <= 99) std::snprintf(e, sizeof e, "E%c%02d", ex < 0 ? '-' : '+', std::abs(ex));
  else std::snprintf(e, sizeof e, "%c%03d", ex < 0 ? '-' : '+', std::abs(ex));
  const std::string body = std::string(neg ? "-" : "") + "0." + digits + e;
  std::snprintf(out, sizeof out, "%14s", body.c_str());
  return out;
}

extern "C" int kmm_write_particles(kmm_ctx* c, const char* path, int it) {
  if (!c || !path) return fail(KMM_E_ARG, "null argument");
  const int n = c->npart;
  std::vector<double> a((size_t)7 * n, 0.0);
  if (n) {
    int r = kmm_get_particles(c, n, a.data(), a.data() + n, a.data() + 2 * (size_t)n, a.data() + 3 * (size_t)n,
                              a.data() + 4 * (size_t)n, a.data() + 5 * (size_t)n);
    if (!r && c->part_swirl) r = kmm_get_particle_swirl(c, n, a.data() + 6 * (size_t)n);
    if (r) return r;
  }
  FILE* f = std::fopen(path, "w");
  if (!f) return fail(KMM_E_IO, std::string("cannot write ") + path);
  std::fprintf(f, " variables = \"x\", \"y\", \"z\", \"u\", \"v\", \"w\", \"swirl\"\n");
  std::fprintf(f, " zone f=point t=\"Particles\", solutiontime=%12d \n\n", it);
  for (int j = 0; j < n; ++j) {
    std::string line;
    for (int k = 0; k < 6; ++k) line += fortran_e14_6(a[(size_t)k * n + j]) + " ";
    line += fortran_e14_6(a[(size_t)6 * n + j]) + " ";
    std::fprintf(f, "%s\n", line.c_str());
  }
  const bool ok = std::fclose(f) == 0;
  return ok ? KMM_OK : fail(KMM_E_IO, std::string("cannot write ") + path);
}

extern "C" int kmm_fluid_interp(kmm_ctx* c, int n, const double* xp, const double* yp, const double* zp, int nf,
                                const double* fields, double* out, int* cells) {
  if (!c || n < 0 || nf < 1 || (n > 0 && (!xp || !yp || !zp || !fields || !out))) return fail(KMM_E_ARG, "bad argument");
  if (n == 0) return KMM_OK;
  KMM_HIP(hipSetDevice(c->device));
  const size_t pf = (size_t)c->nyp * c->mz * c->mx;
  double *df = nullptr, *dp = nullptr, *dout = nullptr;
  int *dcell = nullptr, *derr = nullptr;
  hipError_t e = hipMalloc((void**)&df, (size_t)nf * pf * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&dp, (size_t)3 * n * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&dout, (size_t)nf * n * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&dcell, (size_t)3 * n * sizeof(int));
  if (e == hipSuccess) e = hipMalloc((void**)&derr, sizeof(int));
  if (e == hipSuccess) e = hipMemcpyAsync(df, fields, (size_t)nf * pf * sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dp, xp, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dp + n, yp, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dp + 2 * (size_t)n, zp, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(derr, 0, sizeof(int), c->stream);
  int herr = 0;
  if (e == hipSuccess) {
    kmm::TracerArgs t = tracer_args(c);
    t.n = n; t.nf = nf;
    t.fld = df; t.fstride = (long long)pf;
    t.sy = 1; t.sz = c->nyp; t.sx = (long long)c->nyp * c->mz;  // Fortran (nyp, mz, mx)
    t.px = dp; t.py = dp + n; t.pz = dp + 2 * (size_t)n;
    t.out = dout; t.cells = dcell; t.err = derr;
    e = kmm::launch_tracer(t, c->stream);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, (size_t)nf * n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && cells)
    e = hipMemcpyAsync(cells, dcell, (size_t)3 * n * sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&herr, derr, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(df); hipFree(dp); hipFree(dout); hipFree(dcell); hipFree(derr);
  if (e != hipSuccess) return fail(KMM_E_HIP, hipGetErrorString(e));
  if (herr) return fail(KMM_E_PARTICLE, "a point lies outside the grid");
  return KMM_OK;
}

extern "C" int kmm_get_physical(kmm_ctx* c, int comp, double* out) {
  if (!c || !out || comp < 0 || comp > 2) return fail(KMM_E_ARG, "bad argument");
  if (!c->cap_valid) return fail(KMM_E_ARG, "no physical velocity kept (set particles and evaluate the flow first)");
  KMM_HIP(hipSetDevice(c->device));
  const size_t pf = (size_t)c->nyp * c->mz * c->mx;
  std::vector<double> h(pf);
  KMM_HIP(hipMemcpyAsync(h.data(), c->cap + (size_t)comp * pf, pf * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  KMM_HIP(hipStreamSynchronize(c->stream));
  for (int k = 0; k < c->nyp; ++k)
    for (int jz = 0; jz < c->mz; ++jz)
      for (int ix = 0; ix < c->mx; ++ix)
        out[k + (size_t)c->nyp * (jz + (size_t)c->mz * ix)] = h[((size_t)k * c->mz + jz) * c->mx + ix];
  return KMM_OK;
}

extern "C" int kmm_get_dpdx(kmm_ctx* c, double* dpdx) {
  if (!c || !dpdx) return fail(KMM_E_ARG, "null argument");
  KMM_HIP(hipSetDevice(c->device));
  KMM_HIP(hipMemcpyAsync(dpdx, c->dpdx_d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  KMM_HIP(hipStreamSynchronize(c->stream));
  return KMM_OK;
}

extern "C" int kmm_set_dpdx(kmm_ctx* c, double dpdx) {
  if (!c) return fail(KMM_E_ARG, "null argument");
  KMM_HIP(hipSetDevice(c->device));
  c->dPdx = dpdx;
  KMM_HIP(hipMemcpyAsync(c->dpdx_d, &c->dPdx, sizeof(double), hipMemcpyHostToDevice, c->stream));
  KMM_HIP(hipStreamSynchronize(c->stream));
  return KMM_OK;
}

extern "C" int kmm_gradv(kmm_ctx* c, const double* u, const double* v, const double* w, double* u11, double* u12,
                         double* u13, double* u21, double* u22, double* u23, double* u31, double* u32, double* u33,
                         double* lu, double* lv, double* lw) {
  double* outs[12] = {u11, u12, u13, u21, u22, u23, u31, u32, u33, lu, lv, lw};
  if (!c || !u || !v || !w) return fail(KMM_E_ARG, "null argument");
  for (double* o : outs)
    if (!o) return fail(KMM_E_ARG, "null argument");
  KMM_HIP(hipSetDevice(c->device));
  const size_t fb = c->S * sizeof(double2);
  double2* d = nullptr;
  double2** dptr = nullptr;
  hipError_t e = hipMalloc((void**)&d, 15 * fb);
  if (e == hipSuccess) e = hipMalloc((void**)&dptr, 12 * sizeof(double2*));
  if (e != hipSuccess) {
    hipFree(d);
    return fail(KMM_E_HIP, hipGetErrorString(e));
  }
  int r = upload_field(c, d, u);
  if (!r) r = upload_field(c, d + c->S, v);
  if (!r) r = upload_field(c, d + 2 * c->S, w);
  double2* hp[12];
  for (int q = 0; q < 12; ++q) hp[q] = d + (3 + q) * c->S;
  if (!r) {
    e = hipMemcpyAsync(dptr, hp, sizeof(hp), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
      e = kmm::launch_gradv(c->ny, c->P, c->nz, c->alpha, c->beta, c->dyde, d, d + c->S, d + 2 * c->S, dptr,
                            c->stream);
    if (e != hipSuccess) r = fail(KMM_E_HIP, hipGetErrorString(e));
  }
  for (int q = 0; q < 12 && !r; ++q) r = download_field(c, hp[q], outs[q]);
  hipFree(d);
  hipFree(dptr);
  return r;
}

extern "C" int kmm_calcswirl(kmm_ctx* c, const double* u11, const double* u21, const double* u31,
                             const double* u12, const double* u22, const double* u32, const double* u13,
                             const double* u23, const double* u33, double* swirl) {
  // tensor components in u11 u12 u13 u21 u22 u23 u31 u32 u33 order
  const double* in[9] = {u11, u12, u13, u21, u22, u23, u31, u32, u33};
  if (!c || !swirl) return fail(KMM_E_ARG, "null argument");
  for (const double* a : in)
    if (!a) return fail(KMM_E_ARG, "null argument");
  KMM_HIP(hipSetDevice(c->device));
  const size_t n = (size_t)c->nyp * c->mz * c->mx;
  double* d = nullptr;
  hipError_t e = hipMalloc((void**)&d, 10 * n * sizeof(double));
  for (int k = 0; k < 9 && e == hipSuccess; ++k)
    e = hipMemcpyAsync(d + k * n, in[k], n * sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = kmm::launch_swirl((int)n, d, d + 9 * n, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(swirl, d + 9 * n, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(d);
  return e == hipSuccess ? KMM_OK : fail(KMM_E_HIP, hipGetErrorString(e));
}

// ------------------------------------------------------ learned closure --
extern "C" int kmm_set_poly_capture(kmm_ctx* c, int on) {
  if (!c) return fail(KMM_E_ARG, "null argument");
  if (!c->cfg.polymer) return fail(KMM_E_CONFIG, "capture needs the polymer (polymer = 1)");
  KMM_HIP(hipSetDevice(c->device));
  if (on && !c->capg) {
    const size_t P = (size_t)c->nyp * c->mz * c->mx;
    KMM_HIP(hipMalloc((void**)&c->capg, 9 * P * sizeof(float)));
    c->allocs.push_back(c->capg);
    for (int q = 0; q < 3; ++q) {
      KMM_HIP(hipMalloc((void**)&c->hist[q], 6 * P * sizeof(float)));
      c->allocs.push_back(c->hist[q]);
    }
  }
  c->poly_capture = on != 0;
  drop_graph(c);
  return KMM_OK;
}

extern "C" int kmm_poly_samples(kmm_ctx* c, int n, unsigned long long seed, float* X, float* T) {
  if (!c || n <= 0 || !X || !T) return fail(KMM_E_ARG, "bad argument");
  if (!c->capg || c->hist_n < 3)
    return fail(KMM_E_ARG, "needs three consecutive captured evaluations (kmm_set_poly_capture, then step)");
  KMM_HIP(hipSetDevice(c->device));
  const int it = c->hist_it;
  kmm::MlpFieldIO io;
  io.P = (size_t)c->nyp * c->mz * c->mx;
  io.plane_pts = c->mz * c->mx;
  io.half = 0.5 * c->cfg.yl;
  io.ycoord = c->ycoord_d;
  io.g = c->capg;
  io.h1 = c->hist[hslot(it - 1)];
  io.h2 = c->hist[hslot(it - 2)];
  io.out = nullptr;
  float *dX = nullptr, *dT = nullptr;
  hipError_t e = hipMalloc((void**)&dX, (size_t)n * 22 * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&dT, (size_t)n * 6 * sizeof(float));
  if (e == hipSuccess) e = kmm::launch_mlp_sample(io, c->hist[hslot(it)], n, seed, dX, dT, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(X, dX, (size_t)n * 22 * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(T, dT, (size_t)n * 6 * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(dX);
  hipFree(dT);
  return e == hipSuccess ? KMM_OK : fail(KMM_E_HIP, hipGetErrorString(e));
}

extern "C" int kmm_set_closure(kmm_ctx* c, kmm_mlp* mlp) {
  if (!c) return fail(KMM_E_ARG, "null argument");
  if (mlp) {
    if (!c->cfg.polymer) return fail(KMM_E_CONFIG, "the closure replaces the polymer (polymer = 1)");
    const void* wb;
    const float *par, *nrm;
    int dev = -1;
    kmm_mlp_view(mlp, &wb, &par, &nrm, &dev);
    if (dev != c->device) return fail(KMM_E_ARG, "network and channel on different devices");
    if (!c->capg || c->hist_n < 2 || c->hist_it != c->it)
      return fail(KMM_E_ARG, "the closure needs the stress of the last two steps (capture on for two steps)");
  }
  c->closure = mlp;
  c->closure_active = mlp != nullptr;
  c->closure_it0 = c->it + 1;
  c->audit_w = -1;
  c->fallback_it = -1;
  drop_graph(c);
  return KMM_OK;
}

extern "C" int kmm_set_closure_audit(kmm_ctx* c, int every, int window, double gate) {
  if (!c) return fail(KMM_E_ARG, "null argument");
  if (every < 0 || (every > 0 && (window < 3 || window >= every)) || !(gate >= 0.0))
    return fail(KMM_E_ARG, "audit: every = 0 (off) or window >= 3 and every > window; gate >= 0");
  if (every > 0) {
    if (!c->cfg.polymer) return fail(KMM_E_CONFIG, "audits need the polymer (polymer = 1)");
    if (!c->capg) return fail(KMM_E_ARG, "audits need the captures (kmm_set_poly_capture)");
    KMM_HIP(hipSetDevice(c->device));
    int r = audit_buffers(c);
    if (r) return r;
  }
  if (c->audit_w >= 0) return fail(KMM_E_ARG, "an audit window is running");
  c->audit_every = every;
  c->audit_window = window;
  c->audit_gate = gate;
  drop_graph(c);
  return KMM_OK;
}

extern "C" int kmm_set_closure_retrain(kmm_ctx* c, int steps, int batch, float lr, unsigned long long seed) {
  if (!c || steps < 0 || (steps > 0 && (batch <= 0 || !(lr > 0.0f))))
    return fail(KMM_E_ARG, "retrain: steps >= 0; batch > 0 and lr > 0 when steps > 0");
  KMM_HIP(hipSetDevice(c->device));
  const size_t n = (size_t)steps * batch;
  if (steps > 0 && n > (size_t)(c->retrain_steps) * c->retrain_batch) {
    int r;
    if ((r = dalloc(c, (void**)&c->rX, n * 22 * sizeof(float)))) return r;
    if ((r = dalloc(c, (void**)&c->rT, n * 6 * sizeof(float)))) return r;
  }
  c->retrain_steps = steps;
  c->retrain_batch = batch;
  c->retrain_lr = lr;
  c->retrain_seed = seed;
  return KMM_OK;
}

extern "C" int kmm_closure_status(kmm_ctx* c, int* active, int* audits, double* last_residual, int* fallback_it) {
  if (!c || !active || !audits || !last_residual || !fallback_it) return fail(KMM_E_ARG, "null argument");
  *active = (c->closure && c->closure_active) ? 1 : 0;
  *audits = c->audits;
  *last_residual = c->audit_last;
  *fallback_it = c->fallback_it;
  return KMM_OK;
}

extern "C" int kmm_poly_conf_from_stress(kmm_ctx* c) {
  if (!c) return fail(KMM_E_ARG, "null argument");
  if (!c->cfg.polymer) return fail(KMM_E_CONFIG, "needs the polymer (polymer = 1)");
  if (!c->capg || c->hist_n < 1) return fail(KMM_E_ARG, "no captured stress");
  KMM_HIP(hipSetDevice(c->device));
  int r = rebuild_conf(c, c->hist[hslot(c->hist_it)]);
  if (r) return r;
  KMM_HIP(hipStreamSynchronize(c->stream));
  return KMM_OK;
}

// stress of the last evaluation on the 3/2 grid, reference layout (nyp, mz, mx)
// per component (11 12 13 22 23 33)
extern "C" int kmm_get_poly_stress(kmm_ctx* c, int comp, double* out) {
  if (!c || !out || comp < 0 || comp > 5) return fail(KMM_E_ARG, "bad argument");
  if (!c->capg || c->hist_n < 1) return fail(KMM_E_ARG, "no captured stress");
  KMM_HIP(hipSetDevice(c->device));
  const size_t P = (size_t)c->nyp * c->mz * c->mx;
  std::vector<float> h(P);
  KMM_HIP(hipMemcpyAsync(h.data(), c->hist[hslot(c->hist_it)] + (size_t)comp * P, P * sizeof(float),
                         hipMemcpyDeviceToHost, c->stream));
  KMM_HIP(hipStreamSynchronize(c->stream));
  for (int k = 0; k < c->nyp; ++k)
    for (int jz = 0; jz < c->mz; ++jz)
      for (int ix = 0; ix < c->mx; ++ix)
        out[k + (size_t)c->nyp * (jz + (size_t)c->mz * ix)] = h[((size_t)k * c->mz + jz) * c->mx + ix];
  return KMM_OK;
}

// ------------------------------------------------------------ autotuner --
// Occupancy variants of the two heaviest kernels (the pencil step kernel and
// the x-plane kernel) are timed on a twin of the context -- same
// configuration, geometry, tracers and state, its own buffers -- and the
// faster variant of each is kept (the role FFTW's PATIENT planning plays for
// the reference's transforms, solvers.f90:2057-2069).
extern "C" int kmm_set_tuning(kmm_ctx* c, int pencil, int xplane) {
  if (!c || pencil < 0 || pencil > 1 || xplane < 0 || xplane > 1) return fail(KMM_E_ARG, "variants are 0 or 1");
  c->tune_pencil = pencil;
  c->tune_xplane = xplane;
  drop_graph(c);
  return KMM_OK;
}

extern "C" int kmm_get_tuning(kmm_ctx* c, int* pencil, int* xplane) {
  if (!c || !pencil || !xplane) return fail(KMM_E_ARG, "null argument");
  *pencil = c->tune_pencil;
  *xplane = c->tune_xplane;
  return KMM_OK;
}

extern "C" int kmm_autotune(kmm_ctx* c, int steps, double* ms) {
  if (!c || steps < 1 || steps > 1000) return fail(KMM_E_ARG, "steps must be in 1..1000");
  KMM_HIP(hipSetDevice(c->device));
  KMM_HIP(hipStreamSynchronize(c->stream));
  kmm_ctx* t = nullptr;
  int r = kmm_create(&c->cfg, c->device, &t);
  if (r) return r;
  if (!c->imatrix_h.empty()) r = kmm_set_geometry(t, c->imatrix_h.data());
  if (!r && c->npart > 0) {
    std::vector<double> x(c->npart), y(c->npart), z(c->npart);
    r = kmm_get_particles(c, c->npart, x.data(), y.data(), z.data(), nullptr, nullptr, nullptr);
    if (!r) r = kmm_set_particles(t, c->npart, x.data(), y.data(), z.data());
  }
  std::vector<int> ids = {KMM_V, KMM_OMY, KMM_U0, KMM_W0};
  if (c->cfg.scalar) ids.push_back(KMM_SCALAR);
  if (c->cfg.polymer)
    for (int q = KMM_C11; q <= KMM_C33; ++q) ids.push_back(q);
  std::vector<double> buf(2 * c->S + 16);
  for (size_t k = 0; k < ids.size() && !r; ++k) {
    r = kmm_get_field(c, ids[k], buf.data());
    if (!r) r = kmm_put_field(t, ids[k], buf.data());
  }
  if (!r) r = kmm_startup(t);
  double tm[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
  for (int v = 0; v < 2 && !r; ++v) {
    t->tune_pencil = v;
    t->tune_xplane = v;
    t->timing = 0;
    r = kmm_step(t, 2);
    if (r) break;
    t->timing = 1;
    r = kmm_step(t, steps);
    if (r) break;
    double kms[8];
    kmm_kernel_times(t, kms, nullptr);
    tm[0][v] = kms[0];  // k_pencil
    tm[1][v] = kms[3];  // k_xplane
  }
  const std::string why = r ? g_last_error : std::string();
  kmm_destroy(t);
  if (r) return fail(r, "autotune: " + why);
  c->tune_pencil = tm[0][1] < tm[0][0] ? 1 : 0;
  c->tune_xplane = tm[1][1] < tm[1][0] ? 1 : 0;
  drop_graph(c);
  if (ms) {
    ms[0] = tm[0][0]; ms[1] = tm[0][1]; ms[2] = tm[1][0]; ms[3] = tm[1][1];
  }
  return KMM_OK;
}

// ----------------------------------------------------------- statistics --
// Chebyshev sum f(y_k) = sum_n a_n cos(n pi k / ny), y_k = cos(pi k / ny)
// (the inverse transform with fac = c(i)/2, solvers.f90:2241-2546)
static std::vector<double> cheb_eval(const std::vector<double>& a, int ny) {
  std::vector<double> f(ny + 1, 0.0);
  const double pi = 2.0 * std::acos(0.0);
  for (int k = 0; k <= ny; ++k) {
    double s = 0.0;
    for (int n = 0; n <= ny; ++n) s += a[n] * std::cos(pi * (double)n * (double)k / (double)ny);
    f[k] = s;
  }
  return f;
}

// cderiv (derivatives.f90:258-314) of one Chebyshev series
static std::vector<double> cheb_deriv(const std::vector<double>& a, int ny, double dyde) {
  std::vector<double> d(ny + 1, 0.0);
  if (ny >= 1) d[ny - 1] = 2.0 * ny * a[ny] * dyde;
  for (int n = ny - 2; n >= 0; --n) d[n] = d[n + 2] + 2.0 * (n + 1) * a[n + 1] * dyde;
  d[0] *= 0.5;
  return d;
}

extern "C" int kmm_statistics(kmm_ctx* c, double* prof, double* scal) {
  if (!c || !prof || !scal) return fail(KMM_E_ARG, "null argument");
  if (!c->nl_valid) return fail(KMM_E_ARG, "no evaluation yet (kmm_startup first)");
  KMM_HIP(hipSetDevice(c->device));
  const int nyp = c->nyp, ny = c->ny;
  const kmm_config& g = c->cfg;
  if (!c->stats_d) {
    KMM_HIP(hipMalloc((void**)&c->stats_d, (size_t)8 * nyp * sizeof(double)));
    c->allocs.push_back(c->stats_d);
  }
  kmm::StatsArgs a;
  std::memset(&a, 0, sizeof(a));
  a.P = c->P; a.nz = c->nz; a.nyp = nyp; a.mixed = g.precision == 1;
  a.u = c->T[0]; a.v = c->T[1]; a.w = c->T[2];
  const bool poly = g.polymer != 0;
  if (poly) a.q = c->NLP[7];                                   // str12 of the last evaluation
  if (poly && c->pprev.evaluate && !c->nlp_closure) { a.c11 = c->TP[0]; a.c22 = c->TP[3]; a.c33 = c->TP[5]; }
  a.out = c->stats_d;
  hipError_t e = kmm::launch_stats(a, c->stream);
  if (e != hipSuccess) return fail(KMM_E_HIP, std::string("statistics: ") + hipGetErrorString(e));
  std::vector<double> out((size_t)8 * nyp);
  KMM_HIP(hipMemcpyAsync(out.data(), c->stats_d, out.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  // the (0,0) modes of the state: u0 (Chebyshev) and the scalar
  std::vector<double> u0(nyp), th(nyp, 0.0);
  KMM_HIP(hipMemcpyAsync(u0.data(), c->u0, nyp * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (g.scalar)
    KMM_HIP(hipMemcpy2DAsync(th.data(), sizeof(double), c->TH, sizeof(double2), sizeof(double), nyp,
                             hipMemcpyDeviceToHost, c->stream));  // pencil 0, pencil-major
  // swirl strength: the tracers' field when this evaluation made one, else
  // computed here from the same z intermediates
  std::vector<double> swy(nyp, 0.0);
  const double* swf = nullptr;
  if (c->npart > 0 && c->part_swirl && c->cap_valid && c->cap) {
    swf = c->cap + (size_t)3 * nyp * c->mz * c->mx;
  } else {
    const size_t pp = (size_t)nyp * c->mz * c->mx;
    if (!c->swirl_tmp) {
      KMM_HIP(hipMalloc((void**)&c->swirl_tmp, pp * sizeof(double)));
      c->allocs.push_back(c->swirl_tmp);
    }
    if (!c->ZP) {
      KMM_HIP(hipMalloc(&c->ZP, 2 * (size_t)nyp * c->mz * c->nxh * c->vsz));
      c->allocs.push_back(c->ZP);
    }
    kmm::PolyPlaneArgs sa;
    std::memset(&sa, 0, sizeof(sa));
    sa.nxh = c->nxh; sa.nz = c->nz; sa.nzh = c->nzh; sa.nyp = nyp; sa.mx = c->mx; sa.mz = c->mz;
    sa.mixed = g.precision == 1;
    sa.alpha = c->alpha; sa.beta = c->beta;
    sa.twZ = c->twZ; sa.twX = c->twX;
    sa.U = c->Z;
    sa.in[0] = c->T[1];
    sa.in[1] = c->T[2];
    sa.gz = 0;
    sa.ZP = c->ZP;
    sa.swirl = c->swirl_tmp;
    if ((e = kmm::launch_swirl_plane(sa, true, c->stream)) != hipSuccess)
      return fail(KMM_E_HIP, std::string("statistics swirl: ") + hipGetErrorString(e));
    swf = c->swirl_tmp;
  }
  if ((e = kmm::launch_plane_mean(swf, nyp, c->mz * c->mx, c->stats_d, c->stream)) != hipSuccess)
    return fail(KMM_E_HIP, std::string("statistics swirl mean: ") + hipGetErrorString(e));
  KMM_HIP(hipMemcpyAsync(swy.data(), c->stats_d, nyp * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  KMM_HIP(hipStreamSynchronize(c->stream));
  const std::vector<double> du = cheb_eval(cheb_deriv(u0, ny, c->dyde), ny);
  const std::vector<double> thp = cheb_eval(th, ny);
  const double zbeta1 = poly ? (1.0 - g.qbeta) / (g.re * g.qbeta * g.tpoly) : 0.0;
  double* umean = prof;
  double* urms = prof + nyp;
  double* vrms = prof + 2 * nyp;
  double* wrms = prof + 3 * nyp;
  double* rss = prof + 4 * nyp;
  double* txy = prof + 5 * nyp;
  double* betay = prof + 6 * nyp;
  double* swirly = prof + 7 * nyp;
  double* pss = prof + 8 * nyp;
  for (int i = 0; i < nyp; ++i) {
    umean[i] = out[(size_t)4 * nyp + i];
    urms[i] = std::sqrt(out[i]);
    vrms[i] = std::sqrt(out[(size_t)nyp + i]);
    wrms[i] = std::sqrt(out[(size_t)2 * nyp + i]);
    rss[i] = out[(size_t)3 * nyp + i];
    txy[i] = du[i] / g.re;
    // writeoutputs gets beta_poly in POLYMER builds (qbeta for itarget 0), the scalar otherwise
    betay[i] = poly ? g.qbeta : (g.scalar ? thp[i] : 0.0);
    swirly[i] = swy[i];
    pss[i] = poly ? zbeta1 * out[(size_t)6 * nyp + i] : 0.0;
  }
  double mf = 0.0, ckk = 0.0;
  for (int i = 0; i + 1 < nyp; ++i)
    mf += 1.0 / g.yl * 0.5 * (umean[i + 1] + umean[i]) * std::fabs(c->ycoord[i + 1] - c->ycoord[i]);
  if (a.c11)
    for (int i = 0; i < nyp; ++i) {
      const double ly = (i == 0) ? std::fabs(c->ycoord[1] - c->ycoord[0]) / 2.0
                      : (i == nyp - 1) ? std::fabs(c->ycoord[nyp - 1] - c->ycoord[nyp - 2]) / 2.0
                      : std::fabs(c->ycoord[i + 1] - c->ycoord[i]) / 2.0 + std::fabs(c->ycoord[i] - c->ycoord[i - 1]) / 2.0;
      ckk += out[(size_t)7 * nyp + i] * ly / g.yl;
    }
  scal[0] = mf;
  scal[1] = txy[0];
  scal[2] = txy[nyp - 1];
  scal[3] = pss[0];
  scal[4] = pss[nyp - 1];
  scal[5] = a.c11 ? ckk : std::nan("");
  return KMM_OK;
}

// Physical fields of the last evaluation on the 3/2 grid, the arguments of
// write_flowfield_ascii (helpers.f90:1671-1767): u, v, w, wx, wy, wz, swirl
// and, with the scalar, the scalar; FP64, reference layout (nyp, mz, mx) per
// field.  The velocity and vorticity come from the transfer fields through
// the z and x inverse transforms, the swirl from the gradient/swirl plane on
// the same z intermediates, the scalar from an inverse Chebyshev transform of
// the state first.
extern "C" int kmm_flowfield(kmm_ctx* c, double* out) {
  if (!c || !out) return fail(KMM_E_ARG, "null argument");
  if (!c->nl_valid) return fail(KMM_E_ARG, "no evaluation yet (kmm_startup first)");
  KMM_HIP(hipSetDevice(c->device));
  const int nyp = c->nyp;
  const size_t P = (size_t)nyp * c->mz * c->mx;
  const int nf = c->cfg.scalar ? 8 : 7;
  const size_t zb = (size_t)nyp * c->mz * c->nxh * c->vsz;
  const size_t tb = (size_t)nyp * c->P * c->vsz;
  double* d = nullptr;
  void *lz = nullptr, *lt = nullptr;
  hipError_t e = hipMalloc((void**)&d, (size_t)(nf + 2) * P * sizeof(double));  // the scalar pass writes 3 fields
  if (e == hipSuccess) e = hipMalloc(&lz, 3 * zb);
  if (e == hipSuccess && c->cfg.scalar) e = hipMalloc(&lt, tb);
  if (e == hipSuccess && !c->ZP) {
    e = hipMalloc(&c->ZP, 2 * zb);
    if (e == hipSuccess) c->allocs.push_back(c->ZP);
  }
  kmm::MrLapArgs la;
  std::memset(&la, 0, sizeof(la));
  la.P = (int)c->P; la.nz = c->nz; la.nzh = c->nzh; la.nxh = c->nxh; la.nyp = nyp; la.mx = c->mx; la.mz = c->mz;
  la.mixed = c->cfg.precision == 1;
  la.alpha = c->alpha; la.beta = c->beta; la.dyde = c->dyde;
  la.twM = c->twM; la.twZ = c->twZ; la.twX = c->twX;
  la.LZ = lz;
  for (int g3 = 0; g3 < 2 && e == hipSuccess; ++g3) {  // u, v, w then wx, wy, wz
    for (int f = 0; f < 3; ++f) la.LT[f] = c->T[3 * g3 + f];
    la.out = d + (size_t)3 * g3 * P;
    e = kmm::launch_mr_lap_plane(la, c->stream);
  }
  if (e == hipSuccess) {  // swirl_3d (calcswirl) from the NS plane's z intermediates
    kmm::PolyPlaneArgs sa;
    std::memset(&sa, 0, sizeof(sa));
    sa.nxh = c->nxh; sa.nz = c->nz; sa.nzh = c->nzh; sa.nyp = nyp; sa.mx = c->mx; sa.mz = c->mz;
    sa.mixed = c->cfg.precision == 1;
    sa.alpha = c->alpha; sa.beta = c->beta;
    sa.twZ = c->twZ; sa.twX = c->twX;
    sa.U = c->Z;
    sa.in[0] = c->T[1];
    sa.in[1] = c->T[2];
    sa.gz = 0;
    sa.ZP = c->ZP;
    sa.swirl = d + 6 * P;
    e = kmm::launch_swirl_plane(sa, true, c->stream);
  }
  if (e == hipSuccess && c->cfg.scalar) {
    e = kmm::launch_cheb(c->ny, c->P, c->twM, c->TH, lt, 2, c->cfg.precision, c->stream);  // pencil-major input
    for (int f = 0; f < 3; ++f) la.LT[f] = lt;
    la.out = d + 7 * P;
    if (e == hipSuccess) e = kmm::launch_mr_lap_plane(la, c->stream);
  }
  std::vector<double> h;
  if (e == hipSuccess) {
    h.resize((size_t)nf * P);
    e = hipMemcpyAsync(h.data(), d, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(d);
  hipFree(lz);
  if (lt) hipFree(lt);
  if (e != hipSuccess) return fail(KMM_E_HIP, std::string("flowfield: ") + hipGetErrorString(e));
  for (int f = 0; f < nf; ++f)
    for (int k = 0; k < nyp; ++k)
      for (int jz = 0; jz < c->mz; ++jz)
        for (int ix = 0; ix < c->mx; ++ix)
          out[(size_t)f * P + k + (size_t)nyp * (jz + (size_t)c->mz * ix)] =
              h[(size_t)f * P + ((size_t)k * c->mz + jz) * c->mx + ix];
  return KMM_OK;
}

// Fortran Ew.d output (d significant digits as 0.ddd...E+ee, right-justified
// in w columns; exponents beyond two digits drop the E)
static std::string fortran_e(double v, int w, int dgt) {
  char m[64];
  if (v == 0.0) {
    std::snprintf(m, sizeof m, "0.%0*dE+00", dgt, 0);
  } else if (!std::isfinite(v)) {
    std::snprintf(m, sizeof m, "%s", std::isnan(v) ? "NaN" : (v > 0 ? "Infinity" : "-Infinity"));
  } else {
    char t[64];
    std::snprintf(t, sizeof t, "%.*E", dgt - 1, std::fabs(v));  // d.ddd E+xx
    const char* ep = std::strchr(t, 'E');
    const int ex = std::atoi(ep + 1) + 1;
    std::string dig(1, t[0]);
    dig.append(t + 2, (size_t)(ep - (t + 2)));
    const int ax = ex < 0 ? -ex : ex;
    if (ax < 100) std::snprintf(m, sizeof m, "%s0.%sE%c%02d", v < 0 ? "-" : "", dig.c_str(), ex < 0 ? '-' : '+', ax);
    else std::snprintf(m, sizeof m, "%s0.%s%c%03d", v < 0 ? "-" : "", dig.c_str(), ex < 0 ? '-' : '+', ax);
  }
  std::string r(m);
  return r.size() >= (size_t)w ? r : std::string((size_t)w - r.size(), ' ') + r;
}

// Fortran Aw output of a shorter string: right-justified
static std::string fortran_a(const std::string& s, size_t w) {
  return s.size() >= w ? s.substr(0, w) : std::string(w - s.size(), ' ') + s;
}

// write_flowfield_ascii (helpers.f90:1671-1767): dir/time-XXXXXX.dat with the
// fields of kmm_flowfield, one point a line (8(e14.6,1x)), y outermost, then
// z (all mz planes when nz > 10, else the first), then x; at it = 1 also
// dir/grid.dat with x, y, z
extern "C" int kmm_write_flowfield(kmm_ctx* c, const char* dir, int it, int iprnfrq) {
  if (!c || !dir || iprnfrq <= 0) return fail(KMM_E_ARG, "bad argument");
  const int nyp = c->nyp, mx = c->mx, mz = c->mz;
  const size_t P = (size_t)nyp * mz * mx;
  const int nf = c->cfg.scalar ? 8 : 7;
  std::vector<double> f((size_t)nf * P);
  int r = kmm_flowfield(c, f.data());
  if (r) return r;
  std::vector<double> y(nyp);
  KMM_HIP(hipMemcpy(y.data(), c->ycoord_d, nyp * sizeof(double), hipMemcpyDeviceToHost));
  const int mzc = c->nz > 10 ? mz : 1;
  const double delxm = c->cfg.xl / mx, delzm = c->cfg.zl / mz;
  char name[64];
  std::snprintf(name, sizeof name, "/time-%06d.dat", it);
  const std::string path = std::string(dir) + name;
  FILE* fp = std::fopen(path.c_str(), "w");
  if (!fp) return fail(KMM_E_IO, "cannot open " + path);
  const std::string vars = c->cfg.scalar
      ? "filetype = solution, variables = \"swirl\", \"u\", \"v\", \"w\", \"wx\", \"wy\", \"wz\", \"scl\""
      : "filetype = solution, variables = \"swirl\", \"u\", \"v\", \"w\", \"wx\", \"wy\", \"wz\"";
  std::fprintf(fp, "%s\n", fortran_a(vars, 93).c_str());
  std::fprintf(fp, "%s%3d%s%4d%s%4d%s%4d%s\n", fortran_a("zone f=point t=\"Field\", solutiontime=", 40).c_str(),
               it / iprnfrq, fortran_a(",i=", 5).c_str(), mx, fortran_a("j=", 5).c_str(), 1,
               fortran_a("k=", 5).c_str(), nyp, " \n");
  static const int order[8] = {6, 0, 1, 2, 3, 4, 5, 7};  // swirl, u, v, w, wx, wy, wz, scl
  std::string line;
  for (int k = 0; k < nyp; ++k)
    for (int jz = 0; jz < mzc; ++jz)
      for (int ix = 0; ix < mx; ++ix) {
        line.clear();
        const size_t o = k + (size_t)nyp * (jz + (size_t)mz * ix);
        for (int q = 0; q < nf; ++q) {
          line += fortran_e(f[(size_t)order[q] * P + o], 14, 6);
          line += ' ';
        }
        line += '\n';
        std::fputs(line.c_str(), fp);
      }
  std::fclose(fp);
  if (it == 1) {
    const std::string gpath = std::string(dir) + "/grid.dat";
    FILE* gp = std::fopen(gpath.c_str(), "w");
    if (!gp) return fail(KMM_E_IO, "cannot open " + gpath);
    std::fprintf(gp, " filetype = grid, variables = \"x\", \"y\", \"z\"\n");
    std::fprintf(gp, " zone f=point t=\"Grid\", i=%12d j=%12d k=%12d\n", mx, 1, nyp);
    for (int k = 0; k < nyp; ++k)
      for (int jz = 0; jz < mzc; ++jz)
        for (int ix = 0; ix < mx; ++ix)
          std::fprintf(gp, "%s %s %s \n", fortran_e(ix * delxm, 14, 6).c_str(), fortran_e(y[k], 14, 6).c_str(),
                       fortran_e(jz * delzm, 14, 6).c_str());
    std::fclose(gp);
  }
  return KMM_OK;
}

extern "C" int kmm_write_outputs(kmm_ctx* c, const char* dir, int first) {
  if (!c || !dir) return fail(KMM_E_ARG, "null argument");
  const int nyp = c->nyp;
  std::vector<double> prof((size_t)9 * nyp), scal(6);
  int r = kmm_statistics(c, prof.data(), scal.data());
  if (r) return r;
  // file names and contents of writeoutputs (helpers.f90:1362-1436): one
  // value per line, y profiles top to bottom, appended on later calls
  struct Out { const char* name; int prof; int scal; };
  static const Out files[] = {{"Txy0", 5, -1},   {"urms0", 1, -1},  {"RSS0", 4, -1},   {"vrms0", 2, -1},
                              {"wrms0", 3, -1},  {"massflux0", -1, 0}, {"mean_u0", 0, -1}, {"betay0", 6, -1},
                              {"swirly0", 7, -1}, {"twss0", -1, 1},  {"bwss0", -1, 2},  {"tpss0", -1, 3},
                              {"bpss0", -1, 4},  {"Ckk0", -1, 5},   {"PSS0", 8, -1}};
  for (const Out& o : files) {
    const std::string path = std::string(dir) + "/" + o.name;
    FILE* fp = std::fopen(path.c_str(), first ? "w" : "a");
    if (!fp) return fail(KMM_E_IO, "cannot open " + path);
    if (o.prof >= 0)
      for (int i = 0; i < nyp; ++i) std::fprintf(fp, " %24.16E\n", prof[(size_t)o.prof * nyp + i]);
    else
      std::fprintf(fp, " %24.16E\n", scal[o.scal]);
    std::fclose(fp);
  }
  return KMM_OK;
}
