// Kernel argument blocks and launch entry points shared by kmm_api.cpp and
// the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

namespace kmm {

// Per-pencil table of the 1/d reciprocals of eqsolve/pntsolve
// (solvers.f90:500-553, 931-935): entry g (x: even, y: odd) for the
// coefficient pairs g = 0..ny/2-1, and the even extra row's 1/d in .x of the
// last entry; 65 entries for ny <= 128.
__host__ __device__ constexpr int tab_width(int ny) { return ny / 2 <= 64 ? 65 : ny / 2 + 1; }

// Pencil-stage arguments.  Field buffers are [y][kx][kz] complex128; a
// "pencil" p = kx*nz + kz; P = nxh*nz pencils per y-plane.
struct PencilArgs {
  int nxh, nz, nzh, P, ny;
  int tune;             // occupancy variant of the step kernel (autotuner)
  double alpha, beta, dyde, re, dt, theta, diff;
  double gH, gS;  // re/(dt theta), re diff/(dt theta)
  // mean-mode Robin data (setbcs, main_dns.f90:1175-1418)
  double atu, btu, gtu, abu, bbu, gbu, atw, btw, gtw, abw, bbw, gbw;
  int scalar;
  int smooth;       // apply spectral smoothing to the NL terms
  int first_nl;     // NL terms come from the NL buffers (else from FN/GN/SCN)
  int mixed;           // 1: transfer/NL fields and their transforms in FP32
  const double* smz;   // [nz] z smoothing factors
  const double* smy;   // [nyp] y smoothing factors
  const void* twM;     // FFT twiddles for length 2*ny (transfer precision)
  const double2* tabH; // [P][tab_width(ny)] 1/d tables, wn = k2 + gH
  const double2* tabP; // wn = k2
  const double2* tabS; // wn = k2 + gS
  const double2* G1;   // [P][ny/2+1] Green's functions (real, lane-pair packed)
  const double2* G2;
  const double4* gk;   // [P] influence coefficients (k1b,k1t,k2b,k2t)
  const void *NL1, *NL2, *NL3, *NLS;   // physical-y NL: (u x w)_x,_y,_z, scalar (transfer precision)
  double2 *V, *OMY, *FN, *FNM1, *GN, *GNM1, *TH, *SCN, *SCNM1;
  double *u0, *w0, *h1n, *h1nm1, *h3n, *h3nm1;
  void* T[9];          // transfer: u v w omx omy omz thx thy thz (physical y, transfer precision)
  // FENE-P coupling (nullptr when off)
  const double2* TF[3];  // polymer force div(qp) (spectral) added to NL1..NL3 before smoothing
  void* TPD[3];          // transfer: D u, D v, D w (physical y, transfer precision)
  // constant mass flux: correct u0 and the device-resident dPdx
  int cmf;
  double ubulk;
  double* dpdx;
};

// Per-routine entries (kmm_routines.hip): one spectral helper procedure of
// the reference on whole FP64 fields [y][kx][kz].
enum RoutineOp {
  RT_PHIRHS = 0, RT_NONLIN, RT_VELOC, RT_VORT, RT_GRAD, RT_POLYFORCE, RT_SUBFORCE, RT_POLYNL, RT_NORM
};
struct RoutineArgs {
  int op;
  int nxh, nz, nzh, P;
  double alpha, beta, dyde, re, dt, theta;
  double par;                // phirhs: var; polyNL: 1/tpoly
  const double2* in[6];
  double2* out[3];
  double *m0, *m1, *m2, *m3;  // real mean profiles (nyp)
  double rb[6];              // uzero Robin rows: at, bt, gt, ab, bb, gb
};
hipError_t launch_routine(int ny, const RoutineArgs& a, hipStream_t s);
hipError_t launch_uzero(int ny, const RoutineArgs& a, const double2* tab, hipStream_t s);

// Per-pencil FENE-P stage (main_dns.f90:206-321, helpers.f90:498-674 and the
// spectral half of the polymer terms of vcw3dp).
struct PolyArgs {
  int nxh, nz, nzh, P, ny;
  double alpha, beta, dyde, re, dt, theta;
  double diffpoly, tpoly, zbeta1, gC;  // gC = re diffpoly/(dt theta)
  double c0[3];                        // c11z, c22z, c33z
  int partB;     // forward-transform NLP into c_n, str_n (else load CN, STRN)
  int closure;   // learned closure: NLP holds only str_n (no c_n, no conformation)
  int force;     // write TF = div(zbeta1 str_n) (unsmoothed)
  int copy;      // CNM1/STRNM1 = unsmoothed c_n/str_n (n -> n-1 before smoothing)
  int smooth;
  int init;      // C = diag(c11z, c22z, c33z) before the solve
  int solve;     // polyNL + phirhs + Neumann penta
  int evaluate;  // write the transfer fields C, D C
  const double* smz;
  const double* smy;
  const void* twM;
  const double2* tabC;  // [P][65] 1/d tables, wn = k2 + gC
  const void* NLP[12];  // c_n (6), str_n (6): physical y, transfer precision
  double2* C[6];
  double2* CN[6];
  double2* CNM1[6];
  double2* STRN[6];
  double2* STRNM1[6];
  double2* TF[3];
  void* TP[12];         // transfer: C (6), D C (6)
};
// mode 0: step, 1: startup (evaluate), 2: finish (part B), 3: C from the
// physical-y fields in NLP[0..5] (forward Chebyshev transform only), 4: the
// force of a closure stress in NLP[6..11] (part B without C)
hipError_t launch_poly_pencil(int ny, int mode, const PolyArgs& a, bool mixed, hipStream_t s);

// Plane-stage arguments.
struct PlaneArgs {
  int nxh, nz, nzh, nyp, mx, mz, P;
  int nfin, nfout;     // input / output field counts (6/3 NS, 9/4 with scalar)
  int mixed;           // 1: FP32 transforms and products
  double dt, dPdx, delxm, delzm, src_uniform;
  const double* dpdx_d;  // device-resident dPdx (read instead of dPdx when set)
  int flow_select;
  const double* seght;
  const void* twZ;     // length mz twiddles
  const void* twX;     // length mx twiddles
  const void* in[9];   // spectral-xz, physical-y inputs [y][kx][kz]
  void* Z;             // [nfin][nyp][mz][nxh]
  void* ZF;            // [nfout][nyp][mz][nxh]
  void* out[4];        // NL [y][kx][kz]
  unsigned long long* cfl;  // bit pattern of the max CFL number for this call
  // immersed boundary (Goldstein feedback, solvers.f90:2996-3108)
  int nsrc;
  const int4* src;        // (y, z, x, imatrix type) per solid cell
  const double* srcp;     // [nsrc][6] per-type constants (buffer gains, initial
                          // profile, suction and wall velocities)
  double* fint;           // [3][nsrc] feedback integrals
  double* drag;           // [3][nsrc]
  double gain, ugain;
  int it, perturbtime;
  const int* trow;        // [nyp*mz+1] first target of each (y,z) row
  const int* tx;          // x index per target
  const int* tptr;        // [ntgt+1] first entry per target
  const int* ent_src;     // source index per entry
  const double* ent_w;    // fspread weight per entry
  const int* fixrow;      // [nyp*mz+1] solid cells per (y,z) row (no -dPdx)
  const int* fixx;        // x index per solid cell
  double* cap;            // nullptr, or [3][nyp][mz][mx] physical u, v, w (FP64)
  const double* scsrc;    // nullptr, or [nyp][mz][mx] blob source of the scalar (added / dt)
  unsigned dzmask;        // k_zinv: bit f set -> input f is multiplied by i kz first
  int tune;               // occupancy variant of k_xplane (autotuner)
  double beta;            // z wavenumber unit
};

// Plane stage of the FENE-P terms (solvers.f90:2901-2983 and the transforms
// around them).  Z-transformed inputs, [f][nyp][mz][nxh] at transfer precision.
struct PolyPlaneArgs {
  int nxh, nz, nzh, nyp, mx, mz;
  int mixed;
  double alpha, beta;
  int ipeter;
  double L2;              // zlmax^2
  const void* twZ;
  const void* twX;
  const void* U;          // NS z intermediates: u, v, w, wx, wy, wz (fields 0..5)
  const void* in[24];     // spectral-xz inputs of the z stage (see kmm_api.cpp)
  void* ZP;               // [24]: dz u(3), dy u(3), C(6), dz C(6), dy C(6)
  void* ZFP;              // [12]: c_n(6), str_n(6) after the forward x transform
  void* out[12];          // NLP outputs [y][kx][kz]
  double* swirl;          // [nyp][mz][mx] swirl strength (launch_swirl_plane)
  float* capg;            // nullptr or [9][P] physical du_i/dx_k, P = nyp mz mx, [y][z][x]
  float* caps;            // nullptr or [6][P] physical Peterlin stress (11 12 13 22 23 33)
  int gz;                 // ZP row of dz v (dz w follows) for the gradient/swirl plane
  double* mrf;            // nullptr or [12][P] FP64: wx, wy, wz, then du_i/dx_k at 3 + 3 i + k
};
hipError_t launch_poly_plane(const PolyPlaneArgs& a, hipStream_t s);
// velocity gradient (capg) and/or swirl strength on the 3/2 grid; with zstage
// the z transforms of dz v, dz w (from in[0], in[1]) into ZP rows 0, 1 run
// first (gz = 0); after the polymer stage they are its ZP rows 1, 2 (gz = 1)
hipError_t launch_swirl_plane(const PolyPlaneArgs& a, bool zstage, hipStream_t s);
// learned closure: forward transforms of the stress fields caps[6][P] into out[6..11]
hipError_t launch_closure_fwd(const PolyPlaneArgs& a, hipStream_t s);

// MLP on field-major inputs: features of point p are g[k][p] (k < 9), the
// wall distance 1 - |ycoord[p / (mz mx)]| / (yl/2), h1[j][p] and
// h1[j][p] - h2[j][p] (j < 6); the network gives the correction to linear
// extrapolation: out[j][p] = 2 h1[j][p] - h2[j][p] + y_j
struct MlpFieldIO {
  size_t P;
  int plane_pts;           // mz * mx
  double half;             // yl / 2
  const double* ycoord;
  const float* g;
  const float* h1;
  const float* h2;
  float* out;
};
// act: 0 ReLU, 1 GELU (sigmoid form)
hipError_t launch_mlp_field(const MlpFieldIO& io, const void* wb, const float* par, const float* norm, int act,
                            hipStream_t s);
// training samples from captured fields: for i < n, point p_i = hash(seed, i)
// mod P; X[i] = features of p_i, T[i] = t - 2 h1 + h2 at p_i
hipError_t launch_mlp_sample(const MlpFieldIO& io, const float* t, int n, unsigned long long seed, float* X,
                             float* T, hipStream_t s);

// Lagrangian tracers (part_track.f90): particle j has x, y, z, u, v, w at
// part[0..5][j]; fields are read at fld[f*fstride + k*sy + jz*sz + ix*sx]
// (k: y index, jz: z index, ix: x index on the 3/2 grid).
struct TracerArgs {
  int n, mx, mz, ny;
  double xl, yl, zl, delxm, delzm, pi;
  double dt, ap;            // tracer step dt/substep; particle radius
  const double* ycoord;     // [ny+1]
  const double* fld;
  long long fstride, sy, sz, sx;
  int nf;                   // fields interpolated (3 for tracers)
  int update;               // 1: Euler position update before interpolation
  double* part;             // [6][n] (tracer mode)
  const double* px;         // query positions (interpolation mode, part == nullptr)
  const double* py;
  const double* pz;
  double* out;              // [nf][n] (interpolation mode)
  int* cells;               // nullptr or [3][n]: imin, jmin, kmax (0-based)
  int* err;                 // set to 1 when a particle leaves the grid
};
hipError_t launch_tracer(const TracerArgs& a, hipStream_t s);
// Laplacians of u, v, w (gradv's Lu, Lv, Lw, derivatives.f90:77-79) on the 3/2
// grid for inertial particles: the pencil stage forms u, w from v, omega_y
// (veloc) and D^2 f - k^2 f into LT (physical y, transfer precision); the
// plane stage transforms them in z (LZ) and x into out[3][P] (FP64, [y][z][x])
struct MrLapArgs {
  int P, nz, nzh, nxh, nyp, mx, mz, mixed;
  double alpha, beta, dyde;
  const void* twM;
  const void* twZ;
  const void* twX;
  const double2* V;         // v, omega_y state (Chebyshev, [y][kx][kz])
  const double2* OMY;
  const double* u0;
  const double* w0;
  void* LT[3];
  void* LZ;                 // [3][nyp][mz][nxh]
  double* out;
};
hipError_t launch_mr_lap_pencil(int ny, const MrLapArgs& a, hipStream_t s);
hipError_t launch_mr_lap_plane(const MrLapArgs& a, hipStream_t s);
// Inertial (Maxey-Riley) particles, RK4 over substep sub-steps of t.dt
// (part_track.f90:131-150, 617-1075): one thread per particle; fields on the
// 3/2 grid [y][z][x]; err bit 1: a particle left the grid, bit 2: particle
// CFL (max |u_f - u_p| dt / a >= 10, part_track.f90:211-221)
struct MrArgs {
  TracerArgs t;             // geometry, part [7][n], err, ycoord, dt / substep, ap
  const double* cap;        // u, v, w, swirl [4][P]
  const double* mrf;        // wx, wy, wz, du_i/dx_k [12][P]
  const double* lap;        // Laplacians [3][P]
  const double* uold;       // u, v, w of the previous evaluation [3][P]; replaced by the
                            // rates (u - u_old)/dt (dt = t.dt) before the integration
  const double* lold;       // Laplacians of the previous evaluation [3][P]; replaced by
                            // their rates likewise
  int substep, cd_switch, swirl;
  double re, ratio, c_mu, gravity, dtfull;
};
hipError_t launch_inertial(const MrArgs& a, hipStream_t s);
// scalar source of Gaussian blobs at the particles (scl_flag 2), summed into
// src[nyp][mz][mx] (cleared first)
struct BlobArgs {
  int n, mx, mz, nyp;
  double delxm, delzm, sigmax, sigmay, sigmaz, deltaT;
  const double* ycoord;
  const double* part;       // [6][n]: x, y, z first
  double* src;
};
hipError_t launch_blob_source(const BlobArgs& a, hipStream_t s);
// writeoutputs statistics (kmm_stats.hip): out[8][nyp], see the kernel file
struct StatsArgs {
  int P, nz, nyp, mixed;
  const void* u;
  const void* v;
  const void* w;
  const void* s;          // optional (0,0)-mode fields, transfer layout
  const void* q;
  const void* c11;
  const void* c22;
  const void* c33;
  double* out;
};
hipError_t launch_stats(const StatsArgs& a, hipStream_t s);
hipError_t launch_plane_mean(const double* f, int nyp, int n, double* out, hipStream_t s);
// closure audits (kmm_audit.hip): C = (str + I)/f over P points of [6][P]
// arrays (str replaced by 2 str - prev when prev is given); residual sums acc[0] += sum (pred - ref)^2, acc[1] += sum ref^2
// over n values (part: 2 * audit_part_count() doubles)
hipError_t launch_conf_from_stress(const float* str, const float* prev, float* C, size_t P, double L2, int ipeter,
                                  hipStream_t s);
int audit_part_count();
hipError_t launch_audit_residual(const float* pred, const float* ref, size_t n, double* part, double* acc,
                                 hipStream_t s);
// learned closure (kmm_mlp.hip)
int mlp_param_count();
int mlp_wb_count();
int mlp_fb_blocks(int n);
hipError_t launch_mlp_infer(int n, const float* X, float* Y, const void* wb, const float* par, const float* norm,
                            int act, hipStream_t s);
hipError_t launch_mlp_pack(const float* par, void* wb, hipStream_t s);
// mode bit 0: bf16 rounding as in inference; bit 1: GELU (else ReLU)
hipError_t launch_mlp_grad(int n, const float* X, const float* T, const float* par, const float* norm, int mode,
                           float* gpart, float* lpart, float* g, float* loss, hipStream_t s);
hipError_t launch_mlp_adam(float* par, float* m, float* v, const float* g, float lr, int t, void* wb,
                           hipStream_t s);
hipError_t launch_gradv(int ny, int P, int nz, double alpha, double beta, double dyde, const double2* u,
                        const double2* v, const double2* w, double2* const* out, hipStream_t s);
// swirl strength of n velocity-gradient tensors; g[k*n + j] = component k
// (u11 u12 u13 u21 ... u33 order) of point j
hipError_t launch_swirl(int n, const double* g, double* swirl, hipStream_t s);

hipError_t launch_tau_tables(int ny, const PencilArgs& a, double g, double2* tab, hipStream_t s);
hipError_t launch_green(int ny, const PencilArgs& a, double pb1, double pt1, double pb2, double pt2,
                        double2* G1, double2* G2, double4* gk, hipStream_t s);
hipError_t launch_pencil_step(int ny, const PencilArgs& a, hipStream_t s);
hipError_t launch_pencil_startup(int ny, const PencilArgs& a, hipStream_t s);
hipError_t launch_pencil_finish(int ny, const PencilArgs& a, hipStream_t s);
hipError_t launch_plane(const PlaneArgs& a, hipStream_t s);
hipError_t launch_plane_fwd_only(const PlaneArgs& a, hipStream_t s);
// ev[0..3] recorded after k_zinv, k_ibdrag, k_xplane, k_zfwd
hipError_t launch_plane_timed(const PlaneArgs& a, hipStream_t s, hipEvent_t* ev);
// one step of nmem channels of the same size in one launch per stage: dA
// (device) holds each channel's arguments; a0 is a host copy of the first
// (sizes, precision, scalar flag and occupancy variants).  Channels without
// IB cells (k_ibdrag is not part of the batched plane stage).
hipError_t launch_pencil_step_batch(int ny, const PencilArgs& a0, const PencilArgs* dA, int nmem, hipStream_t s);
hipError_t launch_plane_batch(const PlaneArgs& a0, const PlaneArgs* dA, int nmem, hipStream_t s);

// per-routine (parity) kernels
hipError_t launch_cderiv(int ny, int P, double dyde, const double2* in, double2* out, hipStream_t s);
hipError_t launch_solve(int ny, const PencilArgs& a, const double2* tab, const double* wn, double2* u,
                        const double2* bctop, const double2* bcbot, int ib, const int* kzero, hipStream_t s);
hipError_t launch_penta(int ny, const PencilArgs& a, const double2* tab, const double* wn, double2* u,
                        double at, double bt, const double2* gtop, double ab, double bb, const double2* gbot,
                        int bottom_sign, const int* kzero, hipStream_t s);
hipError_t launch_tau_tables_wn(int ny, int P, double dyde, const double* wn, double2* tab, hipStream_t s);
hipError_t launch_cheb(int ny, int P, const void* tw, const void* in, void* out, int inverse, int mixed,
                       hipStream_t s);

}  // namespace kmm

// device view of a learned-closure network (kmm_mlp_api.cpp)
struct kmm_mlp;
bool kmm_mlp_view(kmm_mlp* m, const void** wb, const float** par, const float** norm, int* device);
