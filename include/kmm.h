/*
 * kmm.h -- C ABI of the MI355X KMM channel-step library (libkmm.so).
 *
 * The library replaces the time-step path of the Polymer_DNS reference
 * (FFTW_DNS/code/src/dns): the body of the time loop of ``program dns``
 * (main_dns.f90:198-811) together with the solver/derivative/helper routines
 * it calls and ``vcw3dp`` (solvers.f90:1726-3538).  Fortran hosts bind these
 * entry points through iso_c_binding; every argument is a plain pointer or
 * scalar.  Arrays passed in "reference layout" are Fortran column-major:
 *   spectral field  complex(8)  (nyp, nz, nxh)   -- y fastest
 *   mean profile    real(8)     (nyp)
 *   IB integrals    real(8)     (nyp, mz, mx)
 *   geometry        integer(4)  (nyp, mz, mx)
 * Complex numbers are interleaved (re, im) doubles.
 *
 * Every function returns a kmm_status (0 on success).  The library keeps
 * the state on the GPU between calls; one context per host thread.
 */
#ifndef KMM_H
#define KMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum kmm_status {
  KMM_OK = 0,
  KMM_E_CFL = 1,      /* CFL > 1 (solvers.f90:3503-3514 "CFL failure"), or the inertial
                         particle CFL >= 10 (part_track.f90:216-221) */
  KMM_E_PARTICLE = 2, /* particle left the domain (part_track.f90:498-504) */
  KMM_E_HIP = 3,      /* HIP runtime error */
  KMM_E_CONFIG = 4,   /* unsupported or inconsistent configuration */
  KMM_E_IO = 5,       /* file missing, short or malformed */
  KMM_E_ARG = 6       /* bad argument */
} kmm_status;

/* Runtime equivalent of dns.config (main_dns.f90:934-1068) plus the
 * compile-time sizes of grid_size.f90:2. */
typedef struct kmm_config {
  int nx, ny, nz;         /* NX1, NY1, NZ1 */
  int nsteps, iprnfrq;    /* NSTEPS, IPRNFRQ */
  double dt;              /* DT */
  double gain, ugain;     /* Goldstein feedback gains */
  double theta;           /* THETA (CN implicitness, <= 0.99) */
  double yl, zl, xl;      /* YL, ZL, XL */
  double re;              /* RE = 1/nu */
  double uinf;            /* Uinf */
  int kwall;              /* KWALL */
  int perturbtime;        /* perturbation staying time (imatrix type 4) */
  double r_tau;           /* R_tau (sets dPdx, init_flow.f90:117) */
  int irstrt, crstrt;     /* restart flags */
  int flow_select;        /* 1 channel, 2 2-D Couette, 3 3-D Couette */
  int scalar;             /* 1 = SCALAR build (passive scalar) */
  int scl_flag;           /* 0,1,3: no source; 2: Gaussian blobs of rate deltaT at the
                             particles between src_start and src_stop; 6: uniform source deltaT */
  double diff;            /* DIFF (Prandtl/Schmidt number) */
  double deltaT;          /* scalar source rate for scl_flag 6 */
  int src_start, src_stop;
  int cadence;
  int precision;          /* 0: FP64 throughout; 1: mixed (FP32 transforms, products
                             and transfer fields; FP64 state and tau solves) */
  int particle_flag;      /* 0: tracers; > 0: inertial Maxey-Riley particles (RK4,
                             part_track.f90:617-1075) */
  double ratio, ap, c_mu; /* particle density ratio, radius, C_mu */
  int cd_switch;          /* drag law 0-3 (part_track.f90:826-840); sets the particle
                             sub-step (part_track.f90:338-346) */
  /* FENE-P conformation tensor (POLYMER build; dns.config lines 35-37, 55-62) */
  int polymer;            /* 1 = POLYMER build */
  int itarget;            /* 0: uniform beta = qbeta (the only mode built) */
  int ipeter;             /* 1: Peterlin function (FENE-P), 0: Oldroyd-B */
  int ipolyflag;
  double zlmax;           /* maximum extensibility L */
  double tpoly;           /* relaxation time lambda */
  double alpha_poly;
  double diffpoly;        /* artificial diffusivity 1/(re diffpoly) */
  double c11z, c22z, c33z;/* initial diagonal of C */
  double qbeta;           /* solvent viscosity ratio beta */
  /* constant mass flux (no reference counterpart): every step the mean
   * pressure gradient is corrected so that the bulk velocity stays ubulk
   * (ubulk <= 0: the bulk velocity of the state at kmm_startup) */
  int cmf;
  double ubulk;
  int particle_swirl;     /* 1: swirl strength at the particles every step (swirl_part,
                             solvers.f90:3479-3481); kmm_config_read sets 1 */
  double sigmax, sigmay, sigmaz;  /* Gaussian blob widths of the scl_flag 2 source
                                     (dns.config line 65) */
  double gravity;         /* buoyancy on inertial particles, -gravity (Pi_f - Pi_p) on
                             dv/dt (dns.config line 23, part_track.f90:873) */
  int bftail;             /* BFTAIL: last x index of the buffer region (imatrix 6, 8),
                             dns.config line 13, forcepre main_dns.f90:1495-1511 */
  double xstart;          /* XSTART: x origin of the suction profile (imatrix 7),
                             dns.config line 18, main_dns.f90:1537-1546 */
} kmm_config;

typedef struct kmm_ctx kmm_ctx;

/* field identifiers for kmm_put_field / kmm_get_field */
enum {
  KMM_V = 0, KMM_FN, KMM_FNM1, KMM_OMY, KMM_GN, KMM_GNM1,    /* restart record 1 */
  KMM_U0, KMM_H1N, KMM_H1NM1, KMM_W0, KMM_H3N, KMM_H3NM1,   /* mean profiles */
  KMM_FXINTG, KMM_FYINTG, KMM_FZINTG,                       /* restart record 2 */
  KMM_SCALAR, KMM_SCN, KMM_SCNM1,                           /* SCALAR build */
  /* POLYMER build: independent components 11, 12, 13, 22, 23, 33 each */
  KMM_C11, KMM_C12, KMM_C13, KMM_C22, KMM_C23, KMM_C33,     /* conformation C */
  KMM_C11N, KMM_C12N, KMM_C13N, KMM_C22N, KMM_C23N, KMM_C33N,
  KMM_C11NM1, KMM_C12NM1, KMM_C13NM1, KMM_C22NM1, KMM_C23NM1, KMM_C33NM1,
  KMM_STR11N, KMM_STR12N, KMM_STR13N, KMM_STR22N, KMM_STR23N, KMM_STR33N,
  KMM_STR11NM1, KMM_STR12NM1, KMM_STR13NM1, KMM_STR22NM1, KMM_STR23NM1, KMM_STR33NM1,
  KMM_NFIELDS
};

const char* kmm_status_string(int status);

/* Positional dns.config reader with the reference's line semantics
 * (main_dns.f90:934-1068; replaces setstuff's read block). */
int kmm_config_read(const char* path, kmm_config* cfg);

/* Allocate device state for cfg on HIP device `device` (setstuff, gfcn,
 * denom and forcepre: main_dns.f90:148-163). */
int kmm_create(const kmm_config* cfg, int device, kmm_ctx** ctx);
int kmm_destroy(kmm_ctx* ctx);

/* imatrix cell types (main_dns.f90:1120-1122, read as (nyp,mz,mx));
 * types 0, 1 and 4 are supported. */
int kmm_set_geometry(kmm_ctx* ctx, const int32_t* imatrix);

int kmm_put_field(kmm_ctx* ctx, int field, const double* host);
int kmm_get_field(kmm_ctx* ctx, int field, double* host);

/* New-run start-up of ``initial`` (main_dns.f90:1923-2229) from the current
 * v, omega_y, u0, w0 (and scalar): nonlinear terms at it=0 and the AB2
 * history set equal to them. */
int kmm_startup(kmm_ctx* ctx);

/* Advance nsteps time steps (main_dns.f90:198-811).  Returns KMM_E_CFL with
 * the failing step and its CFL number in kmm_last_cfl if the CFL number
 * exceeds 1 (solvers.f90:3503-3514, where the reference stops).  Steps are
 * queued in chunks checked after they complete (64 steps, or one graph
 * replay of kmm_set_graph's length), so on KMM_E_CFL the state has advanced
 * to the end of the chunk that holds the failing step. */
int kmm_step(kmm_ctx* ctx, int nsteps);

/* gradv (derivatives.f90:6-86): the nine gradients u_ij = d u_i / d x_j
 * and the Laplacians of u, v, w, all spectral (nyp, nz, nxh). */
int kmm_gradv(kmm_ctx* ctx, const double* u, const double* v, const double* w, double* u11, double* u12,
              double* u13, double* u21, double* u22, double* u23, double* u31, double* u32, double* u33,
              double* lu, double* lv, double* lw);
/* calcswirl (helpers.f90:360-404): swirl strength sqrt(sum Im(lambda)^2) of
 * the velocity-gradient tensor at every point of physical arrays
 * real(8) (nyp, mz, mx), argument order as the reference's dummy list; 0
 * where the eigenvalues are all real (the reference returns rounding noise
 * there) or the result is not finite. */
int kmm_calcswirl(kmm_ctx* ctx, const double* u11, const double* u21, const double* u31, const double* u12,
                  const double* u22, const double* u32, const double* u13, const double* u23, const double* u33,
                  double* swirl);

/* Ensemble of independent channels (any grid sizes, one context each, same
 * device): advances every member by nsteps.  Members of one shape (grid,
 * precision, scalar flag, occupancy variants) whose steps are plain channel
 * steps (no polymer, particles, IB cells, scalar source window, closure or
 * timing) run together: one launch per stage and step for the whole group,
 * each channel a slice of the grid with its own argument block.  The other
 * members queue one chunk of steps (a graph replay where possible) on their
 * own streams, so groups and members run concurrently.  Returns the first
 * error; its message names the member or group.  No reference counterpart. */
int kmm_ensemble_step(kmm_ctx** ctxs, int n, int nsteps);
/* kmm_ensemble_step with flags: KMM_ENSEMBLE_NO_BATCH runs every member on
 * its own stream (no shared launches). */
#define KMM_ENSEMBLE_NO_BATCH 1
int kmm_ensemble_step_ex(kmm_ctx** ctxs, int n, int nsteps, int flags);
/* Number of this channel's steps that ran in a group's shared launches. */
int kmm_batched_steps(kmm_ctx* ctx, long long* n);

/* Capture `steps` consecutive time steps into a HIP graph and replay it
 * whenever kmm_step runs steps whose launches are all alike (NL terms
 * valid, polymer and IB time-lines past their switch points, tracers
 * tracking); the rest run as ordinary launches.  0 turns it off.  The CFL
 * buffer is cleared outside the graph.  No reference counterpart. */
int kmm_set_graph(kmm_ctx* ctx, int steps);

/* Mean pressure gradient dPdx (init_flow.f90:117; updated every step when
 * cmf is set) and its override. */
int kmm_get_dpdx(kmm_ctx* ctx, double* dpdx);
int kmm_set_dpdx(kmm_ctx* ctx, double dpdx);

/* Largest CFL number of the last step run and the step counter it. */
int kmm_last_cfl(kmm_ctx* ctx, double* cflmax, int* it);

/* outputs/last-restart (main_dns.f90:776-781) / setup/restart (1744-1747).
 * With irstrt = 2 in the configuration, kmm_read_restart reads a restart
 * written at half the streamwise resolution (nx/2; main_dns.f90:1748-1809)
 * into the low kx modes and the first mx/2 columns of the IB integrals, the
 * rest zero. */
int kmm_write_restart(kmm_ctx* ctx, const char* path);
int kmm_read_restart(kmm_ctx* ctx, const char* path);
/* c-restart (main_dns.f90:782-793, 1910-1916).  SCALAR only: the three
 * records scalar / scn / scnm1.  POLYMER: the reference's five records
 * (scalar, c11..c33 | scn, c11n..c33n | scnm1, c11nm1..c33nm1 |
 * str11n..str33n | str11nm1..str33nm1), the scalar fields written as zeros
 * when there is no scalar; c21, c31, c32 are written as c12, c13, c23 and
 * ignored on reading. */
int kmm_write_crestart(kmm_ctx* ctx, const char* path);
int kmm_read_crestart(kmm_ctx* ctx, const char* path);

/* --- per-routine entry points, reference argument semantics ----------- */

/* cderiv(f,df), derivatives.f90:258: f, df complex (nyp,nz,nxh) */
int kmm_cderiv(kmm_ctx* ctx, const double* f, double* df);

/* solve(u,x,g,dyde,bctop,bcbot,ib,a,wrkc), solvers.f90:174: one kx slab
 * u complex (nyp,nz) in/out, bctop/bcbot complex (nz); a/wrkc are internal */
int kmm_solve(kmm_ctx* ctx, double* u, double x, double g, double dyde, const double* bctop,
              const double* bcbot, int ib);

/* kmm_solve without allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, hosts that forbid hipMalloc).  Same arguments,
 * kernels and results as kmm_solve, with every device buffer taken from a
 * fixed array each GPU holds until the process ends (about 5.8 MB): slabs
 * up to nz <= 1024 and nyp*nz <= 131072 complex values; larger slabs return
 * KMM_E_ARG.  Calls that overlap on one GPU share the array and race. */
int kmm_solve_noalloc(kmm_ctx* ctx, double* u, double x, double g, double dyde, const double* bctop,
                      const double* bcbot, int ib);

/* penta(u,x,g,dyde,ib,at,bt,gtop,ab,bb,gbot,a,wrkc), solvers.f90:680 */
int kmm_penta(kmm_ctx* ctx, double* u, double x, double g, double dyde, int ib, double at, double bt,
              const double* gtop, double ab, double bb, const double* gbot);

/* vcw3dp(u,v,w,omx,omy,omz,fn,gn,u11..u33,[scalar,sclx,scly,sclz,scn],..),
 * solvers.f90:1726-1740, NS (+SCALAR) argument list in the reference's
 * order.  Inputs u, v, w, omx, omy, omz (+ sclx, scly, sclz) spectral;
 * outputs fn = (u x om)_y, gn = (u x om)_x, and omz, overwritten with
 * (u x om)_z as in the reference (+ scn = -u.grad(scalar) + source).  uij
 * (the nine velocity gradients u11, u12, ... u33) and scalar may be NULL:
 * they feed calcswirl and part_track, which kmm_step runs itself.  Updates
 * the IB integrals; returns KMM_E_CFL if the CFL number exceeds 1. */
int kmm_vcw3dp(kmm_ctx* ctx, const double* u, const double* v, const double* w, const double* omx,
               const double* omy, double* omz, double* fn, double* gn, const double* const* uij,
               const double* scalar, const double* sclx, const double* scly, const double* sclz, double* scn,
               int it);

/* vcw3dp with the POLYMER argument list (solvers.f90:1726-1740): as
 * kmm_vcw3dp, plus cij[9] = c11 c12 c13 c21 c22 c23 c31 c32 c33 and dc[27]
 * (derivscji's dummy order, see kmm_derivscji) in; cn[6], strn[6], qp[6]
 * (components 11 12 13 22 23 33: c_ij^n, str_ij^n, qp_ij) out, for it >=
 * src_start - 1 as in the reference (otherwise they are left as they are).
 * uij (u11 ... u33) is required: u12, u22, u32 enter the products.  C is
 * taken symmetric (c21 = c12, ...), as the step carries it. */
int kmm_vcw3dp_poly(kmm_ctx* ctx, const double* u, const double* v, const double* w, const double* omx,
                    const double* omy, double* omz, double* fn, double* gn, const double* const* uij,
                    const double* scalar, const double* sclx, const double* scly, const double* sclz, double* scn,
                    const double* const* cij, const double* const* dc, double* const* cn, double* const* strn,
                    double* const* qp, int it);

/* Spectral helper procedures on whole fields (complex (nyp,nz,nxh) each,
 * the reference's layout), with the reference's argument meaning.  Work
 * arrays of the reference (wrkc, wrk1) are internal. */
/* phirhs(v,wrkc,wrk1,fn,fnm1,bcbot,bctop,var), solvers.f90:555-676
 * (THETA <= 0.99): the CN/AB2 right-hand side replaces v; bcbot and bctop
 * (complex (nz,nxh), may be NULL) are zeroed as the reference does. */
int kmm_phirhs(kmm_ctx* ctx, double* v, const double* fn, const double* fnm1, double* bcbot, double* bctop,
               double var);
/* nonlin(fn,omz,gn,wrkc,wrk1), solvers.f90:1497-1572: on entry fn, omz, gn
 * hold (u x om)_y, _z, _x; on return fn and gn hold the nonlinear terms of
 * the v and omega_y equations (omz unchanged). */
int kmm_nonlin(kmm_ctx* ctx, double* fn, const double* omz, double* gn);
/* veloc(u,v,w,u0,w0,omy,wrkc), solvers.f90:1576-1659: u, w from continuity
 * and omega_y; the mean mode takes u0, w0 (real (nyp)); v(:,1,1) = 0. */
int kmm_veloc(kmm_ctx* ctx, double* u, double* v, double* w, const double* u0, const double* w0,
              const double* omy);
/* vort(u,v,w,omx,omz), solvers.f90:1663-1722 */
int kmm_vort(kmm_ctx* ctx, const double* u, const double* v, const double* w, double* omx, double* omz);
/* gradscl(scalar,sclx,scly,sclz), derivatives.f90:90-144 */
int kmm_gradscl(kmm_ctx* ctx, const double* scalar, double* sclx, double* scly, double* sclz);
/* derivscji(c11..c33, dc111..dc333), derivatives.f90:148-254: cij[9] =
 * c11 c12 c13 c21 c22 c23 c31 c32 c33; dc[27] in the reference's dummy
 * order dc111 dc112 dc113 dc211 dc212 dc213 dc311 dc312 dc313 dc121 ...
 * dc323 dc131 ... dc333, with dc_ijk = d c_ij / d x_k. */
int kmm_derivscji(kmm_ctx* ctx, const double* const* cij, double* const* dc);
/* polyforce(s11,s12,s13,s22,s23,s33,t1,t2,t3), helpers.f90:498-560 */
int kmm_polyforce(kmm_ctx* ctx, const double* s11, const double* s12, const double* s13, const double* s22,
                  const double* s23, const double* s33, double* t1, double* t2, double* t3);
/* subforce(gn,fn,omz,t1,t2,t3), helpers.f90:564-606: gn += t1, fn += t2,
 * omz += t3 */
int kmm_subforce(kmm_ctx* ctx, double* gn, double* fn, double* omz, const double* t1, const double* t2,
                 const double* t3);
/* polyNL(c11NL..c33NL, c11n.., c11nm1.., str11n.., str11nm1..),
 * helpers.f90:610-674: each argument is an array of the six components
 * 11 12 13 22 23 33; cNL = 3 (c_n - str_n/tpoly) - (c_nm1 - str_nm1/tpoly). */
int kmm_polyNL(kmm_ctx* ctx, double* const* cNL, const double* const* cn, const double* const* cnm1,
               const double* const* strn, const double* const* strnm1);
/* norm(a), helpers.f90:91-158: reality of the mean mode, Hermitian
 * symmetry of the kx = 0 plane, zero z-Nyquist row; in place. */
int kmm_norm(kmm_ctx* ctx, double* a);
/* uzero(u0,h1,h1l,uflag), solvers.f90:1065-1176: the mean-mode CN/AB2
 * Robin solve; u0, h1, h1l real (nyp); uflag 1 (u0) or 3 (w0, with the
 * reference's btu in the w rows, solvers.f90:1173).  u0 is replaced. */
int kmm_uzero(kmm_ctx* ctx, double* u0, const double* h1, const double* h1l, int uflag);

/* time spent in the first step of the last kmm_step call (ms): pencil
 * stage, plane stage */
int kmm_stage_times(kmm_ctx* ctx, double* pencil_ms, double* plane_ms);

/* Per-kernel HIP-event timing of every step of later kmm_step calls:
 * average ms per step of k_pencil, k_zinv, k_ibdrag, k_xplane, k_zfwd,
 * the polymer pencil kernel, the polymer plane kernels and the tracer kernel
 * (ms[8]) over the last kmm_step call.  on = 1: events around every kernel;
 * on = 2: around the pencil step kernel only (two events per step; the other
 * entries read 0), which keeps the events' own cost out of the step time. */
int kmm_set_timing(kmm_ctx* ctx, int on);
int kmm_kernel_times(kmm_ctx* ctx, double* ms, int* nsamples);

/* ---- Lagrangian tracers (part_track.f90) ----
 * Positions are on the 3/2 physical grid: x in [0, xl), y in [-yl/2, yl/2],
 * z in [0, zl).  After kmm_set_particles / kmm_read_particles the next flow
 * evaluation (kmm_startup, or the next step) sets each particle velocity to
 * the fluid velocity there (init_part, part_track.f90:229-292, run when
 * it == irstrt); every later step moves the particles by one explicit Euler
 * step of dt/substep and re-interpolates (tracer_update, :360-433, called from
 * vcw3dp, solvers.f90:3463-3466).  A particle leaving the grid makes
 * kmm_startup / kmm_step return KMM_E_PARTICLE. */

/* Replaces the particles.dat read of init_part (part_track.f90:259-267):
 * a header line, then one particle per line "x y z u v w tp" (u, v, w and
 * the release time tp are read and ignored, as in the reference). */
int kmm_read_particles(kmm_ctx* ctx, const char* path);
/* Same from arrays (xpart, ypart, zpart of grid_size.f90:39); n = 0 removes
 * all particles. */
int kmm_set_particles(kmm_ctx* ctx, int n, const double* x, const double* y, const double* z);
/* Number of particles set (npart of grid_size.f90:2, a runtime value here). */
int kmm_particle_count(kmm_ctx* ctx, int* n);
/* Current xpart..wpart (any pointer may be NULL). */
int kmm_get_particles(kmm_ctx* ctx, int n, double* x, double* y, double* z, double* u, double* v, double* w);
/* swirl_part: swirl strength at the particles (particle_swirl = 1) */
int kmm_get_particle_swirl(kmm_ctx* ctx, int n, double* swirl);
/* part-t-XXXXXX.dat writer of part_track (part_track.f90:196-206): Tecplot
 * point format, columns x y z u v w swirl in E14.6 (swirl 0 when
 * particle_swirl is off). */
int kmm_write_particles(kmm_ctx* ctx, const char* path, int it);
/* fluid_interp / fluid_interp1 (part_track.f90:437-613) + interpolate3
 * (:1184-1222): trilinear interpolation of nf physical fields given in
 * reference layout real(8) (nyp, mz, mx), stored back to back, at n points.
 * out is (n, nf) Fortran order (out[f*n + j]).  cells (may be NULL) receives
 * the 0-based imin, jmin, kmax of each point as (n, 3) Fortran order. */
int kmm_fluid_interp(kmm_ctx* ctx, int n, const double* xp, const double* yp, const double* zp, int nf,
                     const double* fields, double* out, int* cells);
/* up3d / vp3d / wp3d (solvers.f90:3113-3115) of the last flow evaluation in
 * reference layout real(8) (nyp, mz, mx); comp 0, 1, 2 = u, v, w.  Kept only
 * while particles are set. */
int kmm_get_physical(kmm_ctx* ctx, int comp, double* out);

/* ---- learned polymer-stress closure (no reference counterpart) ----
 * A 22 -> 128 -> 128 -> 128 -> 128 -> 6 perceptron, ReLU after the hidden
 * layers, bf16 weights and FP32 accumulation on the matrix cores.  Inputs per point: the
 * nine velocity gradients du_i/dx_j (row-major), the wall distance 1 - |y|,
 * the six stress components (11 12 13 22 23 33) of the previous step and their
 * change over the step before; output: six values (in the channel closure,
 * the correction to linear extrapolation of the stress).  Arrays are
 * point-major: X[n][22], Y[n][6].  Parameters: FP32 master vector of
 * kmm_mlp_param_count() floats, W_l (n_out x n_in, row-major) then b_l for
 * each layer. */
typedef struct kmm_mlp kmm_mlp;
int kmm_mlp_param_count(void);
int kmm_mlp_create(int device, unsigned long long seed, kmm_mlp** out);
int kmm_mlp_destroy(kmm_mlp* mlp);
int kmm_mlp_set_params(kmm_mlp* mlp, int np, const float* params);
int kmm_mlp_get_params(kmm_mlp* mlp, int np, float* params);
/* input/output normalisation x_n = (x - mu_x)/sd_x, y = y_n sd_y + mu_y */
int kmm_mlp_set_norm(kmm_mlp* mlp, const float* mu_x, const float* sd_x, const float* mu_y, const float* sd_y);
int kmm_mlp_infer(kmm_mlp* mlp, int n, const float* X, float* Y);
/* loss (mean over points of the summed squared normalised error) and its
 * gradient; mode 0: FP32, mode 1: bf16 rounding as in inference */
int kmm_mlp_grad(kmm_mlp* mlp, int n, const float* X, const float* T, int mode, float* grad, float* loss);
/* Adam (lr, beta 0.9/0.999) on seeded mini-batches; losses[steps] optional */
/* hidden-layer activation: 0 ReLU (default), 1 GELU as x sigma(1.702 x),
 * 2 GELU as x s(t) with t = clamp(1/2 + x/5.28, 0, 1) and the quintic
 * smoothstep s(t) = t^3 (t (6 t - 15) + 10) (no transcendental; within
 * 0.0195 of the exact GELU); applied in FP32 before the bf16 rounding of the
 * next layer's input */
int kmm_mlp_set_activation(kmm_mlp* mlp, int act);
int kmm_mlp_train(kmm_mlp* mlp, int n, const float* X, const float* T, int steps, int batch, float lr,
                  unsigned long long seed, float* losses);

/* Learned closure in the channel step (polymer = 1):
 * kmm_set_poly_capture keeps, at every FENE-P evaluation, the physical
 * velocity gradient and Peterlin stress (FP32, 3/2 grid) of the last three
 * steps; kmm_poly_samples draws n seeded (features, target) pairs from the
 * last three captured steps: features = du_i/dx_j (9), wall distance,
 * str(n-1) (6) and str(n-1) - str(n-2) (6, formed in FP32), target = the
 * correction to linear extrapolation str(n) - 2 str(n-1) + str(n-2); in the
 * step the network's output is added to 2 str(n-1) - str(n-2) in FP32.  kmm_set_closure(ctx, mlp) replaces the conformation
 * transport by the network from the next step on (needs the stress of the
 * two previous steps), NULL returns to FENE-P; kmm_get_poly_stress returns
 * the last stress component on the 3/2 grid (nyp, mz, mx). */
int kmm_set_poly_capture(kmm_ctx* ctx, int on);
int kmm_poly_samples(kmm_ctx* ctx, int n, unsigned long long seed, float* X, float* T);
int kmm_set_closure(kmm_ctx* ctx, kmm_mlp* mlp);
int kmm_get_poly_stress(kmm_ctx* ctx, int comp, double* out);

/* Residual gate and audit steps of the learned closure (no reference
 * counterpart).  With every > 0, step closure_start + m*every (m >= 1) opens
 * an audit window of `window` (>= 3, < every) steps that run the full FENE-P
 * transport instead of the network: C is rebuilt from the closure's last
 * stress (the inverse of str = f C - I: f = (L^2 + tr str)/L^2 for ipeter 1,
 * 1 otherwise) and stands for C at the window's first step; the next step
 * starts the transport history afresh.  From the window's third step on the
 * network predicts each step's stress from the FENE-P history and gradient;
 * the relative residual ||pred - str|| / ||str|| over those steps is the gate:
 * at or below `gate` the closure resumes from the FENE-P stress history, above
 * it the run stays on FENE-P (fallback) until kmm_set_closure is called again.
 * every = 0 turns audits off.  Needs the captures (kmm_set_poly_capture). */
int kmm_set_closure_audit(kmm_ctx* ctx, int every, int window, double gate);
/* Retraining of the closure at the end of every passed audit window: `steps`
 * Adam steps of `batch` pairs, learning rate lr, drawn (seeded) from the
 * window's FENE-P evaluations; steps = 0 turns it off.  No reference
 * counterpart. */
int kmm_set_closure_retrain(kmm_ctx* ctx, int steps, int batch, float lr, unsigned long long seed);
/* active: 1 while the network carries the stress; audits: completed windows;
 * last_residual: the last window's residual (-1 before the first);
 * fallback_it: the first FENE-P step after a failed audit, -1 if none */
int kmm_closure_status(kmm_ctx* ctx, int* active, int* audits, double* last_residual, int* fallback_it);
/* C replaced by the tensor rebuilt from the last captured stress (the audit's
 * restart of the transport); the pending plane outputs are kept */
int kmm_poly_conf_from_stress(kmm_ctx* ctx);

/* writeoutputs (helpers.f90:1227-1436, with the mean u of solvers.f90:
 * 3487-3495) for the last evaluation: prof[9][nyp] = mean u, u'_rms, v'_rms,
 * w'_rms, Reynolds shear stress <u'v>, mean viscous shear stress <du/dy>/re,
 * betay (beta_poly = qbeta in POLYMER builds, else the mean scalar), mean
 * swirl strength, polymer shear stress <qp12>; y from +1 to -1.  scal[6] =
 * mass flux (bulk velocity), top and bottom wall shear stress, top and bottom
 * polymer shear stress, volume mean of tr C (NaN when no conformation tensor
 * was evaluated).  Second moments come from the spectral coefficients
 * (Parseval on the 3/2 grid is exact), means from the (0,0) modes. */
int kmm_statistics(kmm_ctx* ctx, double* prof, double* scal);
/* the same statistics appended to the reference's output files in dir
 * (Txy0 urms0 RSS0 vrms0 wrms0 massflux0 mean_u0 betay0 swirly0 twss0 bwss0
 * tpss0 bpss0 Ckk0 PSS0), created when first != 0 */
int kmm_write_outputs(kmm_ctx* ctx, const char* dir, int first);

/* Physical fields of the last evaluation on the 3/2 grid, the arguments of
 * write_flowfield_ascii (helpers.f90:1671-1767): u, v, w, wx, wy, wz, swirl,
 * and the scalar when scalar = 1; out: 7 or 8 fields of nyp*mz*mx doubles,
 * each in the reference layout (nyp, mz, mx). */
int kmm_flowfield(kmm_ctx* ctx, double* out);
/* write_flowfield_ascii: dir/time-XXXXXX.dat (it = XXXXXX, solution time
 * it/iprnfrq) in the reference's Tecplot ASCII layout, and dir/grid.dat when
 * it = 1 */
int kmm_write_flowfield(kmm_ctx* ctx, const char* dir, int it, int iprnfrq);

/* Autotuner (no reference counterpart; the role of FFTW's PATIENT planning,
 * solvers.f90:2057-2069): occupancy variants of the pencil step kernel and of
 * the x-plane kernel (0: the compiler's register budget, 1: one more wave per
 * SIMD) are timed over `steps` steps on a twin of the context (same
 * configuration, geometry, tracers and state; the context itself does not
 * advance) and the faster of each is kept.  ms (or NULL) receives
 * {pencil v0, pencil v1, xplane v0, xplane v1} in ms per step. */
int kmm_autotune(kmm_ctx* ctx, int steps, double* ms);
int kmm_set_tuning(kmm_ctx* ctx, int pencil, int xplane);
int kmm_get_tuning(kmm_ctx* ctx, int* pencil, int* xplane);

#ifdef __cplusplus
}
#endif
#endif /* KMM_H */
