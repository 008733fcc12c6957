// Pencil-stage kernels of the KMM channel step (wall-normal direction).
//
// k_pencil_step fuses, per (kx,kz) pencil, the work the reference spreads
// over FFTW_DNS/code/src/dns/main_dns.f90:241-811 and the spectral half of
// vcw3dp: forward Chebyshev transform of the previous step's nonlinear terms
// (solvers.f90:3279-3416), smoothing (main_dns.f90:710-745), nonlin
// (solvers.f90:1497-1572), the scalar and phi/v/omega_y CN/AB2 solves
// (phirhs/solve/penta, main_dns.f90:324-484), the influence-matrix
// correction (main_dns.f90:433-466), the mean-flow solves (uzero,
// solvers.f90:1065-1176), veloc/vort/gradscl and the inverse Chebyshev
// transforms (solvers.f90:2241-2546).  Projections of ``norm`` that couple
// +kz and -kz are applied once per step in the plane stage (they commute
// with every pencil operation); the pencil-local ones are applied here.
#include "kmm_kernels.h"
#include "kmm_pencil.h"

namespace kmm {

// ---------------------------------------------------------- 1/d tables --
// d recurrence of eqsolve/pntsolve (solvers.f90:500-553, 931-935) per pencil
// and parity; thread per pencil.
template <int NY>
__global__ void k_tau_tables(int P, int nz, double alpha, double beta, double dyde, double g,
                             const double* __restrict__ wnv, double2* __restrict__ tab) {
  constexpr int L = NY / 2;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  double wn;
  if (wnv) {
    wn = wnv[p];
  } else {
    const int kx = p / nz, kzi = p % nz;
    const int kzs = (kzi < nz / 2) ? kzi : kzi - nz;
    const double kxw = kx * alpha, kzw = kzs * beta;
    wn = kxw * kxw + kzw * kzw + g;
  }
  constexpr int TW = tab_width(NY);
  double2* t = tab + (size_t)p * TW;
  for (int i = 0; i < TW; ++i) t[i] = make_double2(0.0, 0.0);
  // even: m = L
  double a1, a2, a3, b1, b2, b3;
  coef_even<NY>(L, wn, dyde, a1, a2, a3);
  double d = a2;
  t[TW - 1].x = 1.0 / d;
  double dn = d;  // d_{i+1}
  for (int i = L - 1; i >= 1; --i) {
    coef_even<NY>(i, wn, dyde, a1, a2, a3);
    coef_even<NY>(i + 1, wn, dyde, b1, b2, b3);
    d = a2 - a3 * b1 / dn;
    t[i].x = 1.0 / d;
    dn = d;
  }
  // odd: m = L-1
  coef_odd<NY>(L - 1, wn, dyde, a1, a2, a3);
  d = a2;
  t[L - 1].y = 1.0 / d;
  dn = d;
  for (int i = L - 2; i >= 1; --i) {
    coef_odd<NY>(i, wn, dyde, a1, a2, a3);
    coef_odd<NY>(i + 1, wn, dyde, b1, b2, b3);
    d = a2 - a3 * b1 / dn;
    t[i].y = 1.0 / d;
    dn = d;
  }
}


// ------------------------------------------------- Green's functions --
// gfcn (solvers.f90:5-172) with real boundary data: the reference's complex
// data (pt,pt) only scales each pencil's solution by a constant that
// cancels in c1*gf1 + c2*gf2, so the real solution and real influence
// coefficients give the same composite v.
template <int NY>
__global__ __launch_bounds__(PW * 64) void k_green(PencilArgs A, double pb1, double pt1, double pb2, double pt2,
                                                   double2* __restrict__ G1, double2* __restrict__ G2,
                                                   double4* __restrict__ gk) {
  constexpr int L = NY / 2;
  using PF = PFN<NY>;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int p = blockIdx.x * PW + wv;
  if (p >= A.P) return;
  double kx, kz;
  int kxi, kzi;
  pencil_k(A, p, kx, kz, kxi, kzi);
  const double k2 = kx * kx + kz * kz;
  const double2* tH = A.tabH + (size_t)p * tab_width(NY);
  const double2* tP = A.tabP + (size_t)p * tab_width(NY);
  double2 gt[2], gb[2];
  for (int w = 0; w < 2; ++w) {
    const double pb = w ? pb2 : pb1, pt = w ? pt2 : pt1;
    PF z = pf_zero<NY>();
    PF phi = tau_solve<NY>(z, k2 + A.gH, A.dyde, tH, make_double2(pt, 0.0), make_double2(pb, 0.0), 0, lane);
    PF g = tau_solve<NY>(phi, k2, A.dyde, tP, czero(), czero(), 0, lane);
    g = pf_real(g);
    double2* G = (w ? G2 : G1) + (size_t)p * (L + 1);
#pragma unroll
    for (int q = 0; q < PLay<NY>::S; ++q) {
      const int gi = q * WAVE + lane;
      if (gi < L) G[gi] = make_double2(g.a[q].x, g.b[q].x);
    }
    if (lane == 0) G[L] = make_double2(g.e.x, 0.0);
    wall_derivs<NY>(g, A.dyde, lane, gt[w], gb[w]);
  }
  if (lane == 0) {
    const double g1b = gb[0].x, g1t = gt[0].x, g2b = gb[1].x, g2t = gt[1].x;
    const double den = g1t * g2b - g2t * g1b;
    // c1 = (vb g2t - vt g2b)/den, c2 = (vt g1b - vb g1t)/den
    gk[p] = make_double4(g2t / den, -g2b / den, -g1t / den, g1b / den);
  }
}

// ----------------------------------------------------------- the step --
template <int NY>
__device__ __forceinline__ PFN<NY> green_load(const double2* __restrict__ G, int p, int lane) {
  constexpr int L = NY / 2;
  const double2* g = G + (size_t)p * (L + 1);
  PFN<NY> f = pf_zero<NY>();
#pragma unroll
  for (int q = 0; q < PLay<NY>::S; ++q) {
    const int gi = q * WAVE + lane;
    if (gi < L) { double2 v = g[gi]; f.a[q].x = v.x; f.b[q].x = v.y; }
  }
  f.e.x = g[L].x;
  return f;
}


// MODE 0: full step (NL from buffers or FN/GN/SCN), MODE 1: startup
// (derived fields of the current state only), MODE 2: finish (NL -> FN, GN,
// SCN, h1n, h3n).
// TV: occupancy variant picked by the autotuner (0: the compiler's register
// budget, 2 waves/SIMD at NY = 128; 1: at least 3 waves/SIMD, <= 168 VGPRs)
// The body of one block of pencils (blk: the block's index along its
// channel's pencils).
template <int NY, bool SCALAR, int MODE, typename V>
__device__ __forceinline__ void pencil_body(const PencilArgs& A, int blk) {
  using PF = PFN<NY>;
  extern __shared__ __align__(16) unsigned char smraw[];
  V* smem = reinterpret_cast<V*>(smraw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int p = blk * PW + wv;
  const V* twl = lds_twiddles<NY, V>(smem, static_cast<const V*>(A.twM));
  if (p >= A.P) return;
  V* buf = wave_lds<NY, V>(smem, wv);
  V* scr = buf + FFTBuf<2 * NY>::len;
  V* const* TT = reinterpret_cast<V* const*>(A.T);
  const size_t P = (size_t)A.P;
  double kx, kz;
  int kxi, kzi;
  pencil_k(A, p, kx, kz, kxi, kzi);
  const double k2 = kx * kx + kz * kz;
  const bool is00 = (kxi == 0 && kzi == 0);
  const bool nyq = (kzi == A.nzh);
  const double dy = A.dyde;
  const int nT = SCALAR ? 9 : 6;

  if (nyq) {  // norm zeroes the z-Nyquist row (helpers.f90:149-156)
    PF z = pf_zero<NY>();
    if (MODE == 0) {
      pf_store_pm<NY>(A.V, p, lane, z); pf_store_pm<NY>(A.OMY, p, lane, z);
      pf_store_pm<NY>(A.FNM1, p, lane, z); pf_store_pm<NY>(A.GNM1, p, lane, z);
      if (SCALAR) { pf_store_pm<NY>(A.TH, p, lane, z); pf_store_pm<NY>(A.SCNM1, p, lane, z); }
    }
    if (MODE == 2) {
      pf_store_pm<NY>(A.FN, p, lane, z); pf_store_pm<NY>(A.GN, p, lane, z);
      if (SCALAR) pf_store_pm<NY>(A.SCN, p, lane, z);
    }
    if (MODE != 2)
      for (int t = 0; t < nT + 3; ++t) {
        V* d = (t < nT) ? TT[t] : static_cast<V*>(A.TPD[t - nT]);
        if (!d) continue;
        pf_store_zero<NY, V>(d, P, p, lane);
      }
    return;
  }

  // ---------------- part B: nonlinear terms -> fn, gn, scn (spectral) --
  PF fn, gn, sn;
  PF h1m, h3m;  // mean-mode profiles (only used by the (0,0) pencil)
  if (MODE == 0 || MODE == 2) {
    if (MODE == 2 || A.first_nl) {
      const double sz = A.smooth ? A.smz[kzi] : 1.0;
      // each pencil's load is issued one transform ahead
      PhysN<NY, V> q1 = phys_load<NY, V>(static_cast<const V*>(A.NL1), P, p, lane);
      PhysN<NY, V> q3 = phys_load<NY, V>(static_cast<const V*>(A.NL3), P, p, lane);
      PF h1 = phys_to_cheb<NY, V>(q1, buf, scr, twl, lane);
      PhysN<NY, V> q2 = phys_load<NY, V>(static_cast<const V*>(A.NL2), P, p, lane);
      PF h3 = phys_to_cheb<NY, V>(q3, buf, scr, twl, lane);
      PhysN<NY, V> qs;
      if (SCALAR) qs = phys_load<NY, V>(static_cast<const V*>(A.NLS), P, p, lane);
      PF h2 = phys_to_cheb<NY, V>(q2, buf, scr, twl, lane);
      if (A.TF[0]) {  // polymer force, subforce (helpers.f90:564-606), before smoothing
        h1 = pf_add(h1, pf_load<NY>(A.TF[0], P, p, lane));
        h2 = pf_add(h2, pf_load<NY>(A.TF[1], P, p, lane));
        h3 = pf_add(h3, pf_load<NY>(A.TF[2], P, p, lane));
      }
      if (A.smooth) {
        h1 = smooth_apply<NY>(h1, A.smy, sz, lane);
        h2 = smooth_apply<NY>(h2, A.smy, sz, lane);
        h3 = smooth_apply<NY>(h3, A.smy, sz, lane);
      }
      if (is00) { h1 = pf_real(h1); h2 = pf_real(h2); h3 = pf_real(h3); }
      h1m = h1;
      h3m = h3;
      // nonlin (solvers.f90:1497-1572)
      PF x = pf_add(pf_imul(h3, kz), pf_imul(h1, kx));
      PF dx = cderiv<NY>(x, dy, lane);
      fn = pf_sub(pf_scale(h2, -k2), dx);
      gn = pf_sub(pf_imul(h1, kz), pf_imul(h3, kx));
      if (SCALAR) {
        sn = phys_to_cheb<NY, V>(qs, buf, scr, twl, lane);
        if (A.smooth) sn = smooth_apply<NY>(sn, A.smy, sz, lane);
        if (is00) sn = pf_real(sn);
      }
      if (is00) {
        // main_dns.f90:748-751 (h1n = Re gn(:,1,1), h3n = Re omz(:,1,1))
        if (MODE == 2) {
          pf_store_real<NY>(A.h1n, lane, h1m);
          pf_store_real<NY>(A.h3n, lane, h3m);
        }
      }
      if (MODE == 2) {
        pf_store_pm<NY>(A.FN, p, lane, fn);
        pf_store_pm<NY>(A.GN, p, lane, gn);
        if (SCALAR) pf_store_pm<NY>(A.SCN, p, lane, sn);
        return;
      }
    } else {
      fn = pf_load_pm<NY>(A.FN, p, lane);
      gn = pf_load_pm<NY>(A.GN, p, lane);
      if (SCALAR) sn = pf_load_pm<NY>(A.SCN, p, lane);
      if (is00) { h1m = pf_load_real<NY>(A.h1n, lane); h3m = pf_load_real<NY>(A.h3n, lane); }
    }
  }

  const double th = A.theta;
  // ---------------- scalar (main_dns.f90:324-358) ----------------------
  PF sc;
  if (SCALAR) {
    sc = pf_load_pm<NY>(A.TH, p, lane);
    if (MODE == 0) {
      PF scnm1 = pf_load_pm<NY>(A.SCNM1, p, lane);
      PF d1 = cderiv<NY>(sc, dy, lane);
      PF d2 = cderiv<NY>(d1, dy, lane);
      PF rhs = phirhs<NY>(sc, d2, sn, scnm1, k2, A.re, A.diff, A.dt, th);
      pf_store_pm<NY>(A.SCNM1, p, lane, sn);
      auto tt = [dy](int n) -> double { return dy * (double)n * (double)n; };
      auto tb = [dy](int n) -> double { double v = dy * (double)n * (double)n; return (n & 1) ? -v : v; };
      sc = tau_penta<NY>(rhs, k2 + A.gS, dy, A.tabS + (size_t)p * tab_width(NY), tt, tb, czero(), czero(), lane);
      if (is00) sc = pf_real(sc);
      pf_store_pm<NY>(A.TH, p, lane, sc);
    }
  }

  // ---------------- v and omega_y (main_dns.f90:360-502) ---------------
  PF v, omy;
  PF u0p, w0p;
  if (MODE == 0) {
    if (!is00) {
      v = pf_load_pm<NY>(A.V, p, lane);
      PF fnm1 = pf_load_pm<NY>(A.FNM1, p, lane);
      PF d1 = cderiv<NY>(v, dy, lane);
      PF d2 = cderiv<NY>(d1, dy, lane);
      PF phi = pf_sub(d2, pf_scale(v, k2));
      d1 = cderiv<NY>(phi, dy, lane);
      d2 = cderiv<NY>(d1, dy, lane);
      PF rhs = phirhs<NY>(phi, d2, fn, fnm1, k2, A.re, 1.0, A.dt, th);
      pf_store_pm<NY>(A.FNM1, p, lane, fn);
      phi = tau_solve<NY>(rhs, k2 + A.gH, dy, A.tabH + (size_t)p * tab_width(NY), czero(), czero(), 0, lane);
      PF vp = tau_solve<NY>(phi, k2, dy, A.tabP + (size_t)p * tab_width(NY), czero(), czero(), 0, lane);
      double2 vt, vb;
      wall_derivs<NY>(vp, dy, lane, vt, vb);
      const double4 g = A.gk[p];
      double2 c1 = cadd(cscale(vb, g.x), cscale(vt, g.y));
      double2 c2 = cadd(cscale(vb, g.z), cscale(vt, g.w));
      v = pf_add(vp, pf_add(pf_cmul_real(c1, green_load<NY>(A.G1, p, lane)),
                            pf_cmul_real(c2, green_load<NY>(A.G2, p, lane))));
      omy = pf_load_pm<NY>(A.OMY, p, lane);
      PF gnm1 = pf_load_pm<NY>(A.GNM1, p, lane);
      PF e1 = cderiv<NY>(omy, dy, lane);
      PF e2 = cderiv<NY>(e1, dy, lane);
      PF orhs = phirhs<NY>(omy, e2, gn, gnm1, k2, A.re, 1.0, A.dt, th);
      pf_store_pm<NY>(A.GNM1, p, lane, gn);
      const double atu = A.atu, btu = A.btu, abu = A.abu, bbu = A.bbu;
      auto tt = [=](int n) -> double { return atu + dy * btu * (double)n * (double)n; };
      auto tb = [=](int n) -> double { double t = abu + dy * bbu * (double)n * (double)n; return (n & 1) ? -t : t; };
      omy = tau_penta<NY>(orhs, k2 + A.gH, dy, A.tabH + (size_t)p * tab_width(NY), tt, tb, czero(), czero(), lane);
      pf_store_pm<NY>(A.V, p, lane, v);
      pf_store_pm<NY>(A.OMY, p, lane, omy);
    } else {
      // (0,0): v = 0 (veloc), omega_y = 0, fnm1 = fn = 0, gnm1 = gn = 0,
      // and the mean-flow solves (uzero, solvers.f90:1065-1176).
      v = pf_zero<NY>();
      omy = pf_zero<NY>();
      pf_store_pm<NY>(A.V, p, lane, v);
      pf_store_pm<NY>(A.OMY, p, lane, omy);
      pf_store_pm<NY>(A.FNM1, p, lane, pf_zero<NY>());
      pf_store_pm<NY>(A.GNM1, p, lane, pf_zero<NY>());
      for (int w = 0; w < 2; ++w) {
        double* um = w ? A.w0 : A.u0;
        double* hl = w ? A.h3nm1 : A.h1nm1;
        double* hn = w ? A.h3n : A.h1n;
        const PF& hcur = w ? h3m : h1m;
        PF u = pf_load_real<NY>(um, lane);
        PF hlp = pf_load_real<NY>(hl, lane);
        PF d1 = cderiv<NY>(u, dy, lane);
        PF d2 = cderiv<NY>(d1, dy, lane);
        // wrk1 = -re u0/(dt th) - re (3 h1 - h1l)/(2 th) - (1-th)/th u0''
        PF r = pf_scale(u, -A.re / (A.dt * th));
        r = pf_axpy(-3.0 * A.re / (2.0 * th), hcur, r);
        r = pf_axpy(A.re / (2.0 * th), hlp, r);
        r = pf_axpy(-(1.0 - th) / th, d2, r);
        // psolve1d with mixedbc1d rows (correct bottom sign); w0 uses btu
        // (quirk of solvers.f90:1173).
        const double at = w ? A.atw : A.atu, bt = A.btu, ab = w ? A.abw : A.abu, bb = w ? A.bbw : A.bbu;
        const double gt = w ? A.gtw : A.gtu, gb = w ? A.gbw : A.gbu;
        auto tt = [=](int n) -> double { return at + dy * bt * (double)n * (double)n; };
        auto tb = [=](int n) -> double { double t = ab - dy * bb * (double)n * (double)n; return (n & 1) ? -t : t; };
        PF un = tau_penta<NY>(r, A.gH, dy, A.tabH + (size_t)p * tab_width(NY), tt, tb, make_double2(gt, 0.0),
                              make_double2(gb, 0.0), lane);
        un = pf_real(un);
        if (w == 0 && A.cmf) {
          // constant mass flux: u0 += f phi, phi = response to a unit forcing
          // held over the step (rhs -re/theta, homogeneous wall data)
          PF unit = pf_zero<NY>();
          if (lane == 0) unit.a[0] = make_double2(-A.re / th, 0.0);
          PF phi = tau_penta<NY>(unit, A.gH, dy, A.tabH + (size_t)p * tab_width(NY), tt, tb, czero(), czero(), lane);
          phi = pf_real(phi);
          const double bu = pf_bulk<NY>(un, lane), bp = pf_bulk<NY>(phi, lane);
          const double fc = (A.ubulk - bu) / bp;
          un = pf_axpy(fc, phi, un);
          if (lane == 0) *A.dpdx = *A.dpdx - fc;
        }
        pf_store_real<NY>(um, lane, un);
        pf_store_real<NY>(hl, lane, hcur);
        pf_store_real<NY>(hn, lane, hcur);
        if (w) w0p = un; else u0p = un;
      }
    }
  } else {  // MODE 1 (startup): current state only
    v = pf_load_pm<NY>(A.V, p, lane);
    omy = pf_load_pm<NY>(A.OMY, p, lane);
    if (is00) {
      v = pf_zero<NY>(); omy = pf_zero<NY>();
      pf_store_pm<NY>(A.V, p, lane, v);
      pf_store_pm<NY>(A.OMY, p, lane, omy);
      u0p = pf_load_real<NY>(A.u0, lane);
      w0p = pf_load_real<NY>(A.w0, lane);
    }
  }

  // ---------------- veloc / vort / gradscl (solvers.f90:1576-1722) ------
  PF u, w;
  if (is00) {
    u = u0p;
    w = w0p;
  } else {
    PF dv = cderiv<NY>(v, dy, lane);
    const double ik2 = 1.0 / k2;
    // u = -(-i kx Dv + i kz omy)/k2 ; w = -(-i kz Dv - i kx omy)/k2
    u = pf_scale(pf_sub(pf_imul(dv, kx), pf_imul(omy, kz)), ik2);
    w = pf_scale(pf_add(pf_imul(dv, kz), pf_imul(omy, kx)), ik2);
  }
  const PF du = cderiv<NY>(u, dy, lane), dw = cderiv<NY>(w, dy, lane);
  PF ox = pf_sub(dw, pf_imul(v, kz));
  PF oz = pf_sub(pf_imul(v, kx), du);
  if (A.TPD[0]) {  // D u, D v, D w for the velocity-gradient tensor (gradv, derivatives.f90:6-86)
    cheb_to_phys_store<NY, V>(du, buf, scr, twl, lane, static_cast<V*>(A.TPD[0]), P, p);
    cheb_to_phys_store<NY, V>(cderiv<NY>(v, dy, lane), buf, scr, twl, lane, static_cast<V*>(A.TPD[1]), P, p);
    cheb_to_phys_store<NY, V>(dw, buf, scr, twl, lane, static_cast<V*>(A.TPD[2]), P, p);
  }
  cheb_to_phys_store<NY, V>(u, buf, scr, twl, lane, TT[0], P, p);
  cheb_to_phys_store<NY, V>(v, buf, scr, twl, lane, TT[1], P, p);
  cheb_to_phys_store<NY, V>(w, buf, scr, twl, lane, TT[2], P, p);
  cheb_to_phys_store<NY, V>(ox, buf, scr, twl, lane, TT[3], P, p);
  cheb_to_phys_store<NY, V>(omy, buf, scr, twl, lane, TT[4], P, p);
  cheb_to_phys_store<NY, V>(oz, buf, scr, twl, lane, TT[5], P, p);
  if (SCALAR) {
    PF tx = pf_imul(sc, kx), tz = pf_imul(sc, kz);
    PF ty = cderiv<NY>(sc, dy, lane);
    cheb_to_phys_store<NY, V>(tx, buf, scr, twl, lane, TT[6], P, p);
    cheb_to_phys_store<NY, V>(ty, buf, scr, twl, lane, TT[7], P, p);
    cheb_to_phys_store<NY, V>(tz, buf, scr, twl, lane, TT[8], P, p);
  }
}

// The full step (MODE 0) with the [y][pencil] fields of the block (the
// nonlinear terms in, the transfer fields out) staged through LDS tiles
// (coalesced rows of PWS pencils); the state, stored pencil-major, is read
// and written by each wave as contiguous memory.  Same operations in the
// same order per pencil as pencil_body, so the results are identical; every
// wave of the block, the z-Nyquist pencil's included, passes every block
// barrier.
template <int NY, bool SCALAR, typename V>
__device__ __forceinline__ void pencil_step_tiled(const PencilArgs& A, int blk) {
  using PF = PFN<NY>;
  using G = TileG<NY>;
  using LP = StepLds<NY, V, SCALAR>;
  constexpr int NTS = LP::NTS;
  extern __shared__ __align__(16) unsigned char smraw[];
  V* smem = reinterpret_cast<V*>(smraw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int p0 = blk * PWS, p = p0 + wv;
  const V* twl = lds_twiddles<NY, V>(smem, static_cast<const V*>(A.twM));
  V* buf = smem + 2 * NY + (size_t)wv * FFTBuf<2 * NY>::len;
  V* scr = buf;  // transforms in place
  unsigned char* region = smraw + LP::TWB;
  V* vtile = reinterpret_cast<V*>(region + LP::FFT);     // NL / transfer tiles after the FFT buffers
  V* const* TT = reinterpret_cast<V* const*>(A.T);
  const size_t P = (size_t)A.P;
  const bool live = (size_t)p < P;
  double kx, kz;
  int kxi, kzi;
  pencil_k(A, live ? p : 0, kx, kz, kxi, kzi);
  const double k2 = kx * kx + kz * kz;
  const bool is00 = live && kxi == 0 && kzi == 0;
  const bool nyq = kzi == A.nzh;
  const bool act = live && !nyq;  // computes; the z-Nyquist pencil is zeroed (helpers.f90:149-156)
  const double dy = A.dyde;
  const double th = A.theta;
  // ---------------- part B: nonlinear terms -> fn, gn, scn (spectral) --
  PF fn, gn, sn;
  PF h1m, h3m;  // mean-mode profiles (only used by the (0,0) pencil)
  if (A.first_nl) {
    V r2[2][G::PER];
    {
      V rn[2][G::PER];
      tile_fetch<NY, V>(static_cast<const V*>(A.NL1), P, p0, rn[0]);
      tile_fetch<NY, V>(static_cast<const V*>(A.NL3), P, p0, rn[1]);
      tile_fetch<NY, V>(static_cast<const V*>(A.NL2), P, p0, r2[0]);
      if (SCALAR) tile_fetch<NY, V>(static_cast<const V*>(A.NLS), P, p0, r2[1]);
      tile_put<NY, V>(vtile, rn[0]);
      tile_put<NY, V>(vtile + G::LEN, rn[1]);
    }
    __syncthreads();
    PF h1, h3, h2;
    if (act) {
      h1 = phys_to_cheb<NY, V>(tile_phys<NY, V>(vtile, wv, lane), buf, scr, twl, lane);
      h3 = phys_to_cheb<NY, V>(tile_phys<NY, V>(vtile + G::LEN, wv, lane), buf, scr, twl, lane);
    }
    __syncthreads();
    tile_put<NY, V>(vtile, r2[0]);
    if (SCALAR) tile_put<NY, V>(vtile + G::LEN, r2[1]);
    __syncthreads();
    if (act) {
      const double sz = A.smooth ? A.smz[kzi] : 1.0;
      h2 = phys_to_cheb<NY, V>(tile_phys<NY, V>(vtile, wv, lane), buf, scr, twl, lane);
      if (A.TF[0]) {  // polymer force, subforce (helpers.f90:564-606), before smoothing
        h1 = pf_add(h1, pf_load<NY>(A.TF[0], P, p, lane));
        h2 = pf_add(h2, pf_load<NY>(A.TF[1], P, p, lane));
        h3 = pf_add(h3, pf_load<NY>(A.TF[2], P, p, lane));
      }
      if (A.smooth) {
        h1 = smooth_apply<NY>(h1, A.smy, sz, lane);
        h2 = smooth_apply<NY>(h2, A.smy, sz, lane);
        h3 = smooth_apply<NY>(h3, A.smy, sz, lane);
      }
      if (is00) { h1 = pf_real(h1); h2 = pf_real(h2); h3 = pf_real(h3); }
      h1m = h1;
      h3m = h3;
      // nonlin (solvers.f90:1497-1572)
      PF x = pf_add(pf_imul(h3, kz), pf_imul(h1, kx));
      PF dx = cderiv<NY>(x, dy, lane);
      fn = pf_sub(pf_scale(h2, -k2), dx);
      gn = pf_sub(pf_imul(h1, kz), pf_imul(h3, kx));
      if (SCALAR) {
        sn = phys_to_cheb<NY, V>(tile_phys<NY, V>(vtile + G::LEN, wv, lane), buf, scr, twl, lane);
        if (A.smooth) sn = smooth_apply<NY>(sn, A.smy, sz, lane);
        if (is00) sn = pf_real(sn);
      }
    }
    __syncthreads();  // the NL tiles become transfer tiles
  } else {
    if (act) {
      fn = pf_load_pm<NY>(A.FN, p, lane);
      gn = pf_load_pm<NY>(A.GN, p, lane);
      if (SCALAR) sn = pf_load_pm<NY>(A.SCN, p, lane);
      if (is00) { h1m = pf_load_real<NY>(A.h1n, lane); h3m = pf_load_real<NY>(A.h3n, lane); }
    }
  }

  // ---------------- solves ----------------------------------------------
  // The state is stored pencil-major: each wave loads its pencils
  // (contiguous) where the solves use them, which keeps the kernel within
  // the registers of three waves per SIMD.
  PF sc, v, omy;
  PF u0p, w0p;
  if (act) {
    // scalar (main_dns.f90:324-358)
    if (SCALAR) {
      sc = pf_load_pm<NY>(A.TH, p, lane);
      const PF scnm1 = pf_load_pm<NY>(A.SCNM1, p, lane);
      PF d1 = cderiv<NY>(sc, dy, lane);
      PF d2 = cderiv<NY>(d1, dy, lane);
      PF rhs = phirhs<NY>(sc, d2, sn, scnm1, k2, A.re, A.diff, A.dt, th);
      pf_store_pm<NY>(A.SCNM1, p, lane, sn);
      auto tt = [dy](int n) -> double { return dy * (double)n * (double)n; };
      auto tb = [dy](int n) -> double { double v = dy * (double)n * (double)n; return (n & 1) ? -v : v; };
      sc = tau_penta<NY>(rhs, k2 + A.gS, dy, A.tabS + (size_t)p * tab_width(NY), tt, tb, czero(), czero(), lane);
      if (is00) sc = pf_real(sc);
      pf_store_pm<NY>(A.TH, p, lane, sc);
    }
    // v and omega_y (main_dns.f90:360-502)
    if (!is00) {
      v = pf_load_pm<NY>(A.V, p, lane);
      const PF fnm1 = pf_load_pm<NY>(A.FNM1, p, lane);
      PF d1 = cderiv<NY>(v, dy, lane);
      PF d2 = cderiv<NY>(d1, dy, lane);
      PF phi = pf_sub(d2, pf_scale(v, k2));
      d1 = cderiv<NY>(phi, dy, lane);
      d2 = cderiv<NY>(d1, dy, lane);
      PF rhs = phirhs<NY>(phi, d2, fn, fnm1, k2, A.re, 1.0, A.dt, th);
      pf_store_pm<NY>(A.FNM1, p, lane, fn);
      phi = tau_solve<NY>(rhs, k2 + A.gH, dy, A.tabH + (size_t)p * tab_width(NY), czero(), czero(), 0, lane);
      PF vp = tau_solve<NY>(phi, k2, dy, A.tabP + (size_t)p * tab_width(NY), czero(), czero(), 0, lane);
      double2 vt, vb;
      wall_derivs<NY>(vp, dy, lane, vt, vb);
      const double4 g = A.gk[p];
      double2 c1 = cadd(cscale(vb, g.x), cscale(vt, g.y));
      double2 c2 = cadd(cscale(vb, g.z), cscale(vt, g.w));
      v = pf_add(vp, pf_add(pf_cmul_real(c1, green_load<NY>(A.G1, p, lane)),
                            pf_cmul_real(c2, green_load<NY>(A.G2, p, lane))));
      omy = pf_load_pm<NY>(A.OMY, p, lane);
      const PF gnm1 = pf_load_pm<NY>(A.GNM1, p, lane);
      PF e1 = cderiv<NY>(omy, dy, lane);
      PF e2 = cderiv<NY>(e1, dy, lane);
      PF orhs = phirhs<NY>(omy, e2, gn, gnm1, k2, A.re, 1.0, A.dt, th);
      pf_store_pm<NY>(A.GNM1, p, lane, gn);
      const double atu = A.atu, btu = A.btu, abu = A.abu, bbu = A.bbu;
      auto tt = [=](int n) -> double { return atu + dy * btu * (double)n * (double)n; };
      auto tb = [=](int n) -> double { double t = abu + dy * bbu * (double)n * (double)n; return (n & 1) ? -t : t; };
      omy = tau_penta<NY>(orhs, k2 + A.gH, dy, A.tabH + (size_t)p * tab_width(NY), tt, tb, czero(), czero(), lane);
      pf_store_pm<NY>(A.V, p, lane, v);
      pf_store_pm<NY>(A.OMY, p, lane, omy);
    } else {
      // (0,0): v = 0 (veloc), omega_y = 0, fnm1 = fn = 0, gnm1 = gn = 0,
      // and the mean-flow solves (uzero, solvers.f90:1065-1176).
      v = pf_zero<NY>();
      omy = pf_zero<NY>();
      pf_store_pm<NY>(A.V, p, lane, v);
      pf_store_pm<NY>(A.OMY, p, lane, omy);
      pf_store_pm<NY>(A.FNM1, p, lane, pf_zero<NY>());
      pf_store_pm<NY>(A.GNM1, p, lane, pf_zero<NY>());
      for (int w = 0; w < 2; ++w) {
        double* um = w ? A.w0 : A.u0;
        double* hl = w ? A.h3nm1 : A.h1nm1;
        double* hn = w ? A.h3n : A.h1n;
        const PF& hcur = w ? h3m : h1m;
        PF u = pf_load_real<NY>(um, lane);
        PF hlp = pf_load_real<NY>(hl, lane);
        PF d1 = cderiv<NY>(u, dy, lane);
        PF d2 = cderiv<NY>(d1, dy, lane);
        PF r = pf_scale(u, -A.re / (A.dt * th));
        r = pf_axpy(-3.0 * A.re / (2.0 * th), hcur, r);
        r = pf_axpy(A.re / (2.0 * th), hlp, r);
        r = pf_axpy(-(1.0 - th) / th, d2, r);
        const double at = w ? A.atw : A.atu, bt = A.btu, ab = w ? A.abw : A.abu, bb = w ? A.bbw : A.bbu;
        const double gt = w ? A.gtw : A.gtu, gb = w ? A.gbw : A.gbu;
        auto tt = [=](int n) -> double { return at + dy * bt * (double)n * (double)n; };
        auto tb = [=](int n) -> double { double t = ab - dy * bb * (double)n * (double)n; return (n & 1) ? -t : t; };
        PF un = tau_penta<NY>(r, A.gH, dy, A.tabH + (size_t)p * tab_width(NY), tt, tb, make_double2(gt, 0.0),
                              make_double2(gb, 0.0), lane);
        un = pf_real(un);
        if (w == 0 && A.cmf) {
          PF unit = pf_zero<NY>();
          if (lane == 0) unit.a[0] = make_double2(-A.re / th, 0.0);
          PF phi = tau_penta<NY>(unit, A.gH, dy, A.tabH + (size_t)p * tab_width(NY), tt, tb, czero(), czero(), lane);
          phi = pf_real(phi);
          const double bu = pf_bulk<NY>(un, lane), bp = pf_bulk<NY>(phi, lane);
          const double fc = (A.ubulk - bu) / bp;
          un = pf_axpy(fc, phi, un);
          if (lane == 0) *A.dpdx = *A.dpdx - fc;
        }
        pf_store_real<NY>(um, lane, un);
        pf_store_real<NY>(hl, lane, hcur);
        pf_store_real<NY>(hn, lane, hcur);
        if (w) w0p = un; else u0p = un;
      }
    }
  } else if (live) {  // the z-Nyquist pencil: zero state
    const PF z = pf_zero<NY>();
    pf_store_pm<NY>(A.V, p, lane, z);
    pf_store_pm<NY>(A.OMY, p, lane, z);
    pf_store_pm<NY>(A.FNM1, p, lane, z);
    pf_store_pm<NY>(A.GNM1, p, lane, z);
    if (SCALAR) {
      pf_store_pm<NY>(A.TH, p, lane, z);
      pf_store_pm<NY>(A.SCNM1, p, lane, z);
    }
  }

  // ---------------- veloc / vort / gradscl (solvers.f90:1576-1722) ------
  PF u, w;
  if (act) {
    if (is00) {
      u = u0p;
      w = w0p;
    } else {
      PF dv = cderiv<NY>(v, dy, lane);
      const double ik2 = 1.0 / k2;
      u = pf_scale(pf_sub(pf_imul(dv, kx), pf_imul(omy, kz)), ik2);
      w = pf_scale(pf_add(pf_imul(dv, kz), pf_imul(omy, kx)), ik2);
    }
  }
  // inverse transforms into transfer tiles, NTS at a time; out(j) is the
  // j-th transfer field, dst(j) its destination
  auto flush = [&](int j0, int n, V* const* dst) {
    __syncthreads();
    for (int j = 0; j < n; ++j) tile_store<NY, V>(vtile + (size_t)j * G::LEN, dst[j0 + j], P, p0);
    __syncthreads();
  };
  if (A.TPD[0]) {  // D u, D v, D w for the velocity-gradient tensor (gradv, derivatives.f90:6-86)
    V* const D[3] = {static_cast<V*>(A.TPD[0]), static_cast<V*>(A.TPD[1]), static_cast<V*>(A.TPD[2])};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      V* t = vtile + (size_t)j * G::LEN;
      if (act) {
        const PF d = cderiv<NY>(j == 0 ? u : (j == 1 ? v : w), dy, lane);
        cheb_to_phys_tile<NY, V>(d, buf, scr, twl, lane, t, wv);
      } else if (live) {
        tile_col_zero<NY, V>(t, wv, lane);
      }
    }
    flush(0, 3, D);
  }
  // each transfer field is formed just before its transform, and
  // D u, D w right after u, w leave: at most five pencil fields stay live
  PF du, dw;
  constexpr int NOUT = LP::NOUT;
#pragma unroll
  for (int j = 0; j < NOUT; ++j) {
    V* t = vtile + (size_t)(j % NTS) * G::LEN;
    if (act) {
      PF o;
      if (j == 0) o = u;
      else if (j == 1) o = v;
      else if (j == 2) o = w;
      else if (j == 3) o = pf_sub(dw, pf_imul(v, kz));  // omega_x
      else if (j == 4) o = omy;
      else if (j == 5) o = pf_sub(pf_imul(v, kx), du);  // omega_z
      else if (j == 6) o = pf_imul(sc, kx);
      else if (j == 7) o = cderiv<NY>(sc, dy, lane);
      else o = pf_imul(sc, kz);
      cheb_to_phys_tile<NY, V>(o, buf, scr, twl, lane, t, wv);
      if (j == 0) du = cderiv<NY>(u, dy, lane);
      if (j == 2) dw = cderiv<NY>(w, dy, lane);
    } else if (live) {
      tile_col_zero<NY, V>(t, wv, lane);
    }
    if (j % NTS == NTS - 1 || j == NOUT - 1) flush(j - j % NTS, j % NTS + 1, TT);
  }
}

template <int NY, bool SCALAR, int MODE, typename V, int TV = 0>
__global__ __launch_bounds__(PW * 64) __attribute__((amdgpu_waves_per_eu(TV ? 3 : 1))) void k_pencil(PencilArgs A) {
  if constexpr (MODE == 0 && StepLds<NY, V, SCALAR>::ok)
    pencil_step_tiled<NY, SCALAR, V>(A, xcd_group(blockIdx.x, gridDim.x));
  else pencil_body<NY, SCALAR, MODE, V>(A, blockIdx.x);
}

// The step kernel of an ensemble of channels of one size in one launch:
// channel blockIdx.y takes its arguments from AA[blockIdx.y].
template <int NY, bool SCALAR, typename V, int TV>
__global__ __launch_bounds__(PW * 64) __attribute__((amdgpu_waves_per_eu(TV ? 3 : 1)))
void k_pencil_batch(const PencilArgs* __restrict__ AA) {
  if constexpr (StepLds<NY, V, SCALAR>::ok)
    pencil_step_tiled<NY, SCALAR, V>(AA[blockIdx.y], xcd_group(blockIdx.x, gridDim.x));
  else pencil_body<NY, SCALAR, 0, V>(AA[blockIdx.y], blockIdx.x);
}

// -------------------------------------------------- parity kernels --
template <int NY>
__global__ __launch_bounds__(PW * 64) void k_cderiv(int P, double dyde, const double2* __restrict__ in,
                                                    double2* __restrict__ out) {
  using PF = PFN<NY>;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int p = blockIdx.x * PW + wv;
  if (p >= P) return;
  PF f = pf_load<NY>(in, P, p, lane);
  pf_store<NY>(out, P, p, lane, cderiv<NY>(f, dyde, lane));
}

// gradv (derivatives.f90:6-86): out[3 i + k] = d u_i / d x_k for i, k in
// x, y, z, out[9 + i] = Laplacian of u_i (D^2 - k^2); wave per pencil
template <int NY>
__global__ __launch_bounds__(PW * 64) void k_gradv(int P, int nz, int nzh, double alpha, double beta, double dyde,
                                                   const double2* __restrict__ u, const double2* __restrict__ v,
                                                   const double2* __restrict__ w, double2* const* __restrict__ out) {
  using PF = PFN<NY>;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int p = blockIdx.x * PW + wv;
  if (p >= P) return;
  const int kxi = p / nz, kzi = p % nz;
  const double kx = kxi * alpha, kz = ((kzi < nzh) ? kzi : kzi - nz) * beta;
  const double k2 = kx * kx + kz * kz;
  const double2* in[3] = {u, v, w};
#pragma unroll 1
  for (int i = 0; i < 3; ++i) {
    const PF f = pf_load<NY>(in[i], P, p, lane);
    const PF d = cderiv<NY>(f, dyde, lane);
    pf_store<NY>(out[3 * i + 0], P, p, lane, pf_imul(f, kx));
    pf_store<NY>(out[3 * i + 1], P, p, lane, d);
    pf_store<NY>(out[3 * i + 2], P, p, lane, pf_imul(f, kz));
    pf_store<NY>(out[9 + i], P, p, lane, pf_axpy(-k2, f, cderiv<NY>(d, dyde, lane)));
  }
}

// Laplacians of u, v, w for inertial particles (gradv's D^2 f - k^2 f,
// derivatives.f90:77-79, normed as in main_dns.f90:556-558) from the state v,
// omega_y, u0, w0 through veloc (u = -(-i kx Dv + i kz omy)/k2,
// w = -(-i kz Dv - i kx omy)/k2; u0, 0, w0 at kx = kz = 0), inverse Chebyshev
// transformed to physical y; one wave per pencil
template <int NY, typename V>
__global__ __launch_bounds__(PW * 64) void k_mr_lap(MrLapArgs A) {
  using PF = PFN<NY>;
  extern __shared__ __align__(16) unsigned char smraw[];
  V* smem = reinterpret_cast<V*>(smraw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int p = blockIdx.x * PW + wv;
  const V* twl = lds_twiddles<NY, V>(smem, static_cast<const V*>(A.twM));
  if (p >= A.P) return;
  V* buf = wave_lds<NY, V>(smem, wv);
  V* scr = buf + FFTBuf<2 * NY>::len;
  const size_t P = (size_t)A.P;
  double kx, kz;
  int kxi, kzi;
  pencil_k(A, p, kx, kz, kxi, kzi);
  const double k2 = kx * kx + kz * kz;
  const double dy = A.dyde;
  const bool is00 = (kxi == 0 && kzi == 0);
  PF f[3] = {pf_zero<NY>(), pf_zero<NY>(), pf_zero<NY>()};
  if (kzi != A.nzh) {  // norm zeroes the z-Nyquist row
    if (is00) {
      f[0] = pf_load_real<NY>(A.u0, lane);
      f[2] = pf_load_real<NY>(A.w0, lane);
    } else {
      const PF v = pf_load_pm<NY>(A.V, p, lane);
      const PF omy = pf_load_pm<NY>(A.OMY, p, lane);
      const PF dv = cderiv<NY>(v, dy, lane);
      const double ik2 = 1.0 / k2;
      f[0] = pf_scale(pf_sub(pf_imul(dv, kx), pf_imul(omy, kz)), ik2);
      f[1] = v;
      f[2] = pf_scale(pf_add(pf_imul(dv, kz), pf_imul(omy, kx)), ik2);
    }
  }
#pragma unroll 1
  for (int i = 0; i < 3; ++i) {
    PF l = pf_axpy(-k2, f[i], cderiv<NY>(cderiv<NY>(f[i], dy, lane), dy, lane));
    if (is00) l = pf_real(l);
    cheb_to_phys_store<NY, V>(l, buf, scr, twl, lane, static_cast<V*>(A.LT[i]), P, p);
  }
}

// inverse: FP64 spectral in -> V physical-y out; forward: V in -> FP64 out
template <int NY, typename V>
__global__ __launch_bounds__(PW * 64) void k_cheb(int P, const V* __restrict__ tw, const void* in, void* out,
                                                  int inverse) {
  using PF = PFN<NY>;
  extern __shared__ __align__(16) unsigned char smraw[];
  V* smem = reinterpret_cast<V*>(smraw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int p = blockIdx.x * PW + wv;
  const V* twl = lds_twiddles<NY, V>(smem, tw);
  if (p >= P) return;
  V* buf = wave_lds<NY, V>(smem, wv);
  if (inverse) {  // 2: the spectral input is stored pencil-major
    PF f = inverse == 2 ? pf_load_pm<NY>(static_cast<const double2*>(in), p, lane)
                        : pf_load<NY>(static_cast<const double2*>(in), P, p, lane);
    cheb_to_phys_store<NY, V>(f, buf, buf + FFTBuf<2 * NY>::len, twl, lane, static_cast<V*>(out), P, p);
  } else {
    PF r = phys_to_cheb_load<NY, V>(static_cast<const V*>(in), P, p, buf, buf + FFTBuf<2 * NY>::len, twl, lane);
    pf_store<NY>(static_cast<double2*>(out), P, p, lane, r);
  }
}

template <int NY>
__global__ __launch_bounds__(PW * 64) void k_solve(int P, double dyde, const double2* __restrict__ tab,
                                                   const double* __restrict__ wn, double2* __restrict__ u,
                                                   const double2* __restrict__ bct, const double2* __restrict__ bcb,
                                                   int ib, const int* __restrict__ kzero) {
  using PF = PFN<NY>;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int p = blockIdx.x * PW + wv;
  if (p >= P) return;
  PF f = pf_load<NY>(u, P, p, lane);
  PF r = tau_solve<NY>(f, wn[p], dyde, tab + (size_t)p * tab_width(NY), bct[p], bcb[p], ib, lane);
  if (kzero && kzero[p]) r = pf_zero<NY>();
  pf_store<NY>(u, P, p, lane, r);
}

template <int NY>
__global__ __launch_bounds__(PW * 64) void k_penta(int P, double dyde, const double2* __restrict__ tab,
                                                   const double* __restrict__ wn, double2* __restrict__ u,
                                                   double at, double bt, const double2* __restrict__ gtop,
                                                   double ab, double bb, const double2* __restrict__ gbot,
                                                   double bsign, const int* __restrict__ kzero) {
  using PF = PFN<NY>;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int p = blockIdx.x * PW + wv;
  if (p >= P) return;
  PF f = pf_load<NY>(u, P, p, lane);
  auto tt = [=](int n) -> double { return at + dyde * bt * (double)n * (double)n; };
  auto tb = [=](int n) -> double { double t = ab + bsign * dyde * bb * (double)n * (double)n; return (n & 1) ? -t : t; };
  PF r = tau_penta<NY>(f, wn[p], dyde, tab + (size_t)p * tab_width(NY), tt, tb, gtop[p], gbot[p], lane);
  if (kzero && kzero[p]) r = pf_zero<NY>();
  pf_store<NY>(u, P, p, lane, r);
}

// ------------------------------------------------------------ launchers --
hipError_t launch_tau_tables(int ny, const PencilArgs& a, double g, double2* tab, hipStream_t s) {
  KMM_NY_DISPATCH(ny, hipLaunchKernelGGL((k_tau_tables<NY>), dim3((a.P + 127) / 128), dim3(128), 0, s, a.P, a.nz,
                                          a.alpha, a.beta, a.dyde, g, (const double*)nullptr, tab));
  return hipGetLastError();
}
hipError_t launch_tau_tables_wn(int ny, int P, double dyde, const double* wn, double2* tab, hipStream_t s) {
  KMM_NY_DISPATCH(ny, hipLaunchKernelGGL((k_tau_tables<NY>), dim3((P + 127) / 128), dim3(128), 0, s, P, 1, 0.0,
                                          0.0, dyde, 0.0, wn, tab));
  return hipGetLastError();
}
hipError_t launch_green(int ny, const PencilArgs& a, double pb1, double pt1, double pb2, double pt2, double2* G1,
                        double2* G2, double4* gk, hipStream_t s) {
  KMM_NY_DISPATCH(ny, hipLaunchKernelGGL((k_green<NY>), dim3(nblk(a.P)), dim3(PW * 64), 0, s, a, pb1, pt1, pb2,
                                          pt2, G1, G2, gk));
  return hipGetLastError();
}
template <int MODE, typename V, int TV>
static hipError_t launch_pencil_tv(int ny, const PencilArgs& a, hipStream_t s) {
  if (a.scalar) {
    KMM_NY_DISPATCH(ny, hipLaunchKernelGGL((k_pencil<NY, true, MODE, V, TV>),
                                           dim3(step_blocks<NY, V, true, MODE>(a.P)),
                                           dim3(step_threads<NY, V, true, MODE>()),
                                           (step_lds_bytes<NY, V, true, MODE>()), s, a));
  } else {
    KMM_NY_DISPATCH(ny, hipLaunchKernelGGL((k_pencil<NY, false, MODE, V, TV>),
                                           dim3(step_blocks<NY, V, false, MODE>(a.P)),
                                           dim3(step_threads<NY, V, false, MODE>()),
                                           (step_lds_bytes<NY, V, false, MODE>()), s, a));
  }
  return hipGetLastError();
}
template <int MODE, typename V>
static hipError_t launch_pencil_mode_v(int ny, const PencilArgs& a, hipStream_t s) {
  // only the step kernel has tuning variants
  if (MODE == 0 && a.tune == 1) return launch_pencil_tv<MODE, V, 1>(ny, a, s);
  return launch_pencil_tv<MODE, V, 0>(ny, a, s);
}
template <int MODE>
static hipError_t launch_pencil_mode(int ny, const PencilArgs& a, hipStream_t s) {
  return a.mixed ? launch_pencil_mode_v<MODE, float2>(ny, a, s) : launch_pencil_mode_v<MODE, double2>(ny, a, s);
}
hipError_t launch_pencil_step(int ny, const PencilArgs& a, hipStream_t s) { return launch_pencil_mode<0>(ny, a, s); }

template <typename V, int TV>
static hipError_t launch_pencil_batch_v(int ny, const PencilArgs& a0, const PencilArgs* dA, int nmem,
                                        hipStream_t s) {
  if (a0.scalar) {
    KMM_NY_DISPATCH(ny, hipLaunchKernelGGL((k_pencil_batch<NY, true, V, TV>),
                                           dim3(step_blocks<NY, V, true, 0>(a0.P), nmem),
                                           dim3(step_threads<NY, V, true, 0>()),
                                           (step_lds_bytes<NY, V, true, 0>()), s, dA));
  } else {
    KMM_NY_DISPATCH(ny, hipLaunchKernelGGL((k_pencil_batch<NY, false, V, TV>),
                                           dim3(step_blocks<NY, V, false, 0>(a0.P), nmem),
                                           dim3(step_threads<NY, V, false, 0>()),
                                           (step_lds_bytes<NY, V, false, 0>()), s, dA));
  }
  return hipGetLastError();
}
hipError_t launch_pencil_step_batch(int ny, const PencilArgs& a0, const PencilArgs* dA, int nmem, hipStream_t s) {
  if (nmem < 1) return hipErrorInvalidValue;
  if (a0.mixed)
    return a0.tune == 1 ? launch_pencil_batch_v<float2, 1>(ny, a0, dA, nmem, s)
                        : launch_pencil_batch_v<float2, 0>(ny, a0, dA, nmem, s);
  return a0.tune == 1 ? launch_pencil_batch_v<double2, 1>(ny, a0, dA, nmem, s)
                      : launch_pencil_batch_v<double2, 0>(ny, a0, dA, nmem, s);
}
hipError_t launch_pencil_startup(int ny, const PencilArgs& a, hipStream_t s) { return launch_pencil_mode<1>(ny, a, s); }
hipError_t launch_pencil_finish(int ny, const PencilArgs& a, hipStream_t s) { return launch_pencil_mode<2>(ny, a, s); }

hipError_t launch_gradv(int ny, int P, int nz, double alpha, double beta, double dyde, const double2* u,
                        const double2* v, const double2* w, double2* const* out, hipStream_t s) {
  KMM_NY_DISPATCH(ny, hipLaunchKernelGGL((k_gradv<NY>), dim3(nblk(P)), dim3(PW * 64), 0, s, P, nz, nz / 2, alpha,
                                         beta, dyde, u, v, w, out));
  return hipGetLastError();
}
hipError_t launch_mr_lap_pencil(int ny, const MrLapArgs& a, hipStream_t s) {
  if (a.mixed) {
    KMM_NY_DISPATCH(ny, hipLaunchKernelGGL((k_mr_lap<NY, float2>), dim3(nblk(a.P)), dim3(PW * 64),
                                           (lds_bytes<NY, float2>()), s, a));
  } else {
    KMM_NY_DISPATCH(ny, hipLaunchKernelGGL((k_mr_lap<NY, double2>), dim3(nblk(a.P)), dim3(PW * 64),
                                           (lds_bytes<NY, double2>()), s, a));
  }
  return hipGetLastError();
}
hipError_t launch_cderiv(int ny, int P, double dyde, const double2* in, double2* out, hipStream_t s) {
  KMM_NY_DISPATCH(ny, hipLaunchKernelGGL((k_cderiv<NY>), dim3(nblk(P)), dim3(PW * 64), 0, s, P, dyde, in, out));
  return hipGetLastError();
}
hipError_t launch_cheb(int ny, int P, const void* tw, const void* in, void* out, int inverse, int mixed,
                       hipStream_t s) {
  if (mixed) {
    KMM_NY_DISPATCH(ny, hipLaunchKernelGGL((k_cheb<NY, float2>), dim3(nblk(P)), dim3(PW * 64),
                                           (lds_bytes<NY, float2>()), s, P, static_cast<const float2*>(tw), in, out,
                                           inverse));
  } else {
    KMM_NY_DISPATCH(ny, hipLaunchKernelGGL((k_cheb<NY, double2>), dim3(nblk(P)), dim3(PW * 64),
                                           (lds_bytes<NY, double2>()), s, P, static_cast<const double2*>(tw), in,
                                           out, inverse));
  }
  return hipGetLastError();
}
hipError_t launch_solve(int ny, const PencilArgs& a, const double2* tab, const double* wn, double2* u,
                        const double2* bctop, const double2* bcbot, int ib, const int* kzero, hipStream_t s) {
  KMM_NY_DISPATCH(ny, hipLaunchKernelGGL((k_solve<NY>), dim3(nblk(a.P)), dim3(PW * 64), 0, s, a.P, a.dyde, tab,
                                          wn, u, bctop, bcbot, ib, kzero));
  return hipGetLastError();
}
hipError_t launch_penta(int ny, const PencilArgs& a, const double2* tab, const double* wn, double2* u, double at,
                        double bt, const double2* gtop, double ab, double bb, const double2* gbot, int bottom_sign,
                        const int* kzero, hipStream_t s) {
  KMM_NY_DISPATCH(ny, hipLaunchKernelGGL((k_penta<NY>), dim3(nblk(a.P)), dim3(PW * 64), 0, s, a.P, a.dyde, tab,
                                          wn, u, at, bt, gtop, ab, bb, gbot, (double)bottom_sign, kzero));
  return hipGetLastError();
}

}  // namespace kmm
