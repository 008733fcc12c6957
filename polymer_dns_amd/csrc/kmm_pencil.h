// Pencil-stage (wall-normal, Chebyshev) device routines.
//
// One wavefront owns one (kx,kz) pencil.  Lane l holds the Chebyshev
// coefficients (or Gauss-Lobatto values) n = 2l and n = 2l+1 for l < NY/2,
// and every lane carries a copy of n = NY.  With this layout the even and
// odd tau systems of FFTW_DNS/code/src/dns/solvers.f90 (evnc/oddc, rows
// coupling n-2, n, n+2) become nearest-neighbour lane couplings, and the
// sequential recurrences of the reference (cderiv, eqsolve, pntsolve)
// become wave64 affine scans.
#pragma once
#include "kmm_device.h"

namespace kmm {

struct PF {      // one complex pencil in lane-pair layout
  double2 a;     // n = 2*lane
  double2 b;     // n = 2*lane + 1
  double2 e;     // n = NY (uniform across lanes)
};

__device__ __forceinline__ PF pf_zero() { PF f; f.a = f.b = f.e = czero(); return f; }
__device__ __forceinline__ PF pf_axpy(double s, const PF& x, const PF& y) {
  PF r; r.a = cfma(s, x.a, y.a); r.b = cfma(s, x.b, y.b); r.e = cfma(s, x.e, y.e); return r;
}
__device__ __forceinline__ PF pf_scale(const PF& x, double s) {
  PF r; r.a = cscale(x.a, s); r.b = cscale(x.b, s); r.e = cscale(x.e, s); return r;
}
__device__ __forceinline__ PF pf_add(const PF& x, const PF& y) {
  PF r; r.a = cadd(x.a, y.a); r.b = cadd(x.b, y.b); r.e = cadd(x.e, y.e); return r;
}
__device__ __forceinline__ PF pf_sub(const PF& x, const PF& y) {
  PF r; r.a = csub(x.a, y.a); r.b = csub(x.b, y.b); r.e = csub(x.e, y.e); return r;
}
// i*s*x
__device__ __forceinline__ PF pf_imul(const PF& x, double s) {
  PF r; r.a = cimul(x.a, s); r.b = cimul(x.b, s); r.e = cimul(x.e, s); return r;
}
__device__ __forceinline__ PF pf_real(const PF& x) {
  PF r = x; r.a.y = 0.0; r.b.y = 0.0; r.e.y = 0.0; return r;
}

// field F laid out [y][pencil], P pencils per y-plane
template <int NY>
__device__ __forceinline__ PF pf_load(const double2* __restrict__ F, size_t P, size_t p, int lane) {
  constexpr int L = NY / 2;
  PF f;
  if (lane < L) {
    f.a = F[(size_t)(2 * lane) * P + p];
    f.b = F[(size_t)(2 * lane + 1) * P + p];
  } else {
    f.a = f.b = czero();
  }
  f.e = F[(size_t)NY * P + p];
  return f;
}
template <int NY>
__device__ __forceinline__ void pf_store(double2* __restrict__ F, size_t P, size_t p, int lane, const PF& f) {
  constexpr int L = NY / 2;
  if (lane < L) {
    F[(size_t)(2 * lane) * P + p] = f.a;
    F[(size_t)(2 * lane + 1) * P + p] = f.b;
  }
  if (lane == 0) F[(size_t)NY * P + p] = f.e;
}
// real mean profile (nyp doubles)
template <int NY>
__device__ __forceinline__ PF pf_load_real(const double* __restrict__ u, int lane) {
  constexpr int L = NY / 2;
  PF f = pf_zero();
  if (lane < L) { f.a.x = u[2 * lane]; f.b.x = u[2 * lane + 1]; }
  f.e.x = u[NY];
  return f;
}
template <int NY>
__device__ __forceinline__ void pf_store_real(double* __restrict__ u, int lane, const PF& f) {
  constexpr int L = NY / 2;
  if (lane < L) { u[2 * lane] = f.a.x; u[2 * lane + 1] = f.b.x; }
  if (lane == 0) u[NY] = f.e.x;
}

// ------------------------------------------------ Chebyshev derivative --
// derivatives.f90:258-314: df_n = sum_{k>n, k-n odd} 2 k f_k dyde, df_0/2.
template <int NY>
__device__ __forceinline__ PF cderiv(const PF& f, double dyde, int lane) {
  double2 go = cscale(f.b, 2.0 * (2 * lane + 1) * dyde);  // odd k = 2l+1
  double2 ge = cscale(f.a, 2.0 * (2 * lane) * dyde);      // even k = 2l
  PF d;
  d.a = suffix_sum(go, lane);
  double2 se = suffix_sum(ge, lane);
  double2 sx = shdn1(se);  // 0 at lane 63
  d.b = cfma(2.0 * NY * dyde, f.e, sx);
  d.e = czero();
  if (lane == 0) d.a = cscale(d.a, 0.5);
  if (lane >= NY / 2) { d.a = czero(); d.b = czero(); }
  return d;
}

// derivatives.f90:541-595 (y=+1) and 370-427 (y=-1)
template <int NY>
__device__ __forceinline__ void wall_derivs(const PF& f, double dyde, int lane, double2& top, double2& bot) {
  const double n0 = 2.0 * lane, n1 = 2.0 * lane + 1;
  double2 t = cfma(n1 * n1, f.b, cscale(f.a, n0 * n0));
  double2 b = cfma(n1 * n1, f.b, cscale(f.a, -n0 * n0));
  if (lane >= NY / 2) { t = czero(); b = czero(); }
  t = wave_sum(t);
  b = wave_sum(b);
  const double nn = (double)NY * NY;
  top = cscale(cfma(nn, f.e, t), dyde);
  bot = cscale(cfma(-nn, f.e, b), dyde);  // NY even: (-1)^(NY+1) = -1
}

// ----------------------------------------------- Chebyshev transforms --
// FFT plan cost model of the pencils' length-2NY transforms (make_plan): 1,
// fewest vector instructions per lane, since one transform per wave leaves
// every stage at a single round
constexpr int kChebPlan = 1;

// Inverse (solvers.f90:2241-2546): values at y_k = cos(pi k/NY) of
// sum_n a_n T_n via a length-2NY FFT of the even extension of c_n/2 * a_n,
// computed in the transfer precision V and stored to dst[y][pencil].
template <int NY, typename V>
__device__ void cheb_to_phys_store(const PF& f, V* buf, V* scr, const V* tw, int lane, V* __restrict__ dst,
                                   size_t P, size_t p) {
  constexpr int L = NY / 2, M = 2 * NY;
  if (lane < L) {
    V a = to_v<V>((lane == 0) ? f.a : cscale(f.a, 0.5));
    V b = to_v<V>(cscale(f.b, 0.5));
    buf[fpad(2 * lane)] = a;
    buf[fpad(2 * lane + 1)] = b;
    if (lane > 0) buf[fpad(M - 2 * lane)] = a;
    buf[fpad(M - 2 * lane - 1)] = b;
  }
  if (lane == 0) buf[fpad(NY)] = to_v<V>(f.e);
  wave_sync();
  V* r = wave_fft<M, -1, V, kChebPlan>(buf, scr, tw, lane);
  if (lane < L) {
    dst[(size_t)(2 * lane) * P + p] = r[fpad(2 * lane)];
    dst[(size_t)(2 * lane + 1) * P + p] = r[fpad(2 * lane + 1)];
  }
  if (lane == 0) dst[(size_t)NY * P + p] = r[fpad(NY)];
  wave_sync();
}
// A pencil's physical-y values held in registers (lane l: y = 2l, 2l+1; lane
// 0 also y = NY), so that a load can be issued ahead of the transform that
// consumes it.
template <typename V>
struct PhysPencil {
  V a, b, e;
};
template <int NY, typename V>
__device__ __forceinline__ PhysPencil<V> phys_load(const V* __restrict__ src, size_t P, size_t p, int lane) {
  constexpr int L = NY / 2;
  PhysPencil<V> q;
  const int l = lane < L ? lane : L - 1;  // lanes >= L reload a valid row and ignore it
  q.a = src[(size_t)(2 * l) * P + p];
  q.b = src[(size_t)(2 * l + 1) * P + p];
  q.e = src[(size_t)NY * P + p];
  return q;
}
// Forward (solvers.f90:3279-3416): DCT-I through the length-2NY FFT of the
// even extension, /NY, ends halved; input in the transfer precision V,
// result in FP64.
template <int NY, typename V>
__device__ PF phys_to_cheb(const PhysPencil<V>& q, V* buf, V* scr, const V* tw, int lane) {
  constexpr int L = NY / 2, M = 2 * NY;
  if (lane < L) {
    buf[fpad(2 * lane)] = q.a;
    buf[fpad(2 * lane + 1)] = q.b;
    if (lane > 0) buf[fpad(M - 2 * lane)] = q.a;
    buf[fpad(M - 2 * lane - 1)] = q.b;
  }
  if (lane == 0) buf[fpad(NY)] = q.e;
  wave_sync();
  V* r = wave_fft<M, -1, V, kChebPlan>(buf, scr, tw, lane);
  const double s = 1.0 / NY;
  PF o;
  if (lane < L) {
    o.a = cscale(to_d(r[fpad(2 * lane)]), lane == 0 ? 0.5 * s : s);
    o.b = cscale(to_d(r[fpad(2 * lane + 1)]), s);
  } else { o.a = o.b = czero(); }
  o.e = cscale(to_d(r[fpad(NY)]), 0.5 * s);
  wave_sync();
  return o;
}

// the same, loading the pencil first
template <int NY, typename V>
__device__ PF phys_to_cheb_load(const V* __restrict__ src, size_t P, size_t p, V* buf, V* scr, const V* tw,
                                int lane) {
  constexpr int L = NY / 2, M = 2 * NY;
  if (lane < L) {
    const V a = src[(size_t)(2 * lane) * P + p], b = src[(size_t)(2 * lane + 1) * P + p];
    buf[fpad(2 * lane)] = a;
    buf[fpad(2 * lane + 1)] = b;
    if (lane > 0) buf[fpad(M - 2 * lane)] = a;
    buf[fpad(M - 2 * lane - 1)] = b;
  }
  if (lane == 0) buf[fpad(NY)] = src[(size_t)NY * P + p];
  wave_sync();
  V* r = wave_fft<M, -1, V, kChebPlan>(buf, scr, tw, lane);
  const double s = 1.0 / NY;
  PF o;
  if (lane < L) {
    o.a = cscale(to_d(r[fpad(2 * lane)]), lane == 0 ? 0.5 * s : s);
    o.b = cscale(to_d(r[fpad(2 * lane + 1)]), s);
  } else { o.a = o.b = czero(); }
  o.e = cscale(to_d(r[fpad(NY)]), 0.5 * s);
  wave_sync();
  return o;
}

// -------------------------------------------------------- tau systems --
// Reciprocals of the tau row denominators for lane i (rows of lane 0 repeat
// lane 1's; lane 0 carries no row): even rows j = 2i, odd rows j = 2i + 1,
// 1/(4 j (j-1)), 1/(2 (j^2-1)), 1/(4 j (j+1)).  Folded at compile time with
// IEEE division, so the values equal the run-time quotients bit for bit.
struct TauRcp {
  double r[6][64];
  constexpr TauRcp() : r{} {
    for (int i = 0; i < 64; ++i) {
      const double je = 2.0 * (i < 1 ? 1 : i), jo = je + 1.0;
      r[0][i] = 1.0 / (4.0 * je * (je - 1.0));
      r[1][i] = 1.0 / (2.0 * (je * je - 1.0));
      r[2][i] = 1.0 / (4.0 * je * (je + 1.0));
      r[3][i] = 1.0 / (4.0 * jo * (jo - 1.0));
      r[4][i] = 1.0 / (2.0 * (jo * jo - 1.0));
      r[5][i] = 1.0 / (4.0 * jo * (jo + 1.0));
    }
  }
};
__device__ constexpr TauRcp kTauRcp{};

// Row coefficients a(i,1..3) of evnc/oddc (solvers.f90:324-383) for the row
// owned by lane i (1 <= i); wn = kx^2 + kz^2 + g.
template <int NY>
__device__ __forceinline__ void coef_even(int i, double wn, double dyde, double& a1, double& a2, double& a3) {
  constexpr int L = NY / 2;
  const double j = 2.0 * i;
  a1 = wn / (4.0 * j * (j - 1.0));
  if (i == 1) a1 *= 2.0;
  a2 = (i == L) ? -dyde * dyde : -dyde * dyde - wn / (2.0 * (j * j - 1.0));
  a3 = (i <= L - 2) ? wn / (4.0 * j * (j + 1.0)) : 0.0;
}
template <int NY>
__device__ __forceinline__ void coef_odd(int i, double wn, double dyde, double& a1, double& a2, double& a3) {
  constexpr int L = NY / 2;
  const double j = 2.0 * i + 1.0;
  a1 = wn / (4.0 * j * (j - 1.0));
  a2 = (i == L - 1) ? -dyde * dyde : -dyde * dyde - wn / (2.0 * (j * j - 1.0));
  a3 = (i <= L - 3) ? wn / (4.0 * j * (j + 1.0)) : 0.0;
}

// Tau RHS rows (evns/odds, solvers.f90:387-496) from Chebyshev data s.
// Even row at lane i (1..L-1) and extra row (index L); odd row at lane
// i (1..L-1).  Lane 0 is left for the boundary row.
template <int NY>
__device__ __forceinline__ void tau_rhs(const PF& s, int lane, double2& we, double2& wo, double2& wex) {
  constexpr int L = NY / 2;
  const int i = lane;
  double2 sme = shup1(s.a), spe = shdn1(s.a);
  double2 smo = shup1(s.b), spo = shdn1(s.b);
  we = czero();
  wo = czero();
  if (i >= 1 && i <= L - 1) {
    const double r1 = kTauRcp.r[0][i], r2 = kTauRcp.r[1][i], r3 = kTauRcp.r[2][i];
    if (i <= L - 2) {
      const double cm = (i == 1) ? 2.0 : 1.0;
      we = cadd(cadd(cscale(sme, -cm * r1), cscale(s.a, r2)), cscale(spe, -r3));
    } else {
      we = cadd(cscale(sme, -r1), cscale(s.a, r2));
    }
    const double q1 = kTauRcp.r[3][i], q2 = kTauRcp.r[4][i], q3 = kTauRcp.r[5][i];
    if (i <= L - 3) {
      wo = cadd(cadd(cscale(smo, -q1), cscale(s.b, q2)), cscale(spo, -q3));
    } else if (i == L - 2) {
      wo = cadd(cscale(smo, -q1), cscale(s.b, q2));
    } else {
      wo = cscale(smo, -q1);
    }
  }
  const double jx = (double)NY;
  wex = cscale(bcast(s.a, L - 1), -1.0 / (4.0 * jx * (jx - 1.0)));
}

// One parity of the elimination of eqsolve/pntsolve (solvers.f90:500-553,
// 906-1017).  Lane i (1..L-1) holds row i+1: RHS w, a1(i), a3(i), 1/d_i.
// Even parity also has the extra row (index L, unknown L+1).
struct TauPar {
  double2 wp;    // w' (eliminated RHS) at lane i, row i+1
  double a1, invd;
  double2 wpx;   // even extra w'
  double a1x, invdx;
  bool even;
};

template <int NY>
__device__ __forceinline__ TauPar tau_eliminate(bool even, double wn, double dyde, double invd, double invdx,
                                                double2 w, double2 wx, int lane) {
  constexpr int L = NY / 2;
  TauPar t;
  t.even = even;
  // coef_even / coef_odd of the lane's row with the tabulated reciprocals
  // (wn r in place of wn / (1/r): within one rounding of the quotient)
  const int li = lane < 1 ? 1 : lane;
  double a1, a3;
  if (even) {
    a1 = wn * kTauRcp.r[0][li];
    if (li == 1) a1 *= 2.0;
    a3 = (li <= L - 2) ? wn * kTauRcp.r[2][li] : 0.0;
  } else {
    a1 = wn * kTauRcp.r[3][li];
    a3 = (li <= L - 3) ? wn * kTauRcp.r[5][li] : 0.0;
  }
  const bool row = (lane >= 1 && lane <= L - 1);
  double invd_next = shdn1(invd);
  double rho = row ? a3 * invd_next : 0.0;
  double2 wv = row ? w : czero();
  t.wp = suffix_affine(-rho, wv, lane);
  if (!row) t.wp = czero();
  t.a1 = row ? a1 : 0.0;
  t.invd = row ? invd : 0.0;
  if (even) {
    constexpr double jx = 2.0 * L;
    constexpr double rx1 = 1.0 / (4.0 * jx * (jx - 1.0));
    t.a1x = wn * rx1;  // coef_even's a1 of the extra row
    t.invdx = invdx;
    t.wpx = wx;
  } else {
    t.a1x = 0.0; t.invdx = 0.0; t.wpx = czero();
  }
  return t;
}

// Boundary-row reduction of one parity for BC weights T(q) = tq(n_q).
// Returns sum_i F_{i+1} w'_{i+1} / d_i and sets F1.
template <int NY, typename TW>
__device__ __forceinline__ double2 tau_bcsum(const TauPar& t, TW tq, int lane, double& F1) {
  constexpr int L = NY / 2;
  // Chebyshev index of unknown q: even 2(q-1), odd 2q-1
  auto T = [&](int q) -> double { return t.even ? tq(2 * (q - 1)) : tq(2 * q - 1); };
  double carry;
  if (t.even) carry = T(L) - T(L + 1) * t.a1x * t.invdx;   // F_L
  else carry = T(L);                                          // F_{m+1}, m = L-1
  const bool row = (lane >= 1 && lane <= L - 1);
  double k = -t.a1 * t.invd;
  double c = row ? k : 0.0;
  double wv = row ? T(lane) : 0.0;
  if (lane == L - 1) { wv += k * carry; c = 0.0; }
  double F = suffix_affine_r(c, wv, lane);
  double Fn = shdn1(F);
  if (lane == L - 1) Fn = carry;
  double2 term = row ? cscale(t.wp, Fn * t.invd) : czero();
  double2 s = wave_sum(term);
  if (t.even) s = cfma(T(L + 1) * t.invdx, t.wpx, s);
  F1 = bcast(F, 1);
  return s;
}

// Forward substitution x_{i+1} = (w'_{i+1} - a1_i x_i)/d_i from x_1.
template <int NY>
__device__ __forceinline__ void tau_forward(const TauPar& t, double2 x1, int lane, double2& x, double2& xx) {
  constexpr int L = NY / 2;
  const bool row = (lane >= 1 && lane <= L - 1);
  double b = row ? -t.a1 * t.invd : 0.0;
  double2 a = row ? cscale(t.wp, t.invd) : (lane == 0 ? x1 : czero());
  x = prefix_affine(b, a, lane);
  if (lane >= L) x = czero();
  if (t.even) {
    double2 xl = bcast(x, L - 1);
    xx = cscale(cfma(-t.a1x, xl, t.wpx), t.invdx);
  } else {
    xx = czero();
  }
}

// solve (solvers.f90:174-283): (dyde^2 D^2 - wn) u = f, Dirichlet rows
// (ib=0) or Neumann rows (ib!=0, wn already guarded by the caller).
// tab: per-pencil 1/d table [65] (x: even, y: odd; [64].x even extra).
template <int NY>
__device__ PF tau_solve(const PF& f, double wn, double dyde, const double2* __restrict__ tab,
                        double2 bctop, double2 bcbot, int ib, int lane) {
  constexpr int L = NY / 2;
  double2 we, wo, wex;
  tau_rhs<NY>(f, lane, we, wo, wex);
  const double2 te = (lane < 64) ? tab[lane] : make_double2(0.0, 0.0);
  const double invdx = tab[64].x;
  TauPar E = tau_eliminate<NY>(true, wn, dyde, te.x, invdx, we, wex, lane);
  TauPar O = tau_eliminate<NY>(false, wn, dyde, te.y, 0.0, wo, czero(), lane);
  double2 bce, bco;
  if (ib == 0) { bce = cscale(cadd(bctop, bcbot), 0.5); bco = cscale(csub(bctop, bcbot), 0.5); }
  else { bce = cscale(csub(bctop, bcbot), 0.5); bco = cscale(cadd(bctop, bcbot), 0.5); }
  double F1e, F1o;
  double2 se, so;
  if (ib == 0) {
    auto one = [](int) -> double { return 1.0; };
    se = tau_bcsum<NY>(E, one, lane, F1e);
    so = tau_bcsum<NY>(O, one, lane, F1o);
  } else {
    auto nn = [dyde](int n) -> double { return dyde * (double)n * (double)n; };
    se = tau_bcsum<NY>(E, nn, lane, F1e);
    so = tau_bcsum<NY>(O, nn, lane, F1o);
  }
  double2 x1e = cscale(csub(bce, se), 1.0 / F1e);
  double2 x1o = cscale(csub(bco, so), 1.0 / F1o);
  PF u;
  double2 xx, dummy;
  tau_forward<NY>(E, x1e, lane, u.a, xx);
  tau_forward<NY>(O, x1o, lane, u.b, dummy);
  u.e = xx;
  (void)L;
  return u;
}

// penta (solvers.f90:680-1061) with Robin rows sum tt(n) u_n = gtop,
// sum tb(n) u_n = gbot; tt/tb are closed-form functions of n.
template <int NY, typename TT, typename TB>
__device__ PF tau_penta(const PF& f, double wn, double dyde, const double2* __restrict__ tab,
                        TT tt, TB tb, double2 gtop, double2 gbot, int lane) {
  double2 we, wo, wex;
  tau_rhs<NY>(f, lane, we, wo, wex);
  const double2 te = tab[lane];
  const double invdx = tab[64].x;
  TauPar E = tau_eliminate<NY>(true, wn, dyde, te.x, invdx, we, wex, lane);
  TauPar O = tau_eliminate<NY>(false, wn, dyde, te.y, 0.0, wo, czero(), lane);
  double f2e, f1e, f2o, f1o;
  double2 s2e = tau_bcsum<NY>(E, tb, lane, f2e);   // js = 2 (bottom row) first
  double2 s1e = tau_bcsum<NY>(E, tt, lane, f1e);
  double2 s2o = tau_bcsum<NY>(O, tb, lane, f2o);
  double2 s1o = tau_bcsum<NY>(O, tt, lane, f1o);
  double2 s1 = csub(csub(gtop, s1e), s1o);
  double2 s2 = csub(csub(gbot, s2e), s2o);
  const double det = f1e * f2o - f2e * f1o;
  double2 u0 = cscale(csub(cscale(s1, f2o), cscale(s2, f1o)), 1.0 / det);
  double2 u1 = cscale(csub(cscale(s2, f1e), cscale(s1, f2e)), 1.0 / det);
  PF u;
  double2 xx, dummy;
  tau_forward<NY>(E, u0, lane, u.a, xx);
  tau_forward<NY>(O, u1, lane, u.b, dummy);
  u.e = xx;
  return u;
}

// phirhs (solvers.f90:555-676) for theta <= 0.99:
// -re var/(dt th) v - re var/(2 th)(3 fn - fnm1) + (1-th)/th (k2 v - D2 v)
template <int NY>
__device__ __forceinline__ PF phirhs(const PF& v, const PF& d2v, const PF& fn, const PF& fnm1,
                                     double k2, double re, double var, double dt, double th) {
  const double c0 = -re * var / (dt * th);
  const double c1 = -re * var / (2.0 * th);
  const double c2 = (1.0 - th) / th;
  PF r = pf_scale(v, c0);
  r = pf_axpy(3.0 * c1, fn, r);
  r = pf_axpy(-c1, fnm1, r);
  r = pf_axpy(c2 * k2, v, r);
  r = pf_axpy(-c2, d2v, r);
  return r;
}


// mean over y of a Chebyshev series: sum over even n of a_n / (1 - n^2)
template <int NY>
__device__ __forceinline__ double pf_bulk(const PF& f, int lane) {
  constexpr int L = NY / 2;
  const double n = 2.0 * lane;
  double v = (lane < L) ? f.a.x / (1.0 - n * n) : 0.0;
  v = wave_sum(v);
  return v + f.e.x / (1.0 - (double)NY * (double)NY);
}

// ------------------------------------------------ shared by the pencil kernels --
constexpr int PW = 8;  // wavefronts (pencils) per workgroup

template <int NY, typename V>
__device__ __forceinline__ V* wave_lds(V* smem, int wv) {
  // twiddles [2NY], then two padded FFT buffers of length 2NY per wave
  return smem + 2 * NY + (size_t)wv * 2 * FFTBuf<2 * NY>::len;
}
// block-wide copy of the length-2NY twiddles into LDS (before any early exit)
template <int NY, typename V>
__device__ __forceinline__ const V* lds_twiddles(V* smem, const V* __restrict__ tw) {
  for (int e = threadIdx.x; e < 2 * NY; e += blockDim.x) smem[e] = tw[e];
  __syncthreads();
  return smem;
}

// (kx, kz) wavenumbers of pencil p = kx_index * nz + kz_index
template <class Args>
__device__ __forceinline__ void pencil_k(const Args& A, int p, double& kx, double& kz, int& kxi, int& kzi) {
  kxi = p / A.nz;
  kzi = p % A.nz;
  const int kzs = (kzi < A.nzh) ? kzi : kzi - A.nz;
  kx = kxi * A.alpha;
  kz = kzs * A.beta;
}

template <int NY>
__device__ __forceinline__ PF smooth_apply(const PF& f, const double* __restrict__ smy, double sz, int lane) {
  constexpr int L = NY / 2;
  PF r = f;
  if (lane < L) {
    r.a = cscale(f.a, sz * smy[2 * lane]);
    r.b = cscale(f.b, sz * smy[2 * lane + 1]);
  }
  r.e = cscale(f.e, sz * smy[NY]);
  return r;
}

// host-side helpers of the pencil launchers
#define KMM_NY_DISPATCH(ny, ...) \
  switch (ny) {                   \
    case 32: { constexpr int NY = 32; __VA_ARGS__; } break; \
    case 48: { constexpr int NY = 48; __VA_ARGS__; } break; \
    case 64: { constexpr int NY = 64; __VA_ARGS__; } break; \
    case 80: { constexpr int NY = 80; __VA_ARGS__; } break; \
    case 96: { constexpr int NY = 96; __VA_ARGS__; } break; \
    case 128: { constexpr int NY = 128; __VA_ARGS__; } break; \
    default: return hipErrorInvalidValue; \
  }

static inline int nblk(int P) { return (P + PW - 1) / PW; }
template <int NY, typename V>
static inline size_t lds_bytes() { return (size_t)(PW * 2 * FFTBuf<2 * NY>::len + 2 * NY) * sizeof(V); }

}  // namespace kmm
