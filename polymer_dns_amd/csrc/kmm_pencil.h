// Pencil-stage (wall-normal, Chebyshev) device routines.
//
// One wavefront owns one (kx,kz) pencil.  The Chebyshev coefficients (or
// Gauss-Lobatto values) n = 0..NY-1 are held as pairs: pair g holds n = 2g
// and n = 2g+1, and pair g sits on lane g % 64 of segment g / 64, so a pencil
// of NY <= 128 is one segment (lane l: n = 2l, 2l+1) and NY = 192 two; every
// lane carries a copy of n = NY.  With this layout the even and odd tau
// systems of FFTW_DNS/code/src/dns/solvers.f90 (evnc/oddc, rows coupling
// n-2, n, n+2) become nearest-neighbour lane couplings, and the sequential
// recurrences of the reference (cderiv, eqsolve, pntsolve) become wave64
// affine scans, chained from segment to segment.
#pragma once
#include "kmm_device.h"
#include "kmm_kernels.h"

namespace kmm {

// pairs and segments of an NY pencil
template <int NY>
struct PLay {
  static constexpr int L = NY / 2;
  static constexpr int S = (L + WAVE - 1) / WAVE;
};

template <int S>
struct PFT {         // one complex pencil in pair layout
  double2 a[S];      // n = 2g, g = 64 s + lane
  double2 b[S];      // n = 2g + 1
  double2 e;         // n = NY (uniform across lanes)
};
template <int NY>
using PFN = PFT<PLay<NY>::S>;

template <int NY>
__device__ __forceinline__ PFN<NY> pf_zero() {
  PFN<NY> f;
#pragma unroll
  for (int s = 0; s < PLay<NY>::S; ++s) f.a[s] = f.b[s] = czero();
  f.e = czero();
  return f;
}
template <int S>
__device__ __forceinline__ PFT<S> pf_axpy(double c, const PFT<S>& x, const PFT<S>& y) {
  PFT<S> r;
#pragma unroll
  for (int s = 0; s < S; ++s) { r.a[s] = cfma(c, x.a[s], y.a[s]); r.b[s] = cfma(c, x.b[s], y.b[s]); }
  r.e = cfma(c, x.e, y.e);
  return r;
}
template <int S>
__device__ __forceinline__ PFT<S> pf_scale(const PFT<S>& x, double c) {
  PFT<S> r;
#pragma unroll
  for (int s = 0; s < S; ++s) { r.a[s] = cscale(x.a[s], c); r.b[s] = cscale(x.b[s], c); }
  r.e = cscale(x.e, c);
  return r;
}
template <int S>
__device__ __forceinline__ PFT<S> pf_add(const PFT<S>& x, const PFT<S>& y) {
  PFT<S> r;
#pragma unroll
  for (int s = 0; s < S; ++s) { r.a[s] = cadd(x.a[s], y.a[s]); r.b[s] = cadd(x.b[s], y.b[s]); }
  r.e = cadd(x.e, y.e);
  return r;
}
template <int S>
__device__ __forceinline__ PFT<S> pf_sub(const PFT<S>& x, const PFT<S>& y) {
  PFT<S> r;
#pragma unroll
  for (int s = 0; s < S; ++s) { r.a[s] = csub(x.a[s], y.a[s]); r.b[s] = csub(x.b[s], y.b[s]); }
  r.e = csub(x.e, y.e);
  return r;
}
// i*c*x
template <int S>
__device__ __forceinline__ PFT<S> pf_imul(const PFT<S>& x, double c) {
  PFT<S> r;
#pragma unroll
  for (int s = 0; s < S; ++s) { r.a[s] = cimul(x.a[s], c); r.b[s] = cimul(x.b[s], c); }
  r.e = cimul(x.e, c);
  return r;
}
template <int S>
__device__ __forceinline__ PFT<S> pf_real(const PFT<S>& x) {
  PFT<S> r = x;
#pragma unroll
  for (int s = 0; s < S; ++s) { r.a[s].y = 0.0; r.b[s].y = 0.0; }
  r.e.y = 0.0;
  return r;
}
// complex number times a real pencil coefficientwise
template <int S>
__device__ __forceinline__ PFT<S> pf_cmul_real(double2 c, const PFT<S>& g) {
  PFT<S> r;
#pragma unroll
  for (int s = 0; s < S; ++s) { r.a[s] = cscale(c, g.a[s].x); r.b[s] = cscale(c, g.b[s].x); }
  r.e = cscale(c, g.e.x);
  return r;
}

// ------------------------------------------- cross-segment lane moves --
// value of pair g+1 (0 past the last segment)
template <int S, typename W>
__device__ __forceinline__ void seg_next(const W (&v)[S], W (&r)[S], int lane) {
#pragma unroll
  for (int s = 0; s < S; ++s) r[s] = shdn1(v[s]);
#pragma unroll
  for (int s = 0; s + 1 < S; ++s) {
    const W n = bcast(v[s + 1], 0);
    if (lane == WAVE - 1) r[s] = n;
  }
}
// value of pair g-1 (0 before the first)
template <int S, typename W>
__device__ __forceinline__ void seg_prev(const W (&v)[S], W (&r)[S], int lane) {
#pragma unroll
  for (int s = 0; s < S; ++s) r[s] = shup1(v[s]);
#pragma unroll
  for (int s = 1; s < S; ++s) {
    const W q = bcast(v[s - 1], WAVE - 1);
    if (lane == 0) r[s] = q;
  }
}
// value of pair G (a compile-time index) on every lane
template <int G, int S, typename W>
__device__ __forceinline__ W seg_bcast(const W (&v)[S]) { return bcast(v[G / WAVE], G % WAVE); }

// inclusive suffix sum over pairs: r_g = sum_{g' >= g} v_g'
template <int S>
__device__ __forceinline__ void seg_suffix_sum(const double2 (&v)[S], double2 (&r)[S], int lane) {
  double2 carry = czero();
#pragma unroll
  for (int s = S - 1; s >= 0; --s) {
    r[s] = suffix_sum(v[s], lane);
    if (s < S - 1) r[s] = cadd(r[s], carry);
    carry = bcast(r[s], 0);
  }
}
// suffix affine y_g = w_g + c_g y_{g+1} over pairs (y past the end = 0): the
// last lane of a segment takes the first value of the next segment as its
// y_{g+1} and starts the segment's own chain
template <int S, typename W>
__device__ __forceinline__ void seg_suffix_affine(const double (&c)[S], const W (&w)[S], W (&r)[S], int lane) {
  W y{};
#pragma unroll
  for (int s = S - 1; s >= 0; --s) {
    double cs = c[s];
    W ws = w[s];
    if (s < S - 1 && lane == WAVE - 1) {
      ws = affine_apply(cs, ws, y);
      cs = 0.0;
    }
    r[s] = suffix_affine_t<W>(cs, ws, lane);
    y = bcast(r[s], 0);
  }
}
// prefix affine y_g = a_g + b_g y_{g-1} over pairs (y before the start = 0)
template <int S>
__device__ __forceinline__ void seg_prefix_affine(const double (&b)[S], const double2 (&a)[S], double2 (&r)[S],
                                                  int lane) {
  double2 y = czero();
#pragma unroll
  for (int s = 0; s < S; ++s) {
    double bs = b[s];
    double2 as = a[s];
    if (s > 0 && lane == 0) {
      as = cfma(bs, y, as);
      bs = 0.0;
    }
    r[s] = prefix_affine(bs, as, lane);
    y = bcast(r[s], WAVE - 1);
  }
}

// field F laid out [y][pencil], P pencils per y-plane
template <int NY>
__device__ __forceinline__ PFN<NY> pf_load(const double2* __restrict__ F, size_t P, size_t p, int lane) {
  constexpr int L = PLay<NY>::L, S = PLay<NY>::S;
  PFN<NY> f;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g = s * WAVE + lane;
    if (g < L) {
      f.a[s] = F[(size_t)(2 * g) * P + p];
      f.b[s] = F[(size_t)(2 * g + 1) * P + p];
    } else {
      f.a[s] = f.b[s] = czero();
    }
  }
  f.e = F[(size_t)NY * P + p];
  return f;
}
template <int NY>
__device__ __forceinline__ void pf_store(double2* __restrict__ F, size_t P, size_t p, int lane, const PFN<NY>& f) {
  constexpr int L = PLay<NY>::L, S = PLay<NY>::S;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g = s * WAVE + lane;
    if (g < L) {
      F[(size_t)(2 * g) * P + p] = f.a[s];
      F[(size_t)(2 * g + 1) * P + p] = f.b[s];
    }
  }
  if (lane == 0) F[(size_t)NY * P + p] = f.e;
}
// the same for a field laid out pencil-major, [pencil][y] (the reference's
// own order of a spectral field): a pencil is NY+1 contiguous values
template <int NY>
__device__ __forceinline__ PFN<NY> pf_load_pm(const double2* __restrict__ F, size_t p, int lane) {
  constexpr int L = PLay<NY>::L, S = PLay<NY>::S;
  const double2* f0 = F + p * (size_t)(NY + 1);
  PFN<NY> f;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g = s * WAVE + lane;
    if (g < L) {
      f.a[s] = f0[2 * g];
      f.b[s] = f0[2 * g + 1];
    } else {
      f.a[s] = f.b[s] = czero();
    }
  }
  f.e = f0[NY];
  return f;
}
template <int NY>
__device__ __forceinline__ void pf_store_pm(double2* __restrict__ F, size_t p, int lane, const PFN<NY>& f) {
  constexpr int L = PLay<NY>::L, S = PLay<NY>::S;
  double2* f0 = F + p * (size_t)(NY + 1);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g = s * WAVE + lane;
    if (g < L) {
      f0[2 * g] = f.a[s];
      f0[2 * g + 1] = f.b[s];
    }
  }
  if (lane == 0) f0[NY] = f.e;
}
// zero rows 0..NY of pencil p of a field in the transfer precision V
template <int NY, typename V>
__device__ __forceinline__ void pf_store_zero(V* __restrict__ F, size_t P, size_t p, int lane) {
  constexpr int L = PLay<NY>::L, S = PLay<NY>::S;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g = s * WAVE + lane;
    if (g < L) { F[(size_t)(2 * g) * P + p] = vzero<V>(); F[(size_t)(2 * g + 1) * P + p] = vzero<V>(); }
  }
  if (lane == 0) F[(size_t)NY * P + p] = vzero<V>();
}
// real mean profile (nyp doubles)
template <int NY>
__device__ __forceinline__ PFN<NY> pf_load_real(const double* __restrict__ u, int lane) {
  constexpr int L = PLay<NY>::L, S = PLay<NY>::S;
  PFN<NY> f = pf_zero<NY>();
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g = s * WAVE + lane;
    if (g < L) { f.a[s].x = u[2 * g]; f.b[s].x = u[2 * g + 1]; }
  }
  f.e.x = u[NY];
  return f;
}
template <int NY>
__device__ __forceinline__ void pf_store_real(double* __restrict__ u, int lane, const PFN<NY>& f) {
  constexpr int L = PLay<NY>::L, S = PLay<NY>::S;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g = s * WAVE + lane;
    if (g < L) { u[2 * g] = f.a[s].x; u[2 * g + 1] = f.b[s].x; }
  }
  if (lane == 0) u[NY] = f.e.x;
}

// ------------------------------------------------ Chebyshev derivative --
// derivatives.f90:258-314: df_n = sum_{k>n, k-n odd} 2 k f_k dyde, df_0/2.
template <int NY>
__device__ __forceinline__ PFN<NY> cderiv(const PFN<NY>& f, double dyde, int lane) {
  constexpr int L = PLay<NY>::L, S = PLay<NY>::S;
  double2 go[S], ge[S], so[S], se[S], sx[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g = s * WAVE + lane;
    go[s] = cscale(f.b[s], 2.0 * (2 * g + 1) * dyde);  // odd k = 2g+1
    ge[s] = cscale(f.a[s], 2.0 * (2 * g) * dyde);      // even k = 2g
  }
  seg_suffix_sum<S>(go, so, lane);
  seg_suffix_sum<S>(ge, se, lane);
  seg_next<S>(se, sx, lane);  // 0 past the last pair
  PFN<NY> d;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g = s * WAVE + lane;
    d.a[s] = so[s];
    d.b[s] = cfma(2.0 * NY * dyde, f.e, sx[s]);
    if (g >= L) { d.a[s] = czero(); d.b[s] = czero(); }
  }
  d.e = czero();
  if (lane == 0) d.a[0] = cscale(d.a[0], 0.5);
  return d;
}

// derivatives.f90:541-595 (y=+1) and 370-427 (y=-1)
template <int NY>
__device__ __forceinline__ void wall_derivs(const PFN<NY>& f, double dyde, int lane, double2& top, double2& bot) {
  constexpr int L = PLay<NY>::L, S = PLay<NY>::S;
  double2 t = czero(), b = czero();
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g = s * WAVE + lane;
    const double n0 = 2.0 * g, n1 = 2.0 * g + 1;
    double2 ts = cfma(n1 * n1, f.b[s], cscale(f.a[s], n0 * n0));
    double2 bs = cfma(n1 * n1, f.b[s], cscale(f.a[s], -n0 * n0));
    if (g >= L) { ts = czero(); bs = czero(); }
    t = s ? cadd(t, ts) : ts;
    b = s ? cadd(b, bs) : bs;
  }
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
__device__ __forceinline__ V* cheb_to_phys(const PFN<NY>& f, V* buf, V* scr, const V* tw, int lane) {
  constexpr int L = PLay<NY>::L, S = PLay<NY>::S, M = 2 * NY;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g = s * WAVE + lane;
    if (g < L) {
      V a = to_v<V>((g == 0) ? f.a[s] : cscale(f.a[s], 0.5));
      V b = to_v<V>(cscale(f.b[s], 0.5));
      buf[fpad(2 * g)] = a;
      buf[fpad(2 * g + 1)] = b;
      if (g > 0) buf[fpad(M - 2 * g)] = a;
      buf[fpad(M - 2 * g - 1)] = b;
    }
  }
  if (lane == 0) buf[fpad(NY)] = to_v<V>(f.e);
  wave_sync();
  return wave_fft<M, -1, V, kChebPlan>(buf, scr, tw, lane);
}
template <int NY, typename V>
__device__ void cheb_to_phys_store(const PFN<NY>& f, V* buf, V* scr, const V* tw, int lane, V* __restrict__ dst,
                                   size_t P, size_t p) {
  constexpr int L = PLay<NY>::L, S = PLay<NY>::S;
  const V* r = cheb_to_phys<NY, V>(f, buf, scr, tw, lane);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g = s * WAVE + lane;
    if (g < L) {
      dst[(size_t)(2 * g) * P + p] = r[fpad(2 * g)];
      dst[(size_t)(2 * g + 1) * P + p] = r[fpad(2 * g + 1)];
    }
  }
  if (lane == 0) dst[(size_t)NY * P + p] = r[fpad(NY)];
  wave_sync();
}
// A pencil's physical-y values held in registers (pair g: y = 2g, 2g+1;
// every lane also y = NY), so that a load can be issued ahead of the
// transform that consumes it.
template <typename V, int S>
struct PhysPencil {
  V a[S], b[S], e;
};
template <int NY, typename V>
using PhysN = PhysPencil<V, PLay<NY>::S>;
template <int NY, typename V>
__device__ __forceinline__ PhysN<NY, V> phys_load(const V* __restrict__ src, size_t P, size_t p, int lane) {
  constexpr int L = PLay<NY>::L, S = PLay<NY>::S;
  PhysN<NY, V> q;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g0 = s * WAVE + lane;
    const int g = g0 < L ? g0 : L - 1;  // pairs >= L reload a valid row and ignore it
    q.a[s] = src[(size_t)(2 * g) * P + p];
    q.b[s] = src[(size_t)(2 * g + 1) * P + p];
  }
  q.e = src[(size_t)NY * P + p];
  return q;
}
// Forward (solvers.f90:3279-3416): DCT-I through the length-2NY FFT of the
// even extension, /NY, ends halved; input in the transfer precision V,
// result in FP64.
template <int NY, typename V>
__device__ __forceinline__ void cheb_fill(const PhysN<NY, V>& q, V* buf, int lane) {
  constexpr int L = PLay<NY>::L, S = PLay<NY>::S, M = 2 * NY;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g = s * WAVE + lane;
    if (g < L) {
      buf[fpad(2 * g)] = q.a[s];
      buf[fpad(2 * g + 1)] = q.b[s];
      if (g > 0) buf[fpad(M - 2 * g)] = q.a[s];
      buf[fpad(M - 2 * g - 1)] = q.b[s];
    }
  }
  if (lane == 0) buf[fpad(NY)] = q.e;
}
template <int NY, typename V>
__device__ __forceinline__ PFN<NY> cheb_take(const V* r, int lane) {
  constexpr int L = PLay<NY>::L, S = PLay<NY>::S;
  const double sc = 1.0 / NY;
  PFN<NY> o;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g = s * WAVE + lane;
    if (g < L) {
      o.a[s] = cscale(to_d(r[fpad(2 * g)]), g == 0 ? 0.5 * sc : sc);
      o.b[s] = cscale(to_d(r[fpad(2 * g + 1)]), sc);
    } else {
      o.a[s] = o.b[s] = czero();
    }
  }
  o.e = cscale(to_d(r[fpad(NY)]), 0.5 * sc);
  return o;
}
template <int NY, typename V>
__device__ PFN<NY> phys_to_cheb(const PhysN<NY, V>& q, V* buf, V* scr, const V* tw, int lane) {
  cheb_fill<NY, V>(q, buf, lane);
  wave_sync();
  V* r = wave_fft<2 * NY, -1, V, kChebPlan>(buf, scr, tw, lane);
  PFN<NY> o = cheb_take<NY, V>(r, lane);
  wave_sync();
  return o;
}

// the same, loading the pencil first
template <int NY, typename V>
__device__ PFN<NY> phys_to_cheb_load(const V* __restrict__ src, size_t P, size_t p, V* buf, V* scr, const V* tw,
                                     int lane) {
  return phys_to_cheb<NY, V>(phys_load<NY, V>(src, P, p, lane), buf, scr, tw, lane);
}

// -------------------------------------------------------- tau systems --
// Reciprocals of the tau row denominators for pair i (rows of pair 0 repeat
// pair 1's; pair 0 carries no row): even rows j = 2i, odd rows j = 2i + 1,
// 1/(4 j (j-1)), 1/(2 (j^2-1)), 1/(4 j (j+1)).  Folded at compile time with
// IEEE division, so the values equal the run-time quotients bit for bit.
constexpr int kMaxPairs = 128;
struct TauRcp {
  double r[6][kMaxPairs];
  constexpr TauRcp() : r{} {
    for (int i = 0; i < kMaxPairs; ++i) {
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
// owned by pair i (1 <= i); wn = kx^2 + kz^2 + g.
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
// Even row at pair i (1..L-1) and extra row (index L); odd row at pair
// i (1..L-1).  Pair 0 is left for the boundary row.
template <int NY>
__device__ __forceinline__ void tau_rhs(const PFN<NY>& f, int lane, double2 (&we)[PLay<NY>::S],
                                        double2 (&wo)[PLay<NY>::S], double2& wex) {
  constexpr int L = PLay<NY>::L, S = PLay<NY>::S;
  double2 sme[S], spe[S], smo[S], spo[S];
  seg_prev<S>(f.a, sme, lane);
  seg_next<S>(f.a, spe, lane);
  seg_prev<S>(f.b, smo, lane);
  seg_next<S>(f.b, spo, lane);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int i = s * WAVE + lane;
    we[s] = czero();
    wo[s] = czero();
    if (i >= 1 && i <= L - 1) {
      const double r1 = kTauRcp.r[0][i], r2 = kTauRcp.r[1][i], r3 = kTauRcp.r[2][i];
      if (i <= L - 2) {
        const double cm = (i == 1) ? 2.0 : 1.0;
        we[s] = cadd(cadd(cscale(sme[s], -cm * r1), cscale(f.a[s], r2)), cscale(spe[s], -r3));
      } else {
        we[s] = cadd(cscale(sme[s], -r1), cscale(f.a[s], r2));
      }
      const double q1 = kTauRcp.r[3][i], q2 = kTauRcp.r[4][i], q3 = kTauRcp.r[5][i];
      if (i <= L - 3) {
        wo[s] = cadd(cadd(cscale(smo[s], -q1), cscale(f.b[s], q2)), cscale(spo[s], -q3));
      } else if (i == L - 2) {
        wo[s] = cadd(cscale(smo[s], -q1), cscale(f.b[s], q2));
      } else {
        wo[s] = cscale(smo[s], -q1);
      }
    }
  }
  const double jx = (double)NY;
  wex = cscale(seg_bcast<L - 1, S>(f.a), -1.0 / (4.0 * jx * (jx - 1.0)));
}

// One parity of the elimination of eqsolve/pntsolve (solvers.f90:500-553,
// 906-1017).  Pair i (1..L-1) holds row i+1: RHS w, a1(i), a3(i), 1/d_i.
// Even parity also has the extra row (index L, unknown L+1).
template <int S>
struct TauPar {
  double2 wp[S];  // w' (eliminated RHS) at pair i, row i+1
  double a1[S], invd[S];
  double2 wpx;    // even extra w'
  double a1x, invdx;
  bool even;
};

template <int NY>
__device__ __forceinline__ TauPar<PLay<NY>::S> tau_eliminate(bool even, double wn, double dyde,
                                                             const double (&invd)[PLay<NY>::S], double invdx,
                                                             const double2 (&w)[PLay<NY>::S], double2 wx, int lane) {
  constexpr int L = PLay<NY>::L, S = PLay<NY>::S;
  TauPar<S> t;
  t.even = even;
  double invd_next[S], nrho[S];
  double2 wv[S];
  seg_next<S>(invd, invd_next, lane);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g = s * WAVE + lane;
    // coef_even / coef_odd of the pair's row with the tabulated reciprocals
    // (wn r in place of wn / (1/r): within one rounding of the quotient)
    const int li = g < 1 ? 1 : g;
    double a1, a3;
    if (even) {
      a1 = wn * kTauRcp.r[0][li];
      if (li == 1) a1 *= 2.0;
      a3 = (li <= L - 2) ? wn * kTauRcp.r[2][li] : 0.0;
    } else {
      a1 = wn * kTauRcp.r[3][li];
      a3 = (li <= L - 3) ? wn * kTauRcp.r[5][li] : 0.0;
    }
    const bool row = (g >= 1 && g <= L - 1);
    const double rho = row ? a3 * invd_next[s] : 0.0;
    nrho[s] = -rho;
    wv[s] = row ? w[s] : czero();
    t.a1[s] = row ? a1 : 0.0;
    t.invd[s] = row ? invd[s] : 0.0;
  }
  seg_suffix_affine<S, double2>(nrho, wv, t.wp, lane);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g = s * WAVE + lane;
    if (!(g >= 1 && g <= L - 1)) t.wp[s] = czero();
  }
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
__device__ __forceinline__ double2 tau_bcsum(const TauPar<PLay<NY>::S>& t, TW tq, int lane, double& F1) {
  constexpr int L = PLay<NY>::L, S = PLay<NY>::S;
  // Chebyshev index of unknown q: even 2(q-1), odd 2q-1
  auto T = [&](int q) -> double { return t.even ? tq(2 * (q - 1)) : tq(2 * q - 1); };
  double carry;
  if (t.even) carry = T(L) - T(L + 1) * t.a1x * t.invdx;   // F_L
  else carry = T(L);                                          // F_{m+1}, m = L-1
  double c[S], wv[S], F[S], Fn[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g = s * WAVE + lane;
    const bool row = (g >= 1 && g <= L - 1);
    const double k = -t.a1[s] * t.invd[s];
    c[s] = row ? k : 0.0;
    wv[s] = row ? T(g) : 0.0;
    if (g == L - 1) { wv[s] += k * carry; c[s] = 0.0; }
  }
  seg_suffix_affine<S, double>(c, wv, F, lane);
  seg_next<S>(F, Fn, lane);
  double2 term = czero();
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g = s * WAVE + lane;
    const bool row = (g >= 1 && g <= L - 1);
    if (g == L - 1) Fn[s] = carry;
    const double2 ts = row ? cscale(t.wp[s], Fn[s] * t.invd[s]) : czero();
    term = s ? cadd(term, ts) : ts;
  }
  double2 sum = wave_sum(term);
  if (t.even) sum = cfma(T(L + 1) * t.invdx, t.wpx, sum);
  F1 = bcast(F[0], 1);
  return sum;
}

// Forward substitution x_{i+1} = (w'_{i+1} - a1_i x_i)/d_i from x_1.
template <int NY>
__device__ __forceinline__ void tau_forward(const TauPar<PLay<NY>::S>& t, double2 x1, int lane,
                                            double2 (&x)[PLay<NY>::S], double2& xx) {
  constexpr int L = PLay<NY>::L, S = PLay<NY>::S;
  double b[S];
  double2 a[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g = s * WAVE + lane;
    const bool row = (g >= 1 && g <= L - 1);
    b[s] = row ? -t.a1[s] * t.invd[s] : 0.0;
    a[s] = row ? cscale(t.wp[s], t.invd[s]) : (g == 0 ? x1 : czero());
  }
  seg_prefix_affine<S>(b, a, x, lane);
#pragma unroll
  for (int s = 0; s < S; ++s)
    if (s * WAVE + lane >= L) x[s] = czero();
  if (t.even) {
    const double2 xl = seg_bcast<L - 1, S>(x);
    xx = cscale(cfma(-t.a1x, xl, t.wpx), t.invdx);
  } else {
    xx = czero();
  }
}

// the pencil's 1/d table entries of its pairs
template <int NY>
__device__ __forceinline__ void tab_load(const double2* __restrict__ tab, int lane, double (&te)[PLay<NY>::S],
                                         double (&to)[PLay<NY>::S], double& invdx) {
  constexpr int L = PLay<NY>::L, S = PLay<NY>::S;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g = s * WAVE + lane;
    const double2 v = (S == 1 || g < L) ? tab[g] : make_double2(0.0, 0.0);
    te[s] = v.x;
    to[s] = v.y;
  }
  invdx = tab[tab_width(NY) - 1].x;
}

// a pencil's 1/d table entries held in registers (loaded ahead of the solve)
template <int S>
struct TabV {
  double te[S], to[S], invdx;
};
template <int NY>
using TabN = TabV<PLay<NY>::S>;
template <int NY>
__device__ __forceinline__ TabN<NY> tab_get(const double2* __restrict__ tab, int lane) {
  TabN<NY> t;
  tab_load<NY>(tab, lane, t.te, t.to, t.invdx);
  return t;
}

// solve (solvers.f90:174-283): (dyde^2 D^2 - wn) u = f, Dirichlet rows
// (ib=0) or Neumann rows (ib!=0, wn already guarded by the caller).
// tb: the pencil's 1/d table (tab_width entries).
template <int NY>
__device__ PFN<NY> tau_solve(const PFN<NY>& f, double wn, double dyde, const TabN<NY>& tb, double2 bctop,
                             double2 bcbot, int ib, int lane) {
  constexpr int S = PLay<NY>::S;
  double2 we[S], wo[S], wex;
  tau_rhs<NY>(f, lane, we, wo, wex);
  const TauPar<S> E = tau_eliminate<NY>(true, wn, dyde, tb.te, tb.invdx, we, wex, lane);
  const TauPar<S> O = tau_eliminate<NY>(false, wn, dyde, tb.to, 0.0, wo, czero(), lane);
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
  PFN<NY> u;
  double2 xx, dummy;
  tau_forward<NY>(E, x1e, lane, u.a, xx);
  tau_forward<NY>(O, x1o, lane, u.b, dummy);
  u.e = xx;
  return u;
}
template <int NY>
__device__ __forceinline__ PFN<NY> tau_solve(const PFN<NY>& f, double wn, double dyde,
                                             const double2* __restrict__ tab, double2 bctop, double2 bcbot, int ib,
                                             int lane) {
  return tau_solve<NY>(f, wn, dyde, tab_get<NY>(tab, lane), bctop, bcbot, ib, lane);
}

// penta (solvers.f90:680-1061) with Robin rows sum tt(n) u_n = gtop,
// sum tb(n) u_n = gbot; tt/tb are closed-form functions of n.
template <int NY, typename TT, typename TB>
__device__ PFN<NY> tau_penta(const PFN<NY>& f, double wn, double dyde, const TabN<NY>& tab, TT tt, TB tb,
                             double2 gtop, double2 gbot, int lane) {
  constexpr int S = PLay<NY>::S;
  double2 we[S], wo[S], wex;
  tau_rhs<NY>(f, lane, we, wo, wex);
  const TauPar<S> E = tau_eliminate<NY>(true, wn, dyde, tab.te, tab.invdx, we, wex, lane);
  const TauPar<S> O = tau_eliminate<NY>(false, wn, dyde, tab.to, 0.0, wo, czero(), lane);
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
  PFN<NY> u;
  double2 xx, dummy;
  tau_forward<NY>(E, u0, lane, u.a, xx);
  tau_forward<NY>(O, u1, lane, u.b, dummy);
  u.e = xx;
  return u;
}
template <int NY, typename TT, typename TB>
__device__ __forceinline__ PFN<NY> tau_penta(const PFN<NY>& f, double wn, double dyde,
                                             const double2* __restrict__ tab, TT tt, TB tb, double2 gtop,
                                             double2 gbot, int lane) {
  return tau_penta<NY>(f, wn, dyde, tab_get<NY>(tab, lane), tt, tb, gtop, gbot, lane);
}

// phirhs (solvers.f90:555-676) for theta <= 0.99:
// -re var/(dt th) v - re var/(2 th)(3 fn - fnm1) + (1-th)/th (k2 v - D2 v)
template <int NY>
__device__ __forceinline__ PFN<NY> phirhs(const PFN<NY>& v, const PFN<NY>& d2v, const PFN<NY>& fn,
                                          const PFN<NY>& fnm1, double k2, double re, double var, double dt,
                                          double th) {
  const double c0 = -re * var / (dt * th);
  const double c1 = -re * var / (2.0 * th);
  const double c2 = (1.0 - th) / th;
  PFN<NY> r = pf_scale(v, c0);
  r = pf_axpy(3.0 * c1, fn, r);
  r = pf_axpy(-c1, fnm1, r);
  r = pf_axpy(c2 * k2, v, r);
  r = pf_axpy(-c2, d2v, r);
  return r;
}


// mean over y of a Chebyshev series: sum over even n of a_n / (1 - n^2)
template <int NY>
__device__ __forceinline__ double pf_bulk(const PFN<NY>& f, int lane) {
  constexpr int L = PLay<NY>::L, S = PLay<NY>::S;
  double v = 0.0;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g = s * WAVE + lane;
    const double n = 2.0 * g;
    const double vs = (g < L) ? f.a[s].x / (1.0 - n * n) : 0.0;
    v = s ? v + vs : vs;
  }
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
// ------------------------------------------------ block-staged tiles --
// The PWS consecutive pencils p0 .. p0+PWS-1 of a block, rows y = 0..NY of
// a field laid out [y][pencil], staged through an LDS tile [NY+1][PWS+1]
// (one pad element per row): the block's global accesses then cover whole
// rows of PWS pencils (64 bytes per row for FP64), and each wave takes its
// own pencil's column from the tile.  Thread t moves elements
// e = t + k PWS*64 (row e / PWS, pencil e % PWS).  Four pencils per block
// keep the staged step kernel within half the CU's LDS, so that two blocks
// share a CU and one computes while the other waits at a block barrier.
constexpr int PWS = 4;
constexpr int TLD = PWS + 1;
template <int NY>
struct TileG {
  static constexpr int ROWS = NY + 1, NEL = ROWS * PWS, LEN = ROWS * TLD;
  static constexpr int PER = (NEL + PWS * WAVE - 1) / (PWS * WAVE);
};
template <int NY, typename E>
__device__ __forceinline__ void tile_fetch(const E* __restrict__ F, size_t P, int p0, E (&r)[TileG<NY>::PER]) {
  using G = TileG<NY>;
#pragma unroll
  for (int k = 0; k < G::PER; ++k) {
    const int e = (int)threadIdx.x + k * PWS * WAVE;
    const int y = e / PWS, w = e % PWS;
    r[k] = (e < G::NEL && (size_t)(p0 + w) < P) ? F[(size_t)y * P + p0 + w] : vzero<E>();
  }
}
template <int NY, typename E>
__device__ __forceinline__ void tile_put(E* t, const E (&r)[TileG<NY>::PER]) {
  using G = TileG<NY>;
#pragma unroll
  for (int k = 0; k < G::PER; ++k) {
    const int e = (int)threadIdx.x + k * PWS * WAVE;
    if (e < G::NEL) t[(e / PWS) * TLD + e % PWS] = r[k];
  }
}
template <int NY, typename E>
__device__ __forceinline__ void tile_store(const E* t, E* __restrict__ F, size_t P, int p0) {
  using G = TileG<NY>;
#pragma unroll
  for (int k = 0; k < G::PER; ++k) {
    const int e = (int)threadIdx.x + k * PWS * WAVE;
    const int y = e / PWS, w = e % PWS;
    if (e < G::NEL && (size_t)(p0 + w) < P) F[(size_t)y * P + p0 + w] = t[y * TLD + w];
  }
}
// physical values of column wv of a tile in the transfer precision
template <int NY, typename V>
__device__ __forceinline__ PhysN<NY, V> tile_phys(const V* t, int wv, int lane) {
  constexpr int L = PLay<NY>::L, S = PLay<NY>::S;
  PhysN<NY, V> q;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g0 = s * WAVE + lane;
    const int g = g0 < L ? g0 : L - 1;
    q.a[s] = t[(2 * g) * TLD + wv];
    q.b[s] = t[(2 * g + 1) * TLD + wv];
  }
  q.e = t[NY * TLD + wv];
  return q;
}
// inverse transform of f into column wv of a transfer tile (zeros: zero column)
template <int NY, typename V>
__device__ __forceinline__ void cheb_to_phys_tile(const PFN<NY>& f, V* buf, V* scr, const V* tw, int lane, V* t,
                                                  int wv) {
  constexpr int L = PLay<NY>::L, S = PLay<NY>::S;
  const V* r = cheb_to_phys<NY, V>(f, buf, scr, tw, lane);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g = s * WAVE + lane;
    if (g < L) {
      t[(2 * g) * TLD + wv] = r[fpad(2 * g)];
      t[(2 * g + 1) * TLD + wv] = r[fpad(2 * g + 1)];
    }
  }
  if (lane == 0) t[NY * TLD + wv] = r[fpad(NY)];
  wave_sync();
}
template <int NY, typename V>
__device__ __forceinline__ void tile_col_zero(V* t, int wv, int lane) {
  for (int y = lane; y <= NY; y += WAVE) t[y * TLD + wv] = vzero<V>();
}

// Pencil group of block b of nb: blocks are dispatched round-robin over the
// 8 XCDs (block b on XCD b % 8), each with its own L2; groups 4k .. 4k+3
// share the cache lines of every [y][pencil] row (64 bytes per group for
// FP64, 32 for FP32), so they go to blocks b, b+8, b+16, b+24 of one XCD.
// A bijection of 0..nb-1 (blocks past the last multiple of 32 keep their
// index).
__device__ __forceinline__ int xcd_group(int b, int nb) {
  if (b >= nb / 32 * 32) return b;
  const int x = b % 8, s = b / 8;
  return (s / 4) * 32 + x * 4 + s % 4;
}

// LDS plan of the staged step kernel (bytes): twiddles, then a region used
// as [FFT buffers | two nonlinear-term tiles] while the nonlinear terms are
// transformed (two rounds of two) and as [FFT buffers | transfer tiles] for
// the inverse transforms; at most 54 KB, so that three blocks of four waves
// share a CU (three waves per SIMD at the step kernel's 168 registers).
template <int NY, typename V, bool SCALAR>
struct StepLds {
  // three blocks per CU up to ny = 128, two at ny = 192
  static constexpr size_t CAP = (NY <= 128 ? 54 : 80) * 1024;
  static constexpr size_t TWB = (size_t)2 * NY * sizeof(V);
  static constexpr size_t FFT = (size_t)PWS * FFTBuf<2 * NY>::len * sizeof(V);  // one in-place buffer per wave
  static constexpr size_t TV = (size_t)TileG<NY>::LEN * sizeof(V);
  static constexpr int NOUT = SCALAR ? 9 : 6;
  static constexpr int fit = (CAP > TWB + FFT) ? (int)((CAP - TWB - FFT) / TV) : 0;
  static constexpr int NTS = fit < NOUT ? fit : NOUT;  // transfer tiles per flush
  static constexpr size_t A1 = FFT + 2 * TV, A3 = FFT + (size_t)NTS * TV;
  static constexpr size_t REGION = A1 > A3 ? A1 : A3;
  static constexpr bool ok = NTS >= 3 && TWB + REGION <= CAP && fft_in_place<2 * NY, kChebPlan>();
  static constexpr size_t bytes = TWB + REGION;
};

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
__device__ __forceinline__ PFN<NY> smooth_apply(const PFN<NY>& f, const double* __restrict__ smy, double sz,
                                                int lane) {
  constexpr int L = PLay<NY>::L, S = PLay<NY>::S;
  PFN<NY> r = f;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int g = s * WAVE + lane;
    if (g < L) {
      r.a[s] = cscale(f.a[s], sz * smy[2 * g]);
      r.b[s] = cscale(f.b[s], sz * smy[2 * g + 1]);
    }
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
    case 192: { constexpr int NY = 192; __VA_ARGS__; } break; \
    default: return hipErrorInvalidValue; \
  }

static inline int nblk(int P) { return (P + PW - 1) / PW; }
template <int NY, typename V>
static inline size_t lds_bytes(int waves = PW) { return (size_t)(waves * 2 * FFTBuf<2 * NY>::len + 2 * NY) * sizeof(V); }
// the step kernel's launch: the staged plan (PWS pencils per block) where
// it fits, else PW pencils per block with their FFT buffers
template <int NY, typename V, bool SCALAR, int MODE>
constexpr bool step_tiled() { return MODE == 0 && StepLds<NY, V, SCALAR>::ok; }
template <int NY, typename V, bool SCALAR, int MODE>
static inline size_t step_lds_bytes() {
  if constexpr (step_tiled<NY, V, SCALAR, MODE>()) return StepLds<NY, V, SCALAR>::bytes;
  else return lds_bytes<NY, V>();
}
template <int NY, typename V, bool SCALAR, int MODE>
static inline int step_blocks(int P) {
  return step_tiled<NY, V, SCALAR, MODE>() ? (P + PWS - 1) / PWS : nblk(P);
}
template <int NY, typename V, bool SCALAR, int MODE>
static inline int step_threads() { return (step_tiled<NY, V, SCALAR, MODE>() ? PWS : PW) * 64; }

}  // namespace kmm
