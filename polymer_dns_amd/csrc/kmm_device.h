// Device-side primitives shared by the pencil (y) and plane (x/z) kernels of
// the KMM channel step: complex FP64 arithmetic, wave64 scans/reductions, and
// an LDS Stockham FFT for lengths 2^a 3^b 5^c.  Written for gfx950 (CDNA4):
// 64-wide wavefronts, one wavefront owns one FFT line / one Chebyshev pencil.
#pragma once
#include <hip/hip_runtime.h>

namespace kmm {

constexpr int WAVE = 64;

// ---------------------------------------------------------------- complex --
// Complex arithmetic on double2 (FP64 path) and float2 (FP32 transforms of
// the mixed-precision path).
template <typename V> struct Scal;
template <> struct Scal<double2> { using T = double; };
template <> struct Scal<float2> { using T = float; };
template <typename V>
__device__ __forceinline__ V mkv(typename Scal<V>::T x, typename Scal<V>::T y) { V r; r.x = x; r.y = y; return r; }

template <typename V> __device__ __forceinline__ V cadd(V a, V b) { return mkv<V>(a.x + b.x, a.y + b.y); }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { return mkv<V>(a.x - b.x, a.y - b.y); }
template <typename V> __device__ __forceinline__ V cscale(V a, typename Scal<V>::T s) { return mkv<V>(a.x * s, a.y * s); }
template <typename V> __device__ __forceinline__ V cmul(V a, V b) {
  return mkv<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename V> __device__ __forceinline__ V cconj(V a) { return mkv<V>(a.x, -a.y); }
// i*s*a for real s
template <typename V> __device__ __forceinline__ V cimul(V a, typename Scal<V>::T s) { return mkv<V>(-a.y * s, a.x * s); }
__device__ __forceinline__ double2 czero() { return make_double2(0.0, 0.0); }
template <typename V> __device__ __forceinline__ V vzero() { return mkv<V>(0, 0); }
// a + s*b
__device__ __forceinline__ double2 cfma(double s, double2 b, double2 a) {
  return make_double2(fma(s, b.x, a.x), fma(s, b.y, a.y));
}
// precision conversion
template <typename V> __device__ __forceinline__ V to_v(double2 a) {
  return mkv<V>((typename Scal<V>::T)a.x, (typename Scal<V>::T)a.y);
}
template <typename V> __device__ __forceinline__ double2 to_d(V a) { return make_double2((double)a.x, (double)a.y); }

// ------------------------------------------------------------ wave sync --
// Orders LDS traffic between lanes of one wavefront (each wave owns its own
// LDS region, so no workgroup barrier is needed).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// ------------------------------------------------------------- shuffles --
// DPP moves (register-to-register lane exchange, no LDS): lanes whose
// source is outside the row/wave read 0.
constexpr int DPP_ROW_SHL = 0x100;   // lane i <- lane i+k within a row of 16
constexpr int DPP_ROW_SHR = 0x110;   // lane i <- lane i-k within a row of 16
constexpr int DPP_WAVE_SHL1 = 0x130; // lane i <- lane i+1 across the wave
constexpr int DPP_WAVE_SHR1 = 0x138; // lane i <- lane i-1 across the wave

template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  const long long b = __double_as_longlong(v);
  int lo = (int)(unsigned long long)b, hi = (int)((unsigned long long)b >> 32);
  lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xF, 0xF, true);
  hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, 0xF, 0xF, true);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <int CTRL>
__device__ __forceinline__ double2 dpp(double2 v) { return make_double2(dpp<CTRL>(v.x), dpp<CTRL>(v.y)); }

__device__ __forceinline__ double rdlane(double v, int l) {
  const long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_readlane((int)(unsigned long long)b, l);
  int hi = __builtin_amdgcn_readlane((int)((unsigned long long)b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double2 rdlane(double2 v, int l) { return make_double2(rdlane(v.x, l), rdlane(v.y, l)); }

// neighbour exchange (lane+1 / lane-1; 0 beyond the wave ends)
__device__ __forceinline__ double shdn1(double v) { return dpp<DPP_WAVE_SHL1>(v); }
__device__ __forceinline__ double2 shdn1(double2 v) { return dpp<DPP_WAVE_SHL1>(v); }
__device__ __forceinline__ double shup1(double v) { return dpp<DPP_WAVE_SHR1>(v); }
__device__ __forceinline__ double2 shup1(double2 v) { return dpp<DPP_WAVE_SHR1>(v); }

__device__ __forceinline__ double shdn(double v, int off) { return __shfl_down(v, off, WAVE); }
__device__ __forceinline__ double shup(double v, int off) { return __shfl_up(v, off, WAVE); }
__device__ __forceinline__ double2 shdn(double2 v, int off) {
  return make_double2(__shfl_down(v.x, off, WAVE), __shfl_down(v.y, off, WAVE));
}
__device__ __forceinline__ double2 shup(double2 v, int off) {
  return make_double2(__shfl_up(v.x, off, WAVE), __shfl_up(v.y, off, WAVE));
}
__device__ __forceinline__ double bcast(double v, int src) { return rdlane(v, src); }
__device__ __forceinline__ double2 bcast(double2 v, int src) { return rdlane(v, src); }

// ---------------------------------------------------------------- scans --
// All scans run four DPP levels inside each row of 16 lanes and combine the
// four row aggregates read with v_readlane.  A DPP move reads 0 where its
// source lane lies outside the row, which makes the additive updates of
// those lanes no-ops without a select.

// inclusive suffix sum: r_l = sum_{l' >= l} v_l'
__device__ __forceinline__ double2 suffix_sum(double2 v, int lane) {
  v = cadd(v, dpp<DPP_ROW_SHL | 1>(v));
  v = cadd(v, dpp<DPP_ROW_SHL | 2>(v));
  v = cadd(v, dpp<DPP_ROW_SHL | 4>(v));
  v = cadd(v, dpp<DPP_ROW_SHL | 8>(v));
  const double2 a1 = rdlane(v, 16), a2 = rdlane(v, 32), a3 = rdlane(v, 48);
  const int row = lane >> 4;
  double2 add = czero();
  if (row < 3) add = cadd(add, a3);
  if (row < 2) add = cadd(add, a2);
  if (row < 1) add = cadd(add, a1);
  return cadd(v, add);
}

__device__ __forceinline__ double wave_sum(double v) {
  v += dpp<DPP_ROW_SHL | 1>(v);
  v += dpp<DPP_ROW_SHL | 2>(v);
  v += dpp<DPP_ROW_SHL | 4>(v);
  v += dpp<DPP_ROW_SHL | 8>(v);
  return (rdlane(v, 0) + rdlane(v, 16)) + (rdlane(v, 32) + rdlane(v, 48));
}
__device__ __forceinline__ double2 wave_sum(double2 v) { return make_double2(wave_sum(v.x), wave_sum(v.y)); }
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, WAVE));
  return v;
}

// Suffix affine recurrence y_l = w_l + c_l * y_{l+1} over all 64 lanes
// (y_64 = 0).  A lane with c_l = 0 starts a chain.  The map of lane l is
// x -> w_l + c_l x; the scan composes maps over increasing lanes.
template <typename W>
__device__ __forceinline__ W affine_apply(double c, W w, W x);
template <>
__device__ __forceinline__ double2 affine_apply<double2>(double c, double2 w, double2 x) { return cfma(c, x, w); }
template <>
__device__ __forceinline__ double affine_apply<double>(double c, double w, double x) { return fma(c, x, w); }

template <typename W>
__device__ __forceinline__ W suffix_affine_t(double c, W w, int lane) {
  const int r = lane & 15;
#define KMM_SFX_LEVEL(K)                                  \
  {                                                       \
    const double c2 = dpp<DPP_ROW_SHL | K>(c);            \
    w = affine_apply(c, w, dpp<DPP_ROW_SHL | K>(w));      \
    if (r + K < 16) c = c * c2;                           \
  }
  KMM_SFX_LEVEL(1) KMM_SFX_LEVEL(2) KMM_SFX_LEVEL(4) KMM_SFX_LEVEL(8)
#undef KMM_SFX_LEVEL
  // row aggregates (maps of whole rows 1..3), composed from the top
  const double c1 = rdlane(c, 16), c2 = rdlane(c, 32), c3 = rdlane(c, 48);
  const W w1 = rdlane(w, 16), w2 = rdlane(w, 32), w3 = rdlane(w, 48);
  // value entering row k from above: y at lane 16k = row map k applied to the rows after it
  const W y3 = w3;                              // y_64 = 0
  const W y2 = affine_apply(c2, w2, y3);
  const W y1 = affine_apply(c1, w1, y2);
  const int row = lane >> 4;
  if (row == 3) return w;
  const W yin = (row == 0) ? y1 : ((row == 1) ? y2 : y3);
  return affine_apply(c, w, yin);
}
__device__ __forceinline__ double2 suffix_affine(double c, double2 w, int lane) { return suffix_affine_t<double2>(c, w, lane); }
__device__ __forceinline__ double suffix_affine_r(double c, double w, int lane) { return suffix_affine_t<double>(c, w, lane); }

// Prefix affine recurrence y_l = a_l + b_l * y_{l-1} (y_{-1} = 0).
__device__ __forceinline__ double2 prefix_affine(double b, double2 a, int lane) {
  const int r = lane & 15;
#define KMM_PFX_LEVEL(K)                                  \
  {                                                       \
    const double b2 = dpp<DPP_ROW_SHR | K>(b);            \
    a = cfma(b, dpp<DPP_ROW_SHR | K>(a), a);              \
    if (r >= K) b = b * b2;                               \
  }
  KMM_PFX_LEVEL(1) KMM_PFX_LEVEL(2) KMM_PFX_LEVEL(4) KMM_PFX_LEVEL(8)
#undef KMM_PFX_LEVEL
  const double b0 = rdlane(b, 15), b1 = rdlane(b, 31), b2 = rdlane(b, 47);
  const double2 a0 = rdlane(a, 15), a1 = rdlane(a, 31), a2 = rdlane(a, 47);
  const double2 y0 = a0;                       // y at lane 15 (y_{-1} = 0)
  const double2 y1 = cfma(b1, y0, a1);         // y at lane 31
  const double2 y2 = cfma(b2, y1, a2);         // y at lane 47
  (void)b0;
  const int row = lane >> 4;
  if (row == 0) return a;
  const double2 yin = (row == 1) ? y0 : (row == 2 ? y1 : y2);
  return cfma(b, yin, a);
}

// ------------------------------------------------------------------ FFT --
// Radix plan for N = 2^a 3^b 5^c (the radices of its stages, in order).
struct FFTPlan {
  int n;
  int nst;
  int r[16];
};
// Radix plan for one wavefront: each stage runs ceil((N/R)/64) butterfly
// rounds.  Cost model 0 minimises the total number of rounds (then the
// number of stages); model 1 minimises an estimate of the instructions a
// lane issues: per round of a radix-R stage, the butterfly's adds and
// multiplies, R-1 twiddle multiplies (none in the first stage), R LDS loads
// and stores and the index arithmetic, plus a fixed cost per stage for its
// wave barrier and LDS latency.  Radices 8, 6, 5, 4, 3, 2.
struct PlanSearch {
  int best_cost, best_nst, cur_nst;
  int best[16], cur[16];
};
__host__ __device__ constexpr int radix_cost(int r, bool first) {
  // butterfly arithmetic (real ops) + twiddles (6 per complex multiply) +
  // 2R LDS ops + 10 index operations
  const int bf = r == 2 ? 4 : r == 3 ? 12 : r == 4 ? 16 : r == 5 ? 34 : r == 6 ? 40 : 60;
  return bf + (first ? 0 : 6 * (r - 1)) + 2 * r + 10;
}
__host__ __device__ constexpr void plan_dfs(PlanSearch& s, int n, int m, int cost, int model) {
  if (m == 1) {
    if (cost < s.best_cost || (cost == s.best_cost && s.cur_nst < s.best_nst)) {
      s.best_cost = cost;
      s.best_nst = s.cur_nst;
      for (int i = 0; i < s.cur_nst; ++i) s.best[i] = s.cur[i];
    }
    return;
  }
  if (s.cur_nst >= 16 || cost >= s.best_cost) return;
  const int rad[6] = {8, 6, 5, 4, 3, 2};
  for (int i = 0; i < 6; ++i) {
    const int r = rad[i];
    if (m % r) continue;
    const int rounds = (n / r + 63) / 64;
    const int c = model == 0 ? rounds : rounds * radix_cost(r, s.cur_nst == 0) + 24;
    s.cur[s.cur_nst++] = r;
    plan_dfs(s, n, m / r, cost + c, model);
    --s.cur_nst;
  }
}
__host__ __device__ constexpr FFTPlan make_plan(int n, int model = 0) {
  PlanSearch s{1 << 30, 99, 0, {0}, {0}};
  plan_dfs(s, n, n, 0, model);
  FFTPlan p{n, s.best_nst, {0}};
  for (int i = 0; i < s.best_nst; ++i) p.r[i] = s.best[i];
  return p;
}

// A wave's FFT can run in place (one buffer) when every stage of its plan
// is one round (N / R <= 64): a stage's LDS reads of all lanes then precede
// its writes, which depend on them.
template <int N, int MODEL>
constexpr bool fft_in_place() {
  constexpr FFTPlan P = make_plan(N, MODEL);
  for (int s = 0; s < P.nst; ++s)
    if (N / P.r[s] > WAVE) return false;
  return true;
}

// twiddle w^k with w = exp(-2 pi i / N) from a table of N entries; DIR=+1
// (FFTW_BACKWARD) uses the conjugate.
template <int DIR, typename V>
__device__ __forceinline__ V twid(const V* __restrict__ tw, int k) {
  V t = tw[k];
  if (DIR > 0) t.y = -t.y;
  return t;
}

template <int DIR, typename V>
__device__ __forceinline__ void dft2(V& a, V& b) {
  V t = a;
  a = cadd(t, b);
  b = csub(t, b);
}
template <int DIR, typename V>
__device__ __forceinline__ void dft3(V& a, V& b, V& c) {
  using T = typename Scal<V>::T;
  // forward: w = exp(-2 pi i/3); backward conjugate
  const T s3 = (T)0.86602540378443864676 * (DIR < 0 ? (T)1 : (T)-1);
  V t1 = cadd(b, c);
  V t2 = mkv<V>(a.x - (T)0.5 * t1.x, a.y - (T)0.5 * t1.y);
  V d = csub(b, c);
  V t3 = mkv<V>(s3 * d.y, -s3 * d.x);
  a = cadd(a, t1);
  b = cadd(t2, t3);
  c = csub(t2, t3);
}
template <int DIR, typename V>
__device__ __forceinline__ void dft4(V& a, V& b, V& c, V& d) {
  V t0 = cadd(a, c), t1 = csub(a, c);
  V t2 = cadd(b, d), t3 = csub(b, d);
  // forward: y1 = t1 - i t3, y3 = t1 + i t3
  V it3 = (DIR < 0) ? mkv<V>(t3.y, -t3.x) : mkv<V>(-t3.y, t3.x);
  a = cadd(t0, t2);
  c = csub(t0, t2);
  b = cadd(t1, it3);
  d = csub(t1, it3);
}
// 5-point DFT from its definition, with c1 = cos(2 pi/5), c2 = cos(4 pi/5),
// s1 = sin(2 pi/5), s2 = sin(4 pi/5): with t1 = x1 + x4, t2 = x2 + x3,
// t3 = x1 - x4, t4 = x2 - x3, X_1,4 = x0 + c1 t1 + c2 t2 -/+ i (s1 t3 + s2 t4)
// and X_2,3 = x0 + c2 t1 + c1 t2 -/+ i (s2 t3 - s1 t4) for the forward sign
template <int DIR, typename V>
__device__ __forceinline__ void dft5(V& x0, V& x1, V& x2, V& x3, V& x4) {
  using T = typename Scal<V>::T;
  const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
  const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
  const V t1 = cadd(x1, x4), t2 = cadd(x2, x3), t3 = csub(x1, x4), t4 = csub(x2, x3);
  const V a1 = mkv<V>(x0.x + c1 * t1.x + c2 * t2.x, x0.y + c1 * t1.y + c2 * t2.y);
  const V a2 = mkv<V>(x0.x + c2 * t1.x + c1 * t2.x, x0.y + c2 * t1.y + c1 * t2.y);
  const V b1 = mkv<V>(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
  const V b2 = mkv<V>(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
  // i b = (-b.y, b.x); forward: X1 = a1 - i b1
  const T sg = (DIR < 0) ? (T)1 : (T)-1;
  const V ib1 = mkv<V>(-sg * b1.y, sg * b1.x), ib2 = mkv<V>(-sg * b2.y, sg * b2.x);
  x0 = cadd(x0, cadd(t1, t2));
  x1 = csub(a1, ib1);
  x4 = cadd(a1, ib1);
  x2 = csub(a2, ib2);
  x3 = cadd(a2, ib2);
}
// 6-point DFT, natural order: two 3-point DFTs of the even/odd samples
template <int DIR, typename V>
__device__ __forceinline__ void dft6(V& v0, V& v1, V& v2, V& v3, V& v4, V& v5) {
  using T = typename Scal<V>::T;
  const T sg = (DIR < 0) ? (T)-1 : (T)1;
  const T h = (T)0.5, s3 = (T)0.86602540378443864676 * sg;
  V e0 = v0, e1 = v2, e2 = v4, o0 = v1, o1 = v3, o2 = v5;
  dft3<DIR>(e0, e1, e2);
  dft3<DIR>(o0, o1, o2);
  const V w1o1 = cmul(o1, mkv<V>(h, s3));    // W6^1 o1
  const V w2o2 = cmul(o2, mkv<V>(-h, s3));   // W6^2 o2
  v0 = cadd(e0, o0);
  v3 = csub(e0, o0);
  v1 = cadd(e1, w1o1);
  v4 = csub(e1, w1o1);
  v2 = cadd(e2, w2o2);
  v5 = csub(e2, w2o2);
}
// 8-point DFT, natural order: two 4-point DFTs of the even/odd samples
template <int DIR, typename V>
__device__ __forceinline__ void dft8(V& v0, V& v1, V& v2, V& v3, V& v4, V& v5, V& v6, V& v7) {
  using T = typename Scal<V>::T;
  const T sg = (DIR < 0) ? (T)-1 : (T)1;
  const T r = (T)0.70710678118654752440;
  V e0 = v0, e1 = v2, e2 = v4, e3 = v6, o0 = v1, o1 = v3, o2 = v5, o3 = v7;
  dft4<DIR>(e0, e1, e2, e3);
  dft4<DIR>(o0, o1, o2, o3);
  const V w1 = cmul(o1, mkv<V>(r, sg * r));          // W8^1 o1
  const V w2 = mkv<V>(-sg * o2.y, sg * o2.x);        // W8^2 o2 = sg i o2
  const V w3 = cmul(o3, mkv<V>(-r, sg * r));         // W8^3 o3
  v0 = cadd(e0, o0);
  v4 = csub(e0, o0);
  v1 = cadd(e1, w1);
  v5 = csub(e1, w1);
  v2 = cadd(e2, w2);
  v6 = csub(e2, w2);
  v3 = cadd(e3, w3);
  v7 = csub(e3, w3);
}

// FFT buffers in LDS are padded by one element every 16 so that the
// stride-R writes of the Stockham passes do not pile onto the same banks.
__device__ __forceinline__ int fpad(int i) { return i + (i >> 4); }
template <int N> struct FFTBuf { static constexpr int len = N + N / 16 + 1; };

// One-wavefront Stockham FFT of length N held in LDS buffer x (scratch y of
// the same padded size FFTBuf<N>::len, indexed through fpad).  Unnormalised; DIR=-1 is FFTW_FORWARD (exp(-i..)),
// DIR=+1 is FFTW_BACKWARD.  Returns the buffer holding the result.
template <int N, int DIR, typename V, int MODEL = 0>
__device__ V* wave_fft(V* x, V* y, const V* __restrict__ tw, int lane) {
  constexpr FFTPlan P = make_plan(N, MODEL);
  int Ns = 1;
#pragma unroll
  for (int s = 0; s < P.nst; ++s) {
    const int R = P.r[s];
    const int nb = N / R;
    const int tstride = N / (Ns * R);
    for (int j = lane; j < nb; j += WAVE) {
      const int k = j % Ns;
      const int idxD = (j / Ns) * Ns * R + k;
      if (R == 8) {
        V v[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) v[m] = x[fpad(j + m * nb)];
        if (Ns > 1) {
#pragma unroll
          for (int m = 1; m < 8; ++m) v[m] = cmul(v[m], twid<DIR>(tw, m * k * tstride));
        }
        dft8<DIR>(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]);
#pragma unroll
        for (int m = 0; m < 8; ++m) y[fpad(idxD + m * Ns)] = v[m];
      } else if (R == 6) {
        V v[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) v[m] = x[fpad(j + m * nb)];
        if (Ns > 1) {
#pragma unroll
          for (int m = 1; m < 6; ++m) v[m] = cmul(v[m], twid<DIR>(tw, m * k * tstride));
        }
        dft6<DIR>(v[0], v[1], v[2], v[3], v[4], v[5]);
#pragma unroll
        for (int m = 0; m < 6; ++m) y[fpad(idxD + m * Ns)] = v[m];
      } else if (R == 5) {
        V v[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) v[m] = x[fpad(j + m * nb)];
        if (Ns > 1) {
#pragma unroll
          for (int m = 1; m < 5; ++m) v[m] = cmul(v[m], twid<DIR>(tw, m * k * tstride));
        }
        dft5<DIR>(v[0], v[1], v[2], v[3], v[4]);
#pragma unroll
        for (int m = 0; m < 5; ++m) y[fpad(idxD + m * Ns)] = v[m];
      } else if (R == 4) {
        V v0 = x[fpad(j)], v1 = x[fpad(j + nb)], v2 = x[fpad(j + 2 * nb)], v3 = x[fpad(j + 3 * nb)];
        if (Ns > 1) {
          v1 = cmul(v1, twid<DIR>(tw, k * tstride));
          v2 = cmul(v2, twid<DIR>(tw, 2 * k * tstride));
          v3 = cmul(v3, twid<DIR>(tw, 3 * k * tstride));
        }
        dft4<DIR>(v0, v1, v2, v3);
        y[fpad(idxD)] = v0; y[fpad(idxD + Ns)] = v1; y[fpad(idxD + 2 * Ns)] = v2; y[fpad(idxD + 3 * Ns)] = v3;
      } else if (R == 3) {
        V v0 = x[fpad(j)], v1 = x[fpad(j + nb)], v2 = x[fpad(j + 2 * nb)];
        if (Ns > 1) {
          v1 = cmul(v1, twid<DIR>(tw, k * tstride));
          v2 = cmul(v2, twid<DIR>(tw, 2 * k * tstride));
        }
        dft3<DIR>(v0, v1, v2);
        y[fpad(idxD)] = v0; y[fpad(idxD + Ns)] = v1; y[fpad(idxD + 2 * Ns)] = v2;
      } else {
        V v0 = x[fpad(j)], v1 = x[fpad(j + nb)];
        if (Ns > 1) v1 = cmul(v1, twid<DIR>(tw, k * tstride));
        dft2<DIR>(v0, v1);
        y[fpad(idxD)] = v0; y[fpad(idxD + Ns)] = v1;
      }
    }
    wave_sync();
    V* t = x; x = y; y = t;
    Ns *= R;
  }
  return x;
}

// radix-R DFT of v[0..R-1] in place (natural order)
template <int R, int DIR, typename V>
__device__ __forceinline__ void dft_r(V (&v)[8]) {
  if constexpr (R == 8) dft8<DIR>(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]);
  else if constexpr (R == 6) dft6<DIR>(v[0], v[1], v[2], v[3], v[4], v[5]);
  else if constexpr (R == 5) dft5<DIR>(v[0], v[1], v[2], v[3], v[4]);
  else if constexpr (R == 4) dft4<DIR>(v[0], v[1], v[2], v[3]);
  else if constexpr (R == 3) dft3<DIR>(v[0], v[1], v[2]);
  else dft2<DIR>(v[0], v[1]);
}
__host__ __device__ constexpr FFTPlan reverse_plan(FFTPlan p) {
  FFTPlan q{p.n, p.nst, {0}};
  for (int i = 0; i < p.nst; ++i) q.r[i] = p.r[p.nst - 1 - i];
  return q;
}
// One stage s of the Stockham FFT of plan PL (the radix is a compile-time
// constant), with its input taken from registers (FIRST: lane j holds
// x[j + m N/R], one round) or its output left in registers (LAST: lane j
// keeps y[j + m N/R], one round).  TWC: the twiddles w^2 .. w^(R-1) of a
// butterfly are powers of the one read from the table (fewer LDS loads).
template <int N, int DIR, typename V, int R, int NS, bool FIRST, bool LAST, bool TWC>
__device__ __forceinline__ void fft_stage(V* x, V* y, const V* __restrict__ tw, int lane, const V (&in)[8],
                                          V (&out)[8]) {
  constexpr int nb = N / R;
  constexpr int tstride = N / (NS * R);
  static_assert(!(FIRST || LAST) || nb == WAVE, "register stages need N / R == 64");
  for (int j = lane; j < nb; j += WAVE) {
    const int k = j % NS;
    const int idxD = (j / NS) * NS * R + k;
    V v[8];
#pragma unroll
    for (int m = 0; m < R; ++m) v[m] = FIRST ? in[m] : x[fpad(j + m * nb)];
    if constexpr (NS > 1) {
      if constexpr (TWC) {
        const V w1 = twid<DIR>(tw, k * tstride);
        V wm = w1;
#pragma unroll
        for (int m = 1; m < R; ++m) {
          v[m] = cmul(v[m], wm);
          if (m + 1 < R) wm = cmul(wm, w1);
        }
      } else {
#pragma unroll
        for (int m = 1; m < R; ++m) v[m] = cmul(v[m], twid<DIR>(tw, m * k * tstride));
      }
    }
    dft_r<R, DIR, V>(v);
#pragma unroll
    for (int m = 0; m < R; ++m) {
      if (LAST) out[m] = v[m];
      else y[fpad(idxD + m * NS)] = v[m];
    }
  }
}
template <int N, int DIR, typename V, int MODEL, bool REV, bool FIRST, bool LAST, bool TWC, int S, int NS>
__device__ __forceinline__ V* fft_stages(V* x, V* y, const V* __restrict__ tw, int lane, const V (&in)[8],
                                         V (&out)[8]) {
  constexpr FFTPlan P = REV ? reverse_plan(make_plan(N, MODEL)) : make_plan(N, MODEL);
  if constexpr (S == P.nst) {
    return x;
  } else {
    constexpr int R = P.r[S];
    constexpr bool last = LAST && S == P.nst - 1;
    fft_stage<N, DIR, V, R, NS, FIRST && S == 0, last, TWC>(x, y, tw, lane, in, out);
    if constexpr (last) return y;
    wave_sync();
    return fft_stages<N, DIR, V, MODEL, REV, FIRST, LAST, TWC, S + 1, NS * R>(y, x, tw, lane, in, out);
  }
}
// The one-wavefront FFT of wave_fft with the plan optionally reversed, the
// first stage optionally fed from registers and the last stage's outputs
// optionally kept in registers (then nothing is written by it and the
// returned buffer is not the result).
template <int N, int DIR, typename V, int MODEL, bool REV, bool FIRST, bool LAST, bool TWC = false>
__device__ __forceinline__ V* wave_fft_ex(V* x, V* y, const V* __restrict__ tw, int lane, const V (&in)[8],
                                          V (&out)[8]) {
  return fft_stages<N, DIR, V, MODEL, REV, FIRST, LAST, TWC, 0, 1>(x, y, tw, lane, in, out);
}
// first / last radix of a plan
template <int N, int MODEL>
constexpr int plan_last_radix() { return make_plan(N, MODEL).r[make_plan(N, MODEL).nst - 1]; }

// Swirl strength (calcswirl + eig, helpers.f90:197-319, 360-404): the
// reference evaluates the three eigenvalues of the velocity-gradient tensor
// with Cardano's closed form in complex arithmetic and returns
// sqrt(sum Im(lambda)^2), 0 where that is 0 or NaN.  For a real 3x3 matrix
// with one real root and a complex pair that is sqrt(2) |Im| of the pair.
// Here: characteristic polynomial t^3 + p t + q of the deviatoric part,
// discriminant D = (q/2)^2 + (p/3)^3; D > 0: Im of the pair is
// (sqrt(3)/2) |u - v| with u = cbrt(-q/2 - sign(q) sqrt(D)) and
// v = -p/(3u) (no cancellation, unlike the reference's complex form, whose
// S = (B + sqrt(B^2 - A^3))^(1/3) loses digits when B < 0), so
// swirl = sqrt(3/2) |u - v|; D <= 0 (three real roots): 0.
__device__ __forceinline__ double swirl_of(const double (&a)[9]) {
  const double tr = a[0] + a[4] + a[8];
  // sum of principal 2x2 minors and determinant
  const double m2 = a[0] * a[4] - a[1] * a[3] + a[0] * a[8] - a[2] * a[6] + a[4] * a[8] - a[5] * a[7];
  const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                     a[2] * (a[3] * a[7] - a[4] * a[6]);
  // lambda^3 - tr lambda^2 + m2 lambda - det = 0, lambda = t + tr/3
  const double pp = m2 - tr * tr / 3.0;
  const double qq = -2.0 * tr * tr * tr / 27.0 + tr * m2 / 3.0 - det;
  const double disc = 0.25 * qq * qq + pp * pp * pp / 27.0;
  if (!(disc > 0.0)) return 0.0;
  // the cube root whose argument does not cancel, then v = -p/(3u)
  const double sd = sqrt(disc);
  const double uu = cbrt(-0.5 * qq + (qq > 0.0 ? -sd : sd));
  const double vv = -pp / (3.0 * uu);
  const double sw = 1.2247448713915890491 * fabs(uu - vv);  // sqrt(3/2)
  return (sw > 0.0 && isfinite(sw)) ? sw : 0.0;
}

}  // namespace kmm
