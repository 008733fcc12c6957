// Device-side primitives shared by the pencil (y) and plane (x/z) kernels of
// the KMM channel step: complex FP64 arithmetic, wave64 scans/reductions, and
// an LDS Stockham FFT for lengths 2^a 3^b.  Written for gfx950 (CDNA4):
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
__device__ __forceinline__ double shdn(double v, int off) { return __shfl_down(v, off, WAVE); }
__device__ __forceinline__ double shup(double v, int off) { return __shfl_up(v, off, WAVE); }
__device__ __forceinline__ double2 shdn(double2 v, int off) {
  return make_double2(__shfl_down(v.x, off, WAVE), __shfl_down(v.y, off, WAVE));
}
__device__ __forceinline__ double2 shup(double2 v, int off) {
  return make_double2(__shfl_up(v.x, off, WAVE), __shfl_up(v.y, off, WAVE));
}
__device__ __forceinline__ double bcast(double v, int src) { return __shfl(v, src, WAVE); }
__device__ __forceinline__ double2 bcast(double2 v, int src) {
  return make_double2(__shfl(v.x, src, WAVE), __shfl(v.y, src, WAVE));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, WAVE);
  return v;
}
__device__ __forceinline__ double2 wave_sum(double2 v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    v.x += __shfl_xor(v.x, off, WAVE);
    v.y += __shfl_xor(v.y, off, WAVE);
  }
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, WAVE));
  return v;
}

// inclusive suffix sum: r_l = sum_{l' >= l} v_l'
__device__ __forceinline__ double2 suffix_sum(double2 v, int lane) {
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    double2 t = shdn(v, off);
    if (lane + off < WAVE) v = cadd(v, t);
  }
  return v;
}

// Suffix affine recurrence y_l = w_l + c_l * y_{l+1} over all 64 lanes
// (y_64 = 0).  A lane with c_l = 0 starts a chain.
__device__ __forceinline__ double2 suffix_affine(double c, double2 w, int lane) {
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    double c2 = shdn(c, off);
    double2 w2 = shdn(w, off);
    if (lane + off < WAVE) {
      w = cfma(c, w2, w);
      c = c * c2;
    }
  }
  return w;
}
__device__ __forceinline__ double suffix_affine_r(double c, double w, int lane) {
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    double c2 = shdn(c, off);
    double w2 = shdn(w, off);
    if (lane + off < WAVE) {
      w = fma(c, w2, w);
      c = c * c2;
    }
  }
  return w;
}
// Prefix affine recurrence y_l = a_l + b_l * y_{l-1} (y_{-1} = 0).
__device__ __forceinline__ double2 prefix_affine(double b, double2 a, int lane) {
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    double b2 = shup(b, off);
    double2 a2 = shup(a, off);
    if (lane >= off) {
      a = cfma(b, a2, a);
      b = b * b2;
    }
  }
  return a;
}

// ------------------------------------------------------------------ FFT --
// Radix plan for N = 2^a 3^b: radix-3 stages first, then radix-4, one
// radix-2 if a is odd.
struct FFTPlan {
  int n;
  int nst;
  int r[16];
};
__host__ __device__ constexpr FFTPlan make_plan(int n) {
  FFTPlan p{n, 0, {0}};
  int m = n;
  while (m % 3 == 0) { p.r[p.nst++] = 3; m /= 3; }
  while (m % 4 == 0) { p.r[p.nst++] = 4; m /= 4; }
  while (m % 2 == 0) { p.r[p.nst++] = 2; m /= 2; }
  return p;
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

// One-wavefront Stockham FFT of length N held in LDS buffer x (scratch y of
// the same size).  Unnormalised; DIR=-1 is FFTW_FORWARD (exp(-i..)),
// DIR=+1 is FFTW_BACKWARD.  Returns the buffer holding the result.
template <int N, int DIR, typename V>
__device__ V* wave_fft(V* x, V* y, const V* __restrict__ tw, int lane) {
  constexpr FFTPlan P = make_plan(N);
  int Ns = 1;
#pragma unroll
  for (int s = 0; s < P.nst; ++s) {
    const int R = P.r[s];
    const int nb = N / R;
    const int tstride = N / (Ns * R);
    for (int j = lane; j < nb; j += WAVE) {
      const int k = j % Ns;
      const int idxD = (j / Ns) * Ns * R + k;
      if (R == 4) {
        V v0 = x[j], v1 = x[j + nb], v2 = x[j + 2 * nb], v3 = x[j + 3 * nb];
        if (Ns > 1) {
          v1 = cmul(v1, twid<DIR>(tw, k * tstride));
          v2 = cmul(v2, twid<DIR>(tw, 2 * k * tstride));
          v3 = cmul(v3, twid<DIR>(tw, 3 * k * tstride));
        }
        dft4<DIR>(v0, v1, v2, v3);
        y[idxD] = v0; y[idxD + Ns] = v1; y[idxD + 2 * Ns] = v2; y[idxD + 3 * Ns] = v3;
      } else if (R == 3) {
        V v0 = x[j], v1 = x[j + nb], v2 = x[j + 2 * nb];
        if (Ns > 1) {
          v1 = cmul(v1, twid<DIR>(tw, k * tstride));
          v2 = cmul(v2, twid<DIR>(tw, 2 * k * tstride));
        }
        dft3<DIR>(v0, v1, v2);
        y[idxD] = v0; y[idxD + Ns] = v1; y[idxD + 2 * Ns] = v2;
      } else {
        V v0 = x[j], v1 = x[j + nb];
        if (Ns > 1) v1 = cmul(v1, twid<DIR>(tw, k * tstride));
        dft2<DIR>(v0, v1);
        y[idxD] = v0; y[idxD + Ns] = v1;
      }
    }
    wave_sync();
    V* t = x; x = y; y = t;
    Ns *= R;
  }
  return x;
}

}  // namespace kmm
