// Pencil stage of the FENE-P conformation tensor (POLYMER build of the
// reference), one wavefront per (kx,kz) pencil, the six independent
// components c11, c12, c13, c22, c23, c33 one after another:
//
//   part B  forward Chebyshev transform of the plane stage's c_n and str_n
//           (solvers.f90:3279-3416); the polymer force
//           t = div(qp), qp = zbeta1 str_n (polyforce, helpers.f90:498-560,
//           before smoothing as in main_dns.f90:674-680); the n -> n-1 copy
//           that the reference makes before smoothing (main_dns.f90:682-703);
//           smoothing (main_dns.f90:726-737)
//   init    C = diag(c11z, c22z, c33z) (main_dns.f90:206-236; the constant's
//           only Chebyshev-Fourier coefficient is the (0,0,0) one)
//   solve   polyNL (helpers.f90:610-674), phirhs with fn = cNL,
//           fnm1 = 2 cNL (main_dns.f90:249-261), n -> n-1 copy (:263-281)
//           and the Neumann penta solve with g = re diffpoly/(dt theta)
//           (:286-304)
//   evaluate  C and dC/dy to physical y for the plane stage (derivscji's
//           y-derivatives, derivatives.f90:148-254; the x and z derivatives
//           are taken in the plane stage).
//
// The conformation tensor and its histories (C, CN, CNM1, STRN, STRNM1) are
// stored pencil-major, [pencil][y] (the reference's order of a spectral
// field): a wave reads and writes its pencil as contiguous memory.  The
// fields shared with the plane stage (the physical-y NLP and TP fields, the
// force TF) keep the [y][pencil] layout of the transforms.
//
// MODE 0: step, MODE 1: startup (evaluate only), MODE 2: finish (part B only),
// MODE 3: C from the physical-y fields in NLP[0..5] (a conformation tensor
// rebuilt on the 3/2 grid and forward transformed in x and z), MODE 4: part B
// of a learned-closure stress (the force only: a closure stress is neither
// smoothed nor stored, and there is no conformation to transport), MODE 5:
// the step after a separate MODE 2 pass (C init, solve and evaluate, reading
// the smoothed c_n and str_n that pass stored; at ny = 192 the two passes
// hold fewer registers than the fused step).
#include "kmm_kernels.h"
#include "kmm_pencil.h"

namespace kmm {

// Pencils (wavefronts) per workgroup. At ny = 192 the two-segment solve needs
// more than 256 registers a lane: a 4-wave workgroup lets the compiler hold
// them in the unified VGPR/AGPR file at one wave per SIMD instead of spilling
// about 300 bytes a lane to scratch at two (4.5 % faster in FP64).
template <int NY> constexpr int poly_pw() { return NY > 128 ? 4 : PW; }

template <int NY, int MODE, typename V>
__global__ __launch_bounds__(poly_pw<NY>() * 64) void k_poly_pencil(PolyArgs A) {
  using PF = PFN<NY>;
  constexpr bool M0 = (MODE == 0 || MODE == 5);  // mode 5: mode 0 without part B
  extern __shared__ __align__(16) unsigned char smraw[];
  V* smem = reinterpret_cast<V*>(smraw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int p = blockIdx.x * poly_pw<NY>() + wv;
  const V* twl = lds_twiddles<NY, V>(smem, static_cast<const V*>(A.twM));
  if (p >= A.P) return;
  V* buf = wave_lds<NY, V>(smem, wv);
  V* scr = buf + FFTBuf<2 * NY>::len;
  const size_t P = (size_t)A.P;
  double kx, kz;
  int kxi, kzi;
  pencil_k(A, p, kx, kz, kxi, kzi);
  const double k2 = kx * kx + kz * kz;
  const bool is00 = (kxi == 0 && kzi == 0);
  const double dy = A.dyde;

  if (MODE == 3) {
#pragma unroll 1
    for (int c = 0; c < 6; ++c) {
      PF C = pf_zero<NY>();
      if (kzi != A.nzh) {
        C = phys_to_cheb_load<NY, V>(static_cast<const V*>(A.NLP[c]), P, p, buf, scr, twl, lane);
        if (is00) C = pf_real(C);
      }
      pf_store_pm<NY>(A.C[c], p, lane, C);
    }
    return;
  }
  if (MODE == 4) {
    // each component's load is issued before the previous one's transform
    PF t1 = pf_zero<NY>(), t2 = pf_zero<NY>(), t3 = pf_zero<NY>();
    if (kzi != A.nzh) {
      PhysN<NY, V> q = phys_load<NY, V>(static_cast<const V*>(A.NLP[6]), P, p, lane);
#pragma unroll 1
      for (int c = 0; c < 6; ++c) {
        const PhysN<NY, V> cur = q;
        if (c < 5) q = phys_load<NY, V>(static_cast<const V*>(A.NLP[7 + c]), P, p, lane);
        PF sn = phys_to_cheb<NY, V>(cur, buf, scr, twl, lane);
        if (is00) sn = pf_real(sn);
        const PF f = pf_scale(sn, A.zbeta1);
        switch (c) {
          case 0: t1 = pf_add(t1, pf_imul(f, kx)); break;
          case 1: t1 = pf_add(t1, cderiv<NY>(f, dy, lane)); t2 = pf_add(t2, pf_imul(f, kx)); break;
          case 2: t1 = pf_add(t1, pf_imul(f, kz)); t3 = pf_add(t3, pf_imul(f, kx)); break;
          case 3: t2 = pf_add(t2, cderiv<NY>(f, dy, lane)); break;
          case 4: t2 = pf_add(t2, pf_imul(f, kz)); t3 = pf_add(t3, cderiv<NY>(f, dy, lane)); break;
          default: t3 = pf_add(t3, pf_imul(f, kz)); break;
        }
      }
      if (is00) { t1 = pf_real(t1); t2 = pf_real(t2); t3 = pf_real(t3); }
    }
    pf_store<NY>(A.TF[0], P, p, lane, t1);
    pf_store<NY>(A.TF[1], P, p, lane, t2);
    pf_store<NY>(A.TF[2], P, p, lane, t3);
    return;
  }
  if (kzi == A.nzh) {  // norm zeroes the z-Nyquist row (helpers.f90:149-156)
    const PF z = pf_zero<NY>();
    for (int c = 0; c < 6; ++c) {
      if (M0 || MODE == 2) {
        if (A.partB || MODE == 2) {
          if (!A.closure) pf_store_pm<NY>(A.CN[c], p, lane, z);
          if (!A.closure || MODE == 2) pf_store_pm<NY>(A.STRN[c], p, lane, z);
        }
        if (A.copy || A.solve) { pf_store_pm<NY>(A.CNM1[c], p, lane, z); pf_store_pm<NY>(A.STRNM1[c], p, lane, z); }
      }
      if (M0 && (A.init || A.solve)) pf_store_pm<NY>(A.C[c], p, lane, z);
      if (MODE != 2 && A.evaluate)
        for (int h = 0; h < 2; ++h) {
          pf_store_zero<NY, V>(static_cast<V*>(A.TP[c + 6 * h]), P, p, lane);
        }
    }
    if (A.force && (MODE == 2 || A.partB))
      for (int t = 0; t < 3; ++t) pf_store<NY>(A.TF[t], P, p, lane, z);
    return;
  }

  const double sz = A.smooth ? A.smz[kzi] : 1.0;
  const double ti = 1.0 / A.tpoly;
  PF t1 = pf_zero<NY>(), t2 = pf_zero<NY>(), t3 = pf_zero<NY>();
#pragma unroll 1
  for (int c = 0; c < 6; ++c) {
    // component c: 11, 12, 13, 22, 23, 33
    PF cn, sn;
    PF cnm1_copy, snm1_copy;
    const bool haveB = (MODE == 2) || (MODE == 0 && A.partB);
    if (haveB) {
      cn = A.closure ? pf_zero<NY>() : phys_to_cheb_load<NY, V>(static_cast<const V*>(A.NLP[c]), P, p, buf, scr, twl, lane);
      sn = phys_to_cheb_load<NY, V>(static_cast<const V*>(A.NLP[6 + c]), P, p, buf, scr, twl, lane);
      if (is00) { cn = pf_real(cn); sn = pf_real(sn); }
      if (A.force) {
        // t_i = d q_ik / d x_k with q = zbeta1 str (symmetric)
        const PF q = pf_scale(sn, A.zbeta1);
        switch (c) {
          case 0: t1 = pf_add(t1, pf_imul(q, kx)); break;
          case 1: t1 = pf_add(t1, cderiv<NY>(q, dy, lane)); t2 = pf_add(t2, pf_imul(q, kx)); break;
          case 2: t1 = pf_add(t1, pf_imul(q, kz)); t3 = pf_add(t3, pf_imul(q, kx)); break;
          case 3: t2 = pf_add(t2, cderiv<NY>(q, dy, lane)); break;
          case 4: t2 = pf_add(t2, pf_imul(q, kz)); t3 = pf_add(t3, cderiv<NY>(q, dy, lane)); break;
          default: t3 = pf_add(t3, pf_imul(q, kz)); break;
        }
      }
      if (A.copy) {
        cnm1_copy = cn;
        snm1_copy = sn;
        pf_store_pm<NY>(A.CNM1[c], p, lane, cn);
        pf_store_pm<NY>(A.STRNM1[c], p, lane, sn);
      }
      // a closure stress feeds only the force: in a step its smoothed copy
      // in STRN has no reader (an explicit finish still stores it)
      const bool keep = !A.closure || MODE == 2;
      if (A.smooth && keep) {
        cn = smooth_apply<NY>(cn, A.smy, sz, lane);
        sn = smooth_apply<NY>(sn, A.smy, sz, lane);
      }
      if (!A.closure) pf_store_pm<NY>(A.CN[c], p, lane, cn);
      if (keep) pf_store_pm<NY>(A.STRN[c], p, lane, sn);
      if (MODE == 2) continue;
    }
    PF C;
    if (M0) {
      if (A.init) {
        C = pf_zero<NY>();
        if (is00 && (c == 0 || c == 3 || c == 5)) {
          const double v0 = A.c0[c == 0 ? 0 : (c == 3 ? 1 : 2)];
          if (lane == 0) C.a[0] = make_double2(v0, 0.0);
        }
      } else {
        C = pf_load_pm<NY>(A.C[c], p, lane);
      }
      if (A.solve) {
        if (!haveB) {
          cn = pf_load_pm<NY>(A.CN[c], p, lane);
          sn = pf_load_pm<NY>(A.STRN[c], p, lane);
        }
        const PF cm = A.copy && haveB ? cnm1_copy : pf_load_pm<NY>(A.CNM1[c], p, lane);
        const PF sm = A.copy && haveB ? snm1_copy : pf_load_pm<NY>(A.STRNM1[c], p, lane);
        // cNL = 3 (c_n - str_n/lambda) - (c_nm1 - str_nm1/lambda)
        const PF nl = pf_sub(pf_scale(pf_axpy(-ti, sn, cn), 3.0), pf_axpy(-ti, sm, cm));
        pf_store_pm<NY>(A.CNM1[c], p, lane, cn);
        pf_store_pm<NY>(A.STRNM1[c], p, lane, sn);
        const PF d1 = cderiv<NY>(C, dy, lane);
        const PF d2 = cderiv<NY>(d1, dy, lane);
        const PF rhs = phirhs<NY>(C, d2, nl, pf_scale(nl, 2.0), k2, A.re, A.diffpoly, A.dt, A.theta);
        // Neumann rows at both walls (at = ab = 0, bt = bb = 1, zero data)
        auto tt = [dy](int n) -> double { return dy * (double)n * (double)n; };
        auto tb = [dy](int n) -> double { double v = dy * (double)n * (double)n; return (n & 1) ? -v : v; };
        C = tau_penta<NY>(rhs, k2 + A.gC, dy, A.tabC + (size_t)p * tab_width(NY), tt, tb, czero(), czero(), lane);
        if (is00) C = pf_real(C);
      }
      if (A.init || A.solve) pf_store_pm<NY>(A.C[c], p, lane, C);
    } else {
      C = pf_load_pm<NY>(A.C[c], p, lane);
    }
    if (A.evaluate) {
      cheb_to_phys_store<NY, V>(C, buf, scr, twl, lane, static_cast<V*>(A.TP[c]), P, p);
      cheb_to_phys_store<NY, V>(cderiv<NY>(C, dy, lane), buf, scr, twl, lane, static_cast<V*>(A.TP[6 + c]), P, p);
    }
  }
  if (A.force && ((MODE == 0 && A.partB) || MODE == 2)) {
    if (is00) { t1 = pf_real(t1); t2 = pf_real(t2); t3 = pf_real(t3); }
    pf_store<NY>(A.TF[0], P, p, lane, t1);
    pf_store<NY>(A.TF[1], P, p, lane, t2);
    pf_store<NY>(A.TF[2], P, p, lane, t3);
  }
}

template <int MODE, typename V>
static hipError_t launch_poly_v(int ny, const PolyArgs& a, hipStream_t s) {
  KMM_NY_DISPATCH(ny, constexpr int W = poly_pw<NY>();
                  hipLaunchKernelGGL((k_poly_pencil<NY, MODE, V>), dim3((a.P + W - 1) / W), dim3(W * 64),
                                     (lds_bytes<NY, V>(W)), s, a));
  return hipGetLastError();
}

hipError_t launch_poly_pencil(int ny, int mode, const PolyArgs& a, bool mixed, hipStream_t s) {
  switch (mode) {
    case 0: return mixed ? launch_poly_v<0, float2>(ny, a, s) : launch_poly_v<0, double2>(ny, a, s);
    case 1: return mixed ? launch_poly_v<1, float2>(ny, a, s) : launch_poly_v<1, double2>(ny, a, s);
    case 2: return mixed ? launch_poly_v<2, float2>(ny, a, s) : launch_poly_v<2, double2>(ny, a, s);
    case 3: return mixed ? launch_poly_v<3, float2>(ny, a, s) : launch_poly_v<3, double2>(ny, a, s);
    case 4: return mixed ? launch_poly_v<4, float2>(ny, a, s) : launch_poly_v<4, double2>(ny, a, s);
    case 5: return mixed ? launch_poly_v<5, float2>(ny, a, s) : launch_poly_v<5, double2>(ny, a, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace kmm
