// Per-routine entry kernels: the reference's spectral helper procedures on
// whole fields, one wavefront per (kx,kz) pencil, for the C ABI entries that
// keep the reference's argument meaning (kmm_phirhs, kmm_nonlin, kmm_veloc,
// kmm_vort, kmm_gradscl, kmm_derivscji, kmm_polyforce, kmm_subforce,
// kmm_polyNL, kmm_norm, kmm_uzero).  The time step itself runs these
// operations fused inside k_pencil / k_poly_pencil; these kernels reuse the
// same device routines (cderiv, phirhs, tau_penta) one operation at a time.
#include "kmm_kernels.h"
#include "kmm_pencil.h"

namespace kmm {

template <int NY>
__global__ __launch_bounds__(PW * 64) void k_routine(RoutineArgs A) {
  using PF = PFN<NY>;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int p = blockIdx.x * PW + wv;
  if (p >= A.P) return;
  const size_t P = (size_t)A.P;
  double kx, kz;
  int kxi, kzi;
  pencil_k(A, p, kx, kz, kxi, kzi);
  const double k2 = kx * kx + kz * kz;
  const double dy = A.dyde;
  auto ld = [&](int i) { return pf_load<NY>(A.in[i], P, p, lane); };
  auto st = [&](int o, const PF& f) { pf_store<NY>(A.out[o], P, p, lane, f); };
  switch (A.op) {
    case RT_PHIRHS: {  // solvers.f90:555-676 (theta <= 0.99): RHS left in v
      const PF v = ld(0);
      const PF d2 = cderiv<NY>(cderiv<NY>(v, dy, lane), dy, lane);
      st(0, phirhs<NY>(v, d2, ld(1), ld(2), k2, A.re, A.par, A.dt, A.theta));
      break;
    }
    case RT_NONLIN: {  // solvers.f90:1497-1572: in fn = H2, omz = H3, gn = H1
      const PF h2 = ld(0), h3 = ld(1), h1 = ld(2);
      const PF x = pf_add(pf_imul(h3, kz), pf_imul(h1, kx));
      st(0, pf_sub(pf_scale(h2, -k2), cderiv<NY>(x, dy, lane)));
      st(1, pf_sub(pf_imul(h1, kz), pf_imul(h3, kx)));
      break;
    }
    case RT_VELOC: {  // solvers.f90:1576-1659: u, w from continuity and omega_y
      PF v = ld(0);
      const PF omy = ld(1);
      PF u, w;
      if (kxi == 0 && kzi == 0) {
        u = pf_load_real<NY>(A.m0, lane);
        w = pf_load_real<NY>(A.m1, lane);
        v = pf_zero<NY>();
      } else {
        const PF dv = cderiv<NY>(v, dy, lane);
        const double ik2 = 1.0 / k2;
        u = pf_scale(pf_sub(pf_imul(dv, kx), pf_imul(omy, kz)), ik2);
        w = pf_scale(pf_add(pf_imul(dv, kz), pf_imul(omy, kx)), ik2);
      }
      st(0, u);
      st(1, w);
      st(2, v);
      break;
    }
    case RT_VORT: {  // solvers.f90:1663-1722: omx = Dw - i kz v, omz = i kx v - Du
      const PF u = ld(0), v = ld(1), w = ld(2);
      st(0, pf_sub(cderiv<NY>(w, dy, lane), pf_imul(v, kz)));
      st(1, pf_sub(pf_imul(v, kx), cderiv<NY>(u, dy, lane)));
      break;
    }
    case RT_GRAD: {  // gradscl (derivatives.f90:90-144), one component of derivscji (:148-254)
      const PF f = ld(0);
      st(0, pf_imul(f, kx));
      st(1, cderiv<NY>(f, dy, lane));
      st(2, pf_imul(f, kz));
      break;
    }
    case RT_POLYFORCE: {  // helpers.f90:498-560: t = div(s) of the symmetric s11..s33
      const PF s11 = ld(0), s12 = ld(1), s13 = ld(2), s22 = ld(3), s23 = ld(4), s33 = ld(5);
      st(0, pf_add(pf_add(pf_imul(s11, kx), pf_imul(s13, kz)), cderiv<NY>(s12, dy, lane)));
      st(1, pf_add(pf_add(pf_imul(s12, kx), pf_imul(s23, kz)), cderiv<NY>(s22, dy, lane)));
      st(2, pf_add(pf_add(pf_imul(s13, kx), pf_imul(s33, kz)), cderiv<NY>(s23, dy, lane)));
      break;
    }
    case RT_SUBFORCE:  // helpers.f90:564-606: gn += t1, fn += t2, omz += t3
      for (int q = 0; q < 3; ++q) st(q, pf_add(ld(q), ld(3 + q)));
      break;
    case RT_POLYNL: {  // helpers.f90:610-674, one component: 3 (cn - strn/l) - (cnm1 - strnm1/l)
      const double ti = A.par;
      const PF a = pf_axpy(-ti, ld(2), ld(0));
      const PF b = pf_axpy(-ti, ld(3), ld(1));
      st(0, pf_sub(pf_scale(a, 3.0), b));
      break;
    }
    case RT_NORM: {  // helpers.f90:91-158
      if (kzi == A.nzh) {
        st(0, pf_zero<NY>());
      } else if (kxi == 0 && kzi == 0) {
        st(0, pf_real(ld(0)));
      } else if (kxi == 0 && kzi < A.nzh) {
        // Hermitian pair kz, -kz of the kx = 0 plane: a(j) = (a(j) + conj a(-j))/2, a(-j) = conj a(j)
        const PF a = ld(0);
        const PF b = pf_load<NY>(A.in[0], P, (size_t)(A.nz - kzi), lane);
        PF h = a, c;
#pragma unroll
        for (int q = 0; q < PLay<NY>::S; ++q) {
          h.a[q] = cscale(cadd(a.a[q], cconj(b.a[q])), 0.5);
          h.b[q] = cscale(cadd(a.b[q], cconj(b.b[q])), 0.5);
          c.a[q] = cconj(h.a[q]);
          c.b[q] = cconj(h.b[q]);
        }
        h.e = cscale(cadd(a.e, cconj(b.e)), 0.5);
        c.e = cconj(h.e);
        st(0, h);
        pf_store<NY>(A.out[0], P, (size_t)(A.nz - kzi), lane, c);
      } else if (kxi != 0) {
        st(0, ld(0));
      }  // kx = 0, kz > nz/2: written by its partner
      break;
    }
    default:
      break;
  }
}

// uzero (solvers.f90:1065-1176): the mean (kx = kz = 0) mode's CN/AB2 Robin
// solve of psolve1d; one wavefront.  m0 = u0, m1 = h1, m2 = h1l (real
// Chebyshev profiles), result in m3.
template <int NY>
__global__ __launch_bounds__(64) void k_uzero(RoutineArgs A, const double2* __restrict__ tab) {
  using PF = PFN<NY>;
  const int lane = threadIdx.x;
  const double th = A.theta, dy = A.dyde;
  const PF u = pf_load_real<NY>(A.m0, lane);
  const PF h = pf_load_real<NY>(A.m1, lane);
  const PF hl = pf_load_real<NY>(A.m2, lane);
  const PF d2 = cderiv<NY>(cderiv<NY>(u, dy, lane), dy, lane);
  PF r = pf_scale(u, -A.re / (A.dt * th));
  r = pf_axpy(-3.0 * A.re / (2.0 * th), h, r);
  r = pf_axpy(A.re / (2.0 * th), hl, r);
  r = pf_axpy(-(1.0 - th) / th, d2, r);
  const double at = A.rb[0], bt = A.rb[1], gt = A.rb[2], ab = A.rb[3], bb = A.rb[4], gb = A.rb[5];
  auto tt = [=](int n) -> double { return at + dy * bt * (double)n * (double)n; };
  auto tb = [=](int n) -> double { double t = ab - dy * bb * (double)n * (double)n; return (n & 1) ? -t : t; };
  const PF un = pf_real(tau_penta<NY>(r, A.re / (th * A.dt), dy, tab, tt, tb, make_double2(gt, 0.0),
                                      make_double2(gb, 0.0), lane));
  pf_store_real<NY>(A.m3, lane, un);
}

hipError_t launch_routine(int ny, const RoutineArgs& a, hipStream_t s) {
  KMM_NY_DISPATCH(ny, hipLaunchKernelGGL((k_routine<NY>), dim3(nblk(a.P)), dim3(PW * 64), 0, s, a));
  return hipGetLastError();
}
hipError_t launch_uzero(int ny, const RoutineArgs& a, const double2* tab, hipStream_t s) {
  KMM_NY_DISPATCH(ny, hipLaunchKernelGGL((k_uzero<NY>), dim3(1), dim3(64), 0, s, a, tab));
  return hipGetLastError();
}

}  // namespace kmm
