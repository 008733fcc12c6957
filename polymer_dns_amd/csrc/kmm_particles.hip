// Lagrangian tracers on the 3/2 physical grid (FFTW_DNS/code/src/dns/part_track.f90).
//
//   k_tracer : one thread per particle.  Tracer mode (part != nullptr): with
//              update = 1 the explicit Euler position update of
//              tracer_update (part_track.f90:404-408), then the fluid velocity
//              at the (new) position becomes the particle velocity
//              (tracer_update :411, init_part :266).  Interpolation mode
//              (part == nullptr): fluid_interp1 / fluid_interp
//              (part_track.f90:437-613) of nf fields at query points.
//
// Cell search and trilinear weights follow fluid_interp (part_track.f90:
// 575-589) and interpolate3 (:1184-1222) operation by operation, with
// floating-point contraction off, so positions, cell indices and weights
// round exactly as the reference's separate multiplies and adds do (acos is
// the only library function on the path).
//
//   k_inertial : one thread per inertial particle, the Maxey-Riley equations
//              by RK4 (part_track.f90:617-1075).
//
// A particle whose cell falls outside the grid (the reference would index out
// of bounds or stop, part_track.f90:582-590) raises *err; its indices are
// clamped so every access stays inside the field.
#include "kmm_kernels.h"
#include "kmm_device.h"

namespace kmm {

__device__ __forceinline__ double interp3(double x, double y, double z, double x0, double x1, double y0,
                                          double y1, double z0, double z1, double c000, double c100,
                                          double c001, double c101, double c010, double c110, double c011,
                                          double c111) {
#pragma clang fp contract(off)
  const double xd = (x1 > x0) ? (x - x0) / (x1 - x0) : 0.0;
  const double yd = (y1 > y0) ? (y - y0) / (y1 - y0) : 0.0;
  const double zd = (z1 > z0) ? (z - z0) / (z1 - z0) : 0.0;
  const double c00 = c000 * (1.0 - xd) + c100 * xd;
  const double c01 = c001 * (1.0 - xd) + c101 * xd;
  const double c10 = c010 * (1.0 - xd) + c110 * xd;
  const double c11 = c011 * (1.0 - xd) + c111 * xd;
  const double c0 = c00 * (1.0 - yd) + c10 * yd;
  const double c1 = c01 * (1.0 - yd) + c11 * yd;
  return c0 * (1.0 - zd) + c1 * zd;
}

// cell search, fluid_interp part_track.f90:575-580 (0-based indices here);
// returns true (with clamped indices) when the point lies outside the grid
__device__ __forceinline__ bool find_cell(const TracerArgs& A, double xp, double yp, double zp, int& imin,
                                          int& jmin, int& kmax) {
#pragma clang fp contract(off)
  const double fx = floor(xp / A.delxm), fz = floor(zp / A.delzm);
  const double fk = floor(1.0 + (double)A.ny / A.pi * acos(2.0 * yp / A.yl));
  bool bad = !(fx > -1e9 && fx < 1e9 && fz > -1e9 && fz < 1e9 && fk >= 1.0 && fk <= (double)A.ny);
  imin = 0;
  jmin = 0;
  kmax = 0;
  if (!bad) {
    imin = ((int)fx + A.mx) % A.mx;
    jmin = ((int)fz + A.mz) % A.mz;
    kmax = (int)fk - 1;
    bad = imin < 0 || jmin < 0;
  }
  if (bad) {
    imin = min(max(imin, 0), A.mx - 1);
    jmin = min(max(jmin, 0), A.mz - 1);
    kmax = min(max(kmax, 0), A.ny - 1);
  }
  return bad;
}

// the eight corner offsets (fluid_interp's order: x fastest, then z, then y
// from kmin = kmax + 1 to kmax) and the cell bounds
struct Cell {
  long long o[8];
  double xmin, xmax, ymin, ymax, zmin, zmax;
};
__device__ __forceinline__ Cell cell_of(const TracerArgs& A, int imin, int jmin, int kmax) {
  Cell c;
  const int imax = (imin + 1) % A.mx, jmax = (jmin + 1) % A.mz, kmin = kmax + 1;
  c.xmin = (double)imin * A.delxm;
  c.xmax = (double)imax * A.delxm;
  c.ymin = A.ycoord[kmin];
  c.ymax = A.ycoord[kmax];
  c.zmin = (double)jmin * A.delzm;
  c.zmax = (double)jmax * A.delzm;
  const long long i0 = imin * A.sx, i1 = imax * A.sx, j0 = jmin * A.sz, j1 = jmax * A.sz;
  const long long k0 = kmin * A.sy, k1 = kmax * A.sy;
  c.o[0] = k0 + j0 + i0; c.o[1] = k0 + j0 + i1; c.o[2] = k0 + j1 + i0; c.o[3] = k0 + j1 + i1;
  c.o[4] = k1 + j0 + i0; c.o[5] = k1 + j0 + i1; c.o[6] = k1 + j1 + i0; c.o[7] = k1 + j1 + i1;
  return c;
}
__device__ __forceinline__ double at(const double* __restrict__ u, const Cell& c, double x, double y, double z) {
  return interp3(x, y, z, c.xmin, c.xmax, c.ymin, c.ymax, c.zmin, c.zmax, u[c.o[0]], u[c.o[1]], u[c.o[2]],
                 u[c.o[3]], u[c.o[4]], u[c.o[5]], u[c.o[6]], u[c.o[7]]);
}

__global__ __launch_bounds__(256) void k_tracer(TracerArgs A) {
#pragma clang fp contract(off)
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= A.n) return;
  const int n = A.n;
  double xp, yp, zp;
  if (A.part) {
    xp = A.part[j];
    yp = A.part[n + j];
    zp = A.part[2 * n + j];
    if (A.update) {
      // tracer_update, part_track.f90:402-408
      const double height = A.ycoord[0] - A.ap;
      const double up = A.part[3 * n + j], vp = A.part[4 * n + j], wp = A.part[5 * n + j];
      xp = fmod(A.xl + xp + up * A.dt, A.xl);
      yp = fmax(fmin(yp + vp * A.dt, height), A.ap + A.ycoord[A.ny]);
      zp = fmod(A.zl + zp + wp * A.dt, A.zl);
    }
  } else {
    xp = A.px[j];
    yp = A.py[j];
    zp = A.pz[j];
  }
  int imin, jmin, kmax;
  if (find_cell(A, xp, yp, zp, imin, jmin, kmax)) atomicOr(A.err, 1);
  if (A.cells) {
    A.cells[j] = imin;
    A.cells[n + j] = jmin;
    A.cells[2 * n + j] = kmax;
  }
  const Cell c = cell_of(A, imin, jmin, kmax);
  for (int f = 0; f < A.nf; ++f) {
    const double r = at(A.fld + f * A.fstride, c, xp, yp, zp);
    if (A.part) A.part[(3 + f) * n + j] = r;
    else A.out[(size_t)f * n + j] = r;
  }
  if (A.part) {
    A.part[j] = xp;
    A.part[n + j] = yp;
    A.part[2 * n + j] = zp;
  }
}

// trilinear weights of interpolate3 (part_track.f90:1184-1222) at a point,
// and the interpolation of one field with them (the same operations as
// interp3, whose weights do not depend on the field)
struct Wts {
  double xd, yd, zd;
};
__device__ __forceinline__ Wts weights(const Cell& c, double x, double y, double z) {
#pragma clang fp contract(off)
  Wts w;
  w.xd = (c.xmax > c.xmin) ? (x - c.xmin) / (c.xmax - c.xmin) : 0.0;
  w.yd = (c.ymax > c.ymin) ? (y - c.ymin) / (c.ymax - c.ymin) : 0.0;
  w.zd = (c.zmax > c.zmin) ? (z - c.zmin) / (c.zmax - c.zmin) : 0.0;
  return w;
}
__device__ __forceinline__ double at_w(const double* __restrict__ u, const Cell& c, const Wts& w) {
#pragma clang fp contract(off)
  const double c00 = u[c.o[0]] * (1.0 - w.xd) + u[c.o[1]] * w.xd;
  const double c01 = u[c.o[2]] * (1.0 - w.xd) + u[c.o[3]] * w.xd;
  const double c10 = u[c.o[4]] * (1.0 - w.xd) + u[c.o[5]] * w.xd;
  const double c11 = u[c.o[6]] * (1.0 - w.xd) + u[c.o[7]] * w.xd;
  const double c0 = c00 * (1.0 - w.yd) + c10 * w.yd;
  const double c1 = c01 * (1.0 - w.yd) + c11 * w.yd;
  return c0 * (1.0 - w.zd) + c1 * w.zd;
}

// Maxey-Riley acceleration at one point (RKstage, part_track.f90:805-873, and
// SubDerivative, :896-1075), operation by operation with contraction off:
// pressure gradient / added mass 3/2 Du_f/Dt, drag C_D 3/(8a) (u_f - u_p +
// a^2/6 Lap u_f), lift C_L (u_f - u_p + a^2/6 Lap u_f) x omega, the Faxen time
// derivative 3/20 a^2 d(Lap u_f)/dt, buoyancy on y.  The fluid time
// derivatives interpolate the rate fields (f - f_old)/dt (k_mr_rates: the
// reference's corner values, computed once per grid point).
__device__ __forceinline__ void mr_accel(const MrArgs& M, double x, double y, double z, double un, double vn,
                                         double wn, double& ku, double& kv, double& kw, int& err) {
#pragma clang fp contract(off)
  const TracerArgs& A = M.t;
  int imin, jmin, kmax;
  if (find_cell(A, x, y, z, imin, jmin, kmax)) err |= 1;
  const Cell c = cell_of(A, imin, jmin, kmax);
  const Wts wt = weights(c, x, y, z);
  const long long P = A.fstride;
  const double uf = at_w(M.cap, c, wt), vf = at_w(M.cap + P, c, wt), wf = at_w(M.cap + 2 * P, c, wt);
  const double sdu = at_w(M.uold, c, wt) + uf * at_w(M.mrf + 3 * P, c, wt) + vf * at_w(M.mrf + 4 * P, c, wt) +
                     wf * at_w(M.mrf + 5 * P, c, wt);
  const double sdv = at_w(M.uold + P, c, wt) + uf * at_w(M.mrf + 6 * P, c, wt) + vf * at_w(M.mrf + 7 * P, c, wt) +
                     wf * at_w(M.mrf + 8 * P, c, wt);
  const double sdw = at_w(M.uold + 2 * P, c, wt) + uf * at_w(M.mrf + 9 * P, c, wt) +
                     vf * at_w(M.mrf + 10 * P, c, wt) + wf * at_w(M.mrf + 11 * P, c, wt);
  const double ap = A.ap, re = M.re;
  const double du = uf - un, dv = vf - vn, dw = wf - wn;
  const double re_p = 2.0 * ap * re * sqrt(du * du + dv * dv + dw * dw);
  double c_d = 0.0;
  if (M.cd_switch == 1) c_d = 16.0 / (2.0 * ap * re) * M.c_mu;
  else if (M.cd_switch == 2) c_d = 24.0 / (2.0 * ap * re) * (1.0 + 0.15 * pow(re_p, 0.687));
  else if (M.cd_switch == 3) c_d = (re_p > 1.0) ? 16.31 * pow(re_p, -0.8331) : 17.28 * pow(re_p, -0.9711);
  const double pi_f = 1.0 / (M.ratio + 0.5), pi_p = M.ratio / (M.ratio + 0.5);
  const double c_l = 0.5;
  const double a2 = ap * ap / 6.0;
  const double ru = uf - un + a2 * at_w(M.lap, c, wt), rv = vf - vn + a2 * at_w(M.lap + P, c, wt);
  const double rw = wf - wn + a2 * at_w(M.lap + 2 * P, c, wt);
  const double wxf = at_w(M.mrf, c, wt), wyf = at_w(M.mrf + P, c, wt), wzf = at_w(M.mrf + 2 * P, c, wt);
  const double dfac = c_d * 3.0 / (8.0 * ap);
  const double ffac = 3.0 / 20.0 * (ap * ap);
  ku = pi_f * (1.5 * sdu + dfac * ru + c_l * (rv * wzf - rw * wyf) + ffac * at_w(M.lold, c, wt));
  kv = pi_f * (1.5 * sdv + dfac * rv + c_l * (rw * wxf - ru * wzf) + ffac * at_w(M.lold + P, c, wt)) -
       M.gravity * (pi_f - pi_p);
  kw = pi_f * (1.5 * sdw + dfac * rw + c_l * (ru * wyf - rv * wxf) + ffac * at_w(M.lold + 2 * P, c, wt));
}

// (f - f_old)/dt of u, v, w and of the Laplacians, in place of the old fields
// (SubDerivative's corner values, part_track.f90:1001-1009, with the
// sub-step dt)
__global__ __launch_bounds__(256) void k_mr_rates(double* __restrict__ uold, double* __restrict__ lold,
                                                  const double* __restrict__ u, const double* __restrict__ lap,
                                                  size_t n3, double dt) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  uold[i] = (u[i] - uold[i]) / dt;
  lold[i] = (lap[i] - lold[i]) / dt;
}

// part_track.f90:131-150 with RK4 (:617-755) per sub-step: stages at dt/2,
// dt/2, dt from the sub-step's start, x and z periodic, y clamped to
// [a + y_min, y_max - a], elastic bounce of v at a clamped wall; then the swirl
// at the new position (solvers.f90:3479-3481)
__global__ __launch_bounds__(128) void k_inertial(MrArgs M) {
#pragma clang fp contract(off)
  const TracerArgs& A = M.t;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= A.n) return;
  const int n = A.n;
  double* part = A.part;
  double x = part[j], y = part[n + j], z = part[2 * n + j];
  double u = part[3 * n + j], v = part[4 * n + j], w = part[5 * n + j];
  const double height = A.ycoord[0] - A.ap, bottom = A.ap + A.ycoord[A.ny];
  const double dt = A.dt;
  int err = 0;
  for (int s = 0; s < M.substep; ++s) {
    double xs = x, ys = y, zs = z, un = u, vn = v, wn = w;
    // RK4 sums accumulated in the reference's order: K1/6 + K2/3 + K3/3 +
    // K4/6 and u/6 + u2/3 + u3/3 + u4/6 (u2..u4: the stage velocities)
    double sku = 0.0, skv = 0.0, skw = 0.0;
    double sxu = u / 6.0, sxv = v / 6.0, sxw = w / 6.0;
#pragma unroll 1
    for (int sn = 0; sn < 4; ++sn) {
      double ku, kv, kw;
      mr_accel(M, xs, ys, zs, un, vn, wn, ku, kv, kw, err);
      const double kd = (sn == 0 || sn == 3) ? 6.0 : 3.0;
      if (sn == 0) { sku = ku / kd; skv = kv / kd; skw = kw / kd; }
      else { sku = sku + ku / kd; skv = skv + kv / kd; skw = skw + kw / kd; }
      const double h = (sn == 2) ? dt : dt / 2.0;
      un = u + h * ku;
      vn = v + h * kv;
      wn = w + h * kw;
      xs = fmod(A.xl + x + h * un, A.xl);
      ys = fmax(fmin(y + h * vn, height), bottom);
      zs = fmod(A.zl + z + h * wn, A.zl);
      if (sn < 3) {
        const double vd = (sn == 2) ? 6.0 : 3.0;
        sxu = sxu + un / vd;
        sxv = sxv + vn / vd;
        sxw = sxw + wn / vd;
      }
    }
    const double xn = fmod(A.xl + x + dt * sxu, A.xl);
    const double yn = fmax(fmin(y + dt * sxv, height), bottom);
    const double zn = fmod(A.zl + z + dt * sxw, A.zl);
    const double unew = u + dt * sku;
    const double vnew = v + dt * skv;
    const double wnew = w + dt * skw;
    const bool bounce = (yn == height && vnew > 0.0) || (yn == bottom && vnew < 0.0);
    x = xn; y = yn; z = zn;
    u = unew;
    v = bounce ? -vnew : vnew;
    w = wnew;
    // relative velocity at the new position for the particle CFL
    int imin, jmin, kmax;
    if (find_cell(A, x, y, z, imin, jmin, kmax)) err |= 1;
    const Cell c = cell_of(A, imin, jmin, kmax);
    const Wts wt = weights(c, x, y, z);
    const long long P = A.fstride;
    const double ru = at_w(M.cap, c, wt) - u, rv = at_w(M.cap + P, c, wt) - v;
    const double rw = at_w(M.cap + 2 * P, c, wt) - w;
    const double umax = sqrt(ru * ru + rv * rv + rw * rw);
    if (umax * M.dtfull / A.ap >= 10.0) err |= 2;
  }
  part[j] = x;
  part[n + j] = y;
  part[2 * n + j] = z;
  part[3 * n + j] = u;
  part[4 * n + j] = v;
  part[5 * n + j] = w;
  if (M.swirl) {
    int imin, jmin, kmax;
    if (find_cell(A, x, y, z, imin, jmin, kmax)) err |= 1;
    part[6 * n + j] = at(M.cap + 3 * A.fstride, cell_of(A, imin, jmin, kmax), x, y, z);
  }
  if (err) atomicOr(A.err, err);
}

hipError_t launch_inertial(const MrArgs& a, hipStream_t s) {
  if (a.t.n <= 0) return hipSuccess;
  // the rate fields replace uold and lold (the callers refresh both after)
  const size_t n3 = 3 * (size_t)a.t.fstride;
  hipLaunchKernelGGL(k_mr_rates, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, s, const_cast<double*>(a.uold),
                     const_cast<double*>(a.lold), a.cap, a.lap, n3, a.t.dt);
  hipLaunchKernelGGL(k_inertial, dim3((a.t.n + 127) / 128), dim3(128), 0, s, a);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_swirl(int n, const double* __restrict__ g, double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double a[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) a[k] = g[(size_t)k * n + j];
  out[j] = swirl_of(a);
}

hipError_t launch_swirl(int n, const double* g, double* swirl, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_swirl, dim3((n + 255) / 256), dim3(256), 0, s, n, g, swirl);
  return hipGetLastError();
}

// Scalar source of Gaussian blobs at the particles (scl_flag 2,
// solvers.f90:2078-2118): scsource(i,j,k) = sum_n deltaT exp(-(bx + by + bz)),
// bx = (x_k - x_n)^2 / (2 sigmax^2) with x_k = k delxm (no periodic wrap), and
// the reference's cut-offs bx < 20, bx + bz < 20, bx + by + bz < 18.  One
// wavefront per particle scatters into its box (FP64 atomics: the sum order
// over particles is free, as in the reference's OpenMP reduction); lanes run
// over y.
__global__ void k_fill0(double* __restrict__ p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

__global__ __launch_bounds__(64) void k_blob(BlobArgs A) {
  const int q = blockIdx.x, lane = threadIdx.x;
  const double xc = A.part[q], yc = A.part[A.n + q], zc = A.part[2 * A.n + q];
  const double sx = 2.0 * (A.sigmax * A.sigmax), sy = 2.0 * (A.sigmay * A.sigmay);
  const double sz = 2.0 * (A.sigmaz * A.sigmaz);
  // index boxes holding every bx < 20 (bz < 20): |x_k - x_n| < sqrt(40) sigmax,
  // widened by one cell; the exact tests below decide
  const double rx = sqrt(40.0) * A.sigmax, rz = sqrt(40.0) * A.sigmaz;
  auto clampi = [](double v, int hi) { return (int)fmin(fmax(v, 0.0), (double)hi); };
  const int k0 = clampi(floor((xc - rx) / A.delxm) - 1.0, A.mx - 1), k1 = clampi(ceil((xc + rx) / A.delxm) + 1.0, A.mx - 1);
  const int j0 = clampi(floor((zc - rz) / A.delzm) - 1.0, A.mz - 1), j1 = clampi(ceil((zc + rz) / A.delzm) + 1.0, A.mz - 1);
  for (int k = k0; k <= k1; ++k) {
    const double dx = (double)k * A.delxm - xc;
    const double bx = (dx * dx) / sx;
    if (!(bx < 20.0)) continue;
    for (int j = j0; j <= j1; ++j) {
      const double dz = (double)j * A.delzm - zc;
      const double bz = (dz * dz) / sz;
      if (!(bx + bz < 20.0)) continue;
      for (int i = lane; i < A.nyp; i += 64) {
        const double dy = A.ycoord[i] - yc;
        const double b = (bx + (dy * dy) / sy) + bz;
        if (b < 18.0) atomicAdd(&A.src[((size_t)i * A.mz + j) * A.mx + k], A.deltaT * exp(-b));
      }
    }
  }
}

hipError_t launch_blob_source(const BlobArgs& a, hipStream_t s) {
  const size_t n = (size_t)a.nyp * a.mz * a.mx;
  hipLaunchKernelGGL(k_fill0, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a.src, n);
  if (a.n > 0) hipLaunchKernelGGL(k_blob, dim3(a.n), dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_tracer(const TracerArgs& a, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  if (a.part && a.nf != 3 && a.nf != 4) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_tracer, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace kmm
