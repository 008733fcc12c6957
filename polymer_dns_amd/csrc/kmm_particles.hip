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
  // cell search, fluid_interp part_track.f90:575-580 (0-based indices here)
  const double fx = floor(xp / A.delxm), fz = floor(zp / A.delzm);
  const double fk = floor(1.0 + (double)A.ny / A.pi * acos(2.0 * yp / A.yl));
  bool bad = !(fx > -1e9 && fx < 1e9 && fz > -1e9 && fz < 1e9 && fk >= 1.0 && fk <= (double)A.ny);
  int imin = 0, jmin = 0, kmax = 0;
  if (!bad) {
    imin = ((int)fx + A.mx) % A.mx;
    jmin = ((int)fz + A.mz) % A.mz;
    kmax = (int)fk - 1;
    bad = imin < 0 || jmin < 0;
  }
  if (bad) {
    atomicOr(A.err, 1);
    imin = min(max(imin, 0), A.mx - 1);
    jmin = min(max(jmin, 0), A.mz - 1);
    kmax = min(max(kmax, 0), A.ny - 1);
  }
  const int imax = (imin + 1) % A.mx, jmax = (jmin + 1) % A.mz, kmin = kmax + 1;
  if (A.cells) {
    A.cells[j] = imin;
    A.cells[n + j] = jmin;
    A.cells[2 * n + j] = kmax;
  }
  const double xmin = (double)imin * A.delxm, xmax = (double)imax * A.delxm;
  const double ymin = A.ycoord[kmin], ymax = A.ycoord[kmax];
  const double zmin = (double)jmin * A.delzm, zmax = (double)jmax * A.delzm;
  const long long i0 = imin * A.sx, i1 = imax * A.sx, j0 = jmin * A.sz, j1 = jmax * A.sz;
  const long long k0 = kmin * A.sy, k1 = kmax * A.sy;
  for (int f = 0; f < A.nf; ++f) {
    const double* u = A.fld + f * A.fstride;
    const double r = interp3(xp, yp, zp, xmin, xmax, ymin, ymax, zmin, zmax, u[k0 + j0 + i0], u[k0 + j0 + i1],
                             u[k0 + j1 + i0], u[k0 + j1 + i1], u[k1 + j0 + i0], u[k1 + j0 + i1],
                             u[k1 + j1 + i0], u[k1 + j1 + i1]);
    if (A.part) A.part[(3 + f) * n + j] = r;
    else A.out[(size_t)f * n + j] = r;
  }
  if (A.part) {
    A.part[j] = xp;
    A.part[n + j] = yp;
    A.part[2 * n + j] = zp;
  }
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
