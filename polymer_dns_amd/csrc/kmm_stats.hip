// Statistics of writeoutputs (helpers.f90:1227-1436) from the transfer fields
// of the last evaluation (physical y, spectral x and z, [y][kx][kz]).
//
// The reference averages products over the 3/2 physical grid.  For fields
// band-limited to |kx| < nx/2, |kz| <= nz/2 the mean over the 3/2 grid of a
// product u v is exactly sum over modes of Re(u_k conj v_k) (Parseval; no pair
// of modes aliases onto the mean on a grid of 3/2 the size), so the second
// moments come from one pass over the spectral coefficients: kx > 0 modes
// count twice (their conjugates are not stored), kx = 0 once.
//
//   out[0..3][y]  sums over modes of |u|^2 and Re(u v*) without the (0,0)
//                 mode, |v|^2 and |w|^2 with it
//   out[4][y]     Re u(0,0)  (the mean over x and z)
//   out[5][y]     Re s(0,0) of an optional field s (scalar), else 0
//   out[6][y]     Re q(0,0) of an optional field q (polymer str12), else 0
//   out[7][y]     Re (c11 + c22 + c33)(0,0) of optional fields (trace of C), else 0
#include "kmm_kernels.h"
#include "kmm_device.h"

namespace kmm {

template <typename V>
__global__ __launch_bounds__(256) void k_stats(StatsArgs A) {
  const int y = blockIdx.x, t = threadIdx.x;
  const size_t base = (size_t)y * A.P;
  const V* u = static_cast<const V*>(A.u) + base;
  const V* v = static_cast<const V*>(A.v) + base;
  const V* w = static_cast<const V*>(A.w) + base;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int p = t; p < A.P; p += 256) {
    const double wt = (p >= A.nz) ? 2.0 : 1.0;  // p = kx nz + kz
    const double2 a = to_d(u[p]), b = to_d(v[p]), c = to_d(w[p]);
    // writeoutputs takes u' = u - <u> but v and w whole (their means count)
    if (p != 0) s0 += wt * (a.x * a.x + a.y * a.y);
    s1 += wt * (b.x * b.x + b.y * b.y);
    s2 += wt * (c.x * c.x + c.y * c.y);
    if (p != 0) s3 += wt * (a.x * b.x + a.y * b.y);
  }
  __shared__ double sh[4][256];
  sh[0][t] = s0; sh[1][t] = s1; sh[2][t] = s2; sh[3][t] = s3;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h)
      for (int k = 0; k < 4; ++k) sh[k][t] += sh[k][t + h];
    __syncthreads();
  }
  if (t < 4) A.out[(size_t)t * A.nyp + y] = sh[t][0];
  if (t == 0) {
    A.out[(size_t)4 * A.nyp + y] = to_d(u[0]).x;
    A.out[(size_t)5 * A.nyp + y] = A.s ? to_d(static_cast<const V*>(A.s)[base]).x : 0.0;
    A.out[(size_t)6 * A.nyp + y] = A.q ? to_d(static_cast<const V*>(A.q)[base]).x : 0.0;
    double tr = 0.0;
    if (A.c11) tr = to_d(static_cast<const V*>(A.c11)[base]).x + to_d(static_cast<const V*>(A.c22)[base]).x +
                    to_d(static_cast<const V*>(A.c33)[base]).x;
    A.out[(size_t)7 * A.nyp + y] = tr;
  }
}

// mean over each y-plane of a physical [nyp][n] field
__global__ __launch_bounds__(256) void k_plane_mean(const double* __restrict__ f, int n, double* __restrict__ out) {
  const int y = blockIdx.x, t = threadIdx.x;
  double s = 0.0;
  for (int i = t; i < n; i += 256) s += f[(size_t)y * n + i];
  __shared__ double sh[256];
  sh[t] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) sh[t] += sh[t + h];
    __syncthreads();
  }
  if (t == 0) out[y] = sh[0] / (double)n;
}

hipError_t launch_stats(const StatsArgs& a, hipStream_t s) {
  if (a.mixed) hipLaunchKernelGGL(k_stats<float2>, dim3(a.nyp), dim3(256), 0, s, a);
  else hipLaunchKernelGGL(k_stats<double2>, dim3(a.nyp), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_plane_mean(const double* f, int nyp, int n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_plane_mean, dim3(nyp), dim3(256), 0, s, f, n, out);
  return hipGetLastError();
}

}  // namespace kmm
