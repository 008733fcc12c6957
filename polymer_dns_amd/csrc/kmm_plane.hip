// Plane-stage kernels of the KMM channel step (the per-y-plane part of
// vcw3dp, FFTW_DNS/code/src/dns/solvers.f90:2550-3275).
//
//   k_zinv   : zero-pad nz -> mz (negative kz on top, solvers.f90:2652-2656)
//              and inverse c2c transform along z for every (y, kx) line.
//   k_ibdrag : Goldstein feedback at solid cells (solvers.f90:2996-3003).
//   k_xplane : two z rows per wavefront packed into one complex x transform
//              (c2r semantics, imaginary part of kx=0 ignored), the pointwise
//              products u x omega and -u.grad(theta) (solvers.f90:2891-2898),
//              CFL, the dPdx body force and the order-dependent 5x5 spreading
//              of the feedback force, then the forward x transform (r2c).
//   k_zfwd   : forward c2c along z, keep nz modes, 1/(mx mz), and the +-kz
//              projection of ``norm`` (helpers.f90:91-158).
// Intermediate layout Z[f][y][z][kx] keeps the x lines contiguous.
#include "kmm_kernels.h"
#include "kmm_device.h"

namespace kmm {

template <int MZ, int NFIN>
__global__ __launch_bounds__(512) void k_zinv(PlaneArgs A) {
  extern __shared__ double2 sm[];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ngx = A.nxh / 8;
  const int y = blockIdx.x / ngx, kx0 = (blockIdx.x % ngx) * 8;
  const int kx = kx0 + wv;
  double2* buf = sm + wv * 2 * MZ;
  double2* scr = buf + MZ;
  double2* tile = sm + 16 * MZ;  // [MZ][8]
#pragma unroll
  for (int f = 0; f < NFIN; ++f) {
    const double2* src = A.in[f] + ((size_t)y * A.nxh + kx) * A.nz;
    for (int z = lane; z < MZ; z += 64) buf[z] = czero();
    wave_sync();
    for (int j = lane; j < A.nz; j += 64) {
      const int jj = (j < A.nzh) ? j : (MZ - A.nz + j);
      buf[jj] = src[j];
    }
    wave_sync();
    double2* r = wave_fft<MZ, +1>(buf, scr, A.twZ, lane);
    for (int z = lane; z < MZ; z += 64) tile[z * 8 + wv] = r[z];
    __syncthreads();
    double2* dst = A.Z + ((size_t)f * A.nyp + y) * MZ * A.nxh;
    for (int e = threadIdx.x; e < MZ * 8; e += 512) {
      const int z = e >> 3, c = e & 7;
      dst[(size_t)z * A.nxh + kx0 + c] = tile[e];
    }
    __syncthreads();
  }
}

template <int MX>
__global__ void k_ibdrag(PlaneArgs A) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= A.nsrc) return;
  const int4 c = A.src[s];
  const bool active = (c.w == 1) || (c.w == 4 && A.it <= A.perturbtime);
  if (!active) {
    for (int d = 0; d < 3; ++d) A.drag[d * A.nsrc + s] = 0.0;
    return;
  }
  for (int d = 0; d < 3; ++d) {
    const double2* row = A.Z + (((size_t)d * A.nyp + c.x) * A.mz + c.y) * A.nxh;
    double acc = 0.0;
    for (int k = 1; k < A.nxh; ++k) {
      const int m = (int)(((long long)k * c.z) % MX);
      const double2 t = A.twX[m];  // exp(-2 pi i m/MX); c2r uses exp(+..)
      const double2 v = row[k];
      acc += v.x * t.x + v.y * t.y;
    }
    const double u = row[0].x + 2.0 * acc;
    double* fi = A.fint + (size_t)d * A.nsrc + s;
    const double fnew = *fi + u;
    *fi = fnew;
    A.drag[d * A.nsrc + s] = -A.ugain * u - A.gain * fnew;
  }
}

template <int MX, int NFIN, int NFOUT, int WPB>
__global__ __launch_bounds__(WPB * 64) void k_xplane(PlaneArgs A) {
  extern __shared__ double2 sm[];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nzb = A.mz / (2 * WPB);
  const int y = blockIdx.x / nzb;
  const int zA = (blockIdx.x % nzb) * 2 * WPB + 2 * wv, zB = zA + 1;
  double2* buf = sm + (size_t)wv * (2 * MX + NFIN * MX);
  double2* scr = buf + MX;
  double* ph = reinterpret_cast<double*>(buf + 2 * MX);  // [NFIN][2][MX]
  const int nxh = A.nxh;
#pragma unroll
  for (int f = 0; f < NFIN; ++f) {
    const double2* zr = A.Z + ((size_t)f * A.nyp + y) * A.mz * nxh;
    const double2* ra = zr + (size_t)zA * nxh;
    const double2* rb = zr + (size_t)zB * nxh;
    for (int x = lane; x < MX; x += 64) buf[x] = czero();
    wave_sync();
    for (int k = lane; k < nxh; k += 64) {
      const double2 ua = ra[k], ub = rb[k];
      if (k == 0) {
        buf[0] = make_double2(ua.x, ub.x);
      } else {
        buf[k] = make_double2(ua.x - ub.y, ua.y + ub.x);
        buf[MX - k] = make_double2(ua.x + ub.y, ub.x - ua.y);
      }
    }
    wave_sync();
    double2* r = wave_fft<MX, +1>(buf, scr, A.twX, lane);
    for (int x = lane; x < MX; x += 64) {
      const double2 v = r[x];
      ph[(f * 2 + 0) * MX + x] = v.x;
      ph[(f * 2 + 1) * MX + x] = v.y;
    }
    wave_sync();
  }
  // pointwise products (solvers.f90:2891-2898), CFL (2985-2986), dPdx (3044)
  const double sgy = A.seght[y];
  const double dpdx = (A.flow_select == 1) ? A.dPdx : 0.0;
  double cmax = 0.0;
  for (int e = lane; e < 2 * MX; e += 64) {
    const int rr = e / MX, x = e % MX;
    const double u = ph[(0 * 2 + rr) * MX + x], v = ph[(1 * 2 + rr) * MX + x], w = ph[(2 * 2 + rr) * MX + x];
    const double ox = ph[(3 * 2 + rr) * MX + x], oy = ph[(4 * 2 + rr) * MX + x], oz = ph[(5 * 2 + rr) * MX + x];
    double vc = 0.0;
    if (NFIN == 9) {
      const double tx = ph[(6 * 2 + rr) * MX + x], ty = ph[(7 * 2 + rr) * MX + x], tz = ph[(8 * 2 + rr) * MX + x];
      vc = -(u * tx + v * ty + w * tz) + A.src_uniform;
    }
    const double c = (fabs(u) / A.delxm + fabs(v) / sgy + fabs(w) / A.delzm) * A.dt;
    cmax = fmax(cmax, c);
    ph[(0 * 2 + rr) * MX + x] = v * oz - w * oy - dpdx;
    ph[(1 * 2 + rr) * MX + x] = -u * oz + w * ox;
    ph[(2 * 2 + rr) * MX + x] = u * oy - v * ox;
    if (NFOUT == 4) ph[(3 * 2 + rr) * MX + x] = vc;
  }
  cmax = wave_max(cmax);
  if (lane == 0) atomicMax(A.cfl, (unsigned long long)__double_as_longlong(cmax));
  wave_sync();
  if (A.nsrc > 0) {
    // spreading of the feedback force onto targets of rows zA, zB
    const int r0 = y * A.mz + zA;
    const int tbeg = A.trow[r0], tmid = A.trow[r0 + 1], tend = A.trow[r0 + 2];
    for (int t = tbeg + lane; t < tend; t += 64) {
      const int rr = (t < tmid) ? 0 : 1;
      const int x = A.tx[t];
      double ax = 0.0, ay = 0.0, az = 0.0;
      for (int q = A.tptr[t]; q < A.tptr[t + 1]; ++q) {
        const int s = A.ent_src[q];
        const double wgt = A.ent_w[q];
        ax += wgt * A.drag[s];
        ay += wgt * A.drag[A.nsrc + s];
        az += wgt * A.drag[2 * A.nsrc + s];
      }
      ph[(0 * 2 + rr) * MX + x] += ax;
      ph[(1 * 2 + rr) * MX + x] += ay;
      ph[(2 * 2 + rr) * MX + x] += az;
    }
    wave_sync();
    if (dpdx != 0.0) {
      const int fbeg = A.fixrow[r0], fmid = A.fixrow[r0 + 1], fend = A.fixrow[r0 + 2];
      for (int t = fbeg + lane; t < fend; t += 64) {
        const int rr = (t < fmid) ? 0 : 1;
        ph[(0 * 2 + rr) * MX + A.fixx[t]] += dpdx;
      }
    }
    wave_sync();
  }
  // forward x transforms (r2c) of the outputs, two rows per transform
#pragma unroll
  for (int o = 0; o < NFOUT; ++o) {
    for (int x = lane; x < MX; x += 64)
      buf[x] = make_double2(ph[(o * 2 + 0) * MX + x], ph[(o * 2 + 1) * MX + x]);
    wave_sync();
    double2* r = wave_fft<MX, -1>(buf, scr, A.twX, lane);
    double2* zo = A.ZF + ((size_t)o * A.nyp + y) * A.mz * nxh;
    for (int k = lane; k < nxh; k += 64) {
      const double2 zk = r[k], zm = r[(MX - k) % MX];
      const double2 ua = make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
      const double2 ub = make_double2(0.5 * (zk.y + zm.y), -0.5 * (zk.x - zm.x));
      zo[(size_t)zA * nxh + k] = ua;
      zo[(size_t)zB * nxh + k] = ub;
    }
    wave_sync();
  }
}

template <int MZ, int NFOUT>
__global__ __launch_bounds__(512) void k_zfwd(PlaneArgs A) {
  extern __shared__ double2 sm[];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ngx = A.nxh / 8;
  const int y = blockIdx.x / ngx, kx0 = (blockIdx.x % ngx) * 8;
  const int kx = kx0 + wv;
  double2* buf = sm + wv * 2 * MZ;
  double2* scr = buf + MZ;
  double2* tile = sm + 16 * MZ;
  const double scale = 1.0 / ((double)A.mx * (double)MZ);
  const int nz = A.nz, nzh = A.nzh;
#pragma unroll
  for (int o = 0; o < NFOUT; ++o) {
    const double2* srcp = A.ZF + ((size_t)o * A.nyp + y) * MZ * A.nxh;
    for (int e = threadIdx.x; e < MZ * 8; e += 512) {
      const int z = e >> 3, c = e & 7;
      tile[e] = srcp[(size_t)z * A.nxh + kx0 + c];
    }
    __syncthreads();
    for (int z = lane; z < MZ; z += 64) buf[z] = tile[z * 8 + wv];
    wave_sync();
    double2* r = wave_fft<MZ, -1>(buf, scr, A.twZ, lane);
    double2* o2 = (r == buf) ? scr : buf;
    for (int j = lane; j < nz; j += 64) {
      const int jj = (j < nzh) ? j : (MZ - nz + j);
      o2[j] = (j == nzh) ? czero() : cscale(r[jj], scale);
    }
    wave_sync();
    double2* dst = A.out[o] + ((size_t)y * A.nxh + kx) * nz;
    for (int j = lane; j < nz; j += 64) {
      double2 v = o2[j];
      if (kx == 0) {
        if (j == 0) v.y = 0.0;
        else if (j < nzh) v = cscale(cadd(o2[j], cconj(o2[nz - j])), 0.5);
        else if (j > nzh) v = cscale(cadd(o2[j], cconj(o2[nz - j])), 0.5);
      }
      dst[j] = v;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------ launchers --
template <int MX, int MZ, int NFIN, int NFOUT>
static hipError_t plane_impl(const PlaneArgs& a, hipStream_t s, bool inverse_part, hipEvent_t* ev) {
  const int ngx = a.nxh / 8;
  const size_t zlds = (size_t)(16 * MZ + 8 * MZ) * sizeof(double2);
  if (inverse_part) {
    hipLaunchKernelGGL((k_zinv<MZ, NFIN>), dim3(a.nyp * ngx), dim3(512), zlds, s, a);
    if (ev) hipEventRecord(ev[0], s);
    if (a.nsrc > 0)
      hipLaunchKernelGGL((k_ibdrag<MX>), dim3((a.nsrc + 127) / 128), dim3(128), 0, s, a);
    if (ev) hipEventRecord(ev[1], s);
    constexpr size_t per_wave = (size_t)(2 * MX + NFIN * MX) * sizeof(double2);
    constexpr int WPB = (4 * per_wave <= 150 * 1024) ? 4 : (2 * per_wave <= 150 * 1024 ? 2 : 1);
    const size_t xlds = (size_t)WPB * per_wave;
    hipLaunchKernelGGL((k_xplane<MX, NFIN, NFOUT, WPB>), dim3(a.nyp * (a.mz / (2 * WPB))), dim3(WPB * 64), xlds, s,
                       a);
    if (ev) hipEventRecord(ev[2], s);
  }
  hipLaunchKernelGGL((k_zfwd<MZ, NFOUT>), dim3(a.nyp * ngx), dim3(512), zlds, s, a);
  if (ev) hipEventRecord(ev[3], s);
  return hipGetLastError();
}

template <int MX, int MZ>
static hipError_t plane_dispatch_f(const PlaneArgs& a, hipStream_t s, bool inv, hipEvent_t* ev) {
  if (a.nfin == 6 && a.nfout == 3) return plane_impl<MX, MZ, 6, 3>(a, s, inv, ev);
  if (a.nfin == 9 && a.nfout == 4) return plane_impl<MX, MZ, 9, 4>(a, s, inv, ev);
  return hipErrorInvalidValue;
}

static hipError_t plane_dispatch(const PlaneArgs& a, hipStream_t s, bool inv, hipEvent_t* ev) {
  if (a.nxh % 8 != 0) return hipErrorInvalidValue;
#define KMM_PLANE_CASE(X, Zz) \
  if (a.mx == X && a.mz == Zz) return plane_dispatch_f<X, Zz>(a, s, inv, ev);
  KMM_PLANE_CASE(48, 48)
  KMM_PLANE_CASE(96, 48)
  KMM_PLANE_CASE(96, 96)
  KMM_PLANE_CASE(144, 144)
  KMM_PLANE_CASE(192, 96)
  KMM_PLANE_CASE(192, 192)
  KMM_PLANE_CASE(288, 288)
  KMM_PLANE_CASE(384, 192)
#undef KMM_PLANE_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_plane(const PlaneArgs& a, hipStream_t s) { return plane_dispatch(a, s, true, nullptr); }
hipError_t launch_plane_timed(const PlaneArgs& a, hipStream_t s, hipEvent_t* ev) {
  return plane_dispatch(a, s, true, ev);
}
hipError_t launch_plane_fwd_only(const PlaneArgs& a, hipStream_t s) { return plane_dispatch(a, s, false, nullptr); }

}  // namespace kmm
