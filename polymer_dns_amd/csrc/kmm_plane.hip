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

template <int MZ, int NFIN, typename V>
__global__ __launch_bounds__(512) void k_zinv(PlaneArgs A) {
  extern __shared__ __align__(16) unsigned char smraw[];
  V* sm = reinterpret_cast<V*>(smraw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ngx = A.nxh / 8;
  const int y = blockIdx.x / ngx, kx0 = (blockIdx.x % ngx) * 8;
  const int kx = kx0 + wv;
  constexpr int BL = FFTBuf<MZ>::len;
  V* buf = sm + wv * 2 * BL;
  V* scr = buf + BL;
  V* tile = sm + 16 * BL;       // [MZ][8]
  V* tw = tile + 8 * MZ;        // [MZ] twiddles
  for (int e = threadIdx.x; e < MZ; e += 512) tw[e] = static_cast<const V*>(A.twZ)[e];
  __syncthreads();
#pragma unroll
  for (int f = 0; f < NFIN; ++f) {
    const V* src = static_cast<const V*>(A.in[f]) + ((size_t)y * A.nxh + kx) * A.nz;
    for (int z = lane; z < MZ; z += 64) buf[fpad(z)] = vzero<V>();
    wave_sync();
    for (int j = lane; j < A.nz; j += 64) {
      const int jj = (j < A.nzh) ? j : (MZ - A.nz + j);
      buf[fpad(jj)] = src[j];
    }
    wave_sync();
    V* r = wave_fft<MZ, +1>(buf, scr, tw, lane);
    for (int z = lane; z < MZ; z += 64) tile[z * 8 + wv] = r[fpad(z)];
    __syncthreads();
    V* dst = static_cast<V*>(A.Z) + ((size_t)f * A.nyp + y) * MZ * A.nxh;
    for (int e = threadIdx.x; e < MZ * 8; e += 512) {
      const int z = e >> 3, c = e & 7;
      dst[(size_t)z * A.nxh + kx0 + c] = tile[e];
    }
    __syncthreads();
  }
}

template <int MX, typename V>
__global__ __launch_bounds__(256) void k_ibdrag(PlaneArgs A) {
  const int lane = threadIdx.x & 63;
  const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (s >= A.nsrc) return;
  const int4 c = A.src[s];
  const bool active = (c.w == 1) || (c.w == 4 && A.it <= A.perturbtime);
  if (!active) {
    if (lane < 3) A.drag[lane * A.nsrc + s] = 0.0;
    return;
  }
  double u3[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    // c2r value at x: Re Z_0 + 2 sum_{k>=1} Re(Z_k exp(+2 pi i k x/MX))
    const V* row = static_cast<const V*>(A.Z) + (((size_t)d * A.nyp + c.x) * A.mz + c.y) * A.nxh;
    double acc = 0.0;
    for (int k = lane; k < A.nxh; k += 64) {
      const double2 v = to_d(row[k]);
      if (k == 0) {
        acc += v.x;
      } else {
        const double2 t = to_d(static_cast<const V*>(A.twX)[(int)(((long long)k * c.z) % MX)]);  // exp(-2 pi i m/MX)
        acc += 2.0 * (v.x * t.x + v.y * t.y);
      }
    }
    u3[d] = wave_sum(acc);
  }
  if (lane < 3) {
    const double u = (lane == 0) ? u3[0] : (lane == 1 ? u3[1] : u3[2]);
    double* fi = A.fint + (size_t)lane * A.nsrc + s;
    const double fnew = *fi + u;
    *fi = fnew;
    A.drag[lane * A.nsrc + s] = -A.ugain * u - A.gain * fnew;
  }
}

template <int MX, int NFIN, int NFOUT, int WPB, typename V>
__global__ __launch_bounds__(WPB * 64) void k_xplane(PlaneArgs A) {
  using T = typename Scal<V>::T;
  constexpr int XPL = (MX + 63) / 64;  // x positions per lane: x = lane + 64 m
  extern __shared__ __align__(16) unsigned char smraw[];
  V* sm = reinterpret_cast<V*>(smraw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nzb = A.mz / (2 * WPB);
  const int y = blockIdx.x / nzb;
  const int zA = (blockIdx.x % nzb) * 2 * WPB + 2 * wv, zB = zA + 1;
  constexpr int BL = FFTBuf<MX>::len;
  V* tw = sm;                              // [MX] twiddles
  V* buf = sm + MX + (size_t)wv * 2 * BL;
  V* scr = buf + BL;
  for (int e = threadIdx.x; e < MX; e += WPB * 64) tw[e] = static_cast<const V*>(A.twX)[e];
  __syncthreads();
  const int nxh = A.nxh;
  // ph[f][m] = (row A, row B) physical values at x = lane + 64 m
  V ph[NFIN][XPL];
#pragma unroll
  for (int f = 0; f < NFIN; ++f) {
    const V* zr = static_cast<const V*>(A.Z) + ((size_t)f * A.nyp + y) * A.mz * nxh;
    const V* ra = zr + (size_t)zA * nxh;
    const V* rb = zr + (size_t)zB * nxh;
    for (int x = lane; x < MX; x += 64) buf[fpad(x)] = vzero<V>();
    wave_sync();
    for (int k = lane; k < nxh; k += 64) {
      const V ua = ra[k], ub = rb[k];
      if (k == 0) {
        buf[0] = mkv<V>(ua.x, ub.x);
      } else {
        buf[fpad(k)] = mkv<V>(ua.x - ub.y, ua.y + ub.x);
        buf[fpad(MX - k)] = mkv<V>(ua.x + ub.y, ub.x - ua.y);
      }
    }
    wave_sync();
    V* r = wave_fft<MX, +1>(buf, scr, tw, lane);
#pragma unroll
    for (int m = 0; m < XPL; ++m) {
      const int x = lane + 64 * m;
      ph[f][m] = (x < MX) ? r[fpad(x)] : vzero<V>();
    }
    wave_sync();
  }
  // physical u, v, w for particle tracking (up3d/vp3d/wp3d, solvers.f90:3113-3115)
  if (A.cap) {
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      double* ca = A.cap + (((size_t)f * A.nyp + y) * A.mz + zA) * MX;
#pragma unroll
      for (int m = 0; m < XPL; ++m) {
        const int x = lane + 64 * m;
        if (x < MX) {
          ca[x] = (double)ph[f][m].x;
          ca[MX + x] = (double)ph[f][m].y;
        }
      }
    }
  }
  // pointwise products (solvers.f90:2891-2898), CFL (2985-2986), dPdx (3044)
  const T sgy = (T)A.seght[y];
  const T dpdx = (T)((A.flow_select == 1) ? A.dPdx : 0.0);
  const T delxm = (T)A.delxm, delzm = (T)A.delzm, dt = (T)A.dt, src_u = (T)A.src_uniform;
  double cmax = 0.0;
#pragma unroll
  for (int m = 0; m < XPL; ++m) {
    const bool ok = (lane + 64 * m) < MX;
#pragma unroll
    for (int rr = 0; rr < 2; ++rr) {
      const T u = rr ? ph[0][m].y : ph[0][m].x, v = rr ? ph[1][m].y : ph[1][m].x;
      const T w = rr ? ph[2][m].y : ph[2][m].x, ox = rr ? ph[3][m].y : ph[3][m].x;
      const T oy = rr ? ph[4][m].y : ph[4][m].x, oz = rr ? ph[5][m].y : ph[5][m].x;
      T vc = 0;
      if (NFIN == 9) {
        const T tx = rr ? ph[6][m].y : ph[6][m].x, ty = rr ? ph[7][m].y : ph[7][m].x;
        const T tz = rr ? ph[8][m].y : ph[8][m].x;
        vc = -(u * tx + v * ty + w * tz) + src_u;
      }
      const double c = (double)((fabs(u) / delxm + fabs(v) / sgy + fabs(w) / delzm) * dt);
      if (ok) cmax = fmax(cmax, c);
      const T o0 = v * oz - w * oy - dpdx, o1 = -u * oz + w * ox, o2 = u * oy - v * ox;
      if (rr) { ph[0][m].y = o0; ph[1][m].y = o1; ph[2][m].y = o2; if (NFOUT == 4) ph[3][m].y = vc; }
      else    { ph[0][m].x = o0; ph[1][m].x = o1; ph[2][m].x = o2; if (NFOUT == 4) ph[3][m].x = vc; }
    }
  }
  cmax = wave_max(cmax);
  if (lane == 0) atomicMax(A.cfl, (unsigned long long)__double_as_longlong(cmax));
  if (A.nsrc > 0) {
    // gathered spreading of the feedback force onto targets of rows zA, zB;
    // the lane owning x adds (x = lane + 64 m)
    const int r0 = y * A.mz + zA;
    const int tbeg = A.trow[r0], tmid = A.trow[r0 + 1], tend = A.trow[r0 + 2];
    for (int t = tbeg; t < tend; ++t) {
      const int x = A.tx[t];
      if ((x & 63) != lane) continue;
      const int rr = (t < tmid) ? 0 : 1;
      double ax = 0.0, ay = 0.0, az = 0.0;
      for (int q = A.tptr[t]; q < A.tptr[t + 1]; ++q) {
        const int s = A.ent_src[q];
        const double wgt = A.ent_w[q];
        ax += wgt * A.drag[s];
        ay += wgt * A.drag[A.nsrc + s];
        az += wgt * A.drag[2 * A.nsrc + s];
      }
#pragma unroll
      for (int m = 0; m < XPL; ++m) {
        if ((x >> 6) != m) continue;
        if (rr) { ph[0][m].y += (T)ax; ph[1][m].y += (T)ay; ph[2][m].y += (T)az; }
        else    { ph[0][m].x += (T)ax; ph[1][m].x += (T)ay; ph[2][m].x += (T)az; }
      }
    }
    if (dpdx != 0.0) {
      const int fbeg = A.fixrow[r0], fmid = A.fixrow[r0 + 1], fend = A.fixrow[r0 + 2];
      for (int t = fbeg; t < fend; ++t) {
        const int x = A.fixx[t];
        if ((x & 63) != lane) continue;
        const int rr = (t < fmid) ? 0 : 1;
#pragma unroll
        for (int m = 0; m < XPL; ++m) {
          if ((x >> 6) != m) continue;
          if (rr) ph[0][m].y += dpdx; else ph[0][m].x += dpdx;
        }
      }
    }
  }
  // forward x transforms (r2c) of the outputs, two rows per transform
#pragma unroll
  for (int o = 0; o < NFOUT; ++o) {
#pragma unroll
    for (int m = 0; m < XPL; ++m) {
      const int x = lane + 64 * m;
      if (x < MX) buf[fpad(x)] = ph[o][m];
    }
    wave_sync();
    V* r = wave_fft<MX, -1>(buf, scr, tw, lane);
    V* zo = static_cast<V*>(A.ZF) + ((size_t)o * A.nyp + y) * A.mz * nxh;
    for (int k = lane; k < nxh; k += 64) {
      const V zk = r[fpad(k)], zm = r[fpad((MX - k) % MX)];
      const V ua = mkv<V>((T)0.5 * (zk.x + zm.x), (T)0.5 * (zk.y - zm.y));
      const V ub = mkv<V>((T)0.5 * (zk.y + zm.y), (T)-0.5 * (zk.x - zm.x));
      zo[(size_t)zA * nxh + k] = ua;
      zo[(size_t)zB * nxh + k] = ub;
    }
    wave_sync();
  }
}

template <int MZ, int NFOUT, typename V>
__global__ __launch_bounds__(512) void k_zfwd(PlaneArgs A) {
  using T = typename Scal<V>::T;
  extern __shared__ __align__(16) unsigned char smraw[];
  V* sm = reinterpret_cast<V*>(smraw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ngx = A.nxh / 8;
  const int y = blockIdx.x / ngx, kx0 = (blockIdx.x % ngx) * 8;
  const int kx = kx0 + wv;
  constexpr int BL = FFTBuf<MZ>::len;
  V* buf = sm + wv * 2 * BL;
  V* scr = buf + BL;
  V* tile = sm + 16 * BL;
  V* tw = tile + 8 * MZ;
  for (int e = threadIdx.x; e < MZ; e += 512) tw[e] = static_cast<const V*>(A.twZ)[e];
  __syncthreads();
  const T scale = (T)(1.0 / ((double)A.mx * (double)MZ));
  const int nz = A.nz, nzh = A.nzh;
#pragma unroll
  for (int o = 0; o < NFOUT; ++o) {
    const V* srcp = static_cast<const V*>(A.ZF) + ((size_t)o * A.nyp + y) * MZ * A.nxh;
    for (int e = threadIdx.x; e < MZ * 8; e += 512) {
      const int z = e >> 3, c = e & 7;
      tile[e] = srcp[(size_t)z * A.nxh + kx0 + c];
    }
    __syncthreads();
    for (int z = lane; z < MZ; z += 64) buf[fpad(z)] = tile[z * 8 + wv];
    wave_sync();
    V* r = wave_fft<MZ, -1>(buf, scr, tw, lane);
    V* o2 = (r == buf) ? scr : buf;
    for (int j = lane; j < nz; j += 64) {
      const int jj = (j < nzh) ? j : (MZ - nz + j);
      o2[j] = (j == nzh) ? vzero<V>() : cscale(r[fpad(jj)], scale);
    }
    wave_sync();
    V* dst = static_cast<V*>(A.out[o]) + ((size_t)y * A.nxh + kx) * nz;
    for (int j = lane; j < nz; j += 64) {
      V v = o2[j];
      if (kx == 0) {
        if (j == 0) v.y = 0;
        else if (j != nzh) v = cscale(cadd(o2[j], cconj(o2[nz - j])), (T)0.5);
      }
      dst[j] = v;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------ launchers --
template <int MX, int MZ, int NFIN, int NFOUT, typename V>
static hipError_t plane_impl(const PlaneArgs& a, hipStream_t s, bool inverse_part, hipEvent_t* ev) {
  const int ngx = a.nxh / 8;
  const size_t zlds = (size_t)(16 * FFTBuf<MZ>::len + 9 * MZ) * sizeof(V);
  if (inverse_part) {
    hipLaunchKernelGGL((k_zinv<MZ, NFIN, V>), dim3(a.nyp * ngx), dim3(512), zlds, s, a);
    if (ev) hipEventRecord(ev[0], s);
    if (a.nsrc > 0)
      hipLaunchKernelGGL((k_ibdrag<MX, V>), dim3((a.nsrc + 3) / 4), dim3(256), 0, s, a);
    if (ev) hipEventRecord(ev[1], s);
    constexpr int WPB = 4;
    const size_t xlds = (size_t)(MX + WPB * 2 * FFTBuf<MX>::len) * sizeof(V);
    hipLaunchKernelGGL((k_xplane<MX, NFIN, NFOUT, WPB, V>), dim3(a.nyp * (a.mz / (2 * WPB))), dim3(WPB * 64), xlds,
                       s, a);
    if (ev) hipEventRecord(ev[2], s);
  }
  hipLaunchKernelGGL((k_zfwd<MZ, NFOUT, V>), dim3(a.nyp * ngx), dim3(512), zlds, s, a);
  if (ev) hipEventRecord(ev[3], s);
  return hipGetLastError();
}

template <int MX, int MZ>
static hipError_t plane_dispatch_f(const PlaneArgs& a, hipStream_t s, bool inv, hipEvent_t* ev) {
  if (a.mixed) {
    if (a.nfin == 6 && a.nfout == 3) return plane_impl<MX, MZ, 6, 3, float2>(a, s, inv, ev);
    if (a.nfin == 9 && a.nfout == 4) return plane_impl<MX, MZ, 9, 4, float2>(a, s, inv, ev);
  } else {
    if (a.nfin == 6 && a.nfout == 3) return plane_impl<MX, MZ, 6, 3, double2>(a, s, inv, ev);
    if (a.nfin == 9 && a.nfout == 4) return plane_impl<MX, MZ, 9, 4, double2>(a, s, inv, ev);
  }
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
