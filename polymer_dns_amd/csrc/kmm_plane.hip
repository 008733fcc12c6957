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
#include <algorithm>
#include <cstring>
#include <utility>

#include "kmm_kernels.h"
#include "kmm_device.h"

namespace kmm {

// FFT plan cost model of the x and z transforms (make_plan): 1, fewest
// vector instructions per lane.  Measured against fewest rounds at 128^3
// FP64: k_zinv -11 %, k_zfwd -9 %, k_xplane -3 % (more LDS passes, fewer
// idle lanes in the radix-8 stages).
constexpr int kPlanePlan = 1;

// one wave's line of nz spectral values, NJ per lane
template <int NJ, typename V>
__device__ __forceinline__ void zinv_line_load(V (&pre)[NJ], const V* __restrict__ src, int nz, int lane) {
#pragma unroll
  for (int q = 0; q < NJ; ++q) {
    const int j = lane + 64 * q;
    pre[q] = (j < nz) ? src[j] : vzero<V>();
  }
}

// LDS of the z kernels: per wave one FFT buffer (the transforms run in
// place: every stage of the z plans is one round of butterflies, so a
// stage's LDS reads precede its writes; the waves' results are read in
// place by the block's store, and k_zfwd's block load writes straight into
// them), then the MZ twiddles: 43.9 KB for FP64 lines of 288, 29.3 KB at 192.
template <int MZ, typename V>
struct ZLds {
  static_assert(fft_in_place<MZ, kPlanePlan>(), "z transforms run in place");
  static constexpr size_t bytes = (size_t)(8 * FFTBuf<MZ>::len + MZ) * sizeof(V);
};

// Work items are (field, y, group of 8 kx lines), walked by a grid sized to
// the blocks the CUs hold at once (launch_zitems): with 9 fields per
// (y, group) the last round of blocks is a small fraction of the work, where a
// grid of one block per (y, group) would leave a nearly empty third round.
template <int MZ, int NFIN, typename V>
__device__ __forceinline__ void zinv_body(const PlaneArgs& A, int bid) {
  extern __shared__ __align__(16) unsigned char smraw[];
  V* sm = reinterpret_cast<V*>(smraw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ngx = A.nxh / 8;
  const int ngroups = A.nyp * ngx;
  const int nitems = NFIN * ngroups;
  constexpr int BL = FFTBuf<MZ>::len;
  V* buf = sm + wv * BL;
  V* scr = buf;  // in place
  V* twl = sm + 8 * BL;
  for (int e = threadIdx.x; e < MZ; e += 512) twl[e] = static_cast<const V*>(A.twZ)[e];
  __syncthreads();
  const V* tw = twl;
  int it = bid;
  if (it >= nitems) return;
  auto line = [&](int i) -> const V* {
    const int f = i / ngroups, g = i - f * ngroups;
    const int y = g / ngx, kx = (g - y * ngx) * 8 + wv;
    return static_cast<const V*>(A.in[f]) + ((size_t)y * A.nxh + kx) * A.nz;
  };
  // each item's line is loaded into registers one item ahead, so its
  // latency hides behind the previous item's transform and store
  constexpr int NJ = (2 * MZ / 3 + 63) / 64;  // values of a line per lane (nz = 2 mz / 3)
  V pre[NJ];
  zinv_line_load<NJ, V>(pre, line(it), A.nz, lane);
  for (; it < nitems; it += gridDim.x) {
    const int f = it / ngroups, g = it - f * ngroups;
    const int y = g / ngx, kx0 = (g - y * ngx) * 8;
    const int kx = kx0 + wv;
    // zeros at the padding band [nzh, MZ - nzh); the line's modes fill the rest
    for (int z = A.nzh + lane; z < MZ - A.nzh; z += 64) buf[fpad(z)] = vzero<V>();
    wave_sync();
    const bool dz = (A.dzmask >> f) & 1u;  // d/dz = multiplication by i kz
#pragma unroll
    for (int q = 0; q < NJ; ++q) {
      const int j = lane + 64 * q;
      if (j < A.nz) {
        const int jj = (j < A.nzh) ? j : (MZ - A.nz + j);
        V val = pre[q];
        if (dz) {
          using T = typename Scal<V>::T;
          const T wz = (T)(A.beta * (double)((j < A.nzh) ? j : j - A.nz));
          val = mkv<V>(-val.y * wz, val.x * wz);
        }
        buf[fpad(jj)] = val;
      }
    }
    if (it + (int)gridDim.x < nitems) zinv_line_load<NJ, V>(pre, line(it + gridDim.x), A.nz, lane);
    wave_sync();
    // every wave's result sits in its own buffer: the block stores the
    // group's z rows from there
    wave_fft<MZ, +1, V, kPlanePlan>(buf, scr, tw, lane);
    __syncthreads();
    V* dst = static_cast<V*>(A.Z) + ((size_t)f * A.nyp + y) * MZ * A.nxh;
    for (int e = threadIdx.x; e < MZ * 8; e += 512) {
      const int z = e >> 3, c = e & 7;
      dst[(size_t)z * A.nxh + kx0 + c] = sm[c * BL + fpad(z)];
    }
    __syncthreads();
  }
}

template <int MZ, int NFIN, typename V>
__global__ __launch_bounds__(512) void k_zinv(PlaneArgs A) {
  zinv_body<MZ, NFIN, V>(A, blockIdx.x);
}
// channel blockIdx.y of an ensemble of one size (arguments AA[blockIdx.y])
template <int MZ, int NFIN, typename V>
__global__ __launch_bounds__(512) void k_zinv_batch(const PlaneArgs* __restrict__ AA) {
  zinv_body<MZ, NFIN, V>(AA[blockIdx.y], blockIdx.x);
}

template <int MX, typename V>
__global__ __launch_bounds__(256) void k_ibdrag(PlaneArgs A) {
  const int lane = threadIdx.x & 63;
  const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (s >= A.nsrc) return;
  const int4 c = A.src[s];
  const int t = c.w;
  const bool active = (t == 1) || (t == 4 && A.it <= A.perturbtime) || t == 3 || t == 6 || t == 7 || t == 8;
  if (t == 11) {  // constant x-force between the wall and the suction layer
    if (lane < 3) A.drag[lane * A.nsrc + s] = (lane == 0) ? 1.0 : 0.0;
    return;
  }
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
        const double2 tw = to_d(static_cast<const V*>(A.twX)[(int)(((long long)k * c.z) % MX)]);  // exp(-2 pi i m/MX)
        acc += 2.0 * (v.x * tw.x + v.y * tw.y);
      }
    }
    u3[d] = wave_sum(acc);
  }
  if (lane < 3) {
    const double u = (lane == 0) ? u3[0] : (lane == 1 ? u3[1] : u3[2]);
    const double* q = A.srcp + (size_t)6 * s;
    double* fi = A.fint + (size_t)lane * A.nsrc + s;
    double* dr = A.drag + (size_t)lane * A.nsrc + s;
    if (t == 1 || t == 4) {  // solid surface (solvers.f90:2996-3002)
      const double fnew = *fi + u;
      *fi = fnew;
      *dr = -A.ugain * u - A.gain * fnew;
    } else if (t == 3) {  // spanwise-damping texture (:3004-3006)
      if (lane == 2) {
        const double fnew = *fi + u;
        *fi = fnew;
        *dr = -A.ugain * u - A.gain * fnew;
      } else {
        *dr = 0.0;
      }
    } else if (t == 8) {  // spanwise moving wall at wdes (:3008-3021)
      const double e = (lane == 2) ? u - q[0] : u;
      const double fnew = *fi + e;
      *fi = fnew;
      *dr = -A.ugain * e - A.gain * fnew;
    } else if (t == 6) {  // buffer zone, force added at the cell (:3023-3031)
      const double e = u - (lane == 0 ? q[2] : 0.0);
      const double fnew = *fi + e;
      *fi = fnew;
      *dr = -q[0] * e - q[1] * fnew;
    } else {  // 7, suction region (:3033-3039): v to vdes(i), u to Uinf, w to initw
      const double e = u - (lane == 0 ? q[1] : (lane == 1 ? q[0] : 0.0));
      const double fnew = *fi + e;
      *fi = fnew;
      *dr = -A.gain * fnew - A.ugain * e;
    }
  }
}

// the two z rows (A, B) of one field, nxh spectral values each, NK per lane
template <int NK, typename V, typename T>
__device__ __forceinline__ void xrow_pair_load(T (&ax)[NK], T (&ay)[NK], T (&bx)[NK], T (&by)[NK],
                                               const V* __restrict__ ra, const V* __restrict__ rb, int nxh,
                                               int lane) {
#pragma unroll
  for (int q = 0; q < NK; ++q) {
    const int k = min(lane + 64 * q, nxh - 1);  // in range; lanes past nxh do not use theirs
    const V a = ra[k], b = rb[k];
    ax[q] = a.x; ay[q] = a.y; bx[q] = b.x; by[q] = b.y;
  }
}

// TV: occupancy variant picked by the autotuner (0: the compiler's register
// budget; 1: one more wave per SIMD than FP64/FP32 need at MX = 192, i.e.
// >= 3 (FP64) or >= 5 (FP32) waves/SIMD)
template <int MX, int NFIN, int NFOUT, int WPB, typename V>
__device__ __forceinline__ void xplane_body(const PlaneArgs& A, int bid) {
  using T = typename Scal<V>::T;
  constexpr int XPL = (MX + 63) / 64;  // x positions per lane: x = lane + 64 m
  extern __shared__ __align__(16) unsigned char smraw[];
  V* sm = reinterpret_cast<V*>(smraw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nzb = A.mz / (2 * WPB);
  const int y = bid / nzb;
  const int zA = (bid % nzb) * 2 * WPB + 2 * wv, zB = zA + 1;
  constexpr int BL = FFTBuf<MX>::len;
  V* tw = sm;                              // [MX] twiddles
  V* buf = sm + MX + (size_t)wv * 2 * BL;
  V* scr = buf + BL;
  for (int e = threadIdx.x; e < MX; e += WPB * 64) tw[e] = static_cast<const V*>(A.twX)[e];
  __syncthreads();
  const int nxh = A.nxh;
  // ph[f][m] = (row A, row B) physical values at x = lane + 64 m
  V ph[NFIN][XPL];
  // the next field's two rows are loaded into registers during this field's
  // transform (nxh = MX/3 values per row)
  constexpr int NK = (MX / 3 + 63) / 64;
  T pax[NK], pay[NK], pbx[NK], pby[NK];
  auto rows = [&](int f) {
    const V* zr = static_cast<const V*>(A.Z) + ((size_t)f * A.nyp + y) * A.mz * nxh;
    xrow_pair_load<NK, V, T>(pax, pay, pbx, pby, zr + (size_t)zA * nxh, zr + (size_t)zB * nxh, nxh, lane);
  };
  rows(0);
  // MX = 64 R with R the plan's last radix: the last stage of an inverse
  // transform leaves lane j the outputs x = j + 64 m, which can stay in
  // registers, and a forward transform can run the reversed plan, whose
  // first stage reads them there.  Measured at MX = 192: FP32 gains from the
  // inverse side (the forward side costs registers), FP64 from the forward
  // side (the inverse side takes it from 3 to 2 waves per SIMD).
  constexpr bool XOK = MX % 64 == 0 && plan_last_radix<MX, kPlanePlan>() == MX / 64;
  constexpr bool XINV = XOK && sizeof(V) == 8, XFWD = XOK && sizeof(V) == 16;
  // FP64: twiddles as powers of one table value per butterfly (fewer LDS
  // loads in this LDS-bound kernel, -3 % at MX = 192; no gain in FP32)
  constexpr bool XTW = sizeof(V) == 16;
  const V none[8] = {};
  V regs[8];
#pragma unroll
  for (int f = 0; f < NFIN; ++f) {
    // zeros at the dealiasing band [nxh, MX - nxh]; the packed rows fill the rest
    for (int x = nxh + lane; x <= MX - nxh; x += 64) buf[fpad(x)] = vzero<V>();
    wave_sync();
#pragma unroll
    for (int q = 0; q < NK; ++q) {
      const int k = lane + 64 * q;
      if (k >= nxh) continue;
      const V ua = mkv<V>(pax[q], pay[q]), ub = mkv<V>(pbx[q], pby[q]);
      if (k == 0) {
        buf[0] = mkv<V>(ua.x, ub.x);
      } else {
        buf[fpad(k)] = mkv<V>(ua.x - ub.y, ua.y + ub.x);
        buf[fpad(MX - k)] = mkv<V>(ua.x + ub.y, ub.x - ua.y);
      }
    }
    if (f + 1 < NFIN) rows(f + 1);
    wave_sync();
    if constexpr (XINV) {
      wave_fft_ex<MX, +1, V, kPlanePlan, false, false, true, XTW>(buf, scr, tw, lane, none, regs);
#pragma unroll
      for (int m = 0; m < XPL; ++m) ph[f][m] = regs[m];
    } else {
      V* r = wave_fft_ex<MX, +1, V, kPlanePlan, false, false, false, XTW>(buf, scr, tw, lane, none, regs);
#pragma unroll
      for (int m = 0; m < XPL; ++m) {
        const int x = lane + 64 * m;
        ph[f][m] = (x < MX) ? r[fpad(x)] : vzero<V>();
      }
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
  const T dpdx = (T)((A.flow_select == 1) ? (A.dpdx_d ? *A.dpdx_d : A.dPdx) : 0.0);
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
        if (A.scsrc && ok)  // vc += scsource/dt (solvers.f90:3061-3063)
          vc += (T)(A.scsrc[((size_t)y * A.mz + (rr ? zB : zA)) * MX + lane + 64 * m] / A.dt);
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
    V* r;
    if constexpr (XFWD) {
#pragma unroll
      for (int m = 0; m < XPL; ++m) regs[m] = ph[o][m];
      r = wave_fft_ex<MX, -1, V, kPlanePlan, true, true, false, XTW>(buf, scr, tw, lane, regs, regs);
    } else {
#pragma unroll
      for (int m = 0; m < XPL; ++m) {
        const int x = lane + 64 * m;
        if (x < MX) buf[fpad(x)] = ph[o][m];
      }
      wave_sync();
      r = wave_fft_ex<MX, -1, V, kPlanePlan, false, false, false, XTW>(buf, scr, tw, lane, none, regs);
    }
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

template <int MX, int NFIN, int NFOUT, int WPB, typename V, int TV = 0>
__global__ __launch_bounds__(WPB * 64) __attribute__((amdgpu_waves_per_eu(TV ? (sizeof(V) == 16 ? 3 : 5) : 1)))
void k_xplane(PlaneArgs A) {
  xplane_body<MX, NFIN, NFOUT, WPB, V>(A, blockIdx.x);
}
template <int MX, int NFIN, int NFOUT, int WPB, typename V, int TV = 0>
__global__ __launch_bounds__(WPB * 64) __attribute__((amdgpu_waves_per_eu(TV ? (sizeof(V) == 16 ? 3 : 5) : 1)))
void k_xplane_batch(const PlaneArgs* __restrict__ AA) {
  xplane_body<MX, NFIN, NFOUT, WPB, V>(AA[blockIdx.y], blockIdx.x);
}

// one block's [MZ][8] tile of a z-line group, TL values per thread (real
// and imaginary parts kept as scalars so that they stay in registers)
template <int MZ, int TL, typename V, typename T>
__device__ __forceinline__ void zfwd_tile_load(T (&px)[TL], T (&py)[TL], const V* __restrict__ srcp, int nxh,
                                               int kx0) {
#pragma unroll
  for (int q = 0; q < TL; ++q) {
    const int e = min((int)threadIdx.x + 512 * q, MZ * 8 - 1);  // in range; extra values are not stored
    const V v = srcp[(size_t)(e >> 3) * nxh + kx0 + (e & 7)];
    px[q] = v.x;
    py[q] = v.y;
  }
}

// Work items (output field, y, group of 8 kx lines) walked like k_zinv's.
template <int MZ, int NFOUT, typename V>
__device__ __forceinline__ void zfwd_body(const PlaneArgs& A, int bid) {
  using T = typename Scal<V>::T;
  extern __shared__ __align__(16) unsigned char smraw[];
  V* sm = reinterpret_cast<V*>(smraw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ngx = A.nxh / 8;
  const int ngroups = A.nyp * ngx;
  const int nitems = NFOUT * ngroups;
  constexpr int BL = FFTBuf<MZ>::len;
  V* buf = sm + wv * BL;
  V* scr = buf;  // in place
  V* twl = sm + 8 * BL;
  for (int e = threadIdx.x; e < MZ; e += 512) twl[e] = static_cast<const V*>(A.twZ)[e];
  __syncthreads();
  const V* tw = twl;
  int it = bid;
  if (it >= nitems) return;
  const T scale = (T)(1.0 / ((double)A.mx * (double)MZ));
  const int nz = A.nz, nzh = A.nzh;
  auto src = [&](int i, int& kx0) -> const V* {
    const int o = i / ngroups, g = i - o * ngroups;
    const int y = g / ngx;
    kx0 = (g - y * ngx) * 8;
    return static_cast<const V*>(A.ZF) + ((size_t)o * A.nyp + y) * MZ * A.nxh;
  };
  // the next item's tile is loaded into registers during this item's
  // transform
  constexpr int TL = (MZ * 8 + 511) / 512;
  T prx[TL], pry[TL];
  {
    int k0;
    const V* sp = src(it, k0);
    zfwd_tile_load<MZ, TL, V, T>(prx, pry, sp, A.nxh, k0);
  }
  for (; it < nitems; it += gridDim.x) {
    const int o = it / ngroups, g = it - o * ngroups;
    const int y = g / ngx, kx0 = (g - y * ngx) * 8;
    const int kx = kx0 + wv;
    // the block's [MZ][8] group goes straight into the waves' FFT buffers
    // (line c of the group to wave c)
#pragma unroll
    for (int q = 0; q < TL; ++q) {
      const int e = threadIdx.x + 512 * q;
      if (e < MZ * 8) sm[(e & 7) * BL + fpad(e >> 3)] = mkv<V>(prx[q], pry[q]);
    }
    __syncthreads();
    if (it + (int)gridDim.x < nitems) {
      int k0;
      const V* sp = src(it + gridDim.x, k0);
      zfwd_tile_load<MZ, TL, V, T>(prx, pry, sp, A.nxh, k0);
    }
    const V* r = wave_fft<MZ, -1, V, kPlanePlan>(buf, scr, tw, lane);
    // kept mode j of the line, scaled (zero at the z Nyquist mode)
    auto mode = [&](int j) -> V {
      const int jj = (j < nzh) ? j : (MZ - nz + j);
      return (j == nzh) ? vzero<V>() : cscale(r[fpad(jj)], scale);
    };
    V* dst = static_cast<V*>(A.out[o]) + ((size_t)y * A.nxh + kx) * nz;
    for (int j = lane; j < nz; j += 64) {
      V v = mode(j);
      if (kx == 0) {
        if (j == 0) v.y = 0;
        else if (j != nzh) v = cscale(cadd(v, cconj(mode(nz - j))), (T)0.5);
      }
      dst[j] = v;
    }
    __syncthreads();
  }
}

template <int MZ, int NFOUT, typename V>
__global__ __launch_bounds__(512) void k_zfwd(PlaneArgs A) {
  zfwd_body<MZ, NFOUT, V>(A, blockIdx.x);
}
template <int MZ, int NFOUT, typename V>
__global__ __launch_bounds__(512) void k_zfwd_batch(const PlaneArgs* __restrict__ AA) {
  zfwd_body<MZ, NFOUT, V>(AA[blockIdx.y], blockIdx.x);
}

// Grid of the item-walking z kernels: the blocks the device holds at once
// (occupancy x CUs, queried once per kernel instantiation and device), at
// most one per item.
struct ZGridCache {
  int dev = -1, blocks = 0;
};
template <class K>
static int zitems_grid(ZGridCache& cache, K kernel, size_t lds, int nitems) {
  int dev = 0;
  hipGetDevice(&dev);
  if (dev != cache.dev) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 512, lds) != hipSuccess || per_cu < 1)
      per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 1;
    cache.blocks = per_cu * cus;
    cache.dev = dev;
  }
  return nitems < cache.blocks ? nitems : cache.blocks;
}
template <int MZ, int NFIN, typename V>
static void launch_zinv(const PlaneArgs& a, hipStream_t s) {
  const size_t zlds = ZLds<MZ, V>::bytes;
  const int nitems = NFIN * a.nyp * (a.nxh / 8);
  static ZGridCache cache;
  const int grid = zitems_grid(cache, k_zinv<MZ, NFIN, V>, zlds, nitems);
  hipLaunchKernelGGL((k_zinv<MZ, NFIN, V>), dim3(grid), dim3(512), zlds, s, a);
}
template <int MZ, int NFOUT, typename V>
static void launch_zfwd(const PlaneArgs& a, hipStream_t s) {
  const size_t zlds = ZLds<MZ, V>::bytes;
  const int nitems = NFOUT * a.nyp * (a.nxh / 8);
  static ZGridCache cache;
  const int grid = zitems_grid(cache, k_zfwd<MZ, NFOUT, V>, zlds, nitems);
  hipLaunchKernelGGL((k_zfwd<MZ, NFOUT, V>), dim3(grid), dim3(512), zlds, s, a);
}

// ------------------------------------------------------------ launchers --
// ------------------------------------------------------------ FENE-P --
// One wavefront per (y, z) row.  Two real fields of the row share one
// complex x transform: spectra a_k, b_k (k < nxh) of the two fields are
// packed as a + i b with the Hermitian extension; after the inverse
// transform the real part is field a, the imaginary part field b.  With
// dxa/dxb the spectrum is first multiplied by i alpha k (d/dx).
// forward r2c of two real rows through one complex transform; writes the
// nxh kept modes of each (the 1/(mx mz) scale is applied in k_zfwd)
template <int MX, typename V>
__device__ __forceinline__ void xrow_pair_fwd(const typename Scal<V>::T (&a)[(MX + 63) / 64],
                                              const typename Scal<V>::T (&b)[(MX + 63) / 64], V* __restrict__ oa,
                                              V* __restrict__ ob, int nxh, V* buf, V* scr, const V* tw, int lane) {
  using T = typename Scal<V>::T;
  constexpr int XPL = (MX + 63) / 64;
#pragma unroll
  for (int m = 0; m < XPL; ++m) {
    const int x = lane + 64 * m;
    if (x < MX) buf[fpad(x)] = mkv<V>(a[m], b[m]);
  }
  wave_sync();
  // FP64: twiddles as powers of one table value per butterfly (fewer LDS loads)
  const V none[8] = {};
  V unused[8];
  V* r = wave_fft_ex<MX, -1, V, kPlanePlan, false, false, false, sizeof(V) == 16>(buf, scr, tw, lane, none, unused);
  for (int k = lane; k < nxh; k += 64) {
    const V zk = r[fpad(k)], zm = r[fpad((MX - k) % MX)];
    oa[k] = mkv<V>((T)0.5 * (zk.x + zm.x), (T)0.5 * (zk.y - zm.y));
    ob[k] = mkv<V>((T)0.5 * (zk.y + zm.y), (T)-0.5 * (zk.x - zm.x));
  }
  wave_sync();
}

// Spectral inputs of one pair transform, kept in registers so that the
// loads of the next pair overlap the current transform.
template <int MX, typename V>
struct PairIn {
  static constexpr int KPL = (MX / 3 + 63) / 64;  // nxh = MX/3 modes over 64 lanes
  V a[KPL], b[KPL];
};
template <int MX, typename V>
__device__ __forceinline__ void pair_load(PairIn<MX, V>& r, const V* __restrict__ ra, const V* __restrict__ rb,
                                          int nxh, int lane) {
#pragma unroll
  for (int j = 0; j < PairIn<MX, V>::KPL; ++j) {
    const int k = lane + 64 * j;
    if (k < nxh) {
      r.a[j] = ra[k];
      r.b[j] = rb[k];
    }
  }
}
template <int MX, typename V>
__device__ __forceinline__ void pair_inv(const PairIn<MX, V>& r, bool dxa, bool dxb, int nxh, double alpha, V* buf,
                                         V* scr, const V* tw, int lane, typename Scal<V>::T (&oa)[(MX + 63) / 64],
                                         typename Scal<V>::T (&ob)[(MX + 63) / 64]) {
  using T = typename Scal<V>::T;
  constexpr int XPL = (MX + 63) / 64;
  // zeros at the dealiasing band [nxh, MX - nxh]; the packed spectra fill the rest
  for (int x = nxh + lane; x <= MX - nxh; x += 64) buf[fpad(x)] = vzero<V>();
  wave_sync();
#pragma unroll
  for (int j = 0; j < PairIn<MX, V>::KPL; ++j) {
    const int k = lane + 64 * j;
    if (k < nxh) {
      V ua = r.a[j], ub = r.b[j];
      const T wk = (T)(alpha * (double)k);
      if (dxa) ua = mkv<V>(-ua.y * wk, ua.x * wk);
      if (dxb) ub = mkv<V>(-ub.y * wk, ub.x * wk);
      if (k == 0) {
        buf[0] = mkv<V>(ua.x, ub.x);
      } else {
        buf[fpad(k)] = mkv<V>(ua.x - ub.y, ua.y + ub.x);
        buf[fpad(MX - k)] = mkv<V>(ua.x + ub.y, ub.x - ua.y);
      }
    }
  }
  wave_sync();
  const V none[8] = {};
  V unused[8];
  V* res = wave_fft_ex<MX, +1, V, kPlanePlan, false, false, false, sizeof(V) == 16>(buf, scr, tw, lane, none, unused);
#pragma unroll
  for (int m = 0; m < XPL; ++m) {
    const int x = lane + 64 * m;
    const V v = (x < MX) ? res[fpad(x)] : vzero<V>();
    oa[m] = v.x;
    ob[m] = v.y;
  }
  wave_sync();
}

// The 18 pair transforms of a row: field ids >= 0 index the z-stage outputs
// (dz u 0-2, dy u 3-5, C 6-11, dz C 12-17, dy C 18-23), ids < 0 the velocity
// intermediates (-1 - f); dx = multiply by i alpha k first.
struct PairSpec {
  int fa, fb;
  bool dxa, dxb;
};

// Pointwise FENE-P terms of vcw3dp (solvers.f90:2901-2983) on one row:
//   c_ij^n = sum_k C_ik du_j/dx_k + sum_k C_kj du_i/dx_k - u . grad C_ij
//   str_ij = f C_ij - delta_ij,  f = (L^2-3)/(L^2-trC) (ipeter 1) or 1
// Field order of the independent components: 11, 12, 13, 22, 23, 33.
// s_ij += C_ik du_j/dx_k + C_kj du_i/dx_k for one gradient component
// gv = du_a/dx_k (all six independent (i,j); both terms when i = j = a)
template <typename T>
__device__ __forceinline__ void stretch_add(T (&s)[6], const T (&c)[6], int a, int k, T gv) {
  const int I[6] = {0, 0, 0, 1, 1, 2}, J[6] = {0, 1, 2, 1, 2, 2};
  // symmetric C_ik from the independent components
  auto C = [&](int i, int j) -> T {
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    return c[lo == 0 ? hi : (lo == 1 ? 2 + hi : 5)];
  };
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    if (J[e] == a) s[e] += C(I[e], k) * gv;
    if (I[e] == a) s[e] += C(k, J[e]) * gv;
  }
}

// Order chosen for register pressure: C, then the nine velocity-gradient
// components (accumulated straight into the stretching terms; C dies after
// them), u v w, the 18 dC components (advection), and C once more for the
// Peterlin stress after the c_n transforms have freed their registers.
constexpr int kNP = 21;
__host__ __device__ constexpr PairSpec poly_pair(int q) {
  constexpr PairSpec SP[kNP] = {
      {6, 7, false, false},   {8, 9, false, false},   {10, 11, false, false},   // C11 C12 | C13 C22 | C23 C33
      {-1, -2, true, true},   {-3, 3, true, false},   {4, 5, false, false},     // du/dx dv/dx | dw/dx du/dy | dv/dy dw/dy
      {0, 1, false, false},   {2, -1, false, false},  {-2, -3, false, false},   // du/dz dv/dz | dw/dz u | v w
      {6, 7, true, true},     {8, 9, true, true},     {10, 11, true, true},     // dC/dx
      {18, 19, false, false}, {20, 21, false, false}, {22, 23, false, false},   // dC/dy
      {12, 13, false, false}, {14, 15, false, false}, {16, 17, false, false},   // dC/dz
      {6, 7, false, false},   {8, 9, false, false},   {10, 11, false, false}};  // C again
  return SP[q];
}

template <int MX, typename V>
struct PolyRow {
  using T = typename Scal<V>::T;
  static constexpr int XPL = (MX + 63) / 64;
  T u[3][XPL], c[XPL][6], s[XPL][6];
  PairIn<MX, V> pin[2];
  const V* U;
  const V* ZP;
  V* ZF;
  size_t plane, roff;
  int nxh, lane, ipeter;
  double alpha;
  T L2;
  V *buf, *scr;
  const V* tw;
  float* capg;   // nullptr or [9][P] physical du_i/dx_k (FP32)
  float* caps;   // nullptr or [6][P] physical Peterlin stress (FP32)
  size_t P, prow;  // physical points, offset of this row
  __device__ const V* rowp(int id) const {
    return id < 0 ? U + (size_t)(-1 - id) * plane + roff : ZP + (size_t)id * plane + roff;
  }
  __device__ V* orow(int f) const { return ZF + (size_t)f * plane + roff; }
  __device__ void cap(float* base, int f, const T (&v)[XPL]) const {
#pragma unroll
    for (int m = 0; m < XPL; ++m) {
      const int x = lane + 64 * m;
      if (x < MX) base[(size_t)f * P + prow + x] = (float)v[m];
    }
  }

  template <int Q>
  __device__ __forceinline__ void step() {
    constexpr PairSpec sp = poly_pair(Q);
    if constexpr (Q + 1 < kNP) {
      constexpr PairSpec nx = poly_pair(Q + 1);
      pair_load<MX, V>(pin[(Q + 1) & 1], rowp(nx.fa), rowp(nx.fb), nxh, lane);
    }
    if constexpr (Q == 18) {  // c_n complete: forward transforms of its three pairs
      fwd_s<0>();
      fwd_s<2>();
      fwd_s<4>();
    }
    T oa[XPL], ob[XPL];
    pair_inv<MX, V>(pin[Q & 1], sp.dxa, sp.dxb, nxh, alpha, buf, scr, tw, lane, oa, ob);
    if (capg) {  // g[3 i + k] = du_i/dx_k
      if constexpr (Q == 3) { cap(capg, 0, oa); cap(capg, 3, ob); }
      if constexpr (Q == 4) { cap(capg, 6, oa); cap(capg, 1, ob); }
      if constexpr (Q == 5) { cap(capg, 4, oa); cap(capg, 7, ob); }
      if constexpr (Q == 6) { cap(capg, 2, oa); cap(capg, 5, ob); }
      if constexpr (Q == 7) { cap(capg, 8, oa); }
    }
#pragma unroll
    for (int m = 0; m < XPL; ++m) {
      if constexpr (Q == 0 || Q == 18) { c[m][0] = oa[m]; c[m][1] = ob[m]; }
      else if constexpr (Q == 1 || Q == 19) { c[m][2] = oa[m]; c[m][3] = ob[m]; }
      else if constexpr (Q == 2 || Q == 20) { c[m][4] = oa[m]; c[m][5] = ob[m]; }
      else if constexpr (Q == 3) { stretch_add(s[m], c[m], 0, 0, oa[m]); stretch_add(s[m], c[m], 1, 0, ob[m]); }
      else if constexpr (Q == 4) { stretch_add(s[m], c[m], 2, 0, oa[m]); stretch_add(s[m], c[m], 0, 1, ob[m]); }
      else if constexpr (Q == 5) { stretch_add(s[m], c[m], 1, 1, oa[m]); stretch_add(s[m], c[m], 2, 1, ob[m]); }
      else if constexpr (Q == 6) { stretch_add(s[m], c[m], 0, 2, oa[m]); stretch_add(s[m], c[m], 1, 2, ob[m]); }
      else if constexpr (Q == 7) { stretch_add(s[m], c[m], 2, 2, oa[m]); u[0][m] = ob[m]; }
      else if constexpr (Q == 8) { u[1][m] = oa[m]; u[2][m] = ob[m]; }
      else {
        // - u . grad C_ij: pairs 9-11 d/dx, 12-14 d/dy, 15-17 d/dz
        constexpr int dir = (Q - 9) / 3, cq = 2 * ((Q - 9) % 3);
        s[m][cq] -= u[dir][m] * oa[m];
        s[m][cq + 1] -= u[dir][m] * ob[m];
      }
    }
    // this pair's updates are done here: left free, the compiler defers them
    // past the later pairs' transforms and keeps every pending pair's values
    // live (482 registers per lane in FP64 at MX = 288 instead of 256)
#pragma unroll
    for (int m = 0; m < XPL; ++m)
#pragma unroll
      for (int e = 0; e < 6; ++e) asm volatile("" : "+v"(s[m][e])::"memory");
  }
  template <int E>
  __device__ __forceinline__ void fwd_s() {
    T sa[XPL], sb[XPL];
#pragma unroll
    for (int m = 0; m < XPL; ++m) { sa[m] = s[m][E]; sb[m] = s[m][E + 1]; }
    xrow_pair_fwd<MX, V>(sa, sb, orow(E), orow(E + 1), nxh, buf, scr, tw, lane);
  }
  // Peterlin stress str = f C - I and its forward transforms
  template <int E>
  __device__ __forceinline__ void fwd_str() {
    T sa[XPL], sb[XPL];
#pragma unroll
    for (int m = 0; m < XPL; ++m) {
      const T tr = c[m][0] + c[m][3] + c[m][5];
      const T fac = ipeter ? (L2 - (T)3) / (L2 - tr) : (T)1;
      sa[m] = fac * c[m][E] - ((E == 0) ? (T)1 : (T)0);
      sb[m] = fac * c[m][E + 1] - ((E + 1 == 3 || E + 1 == 5) ? (T)1 : (T)0);
    }
    if (caps) { cap(caps, E, sa); cap(caps, E + 1, sb); }
    xrow_pair_fwd<MX, V>(sa, sb, orow(6 + E), orow(7 + E), nxh, buf, scr, tw, lane);
  }
  template <int... Q>
  __device__ __forceinline__ void run(std::integer_sequence<int, Q...>) {
    (step<Q>(), ...);
  }
};

// Rows up to MX = 288 run two waves per SIMD (256 registers: FP64 spills 3
// values at MX = 288, FP32 none). At MX = 384, FP64 also runs two (92 bytes a
// lane spilled: 22 % faster than one wave per SIMD holding 280 registers) and
// FP32 three (36 bytes a lane spilled; 7 % faster than two).
template <int MX, typename V>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MX <= 288 ? 2 : (sizeof(V) == 8 ? 3 : 2))))
void k_xplane_poly(PolyPlaneArgs A) {
  using T = typename Scal<V>::T;
  constexpr int WPB = 4;
  extern __shared__ __align__(16) unsigned char smraw[];
  V* sm = reinterpret_cast<V*>(smraw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  constexpr int BL = FFTBuf<MX>::len;
  V* tw = sm;
  for (int e = threadIdx.x; e < MX; e += WPB * 64) tw[e] = static_cast<const V*>(A.twX)[e];
  __syncthreads();
  const int row = blockIdx.x * WPB + wv;  // grid covers nyp * mz rows exactly
  const int y = row / A.mz, z = row % A.mz;
  PolyRow<MX, V> R;
  R.U = static_cast<const V*>(A.U);
  R.ZP = static_cast<const V*>(A.ZP);
  R.ZF = static_cast<V*>(A.ZFP);
  R.nxh = A.nxh;
  R.plane = (size_t)A.nyp * A.mz * A.nxh;
  R.roff = ((size_t)y * A.mz + z) * A.nxh;
  R.lane = lane;
  R.ipeter = A.ipeter;
  R.alpha = A.alpha;
  R.L2 = (T)A.L2;
  R.buf = sm + MX + (size_t)wv * 2 * BL;
  R.scr = R.buf + BL;
  R.tw = tw;
  R.capg = A.capg;
  R.caps = A.caps;
  R.P = (size_t)A.nyp * A.mz * MX;
  R.prow = ((size_t)y * A.mz + z) * MX;
#pragma unroll
  for (int m = 0; m < PolyRow<MX, V>::XPL; ++m)
#pragma unroll
    for (int e = 0; e < 6; ++e) R.s[m][e] = (T)0;
  {
    constexpr PairSpec p0 = poly_pair(0);
    pair_load<MX, V>(R.pin[0], R.rowp(p0.fa), R.rowp(p0.fb), R.nxh, lane);
  }
  R.run(std::make_integer_sequence<int, kNP>{});
  R.template fwd_str<0>();
  R.template fwd_str<2>();
  R.template fwd_str<4>();
}

// Velocity gradient on one row of the 3/2 grid, for the closure features and
// the swirl strength (calcswirl, helpers.f90:360-404).  Four pair transforms:
// d/dx of u, v, w (the NS plane's z lines of u, v, w times i kx), the
// vorticity rows of the same NS z lines, and dz v, dz w (ZP rows gz, gz+1);
// the other components follow from vort (wx = Dw - i kz v, wy = i kz u -
// i kx w, wz = i kx v - Du) and continuity:
//   du/dy = dv/dx - wz,  du/dz = wy + dw/dx,  dv/dy = -du/dx - dw/dz,
//   dw/dy = wx + dv/dz.
template <int MX, typename V>
__global__ __launch_bounds__(256) void k_xplane_swirl(PolyPlaneArgs A) {
  using T = typename Scal<V>::T;
  constexpr int XPL = (MX + 63) / 64;
  constexpr int WPB = 4;
  extern __shared__ __align__(16) unsigned char smraw[];
  V* sm = reinterpret_cast<V*>(smraw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  constexpr int BL = FFTBuf<MX>::len;
  V* tw = sm;
  V* buf = sm + MX + (size_t)wv * 2 * BL;
  V* scr = buf + BL;
  for (int e = threadIdx.x; e < MX; e += WPB * 64) tw[e] = static_cast<const V*>(A.twX)[e];
  __syncthreads();
  const int row = blockIdx.x * WPB + wv;  // grid covers nyp * mz rows exactly
  const int y = row / A.mz, z = row % A.mz;
  const int nxh = A.nxh;
  const size_t plane = (size_t)A.nyp * A.mz * nxh;
  const size_t roff = ((size_t)y * A.mz + z) * nxh;
  const V* U = static_cast<const V*>(A.U);
  const V* ZP = static_cast<const V*>(A.ZP);
  auto urow = [&](int f) { return U + (size_t)f * plane + roff; };
  auto zrow = [&](int f) { return ZP + (size_t)f * plane + roff; };
  T g[9][XPL];  // g[3 i + k] = d u_i / d x_k
  T ox[XPL], oy[XPL], oz[XPL];
  PairIn<MX, V> pin;
  pair_load<MX, V>(pin, urow(0), urow(1), nxh, lane);
  pair_inv<MX, V>(pin, true, true, nxh, A.alpha, buf, scr, tw, lane, g[0], g[3]);
  pair_load<MX, V>(pin, urow(2), urow(3), nxh, lane);
  pair_inv<MX, V>(pin, true, false, nxh, A.alpha, buf, scr, tw, lane, g[6], ox);
  pair_load<MX, V>(pin, urow(4), urow(5), nxh, lane);
  pair_inv<MX, V>(pin, false, false, nxh, A.alpha, buf, scr, tw, lane, oy, oz);
  pair_load<MX, V>(pin, zrow(A.gz), zrow(A.gz + 1), nxh, lane);
  pair_inv<MX, V>(pin, false, false, nxh, A.alpha, buf, scr, tw, lane, g[5], g[8]);
#pragma unroll
  for (int m = 0; m < XPL; ++m) {
    g[1][m] = g[3][m] - oz[m];
    g[2][m] = oy[m] + g[6][m];
    g[4][m] = -g[0][m] - g[8][m];
    g[7][m] = ox[m] + g[5][m];
  }
  const size_t prow = ((size_t)y * A.mz + z) * MX;
  if (A.capg) {  // du_i/dx_k for the learned closure, [9][P] FP32
    const size_t P = (size_t)A.nyp * A.mz * MX;
#pragma unroll
    for (int k = 0; k < 9; ++k)
#pragma unroll
      for (int m = 0; m < XPL; ++m) {
        const int x = lane + 64 * m;
        if (x < MX) A.capg[(size_t)k * P + prow + x] = (float)g[k][m];
      }
  }
  if (A.mrf) {  // FP64 vorticity and gradient for inertial particles
    const size_t P = (size_t)A.nyp * A.mz * MX;
#pragma unroll
    for (int m = 0; m < XPL; ++m) {
      const int x = lane + 64 * m;
      if (x < MX) {
        A.mrf[prow + x] = (double)ox[m];
        A.mrf[P + prow + x] = (double)oy[m];
        A.mrf[2 * P + prow + x] = (double)oz[m];
#pragma unroll
        for (int k = 0; k < 9; ++k) A.mrf[(size_t)(3 + k) * P + prow + x] = (double)g[k][m];
      }
    }
  }
  if (A.swirl) {
    double* out = A.swirl + prow;
#pragma unroll
    for (int m = 0; m < XPL; ++m) {
      const int x = lane + 64 * m;
      double a[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) a[k] = (double)g[k][m];
      if (x < MX) out[x] = swirl_of(a);
    }
  }
}

// x inverse of the three Laplacian z lines of a row (LZ fields 0..2) into
// FP64 physical rows out[3][P]; the third field is paired with itself
template <int MX, typename V>
__global__ __launch_bounds__(256) void k_mr_xinv(MrLapArgs A) {
  using T = typename Scal<V>::T;
  constexpr int XPL = (MX + 63) / 64;
  constexpr int WPB = 4;
  extern __shared__ __align__(16) unsigned char smraw[];
  V* sm = reinterpret_cast<V*>(smraw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  constexpr int BL = FFTBuf<MX>::len;
  V* tw = sm;
  V* buf = sm + MX + (size_t)wv * 2 * BL;
  V* scr = buf + BL;
  for (int e = threadIdx.x; e < MX; e += WPB * 64) tw[e] = static_cast<const V*>(A.twX)[e];
  __syncthreads();
  const int row = blockIdx.x * WPB + wv;  // grid covers nyp * mz rows exactly
  const int y = row / A.mz, z = row % A.mz;
  const int nxh = A.nxh;
  const size_t plane = (size_t)A.nyp * A.mz * nxh;
  const size_t roff = ((size_t)y * A.mz + z) * nxh;
  const V* LZ = static_cast<const V*>(A.LZ);
  T a[XPL], b[XPL], c[XPL], d[XPL];
  PairIn<MX, V> pin;
  pair_load<MX, V>(pin, LZ + roff, LZ + plane + roff, nxh, lane);
  pair_inv<MX, V>(pin, false, false, nxh, A.alpha, buf, scr, tw, lane, a, b);
  pair_load<MX, V>(pin, LZ + 2 * plane + roff, LZ + 2 * plane + roff, nxh, lane);
  pair_inv<MX, V>(pin, false, false, nxh, A.alpha, buf, scr, tw, lane, c, d);
  const size_t P = (size_t)A.nyp * A.mz * MX;
  const size_t prow = ((size_t)y * A.mz + z) * MX;
#pragma unroll
  for (int m = 0; m < XPL; ++m) {
    const int x = lane + 64 * m;
    if (x < MX) {
      A.out[prow + x] = (double)a[m];
      A.out[P + prow + x] = (double)b[m];
      A.out[2 * P + prow + x] = (double)c[m];
    }
  }
}

// forward x transforms of the six closure stress rows (FP32 [6][P]) into the
// z-stage inputs ZFP[6..11]
template <int MX, typename V>
__global__ __launch_bounds__(256) void k_xrow_fwd_str(PolyPlaneArgs A) {
  using T = typename Scal<V>::T;
  constexpr int XPL = (MX + 63) / 64;
  constexpr int WPB = 4;
  extern __shared__ __align__(16) unsigned char smraw[];
  V* sm = reinterpret_cast<V*>(smraw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  constexpr int BL = FFTBuf<MX>::len;
  V* tw = sm;
  V* buf = sm + MX + (size_t)wv * 2 * BL;
  V* scr = buf + BL;
  for (int e = threadIdx.x; e < MX; e += WPB * 64) tw[e] = static_cast<const V*>(A.twX)[e];
  __syncthreads();
  const int row = blockIdx.x * WPB + wv;
  const int y = row / A.mz, z = row % A.mz;
  const size_t P = (size_t)A.nyp * A.mz * MX, prow = ((size_t)y * A.mz + z) * MX;
  const size_t plane = (size_t)A.nyp * A.mz * A.nxh, roff = ((size_t)y * A.mz + z) * A.nxh;
  V* ZF = static_cast<V*>(A.ZFP);
#pragma unroll
  for (int e = 0; e < 6; e += 2) {
    T sa[XPL], sb[XPL];
#pragma unroll
    for (int m = 0; m < XPL; ++m) {
      const int x = lane + 64 * m;
      sa[m] = (x < MX) ? (T)A.caps[(size_t)e * P + prow + x] : (T)0;
      sb[m] = (x < MX) ? (T)A.caps[(size_t)(e + 1) * P + prow + x] : (T)0;
    }
    xrow_pair_fwd<MX, V>(sa, sb, ZF + (size_t)(6 + e) * plane + roff, ZF + (size_t)(7 + e) * plane + roff, A.nxh,
                         buf, scr, tw, lane);
  }
}

template <int MX, int MZ, typename V>
static hipError_t closure_fwd_impl(const PolyPlaneArgs& pa, hipStream_t s) {
  const size_t xlds = (size_t)(MX + 4 * 2 * FFTBuf<MX>::len) * sizeof(V);
  hipLaunchKernelGGL((k_xrow_fwd_str<MX, V>), dim3(pa.nyp * pa.mz / 4), dim3(256), xlds, s, pa);
  PlaneArgs a;
  std::memset(&a, 0, sizeof(a));
  a.nxh = pa.nxh; a.nz = pa.nz; a.nzh = pa.nzh; a.nyp = pa.nyp; a.mx = pa.mx; a.mz = pa.mz;
  a.twZ = pa.twZ; a.twX = pa.twX;
  const size_t zb = (size_t)pa.nyp * pa.mz * pa.nxh;
  for (int g = 0; g < 2; ++g) {
    a.ZF = static_cast<V*>(pa.ZFP) + (size_t)(6 + 3 * g) * zb;
    for (int o = 0; o < 3; ++o) a.out[o] = pa.out[6 + 3 * g + o];
    launch_zfwd<MZ, 3, V>(a, s);
  }
  return hipGetLastError();
}

template <int MX, int MZ, typename V>
static hipError_t swirl_plane_impl(const PolyPlaneArgs& pa, bool zstage, hipStream_t s) {
  if (zstage) {
    PlaneArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nxh = pa.nxh; a.nz = pa.nz; a.nzh = pa.nzh; a.nyp = pa.nyp; a.mx = pa.mx; a.mz = pa.mz;
    a.twZ = pa.twZ; a.twX = pa.twX; a.beta = pa.beta;
    for (int f = 0; f < 2; ++f) a.in[f] = pa.in[f];
    a.dzmask = 0x3u;  // dz v, dz w into ZP rows 0, 1
    a.Z = pa.ZP;
    launch_zinv<MZ, 2, V>(a, s);
  }
  const size_t xlds = (size_t)(MX + 4 * 2 * FFTBuf<MX>::len) * sizeof(V);
  hipLaunchKernelGGL((k_xplane_swirl<MX, V>), dim3(pa.nyp * pa.mz / 4), dim3(256), xlds, s, pa);
  return hipGetLastError();
}

template <int MX, int MZ, typename V>
static hipError_t mr_lap_plane_impl(const MrLapArgs& la, hipStream_t s) {
  PlaneArgs a;
  std::memset(&a, 0, sizeof(a));
  a.nxh = la.nxh; a.nz = la.nz; a.nzh = la.nzh; a.nyp = la.nyp; a.mx = la.mx; a.mz = la.mz;
  a.twZ = la.twZ; a.twX = la.twX; a.beta = la.beta;
  for (int f = 0; f < 3; ++f) a.in[f] = la.LT[f];
  a.dzmask = 0u;
  a.Z = la.LZ;
  launch_zinv<MZ, 3, V>(a, s);
  const size_t xlds = (size_t)(MX + 4 * 2 * FFTBuf<MX>::len) * sizeof(V);
  hipLaunchKernelGGL((k_mr_xinv<MX, V>), dim3(la.nyp * la.mz / 4), dim3(256), xlds, s, la);
  return hipGetLastError();
}

template <int MX, int MZ, typename V>
static hipError_t poly_plane_impl(const PolyPlaneArgs& pa, hipStream_t s) {
  // z inverse of the 24 polymer inputs (8 per launch, d/dz by mask)
  PlaneArgs a;
  std::memset(&a, 0, sizeof(a));
  a.nxh = pa.nxh; a.nz = pa.nz; a.nzh = pa.nzh; a.nyp = pa.nyp; a.mx = pa.mx; a.mz = pa.mz;
  a.twZ = pa.twZ; a.twX = pa.twX; a.beta = pa.beta;
  const size_t zb = (size_t)pa.nyp * pa.mz * pa.nxh;
  // ZP layout: dz u(3), dy u(3), C(6), dz C(6), dy C(6); the dz fields are
  // [0..2] and [12..17]
  for (int g = 0; g < 3; ++g) {
    a.dzmask = 0;
    for (int f = 0; f < 8; ++f) {
      const int o = 8 * g + f;
      a.in[f] = pa.in[o];
      if (o < 3 || (o >= 12 && o < 18)) a.dzmask |= 1u << f;
    }
    a.Z = static_cast<V*>(pa.ZP) + (size_t)8 * g * zb;
    launch_zinv<MZ, 8, V>(a, s);
  }
  const size_t xlds = (size_t)(MX + 4 * 2 * FFTBuf<MX>::len) * sizeof(V);
  hipLaunchKernelGGL((k_xplane_poly<MX, V>), dim3(pa.nyp * pa.mz / 4), dim3(256), xlds, s, pa);
  // forward z of the 12 outputs, 4 per launch
  for (int g = 0; g < 3; ++g) {
    a.ZF = static_cast<V*>(pa.ZFP) + (size_t)4 * g * zb;
    for (int o = 0; o < 4; ++o) a.out[o] = pa.out[4 * g + o];
    launch_zfwd<MZ, 4, V>(a, s);
  }
  return hipGetLastError();
}

template <int MX, int MZ, int NFIN, int NFOUT, typename V>
static hipError_t plane_impl(const PlaneArgs& a, hipStream_t s, bool inverse_part, hipEvent_t* ev) {
  if (inverse_part) {
    launch_zinv<MZ, NFIN, V>(a, s);
    if (ev) hipEventRecord(ev[0], s);
    if (a.nsrc > 0)
      hipLaunchKernelGGL((k_ibdrag<MX, V>), dim3((a.nsrc + 3) / 4), dim3(256), 0, s, a);
    if (ev) hipEventRecord(ev[1], s);
    constexpr int WPB = 4;
    const size_t xlds = (size_t)(MX + WPB * 2 * FFTBuf<MX>::len) * sizeof(V);
    if (a.tune == 1)
      hipLaunchKernelGGL((k_xplane<MX, NFIN, NFOUT, WPB, V, 1>), dim3(a.nyp * (a.mz / (2 * WPB))), dim3(WPB * 64),
                         xlds, s, a);
    else
      hipLaunchKernelGGL((k_xplane<MX, NFIN, NFOUT, WPB, V, 0>), dim3(a.nyp * (a.mz / (2 * WPB))), dim3(WPB * 64),
                         xlds, s, a);
    if (ev) hipEventRecord(ev[2], s);
  }
  launch_zfwd<MZ, NFOUT, V>(a, s);
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
  KMM_PLANE_CASE(384, 384)
#undef KMM_PLANE_CASE
  return hipErrorInvalidValue;
}

// The plane stage of nmem channels of one size: each kernel's grid gains a
// y dimension over the channels; the item-walking z kernels share the
// blocks the device holds at once among the channels.
template <int MX, int MZ, int NFIN, int NFOUT, typename V>
static hipError_t plane_batch_impl(const PlaneArgs& a, const PlaneArgs* dA, int nmem, hipStream_t s) {
  const size_t zlds = ZLds<MZ, V>::bytes;
  {
    const int nitems = NFIN * a.nyp * (a.nxh / 8);
    static ZGridCache cache;
    const int cap = zitems_grid(cache, k_zinv_batch<MZ, NFIN, V>, zlds, 1 << 30);
    const int grid = std::min(nitems, std::max(1, (cap + nmem - 1) / nmem));
    hipLaunchKernelGGL((k_zinv_batch<MZ, NFIN, V>), dim3(grid, nmem), dim3(512), zlds, s, dA);
  }
  constexpr int WPB = 4;
  const size_t xlds = (size_t)(MX + WPB * 2 * FFTBuf<MX>::len) * sizeof(V);
  const dim3 gx(a.nyp * (a.mz / (2 * WPB)), nmem);
  if (a.tune == 1)
    hipLaunchKernelGGL((k_xplane_batch<MX, NFIN, NFOUT, WPB, V, 1>), gx, dim3(WPB * 64), xlds, s, dA);
  else
    hipLaunchKernelGGL((k_xplane_batch<MX, NFIN, NFOUT, WPB, V, 0>), gx, dim3(WPB * 64), xlds, s, dA);
  {
    const int nitems = NFOUT * a.nyp * (a.nxh / 8);
    static ZGridCache cache;
    const int cap = zitems_grid(cache, k_zfwd_batch<MZ, NFOUT, V>, zlds, 1 << 30);
    const int grid = std::min(nitems, std::max(1, (cap + nmem - 1) / nmem));
    hipLaunchKernelGGL((k_zfwd_batch<MZ, NFOUT, V>), dim3(grid, nmem), dim3(512), zlds, s, dA);
  }
  return hipGetLastError();
}

template <int MX, int MZ>
static hipError_t plane_batch_f(const PlaneArgs& a, const PlaneArgs* dA, int nmem, hipStream_t s) {
  if (a.mixed) {
    if (a.nfin == 6 && a.nfout == 3) return plane_batch_impl<MX, MZ, 6, 3, float2>(a, dA, nmem, s);
    if (a.nfin == 9 && a.nfout == 4) return plane_batch_impl<MX, MZ, 9, 4, float2>(a, dA, nmem, s);
  } else {
    if (a.nfin == 6 && a.nfout == 3) return plane_batch_impl<MX, MZ, 6, 3, double2>(a, dA, nmem, s);
    if (a.nfin == 9 && a.nfout == 4) return plane_batch_impl<MX, MZ, 9, 4, double2>(a, dA, nmem, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_plane_batch(const PlaneArgs& a, const PlaneArgs* dA, int nmem, hipStream_t s) {
  if (a.nxh % 8 != 0 || a.nsrc > 0 || nmem < 1) return hipErrorInvalidValue;
#define KMM_PLANE_CASE(X, Zz) \
  if (a.mx == X && a.mz == Zz) return plane_batch_f<X, Zz>(a, dA, nmem, s);
  KMM_PLANE_CASE(48, 48)
  KMM_PLANE_CASE(96, 48)
  KMM_PLANE_CASE(96, 96)
  KMM_PLANE_CASE(144, 144)
  KMM_PLANE_CASE(192, 96)
  KMM_PLANE_CASE(192, 192)
  KMM_PLANE_CASE(288, 288)
  KMM_PLANE_CASE(384, 192)
  KMM_PLANE_CASE(384, 384)
#undef KMM_PLANE_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_poly_plane(const PolyPlaneArgs& a, hipStream_t s) {
  if (a.nxh % 8 != 0 || (a.nyp * a.mz) % 4 != 0) return hipErrorInvalidValue;
#define KMM_POLY_CASE(X, Zz)                                                              \
  if (a.mx == X && a.mz == Zz)                                                            \
    return a.mixed ? poly_plane_impl<X, Zz, float2>(a, s) : poly_plane_impl<X, Zz, double2>(a, s);
  KMM_POLY_CASE(48, 48)
  KMM_POLY_CASE(96, 48)
  KMM_POLY_CASE(96, 96)
  KMM_POLY_CASE(144, 144)
  KMM_POLY_CASE(192, 96)
  KMM_POLY_CASE(192, 192)
  KMM_POLY_CASE(288, 288)
  KMM_POLY_CASE(384, 192)
  KMM_POLY_CASE(384, 384)
#undef KMM_POLY_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_swirl_plane(const PolyPlaneArgs& a, bool zstage, hipStream_t s) {
  if (a.nxh % 8 != 0 || (a.nyp * a.mz) % 4 != 0) return hipErrorInvalidValue;
#define KMM_SWIRL_CASE(X, Zz)                                                                          \
  if (a.mx == X && a.mz == Zz)                                                                         \
    return a.mixed ? swirl_plane_impl<X, Zz, float2>(a, zstage, s) : swirl_plane_impl<X, Zz, double2>(a, zstage, s);
  KMM_SWIRL_CASE(48, 48)
  KMM_SWIRL_CASE(96, 48)
  KMM_SWIRL_CASE(96, 96)
  KMM_SWIRL_CASE(144, 144)
  KMM_SWIRL_CASE(192, 96)
  KMM_SWIRL_CASE(192, 192)
  KMM_SWIRL_CASE(288, 288)
  KMM_SWIRL_CASE(384, 192)
  KMM_SWIRL_CASE(384, 384)
#undef KMM_SWIRL_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_mr_lap_plane(const MrLapArgs& a, hipStream_t s) {
  if (a.nxh % 8 != 0 || (a.nyp * a.mz) % 4 != 0) return hipErrorInvalidValue;
#define KMM_MRL_CASE(X, Zz)                                                                            \
  if (a.mx == X && a.mz == Zz)                                                                         \
    return a.mixed ? mr_lap_plane_impl<X, Zz, float2>(a, s) : mr_lap_plane_impl<X, Zz, double2>(a, s);
  KMM_MRL_CASE(48, 48)
  KMM_MRL_CASE(96, 48)
  KMM_MRL_CASE(96, 96)
  KMM_MRL_CASE(144, 144)
  KMM_MRL_CASE(192, 96)
  KMM_MRL_CASE(192, 192)
  KMM_MRL_CASE(288, 288)
  KMM_MRL_CASE(384, 192)
  KMM_MRL_CASE(384, 384)
#undef KMM_MRL_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_closure_fwd(const PolyPlaneArgs& a, hipStream_t s) {
  if (a.nxh % 8 != 0 || (a.nyp * a.mz) % 4 != 0) return hipErrorInvalidValue;
#define KMM_CLO_CASE(X, Zz) \
  if (a.mx == X && a.mz == Zz) return a.mixed ? closure_fwd_impl<X, Zz, float2>(a, s) : closure_fwd_impl<X, Zz, double2>(a, s);
  KMM_CLO_CASE(48, 48)
  KMM_CLO_CASE(96, 48)
  KMM_CLO_CASE(96, 96)
  KMM_CLO_CASE(144, 144)
  KMM_CLO_CASE(192, 96)
  KMM_CLO_CASE(192, 192)
  KMM_CLO_CASE(288, 288)
  KMM_CLO_CASE(384, 192)
  KMM_CLO_CASE(384, 384)
#undef KMM_CLO_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_plane(const PlaneArgs& a, hipStream_t s) { return plane_dispatch(a, s, true, nullptr); }
hipError_t launch_plane_timed(const PlaneArgs& a, hipStream_t s, hipEvent_t* ev) {
  return plane_dispatch(a, s, true, ev);
}
hipError_t launch_plane_fwd_only(const PlaneArgs& a, hipStream_t s) { return plane_dispatch(a, s, false, nullptr); }

}  // namespace kmm
