// Audit steps of the learned closure (no reference counterpart): the
// conformation tensor rebuilt from a stress field, and the residual sums that
// gate the closure against the full FENE-P transport.
//
//   k_conf_from_stress  inverse of the FENE-P stress map str = f C - I,
//                       f = (L^2 - 3)/(L^2 - tr C) (ipeter 1) or 1
//                       (the forward map, solvers.f90:2925-2949): with
//                       s = tr str, f = (L^2 + s)/L^2 and C = (str + I)/f.
//                       Pointwise over the 3/2 grid, FP32 in and out;
//                       optionally from the one-step extrapolation
//                       2 str(n-1) - str(n-2).
//   k_audit_partial     per-block sums of (pred - ref)^2 and ref^2 (FP64),
//   k_audit_final       summed in a fixed order and added to acc[0], acc[1].
#include "kmm_kernels.h"

namespace kmm {

// component order 11 12 13 22 23 33; diagonal entries 0, 3, 5
// With prev != nullptr the stress is first extrapolated one step ahead,
// 2 str - prev.
__global__ void k_conf_from_stress(const float* __restrict__ str, const float* __restrict__ prev,
                                   float* __restrict__ C, size_t P, float L2, int ipeter) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  float s[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) s[q] = str[(size_t)q * P + p];
  if (prev) {
#pragma unroll
    for (int q = 0; q < 6; ++q) s[q] = 2.0f * s[q] - prev[(size_t)q * P + p];
  }
  const float tr = s[0] + s[3] + s[5];
  const float inv = ipeter ? L2 / (L2 + tr) : 1.0f;  // 1/f
#pragma unroll
  for (int q = 0; q < 6; ++q) C[(size_t)q * P + p] = (s[q] + ((q == 0 || q == 3 || q == 5) ? 1.0f : 0.0f)) * inv;
}

hipError_t launch_conf_from_stress(const float* str, const float* prev, float* C, size_t P, double L2, int ipeter,
                                  hipStream_t s) {
  if (P == 0) return hipSuccess;
  hipLaunchKernelGGL(k_conf_from_stress, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, str, prev, C, P,
                     (float)L2, ipeter);
  return hipGetLastError();
}

constexpr int kAuditBlocks = 1024;

__global__ __launch_bounds__(256) void k_audit_partial(const float* __restrict__ pred, const float* __restrict__ ref,
                                                       size_t n, double* __restrict__ part) {
  __shared__ double se[256], ss[256];
  double e = 0.0, r = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double d = (double)pred[i] - (double)ref[i];
    e += d * d;
    r += (double)ref[i] * (double)ref[i];
  }
  se[threadIdx.x] = e;
  ss[threadIdx.x] = r;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      se[threadIdx.x] += se[threadIdx.x + h];
      ss[threadIdx.x] += ss[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = se[0];
    part[2 * blockIdx.x + 1] = ss[0];
  }
}

__global__ void k_audit_final(const double* __restrict__ part, int nb, double* __restrict__ acc) {
  if (threadIdx.x != 0) return;
  double e = 0.0, r = 0.0;
  for (int b = 0; b < nb; ++b) {
    e += part[2 * b];
    r += part[2 * b + 1];
  }
  acc[0] += e;
  acc[1] += r;
}

int audit_part_count() { return kAuditBlocks; }

hipError_t launch_audit_residual(const float* pred, const float* ref, size_t n, double* part, double* acc,
                                 hipStream_t s) {
  hipLaunchKernelGGL(k_audit_partial, dim3(kAuditBlocks), dim3(256), 0, s, pred, ref, n, part);
  hipLaunchKernelGGL(k_audit_final, dim3(1), dim3(64), 0, s, part, kAuditBlocks, acc);
  return hipGetLastError();
}

}  // namespace kmm
