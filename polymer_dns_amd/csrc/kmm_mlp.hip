// Learned polymer-stress closure: a 22 -> 4 x 128 (GELU) -> 6 multilayer
// perceptron with bfloat16 weights and FP32 accumulation (no reference
// counterpart).
//
//   k_mlp_infer   inference on the matrix cores: v_mfma_f32_16x16x32_bf16,
//                 one wavefront per tile of 16 points, all bf16 weights staged
//                 in LDS (rows padded by 8 to spread banks), activations of
//                 a tile transposed through a per-wave LDS buffer between
//                 layers.
//   k_mlp_fb      training forward + backward on FP32 VALUs, 16 samples per
//                 block; writes per-block gradient partials (deterministic
//                 reduction afterwards, no atomics).  mode 1 rounds weights
//                 and layer inputs to bf16 exactly where inference does;
//                 mode 0 is plain FP32 (gradient checks).
//   k_mlp_reduce  sums the partials in a fixed order;  k_mlp_adam applies
//                 Adam to the FP32 master parameters and refreshes the bf16
//                 copies.
//
// Parameter vector (FP32 master, `kMlpParams` floats): for l = 0..4,
// W_l (n_out x n_in row-major) then b_l; sizes 22,128,128,128,128,6.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "kmm_kernels.h"
#include "kmm_device.h"

namespace kmm {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 22, kH = 128, kOut = 6, kNL = 5;
__host__ __device__ constexpr int lay_in(int l) { return l == 0 ? kIn : kH; }
__host__ __device__ constexpr int lay_out(int l) { return l == kNL - 1 ? kOut : kH; }
__host__ __device__ constexpr int w_off(int l) {
  int o = 0;
  for (int i = 0; i < l; ++i) o += lay_out(i) * lay_in(i) + lay_out(i);
  return o;
}
__host__ __device__ constexpr int b_off(int l) { return w_off(l) + lay_out(l) * lay_in(l); }
constexpr int kParams = w_off(kNL);

// padded bf16 layout for the matrix cores: K padded to 32, N to 16, rows of
// K + 8 elements
__host__ __device__ constexpr int kin_pad(int l) { return l == 0 ? 32 : kH; }
__host__ __device__ constexpr int nout_pad(int l) { return l == kNL - 1 ? 16 : kH; }
__host__ __device__ constexpr int wrow(int l) { return kin_pad(l) + 8; }
__host__ __device__ constexpr int wb_off(int l) {
  int o = 0;
  for (int i = 0; i < l; ++i) o += nout_pad(i) * wrow(i);
  return o;
}
constexpr int kWb = wb_off(kNL);  // bf16 elements of the padded weights
constexpr int kBias = 4 * kH + 16;

__device__ __forceinline__ float gelu_f(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_df(float x) {
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * __expf(-0.5f * x * x);
}
// round to bfloat16 (nearest even) and back
__device__ __forceinline__ float bf16r(float x) {
  unsigned u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return x;
  u = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
  return __uint_as_float(u);
}

}  // namespace

int mlp_param_count() { return kParams; }
int mlp_wb_count() { return kWb; }

// -------------------------------------------------------------- inference --
// feature k of point p from field-major captures (learned-closure layout)
__device__ __forceinline__ float field_feature(const MlpFieldIO& io, size_t p, int k) {
  if (k < 9) return io.g[(size_t)k * io.P + p];
  if (k == 9) return (float)(1.0 - fabs(io.ycoord[p / (size_t)io.plane_pts]) / io.half);
  if (k < 16) return io.h1[(size_t)(k - 10) * io.P + p];
  // the last step's change, formed in FP32 before the bf16 rounding of the inputs
  return io.h1[(size_t)(k - 16) * io.P + p] - io.h2[(size_t)(k - 16) * io.P + p];
}

// FIELD = false: X[n][22] -> Y[n][6]; FIELD = true: io fields -> io.out[6][P]
template <bool FIELD>
__global__ __launch_bounds__(256) void k_mlp_infer(int n, const float* __restrict__ X, float* __restrict__ Y,
                                                   MlpFieldIO io, const __bf16* __restrict__ wb,
                                                   const float* __restrict__ par, const float* __restrict__ norm) {
  extern __shared__ __align__(16) unsigned char smraw[];
  __bf16* sw = reinterpret_cast<__bf16*>(smraw);                 // [kWb]
  float* sb = reinterpret_cast<float*>(sw + kWb);                 // biases [kBias]
  __bf16* sact = reinterpret_cast<__bf16*>(sb + kBias);          // per wave [16][kH + 8]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int e = tid; e < kWb; e += blockDim.x) sw[e] = wb[e];
  // biases: layer l at 128 l (the last layer's 16 padded entries at 512)
  for (int l = 0; l < kNL; ++l)
    for (int j = tid; j < nout_pad(l); j += blockDim.x) sb[128 * l + j] = (j < lay_out(l)) ? par[b_off(l) + j] : 0.0f;
  __syncthreads();
  __bf16* act = sact + (size_t)wv * 16 * (kH + 8);
  const float* mu_x = norm;
  const float* sd_x = norm + kIn;
  const float* mu_y = norm + 2 * kIn;
  const float* sd_y = norm + 2 * kIn + kOut;
  const int r16 = lane & 15, kq = lane >> 4;  // operand row / k-group
  const int ntiles = (n + 15) / 16;
  const int wstride = gridDim.x * (blockDim.x >> 6);
  for (int t = blockIdx.x * (blockDim.x >> 6) + wv; t < ntiles; t += wstride) {
    const int p0 = t * 16;
    // input tile: 16 points x 32 (22 used), normalised, bf16; field inputs
    // are read with consecutive lanes on consecutive points
    for (int e = lane; e < 16 * 32; e += 64) {
      int r, k;
      if (FIELD) { k = e >> 4; r = e & 15; }
      else { r = e >> 5; k = e & 31; }
      float v = 0.0f;
      if (k < kIn && p0 + r < n) {
        const float xv = FIELD ? field_feature(io, (size_t)(p0 + r), k) : X[(size_t)(p0 + r) * kIn + k];
        v = (xv - mu_x[k]) / sd_x[k];
      }
      act[r * (kH + 8) + k] = (__bf16)v;
    }
    wave_sync();
#pragma unroll 1
    for (int l = 0; l < kNL; ++l) {
      const int K = kin_pad(l), NT = nout_pad(l) / 16;
      const __bf16* W = sw + wb_off(l);
      const int wr = wrow(l);
      f32x4 acc[8];
#pragma unroll
      for (int nt = 0; nt < 8; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
      for (int ks = 0; ks < K / 32; ++ks) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(act + r16 * (kH + 8) + ks * 32 + kq * 8);
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) {
          if (nt < NT) {
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(W + (nt * 16 + r16) * wr + ks * 32 + kq * 8);
            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[nt], 0, 0, 0);
          }
        }
      }
      wave_sync();
      // accumulator layout: lane -> column lane%16 of the tile, rows 4 (lane/16) + i
      if (l < kNL - 1) {
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) {
          const int col = nt * 16 + r16;
          const float bias = sb[128 * l + col];
#pragma unroll
          for (int i = 0; i < 4; ++i) act[(4 * kq + i) * (kH + 8) + col] = (__bf16)gelu_f(acc[nt][i] + bias);
        }
      } else {
        const int col = r16;
        if (col < kOut) {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int p = p0 + 4 * kq + i;
            if (p < n) {
              const float yv = (acc[0][i] + sb[512 + col]) * sd_y[col] + mu_y[col];
              // the closure network predicts the correction to linear extrapolation
              // of the two previous steps: out = 2 h1 - h2 + y
              if (FIELD) {
                const size_t q = (size_t)col * io.P + p;
                io.out[q] = io.h1[q] + (io.h1[q] - io.h2[q]) + yv;
              }
              else Y[(size_t)p * kOut + col] = yv;
            }
          }
        }
      }
      wave_sync();
    }
  }
}

size_t mlp_infer_lds() { return (size_t)kWb * 2 + (size_t)kBias * 4 + (size_t)4 * 16 * (kH + 8) * 2; }

template <bool FIELD>
static hipError_t mlp_infer_launch(int n, const float* X, float* Y, const MlpFieldIO& io, const void* wb,
                                   const float* par, const float* norm, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int tiles = (n + 15) / 16;
  const int blocks = std::min((tiles + 3) / 4, 256 * 8);
  static bool attr = false;  // > 64 KB of dynamic LDS (160 KB per CU on gfx950)
  if (!attr) {
    hipError_t e = hipFuncSetAttribute((const void*)k_mlp_infer<FIELD>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)mlp_infer_lds());
    if (e != hipSuccess) return e;
    attr = true;
  }
  hipLaunchKernelGGL((k_mlp_infer<FIELD>), dim3(blocks), dim3(256), mlp_infer_lds(), s, n, X, Y, io,
                     static_cast<const __bf16*>(wb), par, norm);
  return hipGetLastError();
}

hipError_t launch_mlp_infer(int n, const float* X, float* Y, const void* wb, const float* par, const float* norm,
                            hipStream_t s) {
  MlpFieldIO io;
  std::memset(&io, 0, sizeof(io));
  return mlp_infer_launch<false>(n, X, Y, io, wb, par, norm, s);
}

hipError_t launch_mlp_field(const MlpFieldIO& io, const void* wb, const float* par, const float* norm,
                            hipStream_t s) {
  return mlp_infer_launch<true>((int)io.P, nullptr, nullptr, io, wb, par, norm, s);
}

// SplitMix64 of (seed, i): point index of sample i
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void k_mlp_sample(MlpFieldIO io, const float* __restrict__ t, int n, unsigned long long seed,
                             float* __restrict__ X, float* __restrict__ T) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t p = (size_t)(mix64(seed ^ mix64((unsigned long long)i)) % (unsigned long long)io.P);
  for (int k = 0; k < kIn; ++k) X[(size_t)i * kIn + k] = field_feature(io, p, k);
  for (int j = 0; j < kOut; ++j) {
    const size_t q = (size_t)j * io.P + p;
    T[(size_t)i * kOut + j] = (t[q] - io.h1[q]) - (io.h1[q] - io.h2[q]);
  }
}

hipError_t launch_mlp_sample(const MlpFieldIO& io, const float* t, int n, unsigned long long seed, float* X,
                             float* T, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_sample, dim3((n + 255) / 256), dim3(256), 0, s, io, t, n, seed, X, T);
  return hipGetLastError();
}

// -------------------------------------------------------------- training --
constexpr int kBS = 16;  // samples per block (static LDS ~51 KB)

// forward + backward for kBS samples; grad partials [gridDim.x][kParams],
// loss partials [gridDim.x]; scale = 1/(total samples)
__global__ __launch_bounds__(256) void k_mlp_fb(int n, const float* __restrict__ X, const float* __restrict__ T,
                                                const float* __restrict__ par, const float* __restrict__ norm,
                                                int mode, float scale, float* __restrict__ gpart,
                                                float* __restrict__ lpart) {
  __shared__ float a0[kBS][kIn];
  __shared__ float z[4][kBS][kH];      // pre-activations of the hidden layers
  __shared__ float zo[kBS][kOut];
  __shared__ float d[2][kBS][kH];      // backward deltas (ping-pong)
  __shared__ float lsum[256];
  const int tid = threadIdx.x;
  const int s0 = blockIdx.x * kBS;
  const float* mu_x = norm;
  const float* sd_x = norm + kIn;
  const float* mu_y = norm + 2 * kIn;
  const float* sd_y = norm + 2 * kIn + kOut;
  const bool rb = mode == 1;
  auto Wv = [&](int l, int j, int k) -> float {
    const float w = par[w_off(l) + j * lay_in(l) + k];
    return rb ? bf16r(w) : w;
  };
  // layer input value a_l[s][k] (layer l's input): a0 or gelu(z_{l-1}), bf16-rounded in mode 1
  auto Ain = [&](int l, int s, int k) -> float {
    const float v = (l == 0) ? a0[s][k] : gelu_f(z[l - 1][s][k]);
    return rb ? bf16r(v) : v;
  };
  for (int e = tid; e < kBS * kIn; e += 256) {
    const int s = e / kIn, k = e % kIn;
    a0[s][k] = (s0 + s < n) ? (X[(size_t)(s0 + s) * kIn + k] - mu_x[k]) / sd_x[k] : 0.0f;
  }
  __syncthreads();
  // forward
  for (int l = 0; l < kNL; ++l) {
    const int N = lay_out(l), K = lay_in(l);
    for (int e = tid; e < kBS * N; e += 256) {
      const int s = e / N, j = e % N;
      float acc = par[b_off(l) + j];
      for (int k = 0; k < K; ++k) acc += Ain(l, s, k) * Wv(l, j, k);
      if (l < kNL - 1) z[l][s][j] = acc;
      else zo[s][j] = acc;
    }
    __syncthreads();
  }
  // loss and output delta: L = scale sum (yn - tn)^2
  float lacc = 0.0f;
  for (int e = tid; e < kBS * kOut; e += 256) {
    const int s = e / kOut, j = e % kOut;
    float dv = 0.0f;
    if (s0 + s < n) {
      const float tn = (T[(size_t)(s0 + s) * kOut + j] - mu_y[j]) / sd_y[j];
      const float r = zo[s][j] - tn;
      lacc += r * r;
      dv = 2.0f * scale * r;
    }
    d[0][s][j] = dv;
  }
  lsum[tid] = lacc;
  __syncthreads();
  if (tid == 0) {
    float t = 0.0f;
    for (int i = 0; i < 256; ++i) t += lsum[i];
    lpart[blockIdx.x] = t * scale;
  }
  // backward
  float* gp = gpart + (size_t)blockIdx.x * kParams;
  int cur = 0;
  for (int l = kNL - 1; l >= 0; --l) {
    const int N = lay_out(l), K = lay_in(l);
    // weight and bias gradient partials of this block
    for (int e = tid; e < N * K; e += 256) {
      const int j = e / K, k = e % K;
      float acc = 0.0f;
      for (int s = 0; s < kBS; ++s) acc += d[cur][s][j] * Ain(l, s, k);
      gp[w_off(l) + e] = acc;
    }
    for (int j = tid; j < N; j += 256) {
      float acc = 0.0f;
      for (int s = 0; s < kBS; ++s) acc += d[cur][s][j];
      gp[b_off(l) + j] = acc;
    }
    if (l == 0) break;
    // delta of the previous layer: (d W) * gelu'(z)
    for (int e = tid; e < kBS * K; e += 256) {
      const int s = e / K, k = e % K;
      float acc = 0.0f;
      for (int j = 0; j < N; ++j) acc += d[cur][s][j] * Wv(l, j, k);
      d[cur ^ 1][s][k] = acc * gelu_df(z[l - 1][s][k]);
    }
    __syncthreads();
    cur ^= 1;
  }
}

// g[i] = sum over blocks of gpart[b][i] (fixed order); loss = sum of lpart
__global__ void k_mlp_reduce(int nblk, const float* __restrict__ gpart, const float* __restrict__ lpart,
                             float* __restrict__ g, float* __restrict__ loss) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kParams) {
    float acc = 0.0f;
    for (int b = 0; b < nblk; ++b) acc += gpart[(size_t)b * kParams + i];
    g[i] = acc;
  }
  if (i == 0) {
    float t = 0.0f;
    for (int b = 0; b < nblk; ++b) t += lpart[b];
    *loss = t;
  }
}

// Adam on the FP32 master parameters; refresh the padded bf16 copies
__global__ void k_mlp_adam(float* __restrict__ par, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ g, float lr, float b1, float b2, float eps, float c1, float c2,
                           __bf16* __restrict__ wb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kParams) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float p = par[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  par[i] = p;
  // padded bf16 copy of the weights
  for (int l = 0; l < kNL; ++l) {
    const int w0 = w_off(l), wn = lay_out(l) * lay_in(l);
    if (i >= w0 && i < w0 + wn) {
      const int j = (i - w0) / lay_in(l), k = (i - w0) % lay_in(l);
      wb[wb_off(l) + j * wrow(l) + k] = (__bf16)p;
    }
  }
}

// rebuild the padded bf16 weights (zero padding) from the master parameters
__global__ void k_mlp_pack(const float* __restrict__ par, __bf16* __restrict__ wb) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kWb) return;
  int l = 0;
  while (l + 1 < kNL && e >= wb_off(l + 1)) ++l;
  const int r = (e - wb_off(l)) / wrow(l), k = (e - wb_off(l)) % wrow(l);
  float v = 0.0f;
  if (r < lay_out(l) && k < lay_in(l)) v = par[w_off(l) + r * lay_in(l) + k];
  wb[e] = (__bf16)v;
}

hipError_t launch_mlp_pack(const float* par, void* wb, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_pack, dim3((kWb + 255) / 256), dim3(256), 0, s, par, static_cast<__bf16*>(wb));
  return hipGetLastError();
}

int mlp_fb_blocks(int n) { return (n + kBS - 1) / kBS; }

hipError_t launch_mlp_grad(int n, const float* X, const float* T, const float* par, const float* norm, int mode,
                           float* gpart, float* lpart, float* g, float* loss, hipStream_t s) {
  const int nb = mlp_fb_blocks(n);
  hipLaunchKernelGGL(k_mlp_fb, dim3(nb), dim3(256), 0, s, n, X, T, par, norm, mode, 1.0f / (float)n, gpart,
                     lpart);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kParams + 255) / 256), dim3(256), 0, s, nb, gpart, lpart, g, loss);
  return hipGetLastError();
}

hipError_t launch_mlp_adam(float* par, float* m, float* v, const float* g, float lr, int t, void* wb,
                           hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
  hipLaunchKernelGGL(k_mlp_adam, dim3((kParams + 255) / 256), dim3(256), 0, s, par, m, v, g, lr, b1, b2, eps, c1,
                     c2, static_cast<__bf16*>(wb));
  return hipGetLastError();
}

}  // namespace kmm
