// Learned polymer-stress closure: a 22 -> 4 x 128 -> 6 multilayer
// perceptron with bfloat16 weights and FP32 accumulation (no reference
// counterpart); hidden activation ReLU, or GELU with a sigmoid or a
// polynomial-sigmoid factor
// x sigma(1.702 x), chosen per network.
//
//   k_mlp_infer   inference on the matrix cores: v_mfma_f32_16x16x32_bf16
//                 with the weights as the A operand (staged once per block in
//                 LDS, input index permuted so that each layer's accumulators
//                 are the next layer's B operands) and the activations of
//                 kT tiles of 16 points in registers; 8 wavefronts per block,
//                 one block per CU, persistent over passes of 64 points.
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
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef short i16x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));  // eight packed bf16 (an MFMA operand)

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

// activation: ReLU.  On the matrix-core path it is applied to the packed bf16
// pair after rounding (a negative bf16 is a negative int16, so one
// v_pk_max_i16 against 0 serves two activations; rounding and ReLU commute).
// On gfx950 the vector ALU and the matrix cores do not overlap much (each
// vector instruction adds about 3.5 cycles to a 16x16x32 product's 17), so
// the activation's instruction count sets a large part of inference time.
__device__ __forceinline__ float act_f(float x) { return fmaxf(x, 0.0f); }
__device__ __forceinline__ float act_df(float x) { return x > 0.0f ? 1.0f : 0.0f; }
// GELU in the sigmoid form x sigma(1.702 x) (Hendrycks and Gimpel, "Gaussian
// Error Linear Units", 2016, section 2; within 0.02 of x Phi(x)): one
// exponential, one reciprocal, three multiply-adds per activation
constexpr float kGeluA = 1.702f;
__device__ __forceinline__ float gelu_sig(float x) {
  return __builtin_amdgcn_rcpf(1.0f + __expf(-kGeluA * x));
}
__device__ __forceinline__ float gelu_f(float x) { return x * gelu_sig(x); }
__device__ __forceinline__ float gelu_df(float x) {
  const float sg = gelu_sig(x);
  return sg + kGeluA * x * sg * (1.0f - sg);
}
// GELU with a polynomial sigmoid (no exponential, no reciprocal: only
// full-rate vector operations): x s(t), t = clamp(1/2 + x/(2c), 0, 1),
// s(t) = t^3 (t (6 t - 15) + 10) (the quintic smoothstep, s' = 30 t^2 (t-1)^2
// vanishes at both clamps), c = 2.64: within 0.0195 of x Phi(x) (the sigmoid
// form above: 0.0204)
constexpr float kGpK = 1.0f / (2.0f * 2.64f);
__device__ __forceinline__ float gelup_t(float x) { return fminf(fmaxf(fmaf(x, kGpK, 0.5f), 0.0f), 1.0f); }
__device__ __forceinline__ float gelup_f(float x) {
  const float t = gelup_t(x);
  return x * (t * t * t * fmaf(t, fmaf(t, 6.0f, -15.0f), 10.0f));
}
__device__ __forceinline__ float gelup_df(float x) {
  const float t = gelup_t(x), u = t * (t - 1.0f);
  return t * t * t * fmaf(t, fmaf(t, 6.0f, -15.0f), 10.0f) + x * kGpK * 30.0f * u * u;
}
// activation a (0 ReLU, 1 GELU, 2 GELU with the polynomial sigmoid) and its derivative
__device__ __forceinline__ float act_a(int a, float x) {
  return a == 2 ? gelup_f(x) : (a ? gelu_f(x) : act_f(x));
}
__device__ __forceinline__ float act_da(int a, float x) {
  return a == 2 ? gelup_df(x) : (a ? gelu_df(x) : act_df(x));
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
// Layout for the matrix cores.  One v_mfma_f32_16x16x32_bf16 computes
// D[m][n] = sum_k A[m][k] B[k][n] + C[m][n] with lane l holding A row l%16,
// B column l%16, both at k = 8 (l/16) + i (i < 8), and D column l%16, rows
// 4 (l/16) + i (i < 4).  Inference puts the weights in A (m = output neuron)
// and the activations in B (n = point), so a lane's D values for neuron tiles
// 2j and 2j+1 are exactly the eight B values it needs at k-step j of the next
// layer once the next layer's input index is permuted:
//   slot s = 32 j + 8 q + i  <->  neuron 32 j + 16 (i/4) + 4 q + i%4.
// The permutation is applied to the packed bf16 weights, so activations never
// leave registers between layers.  The first layer's slots hold the features
// in the order that lets lanes q = 0, 1 keep 2 h1 - h2 for the outputs they
// produce (rows 0-3 and 4-5 of the last layer):
//   q=0: h1[0..3], (h1-h2)[0..3]     q=1: h1[4,5], (h1-h2)[4,5], g[0..3]
//   q=2: g[4..8], wall distance, 0, 0     q=3: 0
__host__ __device__ constexpr int feat_of_slot(int s) {
  return s < 4 ? 10 + s : s < 8 ? 12 + s : s < 10 ? 6 + s : s < 12 ? 10 + s : s < 21 ? s - 12 : s == 21 ? 9 : -1;
}
__host__ __device__ constexpr int slot_of_feat(int k) {
  return k < 9 ? 12 + k : k == 9 ? 21 : k < 14 ? k - 10 : k < 16 ? k - 6 : k < 20 ? k - 12 : k - 10;
}
__host__ __device__ constexpr int neuron_of_slot(int s) {
  return (s & ~31) | ((s & 4) << 2) | ((s & 24) >> 1) | (s & 3);
}
__host__ __device__ constexpr int slot_of_neuron(int k) {
  return (k & ~31) | ((k & 12) << 1) | ((k & 16) >> 2) | (k & 3);
}
// packed position of input index k of layer l
__host__ __device__ constexpr int slot_of(int l, int k) { return l == 0 ? slot_of_feat(k) : slot_of_neuron(k); }
// input index of packed slot s of layer l (-1: padding)
__host__ __device__ constexpr int input_of(int l, int s) { return l == 0 ? feat_of_slot(s) : neuron_of_slot(s); }

constexpr int kT = 4;             // 16-point tiles per wavefront pass (each weight read feeds kT products)
constexpr int kChunk = 16 * kT;   // points per pass
constexpr int kInferThreads = 512;  // 8 wavefronts: 2 per SIMD

__device__ __forceinline__ float wall_feature(const MlpFieldIO& io, size_t p) {
  return (float)(1.0 - fabs(io.ycoord[(unsigned)p / (unsigned)io.plane_pts]) / io.half);
}

// feature k of point p from field-major captures (learned-closure layout)
__device__ __forceinline__ float field_feature(const MlpFieldIO& io, size_t p, int k) {
  if (k < 9) return io.g[(size_t)k * io.P + p];
  if (k == 9) return wall_feature(io, p);
  if (k < 16) return io.h1[(size_t)(k - 10) * io.P + p];
  // the last step's change, formed in FP32 before the bf16 rounding of the inputs
  return io.h1[(size_t)(k - 16) * io.P + p] - io.h2[(size_t)(k - 16) * io.P + p];
}

// activation of two FP32 values, rounded to a packed bf16 pair: ReLU after
// the rounding (ACT 0), GELU before it (ACT 1, 2)
template <int ACT>
__device__ __forceinline__ unsigned act_pk(float a, float b) {
  if constexpr (ACT == 1) {
    const bf16x2 p = __builtin_convertvector(f32x2{gelu_f(a), gelu_f(b)}, bf16x2);
    return __builtin_bit_cast(unsigned, p);
  } else if constexpr (ACT == 2) {
    const bf16x2 p = __builtin_convertvector(f32x2{gelup_f(a), gelup_f(b)}, bf16x2);
    return __builtin_bit_cast(unsigned, p);
  } else {
    const bf16x2 p = __builtin_convertvector(f32x2{a, b}, bf16x2);  // one v_cvt_pk_bf16_f32
    return __builtin_bit_cast(unsigned, __builtin_elementwise_max(__builtin_bit_cast(i16x2, p), i16x2{0, 0}));
  }
}

__device__ __forceinline__ u32x4 pack8(const float (&v)[8]) {
  bf16x8 o;
#pragma unroll
  for (int i = 0; i < 8; ++i) o[i] = (__bf16)v[i];
  return __builtin_bit_cast(u32x4, o);
}

// One layer with KS input k-steps and 128 outputs, eight neuron tiles of 16:
// A operands (weights, rows of wr bf16) from LDS, each read once per pass and
// used for kT point tiles; the bias enters as the accumulator of the first
// k-step; outputs activated, rounded to bf16 and laid out as the next layer's
// B operands (tile nt fills elements 4 (nt%2) .. 4 (nt%2) + 3 of out[nt/2]).
// Software pipelined: tile nt's products are interleaved with the activation
// of tile nt-1 (matrix and vector pipes busy together) and with the weight
// reads of tile nt+1.
template <int KS, int ACT>
__device__ __forceinline__ void mlp_layer(const u32x4 (&in)[KS][kT], u32x4 (&out)[4][kT], const __bf16* W, int wr,
                                          const float* bias, int r16, int kq) {
  const __bf16* wl = W + r16 * wr + kq * 8;  // this lane's A row in tile 0
  const float* bl = bias + 4 * kq;            // this lane's four accumulator rows in tile 0
  f32x4 acc[2][kT];
  bf16x8 a[2][KS];
  f32x4 bb[2];
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) a[0][ks] = *reinterpret_cast<const bf16x8*>(wl + ks * 32);
  bb[0] = *reinterpret_cast<const f32x4*>(bl);
#pragma unroll
  for (int nt = 0; nt <= 8; ++nt) {
    if (nt < 7) {
#pragma unroll
      for (int ks = 0; ks < KS; ++ks)
        a[(nt + 1) & 1][ks] = *reinterpret_cast<const bf16x8*>(wl + (nt + 1) * 16 * wr + ks * 32);
      bb[(nt + 1) & 1] = *reinterpret_cast<const f32x4*>(bl + (nt + 1) * 16);
    }
    if (nt < 8) {
      const f32x4 b4 = bb[nt & 1];
#pragma unroll
      for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int t = 0; t < kT; ++t)
          acc[nt & 1][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
              a[nt & 1][ks], __builtin_bit_cast(bf16x8, in[ks][t]), ks == 0 ? b4 : acc[nt & 1][t], 0, 0, 0);
    }
    if (nt > 0) {
      const int m = nt - 1;
#pragma unroll
      for (int t = 0; t < kT; ++t) {
        out[m >> 1][t][2 * (m & 1)] = act_pk<ACT>(acc[m & 1][t][0], acc[m & 1][t][1]);
        out[m >> 1][t][2 * (m & 1) + 1] = act_pk<ACT>(acc[m & 1][t][2], acc[m & 1][t][3]);
      }
    }
    if (nt > 0 && nt < 8) {
      // one product, then the vector work that fits in its 16 cycles
#pragma unroll
      for (int i = 0; i < KS * kT; ++i) {
        if (i <= KS) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x002, 64 / (KS * kT) > 4 ? 64 / (KS * kT) : 4, 0);
      }
    }
    // keep each stage's registers from being hoisted into the previous one
    __builtin_amdgcn_sched_barrier(0);
  }
}

// FIELD = false: X[n][22] -> Y[n][6]; FIELD = true: io fields -> io.out[6][P].
// Persistent: wavefront w takes passes w, w + (all wavefronts), ...
template <bool FIELD, int ACT>
__global__ __launch_bounds__(kInferThreads) void k_mlp_infer(int n, const float* __restrict__ X,
                                                             float* __restrict__ Y, MlpFieldIO io,
                                                             const __bf16* __restrict__ wb,
                                                             const float* __restrict__ par,
                                                             const float* __restrict__ norm) {
  extern __shared__ __align__(16) unsigned char smraw[];
  __bf16* sw = reinterpret_cast<__bf16*>(smraw);           // [kWb]
  float* sb = reinterpret_cast<float*>(sw + kWb);           // biases [kBias]
  float* sxs = sb + kBias;                                  // input scale per slot [32]
  float* sxo = sxs + 32;                                    // input shift per slot [32]
  const int tid = threadIdx.x, lane = tid & 63;
  {
    const uint4* src = reinterpret_cast<const uint4*>(wb);
    uint4* dst = reinterpret_cast<uint4*>(sw);
    for (int e = tid; e < kWb / 8; e += blockDim.x) dst[e] = src[e];
  }
  // biases: layer l at 128 l (the last layer's 16 padded entries at 512)
  for (int l = 0; l < kNL; ++l)
    for (int j = tid; j < nout_pad(l); j += blockDim.x) sb[128 * l + j] = (j < lay_out(l)) ? par[b_off(l) + j] : 0.0f;
  if (tid < 32) {
    const int k = feat_of_slot(tid);
    sxs[tid] = k < 0 ? 0.0f : 1.0f / norm[kIn + k];
    sxo[tid] = k < 0 ? 0.0f : -norm[k] / norm[kIn + k];
  }
  __syncthreads();
  const float* mu_y = norm + 2 * kIn;
  const float* sd_y = norm + 2 * kIn + kOut;
  const int r16 = lane & 15, kq = lane >> 4;
  const int npass = (n + kChunk - 1) / kChunk;
  const int nwave = gridDim.x * (blockDim.x >> 6);
  for (int c = blockIdx.x * (blockDim.x >> 6) + (tid >> 6); c < npass; c += nwave) {
    const int p0 = c * kChunk;
    u32x4 x0[1][kT];
    float ex[kT][4];  // 2 h1 - h2 of the outputs this lane produces
    // all of the pass's loads first, then one wait.  Every load instruction
    // reads one array for the whole wavefront (rows chosen per lane group,
    // always in range; the values a group does not need are discarded):
    // A = h1 rows 4q + j (q = 0: 0-3, q = 1: 4-5), B = h2 likewise,
    // G = g rows 4 (q >= 2) + j (q = 1: 0-3, q >= 2: 4-8).
    float A[kT][4], B[kT][4], G[kT][5];
    double yc[kT];
    const unsigned P = FIELD ? (unsigned)io.P : 0u;
#pragma unroll
    for (int t = 0; t < kT; ++t) {
      const unsigned p = (unsigned)min(p0 + 16 * t + r16, n - 1);
      if (FIELD) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const unsigned row = (unsigned)min(4 * kq + j, 5);
          A[t][j] = io.h1[row * P + p];
          B[t][j] = io.h2[row * P + p];
        }
#pragma unroll
        for (int j = 0; j < 5; ++j) G[t][j] = io.g[(unsigned)min((kq >= 2 ? 4 : 0) + j, 8) * P + p];
        yc[t] = io.ycoord[p / (unsigned)io.plane_pts];
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int k = feat_of_slot(8 * kq + j);
          const float x = X[(size_t)p * kIn + (k < 0 ? 0 : k)];
          if (j < 4) A[t][j] = x;
          else B[t][j - 4] = x;
        }
      }
    }
    const bool q0 = kq == 0, q1 = kq == 1;
#pragma unroll
    for (int t = 0; t < kT; ++t) {
      float v[8];
      if (FIELD) {
        // slots of this lane's group: q=0 h1[0..3], h1[0..3]-h2[0..3];
        // q=1 h1[4,5], h1[4,5]-h2[4,5], g[0..3]; q=2 g[4..8], wall, 0, 0
        const float wall = (float)(1.0 - fabs(yc[t]) / io.half);
        float d[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          d[i] = A[t][i] - B[t][i];  // the last step's change, in FP32
          ex[t][i] = A[t][i] + d[i];
        }
        v[0] = q0 || q1 ? A[t][0] : G[t][0];
        v[1] = q0 || q1 ? A[t][1] : G[t][1];
        v[2] = q0 ? A[t][2] : q1 ? d[0] : G[t][2];
        v[3] = q0 ? A[t][3] : q1 ? d[1] : G[t][3];
        v[4] = q0 ? d[0] : q1 ? G[t][0] : G[t][4];
        v[5] = q0 ? d[1] : q1 ? G[t][1] : wall;
        v[6] = q0 ? d[2] : q1 ? G[t][2] : 0.0f;
        v[7] = q0 ? d[3] : q1 ? G[t][3] : 0.0f;
      } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          v[i] = A[t][i];
          v[4 + i] = B[t][i];
        }
      }
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = fmaf(v[i], sxs[8 * kq + i], sxo[8 * kq + i]);
      x0[0][t] = pack8(v);
    }
    u32x4 a1[4][kT], a2[4][kT];
    mlp_layer<1, ACT>(x0, a1, sw + wb_off(0), wrow(0), sb, r16, kq);
    mlp_layer<4, ACT>(a1, a2, sw + wb_off(1), wrow(1), sb + 128, r16, kq);
    mlp_layer<4, ACT>(a2, a1, sw + wb_off(2), wrow(2), sb + 256, r16, kq);
    mlp_layer<4, ACT>(a1, a2, sw + wb_off(3), wrow(3), sb + 384, r16, kq);
    // output layer: one neuron tile (rows 0-5 used)
    f32x4 acc[kT];
    {
      const __bf16* W = sw + wb_off(4);
      const f32x4 b4 = *reinterpret_cast<const f32x4*>(sb + 512 + 4 * kq);
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(W + r16 * wrow(4) + ks * 32 + kq * 8);
#pragma unroll
        for (int t = 0; t < kT; ++t)
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, __builtin_bit_cast(bf16x8, a2[ks][t]), ks == 0 ? b4 : acc[t],
                                                           0, 0, 0);
      }
    }
    if (kq < 2) {
      const int nout = kq == 0 ? 4 : 2;
#pragma unroll
      for (int t = 0; t < kT; ++t) {
        const int p = p0 + 16 * t + r16;
        if (p < n) {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int col = 4 * kq + i;
            if (i < nout) {
              const float yv = acc[t][i] * sd_y[col] + mu_y[col];
              // the closure network predicts the correction to linear
              // extrapolation of the two previous steps: out = 2 h1 - h2 + y
              if (FIELD) io.out[(size_t)col * io.P + p] = ex[t][i] + yv;
              else Y[(size_t)p * kOut + col] = yv;
            }
          }
        }
      }
    }
  }
}

size_t mlp_infer_lds() { return (size_t)kWb * 2 + (size_t)kBias * 4 + 64 * 4; }

static int device_cus() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                              hipSuccess || cus <= 0)
    cus = 256;
  return cus;
}

template <bool FIELD, int ACT>
static hipError_t mlp_infer_launch_a(int n, const float* X, float* Y, const MlpFieldIO& io, const void* wb,
                                     const float* par, const float* norm, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int passes = (n + kChunk - 1) / kChunk;
  static const int cus = device_cus();
  const int blocks = std::min((passes + kInferThreads / 64 - 1) / (kInferThreads / 64), cus);
  static bool attr = false;  // > 64 KB of dynamic LDS (160 KB per CU on gfx950)
  if (!attr) {
    hipError_t e = hipFuncSetAttribute((const void*)k_mlp_infer<FIELD, ACT>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)mlp_infer_lds());
    if (e != hipSuccess) return e;
    attr = true;
  }
  hipLaunchKernelGGL((k_mlp_infer<FIELD, ACT>), dim3(blocks), dim3(kInferThreads), mlp_infer_lds(), s, n, X, Y, io,
                     static_cast<const __bf16*>(wb), par, norm);
  return hipGetLastError();
}
template <bool FIELD>
static hipError_t mlp_infer_launch(int n, const float* X, float* Y, const MlpFieldIO& io, const void* wb,
                                   const float* par, const float* norm, int act, hipStream_t s) {
  if (act == 2) return mlp_infer_launch_a<FIELD, 2>(n, X, Y, io, wb, par, norm, s);
  return act ? mlp_infer_launch_a<FIELD, 1>(n, X, Y, io, wb, par, norm, s)
             : mlp_infer_launch_a<FIELD, 0>(n, X, Y, io, wb, par, norm, s);
}

hipError_t launch_mlp_infer(int n, const float* X, float* Y, const void* wb, const float* par, const float* norm,
                            int act, hipStream_t s) {
  MlpFieldIO io;
  std::memset(&io, 0, sizeof(io));
  return mlp_infer_launch<false>(n, X, Y, io, wb, par, norm, act, s);
}

hipError_t launch_mlp_field(const MlpFieldIO& io, const void* wb, const float* par, const float* norm, int act,
                            hipStream_t s) {
  // 32-bit point indices inside the kernel
  if (io.P >= (size_t)0x7fffffff || io.plane_pts <= 0) return hipErrorInvalidValue;
  return mlp_infer_launch<true>((int)io.P, nullptr, nullptr, io, wb, par, norm, act, s);
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
  const bool rb = (mode & 1) != 0;   // bf16 rounding where inference rounds
  const int act = (mode >> 1) & 3;    // 0 ReLU, 1 GELU, 2 GELU (polynomial sigmoid)
  auto Wv = [&](int l, int j, int k) -> float {
    const float w = par[w_off(l) + j * lay_in(l) + k];
    return rb ? bf16r(w) : w;
  };
  // layer input value a_l[s][k] (layer l's input): a0 or act(z_{l-1}), bf16-rounded in mode 1
  auto Ain = [&](int l, int s, int k) -> float {
    const float v = (l == 0) ? a0[s][k] : act_a(act, z[l - 1][s][k]);
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
    // delta of the previous layer: (d W) * act'(z)
    for (int e = tid; e < kBS * K; e += 256) {
      const int s = e / K, k = e % K;
      float acc = 0.0f;
      for (int j = 0; j < N; ++j) acc += d[cur][s][j] * Wv(l, j, k);
      d[cur ^ 1][s][k] = acc * act_da(act, z[l - 1][s][k]);
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
      wb[wb_off(l) + j * wrow(l) + slot_of(l, k)] = (__bf16)p;
    }
  }
}

// rebuild the padded, slot-permuted bf16 weights (zero padding) from the
// master parameters
__global__ void k_mlp_pack(const float* __restrict__ par, __bf16* __restrict__ wb) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kWb) return;
  int l = 0;
  while (l + 1 < kNL && e >= wb_off(l + 1)) ++l;
  const int r = (e - wb_off(l)) / wrow(l), s = (e - wb_off(l)) % wrow(l);
  const int k = s < kin_pad(l) ? input_of(l, s) : -1;
  float v = 0.0f;
  if (r < lay_out(l) && k >= 0 && k < lay_in(l)) v = par[w_off(l) + r * lay_in(l) + k];
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
