// C ABI of the learned polymer-stress closure (kmm_mlp_*): parameter
// handling, normalisation, GPU inference and GPU training with Adam on
// mini-batches.  No reference counterpart (the network is new).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../include/kmm.h"
#include "kmm_kernels.h"

const char* kmm_last_error_set(const std::string& msg);  // kmm_api.cpp

struct kmm_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  int np = 0;                       // parameter count
  float *par = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;
  float* norm = nullptr;            // mu_x[22], sd_x[22], mu_y[6], sd_y[6]
  void* wb = nullptr;               // padded bf16 weights
  int adam_t = 0;
  int act = 0;                      // hidden-layer activation: 0 ReLU, 1 GELU, 2 GELU (polynomial sigmoid)
  // scratch (grown on demand)
  float *X = nullptr, *T = nullptr, *Y = nullptr, *gpart = nullptr, *lpart = nullptr, *loss = nullptr;
  size_t capX = 0, capT = 0, capY = 0, capG = 0;
};

static int mfail(int code, const std::string& msg) {
  kmm_last_error_set(msg);
  return code;
}
#define MHIP(call)                                                                    \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) return mfail(KMM_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

static int grow(float** p, size_t* cap, size_t n) {
  if (n <= *cap) return KMM_OK;
  if (*p) hipFree(*p);
  *p = nullptr;
  MHIP(hipMalloc((void**)p, n * sizeof(float)));
  *cap = n;
  return KMM_OK;
}

extern "C" int kmm_mlp_param_count(void) { return kmm::mlp_param_count(); }

extern "C" int kmm_mlp_create(int device, unsigned long long seed, kmm_mlp** out) {
  if (!out) return mfail(KMM_E_ARG, "null argument");
  *out = nullptr;
  if (hipSetDevice(device) != hipSuccess) return mfail(KMM_E_HIP, "hipSetDevice failed");
  kmm_mlp* n = new kmm_mlp();
  n->device = device;
  n->np = kmm::mlp_param_count();
  if (hipStreamCreate(&n->stream) != hipSuccess) { delete n; return mfail(KMM_E_HIP, "hipStreamCreate failed"); }
  const size_t pb = (size_t)n->np * sizeof(float);
  if (hipMalloc((void**)&n->par, pb) != hipSuccess || hipMalloc((void**)&n->m, pb) != hipSuccess ||
      hipMalloc((void**)&n->v, pb) != hipSuccess || hipMalloc((void**)&n->g, pb) != hipSuccess ||
      hipMalloc((void**)&n->norm, 56 * sizeof(float)) != hipSuccess ||
      hipMalloc(&n->wb, (size_t)kmm::mlp_wb_count() * 2) != hipSuccess ||
      hipMalloc((void**)&n->loss, sizeof(float)) != hipSuccess) {
    kmm_mlp_destroy(n);
    return mfail(KMM_E_HIP, "hipMalloc failed");
  }
  // He-style uniform initialisation U(-a, a), a = sqrt(6/n_in), zero biases
  static const int sizes[6] = {22, 128, 128, 128, 128, 6};
  std::vector<float> p(n->np, 0.0f);
  std::mt19937_64 rng(seed);
  size_t o = 0;
  for (int l = 0; l < 5; ++l) {
    const double a = std::sqrt(6.0 / sizes[l]);
    std::uniform_real_distribution<double> U(-a, a);
    for (int i = 0; i < sizes[l + 1] * sizes[l]; ++i) p[o++] = (float)U(rng);
    o += sizes[l + 1];
  }
  std::vector<float> nrm(56, 0.0f);
  for (int i = 22; i < 44; ++i) nrm[i] = 1.0f;
  for (int i = 50; i < 56; ++i) nrm[i] = 1.0f;
  int r = kmm_mlp_set_params(n, n->np, p.data());
  if (!r) r = kmm_mlp_set_norm(n, nrm.data(), nrm.data() + 22, nrm.data() + 44, nrm.data() + 50);
  if (r) { kmm_mlp_destroy(n); return r; }
  *out = n;
  return KMM_OK;
}

extern "C" int kmm_mlp_destroy(kmm_mlp* n) {
  if (!n) return KMM_OK;
  hipSetDevice(n->device);
  if (n->stream) hipStreamSynchronize(n->stream);
  for (float* p : {n->par, n->m, n->v, n->g, n->norm, n->X, n->T, n->Y, n->gpart, n->lpart, n->loss})
    if (p) hipFree(p);
  if (n->wb) hipFree(n->wb);
  if (n->stream) hipStreamDestroy(n->stream);
  delete n;
  return KMM_OK;
}

extern "C" int kmm_mlp_set_params(kmm_mlp* n, int np, const float* p) {
  if (!n || !p || np != n->np) return mfail(KMM_E_ARG, "parameter count mismatch");
  MHIP(hipSetDevice(n->device));
  const size_t pb = (size_t)n->np * sizeof(float);
  MHIP(hipMemcpyAsync(n->par, p, pb, hipMemcpyHostToDevice, n->stream));
  MHIP(hipMemsetAsync(n->m, 0, pb, n->stream));
  MHIP(hipMemsetAsync(n->v, 0, pb, n->stream));
  n->adam_t = 0;
  MHIP(kmm::launch_mlp_pack(n->par, n->wb, n->stream));
  MHIP(hipStreamSynchronize(n->stream));
  return KMM_OK;
}

extern "C" int kmm_mlp_get_params(kmm_mlp* n, int np, float* p) {
  if (!n || !p || np != n->np) return mfail(KMM_E_ARG, "parameter count mismatch");
  MHIP(hipSetDevice(n->device));
  MHIP(hipMemcpyAsync(p, n->par, (size_t)n->np * sizeof(float), hipMemcpyDeviceToHost, n->stream));
  MHIP(hipStreamSynchronize(n->stream));
  return KMM_OK;
}

extern "C" int kmm_mlp_set_norm(kmm_mlp* n, const float* mu_x, const float* sd_x, const float* mu_y,
                                const float* sd_y) {
  if (!n || !mu_x || !sd_x || !mu_y || !sd_y) return mfail(KMM_E_ARG, "null argument");
  float h[56];
  std::memcpy(h, mu_x, 22 * sizeof(float));
  std::memcpy(h + 22, sd_x, 22 * sizeof(float));
  std::memcpy(h + 44, mu_y, 6 * sizeof(float));
  std::memcpy(h + 50, sd_y, 6 * sizeof(float));
  for (int i = 22; i < 44; ++i)
    if (!(h[i] > 0.0f)) return mfail(KMM_E_ARG, "input scales must be positive");
  for (int i = 50; i < 56; ++i)
    if (!(h[i] > 0.0f)) return mfail(KMM_E_ARG, "output scales must be positive");
  MHIP(hipSetDevice(n->device));
  MHIP(hipMemcpyAsync(n->norm, h, sizeof(h), hipMemcpyHostToDevice, n->stream));
  MHIP(hipStreamSynchronize(n->stream));
  return KMM_OK;
}

extern "C" int kmm_mlp_infer(kmm_mlp* n, int npts, const float* X, float* Y) {
  if (!n || npts < 0 || (npts > 0 && (!X || !Y))) return mfail(KMM_E_ARG, "bad argument");
  if (npts == 0) return KMM_OK;
  MHIP(hipSetDevice(n->device));
  int r;
  if ((r = grow(&n->X, &n->capX, (size_t)npts * 22)) || (r = grow(&n->Y, &n->capY, (size_t)npts * 6))) return r;
  MHIP(hipMemcpyAsync(n->X, X, (size_t)npts * 22 * sizeof(float), hipMemcpyHostToDevice, n->stream));
  MHIP(kmm::launch_mlp_infer(npts, n->X, n->Y, n->wb, n->par, n->norm, n->act, n->stream));
  MHIP(hipMemcpyAsync(Y, n->Y, (size_t)npts * 6 * sizeof(float), hipMemcpyDeviceToHost, n->stream));
  MHIP(hipStreamSynchronize(n->stream));
  return KMM_OK;
}

// loss and gradient of the current parameters on one batch (mode 0: FP32,
// mode 1: bf16 rounding as in inference)
extern "C" int kmm_mlp_grad(kmm_mlp* n, int npts, const float* X, const float* T, int mode, float* grad,
                            float* loss) {
  if (!n || npts <= 0 || !X || !T || !grad || !loss || (mode != 0 && mode != 1)) return mfail(KMM_E_ARG, "bad argument");
  MHIP(hipSetDevice(n->device));
  const int nb = kmm::mlp_fb_blocks(npts);
  int r;
  if ((r = grow(&n->X, &n->capX, (size_t)npts * 22)) || (r = grow(&n->T, &n->capT, (size_t)npts * 6)) ||
      (r = grow(&n->gpart, &n->capG, (size_t)nb * n->np)))
    return r;
  if (n->lpart) hipFree(n->lpart);
  MHIP(hipMalloc((void**)&n->lpart, (size_t)nb * sizeof(float)));
  MHIP(hipMemcpyAsync(n->X, X, (size_t)npts * 22 * sizeof(float), hipMemcpyHostToDevice, n->stream));
  MHIP(hipMemcpyAsync(n->T, T, (size_t)npts * 6 * sizeof(float), hipMemcpyHostToDevice, n->stream));
  MHIP(kmm::launch_mlp_grad(npts, n->X, n->T, n->par, n->norm, mode | (n->act << 1), n->gpart, n->lpart, n->g,
                            n->loss, n->stream));
  MHIP(hipMemcpyAsync(grad, n->g, (size_t)n->np * sizeof(float), hipMemcpyDeviceToHost, n->stream));
  MHIP(hipMemcpyAsync(loss, n->loss, sizeof(float), hipMemcpyDeviceToHost, n->stream));
  MHIP(hipStreamSynchronize(n->stream));
  return KMM_OK;
}

// Adam training on (X, T): `steps` mini-batches of `batch` samples drawn
// without replacement from a seeded permutation (reshuffled every epoch);
// losses[s] (may be NULL) receives the batch loss before update s.
extern "C" int kmm_mlp_train(kmm_mlp* n, int npts, const float* X, const float* T, int steps, int batch, float lr,
                             unsigned long long seed, float* losses) {
  if (!n || npts <= 0 || !X || !T || steps < 0 || batch <= 0 || batch > npts || !(lr > 0.0f))
    return mfail(KMM_E_ARG, "bad argument");
  MHIP(hipSetDevice(n->device));
  // batches are copied in order from a host-side shuffled copy
  std::vector<int> perm(npts);
  std::iota(perm.begin(), perm.end(), 0);
  std::mt19937_64 rng(seed);
  std::vector<float> bx((size_t)batch * 22), bt((size_t)batch * 6);
  const int nb = kmm::mlp_fb_blocks(batch);
  int r;
  if ((r = grow(&n->X, &n->capX, (size_t)batch * 22)) || (r = grow(&n->T, &n->capT, (size_t)batch * 6)) ||
      (r = grow(&n->gpart, &n->capG, (size_t)nb * n->np)))
    return r;
  if (n->lpart) hipFree(n->lpart);
  MHIP(hipMalloc((void**)&n->lpart, (size_t)nb * sizeof(float)));
  int pos = npts;
  for (int s = 0; s < steps; ++s) {
    if (pos + batch > npts) {
      std::shuffle(perm.begin(), perm.end(), rng);
      pos = 0;
    }
    for (int i = 0; i < batch; ++i) {
      const int q = perm[pos + i];
      std::memcpy(&bx[(size_t)i * 22], X + (size_t)q * 22, 22 * sizeof(float));
      std::memcpy(&bt[(size_t)i * 6], T + (size_t)q * 6, 6 * sizeof(float));
    }
    pos += batch;
    MHIP(hipMemcpyAsync(n->X, bx.data(), bx.size() * sizeof(float), hipMemcpyHostToDevice, n->stream));
    MHIP(hipMemcpyAsync(n->T, bt.data(), bt.size() * sizeof(float), hipMemcpyHostToDevice, n->stream));
    MHIP(kmm::launch_mlp_grad(batch, n->X, n->T, n->par, n->norm, 1 | (n->act << 1), n->gpart, n->lpart, n->g,
                              n->loss, n->stream));
    MHIP(kmm::launch_mlp_adam(n->par, n->m, n->v, n->g, lr, ++n->adam_t, n->wb, n->stream));
    if (losses) MHIP(hipMemcpyAsync(losses + s, n->loss, sizeof(float), hipMemcpyDeviceToHost, n->stream));
    // the host buffers are reused by the next batch
    MHIP(hipStreamSynchronize(n->stream));
  }
  return KMM_OK;
}

// Adam steps on samples already on the device (dX [npts][22], dT [npts][6]),
// queued on stream s without host synchronisation: step k uses the k-th
// consecutive block of `batch` samples (cycling), so the samples must come
// in random order (kmm_poly_samples' sampler draws them so).  Used by the
// channel to retrain the closure on pairs drawn inside audit windows.
int kmm_mlp_train_device(kmm_mlp* n, int npts, const float* dX, const float* dT, int steps, int batch, float lr,
                         hipStream_t s) {
  if (!n || npts <= 0 || !dX || !dT || steps <= 0 || batch <= 0 || batch > npts || !(lr > 0.0f))
    return mfail(KMM_E_ARG, "bad argument");
  const int nb = kmm::mlp_fb_blocks(batch);
  int r;
  if ((r = grow(&n->gpart, &n->capG, (size_t)nb * n->np))) return r;
  if (n->lpart) hipFree(n->lpart);
  MHIP(hipMalloc((void**)&n->lpart, (size_t)nb * sizeof(float)));
  const int nblocks = npts / batch;
  for (int k = 0; k < steps; ++k) {
    const size_t off = (size_t)(k % nblocks) * batch;
    MHIP(kmm::launch_mlp_grad(batch, dX + off * 22, dT + off * 6, n->par, n->norm, 1 | (n->act << 1), n->gpart,
                              n->lpart, n->g,
                              n->loss, s));
    MHIP(kmm::launch_mlp_adam(n->par, n->m, n->v, n->g, lr, ++n->adam_t, n->wb, s));
  }
  return KMM_OK;
}

extern "C" int kmm_mlp_set_activation(kmm_mlp* n, int act) {
  if (!n || act < 0 || act > 2)
    return mfail(KMM_E_ARG, "activation: 0 (ReLU), 1 (GELU, sigmoid form) or 2 (GELU, polynomial sigmoid)");
  n->act = act;
  return KMM_OK;
}

int kmm_mlp_activation(const kmm_mlp* m) { return m ? m->act : 0; }

// loss of the last Adam step queued on stream s (waits for it)
int kmm_mlp_last_loss(kmm_mlp* m, float* loss, hipStream_t s) {
  MHIP(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, s));
  MHIP(hipStreamSynchronize(s));
  return KMM_OK;
}

bool kmm_mlp_view(kmm_mlp* m, const void** wb, const float** par, const float** norm, int* device) {
  if (!m) return false;
  *wb = m->wb;
  *par = m->par;
  *norm = m->norm;
  *device = m->device;
  return true;
}
