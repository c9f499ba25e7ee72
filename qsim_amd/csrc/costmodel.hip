// costmodel.hip -- the pass-time cost model: a small MLP trained and
// evaluated on the GPU with its own kernels.
//
// Network: x (F features) -> 64 -> 64 -> 64 -> 1, ReLU after each hidden
// layer, linear output (a log pass time).  Parameters live on the device in
// one flat FP32 array, layer by layer: W_l (in x out, row-major) then b_l.
// Training minimises the mean squared error with Adam; every step is a fixed
// chain of kernels (forward per layer, loss gradient, backward per layer,
// Adam) on the library stream.  Reductions over the batch are loops in a
// fixed order, so a step is deterministic.
//
// No reference counterpart: qsim's fusion (lib/fuser_mqubit.h:185-526)
// has no cost model; this one scores candidate partitions for the planner.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "internal.h"

namespace qhip {

namespace {

constexpr int kLayers = 4;  // three hidden layers and the output layer

struct CostModel {
  unsigned F = 0, H = 0;
  unsigned in[kLayers], out[kLayers];
  size_t woff[kLayers], boff[kLayers];
  size_t nparams = 0;
  float* params = nullptr;  // device
  float* grads = nullptr;   // device
  float* m = nullptr;       // Adam moments
  float* v = nullptr;
  unsigned long long adam_t = 0;
  // Activation buffers for `cap` samples: Z_l (pre-activation) and A_l.
  unsigned cap = 0;
  float* x = nullptr;
  float* t = nullptr;
  float* z[kLayers] = {};
  float* dz[kLayers] = {};
  float* loss = nullptr;  // per-block partial sums
};

__global__ void LinearForward(const float* __restrict__ X,
                              const float* __restrict__ W,
                              const float* __restrict__ b, float* __restrict__ Z,
                              unsigned B, unsigned I, unsigned O,
                              int relu_input) {
  const unsigned idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * O) return;
  const unsigned s = idx / O, o = idx % O;
  float acc = b[o];
  const float* xr = X + size_t(s) * I;
  for (unsigned i = 0; i < I; ++i) {
    float a = xr[i];
    if (relu_input) a = a > 0.0f ? a : 0.0f;
    acc = fmaf(a, W[size_t(i) * O + o], acc);
  }
  Z[idx] = acc;
}

// dY = 2 (Y - T) / B; per-block partial sums of (Y - T)^2.
__global__ void MseGrad(const float* __restrict__ Y, const float* __restrict__ T,
                        float* __restrict__ dY, float* __restrict__ parts,
                        unsigned B) {
  __shared__ float sh[256];
  const unsigned idx = blockIdx.x * blockDim.x + threadIdx.x;
  float e = 0.0f;
  if (idx < B) {
    const float d = Y[idx] - T[idx];
    dY[idx] = 2.0f * d / float(B);
    e = d * d;
  }
  sh[threadIdx.x] = e;
  __syncthreads();
  for (unsigned w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) parts[blockIdx.x] = sh[0];
}

// dW[i][o] = sum_s a(X[s][i]) dZ[s][o]; db[o] = sum_s dZ[s][o].
__global__ void LinearGradW(const float* __restrict__ X,
                            const float* __restrict__ dZ,
                            float* __restrict__ dW, float* __restrict__ db,
                            unsigned B, unsigned I, unsigned O,
                            int relu_input) {
  const unsigned idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < I * O) {
    const unsigned i = idx / O, o = idx % O;
    float acc = 0.0f;
    for (unsigned s = 0; s < B; ++s) {
      float a = X[size_t(s) * I + i];
      if (relu_input) a = a > 0.0f ? a : 0.0f;
      acc = fmaf(a, dZ[size_t(s) * O + o], acc);
    }
    dW[idx] = acc;
  } else if (idx < I * O + O) {
    const unsigned o = idx - I * O;
    float acc = 0.0f;
    for (unsigned s = 0; s < B; ++s) acc += dZ[size_t(s) * O + o];
    db[o] = acc;
  }
}

// dZprev[s][i] = (sum_o dZ[s][o] W[i][o]) * (Zprev[s][i] > 0).
__global__ void LinearGradX(const float* __restrict__ dZ,
                            const float* __restrict__ W,
                            const float* __restrict__ Zprev,
                            float* __restrict__ dZprev, unsigned B, unsigned I,
                            unsigned O) {
  const unsigned idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * I) return;
  const unsigned s = idx / I, i = idx % I;
  float acc = 0.0f;
  for (unsigned o = 0; o < O; ++o)
    acc = fmaf(dZ[size_t(s) * O + o], W[size_t(i) * O + o], acc);
  dZprev[idx] = Zprev[idx] > 0.0f ? acc : 0.0f;
}

__global__ void AdamStep(float* __restrict__ p, const float* __restrict__ g,
                         float* __restrict__ m, float* __restrict__ v,
                         unsigned N, float lr, float b1, float b2, float eps,
                         float c1, float c2) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void Gather(const float* __restrict__ xall,
                       const float* __restrict__ tall,
                       const unsigned* __restrict__ idx, unsigned B, unsigned F,
                       float* __restrict__ x, float* __restrict__ t) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * F) return;
  const unsigned s = i / F, f = i % F;
  const unsigned r = idx[s];
  x[i] = xall[size_t(r) * F + f];
  if (f == 0) t[s] = tall[r];
}

__global__ void FillZero(float* __restrict__ p, unsigned N) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) p[i] = 0.0f;
}

unsigned Blocks(size_t n) { return unsigned((n + 255) / 256); }

int Reserve(CostModel& c, unsigned B) {
  if (B <= c.cap) return QSIM_HIP_OK;
  (void)hipFree(c.x);
  (void)hipFree(c.t);
  (void)hipFree(c.loss);
  for (int l = 0; l < kLayers; ++l) {
    (void)hipFree(c.z[l]);
    (void)hipFree(c.dz[l]);
    c.z[l] = c.dz[l] = nullptr;
  }
  c.x = c.t = c.loss = nullptr;
  c.cap = 0;
  QHIP_CHECK(hipMalloc(&c.x, size_t(B) * c.F * sizeof(float)));
  QHIP_CHECK(hipMalloc(&c.t, size_t(B) * sizeof(float)));
  QHIP_CHECK(hipMalloc(&c.loss, size_t(Blocks(B)) * sizeof(float)));
  for (int l = 0; l < kLayers; ++l) {
    QHIP_CHECK(hipMalloc(&c.z[l], size_t(B) * c.out[l] * sizeof(float)));
    QHIP_CHECK(hipMalloc(&c.dz[l], size_t(B) * c.out[l] * sizeof(float)));
  }
  c.cap = B;
  return QSIM_HIP_OK;
}

// Forward pass of B samples already in c.x; output in c.z[kLayers - 1].
int Forward(CostModel& c, unsigned B) {
  const float* X = c.x;
  for (int l = 0; l < kLayers; ++l) {
    hipLaunchKernelGGL(LinearForward, dim3(Blocks(size_t(B) * c.out[l])),
                       dim3(256), 0, Stream(), X, c.params + c.woff[l],
                       c.params + c.boff[l], c.z[l], B, c.in[l], c.out[l],
                       l > 0 ? 1 : 0);
    QHIP_LAUNCHED("LinearForward");
    X = c.z[l];
  }
  return QSIM_HIP_OK;
}

// Gradient of the mean squared error of B samples (c.x, c.t) into c.grads.
int Backward(CostModel& c, unsigned B) {
  const int L = kLayers - 1;
  hipLaunchKernelGGL(MseGrad, dim3(Blocks(B)), dim3(256), 0, Stream(), c.z[L],
                     c.t, c.dz[L], c.loss, B);
  QHIP_LAUNCHED("MseGrad");
  for (int l = L; l >= 0; --l) {
    const float* X = l > 0 ? c.z[l - 1] : c.x;
    hipLaunchKernelGGL(LinearGradW,
                       dim3(Blocks(size_t(c.in[l]) * c.out[l] + c.out[l])),
                       dim3(256), 0, Stream(), X, c.dz[l],
                       c.grads + c.woff[l], c.grads + c.boff[l], B, c.in[l],
                       c.out[l], l > 0 ? 1 : 0);
    QHIP_LAUNCHED("LinearGradW");
    if (l > 0) {
      hipLaunchKernelGGL(LinearGradX, dim3(Blocks(size_t(B) * c.in[l])),
                         dim3(256), 0, Stream(), c.dz[l],
                         c.params + c.woff[l], c.z[l - 1], c.dz[l - 1], B,
                         c.in[l], c.out[l]);
      QHIP_LAUNCHED("LinearGradX");
    }
  }
  return QSIM_HIP_OK;
}

int Upload(CostModel& c, const float* x, const float* t, unsigned B) {
  QHIP_TRY(Reserve(c, B));
  QHIP_CHECK(hipMemcpyAsync(c.x, x, size_t(B) * c.F * sizeof(float),
                            hipMemcpyHostToDevice, Stream()));
  if (t) {
    QHIP_CHECK(hipMemcpyAsync(c.t, t, size_t(B) * sizeof(float),
                              hipMemcpyHostToDevice, Stream()));
  }
  return QSIM_HIP_OK;
}

int LossSum(CostModel& c, unsigned B, double* out) {
  std::vector<float> parts(Blocks(B));
  QHIP_CHECK(hipMemcpyAsync(parts.data(), c.loss, parts.size() * sizeof(float),
                            hipMemcpyDeviceToHost, Stream()));
  QHIP_CHECK(hipStreamSynchronize(Stream()));
  double s = 0;
  for (float p : parts) s += p;
  *out = s / B;
  return QSIM_HIP_OK;
}

}  // namespace

}  // namespace qhip

using namespace qhip;

extern "C" {

int qsim_hip_costmodel_create(unsigned num_features, unsigned hidden,
                              void** model) {
  return GuardHost([&]() -> int {
    if (model == nullptr || num_features == 0 || num_features > 256 ||
        hidden == 0 || hidden > 256) {
      SetError("cost model: bad sizes (features %u, hidden %u)", num_features,
               hidden);
      return QSIM_HIP_ERR_ARG;
    }
    *model = nullptr;
    QHIP_TRY(EnsureInit());
    auto* c = new CostModel;
    c->F = num_features;
    c->H = hidden;
    size_t off = 0;
    for (int l = 0; l < kLayers; ++l) {
      c->in[l] = l == 0 ? num_features : hidden;
      c->out[l] = l == kLayers - 1 ? 1 : hidden;
      c->woff[l] = off;
      off += size_t(c->in[l]) * c->out[l];
      c->boff[l] = off;
      off += c->out[l];
    }
    c->nparams = off;
    const size_t bytes = off * sizeof(float);
    hipError_t e = hipMalloc(&c->params, bytes);
    if (e == hipSuccess) e = hipMalloc(&c->grads, bytes);
    if (e == hipSuccess) e = hipMalloc(&c->m, bytes);
    if (e == hipSuccess) e = hipMalloc(&c->v, bytes);
    if (e != hipSuccess) {
      (void)hipFree(c->params);
      (void)hipFree(c->grads);
      (void)hipFree(c->m);
      (void)hipFree(c->v);
      delete c;
      return HipFail(e, "hipMalloc(cost model)");
    }
    // He-style initialisation from a fixed seed; biases zero.
    std::vector<float> p(off, 0.0f);
    std::mt19937 rng(12345);
    for (int l = 0; l < kLayers; ++l) {
      std::normal_distribution<float> d(0.0f, std::sqrt(2.0f / c->in[l]));
      for (size_t i = 0; i < size_t(c->in[l]) * c->out[l]; ++i)
        p[c->woff[l] + i] = d(rng);
    }
    QHIP_CHECK(hipMemcpyAsync(c->params, p.data(), bytes, hipMemcpyHostToDevice,
                              Stream()));
    hipLaunchKernelGGL(FillZero, dim3(Blocks(off)), dim3(256), 0, Stream(), c->m,
                       unsigned(off));
    hipLaunchKernelGGL(FillZero, dim3(Blocks(off)), dim3(256), 0, Stream(), c->v,
                       unsigned(off));
    QHIP_LAUNCHED("FillZero");
    QHIP_CHECK(hipStreamSynchronize(Stream()));
    *model = c;
    return QSIM_HIP_OK;
  });
}

// The process's default model (qsim_hip_costmodel_set_default).
static std::atomic<void*> g_default_model{nullptr};

void* qsim_hip_costmodel_default(void) { return g_default_model.load(); }

int qsim_hip_costmodel_set_default(void* model) {
  g_default_model.store(model);
  return QSIM_HIP_OK;
}

int qsim_hip_costmodel_load(const char* path, void** model) {
  return GuardHost([&]() -> int {
    if (path == nullptr || model == nullptr) {
      SetError("cost model: null path or handle");
      return QSIM_HIP_ERR_ARG;
    }
    *model = nullptr;
    std::FILE* f = std::fopen(path, "rb");
    if (f == nullptr) {
      SetError("cost model: cannot open %s", path);
      return QSIM_HIP_ERR_ARG;
    }
    std::string text;
    char buf[65536];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
    std::fclose(f);
    // {"num_features": F, "hidden": H, "params": [p0, p1, ...], ...}
    auto number_after = [&](const char* key, double* v) {
      const size_t k = text.find(key);
      if (k == std::string::npos) return false;
      const size_t c = text.find(':', k);
      if (c == std::string::npos) return false;
      char* end = nullptr;
      *v = std::strtod(text.c_str() + c + 1, &end);
      return end != text.c_str() + c + 1;
    };
    double F = 0, H = 0;
    const size_t pk = text.find("\"params\"");
    const size_t lb = pk == std::string::npos ? pk : text.find('[', pk);
    if (!number_after("\"num_features\"", &F) ||
        !number_after("\"hidden\"", &H) || lb == std::string::npos) {
      SetError("cost model: %s lacks num_features, hidden or params", path);
      return QSIM_HIP_ERR_ARG;
    }
    std::vector<float> params;
    const char* p = text.c_str() + lb + 1;
    while (true) {
      while (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t' || *p == ',') ++p;
      if (*p == ']' || *p == 0) break;
      char* end = nullptr;
      const double v = std::strtod(p, &end);
      if (end == p) {
        SetError("cost model: bad number in %s", path);
        return QSIM_HIP_ERR_ARG;
      }
      params.push_back(float(v));
      p = end;
    }
    void* m = nullptr;
    QHIP_TRY(qsim_hip_costmodel_create(unsigned(F), unsigned(H), &m));
    const int rc = qsim_hip_costmodel_set_params(m, params.data(), params.size());
    if (rc != QSIM_HIP_OK) {
      qsim_hip_costmodel_free(m);
      return rc;
    }
    *model = m;
    return QSIM_HIP_OK;
  });
}

int qsim_hip_costmodel_free(void* model) {
  if (model == nullptr) return QSIM_HIP_OK;
  {
    void* expected = model;
    g_default_model.compare_exchange_strong(expected, nullptr);
  }
  auto* c = static_cast<CostModel*>(model);
  (void)hipStreamSynchronize(Stream());
  (void)hipFree(c->params);
  (void)hipFree(c->grads);
  (void)hipFree(c->m);
  (void)hipFree(c->v);
  (void)hipFree(c->x);
  (void)hipFree(c->t);
  (void)hipFree(c->loss);
  for (int l = 0; l < kLayers; ++l) {
    (void)hipFree(c->z[l]);
    (void)hipFree(c->dz[l]);
  }
  delete c;
  return QSIM_HIP_OK;
}

uint64_t qsim_hip_costmodel_num_params(void* model) {
  return model ? static_cast<CostModel*>(model)->nparams : 0;
}

int qsim_hip_costmodel_set_params(void* model, const float* params,
                                  uint64_t n) {
  auto* c = static_cast<CostModel*>(model);
  if (c == nullptr || params == nullptr || n != c->nparams) {
    SetError("cost model: expected %zu parameters", c ? c->nparams : 0);
    return QSIM_HIP_ERR_ARG;
  }
  QHIP_CHECK(hipMemcpyAsync(c->params, params, n * sizeof(float),
                            hipMemcpyHostToDevice, Stream()));
  // A new parameter set restarts the optimiser.
  hipLaunchKernelGGL(FillZero, dim3(Blocks(n)), dim3(256), 0, Stream(), c->m,
                     unsigned(n));
  hipLaunchKernelGGL(FillZero, dim3(Blocks(n)), dim3(256), 0, Stream(), c->v,
                     unsigned(n));
  QHIP_LAUNCHED("FillZero");
  c->adam_t = 0;
  QHIP_CHECK(hipStreamSynchronize(Stream()));
  return QSIM_HIP_OK;
}

int qsim_hip_costmodel_get_params(void* model, float* params, uint64_t n) {
  auto* c = static_cast<CostModel*>(model);
  if (c == nullptr || params == nullptr || n != c->nparams) {
    SetError("cost model: expected %zu parameters", c ? c->nparams : 0);
    return QSIM_HIP_ERR_ARG;
  }
  QHIP_CHECK(hipMemcpyAsync(params, c->params, n * sizeof(float),
                            hipMemcpyDeviceToHost, Stream()));
  QHIP_CHECK(hipStreamSynchronize(Stream()));
  return QSIM_HIP_OK;
}

int qsim_hip_costmodel_forward(void* model, const float* x, unsigned batch,
                               float* y) {
  auto* c = static_cast<CostModel*>(model);
  if (c == nullptr || (batch > 0 && (x == nullptr || y == nullptr))) {
    SetError("cost model: null operand");
    return QSIM_HIP_ERR_ARG;
  }
  if (batch == 0) return QSIM_HIP_OK;
  QHIP_TRY(Upload(*c, x, nullptr, batch));
  QHIP_TRY(Forward(*c, batch));
  QHIP_CHECK(hipMemcpyAsync(y, c->z[kLayers - 1], size_t(batch) * sizeof(float),
                            hipMemcpyDeviceToHost, Stream()));
  QHIP_CHECK(hipStreamSynchronize(Stream()));
  return QSIM_HIP_OK;
}

int qsim_hip_costmodel_gradient(void* model, const float* x, const float* t,
                                unsigned batch, double* loss, float* grad) {
  auto* c = static_cast<CostModel*>(model);
  if (c == nullptr || batch == 0 || x == nullptr || t == nullptr) {
    SetError("cost model: null operand or empty batch");
    return QSIM_HIP_ERR_ARG;
  }
  QHIP_TRY(Upload(*c, x, t, batch));
  QHIP_TRY(Forward(*c, batch));
  QHIP_TRY(Backward(*c, batch));
  if (grad) {
    QHIP_CHECK(hipMemcpyAsync(grad, c->grads, c->nparams * sizeof(float),
                              hipMemcpyDeviceToHost, Stream()));
  }
  double l = 0;
  QHIP_TRY(LossSum(*c, batch, &l));
  if (loss) *loss = l;
  return QSIM_HIP_OK;
}

int qsim_hip_costmodel_train(void* model, const float* x, const float* t,
                             unsigned num_samples, unsigned epochs,
                             unsigned batch, float lr, unsigned seed,
                             double* final_loss) {
  return GuardHost([&]() -> int {
    auto* c = static_cast<CostModel*>(model);
    if (c == nullptr || x == nullptr || t == nullptr || num_samples == 0 ||
        batch == 0) {
      SetError("cost model: bad training arguments");
      return QSIM_HIP_ERR_ARG;
    }
    batch = std::min(batch, num_samples);
    QHIP_TRY(Reserve(*c, std::max(batch, num_samples)));
    // The training set stays on the device; each epoch uploads one seeded
    // permutation and every step gathers its minibatch from it.
    float *xall = nullptr, *tall = nullptr;
    unsigned* idx = nullptr;
    auto release = [&]() {
      (void)hipStreamSynchronize(Stream());
      (void)hipFree(xall);
      (void)hipFree(tall);
      (void)hipFree(idx);
    };
    hipError_t e = hipMalloc(&xall, size_t(num_samples) * c->F * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&tall, size_t(num_samples) * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&idx, size_t(num_samples) * sizeof(unsigned));
    if (e == hipSuccess)
      e = hipMemcpyAsync(xall, x, size_t(num_samples) * c->F * sizeof(float),
                         hipMemcpyHostToDevice, Stream());
    if (e == hipSuccess)
      e = hipMemcpyAsync(tall, t, size_t(num_samples) * sizeof(float),
                         hipMemcpyHostToDevice, Stream());
    if (e != hipSuccess) {
      release();
      return HipFail(e, "cost model training buffers");
    }
    std::vector<unsigned> order(num_samples);
    for (unsigned i = 0; i < num_samples; ++i) order[i] = i;
    std::mt19937 rng(seed);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (unsigned ep = 0; ep < epochs; ++ep) {
      std::shuffle(order.begin(), order.end(), rng);
      e = hipMemcpyAsync(idx, order.data(), size_t(num_samples) * sizeof(unsigned),
                         hipMemcpyHostToDevice, Stream());
      if (e != hipSuccess) {
        release();
        return HipFail(e, "cost model permutation");
      }
      for (unsigned s0 = 0; s0 + batch <= num_samples; s0 += batch) {
        hipLaunchKernelGGL(Gather, dim3(Blocks(size_t(batch) * c->F)), dim3(256),
                           0, Stream(), xall, tall, idx + s0, batch, c->F, c->x,
                           c->t);
        int rc = Forward(*c, batch);
        if (rc == QSIM_HIP_OK) rc = Backward(*c, batch);
        if (rc != QSIM_HIP_OK) {
          release();
          return rc;
        }
        ++c->adam_t;
        const float c1 = 1.0f - std::pow(b1, float(c->adam_t));
        const float c2 = 1.0f - std::pow(b2, float(c->adam_t));
        hipLaunchKernelGGL(AdamStep, dim3(Blocks(c->nparams)), dim3(256), 0,
                           Stream(), c->params, c->grads, c->m, c->v,
                           unsigned(c->nparams), lr, b1, b2, eps, c1, c2);
      }
      // `order` is rewritten next epoch: its upload must have finished.
      e = hipStreamSynchronize(Stream());
      if (e != hipSuccess) {
        release();
        return HipFail(e, "cost model epoch");
      }
    }
    // Mean squared error over the whole training set.
    e = hipMemcpyAsync(c->x, xall, size_t(num_samples) * c->F * sizeof(float),
                       hipMemcpyDeviceToDevice, Stream());
    if (e == hipSuccess)
      e = hipMemcpyAsync(c->t, tall, size_t(num_samples) * sizeof(float),
                         hipMemcpyDeviceToDevice, Stream());
    release();
    if (e != hipSuccess) return HipFail(e, "cost model evaluation copy");
    QHIP_TRY(Forward(*c, num_samples));
    hipLaunchKernelGGL(MseGrad, dim3(Blocks(num_samples)), dim3(256), 0, Stream(),
                       c->z[kLayers - 1], c->t, c->dz[kLayers - 1], c->loss,
                       num_samples);
    QHIP_LAUNCHED("MseGrad");
    double l = 0;
    QHIP_TRY(LossSum(*c, num_samples, &l));
    if (final_loss) *final_loss = l;
    return QSIM_HIP_OK;
  });
}

}  // extern "C"
