// Small dense MLP (fp32) trained and evaluated by hand-written HIP kernels: the cost model
// that predicts log(MSM time) from job features and a candidate schedule.
//
// Layers l = 0..L-1 map dims[l] -> dims[l+1] with row-major W_l (out x in) and bias b_l;
// hidden layers use ReLU, the output layer is linear.  Loss: mean squared error.  Optimiser:
// Adam.  All parameters live in one device buffer (W_0, b_0, W_1, b_1, ...).  The network only
// chooses schedules; it never touches field arithmetic.
#include <math.h>
#include <string.h>

#include <vector>

#include "mlp.hpp"
#include "runtime.hpp"

namespace zk {

static constexpr int kT = 256;
static constexpr int kMaxLayers = 6;

// Y[b][o] = act(sum_i W[o][i] X[b][i] + bias[o])
__global__ void dense_fwd_kernel(const float *__restrict__ X, const float *__restrict__ W, const float *__restrict__ bias,
                                 float *__restrict__ Y, int B, int in, int out, int relu) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * out) return;
  const int b = t / out, o = t % out;
  float s = bias[o];
  for (int i = 0; i < in; i++) s += W[o * in + i] * X[b * in + i];
  Y[t] = relu ? fmaxf(s, 0.f) : s;
}

// dZ = dY * act'(Y) (in place on dY)
__global__ void relu_bwd_kernel(float *__restrict__ dY, const float *__restrict__ Y, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n && Y[t] <= 0.f) dY[t] = 0.f;
}

// dW[o][i] = sum_b dZ[b][o] X[b][i];  db[o] = sum_b dZ[b][o]
__global__ void dense_wgrad_kernel(const float *__restrict__ dZ, const float *__restrict__ X, float *__restrict__ dW,
                                   float *__restrict__ db, int B, int in, int out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < out * in) {
    const int o = t / in, i = t % in;
    float s = 0.f;
    for (int b = 0; b < B; b++) s += dZ[b * out + o] * X[b * in + i];
    dW[t] = s;
  } else if (t < out * in + out) {
    const int o = t - out * in;
    float s = 0.f;
    for (int b = 0; b < B; b++) s += dZ[b * out + o];
    db[o] = s;
  }
}

// dX[b][i] = sum_o dZ[b][o] W[o][i]
__global__ void dense_xgrad_kernel(const float *__restrict__ dZ, const float *__restrict__ W, float *__restrict__ dX,
                                   int B, int in, int out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * in) return;
  const int b = t / in, i = t % in;
  float s = 0.f;
  for (int o = 0; o < out; o++) s += dZ[b * out + o] * W[o * in + i];
  dX[t] = s;
}

// squared error over B rows of `outs` outputs: loss = (1/B) sum (p - y)^2, dP = 2 (p - y) / B
__global__ void mse_kernel(const float *__restrict__ P, const float *__restrict__ y, float *__restrict__ dP,
                           float *__restrict__ loss, int B, int outs) {
  __shared__ float sh[kT];
  float acc = 0.f;
  for (int k = threadIdx.x; k < B * outs; k += kT) {
    const float d = P[k] - y[k];
    dP[k] = 2.f * d / (float)B;
    acc += d * d;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = sh[0] / (float)B;
}

__global__ void adam_kernel(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, int n, float lr, float b1, float b2, float c1, float c2, float eps) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gt = g[t];
  m[t] = b1 * m[t] + (1.f - b1) * gt;
  v[t] = b2 * v[t] + (1.f - b2) * gt * gt;
  p[t] -= lr * (m[t] / c1) / (sqrtf(v[t] / c2) + eps);
}

struct Mlp {
  int L = 0;
  std::vector<int> dims;
  std::vector<size_t> woff, boff;
  size_t nparams = 0;
  DevBuf params, grads, m, v;
  int step = 0;
  hipStream_t st = nullptr;
};

static inline unsigned grid(size_t n) { return (unsigned)((n + kT - 1) / kT); }

int mlp_create(int nlayers, const int *dims, unsigned seed, Mlp **out) {
  Context *ctx;
  ZK_TRY(context(&ctx));
  if (nlayers < 1 || nlayers > kMaxLayers) {
    set_error("mlp: 1..6 layers");
    return ZK_EINVAL;
  }
  for (int l = 0; l <= nlayers; l++)
    if (dims[l] < 1 || dims[l] > 4096) {
      set_error("mlp: layer widths must be 1..4096");
      return ZK_EINVAL;
    }
  Mlp *n = new Mlp();
  n->L = nlayers;
  n->dims.assign(dims, dims + nlayers + 1);
  for (int l = 0; l < nlayers; l++) {
    n->woff.push_back(n->nparams);
    n->nparams += (size_t)dims[l] * dims[l + 1];
    n->boff.push_back(n->nparams);
    n->nparams += dims[l + 1];
  }
  if (hipStreamCreateWithFlags(&n->st, hipStreamNonBlocking) != hipSuccess) {
    delete n;
    set_error("mlp: stream creation failed");
    return ZK_EHIP;
  }
  int s;
  if ((s = n->params.alloc(n->nparams * 4)) || (s = n->grads.alloc(n->nparams * 4)) || (s = n->m.alloc(n->nparams * 4)) ||
      (s = n->v.alloc(n->nparams * 4))) {
    delete n;
    return s;
  }
  // He-uniform initialisation from a 64-bit LCG (host), biases zero
  std::vector<float> h(n->nparams, 0.f);
  uint64_t x = 0x9E3779B97F4A7C15ull ^ seed;
  for (int l = 0; l < nlayers; l++) {
    const float lim = sqrtf(6.f / (float)dims[l]);
    for (size_t k = 0; k < (size_t)dims[l] * dims[l + 1]; k++) {
      x = x * 6364136223846793005ull + 1442695040888963407ull;
      const float u = (float)((x >> 40) & 0xffffff) / 16777216.f;
      h[n->woff[l] + k] = (2.f * u - 1.f) * lim;
    }
  }
  ZK_HIP(hipMemcpy(n->params.p, h.data(), n->nparams * 4, hipMemcpyHostToDevice));
  ZK_HIP(hipMemset(n->m.p, 0, n->nparams * 4));
  ZK_HIP(hipMemset(n->v.p, 0, n->nparams * 4));
  *out = n;
  return ZK_OK;
}

void mlp_destroy(Mlp *n) {
  if (n->st) (void)hipStreamDestroy(n->st);
  delete n;
}
size_t mlp_nparams(const Mlp *n) { return n->nparams; }
int mlp_layers(const Mlp *n) { return n->L; }
int mlp_dim(const Mlp *n, int i) { return n->dims[i]; }

int mlp_get_params(Mlp *n, float *host) {
  ZK_HIP(hipMemcpy(host, n->params.p, n->nparams * 4, hipMemcpyDeviceToHost));
  return ZK_OK;
}
int mlp_set_params(Mlp *n, const float *host) {
  ZK_HIP(hipMemcpy(n->params.p, host, n->nparams * 4, hipMemcpyHostToDevice));
  n->step = 0;
  ZK_HIP(hipMemset(n->m.p, 0, n->nparams * 4));
  ZK_HIP(hipMemset(n->v.p, 0, n->nparams * 4));
  return ZK_OK;
}

// forward over B rows of device input X; activations per layer kept in acts (device)
static int forward(Mlp *n, int B, const float *dX, DevBuf *acts) {
  const float *in = dX;
  for (int l = 0; l < n->L; l++) {
    ZK_TRY(acts[l + 1].alloc((size_t)B * n->dims[l + 1] * 4));
    const float *W = n->params.as<float>() + n->woff[l], *b = n->params.as<float>() + n->boff[l];
    hipLaunchKernelGGL(dense_fwd_kernel, dim3(grid((size_t)B * n->dims[l + 1])), dim3(kT), 0, n->st, in, W, b,
                       acts[l + 1].as<float>(), B, n->dims[l], n->dims[l + 1], l + 1 < n->L ? 1 : 0);
    in = acts[l + 1].as<float>();
  }
  ZK_HIP(hipGetLastError());
  return ZK_OK;
}

// loss and gradients (into n->grads) for B rows
static int backward(Mlp *n, int B, const float *dX, const float *dy, float *loss_host) {
  DevBuf acts[kMaxLayers + 1];
  ZK_TRY(forward(n, B, dX, acts));
  DevBuf dcur, dnext, dloss;
  const int outs = n->dims[n->L];
  ZK_TRY(dcur.alloc((size_t)B * outs * 4));
  ZK_TRY(dloss.alloc(4));
  hipLaunchKernelGGL(mse_kernel, dim3(1), dim3(kT), 0, n->st, acts[n->L].as<float>(), dy, dcur.as<float>(),
                     dloss.as<float>(), B, outs);
  for (int l = n->L - 1; l >= 0; l--) {
    const int in = n->dims[l], out = n->dims[l + 1];
    if (l + 1 < n->L)
      hipLaunchKernelGGL(relu_bwd_kernel, dim3(grid((size_t)B * out)), dim3(kT), 0, n->st, dcur.as<float>(),
                         acts[l + 1].as<float>(), B * out);
    const float *X = l == 0 ? dX : acts[l].as<float>();
    hipLaunchKernelGGL(dense_wgrad_kernel, dim3(grid((size_t)out * in + out)), dim3(kT), 0, n->st, dcur.as<float>(), X,
                       n->grads.as<float>() + n->woff[l], n->grads.as<float>() + n->boff[l], B, in, out);
    if (l > 0) {
      ZK_TRY(dnext.alloc((size_t)B * in * 4));
      hipLaunchKernelGGL(dense_xgrad_kernel, dim3(grid((size_t)B * in)), dim3(kT), 0, n->st, dcur.as<float>(),
                         n->params.as<float>() + n->woff[l], dnext.as<float>(), B, in, out);
      std::swap(dcur.p, dnext.p);
      std::swap(dcur.bytes, dnext.bytes);
    }
  }
  ZK_HIP(hipGetLastError());
  if (loss_host) ZK_HIP(hipMemcpyAsync(loss_host, dloss.p, 4, hipMemcpyDeviceToHost, n->st));
  ZK_HIP(hipStreamSynchronize(n->st));
  return ZK_OK;
}

int mlp_forward(Mlp *n, int B, const float *X_host, float *Y_host) {
  DevBuf dX;
  ZK_TRY(dX.alloc((size_t)B * n->dims[0] * 4));
  ZK_HIP(hipMemcpyAsync(dX.p, X_host, (size_t)B * n->dims[0] * 4, hipMemcpyHostToDevice, n->st));
  DevBuf acts[kMaxLayers + 1];
  ZK_TRY(forward(n, B, dX.as<float>(), acts));
  ZK_HIP(hipMemcpyAsync(Y_host, acts[n->L].p, (size_t)B * n->dims[n->L] * 4, hipMemcpyDeviceToHost, n->st));
  ZK_HIP(hipStreamSynchronize(n->st));
  return ZK_OK;
}

int mlp_gradients(Mlp *n, int B, const float *X_host, const float *y_host, float *g_host, float *loss) {
  DevBuf dX, dy;
  ZK_TRY(dX.alloc((size_t)B * n->dims[0] * 4));
  const size_t ybytes = (size_t)B * n->dims[n->L] * 4;
  ZK_TRY(dy.alloc(ybytes));
  ZK_HIP(hipMemcpyAsync(dX.p, X_host, (size_t)B * n->dims[0] * 4, hipMemcpyHostToDevice, n->st));
  ZK_HIP(hipMemcpyAsync(dy.p, y_host, ybytes, hipMemcpyHostToDevice, n->st));
  ZK_TRY(backward(n, B, dX.as<float>(), dy.as<float>(), loss));
  if (g_host) ZK_HIP(hipMemcpy(g_host, n->grads.p, n->nparams * 4, hipMemcpyDeviceToHost));
  return ZK_OK;
}

int mlp_train(Mlp *n, int B, const float *X_host, const float *y_host, int epochs, float lr, float *final_loss) {
  DevBuf dX, dy;
  ZK_TRY(dX.alloc((size_t)B * n->dims[0] * 4));
  const size_t ybytes = (size_t)B * n->dims[n->L] * 4;
  ZK_TRY(dy.alloc(ybytes));
  ZK_HIP(hipMemcpyAsync(dX.p, X_host, (size_t)B * n->dims[0] * 4, hipMemcpyHostToDevice, n->st));
  ZK_HIP(hipMemcpyAsync(dy.p, y_host, ybytes, hipMemcpyHostToDevice, n->st));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float loss = 0.f;
  for (int e = 0; e < epochs; e++) {
    ZK_TRY(backward(n, B, dX.as<float>(), dy.as<float>(), e + 1 == epochs ? &loss : nullptr));
    n->step++;
    const float c1 = 1.f - powf(b1, (float)n->step), c2 = 1.f - powf(b2, (float)n->step);
    hipLaunchKernelGGL(adam_kernel, dim3(grid(n->nparams)), dim3(kT), 0, n->st, n->params.as<float>(),
                       n->grads.as<float>(), n->m.as<float>(), n->v.as<float>(), (int)n->nparams, lr, b1, b2, c1, c2,
                       eps);
  }
  ZK_HIP(hipGetLastError());
  ZK_HIP(hipStreamSynchronize(n->st));
  if (final_loss) *final_loss = loss;
  return ZK_OK;
}

}  // namespace zk
