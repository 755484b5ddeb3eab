// Learned variant selector: an FP32 MLP 24 -> 64 -> 64 -> K (ReLU) that
// predicts the log-runtime (ln ms) of each of the K MTTKRP variants from 24
// statistics of a CSF tensor, with hand-written forward, backward and Adam
// kernels. It stands in for the reference's exhaustive codegen.py / *_tune.sh
// sweep (ebsr_32_tune.sh:1-16, codegen.py:44-51) as the first choice, and
// hands back to the sweep when its predicted margin is small.
//
// Parameters live on the device as one array, row-major [out][in]:
//   W1[64][24] b1[64] W2[64][64] b2[64] W3[K][64] b3[K]
// Inputs are standardised with per-feature (mean, std) and targets with a
// scalar (mean, std), both fixed by the caller from the training set.
// Training minimises the masked mean squared error of the standardised
// prediction (entries with mask 0, e.g. variants that do not fit, are
// ignored). Every reduction runs in a fixed order, so a training run is
// bit-reproducible for a given seed.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "handle.h"
#include "internal.h"

namespace tt {
namespace {

constexpr int kIn = TENTACO_NUM_FEATURES;  // 24
constexpr int kHid = TENTACO_MLP_HIDDEN;   // 64
constexpr int kMaxOut = 64;

__host__ __device__ inline int64_t n_params(int K) {
  return int64_t(kHid) * kIn + kHid + int64_t(kHid) * kHid + kHid + int64_t(K) * kHid + K;
}

struct Layout {
  const float* W1;
  const float* b1;
  const float* W2;
  const float* b2;
  const float* W3;
  const float* b3;
};

__device__ __forceinline__ Layout layout(const float* P, int K) {
  Layout L;
  L.W1 = P;
  L.b1 = L.W1 + kHid * kIn;
  L.W2 = L.b1 + kHid;
  L.b2 = L.W2 + kHid * kHid;
  L.W3 = L.b2 + kHid;
  L.b3 = L.W3 + K * kHid;
  return L;
}

// One 64-thread workgroup per sample; thread j owns hidden unit j of both
// layers and outputs j, j+64, ... . scale = {x_mean[24], x_std[24], y_mean,
// y_std}. Training mode stores the standardised input and both activations.
__global__ __launch_bounds__(64) void mlp_forward_kernel(const float* __restrict__ P,
                                                         const float* __restrict__ scale, int K,
                                                         const float* __restrict__ X, int n,
                                                         float* __restrict__ Xn,
                                                         float* __restrict__ H1,
                                                         float* __restrict__ H2,
                                                         float* __restrict__ Y, int destandardise) {
  __shared__ float sx[kIn];
  __shared__ float s1[kHid];
  __shared__ float s2[kHid];
  const int b = blockIdx.x;
  const int j = threadIdx.x;
  if (b >= n) return;
  const Layout L = layout(P, K);
  if (j < kIn) {
    const float v = (X[int64_t(b) * kIn + j] - scale[j]) / scale[kIn + j];
    sx[j] = v;
    if (Xn) Xn[int64_t(b) * kIn + j] = v;
  }
  __syncthreads();
  float a = L.b1[j];
  for (int i = 0; i < kIn; ++i) a = fmaf(L.W1[j * kIn + i], sx[i], a);
  a = fmaxf(a, 0.f);
  s1[j] = a;
  if (H1) H1[int64_t(b) * kHid + j] = a;
  __syncthreads();
  a = L.b2[j];
  for (int i = 0; i < kHid; ++i) a = fmaf(L.W2[j * kHid + i], s1[i], a);
  a = fmaxf(a, 0.f);
  s2[j] = a;
  if (H2) H2[int64_t(b) * kHid + j] = a;
  __syncthreads();
  const float ym = scale[2 * kIn], ys = scale[2 * kIn + 1];
  for (int o = j; o < K; o += kHid) {
    float y = L.b3[o];
    for (int i = 0; i < kHid; ++i) y = fmaf(L.W3[o * kHid + i], s2[i], y);
    Y[int64_t(b) * K + o] = destandardise ? fmaf(y, ys, ym) : y;
  }
}

// Back-propagated errors per sample (one 64-thread workgroup each):
//   d3 = 2 mask (y - (t - y_mean)/y_std) / count
//   d2 = (W3^T d3) * [h2 > 0],  d1 = (W2^T d2) * [h1 > 0]
// and the sample's squared-error sum (summed in lane order).
__global__ __launch_bounds__(64) void mlp_delta_kernel(const float* __restrict__ P,
                                                       const float* __restrict__ scale, int K,
                                                       const float* __restrict__ Y,
                                                       const float* __restrict__ T,
                                                       const float* __restrict__ M, int n,
                                                       const float* __restrict__ H1,
                                                       const float* __restrict__ H2,
                                                       float inv_count,
                                                       float* __restrict__ D1,
                                                       float* __restrict__ D2,
                                                       float* __restrict__ D3,
                                                       float* __restrict__ sample_loss) {
  __shared__ float s3[kMaxOut];
  __shared__ float s2[kHid];
  __shared__ float red[kHid];
  const int b = blockIdx.x;
  const int j = threadIdx.x;
  if (b >= n) return;
  const Layout L = layout(P, K);
  const float ym = scale[2 * kIn], ys = scale[2 * kIn + 1];
  float l = 0.f;
  for (int o = j; o < K; o += kHid) {
    const int64_t q = int64_t(b) * K + o;
    const float m = M[q];
    const float e = m != 0.f ? Y[q] - (T[q] - ym) / ys : 0.f;
    const float d = 2.f * m * e * inv_count;
    s3[o] = d;
    D3[q] = d;
    l = fmaf(m * e, e, l);
  }
  red[j] = l;
  __syncthreads();
  float a = 0.f;
  for (int o = 0; o < K; ++o) a = fmaf(L.W3[o * kHid + j], s3[o], a);
  a = H2[int64_t(b) * kHid + j] > 0.f ? a : 0.f;
  s2[j] = a;
  D2[int64_t(b) * kHid + j] = a;
  __syncthreads();
  a = 0.f;
  for (int k = 0; k < kHid; ++k) a = fmaf(L.W2[k * kHid + j], s2[k], a);
  a = H1[int64_t(b) * kHid + j] > 0.f ? a : 0.f;
  D1[int64_t(b) * kHid + j] = a;
  if (j == 0) {
    float s = 0.f;
    for (int t = 0; t < kHid; ++t) s += red[t];
    sample_loss[b] = s;
  }
}

// Parameter gradients: one thread per parameter, summing over the batch in
// sample order (deterministic; no atomics).
__global__ __launch_bounds__(256) void mlp_wgrad_kernel(int K, int n,
                                                        const float* __restrict__ Xn,
                                                        const float* __restrict__ H1,
                                                        const float* __restrict__ H2,
                                                        const float* __restrict__ D1,
                                                        const float* __restrict__ D2,
                                                        const float* __restrict__ D3,
                                                        float* __restrict__ G) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n_params(K)) return;
  int64_t q = p;
  float g = 0.f;
  if (q < kHid * kIn) {  // W1[j][i]
    const int j = int(q / kIn), i = int(q % kIn);
    for (int b = 0; b < n; ++b) g = fmaf(D1[int64_t(b) * kHid + j], Xn[int64_t(b) * kIn + i], g);
  } else if ((q -= kHid * kIn) < kHid) {  // b1[j]
    for (int b = 0; b < n; ++b) g += D1[int64_t(b) * kHid + q];
  } else if ((q -= kHid) < kHid * kHid) {  // W2[j][i]
    const int j = int(q / kHid), i = int(q % kHid);
    for (int b = 0; b < n; ++b) g = fmaf(D2[int64_t(b) * kHid + j], H1[int64_t(b) * kHid + i], g);
  } else if ((q -= kHid * kHid) < kHid) {  // b2[j]
    for (int b = 0; b < n; ++b) g += D2[int64_t(b) * kHid + q];
  } else if ((q -= kHid) < int64_t(K) * kHid) {  // W3[o][i]
    const int o = int(q / kHid), i = int(q % kHid);
    for (int b = 0; b < n; ++b) g = fmaf(D3[int64_t(b) * K + o], H2[int64_t(b) * kHid + i], g);
  } else {  // b3[o]
    q -= int64_t(K) * kHid;
    for (int b = 0; b < n; ++b) g += D3[int64_t(b) * K + q];
  }
  G[p] = g;
}

// Adam (Kingma & Ba, 2015) with bias corrections c1 = 1 - beta1^t,
// c2 = 1 - beta2^t passed in.
__global__ __launch_bounds__(256) void mlp_adam_kernel(float* __restrict__ P,
                                                       const float* __restrict__ G,
                                                       float* __restrict__ Mo,
                                                       float* __restrict__ Ve, int64_t np,
                                                       float lr, float beta1, float beta2,
                                                       float eps, float c1, float c2) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const float g = G[p];
  const float m = beta1 * Mo[p] + (1.f - beta1) * g;
  const float v = beta2 * Ve[p] + (1.f - beta2) * g * g;
  Mo[p] = m;
  Ve[p] = v;
  P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// Minibatch rows idx[0..n) of the resident training set into contiguous
// buffers.
__global__ __launch_bounds__(256) void mlp_gather_kernel(const int32_t* __restrict__ idx, int n,
                                                         int K, const float* __restrict__ X,
                                                         const float* __restrict__ T,
                                                         const float* __restrict__ M,
                                                         float* __restrict__ Xb,
                                                         float* __restrict__ Tb,
                                                         float* __restrict__ Mb) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t row = kIn + 2 * K;
  if (t >= int64_t(n) * row) return;
  const int b = int(t / row);
  const int c = int(t % row);
  const int64_t s = idx[b];
  if (c < kIn) Xb[int64_t(b) * kIn + c] = X[s * kIn + c];
  else if (c < kIn + K) Tb[int64_t(b) * K + (c - kIn)] = T[s * K + (c - kIn)];
  else Mb[int64_t(b) * K + (c - kIn - K)] = M[s * K + (c - kIn - K)];
}

// splitmix64 (Steele, Lea & Flood 2014), for initialisation and shuffles.
inline uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline double uniform01(uint64_t& s) { return double(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }
inline double gauss(uint64_t& s) {
  const double u1 = std::max(uniform01(s), 1e-300), u2 = uniform01(s);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

template <class T>
int dalloc(T** p, int64_t n) {
  *p = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(p), size_t(std::max<int64_t>(n, 1)) * sizeof(T)) !=
      hipSuccess) {
    *p = nullptr;
    return fail(TENTACO_ENOMEM, "selector: hipMalloc of %lld elements failed", (long long)n);
  }
  return TENTACO_OK;
}

}  // namespace
}  // namespace tt

using namespace tt;

struct tentaco_mlp {
  int K = 0;
  int64_t np = 0;
  float* P = nullptr;   // parameters
  float* G = nullptr;   // gradients
  float* Mo = nullptr;  // Adam first moment
  float* Ve = nullptr;  // Adam second moment
  int64_t adam_t = 0;
  float* scale = nullptr;  // x_mean[24], x_std[24], y_mean, y_std
  float h_scale[2 * kIn + 2] = {};
  // per-batch work buffers (capacity cap samples)
  int cap = 0;
  float *X = nullptr, *Xn = nullptr, *H1 = nullptr, *H2 = nullptr, *Y = nullptr, *T = nullptr,
        *M = nullptr, *D1 = nullptr, *D2 = nullptr, *D3 = nullptr, *loss = nullptr;

  void free_batch() {
    for (float* p : {X, Xn, H1, H2, Y, T, M, D1, D2, D3, loss}) (void)hipFree(p);
    X = Xn = H1 = H2 = Y = T = M = D1 = D2 = D3 = loss = nullptr;
    cap = 0;
  }
  int ensure(int n) {
    if (n <= cap) return TENTACO_OK;
    free_batch();
    int rc = 0;
    if ((rc = dalloc(&X, int64_t(n) * kIn)) || (rc = dalloc(&Xn, int64_t(n) * kIn)) ||
        (rc = dalloc(&H1, int64_t(n) * kHid)) || (rc = dalloc(&H2, int64_t(n) * kHid)) ||
        (rc = dalloc(&Y, int64_t(n) * K)) || (rc = dalloc(&T, int64_t(n) * K)) ||
        (rc = dalloc(&M, int64_t(n) * K)) || (rc = dalloc(&D1, int64_t(n) * kHid)) ||
        (rc = dalloc(&D2, int64_t(n) * kHid)) || (rc = dalloc(&D3, int64_t(n) * K)) ||
        (rc = dalloc(&loss, n))) {
      free_batch();
      return rc;
    }
    cap = n;
    return TENTACO_OK;
  }
  ~tentaco_mlp() {
    free_batch();
    for (float* p : {P, G, Mo, Ve, scale}) (void)hipFree(p);
  }
};

namespace tt {
namespace {

// forward (training mode) + deltas + parameter gradients on n samples whose
// raw inputs, raw targets and mask already sit in m->X / m->T / m->M.
int train_pass(tentaco_mlp* m, int n, float inv_count, hipStream_t st) {
  mlp_forward_kernel<<<n, kHid, 0, st>>>(m->P, m->scale, m->K, m->X, n, m->Xn, m->H1, m->H2,
                                         m->Y, 0);
  TT_HIP_CHECK(hipGetLastError());
  mlp_delta_kernel<<<n, kHid, 0, st>>>(m->P, m->scale, m->K, m->Y, m->T, m->M, n, m->H1, m->H2,
                                       inv_count, m->D1, m->D2, m->D3, m->loss);
  TT_HIP_CHECK(hipGetLastError());
  const int grid = int((m->np + 255) / 256);
  mlp_wgrad_kernel<<<grid, 256, 0, st>>>(m->K, n, m->Xn, m->H1, m->H2, m->D1, m->D2, m->D3, m->G);
  TT_HIP_CHECK(hipGetLastError());
  return TENTACO_OK;
}

int adam_step(tentaco_mlp* m, float lr, float b1, float b2, float eps, hipStream_t st) {
  ++m->adam_t;
  const float c1 = float(1.0 - std::pow(double(b1), double(m->adam_t)));
  const float c2 = float(1.0 - std::pow(double(b2), double(m->adam_t)));
  const int grid = int((m->np + 255) / 256);
  mlp_adam_kernel<<<grid, 256, 0, st>>>(m->P, m->G, m->Mo, m->Ve, m->np, lr, b1, b2, eps, c1, c2);
  TT_HIP_CHECK(hipGetLastError());
  return TENTACO_OK;
}

}  // namespace

int csf_features(const tentaco_csf* h, int R, float* f) {
  for (int i = 0; i < kIn; ++i) f[i] = 0.f;
  const int N = h->order;
  f[0] = float(std::log(double(std::max<int64_t>(h->stats.nnz, 1))));
  for (int m = 0; m < N; ++m) {
    f[1 + m] = float(std::log(double(std::max(h->dims[m], 1))));
    f[6 + m] = float(std::log(double(std::max(h->nodes[m], 1))));
  }
  for (int m = 1; m < N && m <= 4; ++m) {
    f[11 + 3 * (m - 1) + 0] = float(std::log1p(h->stats.child_mean[m]));
    f[11 + 3 * (m - 1) + 1] = float(std::log1p(h->stats.child_var[m]));
    f[11 + 3 * (m - 1) + 2] = float(std::log1p(h->stats.child_max[m]));
  }
  f[23] = float(std::log(double(std::max(R, 1))));
  return TENTACO_OK;
}

int mlp_select(tentaco_mlp* m, int order, int R, const float* features, int* variant,
               float* margin, float* pred_ms) {
  if (!m || !features) return fail(TENTACO_EINVAL, "selector: NULL argument");
  if (m->K != kNumVariants)
    return fail(TENTACO_EINVAL, "selector has %d outputs, the variant table %d", m->K,
                kNumVariants);
  int rc = m->ensure(1);
  if (rc) return rc;
  TT_HIP_CHECK(hipMemcpy(m->X, features, sizeof(float) * kIn, hipMemcpyHostToDevice));
  mlp_forward_kernel<<<1, kHid, 0, nullptr>>>(m->P, m->scale, m->K, m->X, 1, nullptr, nullptr,
                                              nullptr, m->Y, 1);
  TT_HIP_CHECK(hipGetLastError());
  std::vector<float> y(static_cast<size_t>(m->K));
  TT_HIP_CHECK(hipMemcpy(y.data(), m->Y, sizeof(float) * m->K, hipMemcpyDeviceToHost));
  int best = -1, second = -1;
  for (int v = 0; v < m->K; ++v) {
    if (!variant_fits(order, R, v)) continue;
    if (best < 0 || y[size_t(v)] < y[size_t(best)]) {
      second = best;
      best = v;
    } else if (second < 0 || y[size_t(v)] < y[size_t(second)]) {
      second = v;
    }
  }
  if (best < 0) return fail(TENTACO_EUNSUPPORTED, "selector: no variant fits");
  if (variant) *variant = best;
  if (margin)
    *margin = second < 0 ? INFINITY : float(std::expm1(double(y[size_t(second)] - y[size_t(best)])));
  if (pred_ms)
    for (int v = 0; v < m->K; ++v) pred_ms[v] = float(std::exp(double(y[size_t(v)])));
  return TENTACO_OK;
}

}  // namespace tt

extern "C" {

int64_t tentaco_mlp_num_params(int n_out) { return n_out > 0 ? n_params(n_out) : 0; }

int tentaco_mlp_create(int n_out, uint64_t seed, tentaco_mlp** out) {
  clear_error();
  if (!out) return fail(TENTACO_EINVAL, "selector: out is NULL");
  *out = nullptr;
  if (n_out < 1 || n_out > kMaxOut) return fail(TENTACO_EINVAL, "selector: %d outputs (1..64)", n_out);
  auto* m = new tentaco_mlp();
  m->K = n_out;
  m->np = n_params(n_out);
  int rc = 0;
  if ((rc = dalloc(&m->P, m->np)) || (rc = dalloc(&m->G, m->np)) || (rc = dalloc(&m->Mo, m->np)) ||
      (rc = dalloc(&m->Ve, m->np)) || (rc = dalloc(&m->scale, 2 * kIn + 2))) {
    delete m;
    return rc;
  }
  // He-normal weights (std sqrt(2 / fan_in)), zero biases
  std::vector<float> p(size_t(m->np), 0.f);
  uint64_t s = seed ^ 0x5DEECE66Dull;
  size_t o = 0;
  auto fill = [&](int rows, int cols) {
    const double sd = std::sqrt(2.0 / cols);
    for (int i = 0; i < rows * cols; ++i) p[o++] = float(sd * gauss(s));
    o += size_t(rows);  // bias
  };
  fill(kHid, kIn);
  fill(kHid, kHid);
  fill(n_out, kHid);
  std::vector<float> zero(size_t(m->np), 0.f);
  for (int i = 0; i < kIn; ++i) {
    m->h_scale[i] = 0.f;
    m->h_scale[kIn + i] = 1.f;
  }
  m->h_scale[2 * kIn] = 0.f;
  m->h_scale[2 * kIn + 1] = 1.f;
  if (hipMemcpy(m->P, p.data(), sizeof(float) * p.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->Mo, zero.data(), sizeof(float) * zero.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->Ve, zero.data(), sizeof(float) * zero.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->G, zero.data(), sizeof(float) * zero.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->scale, m->h_scale, sizeof(m->h_scale), hipMemcpyHostToDevice) != hipSuccess) {
    delete m;
    return fail(TENTACO_EHIP, "selector: initial upload failed");
  }
  *out = m;
  return TENTACO_OK;
}

int tentaco_mlp_free(tentaco_mlp* m) {
  delete m;
  return TENTACO_OK;
}

int tentaco_mlp_set_scaling(tentaco_mlp* m, const float* x_mean, const float* x_std, float y_mean,
                            float y_std) {
  clear_error();
  if (!m || !x_mean || !x_std) return fail(TENTACO_EINVAL, "selector: NULL argument");
  for (int i = 0; i < kIn; ++i)
    if (!(x_std[i] > 0.f)) return fail(TENTACO_EINVAL, "selector: x_std[%d] must be > 0", i);
  if (!(y_std > 0.f)) return fail(TENTACO_EINVAL, "selector: y_std must be > 0");
  std::memcpy(m->h_scale, x_mean, sizeof(float) * kIn);
  std::memcpy(m->h_scale + kIn, x_std, sizeof(float) * kIn);
  m->h_scale[2 * kIn] = y_mean;
  m->h_scale[2 * kIn + 1] = y_std;
  TT_HIP_CHECK(hipMemcpy(m->scale, m->h_scale, sizeof(m->h_scale), hipMemcpyHostToDevice));
  return TENTACO_OK;
}

int tentaco_mlp_get_scaling(const tentaco_mlp* m, float* x_mean, float* x_std, float* y_mean,
                            float* y_std) {
  if (!m) return fail(TENTACO_EINVAL, "selector: NULL model");
  if (x_mean) std::memcpy(x_mean, m->h_scale, sizeof(float) * kIn);
  if (x_std) std::memcpy(x_std, m->h_scale + kIn, sizeof(float) * kIn);
  if (y_mean) *y_mean = m->h_scale[2 * kIn];
  if (y_std) *y_std = m->h_scale[2 * kIn + 1];
  return TENTACO_OK;
}

int tentaco_mlp_get_params(const tentaco_mlp* m, float* host, int64_t n) {
  clear_error();
  if (!m || !host) return fail(TENTACO_EINVAL, "selector: NULL argument");
  if (n != m->np) return fail(TENTACO_EINVAL, "selector: %lld params, model has %lld", (long long)n, (long long)m->np);
  TT_HIP_CHECK(hipMemcpy(host, m->P, sizeof(float) * size_t(n), hipMemcpyDeviceToHost));
  return TENTACO_OK;
}

int tentaco_mlp_set_params(tentaco_mlp* m, const float* host, int64_t n) {
  clear_error();
  if (!m || !host) return fail(TENTACO_EINVAL, "selector: NULL argument");
  if (n != m->np) return fail(TENTACO_EINVAL, "selector: %lld params, model has %lld", (long long)n, (long long)m->np);
  TT_HIP_CHECK(hipMemcpy(m->P, host, sizeof(float) * size_t(n), hipMemcpyHostToDevice));
  return TENTACO_OK;
}

int tentaco_mlp_get_grads(const tentaco_mlp* m, float* host, int64_t n) {
  clear_error();
  if (!m || !host) return fail(TENTACO_EINVAL, "selector: NULL argument");
  if (n != m->np) return fail(TENTACO_EINVAL, "selector: %lld grads, model has %lld", (long long)n, (long long)m->np);
  TT_HIP_CHECK(hipMemcpy(host, m->G, sizeof(float) * size_t(n), hipMemcpyDeviceToHost));
  return TENTACO_OK;
}

int tentaco_mlp_forward(tentaco_mlp* m, const float* x_dev, int n, float* y_dev, void* stream) {
  clear_error();
  if (!m || !x_dev || !y_dev) return fail(TENTACO_EINVAL, "selector: NULL argument");
  if (n < 0) return fail(TENTACO_EINVAL, "selector: n = %d", n);
  if (n == 0) return TENTACO_OK;
  mlp_forward_kernel<<<n, kHid, 0, static_cast<hipStream_t>(stream)>>>(
      m->P, m->scale, m->K, x_dev, n, nullptr, nullptr, nullptr, y_dev, 1);
  TT_HIP_CHECK(hipGetLastError());
  return TENTACO_OK;
}

int tentaco_mlp_loss_grad(tentaco_mlp* m, const float* x, const float* t, const float* mask, int n,
                          float* loss) {
  clear_error();
  if (!m || !x || !t || !mask) return fail(TENTACO_EINVAL, "selector: NULL argument");
  if (n < 1) return fail(TENTACO_EINVAL, "selector: n = %d", n);
  int rc = m->ensure(n);
  if (rc) return rc;
  double count = 0.0;
  for (int64_t q = 0; q < int64_t(n) * m->K; ++q) count += mask[q];
  if (count <= 0.0) return fail(TENTACO_EINVAL, "selector: mask selects nothing");
  TT_HIP_CHECK(hipMemcpy(m->X, x, sizeof(float) * size_t(n) * kIn, hipMemcpyHostToDevice));
  TT_HIP_CHECK(hipMemcpy(m->T, t, sizeof(float) * size_t(n) * m->K, hipMemcpyHostToDevice));
  TT_HIP_CHECK(hipMemcpy(m->M, mask, sizeof(float) * size_t(n) * m->K, hipMemcpyHostToDevice));
  rc = train_pass(m, n, float(1.0 / count), nullptr);
  if (rc) return rc;
  std::vector<float> l(static_cast<size_t>(n));
  TT_HIP_CHECK(hipMemcpy(l.data(), m->loss, sizeof(float) * size_t(n), hipMemcpyDeviceToHost));
  double s = 0.0;
  for (float v : l) s += v;
  if (loss) *loss = float(s / count);
  return TENTACO_OK;
}

int tentaco_mlp_adam(tentaco_mlp* m, float lr, float beta1, float beta2, float eps) {
  clear_error();
  if (!m) return fail(TENTACO_EINVAL, "selector: NULL model");
  return adam_step(m, lr, beta1, beta2, eps, nullptr);
}

int tentaco_mlp_train(tentaco_mlp* m, const float* x, const float* t, const float* mask, int n,
                      int epochs, int batch, float lr, uint64_t seed, float* final_loss) {
  clear_error();
  if (!m || !x || !t || !mask) return fail(TENTACO_EINVAL, "selector: NULL argument");
  if (n < 1 || epochs < 1 || batch < 1) return fail(TENTACO_EINVAL, "selector: n, epochs, batch must be >= 1");
  batch = std::min(batch, n);
  const int K = m->K;
  int rc = m->ensure(batch);
  if (rc) return rc;
  float *dX = nullptr, *dT = nullptr, *dM = nullptr;
  int32_t* dIdx = nullptr;
  struct Guard {
    float **a, **b, **c;
    int32_t** d;
    ~Guard() {
      (void)hipFree(*a);
      (void)hipFree(*b);
      (void)hipFree(*c);
      (void)hipFree(*d);
    }
  } guard{&dX, &dT, &dM, &dIdx};
  if ((rc = dalloc(&dX, int64_t(n) * kIn)) || (rc = dalloc(&dT, int64_t(n) * K)) ||
      (rc = dalloc(&dM, int64_t(n) * K)) || (rc = dalloc(&dIdx, n)))
    return rc;
  TT_HIP_CHECK(hipMemcpy(dX, x, sizeof(float) * size_t(n) * kIn, hipMemcpyHostToDevice));
  TT_HIP_CHECK(hipMemcpy(dT, t, sizeof(float) * size_t(n) * K, hipMemcpyHostToDevice));
  TT_HIP_CHECK(hipMemcpy(dM, mask, sizeof(float) * size_t(n) * K, hipMemcpyHostToDevice));
  std::vector<double> row_count(static_cast<size_t>(n), 0.0);
  for (int b = 0; b < n; ++b)
    for (int o = 0; o < K; ++o) row_count[size_t(b)] += mask[int64_t(b) * K + o];
  std::vector<int32_t> perm(static_cast<size_t>(n));
  const int steps_per_epoch = (n + batch - 1) / batch;
  const int64_t total_steps = int64_t(steps_per_epoch) * epochs;
  int64_t step = 0;
  uint64_t s = seed ^ 0xA0761D6478BD642Full;
  double last_loss = 0.0;
  std::vector<float> l(static_cast<size_t>(batch));
  for (int e = 0; e < epochs; ++e) {
    for (int i = 0; i < n; ++i) perm[size_t(i)] = i;
    for (int i = n - 1; i > 0; --i) {  // Fisher-Yates
      const int j = int(splitmix(s) % uint64_t(i + 1));
      std::swap(perm[size_t(i)], perm[size_t(j)]);
    }
    TT_HIP_CHECK(hipMemcpy(dIdx, perm.data(), sizeof(int32_t) * size_t(n), hipMemcpyHostToDevice));
    double epoch_loss = 0.0, epoch_count = 0.0;
    for (int b0 = 0; b0 < n; b0 += batch) {
      const int nb = std::min(batch, n - b0);
      double count = 0.0;
      for (int i = 0; i < nb; ++i) count += row_count[size_t(perm[size_t(b0 + i)])];
      const int64_t work = int64_t(nb) * (kIn + 2 * K);
      mlp_gather_kernel<<<int((work + 255) / 256), 256, 0, nullptr>>>(dIdx + b0, nb, K, dX, dT, dM,
                                                                     m->X, m->T, m->M);
      TT_HIP_CHECK(hipGetLastError());
      if (count > 0.0) {
        rc = train_pass(m, nb, float(1.0 / count), nullptr);
        if (rc) return rc;
        // cosine decay of the step size over the whole run
        const double frac = double(step) / double(std::max<int64_t>(total_steps, 1));
        const float lr_t = float(lr * 0.5 * (1.0 + std::cos(3.141592653589793 * frac)));
        rc = adam_step(m, lr_t, 0.9f, 0.999f, 1e-8f, nullptr);
        if (rc) return rc;
        if (e == epochs - 1) {
          TT_HIP_CHECK(hipMemcpy(l.data(), m->loss, sizeof(float) * size_t(nb), hipMemcpyDeviceToHost));
          for (int i = 0; i < nb; ++i) epoch_loss += l[size_t(i)];
          epoch_count += count;
        }
      }
      ++step;
    }
    if (e == epochs - 1) last_loss = epoch_count > 0.0 ? epoch_loss / epoch_count : 0.0;
  }
  TT_HIP_CHECK(hipDeviceSynchronize());
  if (final_loss) *final_loss = float(last_loss);
  return TENTACO_OK;
}

int tentaco_mlp_select(tentaco_mlp* m, int order, int R, const float* features, int* variant,
                       float* margin, float* pred_ms) {
  clear_error();
  if (order < 2 || order > kMaxOrder) return fail(TENTACO_EINVAL, "selector: order %d", order);
  return mlp_select(m, order, R, features, variant, margin, pred_ms);
}

int tentaco_select_variant(tentaco_mlp* m, const float* features, int order, int R,
                           int* variant) {
  return tentaco_mlp_select(m, order, R, features, variant, nullptr, nullptr);
}

int tentaco_csf_features(const tentaco_csf* h, int R, float* features) {
  clear_error();
  if (!h || !features) return fail(TENTACO_EINVAL, "features: NULL argument");
  if (R <= 0) return fail(TENTACO_EINVAL, "features: R = %d", R);
  return csf_features(h, R, features);
}

}  // extern "C"
