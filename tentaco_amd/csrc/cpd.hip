// CP-ALS around the batched MTTKRP: the caller that splatt_cpd_als stands in
// for in the reference (ds.h:395), for many tensors at once.
//
// One ALS sweep updates every factor of every tensor in mode order
// m = 0 .. 4. Mode step m runs
//   1. one batched, graph-replayed MTTKRP over every tensor with order > m
//      (M_t = B_t x_{k != m} A_t,k, from the tensor's mode-m CSF),
//   2. per tensor: V = Hadamard product of the other modes' Grams, its
//      inverse by Cholesky (FP64, in LDS), A_t,m = M_t V^-1,
//   3. column normalisation (2-norm on the first sweep, max(1, max |a|)
//      afterwards) into lambda_t, and the new Gram A^T A.
// Row-parallel kernels cover every tensor of a step in one launch through a
// job table; their reductions write per-block partials that a second kernel
// sums in block order, so results do not depend on scheduling. The fit
// after a sweep is 1 - ||X - model|| / ||X|| with
// ||X - model||^2 = ||X||^2 + lambda^T (*_k G_k) lambda - 2 <X, model> and
// <X, model> = sum_r lambda_r sum_i M_N[i,r] A_N[i,r] for the last mode N.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "handle.h"
#include "internal.h"

namespace tt {
namespace {

constexpr int kMaxR = 64;
constexpr int kApplyRows = 64;    // rows per block of the apply / scale kernels
constexpr int kStatRows = 2048;   // rows per block of the reduction kernels

struct CpdJob {
  const float* M = nullptr;  // MTTKRP result, rows x R
  float* A = nullptr;        // factor being updated, rows x R
  float* G = nullptr;        // its Gram, R x R
  float* lam = nullptr;      // lambda, R
  float* Vinv = nullptr;     // R x R scratch
  const float* Gk[kMaxOrder - 1] = {};
  int nG = 0;
  int32_t rows = 0;
  int32_t ablk0 = 0, nablk = 0;  // apply blocks (kApplyRows rows each)
  int32_t sblk0 = 0, nsblk = 0;  // reduction blocks (kStatRows rows each)
};

__device__ __forceinline__ int find_job(const int32_t* __restrict__ off, int n, int32_t blk) {
  int lo = 0, hi = n - 1;  // largest j with off[j] <= blk
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= blk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// V = Hadamard product of the job's Grams; Vinv = V^-1 through Cholesky in
// FP64. A non-positive pivot (a rank-deficient V) is lifted to 1e-12 times
// the largest diagonal, which keeps the solve finite.
__global__ __launch_bounds__(256) void hadamard_inverse_kernel(const CpdJob* __restrict__ jobs,
                                                               int R) {
  __shared__ double L[kMaxR][kMaxR + 1];
  __shared__ float W[kMaxR][kMaxR + 1];  // L^-1 (lower triangle)
  __shared__ double dmax;
  const CpdJob& J = jobs[blockIdx.x];
  for (int e = threadIdx.x; e < R * R; e += blockDim.x) {
    const int i = e / R, j = e % R;
    double v = 1.0;
    for (int g = 0; g < J.nG; ++g) v *= double(J.Gk[g][e]);
    L[i][j] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int i = 0; i < R; ++i) m = fmax(m, L[i][i]);
    dmax = m > 0.0 ? m : 1.0;
  }
  __syncthreads();
  for (int k = 0; k < R; ++k) {
    if (threadIdx.x == 0) {
      double d = L[k][k];
      for (int j = 0; j < k; ++j) d -= L[k][j] * L[k][j];
      L[k][k] = sqrt(d > 1e-12 * dmax ? d : 1e-12 * dmax);
    }
    __syncthreads();
    for (int i = k + 1 + threadIdx.x; i < R; i += blockDim.x) {
      double s = L[i][k];
      for (int j = 0; j < k; ++j) s -= L[i][j] * L[k][j];
      L[i][k] = s / L[k][k];
    }
    __syncthreads();
  }
  // column c of L^-1 by forward substitution (one thread per column)
  for (int c = threadIdx.x; c < R; c += blockDim.x) {
    for (int i = 0; i < R; ++i) W[i][c] = 0.f;
    W[c][c] = float(1.0 / L[c][c]);
    for (int i = c + 1; i < R; ++i) {
      double s = 0.0;
      for (int j = c; j < i; ++j) s += L[i][j] * double(W[j][c]);
      W[i][c] = float(-s / L[i][i]);
    }
  }
  __syncthreads();
  // V^-1 = L^-T L^-1
  for (int e = threadIdx.x; e < R * R; e += blockDim.x) {
    const int i = e / R, j = e % R;
    double s = 0.0;
    for (int k = max(i, j); k < R; ++k) s += double(W[k][i]) * double(W[k][j]);
    J.Vinv[e] = float(s);
  }
}

// A = M V^-1 for kApplyRows rows of one job per block.
__global__ __launch_bounds__(256) void apply_kernel(const CpdJob* __restrict__ jobs,
                                                    const int32_t* __restrict__ ablk0, int njobs,
                                                    int R) {
  __shared__ float Vs[kMaxR * kMaxR];
  __shared__ float Ms[kApplyRows][kMaxR + 1];
  const int j = find_job(ablk0, njobs, blockIdx.x);
  const CpdJob& J = jobs[j];
  const int32_t r0 = (int32_t(blockIdx.x) - J.ablk0) * kApplyRows;
  const int nr = min(kApplyRows, J.rows - r0);
  for (int e = threadIdx.x; e < R * R; e += blockDim.x) Vs[e] = J.Vinv[e];
  for (int e = threadIdx.x; e < nr * R; e += blockDim.x)
    Ms[e / R][e % R] = J.M[int64_t(r0 + e / R) * R + e % R];
  __syncthreads();
  for (int e = threadIdx.x; e < nr * R; e += blockDim.x) {
    const int r = e / R, c = e % R;
    float s = 0.f;
    for (int k = 0; k < R; ++k) s = fmaf(Ms[r][k], Vs[k * R + c], s);
    J.A[int64_t(r0 + r) * R + c] = s;
  }
}

// Per reduction block: column statistics of A (sum of squares, or max |a|)
// into part[block * R + c], in FP64.
__global__ __launch_bounds__(256) void colstat_kernel(const CpdJob* __restrict__ jobs,
                                                      const int32_t* __restrict__ sblk0, int njobs,
                                                      int R, int use_max,
                                                      double* __restrict__ part) {
  __shared__ double red[256];
  const int j = find_job(sblk0, njobs, blockIdx.x);
  const CpdJob& J = jobs[j];
  const int32_t r0 = (int32_t(blockIdx.x) - J.sblk0) * kStatRows;
  const int32_t r1 = min(r0 + kStatRows, J.rows);
  const int lanes = blockDim.x / R;  // row lanes per column
  const int c = threadIdx.x % R, l = threadIdx.x / R;
  double acc = 0.0;
  if (l < lanes)
    for (int32_t r = r0 + l; r < r1; r += lanes) {
      const double a = J.A[int64_t(r) * R + c];
      acc = use_max ? fmax(acc, fabs(a)) : acc + a * a;
    }
  red[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < R) {
    double s = red[threadIdx.x];
    for (int q = 1; q < lanes; ++q) {
      const double v = red[q * R + threadIdx.x];
      s = use_max ? fmax(s, v) : s + v;
    }
    part[int64_t(blockIdx.x) * R + threadIdx.x] = s;
  }
}

// lambda from the job's partials (in block order); then A /= lambda.
__global__ __launch_bounds__(64) void lambda_kernel(const CpdJob* __restrict__ jobs, int R,
                                                    int use_max,
                                                    const double* __restrict__ part) {
  const CpdJob& J = jobs[blockIdx.x];
  for (int c = threadIdx.x; c < R; c += blockDim.x) {
    double s = 0.0;
    for (int b = 0; b < J.nsblk; ++b) {
      const double v = part[int64_t(J.sblk0 + b) * R + c];
      s = use_max ? fmax(s, v) : s + v;
    }
    const double lam = use_max ? fmax(s, 1.0) : sqrt(s);
    J.lam[c] = float(lam > 0.0 ? lam : 1.0);
  }
}

__global__ __launch_bounds__(256) void scale_kernel(const CpdJob* __restrict__ jobs,
                                                    const int32_t* __restrict__ ablk0, int njobs,
                                                    int R) {
  const int j = find_job(ablk0, njobs, blockIdx.x);
  const CpdJob& J = jobs[j];
  const int32_t r0 = (int32_t(blockIdx.x) - J.ablk0) * kApplyRows;
  const int nr = min(kApplyRows, J.rows - r0);
  for (int e = threadIdx.x; e < nr * R; e += blockDim.x) {
    const int64_t q = int64_t(r0) * R + e;
    J.A[q] = J.A[q] / J.lam[e % R];
  }
}

// Gram partials: part[block][i*R + j] = sum over the block's rows of
// A[r,i] A[r,j] (upper triangle computed, mirrored), FP64 accumulation.
__global__ __launch_bounds__(256) void gram_partial_kernel(const CpdJob* __restrict__ jobs,
                                                           const int32_t* __restrict__ sblk0,
                                                           int njobs, int R,
                                                           double* __restrict__ part) {
  constexpr int kTile = 32;
  __shared__ float As[kTile][kMaxR + 1];
  const int j = find_job(sblk0, njobs, blockIdx.x);
  const CpdJob& J = jobs[j];
  const int32_t r0 = (int32_t(blockIdx.x) - J.sblk0) * kStatRows;
  const int32_t r1 = min(r0 + kStatRows, J.rows);
  constexpr int kPer = (kMaxR * kMaxR + 255) / 256;
  double acc[kPer];
#pragma unroll
  for (int q = 0; q < kPer; ++q) acc[q] = 0.0;
  for (int32_t t0 = r0; t0 < r1; t0 += kTile) {
    const int nt = min(kTile, r1 - t0);
    __syncthreads();
    for (int e = threadIdx.x; e < kTile * R; e += blockDim.x) {
      const int r = e / R, c = e % R;
      As[r][c] = r < nt ? J.A[int64_t(t0 + r) * R + c] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int e = threadIdx.x + q * 256;
      if (e >= R * R) break;
      const int a = e / R, b = e % R;
      if (a > b) continue;
      double s = 0.0;
      for (int r = 0; r < kTile; ++r) s += double(As[r][a]) * double(As[r][b]);
      acc[q] += s;
    }
  }
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int e = threadIdx.x + q * 256;
    if (e >= R * R) break;
    const int a = e / R, b = e % R;
    if (a > b) continue;
    part[int64_t(blockIdx.x) * R * R + a * R + b] = acc[q];
    part[int64_t(blockIdx.x) * R * R + b * R + a] = acc[q];
  }
}

__global__ __launch_bounds__(256) void gram_reduce_kernel(const CpdJob* __restrict__ jobs, int R,
                                                          const double* __restrict__ part) {
  const CpdJob& J = jobs[blockIdx.x];
  for (int e = threadIdx.x; e < R * R; e += blockDim.x) {
    double s = 0.0;
    for (int b = 0; b < J.nsblk; ++b) s += part[int64_t(J.sblk0 + b) * R * R + e];
    J.G[e] = float(s);
  }
}

// <M, A diag(lambda)> partials per reduction block (fit).
__global__ __launch_bounds__(256) void inner_kernel(const CpdJob* __restrict__ jobs,
                                                    const int32_t* __restrict__ sblk0, int njobs,
                                                    int R, double* __restrict__ part) {
  __shared__ double red[256];
  const int j = find_job(sblk0, njobs, blockIdx.x);
  const CpdJob& J = jobs[j];
  const int32_t r0 = (int32_t(blockIdx.x) - J.sblk0) * kStatRows;
  const int32_t r1 = min(r0 + kStatRows, J.rows);
  double acc = 0.0;
  for (int64_t q = int64_t(r0) * R + threadIdx.x; q < int64_t(r1) * R; q += blockDim.x)
    acc += double(J.M[q]) * double(J.A[q]) * double(J.lam[q % R]);
  red[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int t = 0; t < 256; ++t) s += red[t];
    part[blockIdx.x] = s;
  }
}

template <class T>
int dalloc(T** p, int64_t n) {
  *p = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(p), size_t(std::max<int64_t>(n, 1)) * sizeof(T)) !=
      hipSuccess) {
    *p = nullptr;
    return fail(TENTACO_ENOMEM, "cpd: hipMalloc of %lld elements failed", (long long)n);
  }
  return TENTACO_OK;
}

}  // namespace
}  // namespace tt

using namespace tt;

struct tentaco_cpd {
  int nt = 0, R = 0;
  std::vector<int> order;
  std::vector<std::vector<int32_t>> dims;                  // per tensor, tensor mode
  std::vector<std::vector<std::vector<int>>> lvl;          // [t][m][level] -> tensor mode
  std::vector<std::vector<tentaco_csf*>> h;                // [t][m]
  std::vector<std::vector<float*>> A, G;                   // [t][mode]
  std::vector<float*> lam, M, Vinv;                        // [t]
  std::vector<double> xnorm2;                              // ||X_t||^2
  int sweeps = 0;
  // per mode step
  struct Step {
    std::vector<int> tensors;
    tentaco_batch* batch = nullptr;
    bool captured = false;
    CpdJob* d_jobs = nullptr;
    int32_t *d_ablk0 = nullptr, *d_sblk0 = nullptr;
    int32_t nablk = 0, nsblk = 0;
    std::vector<CpdJob> jobs;
  } step[kMaxOrder];
  // jobs for the Grams of every factor (initialisation)
  double* part = nullptr;
  int64_t part_len = 0;
  hipStream_t own = nullptr;

  ~tentaco_cpd() {
    for (auto& s : step) {
      if (s.batch) tentaco_batch_free(s.batch);
      (void)hipFree(s.d_jobs);
      (void)hipFree(s.d_ablk0);
      (void)hipFree(s.d_sblk0);
    }
    for (auto& v : A)
      for (float* p : v) (void)hipFree(p);
    for (auto& v : G)
      for (float* p : v) (void)hipFree(p);
    for (float* p : lam) (void)hipFree(p);
    for (float* p : M) (void)hipFree(p);
    for (float* p : Vinv) (void)hipFree(p);
    (void)hipFree(part);
    if (own) (void)hipStreamDestroy(own);
  }
};

namespace tt {
namespace {

// Fill a job table for `tensors` updating mode m; out-of-place Grams are the
// other modes' Grams.
int build_jobs(tentaco_cpd* c, int m, const std::vector<int>& tensors, std::vector<CpdJob>& jobs,
               std::vector<int32_t>& ablk0, std::vector<int32_t>& sblk0, int32_t& nablk,
               int32_t& nsblk) {
  jobs.clear();
  ablk0.clear();
  sblk0.clear();
  nablk = nsblk = 0;
  for (int t : tensors) {
    CpdJob J;
    J.M = c->M[size_t(t)];
    J.A = c->A[size_t(t)][size_t(m)];
    J.G = c->G[size_t(t)][size_t(m)];
    J.lam = c->lam[size_t(t)];
    J.Vinv = c->Vinv[size_t(t)];
    for (int k = 0; k < c->order[size_t(t)]; ++k)
      if (k != m) J.Gk[J.nG++] = c->G[size_t(t)][size_t(k)];
    J.rows = c->dims[size_t(t)][size_t(m)];
    J.ablk0 = nablk;
    J.nablk = (J.rows + kApplyRows - 1) / kApplyRows;
    J.sblk0 = nsblk;
    J.nsblk = (J.rows + kStatRows - 1) / kStatRows;
    ablk0.push_back(J.ablk0);
    sblk0.push_back(J.sblk0);
    nablk += J.nablk;
    nsblk += J.nsblk;
    jobs.push_back(J);
  }
  return TENTACO_OK;
}

int upload_jobs(const std::vector<CpdJob>& jobs, const std::vector<int32_t>& ablk0,
                const std::vector<int32_t>& sblk0, CpdJob** dj, int32_t** da, int32_t** ds) {
  (void)hipFree(*dj);
  (void)hipFree(*da);
  (void)hipFree(*ds);
  *dj = nullptr;
  *da = *ds = nullptr;
  int rc = 0;
  if ((rc = dalloc(dj, int64_t(jobs.size()))) || (rc = dalloc(da, int64_t(ablk0.size()))) ||
      (rc = dalloc(ds, int64_t(sblk0.size()))))
    return rc;
  TT_HIP_CHECK(hipMemcpy(*dj, jobs.data(), sizeof(CpdJob) * jobs.size(), hipMemcpyHostToDevice));
  TT_HIP_CHECK(hipMemcpy(*da, ablk0.data(), sizeof(int32_t) * ablk0.size(), hipMemcpyHostToDevice));
  TT_HIP_CHECK(hipMemcpy(*ds, sblk0.data(), sizeof(int32_t) * sblk0.size(), hipMemcpyHostToDevice));
  return TENTACO_OK;
}

int ensure_part(tentaco_cpd* c, int64_t n) {
  if (n <= c->part_len) return TENTACO_OK;
  // only reached before any kernel reads the buffer (create)
  (void)hipFree(c->part);
  c->part = nullptr;
  c->part_len = 0;
  int rc = dalloc(&c->part, n);
  if (rc) return rc;
  c->part_len = n;
  return TENTACO_OK;
}

// Grams of the jobs' A (all rows), on st.
int grams(tentaco_cpd* c, const CpdJob* dj, const int32_t* ds, int njobs, int32_t nsblk,
          hipStream_t st) {
  if (njobs == 0 || nsblk == 0) return TENTACO_OK;
  const int R = c->R;
  int rc = ensure_part(c, int64_t(nsblk) * R * R);
  if (rc) return rc;
  gram_partial_kernel<<<nsblk, 256, 0, st>>>(dj, ds, njobs, R, c->part);
  TT_HIP_CHECK(hipGetLastError());
  gram_reduce_kernel<<<njobs, 256, 0, st>>>(dj, R, c->part);
  TT_HIP_CHECK(hipGetLastError());
  return TENTACO_OK;
}

}  // namespace
}  // namespace tt

extern "C" {

int tentaco_cpd_create(tentaco_csf* const* handles, const int* level_modes, int ntensors, int R,
                       uint64_t seed, tentaco_cpd** out) {
  clear_error();
  if (!handles || !level_modes || !out || ntensors < 1)
    return fail(TENTACO_EINVAL, "cpd: NULL argument or no tensors");
  *out = nullptr;
  if (R != 16 && R != 32 && R != 64) return fail(TENTACO_EUNSUPPORTED, "cpd: R must be 16, 32 or 64");
  auto* c = new tentaco_cpd();
  c->nt = ntensors;
  c->R = R;
  auto bail = [&](int rc) {
    delete c;
    return rc;
  };
  for (int t = 0; t < ntensors; ++t) {
    const tentaco_csf* h0 = handles[size_t(t) * kMaxOrder];
    if (!h0) return bail(fail(TENTACO_EINVAL, "cpd: tensor %d has no mode-0 CSF", t));
    const int N = h0->order;
    c->order.push_back(N);
    std::vector<int32_t> dims(size_t(N), 0);
    std::vector<std::vector<int>> lv(static_cast<size_t>(N));
    std::vector<tentaco_csf*> hs(static_cast<size_t>(N));
    for (int m = 0; m < N; ++m) {
      tentaco_csf* h = handles[size_t(t) * kMaxOrder + m];
      if (!h || h->order != N) return bail(fail(TENTACO_EINVAL, "cpd: tensor %d mode %d CSF missing or of another order", t, m));
      const int* L = level_modes + (size_t(t) * kMaxOrder + m) * kMaxOrder;
      bool seen[kMaxOrder] = {};
      if (L[0] != m) return bail(fail(TENTACO_EINVAL, "cpd: tensor %d mode %d CSF must have mode %d at its root", t, m, m));
      for (int l = 0; l < N; ++l) {
        if (L[l] < 0 || L[l] >= N || seen[L[l]]) return bail(fail(TENTACO_EINVAL, "cpd: tensor %d mode %d level map is not a permutation", t, m));
        seen[L[l]] = true;
        lv[size_t(m)].push_back(L[l]);
      }
      hs[size_t(m)] = h;
    }
    // mode dims from the level dims; every CSF must agree
    for (int m = 0; m < N; ++m)
      for (int l = 0; l < N; ++l) {
        const int md = lv[size_t(m)][size_t(l)];
        const int32_t d = hs[size_t(m)]->dims[l];
        if (dims[size_t(md)] == 0) dims[size_t(md)] = d;
        else if (dims[size_t(md)] != d) return bail(fail(TENTACO_EINVAL, "cpd: tensor %d mode %d dimension disagrees between CSFs", t, md));
      }
    if (hs[0]->nodes[N - 1] <= 0) return bail(fail(TENTACO_EINVAL, "cpd: tensor %d is empty", t));
    c->dims.push_back(dims);
    c->lvl.push_back(lv);
    c->h.push_back(hs);
  }
  // factors (U[0,1) from seed, like gen_factor_matrices, ds.h:386-409),
  // lambda = 1, scratch; ||X||^2 from the mode-0 CSF's values
  uint64_t s = seed ^ 0x2545F4914F6CDD1Dull;
  auto next = [&]() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  c->A.resize(size_t(ntensors));
  c->G.resize(size_t(ntensors));
  for (int t = 0; t < ntensors; ++t) {
    const int N = c->order[size_t(t)];
    int32_t maxd = 0;
    for (int m = 0; m < N; ++m) {
      const int32_t d = c->dims[size_t(t)][size_t(m)];
      maxd = std::max(maxd, d);
      std::vector<float> host(size_t(d) * R);
      for (auto& v : host) v = float(double(next() >> 11) * (1.0 / 9007199254740992.0));
      float *a = nullptr, *g = nullptr;
      int rc = 0;
      if ((rc = dalloc(&a, int64_t(d) * R)) || (rc = dalloc(&g, int64_t(R) * R))) {
        (void)hipFree(a);
        return bail(rc);
      }
      c->A[size_t(t)].push_back(a);
      c->G[size_t(t)].push_back(g);
      if (hipMemcpy(a, host.data(), sizeof(float) * host.size(), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(TENTACO_EHIP, "cpd: factor upload failed"));
    }
    float *lam = nullptr, *Mt = nullptr, *vi = nullptr;
    int rc = 0;
    if ((rc = dalloc(&lam, R)) || (rc = dalloc(&Mt, int64_t(maxd) * R)) || (rc = dalloc(&vi, int64_t(R) * R))) {
      (void)hipFree(lam);
      (void)hipFree(Mt);
      return bail(rc);
    }
    c->lam.push_back(lam);
    c->M.push_back(Mt);
    c->Vinv.push_back(vi);
    std::vector<float> ones(size_t(R), 1.f);
    if (hipMemcpy(lam, ones.data(), sizeof(float) * R, hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(TENTACO_EHIP, "cpd: lambda upload failed"));
    // ||X||^2 (host copy of the values of the mode-0 CSF)
    const tentaco_csf* h0 = c->h[size_t(t)][0];
    const int64_t nnz = h0->nodes[N - 1];
    std::vector<float> vals(static_cast<size_t>(nnz));
    if (hipMemcpy(vals.data(), h0->d_vals, sizeof(float) * size_t(nnz), hipMemcpyDeviceToHost) != hipSuccess)
      return bail(fail(TENTACO_EHIP, "cpd: value download failed"));
    double x2 = 0.0;
    for (float v : vals) x2 += double(v) * double(v);
    c->xnorm2.push_back(x2);
  }
  // mode steps: one batch and one job table each
  for (int m = 0; m < kMaxOrder; ++m) {
    auto& S = c->step[m];
    for (int t = 0; t < ntensors; ++t)
      if (c->order[size_t(t)] > m) S.tensors.push_back(t);
    if (S.tensors.empty()) continue;
    std::vector<tentaco_csf*> tasks;
    std::vector<const float*> facs;
    std::vector<float*> outs;
    std::vector<int32_t> rows;
    for (int t : S.tensors) {
      tasks.push_back(c->h[size_t(t)][size_t(m)]);
      const auto& L = c->lvl[size_t(t)][size_t(m)];
      for (int l = 1; l < kMaxOrder; ++l)
        facs.push_back(l < int(L.size()) ? c->A[size_t(t)][size_t(L[size_t(l)])] : nullptr);
      outs.push_back(c->M[size_t(t)]);
      rows.push_back(c->dims[size_t(t)][size_t(m)]);
    }
    int rc = tentaco_batch_create(tasks.data(), int(tasks.size()), R, &S.batch);
    if (!rc) rc = tentaco_batch_set_operands(S.batch, facs.data(), outs.data(), rows.data());
    if (!rc) rc = tentaco_batch_configure(S.batch, 1, 128, 64);
    if (rc) return bail(rc);
    std::vector<int32_t> ab, sb;
    build_jobs(c, m, S.tensors, S.jobs, ab, sb, S.nablk, S.nsblk);
    rc = upload_jobs(S.jobs, ab, sb, &S.d_jobs, &S.d_ablk0, &S.d_sblk0);
    if (rc) return bail(rc);
  }
  // the reduction partials, sized once for the largest step (never
  // reallocated while kernels may still read them)
  {
    int64_t need = 1;
    for (auto& S : c->step) need = std::max<int64_t>(need, int64_t(S.nsblk) * R * R);
    int rc = ensure_part(c, need);
    if (rc) return bail(rc);
  }
  // initial Grams of every factor
  for (int m = 0; m < kMaxOrder; ++m) {
    auto& S = c->step[m];
    if (S.tensors.empty()) continue;
    int rc = grams(c, S.d_jobs, S.d_sblk0, int(S.jobs.size()), S.nsblk, nullptr);
    if (rc) return bail(rc);
  }
  TT_HIP_CHECK(hipDeviceSynchronize());
  *out = c;
  return TENTACO_OK;
}

int tentaco_cpd_tune(tentaco_cpd* c, int reps) {
  clear_error();
  if (!c) return fail(TENTACO_EINVAL, "cpd: NULL handle");
  for (auto& S : c->step) {
    if (!S.batch) continue;
    int f, b, g;
    float ms;
    int rc = tentaco_batch_autotune(S.batch, reps, nullptr, &f, &b, &g, &ms);
    if (rc) return rc;
    S.captured = false;
  }
  return TENTACO_OK;
}

int tentaco_cpd_iterate(tentaco_cpd* c, int sweeps, int use_graphs, void* stream, double* fits) {
  clear_error();
  if (!c) return fail(TENTACO_EINVAL, "cpd: NULL handle");
  if (sweeps < 1) return fail(TENTACO_EINVAL, "cpd: sweeps must be >= 1");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (use_graphs && !st) {
    if (!c->own) TT_HIP_CHECK(hipStreamCreate(&c->own));
    st = c->own;
  }
  const int R = c->R;
  std::vector<double> inner(size_t(c->nt), 0.0);
  for (int it = 0; it < sweeps; ++it) {
    const bool last = it == sweeps - 1;
    const int use_max = c->sweeps > 0 ? 1 : 0;
    for (int m = 0; m < kMaxOrder; ++m) {
      auto& S = c->step[m];
      if (S.tensors.empty()) continue;
      const int nj = int(S.jobs.size());
      int rc;
      if (use_graphs) {
        if (!S.captured) {
          rc = tentaco_batch_capture(S.batch, st);
          if (rc) return rc;
          S.captured = true;
        }
        rc = tentaco_batch_launch(S.batch, st);
      } else {
        rc = tentaco_batch_run(S.batch, st);
      }
      if (rc) return rc;
      hadamard_inverse_kernel<<<nj, 256, 0, st>>>(S.d_jobs, R);
      TT_HIP_CHECK(hipGetLastError());
      apply_kernel<<<S.nablk, 256, 0, st>>>(S.d_jobs, S.d_ablk0, nj, R);
      TT_HIP_CHECK(hipGetLastError());
      colstat_kernel<<<S.nsblk, 256, 0, st>>>(S.d_jobs, S.d_sblk0, nj, R, use_max, c->part);
      TT_HIP_CHECK(hipGetLastError());
      lambda_kernel<<<nj, 64, 0, st>>>(S.d_jobs, R, use_max, c->part);
      TT_HIP_CHECK(hipGetLastError());
      scale_kernel<<<S.nablk, 256, 0, st>>>(S.d_jobs, S.d_ablk0, nj, R);
      TT_HIP_CHECK(hipGetLastError());
      rc = grams(c, S.d_jobs, S.d_sblk0, nj, S.nsblk, st);
      if (rc) return rc;
      if (last && fits) {
        // <X, model> for the tensors whose last mode is m
        inner_kernel<<<S.nsblk, 256, 0, st>>>(S.d_jobs, S.d_sblk0, nj, R, c->part);
        TT_HIP_CHECK(hipGetLastError());
        std::vector<double> p(size_t(S.nsblk));
        TT_HIP_CHECK(hipMemcpyAsync(p.data(), c->part, sizeof(double) * p.size(),
                                    hipMemcpyDeviceToHost, st));
        TT_HIP_CHECK(hipStreamSynchronize(st));
        for (size_t k = 0; k < S.jobs.size(); ++k) {
          const int t = S.tensors[k];
          if (c->order[size_t(t)] - 1 != m) continue;
          double s = 0.0;
          for (int b = 0; b < S.jobs[k].nsblk; ++b) s += p[size_t(S.jobs[k].sblk0 + b)];
          inner[size_t(t)] = s;
        }
      }
    }
    ++c->sweeps;
  }
  TT_HIP_CHECK(hipStreamSynchronize(st));
  if (fits) {
    std::vector<float> g(static_cast<size_t>(R) * R), lam(static_cast<size_t>(R));
    for (int t = 0; t < c->nt; ++t) {
      std::vector<double> had(size_t(R) * R, 1.0);
      for (int m = 0; m < c->order[size_t(t)]; ++m) {
        TT_HIP_CHECK(hipMemcpy(g.data(), c->G[size_t(t)][size_t(m)], sizeof(float) * g.size(),
                               hipMemcpyDeviceToHost));
        for (size_t e = 0; e < g.size(); ++e) had[e] *= double(g[e]);
      }
      TT_HIP_CHECK(hipMemcpy(lam.data(), c->lam[size_t(t)], sizeof(float) * R, hipMemcpyDeviceToHost));
      double mnorm = 0.0;
      for (int a = 0; a < R; ++a)
        for (int b = 0; b < R; ++b) mnorm += double(lam[size_t(a)]) * double(lam[size_t(b)]) * had[size_t(a) * R + b];
      const double resid = std::max(0.0, c->xnorm2[size_t(t)] + mnorm - 2.0 * inner[size_t(t)]);
      fits[t] = c->xnorm2[size_t(t)] > 0.0 ? 1.0 - std::sqrt(resid) / std::sqrt(c->xnorm2[size_t(t)]) : 0.0;
    }
  }
  return TENTACO_OK;
}

int tentaco_cpd_factor(const tentaco_cpd* c, int t, int m, float* factor, float* lambda) {
  clear_error();
  if (!c || t < 0 || t >= c->nt) return fail(TENTACO_EINVAL, "cpd: bad tensor index");
  if (m < 0 || m >= c->order[size_t(t)]) return fail(TENTACO_EINVAL, "cpd: bad mode");
  if (factor)
    TT_HIP_CHECK(hipMemcpy(factor, c->A[size_t(t)][size_t(m)],
                           sizeof(float) * size_t(c->dims[size_t(t)][size_t(m)]) * c->R,
                           hipMemcpyDeviceToHost));
  if (lambda) TT_HIP_CHECK(hipMemcpy(lambda, c->lam[size_t(t)], sizeof(float) * c->R, hipMemcpyDeviceToHost));
  return TENTACO_OK;
}

int tentaco_cpd_set_factor(tentaco_cpd* c, int t, int m, const float* factor) {
  clear_error();
  if (!c || !factor || t < 0 || t >= c->nt) return fail(TENTACO_EINVAL, "cpd: bad argument");
  if (m < 0 || m >= c->order[size_t(t)]) return fail(TENTACO_EINVAL, "cpd: bad mode");
  TT_HIP_CHECK(hipMemcpy(c->A[size_t(t)][size_t(m)], factor,
                         sizeof(float) * size_t(c->dims[size_t(t)][size_t(m)]) * c->R,
                         hipMemcpyHostToDevice));
  // refresh this factor's Gram through a one-job table
  std::vector<CpdJob> jobs;
  std::vector<int32_t> ab, sb;
  int32_t na = 0, ns = 0;
  std::vector<int> one{t};
  build_jobs(c, m, one, jobs, ab, sb, na, ns);
  CpdJob* dj = nullptr;
  int32_t *da = nullptr, *ds = nullptr;
  int rc = upload_jobs(jobs, ab, sb, &dj, &da, &ds);
  if (!rc) rc = grams(c, dj, ds, 1, ns, nullptr);
  if (!rc && hipDeviceSynchronize() != hipSuccess) rc = fail(TENTACO_EHIP, "cpd: synchronize failed");
  (void)hipFree(dj);
  (void)hipFree(da);
  (void)hipFree(ds);
  return rc;
}

int tentaco_cpd_free(tentaco_cpd* c) {
  delete c;
  return TENTACO_OK;
}

}  // extern "C"
