// CSF TTV for gfx950: y(f) = sum_k B(f,k) * c(k) over the level N-2 positions
// f of B (the (i,j) fibers of a 3-mode tensor), i.e. a CSR SpMV whose rows
// are the innermost fibers. Replaces ttv_kernel / ttv_csf_gpu_taco
// (tensor_kernels/ttv_csf_gpu_taco.h:2-121).
//
// Work split: each workgroup owns 8*BS consecutive nonzeros, each lane 8
// consecutive ones (two dwordx4 loads of crd and of vals per lane, fully
// coalesced per wavefront). The nonzero loads and the c gathers are issued
// first, so their latency overlaps the bracket load and the staging of the
// workgroup's fiber bounds in LDS. A lane sums its nonzeros per fiber; fibers that
// start and end inside a lane are stored directly, the lane's first and last
// partial sums are merged across the wavefront by a segmented shuffle scan
// keyed by fiber id, and the lane holding a fiber's last nonzero writes the
// total. Only the wavefront's first fiber (when it began in an earlier
// wavefront) and its open last fiber are added atomically into the
// zero-filled output. A prep launch zero-fills y and scatters, per fiber, the
// workgroup boundaries that fall inside it (no binary-search pre-pass).

#include <climits>

#include "internal.h"
#include "rowops.h"

namespace tt {
namespace {

constexpr int kTtvBlock = 256;
constexpr int kTtvPerLane = 8;
constexpr int kTtvNB = kTtvBlock * kTtvPerLane;

// y[0..nf) = 0; brk[b] = fiber containing min(b*NB, nnz-1), b in 0..nblk.
__global__ __launch_bounds__(256) void ttv_prep_kernel(
    const int32_t* __restrict__ pos, float* __restrict__ y, int32_t nf,
    int32_t nnz, int32_t* __restrict__ brk, int32_t nblk) {
  const int64_t tid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t nthr = int64_t(gridDim.x) * blockDim.x;
  for (int64_t f = tid; f < nf; f += nthr) {
    y[f] = 0.f;
    const int32_t s = pos[f];
    const int32_t e = pos[f + 1];
    if (s >= e) continue;
    for (int64_t b = (int64_t(s) + kTtvNB - 1) / kTtvNB; b * kTtvNB < e; ++b)
      brk[b] = int32_t(f);
    if (s <= nnz - 1 && nnz - 1 < e) brk[nblk] = int32_t(f);
  }
}

__global__ __launch_bounds__(kTtvBlock) void ttv_kernel(
    const int32_t* __restrict__ pos, const int32_t* __restrict__ crd,
    const float* __restrict__ vals, const float* __restrict__ cv,
    float* __restrict__ y, int32_t nnz, const int32_t* __restrict__ brk) {
  constexpr int U = kTtvPerLane;
  // fiber bounds pos[lo..hi+1] of the workgroup, staged once; a CSF built
  // from coordinates has at most NB+1 non-empty fibers starting in a
  // workgroup, larger brackets (runs of empty fibers) read pos directly
  __shared__ int32_t s_pos[kTtvNB + 2];
  const int32_t b = blockIdx.x;
  const int32_t P = b * kTtvNB;
  const int32_t Pend = min(P + kTtvNB, nnz);
  const int lane = threadIdx.x & 63;
  const int32_t p0 = P + int32_t(threadIdx.x) * U;
  const int32_t p1 = min(p0 + U, Pend);
  const bool active = p0 < p1;
  // nonzeros and c gathers first: independent of the fiber bounds
  int32_t k[U];
  float v[U];
  float cg[U];
  if (active) {
    if (p0 + U <= nnz) {
#pragma unroll
      for (int h = 0; h < U; h += 4) {
        const int4 kk = *reinterpret_cast<const int4*>(crd + p0 + h);
        const float4 vv = *reinterpret_cast<const float4*>(vals + p0 + h);
        k[h] = kk.x; k[h + 1] = kk.y; k[h + 2] = kk.z; k[h + 3] = kk.w;
        v[h] = vv.x; v[h + 1] = vv.y; v[h + 2] = vv.z; v[h + 3] = vv.w;
      }
    } else {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool ok = p0 + u < nnz;
        k[u] = ok ? crd[p0 + u] : 0;
        v[u] = ok ? vals[p0 + u] : 0.f;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) cg[u] = cv[k[u]];
  }
  const int32_t lo = brk[b];
  const int32_t hi = brk[b + 1];
  const int32_t nstage = hi - lo + 2;
  const bool staged = nstage <= kTtvNB + 2;
  if (staged)
    for (int32_t j = threadIdx.x; j < nstage; j += kTtvBlock) s_pos[j] = pos[lo + j];
  __syncthreads();
  auto P_at = [&](int32_t x) -> int32_t { return staged ? s_pos[x - lo] : pos[x]; };
  auto search = [&](int32_t a, int32_t z, int32_t key) {
    while (a < z) {
      const int32_t mid = a + ((z - a + 1) >> 1);
      if (P_at(mid) <= key) a = mid;
      else z = mid - 1;
    }
    return a;
  };

  int32_t f = INT_MAX, nxt = 0, first = INT_MAX;
  float sum = 0.f, vf = 0.f;
  bool closed_first = false, ends_here = false, starts_before = false;
  if (active) {
    f = search(lo, hi, p0);
    nxt = P_at(f + 1);
    first = f;
    starts_before = P_at(f) < p0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int32_t p = p0 + u;
      if (p >= p1) break;
      if (p >= nxt) {
        if (!closed_first) {
          vf = sum;
          closed_first = true;
        } else {
          y[f] = sum;  // started and ended inside this lane
        }
        sum = 0.f;
        int32_t x = f + 1;
        int32_t e = P_at(x + 1);
        if (e <= p) {
          x = search(x + 1, hi, p);
          e = P_at(x + 1);
        }
        f = x;
        nxt = e;
      }
      sum = fmaf(v[u], cg[u], sum);
    }
    ends_here = (nxt == p1);
  }
  // inclusive segmented scan of the lanes' open (last) segments
  const int32_t tkey = f;
  float incl = active ? sum : 0.f;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const float ov = __shfl_up(incl, d);
    const int32_t ok = __shfl_up(tkey, d);
    if (lane >= d && ok == tkey) incl += ov;
  }
  const float pv = __shfl_up(incl, 1);
  const int32_t pk = __shfl_up(tkey, 1);
  const int32_t wfirst = __shfl(first, 0);
  const bool wbefore = __shfl(int(starts_before), 0) != 0;
  if (!active) return;
  const float carry = (lane > 0 && pk == first) ? pv : 0.f;
  const bool is_last = (lane == 63) || (p0 + U >= Pend);
  auto emit = [&](int32_t key, float val, bool shared) {
    if (shared || (key == wfirst && wbefore)) unsafeAtomicAdd(y + key, val);
    else y[key] = val;
  };
  if (closed_first) emit(first, vf + carry, false);
  if (ends_here) emit(f, incl, false);
  else if (is_last) emit(f, incl, true);
}

}  // namespace

size_t ttv_workspace_bytes(const CsfDev& B) {
  const int64_t nblk = (int64_t(B.nnz) + kTtvNB - 1) / kTtvNB;
  return size_t(nblk + 1) * sizeof(int32_t);
}

int launch_ttv(const CsfDev& B, const float* c, float* y, void* ws,
               hipStream_t st) {
  if (B.order < 2 || B.order > kMaxOrder)
    return fail(TENTACO_EUNSUPPORTED, "TTV order %d outside 2..5", B.order);
  const int32_t nf = B.nodes[B.order - 2];
  const int32_t* pos = B.pos[B.order - 1];
  const int32_t* crd = B.crd[B.order - 1];
  const int64_t nblk = (int64_t(B.nnz) + kTtvNB - 1) / kTtvNB;
  int32_t* brk = static_cast<int32_t*>(ws);
  int64_t g = (int64_t(nf) + 255) / 256;
  if (g < 1) g = 1;
  if (g > 16384) g = 16384;
  ttv_prep_kernel<<<int(g), 256, 0, st>>>(pos, y, nf, B.nnz, brk, int32_t(nblk));
  TT_HIP_CHECK(hipGetLastError());
  if (nblk == 0) return TENTACO_OK;
  ktimer_begin(st);
  ttv_kernel<<<int(nblk), kTtvBlock, 0, st>>>(pos, crd, B.vals, c, y, B.nnz, brk);
  TT_HIP_CHECK(hipGetLastError());
  ktimer_end(st, "ttv_kernel");
  return TENTACO_OK;
}

}  // namespace tt
