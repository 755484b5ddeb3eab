// ORACLE - test infrastructure only. CPU restatement of the reference's TACO
// CSF kernels (zhang677/tenTACO tensor_kernels/*_cpu_taco*.h). Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
// library, and only as the checker or the timed CPU baseline; the product path
// (tentaco_amd) never links or calls it.
//
// Every loop nest below restates one reference function (cited per function)
// with the same loop order, the same OpenMP schedule and the same association
// of the floating-point products, so that its rounding matches the reference.
// The reference reads the FP32/int32 taco_tensor_t fields as double/uint64
// (splatt typing, mttkrp3_csf_cpu_taco.h:9-23); here each kernel is a template
// over (index type I, value type V). oracle_*_f64 runs I=int64_t, V=double on
// widened copies of the inputs (the reference's typing); oracle_*_f32 runs
// I=int32_t, V=float on the caller's arrays directly (the GPU's typing).
//
// Root level: the reference iterates i over B1_dimension with a dense root
// (compress_top_level, ds.h:426-436). Here the root is walked as n0 = pos0[1]
// nodes whose output row is crd0[x]; for a dense root crd0[x] == x and the
// walk is identical.

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "../include/taco_tensor_t.h"

namespace {

template <class I, class V>
struct Csf {
  int order = 0;
  I n0 = 0;
  const I* crd0 = nullptr;
  const I* pos[5] = {};
  const I* crd[5] = {};
  const V* vals = nullptr;
};

template <class V>
struct Mat {
  const V* v = nullptr;
  int64_t cols = 0;
};

// Storage for the widened (reference-typed) copy of a CSF and its factors.
struct Wide {
  std::vector<std::vector<int64_t>> pos, crd;
  std::vector<int64_t> crd0;
  std::vector<double> vals;
  std::vector<std::vector<double>> fac;
};

int64_t level_nodes(const taco_tensor_t* B, int m) {
  // number of nodes at level m: n0 for the root, pos_m[n_{m-1}] below it
  int64_t n = B->indices[0][0][1];
  for (int k = 1; k <= m; ++k) n = B->indices[k][0][n];
  return n;
}

bool valid_csf(const taco_tensor_t* B, int order) {
  if (!B || !B->indices || !B->vals) return false;
  for (int m = 0; m < order; ++m)
    if (!B->indices[m] || !B->indices[m][0] || !B->indices[m][1]) return false;
  return true;
}

Csf<int32_t, float> view32(const taco_tensor_t* B, int order) {
  Csf<int32_t, float> t;
  t.order = order;
  t.n0 = B->indices[0][0][1];
  t.crd0 = B->indices[0][1];
  for (int m = 1; m < order; ++m) {
    t.pos[m] = B->indices[m][0];
    t.crd[m] = B->indices[m][1];
  }
  t.vals = B->vals;
  return t;
}

Csf<int64_t, double> widen(const taco_tensor_t* B, int order,
                           const taco_tensor_t* const* F, int R, Wide& w,
                           Mat<double>* mats) {
  Csf<int64_t, double> t;
  t.order = order;
  t.n0 = B->indices[0][0][1];
  w.crd0.assign(B->indices[0][1], B->indices[0][1] + t.n0);
  t.crd0 = w.crd0.data();
  w.pos.resize(order);
  w.crd.resize(order);
  for (int m = 1; m < order; ++m) {
    int64_t parents = level_nodes(B, m - 1);
    int64_t nodes = level_nodes(B, m);
    w.pos[m].assign(B->indices[m][0], B->indices[m][0] + parents + 1);
    w.crd[m].assign(B->indices[m][1], B->indices[m][1] + nodes);
    t.pos[m] = w.pos[m].data();
    t.crd[m] = w.crd[m].data();
  }
  int64_t nnz = level_nodes(B, order - 1);
  w.vals.assign(B->vals, B->vals + nnz);
  t.vals = w.vals.data();
  if (F) {
    w.fac.resize(order);
    for (int m = 1; m < order; ++m) {
      int64_t rows = F[m - 1]->dimensions[0];
      w.fac[m].assign(F[m - 1]->vals, F[m - 1]->vals + rows * R);
      mats[m].v = w.fac[m].data();
      mats[m].cols = R;
    }
  }
  return t;
}

template <class V>
void zero_rows(V* A, int64_t n) {
  // mttkrp3_csf_cpu_taco.h:25-28 (static zero of A)
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < n; ++p) A[p] = V(0);
}

// ---- MTTKRP, 3-mode -------------------------------------------------------
// "taco" schedule: mttkrp3_csf_cpu_taco.h:30-50. dynamic,1 over chunks of 64
// slices; per nonzero A[i,j] += (B * C[k,j]) * D[l,j].
template <class I, class V>
void mttkrp3_taco(V* A, int64_t R, const Csf<I, V>& B, const Mat<V>* M) {
  const I n0 = B.n0;
  const V* C = M[1].v;
  const V* D = M[2].v;
#pragma omp parallel for schedule(dynamic, 1)
  for (I c = 0; c < (n0 + 63) / 64; ++c) {
    for (I s = 0; s < 64; ++s) {
      I x = c * 64 + s;
      if (x >= n0) continue;
      V* Ai = A + (int64_t)B.crd0[x] * R;
      for (I kB = B.pos[1][x]; kB < B.pos[1][x + 1]; ++kB) {
        const V* Ck = C + (int64_t)B.crd[1][kB] * R;
        for (I lB = B.pos[2][kB]; lB < B.pos[2][kB + 1]; ++lB) {
          const V* Dl = D + (int64_t)B.crd[2][lB] * R;
          const V b = B.vals[lB];
          for (int64_t j = 0; j < R; ++j) Ai[j] = Ai[j] + (b * Ck[j]) * Dl[j];
        }
      }
    }
  }
}

// "unscheduled": mttkrp3_csf_cpu_taco_unscheduled.h:30-44. dynamic,16 over
// slices; A[i,j] += B * C[k,j] * D[l,j] (left-to-right: (B*C)*D).
template <class I, class V>
void mttkrp3_unscheduled(V* A, int64_t R, const Csf<I, V>& B,
                         const Mat<V>* M) {
  const V* C = M[1].v;
  const V* D = M[2].v;
#pragma omp parallel for schedule(dynamic, 16)
  for (I x = 0; x < B.n0; ++x) {
    V* Ai = A + (int64_t)B.crd0[x] * R;
    for (I p2 = B.pos[1][x]; p2 < B.pos[1][x + 1]; ++p2) {
      const V* Ck = C + (int64_t)B.crd[1][p2] * R;
      for (I p3 = B.pos[2][p2]; p3 < B.pos[2][p2 + 1]; ++p3) {
        const V* Dl = D + (int64_t)B.crd[2][p3] * R;
        const V b = B.vals[p3];
        for (int64_t j = 0; j < R; ++j) Ai[j] = Ai[j] + b * Ck[j] * Dl[j];
      }
    }
  }
}

// "workspace": mttkrp3_csf_cpu_taco_workspace.h:30-60. Per fiber
// w[j] += B * D[l,j]; at the fiber end A[i,j] += w[j] * C[k,j], w = 0.
template <class I, class V>
void mttkrp3_workspace(V* A, int64_t R, const Csf<I, V>& B, const Mat<V>* M) {
  const I n0 = B.n0;
  const V* C = M[1].v;
  const V* D = M[2].v;
#pragma omp parallel for schedule(dynamic, 1)
  for (I c = 0; c < (n0 + 63) / 64; ++c) {
    std::vector<V> w(R, V(0));
    for (I s = 0; s < 64; ++s) {
      I x = c * 64 + s;
      if (x >= n0) continue;
      V* Ai = A + (int64_t)B.crd0[x] * R;
      for (I kB = B.pos[1][x]; kB < B.pos[1][x + 1]; ++kB) {
        for (I lB = B.pos[2][kB]; lB < B.pos[2][kB + 1]; ++lB) {
          const V* Dl = D + (int64_t)B.crd[2][lB] * R;
          const V b = B.vals[lB];
          for (int64_t j = 0; j < R; ++j) w[j] = w[j] + b * Dl[j];
        }
        const V* Ck = C + (int64_t)B.crd[1][kB] * R;
        for (int64_t j = 0; j < R; ++j) {
          Ai[j] = Ai[j] + w[j] * Ck[j];
          w[j] = V(0);
        }
      }
    }
  }
}

// ---- MTTKRP, 4-mode -------------------------------------------------------
// "taco": mttkrp4_csf_cpu_taco.h:35-58. A += ((B*C)*D)*E per nonzero.
template <class I, class V>
void mttkrp4_taco(V* A, int64_t R, const Csf<I, V>& B, const Mat<V>* M) {
  const I n0 = B.n0;
#pragma omp parallel for schedule(dynamic, 1)
  for (I c = 0; c < (n0 + 63) / 64; ++c) {
    for (I s = 0; s < 64; ++s) {
      I x = c * 64 + s;
      if (x >= n0) continue;
      V* Ai = A + (int64_t)B.crd0[x] * R;
      for (I kB = B.pos[1][x]; kB < B.pos[1][x + 1]; ++kB) {
        const V* Ck = M[1].v + (int64_t)B.crd[1][kB] * R;
        for (I lB = B.pos[2][kB]; lB < B.pos[2][kB + 1]; ++lB) {
          const V* Dl = M[2].v + (int64_t)B.crd[2][lB] * R;
          for (I mB = B.pos[3][lB]; mB < B.pos[3][lB + 1]; ++mB) {
            const V* Em = M[3].v + (int64_t)B.crd[3][mB] * R;
            const V b = B.vals[mB];
            for (int64_t j = 0; j < R; ++j)
              Ai[j] = Ai[j] + ((b * Ck[j]) * Dl[j]) * Em[j];
          }
        }
      }
    }
  }
}

// "unscheduled": mttkrp4_csf_cpu_taco_unscheduled.h (dynamic,16 per slice,
// B*C*D*E left to right).
template <class I, class V>
void mttkrp4_unscheduled(V* A, int64_t R, const Csf<I, V>& B,
                         const Mat<V>* M) {
#pragma omp parallel for schedule(dynamic, 16)
  for (I x = 0; x < B.n0; ++x) {
    V* Ai = A + (int64_t)B.crd0[x] * R;
    for (I p2 = B.pos[1][x]; p2 < B.pos[1][x + 1]; ++p2) {
      const V* Ck = M[1].v + (int64_t)B.crd[1][p2] * R;
      for (I p3 = B.pos[2][p2]; p3 < B.pos[2][p2 + 1]; ++p3) {
        const V* Dl = M[2].v + (int64_t)B.crd[2][p3] * R;
        for (I p4 = B.pos[3][p3]; p4 < B.pos[3][p3 + 1]; ++p4) {
          const V* Em = M[3].v + (int64_t)B.crd[3][p4] * R;
          const V b = B.vals[p4];
          for (int64_t j = 0; j < R; ++j)
            Ai[j] = Ai[j] + b * Ck[j] * Dl[j] * Em[j];
        }
      }
    }
  }
}

// "workspace": mttkrp4_csf_cpu_taco_workspace.h. BE[j] += B*E[m,j] per
// level-2 node; BDE[j] += BE[j]*D[l,j]; A += BDE[j]*C[k,j].
template <class I, class V>
void mttkrp4_workspace(V* A, int64_t R, const Csf<I, V>& B, const Mat<V>* M) {
  const I n0 = B.n0;
#pragma omp parallel for schedule(dynamic, 1)
  for (I c = 0; c < (n0 + 63) / 64; ++c) {
    std::vector<V> be(R, V(0)), bde(R, V(0));
    for (I s = 0; s < 64; ++s) {
      I x = c * 64 + s;
      if (x >= n0) continue;
      V* Ai = A + (int64_t)B.crd0[x] * R;
      for (I kB = B.pos[1][x]; kB < B.pos[1][x + 1]; ++kB) {
        for (I lB = B.pos[2][kB]; lB < B.pos[2][kB + 1]; ++lB) {
          for (I mB = B.pos[3][lB]; mB < B.pos[3][lB + 1]; ++mB) {
            const V* Em = M[3].v + (int64_t)B.crd[3][mB] * R;
            const V b = B.vals[mB];
            for (int64_t j = 0; j < R; ++j) be[j] = be[j] + b * Em[j];
          }
          const V* Dl = M[2].v + (int64_t)B.crd[2][lB] * R;
          for (int64_t j = 0; j < R; ++j) {
            bde[j] = bde[j] + be[j] * Dl[j];
            be[j] = V(0);
          }
        }
        const V* Ck = M[1].v + (int64_t)B.crd[1][kB] * R;
        for (int64_t j = 0; j < R; ++j) {
          Ai[j] = Ai[j] + bde[j] * Ck[j];
          bde[j] = V(0);
        }
      }
    }
  }
}

// ---- MTTKRP, 5-mode -------------------------------------------------------
// "taco": mttkrp5_csf_cpu_taco.h:35-95. Already factorised in the reference:
// BF += B*F; BEF += BF*E; BDEF += BEF*D; A += BDEF*C, with each workspace
// allocated (malloc) and zeroed at the start of its node and freed at its
// end (mttkrp5_csf_cpu_taco.h:49-90). The allocations are kept so that the
// timed CPU baseline carries the same cost the reference pays.
template <class I, class V>
void mttkrp5_taco(V* A, int64_t R, const Csf<I, V>& B, const Mat<V>* M) {
  const I n0 = B.n0;
#pragma omp parallel for schedule(dynamic, 1)
  for (I c = 0; c < (n0 + 63) / 64; ++c) {
    for (I s = 0; s < 64; ++s) {
      I x = c * 64 + s;
      if (x >= n0) continue;
      V* Ai = A + (int64_t)B.crd0[x] * R;
      for (I kB = B.pos[1][x]; kB < B.pos[1][x + 1]; ++kB) {
        V* bdef = static_cast<V*>(std::malloc(sizeof(V) * R));
        for (int64_t j = 0; j < R; ++j) bdef[j] = V(0);
        for (I lB = B.pos[2][kB]; lB < B.pos[2][kB + 1]; ++lB) {
          V* bef = static_cast<V*>(std::malloc(sizeof(V) * R));
          for (int64_t j = 0; j < R; ++j) bef[j] = V(0);
          for (I mB = B.pos[3][lB]; mB < B.pos[3][lB + 1]; ++mB) {
            V* bf = static_cast<V*>(std::malloc(sizeof(V) * R));
            for (int64_t j = 0; j < R; ++j) bf[j] = V(0);
            for (I nB = B.pos[4][mB]; nB < B.pos[4][mB + 1]; ++nB) {
              const V* Fn = M[4].v + (int64_t)B.crd[4][nB] * R;
              const V b = B.vals[nB];
              for (int64_t j = 0; j < R; ++j) bf[j] = bf[j] + b * Fn[j];
            }
            const V* Em = M[3].v + (int64_t)B.crd[3][mB] * R;
            for (int64_t j = 0; j < R; ++j) bef[j] = bef[j] + bf[j] * Em[j];
            std::free(bf);
          }
          const V* Dl = M[2].v + (int64_t)B.crd[2][lB] * R;
          for (int64_t j = 0; j < R; ++j) bdef[j] = bdef[j] + bef[j] * Dl[j];
          std::free(bef);
        }
        const V* Ck = M[1].v + (int64_t)B.crd[1][kB] * R;
        for (int64_t j = 0; j < R; ++j) Ai[j] = Ai[j] + bdef[j] * Ck[j];
        std::free(bdef);
      }
    }
  }
}

// "unscheduled": mttkrp5_csf_cpu_taco_unscheduled.h (B*C*D*E*F per nonzero,
// dynamic,16 per slice).
template <class I, class V>
void mttkrp5_unscheduled(V* A, int64_t R, const Csf<I, V>& B,
                         const Mat<V>* M) {
#pragma omp parallel for schedule(dynamic, 16)
  for (I x = 0; x < B.n0; ++x) {
    V* Ai = A + (int64_t)B.crd0[x] * R;
    for (I p2 = B.pos[1][x]; p2 < B.pos[1][x + 1]; ++p2) {
      const V* Ck = M[1].v + (int64_t)B.crd[1][p2] * R;
      for (I p3 = B.pos[2][p2]; p3 < B.pos[2][p2 + 1]; ++p3) {
        const V* Dl = M[2].v + (int64_t)B.crd[2][p3] * R;
        for (I p4 = B.pos[3][p3]; p4 < B.pos[3][p3 + 1]; ++p4) {
          const V* Em = M[3].v + (int64_t)B.crd[3][p4] * R;
          for (I p5 = B.pos[4][p4]; p5 < B.pos[4][p4 + 1]; ++p5) {
            const V* Fn = M[4].v + (int64_t)B.crd[4][p5] * R;
            const V b = B.vals[p5];
            for (int64_t j = 0; j < R; ++j)
              Ai[j] = Ai[j] + b * Ck[j] * Dl[j] * Em[j] * Fn[j];
          }
        }
      }
    }
  }
}

// "workspace": mttkrp5_csf_cpu_taco_workspace.h:40-98. Same association as
// the taco variant with the three workspaces hoisted to the 64-slice chunk
// and re-zeroed after use.
template <class I, class V>
void mttkrp5_workspace(V* A, int64_t R, const Csf<I, V>& B, const Mat<V>* M) {
  const I n0 = B.n0;
#pragma omp parallel for schedule(dynamic, 1)
  for (I c = 0; c < (n0 + 63) / 64; ++c) {
    std::vector<V> bdef(R, V(0)), bef(R, V(0)), bf(R, V(0));
    for (I s = 0; s < 64; ++s) {
      I x = c * 64 + s;
      if (x >= n0) continue;
      V* Ai = A + (int64_t)B.crd0[x] * R;
      for (I kB = B.pos[1][x]; kB < B.pos[1][x + 1]; ++kB) {
        for (I lB = B.pos[2][kB]; lB < B.pos[2][kB + 1]; ++lB) {
          for (I mB = B.pos[3][lB]; mB < B.pos[3][lB + 1]; ++mB) {
            for (I nB = B.pos[4][mB]; nB < B.pos[4][mB + 1]; ++nB) {
              const V* Fn = M[4].v + (int64_t)B.crd[4][nB] * R;
              const V b = B.vals[nB];
              for (int64_t j = 0; j < R; ++j) bf[j] = bf[j] + b * Fn[j];
            }
            const V* Em = M[3].v + (int64_t)B.crd[3][mB] * R;
            for (int64_t j = 0; j < R; ++j) {
              bef[j] = bef[j] + bf[j] * Em[j];
              bf[j] = V(0);
            }
          }
          const V* Dl = M[2].v + (int64_t)B.crd[2][lB] * R;
          for (int64_t j = 0; j < R; ++j) {
            bdef[j] = bdef[j] + bef[j] * Dl[j];
            bef[j] = V(0);
          }
        }
        const V* Ck = M[1].v + (int64_t)B.crd[1][kB] * R;
        for (int64_t j = 0; j < R; ++j) {
          Ai[j] = Ai[j] + bdef[j] * Ck[j];
          bdef[j] = V(0);
        }
      }
    }
  }
}

// ---- order-2 MTTKRP (sparse-matrix times dense matrix) -------------------
// Not a reference kernel; the same loop shape one level down, used only so
// that the order-2 path of the library has a checker.
template <class I, class V>
void mttkrp2(V* A, int64_t R, const Csf<I, V>& B, const Mat<V>* M) {
#pragma omp parallel for schedule(dynamic, 16)
  for (I x = 0; x < B.n0; ++x) {
    V* Ai = A + (int64_t)B.crd0[x] * R;
    for (I p = B.pos[1][x]; p < B.pos[1][x + 1]; ++p) {
      const V* Ck = M[1].v + (int64_t)B.crd[1][p] * R;
      const V b = B.vals[p];
      for (int64_t j = 0; j < R; ++j) Ai[j] = Ai[j] + b * Ck[j];
    }
  }
}

template <class I, class V>
int run_mttkrp(int order, int schedule, V* A, int64_t a_rows, int64_t R,
               const Csf<I, V>& B, const Mat<V>* M) {
  zero_rows(A, a_rows * R);
  switch (order * 10 + schedule) {
    case 20: case 21: case 22: mttkrp2(A, R, B, M); return 0;
    case 30: mttkrp3_taco(A, R, B, M); return 0;
    case 31: mttkrp3_unscheduled(A, R, B, M); return 0;
    case 32: mttkrp3_workspace(A, R, B, M); return 0;
    case 40: mttkrp4_taco(A, R, B, M); return 0;
    case 41: mttkrp4_unscheduled(A, R, B, M); return 0;
    case 42: mttkrp4_workspace(A, R, B, M); return 0;
    case 50: mttkrp5_taco(A, R, B, M); return 0;
    case 51: mttkrp5_unscheduled(A, R, B, M); return 0;
    case 52: mttkrp5_workspace(A, R, B, M); return 0;
    default: return 1;
  }
}

// ---- TTV ------------------------------------------------------------------
// taco_binarySearchBefore (ttv_csf_cpu_taco.h:2-22): index in
// [arrayStart, arrayEnd] of a slot whose value is <= target; returns
// arrayEnd when array[arrayEnd] <= target, any exact match otherwise.
template <class I>
I search_before(const I* a, I lo, I hi, I target) {
  if (a[hi] <= target) return hi;
  while (hi - lo > 1) {
    I mid = (hi + lo) / 2;
    I v = a[mid];
    if (v < target) lo = mid;
    else if (v > target) hi = mid;
    else return mid;
  }
  return lo;
}

// ttv_csf_cpu_taco (ttv_csf_cpu_taco.h:24-68): dynamic,1 over chunks of 8
// positions f of level N-2; A[f] += B * c[k] over the fiber. The slice
// search and while-advance (:51-58) only locate i, which the output does not
// use; they are kept because they are part of the timed reference cost.
template <class I, class V>
void ttv_taco(V* A, const Csf<I, V>& B, const V* c, int order) {
  const I* pos = B.pos[order - 1];
  const I* crd = B.crd[order - 1];
  const I* ppos = B.pos[order - 2 >= 1 ? order - 2 : 1];
  // number of level N-2 positions
  I nf = B.n0;
  for (int m = 1; m <= order - 2; ++m) nf = B.pos[m][nf];
#pragma omp parallel for schedule(static)
  for (I f = 0; f < nf; ++f) A[f] = V(0);
  const bool has_parent = order == 3;
#pragma omp parallel for schedule(dynamic, 1)
  for (I chunk = 0; chunk < (nf + 7) / 8; ++chunk) {
    I ipos = 0;
    if (has_parent) ipos = search_before(ppos, I(0), B.n0, chunk * 8);
    for (I s = 0; s < 8; ++s) {
      I f = chunk * 8 + s;
      if (f >= nf) continue;
      if (has_parent) {
        while (f == ppos[ipos + 1]) ++ipos;
      }
      for (I kB = pos[f]; kB < pos[f + 1]; ++kB)
        A[f] = A[f] + B.vals[kB] * c[crd[kB]];
    }
    (void)ipos;
  }
}

// ttv_csf_cpu_taco_unscheduled (ttv_csf_cpu_taco_unscheduled.h:49-79):
// omp parallel for over slices, same accumulation per fiber.
template <class I, class V>
void ttv_unscheduled(V* A, const Csf<I, V>& B, const V* c, int order) {
  const I* pos = B.pos[order - 1];
  const I* crd = B.crd[order - 1];
  I parents = B.n0;
  for (int m = 1; m <= order - 2; ++m) parents = B.pos[m][parents];
#pragma omp parallel for schedule(static)
  for (I f = 0; f < parents; ++f) A[f] = V(0);
  if (order == 3) {
#pragma omp parallel for
    for (I x = 0; x < B.n0; ++x)
      for (I f = B.pos[1][x]; f < B.pos[1][x + 1]; ++f)
        for (I kB = pos[f]; kB < pos[f + 1]; ++kB)
          A[f] = A[f] + B.vals[kB] * c[crd[kB]];
  } else {
#pragma omp parallel for
    for (I f = 0; f < parents; ++f)
      for (I kB = pos[f]; kB < pos[f + 1]; ++kB)
        A[f] = A[f] + B.vals[kB] * c[crd[kB]];
  }
}

// Timer::clear_cache (timers.h:62-75): random gather over a buffer much
// larger than the last-level cache before each timed trial (TIME_COLD,
// timers.h:112-117). The reference uses 4 threads over 10M doubles; the
// buffer here is sized by the caller (>= 2x LLC).
struct Flusher {
  std::vector<double> buf;
  std::vector<int32_t> idx;
  explicit Flusher(size_t n) : buf(n), idx(n) {
    std::mt19937_64 rng(12345);
    for (size_t i = 0; i < n; ++i) buf[i] = double(rng() % 1000) / 1000.0;
    std::iota(idx.begin(), idx.end(), 0);
    std::shuffle(idx.begin(), idx.end(), rng);
  }
  double flush() {
    double total = 0.0;
    const size_t n = buf.size();
#pragma omp parallel reduction(+ : total)
    {
      double s = 0.0;
#pragma omp for schedule(static)
      for (size_t i = 0; i < n; ++i) s += buf[idx[i]];
      total += s;
    }
    return total;
  }
};

double now_ms() {
  using clk = std::chrono::steady_clock;
  return std::chrono::duration<double, std::milli>(
             clk::now().time_since_epoch())
      .count();
}

}  // namespace

extern "C" {

// MTTKRP in the reference's typing: 64-bit indices, FP64 values and
// accumulation. factors[m-1] is the FP32 factor of level m (widened here).
// schedule: 0 taco, 1 unscheduled, 2 workspace. A_out: a_rows x R doubles.
int oracle_mttkrp_f64(int order, int schedule, double* A_out, int64_t a_rows,
                      int R, const taco_tensor_t* B,
                      const taco_tensor_t* const* factors) {
  if (order < 2 || order > 5 || !valid_csf(B, order) || !A_out || R <= 0)
    return 1;
  Wide w;
  Mat<double> M[5];
  Csf<int64_t, double> t = widen(B, order, factors, R, w, M);
  return run_mttkrp<int64_t, double>(order, schedule, A_out, a_rows, R, t, M);
}

// MTTKRP in the GPU's typing: int32 indices, FP32 values and accumulation.
int oracle_mttkrp_f32(int order, int schedule, float* A_out, int64_t a_rows,
                      int R, const taco_tensor_t* B,
                      const taco_tensor_t* const* factors) {
  if (order < 2 || order > 5 || !valid_csf(B, order) || !A_out || R <= 0)
    return 1;
  Csf<int32_t, float> t = view32(B, order);
  Mat<float> M[5];
  for (int m = 1; m < order; ++m) {
    M[m].v = factors[m - 1]->vals;
    M[m].cols = R;
  }
  return run_mttkrp<int32_t, float>(order, schedule, A_out, a_rows, R, t, M);
}

// TTV, reference typing. schedule 0 = ttv_csf_cpu_taco, 1 = unscheduled.
// A_out has one double per level N-2 position.
int oracle_ttv_f64(int order, int schedule, double* A_out,
                   const taco_tensor_t* B, const taco_tensor_t* c) {
  if (order < 2 || order > 5 || !valid_csf(B, order) || !A_out || !c) return 1;
  Wide w;
  Csf<int64_t, double> t = widen(B, order, nullptr, 0, w, nullptr);
  int64_t n = c->dimensions[0];
  std::vector<double> cv(c->vals, c->vals + n);
  if (order == 2) {
    // order 2: the "fibers" are the root slices; reuse the unscheduled nest.
    ttv_unscheduled(A_out, t, cv.data(), order);
    return 0;
  }
  if (schedule == 0) ttv_taco(A_out, t, cv.data(), order);
  else ttv_unscheduled(A_out, t, cv.data(), order);
  return 0;
}

int oracle_ttv_f32(int order, int schedule, float* A_out,
                   const taco_tensor_t* B, const taco_tensor_t* c) {
  if (order < 2 || order > 5 || !valid_csf(B, order) || !A_out || !c) return 1;
  Csf<int32_t, float> t = view32(B, order);
  if (order == 2 || schedule != 0) ttv_unscheduled(A_out, t, c->vals, order);
  else ttv_taco(A_out, t, c->vals, order);
  return 0;
}

// Timed CPU baseline (RUN/TIME_COLD, timers.h:112-130): `trials` runs of the
// chosen MTTKRP nest, each after a cache flush of flush_doubles doubles; the
// zeroing of A is inside the timed region as in the reference. Writes the
// minimum and median milliseconds. fp64 selects the reference typing.
int oracle_time_mttkrp(int order, int schedule, int fp64, int64_t a_rows,
                       int R, const taco_tensor_t* B,
                       const taco_tensor_t* const* factors, int trials,
                       int64_t flush_doubles, double* min_ms,
                       double* med_ms) {
  if (order < 2 || order > 5 || !valid_csf(B, order) || trials <= 0) return 1;
  Flusher fl(flush_doubles > 0 ? size_t(flush_doubles) : size_t(1));
  std::vector<double> times;
  volatile double sink = 0.0;
  if (fp64) {
    Wide w;
    Mat<double> M[5];
    Csf<int64_t, double> t = widen(B, order, factors, R, w, M);
    std::vector<double> A(size_t(a_rows) * R);
    for (int k = 0; k < trials; ++k) {
      if (flush_doubles > 0) sink = sink + fl.flush();
      double t0 = now_ms();
      run_mttkrp<int64_t, double>(order, schedule, A.data(), a_rows, R, t, M);
      times.push_back(now_ms() - t0);
    }
  } else {
    Csf<int32_t, float> t = view32(B, order);
    Mat<float> M[5];
    for (int m = 1; m < order; ++m) {
      M[m].v = factors[m - 1]->vals;
      M[m].cols = R;
    }
    std::vector<float> A(size_t(a_rows) * R);
    for (int k = 0; k < trials; ++k) {
      if (flush_doubles > 0) sink = sink + fl.flush();
      double t0 = now_ms();
      run_mttkrp<int32_t, float>(order, schedule, A.data(), a_rows, R, t, M);
      times.push_back(now_ms() - t0);
    }
  }
  std::sort(times.begin(), times.end());
  *min_ms = times.front();
  *med_ms = times[times.size() / 2];
  return 0;
}

int oracle_time_ttv(int order, int fp64, const taco_tensor_t* B,
                    const taco_tensor_t* c, int trials, int64_t flush_doubles,
                    double* min_ms, double* med_ms) {
  if (order < 2 || order > 5 || !valid_csf(B, order) || trials <= 0) return 1;
  Flusher fl(flush_doubles > 0 ? size_t(flush_doubles) : size_t(1));
  std::vector<double> times;
  volatile double sink = 0.0;
  int64_t nf = level_nodes(B, order - 2);
  if (fp64) {
    Wide w;
    Csf<int64_t, double> t = widen(B, order, nullptr, 0, w, nullptr);
    std::vector<double> cv(c->vals, c->vals + c->dimensions[0]);
    std::vector<double> A(nf);
    for (int k = 0; k < trials; ++k) {
      if (flush_doubles > 0) sink = sink + fl.flush();
      double t0 = now_ms();
      ttv_taco(A.data(), t, cv.data(), order);
      times.push_back(now_ms() - t0);
    }
  } else {
    Csf<int32_t, float> t = view32(B, order);
    std::vector<float> A(nf);
    for (int k = 0; k < trials; ++k) {
      if (flush_doubles > 0) sink = sink + fl.flush();
      double t0 = now_ms();
      ttv_taco(A.data(), t, c->vals, order);
      times.push_back(now_ms() - t0);
    }
  }
  std::sort(times.begin(), times.end());
  *min_ms = times.front();
  *med_ms = times[times.size() / 2];
  return 0;
}

int oracle_num_threads(void) { return omp_get_max_threads(); }

// Thread count of the timed nests (the reference drivers run under
// OMP_NUM_THREADS, run_gpu.sh:9).
void oracle_set_num_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}

// taco_binarySearchBefore exposed for the partition tests (ambiguous-match
// semantics of gpu_library.h:185-205, kept to show why the product does not
// use them).
int32_t oracle_search_before(const int32_t* a, int32_t lo, int32_t hi,
                             int32_t target) {
  return search_before<int32_t>(a, lo, hi, target);
}

}  // extern "C"
