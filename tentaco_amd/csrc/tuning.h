// Variant choice for TENTACO_VARIANT_AUTO: a per-handle table filled by the
// autotuner, and a structural heuristic used when no table entry exists.
#pragma once

#include <cstdint>

#include "internal.h"

namespace tt {

// Best variant per rank found by timing every variant on this tensor.
// source: how the entry was chosen (TENTACO_TUNED_SWEEP, _SELECTOR, _SET).
struct TuneTable {
  static constexpr int kSlots = 8;
  int rank[kSlots] = {0, 0, 0, 0, 0, 0, 0, 0};
  int variant[kSlots] = {-1, -1, -1, -1, -1, -1, -1, -1};
  int source[kSlots] = {0, 0, 0, 0, 0, 0, 0, 0};
  int lookup(int R, int* src = nullptr) const {
    for (int i = 0; i < kSlots; ++i)
      if (rank[i] == R && variant[i] >= 0) {
        if (src) *src = source[i];
        return variant[i];
      }
    return -1;
  }
  void set(int R, int v, int src) {
    for (int i = 0; i < kSlots; ++i)
      if (rank[i] == R || variant[i] < 0) {
        rank[i] = R;
        variant[i] = v;
        source[i] = src;
        return;
      }
    rank[kSlots - 1] = R;
    variant[kSlots - 1] = v;
    source[kSlots - 1] = src;
  }
};

// Host-side structure statistics gathered while validating a CSF.
struct CsfStats {
  int64_t nnz = 0;
  int32_t n0 = 0;
  int32_t nonempty_slices = 0;
  int64_t max_slice_nnz = 0;
  // children per level m-1 node at level m (m = 1..order-1): mean, variance, max
  double child_mean[kMaxOrder] = {};
  double child_var[kMaxOrder] = {};
  double child_max[kMaxOrder] = {};
};

// Heuristic choice when the handle has not been autotuned: slice-row when the
// root slices are many and even, edge-parallel otherwise.
int auto_variant(const CsfStats& s, int order, int R);

}  // namespace tt
