// Variant choice for TENTACO_VARIANT_AUTO: a per-handle table filled by the
// autotuner, and a structural heuristic used when no table entry exists.
#pragma once

#include <cstdint>

#include "internal.h"

namespace tt {

// Best variant per rank found by timing every variant on this tensor.
struct TuneTable {
  static constexpr int kSlots = 8;
  int rank[kSlots] = {0, 0, 0, 0, 0, 0, 0, 0};
  int variant[kSlots] = {-1, -1, -1, -1, -1, -1, -1, -1};
  int lookup(int R) const {
    for (int i = 0; i < kSlots; ++i)
      if (rank[i] == R && variant[i] >= 0) return variant[i];
    return -1;
  }
  void set(int R, int v) {
    for (int i = 0; i < kSlots; ++i)
      if (rank[i] == R || variant[i] < 0) {
        rank[i] = R;
        variant[i] = v;
        return;
      }
    rank[kSlots - 1] = R;
    variant[kSlots - 1] = v;
  }
};

// Host-side structure statistics gathered while validating a CSF.
struct CsfStats {
  int64_t nnz = 0;
  int32_t n0 = 0;
  int32_t nonempty_slices = 0;
  int64_t max_slice_nnz = 0;
};

// Heuristic choice when the handle has not been autotuned: slice-row when the
// root slices are many and even, edge-serial otherwise.
int auto_variant(const CsfStats& s, int R);

}  // namespace tt
