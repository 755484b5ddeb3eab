// Owned host-side CSF (builder / generator output) with a taco_tensor_t view.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/taco_tensor_t.h"

namespace tt {

struct HostCsf {
  int order = 0;
  std::vector<int32_t> dims;
  std::vector<std::vector<int32_t>> pos, crd;  // level 0 = {0, dims0} / 0..dims0-1
  std::vector<float> vals;

  taco_tensor_t view{};
  std::vector<int32_t> dims_i32;
  std::vector<int32_t*> idx_ptrs;
  std::vector<int32_t**> lvl_ptrs;
  void make_view();
};

// Lexicographic COO -> CSF (dense root); coords mode-major, duplicates summed
// in input order. Inputs must already be validated (coords inside dims).
int build_csf_from_coo(int order, const int32_t* dims, int64_t n, const int32_t* coords,
                       const float* vals, HostCsf* out);

}  // namespace tt

struct tentaco_host_csf {
  tt::HostCsf h;
};
