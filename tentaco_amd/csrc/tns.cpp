// FROSTT .tns reader: text, one nonzero per line, "i_1 i_2 ... i_N value"
// with 1-based indices; '#' comment lines and blank lines are skipped. It
// replaces the reference's splatt load path (read_tensor / splatt_csf_load,
// ds.h:351-366) for the CSF MTTKRP/TTV path.
//
// The file is read whole and parsed in parallel over line-aligned chunks;
// the per-chunk results are concatenated in file order, so duplicate
// coordinates are summed in input order by the COO -> CSF builder. Values
// are parsed as double and rounded to FP32 (splatt reads doubles;
// to_taco_tensor hands FP32 to the kernels). Modes are permuted into CSF
// levels by mode_order, by default the splatt-like "shortest mode first"
// (ascending dimension, ties by mode number).

#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "host_csf.h"
#include "internal.h"

namespace tt {
namespace {

struct Chunk {
  std::vector<int32_t> c[kMaxOrder];
  std::vector<float> v;
  int32_t max_idx[kMaxOrder] = {};
  size_t err_off = SIZE_MAX;  // byte offset of the first bad line
  const char* err = nullptr;
};

inline bool is_space(char ch) { return ch == ' ' || ch == '\t' || ch == '\r' || ch == '\f' || ch == '\v'; }

// Parse lines in [b, e) (b at a line start, e at a line start or the end).
void parse_chunk(const char* base, size_t b, size_t e, int order, Chunk* out) {
  size_t i = b;
  auto bad = [&](size_t line_start, const char* why) {
    if (out->err_off == SIZE_MAX) {
      out->err_off = line_start;
      out->err = why;
    }
  };
  while (i < e) {
    const size_t line = i;
    while (i < e && is_space(base[i])) ++i;
    if (i >= e) break;
    if (base[i] == '\n') {
      ++i;
      continue;
    }
    if (base[i] == '#') {
      while (i < e && base[i] != '\n') ++i;
      continue;
    }
    int32_t idx[kMaxOrder];
    bool ok = true;
    for (int m = 0; m < order && ok; ++m) {
      while (i < e && is_space(base[i])) ++i;
      if (i >= e || base[i] < '0' || base[i] > '9') {
        bad(line, "expected a positive integer index");
        ok = false;
        break;
      }
      int64_t x = 0;
      while (i < e && base[i] >= '0' && base[i] <= '9') {
        x = x * 10 + (base[i] - '0');
        if (x > INT32_MAX) break;
        ++i;
      }
      if (x < 1 || x > INT32_MAX) {
        bad(line, "index outside 1..2^31-1");
        ok = false;
        break;
      }
      idx[m] = int32_t(x - 1);
    }
    double val = 0.0;
    if (ok) {
      while (i < e && is_space(base[i])) ++i;
      char* endp = nullptr;
      // strtod would skip a newline: a line that ends after its indices
      // must not take its value from the next line
      if (i < e && base[i] != '\n') val = std::strtod(base + i, &endp);
      if (!endp || endp == base + i || size_t(endp - base) > e) {
        bad(line, "expected a value after the indices");
        ok = false;
      } else {
        i = size_t(endp - base);
        while (i < e && is_space(base[i])) ++i;
        if (i < e && base[i] != '\n') {
          bad(line, "wrong number of columns");
          ok = false;
        }
      }
    }
    if (!ok) {
      while (i < e && base[i] != '\n') ++i;
      continue;
    }
    for (int m = 0; m < order; ++m) {
      out->c[m].push_back(idx[m]);
      out->max_idx[m] = std::max(out->max_idx[m], idx[m]);
    }
    out->v.push_back(float(val));
    if (i < e) ++i;  // the newline
  }
}

int count_columns(const char* s, size_t n, size_t* first_line) {
  size_t i = 0;
  while (i < n) {
    const size_t line = i;
    while (i < n && is_space(s[i])) ++i;
    if (i < n && s[i] == '\n') {
      ++i;
      continue;
    }
    if (i < n && s[i] == '#') {
      while (i < n && s[i] != '\n') ++i;
      continue;
    }
    if (i >= n) break;
    int cols = 0;
    while (i < n && s[i] != '\n') {
      while (i < n && is_space(s[i])) ++i;
      if (i >= n || s[i] == '\n') break;
      ++cols;
      while (i < n && !is_space(s[i]) && s[i] != '\n') ++i;
    }
    *first_line = line;
    return cols;
  }
  return 0;
}

int64_t line_number(const char* s, size_t off) {
  return 1 + std::count(s, s + off, '\n');
}

}  // namespace
}  // namespace tt

using namespace tt;

extern "C" {

int tentaco_tns_read(const char* path, const int32_t* dims_in, const int* mode_order,
                     tentaco_host_csf** out, int* order_out, int* mode_order_out) {
  clear_error();
  if (!path || !out) return fail(TENTACO_EINVAL, "tns: NULL argument");
  *out = nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(TENTACO_EINVAL, "tns: cannot open %s: %s", path, std::strerror(errno));
  std::vector<char> buf;
  {
    std::fseek(f, 0, SEEK_END);
    const long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (sz < 0) {
      std::fclose(f);
      return fail(TENTACO_EINVAL, "tns: cannot size %s", path);
    }
    buf.resize(size_t(sz) + 1);
    const size_t got = sz > 0 ? std::fread(buf.data(), 1, size_t(sz), f) : 0;
    std::fclose(f);
    if (got != size_t(sz)) return fail(TENTACO_EINVAL, "tns: short read of %s", path);
    buf[size_t(sz)] = '\0';
  }
  const size_t n = buf.size() - 1;
  const char* s = buf.data();
  size_t first = 0;
  const int cols = count_columns(s, n, &first);
  if (cols == 0) return fail(TENTACO_EINVAL, "tns: %s holds no nonzeros", path);
  const int order = cols - 1;
  if (order < 2 || order > kMaxOrder)
    return fail(TENTACO_EUNSUPPORTED, "tns: %s has order %d (2..5 supported)", path, order);
  // line-aligned chunks, parsed in parallel
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const size_t want = std::max<size_t>(1, std::min<size_t>(hw, n / (1 << 20) + 1));
  std::vector<size_t> cut(want + 1, n);
  cut[0] = 0;
  for (size_t k = 1; k < want; ++k) {
    size_t p = std::max(cut[k - 1], n * k / want);
    while (p < n && s[p - 1] != '\n') ++p;
    cut[k] = p;
  }
  std::vector<Chunk> parts(want);
  {
    std::vector<std::thread> th;
    for (size_t k = 0; k < want; ++k)
      th.emplace_back(parse_chunk, s, cut[k], cut[k + 1], order, &parts[k]);
    for (auto& t : th) t.join();
  }
  for (size_t k = 0; k < want; ++k)
    if (parts[k].err_off != SIZE_MAX)
      return fail(TENTACO_EINVAL, "tns: %s line %lld: %s", path,
                  (long long)line_number(s, parts[k].err_off), parts[k].err);
  int64_t nnz = 0;
  for (auto& p : parts) nnz += int64_t(p.v.size());
  if (nnz >= (int64_t(1) << 31) - (int64_t(1) << 16))
    return fail(TENTACO_EUNSUPPORTED, "tns: %lld nonzeros beyond the int32 index limit", (long long)nnz);
  int32_t dims[kMaxOrder] = {};
  for (int m = 0; m < order; ++m) {
    int32_t mx = 0;
    for (auto& p : parts) mx = std::max(mx, p.max_idx[m]);
    dims[m] = mx + 1;
    if (dims_in) {
      if (dims_in[m] < dims[m])
        return fail(TENTACO_EINVAL, "tns: mode %d index %d exceeds the given dimension %d", m,
                    dims[m], dims_in[m]);
      dims[m] = dims_in[m];
    }
  }
  int perm[kMaxOrder];
  if (mode_order) {
    bool seen[kMaxOrder] = {};
    for (int l = 0; l < order; ++l) {
      const int m = mode_order[l];
      if (m < 0 || m >= order || seen[m]) return fail(TENTACO_EINVAL, "tns: mode_order is not a permutation");
      seen[m] = true;
      perm[l] = m;
    }
  } else {
    std::iota(perm, perm + order, 0);
    std::stable_sort(perm, perm + order, [&](int a, int b) { return dims[a] < dims[b]; });
  }
  std::vector<int32_t> coords(size_t(order) * size_t(nnz));
  std::vector<float> vals(static_cast<size_t>(nnz));
  int32_t pdims[kMaxOrder];
  for (int l = 0; l < order; ++l) pdims[l] = dims[perm[l]];
  {
    size_t q = 0;
    for (auto& p : parts) {
      for (int l = 0; l < order; ++l)
        std::copy(p.c[perm[l]].begin(), p.c[perm[l]].end(), coords.begin() + size_t(l) * nnz + q);
      std::copy(p.v.begin(), p.v.end(), vals.begin() + q);
      q += p.v.size();
    }
  }
  parts.clear();
  buf.clear();
  buf.shrink_to_fit();
  auto* h = new tentaco_host_csf;
  const int rc = build_csf_from_coo(order, pdims, nnz, coords.data(), vals.data(), &h->h);
  if (rc) {
    delete h;
    return rc;
  }
  if (order_out) *order_out = order;
  if (mode_order_out)
    for (int l = 0; l < order; ++l) mode_order_out[l] = perm[l];
  *out = h;
  return TENTACO_OK;
}

}  // extern "C"
