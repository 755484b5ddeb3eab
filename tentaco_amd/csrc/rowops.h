// Register-resident factor-row fragments for wave64 MTTKRP kernels.
//
// A factor row of rank R is spread over LPR consecutive lanes of a wavefront
// ("row group"); each lane holds NV float4 fragments covering columns
// (j*LPR + c)*4 .. +3 for lane c of the group, so one global_load_dwordx4 per
// fragment reads LPR*16 contiguous bytes of the row: R=16 -> 4 lanes, R=32 ->
// 8 lanes, R=64 -> 16 lanes, R=128 -> 16 lanes x 2 fragments. The generic rank
// path (R template 0, any runtime R <= 256) uses 16 lanes with one scalar
// column per slot, columns c + 16*j.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tt {

template <int R>
struct Lay {
  static_assert(R % 4 == 0 && R >= 4 && R <= 128, "specialised ranks");
  static constexpr int LPR = (R / 4 >= 16) ? 16 : R / 4;
  static constexpr int NV = R / (4 * LPR);
  static constexpr int kMaxRank = R;
  using V = float4;
};

template <>
struct Lay<0> {
  static constexpr int LPR = 16;
  static constexpr int NV = 16;
  static constexpr int kMaxRank = 256;
  using V = float;
};

template <int R>
struct Row {
  using L = Lay<R>;
  using V = typename L::V;
  static constexpr int NV = L::NV;
  static constexpr int LPR = L::LPR;

  __device__ __forceinline__ static int col(int j, int c) {
    if constexpr (R == 0) return c + LPR * j;
    else return (j * LPR + c) * 4;
  }

  __device__ __forceinline__ static void zero(V (&a)[NV]) {
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      if constexpr (R == 0) a[j] = 0.f;
      else a[j] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }

  // d = M[row, this lane's columns]; Rr is the runtime rank (== R when R>0).
  __device__ __forceinline__ static void load(V (&d)[NV],
                                              const float* __restrict__ M,
                                              int64_t row, int c, int Rr) {
    if constexpr (R == 0) {
      const float* p = M + row * Rr;
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        const int k = col(j, c);
        d[j] = (k < Rr) ? p[k] : 0.f;
      }
    } else {
      const float* p = M + row * R;
#pragma unroll
      for (int j = 0; j < NV; ++j)
        d[j] = *reinterpret_cast<const float4*>(p + col(j, c));
    }
  }

  // a += s * d
  __device__ __forceinline__ static void axpy(V (&a)[NV], float s,
                                              const V (&d)[NV]) {
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      if constexpr (R == 0) {
        a[j] = fmaf(s, d[j], a[j]);
      } else {
        a[j].x = fmaf(s, d[j].x, a[j].x);
        a[j].y = fmaf(s, d[j].y, a[j].y);
        a[j].z = fmaf(s, d[j].z, a[j].z);
        a[j].w = fmaf(s, d[j].w, a[j].w);
      }
    }
  }

  // a += x (.) y   (Hadamard product accumulate)
  __device__ __forceinline__ static void hfma(V (&a)[NV], const V (&x)[NV],
                                              const V (&y)[NV]) {
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      if constexpr (R == 0) {
        a[j] = fmaf(x[j], y[j], a[j]);
      } else {
        a[j].x = fmaf(x[j].x, y[j].x, a[j].x);
        a[j].y = fmaf(x[j].y, y[j].y, a[j].y);
        a[j].z = fmaf(x[j].z, y[j].z, a[j].z);
        a[j].w = fmaf(x[j].w, y[j].w, a[j].w);
      }
    }
  }

  // a = a (.) y
  __device__ __forceinline__ static void hmul(V (&a)[NV], const V (&y)[NV]) {
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      if constexpr (R == 0) {
        a[j] = a[j] * y[j];
      } else {
        a[j].x *= y[j].x;
        a[j].y *= y[j].y;
        a[j].z *= y[j].z;
        a[j].w *= y[j].w;
      }
    }
  }

  __device__ __forceinline__ static void add(V (&a)[NV], const V (&b)[NV]) {
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      if constexpr (R == 0) {
        a[j] += b[j];
      } else {
        a[j].x += b[j].x;
        a[j].y += b[j].y;
        a[j].z += b[j].z;
        a[j].w += b[j].w;
      }
    }
  }

  __device__ __forceinline__ static void store(float* __restrict__ A,
                                               int64_t row, int c, int Rr,
                                               const V (&a)[NV]) {
    if constexpr (R == 0) {
      float* p = A + row * Rr;
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        const int k = col(j, c);
        if (k < Rr) p[k] = a[j];
      }
    } else {
      float* p = A + row * R;
#pragma unroll
      for (int j = 0; j < NV; ++j)
        *reinterpret_cast<float4*>(p + col(j, c)) = a[j];
    }
  }

  // Device-scope float atomics (global_atomic_add_f32) on coarse-grained
  // memory: correct across XCDs, no CAS loop.
  __device__ __forceinline__ static void atomic_add(float* __restrict__ A,
                                                    int64_t row, int c, int Rr,
                                                    const V (&a)[NV]) {
    if constexpr (R == 0) {
      float* p = A + row * Rr;
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        const int k = col(j, c);
        if (k < Rr) unsafeAtomicAdd(p + k, a[j]);
      }
    } else {
      float* p = A + row * R;
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        float* q = p + col(j, c);
        unsafeAtomicAdd(q + 0, a[j].x);
        unsafeAtomicAdd(q + 1, a[j].y);
        unsafeAtomicAdd(q + 2, a[j].z);
        unsafeAtomicAdd(q + 3, a[j].w);
      }
    }
  }

  // FP64 row output (the edge-serial family's accumulation buffer): store,
  // and device-scope FP64 atomic add (global_atomic_add_f64).
  __device__ __forceinline__ static void store_f64(double* __restrict__ W, int64_t row, int c,
                                                   int Rr, const V (&a)[NV]) {
    if constexpr (R == 0) {
      double* p = W + row * Rr;
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        const int k = col(j, c);
        if (k < Rr) p[k] = double(a[j]);
      }
    } else {
      double* p = W + row * R;
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        double2* q = reinterpret_cast<double2*>(p + col(j, c));
        q[0] = make_double2(double(a[j].x), double(a[j].y));
        q[1] = make_double2(double(a[j].z), double(a[j].w));
      }
    }
  }

  __device__ __forceinline__ static void atomic_add_f64(double* __restrict__ W, int64_t row, int c,
                                                        int Rr, const V (&a)[NV]) {
    if constexpr (R == 0) {
      double* p = W + row * Rr;
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        const int k = col(j, c);
        if (k < Rr) unsafeAtomicAdd(p + k, double(a[j]));
      }
    } else {
      double* p = W + row * R;
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        double* q = p + col(j, c);
        unsafeAtomicAdd(q + 0, double(a[j].x));
        unsafeAtomicAdd(q + 1, double(a[j].y));
        unsafeAtomicAdd(q + 2, double(a[j].z));
        unsafeAtomicAdd(q + 3, double(a[j].w));
      }
    }
  }

  // Sum a over the row groups of a wavefront (xor butterfly over the lane
  // bits above log2(LPR)); every lane ends with the wave total of its slots.
  __device__ __forceinline__ static void wave_sum(V (&a)[NV]) {
#pragma unroll
    for (int off = LPR; off < 64; off <<= 1) {
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        if constexpr (R == 0) {
          a[j] += __shfl_xor(a[j], off);
        } else {
          a[j].x += __shfl_xor(a[j].x, off);
          a[j].y += __shfl_xor(a[j].y, off);
          a[j].z += __shfl_xor(a[j].z, off);
          a[j].w += __shfl_xor(a[j].w, off);
        }
      }
    }
  }
};

// Largest x in [lo, hi] with a[x] <= key; requires a[lo] <= key. With runs of
// equal values (empty nodes) this is the last node that starts at or before
// key, i.e. the node that contains key (upper-bound semantics; the
// reference's taco_binarySearchBefore may return any of the equal slots).
__device__ __forceinline__ int32_t search_le(const int32_t* __restrict__ a,
                                             int32_t lo, int32_t hi,
                                             int32_t key) {
  while (lo < hi) {
    const int32_t mid = lo + ((hi - lo + 1) >> 1);
    if (a[mid] <= key) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

}  // namespace tt
