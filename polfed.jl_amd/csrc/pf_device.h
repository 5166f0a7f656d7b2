// Device-side helpers shared by the kernels of libpolfed_hip (ops.hip,
// batch.hip): vector types, vector loads/stores, the XCD-contiguous block
// mapping and a block-wide reduction.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pf {

// ------------------------------------------------------------ vector types
template <typename T, int V> struct VT;
template <> struct VT<double, 1> { typedef double t; };
template <> struct VT<double, 2> { typedef double t __attribute__((ext_vector_type(2))); };
template <> struct VT<double, 4> { typedef double t __attribute__((ext_vector_type(4))); };
template <> struct VT<double, 8> { typedef double t __attribute__((ext_vector_type(8))); };
template <> struct VT<double, 16> { typedef double t __attribute__((ext_vector_type(16))); };
template <> struct VT<float, 1> { typedef float t; };
template <> struct VT<float, 2> { typedef float t __attribute__((ext_vector_type(2))); };
template <> struct VT<float, 4> { typedef float t __attribute__((ext_vector_type(4))); };
template <> struct VT<float, 8> { typedef float t __attribute__((ext_vector_type(8))); };
template <> struct VT<float, 16> { typedef float t __attribute__((ext_vector_type(16))); };
template <> struct VT<float, 32> { typedef float t __attribute__((ext_vector_type(32))); };

// Neighbour gathers issued together before they are accumulated, so their
// memory latencies overlap instead of serializing on one wait per bond.
constexpr int kGatherBatch = 4;  // Quantum Sun outer-spin flips per batch

template <typename T, int V>
__device__ __forceinline__ typename VT<T, V>::t vload(const T* p) {
  return *reinterpret_cast<const typename VT<T, V>::t*>(p);
}
template <typename T, int V>
__device__ __forceinline__ void vstore(T* p, typename VT<T, V>::t v) {
  *reinterpret_cast<typename VT<T, V>::t*>(p) = v;
}
// Streamed (read-once / write-once) accesses with the non-temporal hint, so
// they do not evict the gathered rows from L2.
template <typename T, int V>
__device__ __forceinline__ typename VT<T, V>::t vload_nt(const T* p) {
  return __builtin_nontemporal_load(reinterpret_cast<const typename VT<T, V>::t*>(p));
}
template <typename T, int V>
__device__ __forceinline__ void vstore_nt(T* p, typename VT<T, V>::t v) {
  __builtin_nontemporal_store(v, reinterpret_cast<typename VT<T, V>::t*>(p));
}
// The same pointer in the global address space (for pointers read from
// device-side tables, which are otherwise generic and use flat instructions).
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* as_global(const T* p) {
  return (const __attribute__((address_space(1))) T*)p;
}
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* as_global(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

// Vector load at byte offset off from base, known to be global memory (a
// pointer read from a device-side table is otherwise generic and compiles to
// flat loads with 64-bit address arithmetic).
template <typename T, int V>
__device__ __forceinline__ typename VT<T, V>::t gload(const char* base, uint32_t off) {
  typedef typename VT<T, V>::t vt;
  typedef const __attribute__((address_space(1))) vt* gp;
  typedef const __attribute__((address_space(1))) char* gc;
  return *(gp)((gc)base + off);
}
template <typename T, int V>
__device__ __forceinline__ T vget(typename VT<T, V>::t v, int i) {
  if constexpr (V == 1) {
    return v;
  } else {
    return v[i];
  }
}

// Workgroups are dispatched round-robin over the 8 XCDs; this bijection gives
// XCD x a contiguous range of row groups, so the neighbour gathers of nearby
// rows hit that XCD's own L2.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t n) {
  const int64_t x = b & 7, i = b >> 3, q = n >> 3, r = n & 7;
  return x * q + (x < r ? x : r) + i;
}

// Block-wide sum of one double per thread (256 threads), result in thread 0.
__device__ __forceinline__ double block_sum_256(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
  return t;
}

}  // namespace pf
