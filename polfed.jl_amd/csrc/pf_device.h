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

// Broadcast a 32-bit value from lane Q of this lane's group of G lanes (the
// G lanes that share one row).  G <= 4 uses a DPP quad permutation, G = 8 a
// cross-lane permute.
template <int G, int Q>
__device__ __forceinline__ uint32_t group_bcast(uint32_t v) {
  if constexpr (G <= 1) {
    return v;
  } else if constexpr (G == 2) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, Q | (Q << 2) | ((Q + 2) << 4) | ((Q + 2) << 6),
                                              0xF, 0xF, false);
  } else if constexpr (G == 4) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, Q * 0x55, 0xF, 0xF, false);
  } else {
    const int lane = threadIdx.x & 63;
    return (uint32_t)__shfl((int)v, (lane & ~(G - 1)) + Q, 64);
  }
}

// The G = GM lanes of a row group each hold UB values (lane q's u-th value is
// item q + GM*u); afterwards every lane of the group holds all GM*UB items.
template <int GM, int UB>
__device__ __forceinline__ void group_share(const uint32_t (&mine)[UB],
                                            uint32_t (&out)[GM * UB]) {
#pragma unroll
  for (int u = 0; u < UB; ++u) {
#pragma unroll
    for (int q = 0; q < GM; ++q) {
      uint32_t v;
      if constexpr (GM == 1) v = mine[u];
      else if (q == 0) v = group_bcast<GM, 0>(mine[u]);
      else if (q == 1) v = group_bcast<GM, 1>(mine[u]);
      else if (q == 2) v = group_bcast<GM, 2>(mine[u]);
      else if (q == 3) v = group_bcast<GM, 3>(mine[u]);
      else if (q == 4) v = group_bcast<GM, 4>(mine[u]);
      else if (q == 5) v = group_bcast<GM, 5>(mine[u]);
      else if (q == 6) v = group_bcast<GM, 6>(mine[u]);
      else v = group_bcast<GM, 7>(mine[u]);
      out[q + GM * u] = v;
    }
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
