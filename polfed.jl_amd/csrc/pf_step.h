// Streamed row inputs and the fused output of one Clenshaw step, shared by
// the step kernels of every operator kind (ops.hip, packed.hip).
#pragma once

#include "pf_device.h"
#include "pf_internal.h"

namespace pf {

// Streamed inputs of one row (X and b3), loaded before the neighbour gathers
// so that their latency overlaps them.
template <typename T, int V>
__device__ __forceinline__ void step_load(int flags, int64_t off, const T* __restrict__ X,
                                          const T* __restrict__ B3, typename VT<T, V>::t& xv,
                                          typename VT<T, V>::t& b3v) {
  typedef typename VT<T, V>::t vt;
  const bool nt = flags & kStreamNT;
  xv = (vt)(T)0;
  b3v = (vt)(T)0;
  if (flags & kHasX) xv = nt ? vload_nt<T, V>(X + off) : vload<T, V>(X + off);
  if (flags & kHasB3) b3v = nt ? vload_nt<T, V>(B3 + off) : vload<T, V>(B3 + off);
}

// out = cx*X + ch*h + cb*b3, store (interleaved T, or FP64 column-major with
// kOutCM at row operm[row], the reference rank, when the rows run in the
// partition order), optional KPM dot <P, out>.
template <typename T, int V>
__device__ __forceinline__ void step_store(int flags, int64_t row, int64_t off, int out_col0,
                                           T cx, T ch, T cb, typename VT<T, V>::t h,
                                           typename VT<T, V>::t xv, typename VT<T, V>::t b3v,
                                           void* __restrict__ OUT, int64_t out_ld,
                                           const uint32_t* __restrict__ operm,
                                           const T* __restrict__ Pr, double& dacc) {
  typedef typename VT<T, V>::t vt;
  vt o = (vt)(T)0;
  const bool nt = flags & kStreamNT;
  if (flags & kHasX) o = cx * xv;
  if (flags & kHasH) o += ch * h;
  if (flags & kHasB3) o += cb * b3v;
  if (flags & kOutCM) {
    double* Y = reinterpret_cast<double*>(OUT);
    const int64_t yr = operm ? (int64_t)operm[row] : row;
#pragma unroll
    for (int v = 0; v < V; ++v) Y[yr + (int64_t)(out_col0 + v) * out_ld] = (double)vget<T, V>(o, v);
  } else if (nt) {
    vstore_nt<T, V>(reinterpret_cast<T*>(OUT) + off, o);
  } else {
    vstore<T, V>(reinterpret_cast<T*>(OUT) + off, o);
  }
  if (flags & kDot) {
    const vt p = vload<T, V>(Pr + off);
#pragma unroll
    for (int v = 0; v < V; ++v) dacc += (double)vget<T, V>(p, v) * (double)vget<T, V>(o, v);
  }
}

}  // namespace pf
