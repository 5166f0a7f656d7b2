// Quantum-Sun row kernel body shared by the single-operator step kernel
// (ops.hip) and the multi-instance batch kernel (batch.hip).
//
// H~ b2 at row r (full 2^(Lg+Lloc) space, rank = state) is
//   d(r) b2[r] + sum_{b != a} G~[a][b] b2[base | b] + sum_j c~_j b2[r ^ flip_j],
// a = r mod 2^Lg, base = r - a, d(r) = (Dlo[r mod 2^lb] + Dhi[r >> lb] - shift)
// * scale from the two diagonal tables built once per operator (FP64).  The
// grain weights of a lane's fixed grain index live in registers, the scaled
// outer couplings and flip masks in LDS; gathers use 32-bit byte offsets from
// a uniform base (a shift and an XOR per flip when the row stride in bytes is
// a power of two).
#pragma once

#include "pf_device.h"
#include "pf_internal.h"

namespace pf {

constexpr int kMaxDg = 1 << kMaxGrain;

// Per-block constants of one Quantum-Sun operator in one step.
struct QsunRow {
  const char* b2;        // base of the b2 block (bytes), uniform
  uint32_t cofs_b;       // this lane's column offset (bytes)
  uint32_t rsb;          // row stride (bytes)
  int rsh;               // log2(rsb), or -1 when rsb is not a power of two
  int dg;                // 2^Lg
  int nloc;              // outer spins
  int lb;                // low bits of the diagonal split
  const double* dlo;     // Dlo[2^lb]
  const double* dhi;     // Dhi[2^(Lg+Lloc-lb)]
  double scale, shift;
  const uint32_t* sF;    // flip masks, pre-shifted by rsh when rsh >= 0 (LDS)
  const void* sC;        // scaled outer couplings, type T (LDS)
};

__device__ __forceinline__ uint32_t qsun_off(const QsunRow& q, uint32_t r) {
  return (q.rsh >= 0 ? (r << q.rsh) : r * q.rsb) + q.cofs_b;
}

// Stage the per-block LDS tables: sF (flip masks, shifted), sC (T couplings),
// sG (T grain matrix, dg x dg).  Call from all threads, then __syncthreads().
template <typename T>
__device__ __forceinline__ void qsun_stage(int nloc, int dg, const uint32_t* flip,
                                           const double* coup, const double* G, double scale,
                                           int rsh, uint32_t* sF, T* sC, T* sG) {
  for (int j = threadIdx.x; j < nloc; j += blockDim.x) {
    sF[j] = rsh >= 0 ? (flip[j] << rsh) : flip[j];
    sC[j] = (T)(coup[j] * scale);
  }
  for (int t = threadIdx.x; t < dg * dg; t += blockDim.x) sG[t] = (T)(G[t] * scale);
}

// This lane's grain weights G~[a][b] (0 at b == a, whose term carries d(r)).
template <typename T>
__device__ __forceinline__ void qsun_lane_weights(const T* sG, int dg, int a, T (&wg)[kMaxDg]) {
#pragma unroll
  for (int b = 0; b < kMaxDg; ++b) wg[b] = (b < dg && b != a) ? sG[a * dg + b] : (T)0;
}

// H~ b2 at row r for this lane's V columns.  fixed_a: the lane's grain index
// is the same for every row it visits (rows per block pass divisible by dg),
// so wg holds its weights; otherwise they are read from sG per row.
template <typename T, int V>
__device__ __forceinline__ typename VT<T, V>::t qsun_hrow(const QsunRow& q, uint32_t r,
                                                          bool fixed_a, const T (&wg)[kMaxDg],
                                                          const T* sG) {
  typedef typename VT<T, V>::t vt;
  const T* sC = reinterpret_cast<const T*>(q.sC);
  const int a = (int)(r & (uint32_t)(q.dg - 1));
  const double dd = q.dlo[r & ((1u << q.lb) - 1u)] + q.dhi[r >> q.lb];
  const T d = (T)((dd - q.shift) * q.scale);
  const uint32_t roff = qsun_off(q, r);
  vt acc = d * *reinterpret_cast<const vt*>(q.b2 + roff);
  const uint32_t gbase = qsun_off(q, r & ~(uint32_t)(q.dg - 1));
#pragma unroll
  for (int b = 0; b < kMaxDg; ++b) {
    if (b < q.dg) {  // uniform
      const T w = fixed_a ? wg[b] : (b == a ? (T)0 : sG[a * q.dg + b]);
      acc += w * *reinterpret_cast<const vt*>(q.b2 + gbase + (uint32_t)b * q.rsb);
    }
  }
  const uint32_t rs_r = q.rsh >= 0 ? (r << q.rsh) : r;
  for (int j0 = 0; j0 < q.nloc; j0 += kGatherBatch) {
    vt x[kGatherBatch];
#pragma unroll
    for (int u = 0; u < kGatherBatch; ++u) {
      const int j = j0 + u;
      x[u] = (vt)(T)0;
      if (j < q.nloc) {  // uniform
        const uint32_t o = q.rsh >= 0 ? ((rs_r ^ q.sF[j]) + q.cofs_b)
                                      : ((r ^ q.sF[j]) * q.rsb + q.cofs_b);
        x[u] = *reinterpret_cast<const vt*>(q.b2 + o);
      }
    }
#pragma unroll
    for (int u = 0; u < kGatherBatch; ++u)
      if (j0 + u < q.nloc) acc += sC[j0 + u] * x[u];
  }
  return acc;
}

}  // namespace pf
