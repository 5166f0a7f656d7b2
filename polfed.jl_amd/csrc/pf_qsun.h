// Quantum-Sun row kernel body shared by the single-operator step kernel
// (ops.hip) and the multi-instance batch kernel (batch.hip).
//
// H~ b2 at row r (full 2^(Lg+Lloc) space, rank = state) is
//   d(r) b2[r] + sum_{b != a} G~[a][b] b2[base | b] + sum_j c~_j b2[r ^ flip_j],
// a = r mod 2^Lg, base = r - a, d(r) = (Dlo[r mod 2^lb] + Dhi[r >> lb] - shift)
// * scale from the two diagonal tables built once per operator (FP64).  The
// scaled grain matrix, outer couplings and flip masks sit in LDS; gathers use
// 32-bit byte offsets from a uniform base (a shift and an XOR per flip when the
// row stride in bytes is a power of two).
#pragma once

#include "pf_device.h"
#include "pf_internal.h"

namespace pf {

constexpr int kMaxDg = 1 << kMaxGrain;
constexpr int kMaxLocPad = 32;  // outer spins padded to whole gather batches

// Per-block constants of one Quantum-Sun operator in one step.
struct QsunRow {
  const char* b2;        // base of the b2 block (bytes), uniform
  uint32_t cofs_b;       // this lane's column offset (bytes)
  uint32_t rsb;          // row stride (bytes)
  int rsh;               // log2(rsb), or -1 when rsb is not a power of two
  int dg;                // 2^Lg
  int nloc;              // outer spins, padded to a whole number of gather batches
  int lb;                // low bits of the diagonal split
  const double* dlo;     // Dlo[2^lb]
  const double* dhi;     // Dhi[2^(Lg+Lloc-lb)]
  double scale, shift;
  const uint32_t* sF;    // flip masks, pre-shifted by rsh when rsh >= 0 (LDS)
  const void* sC;        // scaled outer couplings, type T (LDS)
};

// Byte offset of row r (P2: the row stride is 2^rsh bytes).
template <bool P2>
__device__ __forceinline__ uint32_t qsun_off(const QsunRow& q, uint32_t r) {
  return (P2 ? (r << q.rsh) : r * q.rsb) + q.cofs_b;
}

// Stage the per-block LDS tables: sF (flip masks, shifted), sC (T couplings),
// sG (T grain matrix, dg x dg).  The flip list is padded to npad entries with
// zero masks and zero weights (a flip onto the row itself that adds nothing),
// so the gather loop runs whole batches without guards.  Call from all
// threads, then __syncthreads().
template <typename T>
__device__ __forceinline__ void qsun_stage(int nloc, int npad, int dg, const uint32_t* flip,
                                           const double* coup, const double* G, double scale,
                                           int rsh, uint32_t* sF, T* sC, T* sG) {
  for (int j = threadIdx.x; j < npad; j += blockDim.x) {
    const bool real = j < nloc;
    sF[j] = real ? (rsh >= 0 ? (flip[j] << rsh) : flip[j]) : 0u;
    sC[j] = real ? (T)(coup[j] * scale) : (T)0;
  }
  for (int t = threadIdx.x; t < dg * dg; t += blockDim.x) sG[t] = (T)(G[t] * scale);
}

// Outer spins rounded up to whole gather batches of nbf (the padding flips
// are zero-weight reloads of the row itself, cheaper than guarded batches).
__host__ __device__ __forceinline__ int qsun_pad(int nloc, int nbf) {
  return (nloc + nbf - 1) / nbf * nbf;
}

// x from lane (this lane ^ mask), component-wise (all lanes of the group active).
template <typename T, int V>
__device__ __forceinline__ typename VT<T, V>::t shfl_xor_vec(typename VT<T, V>::t x, int mask) {
  if constexpr (V == 1) {
    return __shfl_xor(x, mask, 64);
  } else {
    typename VT<T, V>::t y;
#pragma unroll
    for (int v = 0; v < V; ++v) y[v] = __shfl_xor(x[v], mask, 64);
    return y;
  }
}

// Gathers issued per batch before any is accumulated: about 128 bytes in
// flight per lane and batch, between 2 and 16 gathers.
template <typename T, int V>
constexpr int qsun_flip_batch() {
  return (128 / (V * (int)sizeof(T))) < 2 ? 2
         : (128 / (V * (int)sizeof(T))) > 16 ? 16
                                              : 128 / (V * (int)sizeof(T));
}

// H~ b2 at row r for this lane's V columns, DG = 2^Lg; sG holds the scaled
// grain matrix (LDS).  Every load of a group (diagonal tables, grain block, a
// batch of flips) is issued before the first of them is used; a batch reads
// its flip masks from LDS before one wait, and P2 (row stride a power of two,
// masks pre-shifted) makes each flip offset one XOR and one add, with no
// branch per flip.
template <typename T, int V, int DG, bool P2>
__device__ __forceinline__ typename VT<T, V>::t qsun_hrow(const QsunRow& q, uint32_t r,
                                                          const T* sG, int G) {
  typedef typename VT<T, V>::t vt;
  constexpr int NBF = qsun_flip_batch<T, V>();
  constexpr int GB = DG < NBF ? DG : NBF;  // grain loads per batch (load path)
  const T* sC = reinterpret_cast<const T*>(q.sC);
  const int a = (int)(r & (uint32_t)(DG - 1));
  const double dl = as_global(q.dlo)[r & ((1u << q.lb) - 1u)];
  const double dh = as_global(q.dhi)[r >> q.lb];
  const vt xr = gload<T, V>(q.b2, qsun_off<P2>(q, r));
  const uint32_t rs_r = P2 ? (r << q.rsh) : r;
  vt acc = (vt)(T)0;
  // outer flips: whole batches of NBF (q.nloc is padded to a multiple of NBF)
  for (int j0 = 0; j0 < q.nloc; j0 += NBF) {
    uint32_t f[NBF];
#pragma unroll
    for (int u = 0; u < NBF; ++u) f[u] = q.sF[j0 + u];
    vt x[NBF];
#pragma unroll
    for (int u = 0; u < NBF; ++u) {
      const uint32_t o = P2 ? ((rs_r ^ f[u]) + q.cofs_b) : ((r ^ f[u]) * q.rsb + q.cofs_b);
      x[u] = gload<T, V>(q.b2, o);
    }
#pragma unroll
    for (int u = 0; u < NBF; ++u) acc += sC[j0 + u] * x[u];
  }
  // grain block: with G a power of two its 2^Lg rows sit in lanes whose
  // indices differ by (a ^ a') * G (all active: a pass holds whole blocks), so
  // the rows' own values are shuffled instead of reloaded
  if (G > 0 && (G & (G - 1)) == 0 && DG * G <= 64) {
#pragma unroll
    for (int b = 1; b < DG; ++b) acc += sG[a * DG + (a ^ b)] * shfl_xor_vec<T, V>(xr, b * G);
  } else {
    const uint32_t gbase = qsun_off<P2>(q, r & ~(uint32_t)(DG - 1));
#pragma unroll
    for (int b0 = 0; b0 < DG; b0 += GB) {
      vt xg[GB];
#pragma unroll
      for (int b = 0; b < GB; ++b) xg[b] = gload<T, V>(q.b2, gbase + (uint32_t)(b0 + b) * q.rsb);
#pragma unroll
      for (int b = 0; b < GB; ++b) {
        const int bb = b0 + b;
        if (bb != a) acc += sG[a * DG + bb] * xg[b];
      }
    }
  }
  const T d = (T)((dl + dh - q.shift) * q.scale);
  acc += d * xr;
  return acc;
}

// qsun_hrow for the runtime grain size q.dg (2, 4, 8 or 16) and row-stride
// shape, G lanes per row.  Every lane of a grain block must take part (rows
// per pass are whole blocks and D is a multiple of 2^Lg).
template <typename T, int V, bool P2>
__device__ __forceinline__ typename VT<T, V>::t qsun_hrow_dg2(const QsunRow& q, uint32_t r,
                                                              const T* sG, int G) {
  switch (q.dg) {  // uniform
    case 2: return qsun_hrow<T, V, 2, P2>(q, r, sG, G);
    case 4: return qsun_hrow<T, V, 4, P2>(q, r, sG, G);
    case 16: return qsun_hrow<T, V, 16, P2>(q, r, sG, G);
    default: return qsun_hrow<T, V, 8, P2>(q, r, sG, G);
  }
}
template <typename T, int V>
__device__ __forceinline__ typename VT<T, V>::t qsun_hrow_dg(const QsunRow& q, uint32_t r,
                                                             const T* sG, int G) {
  return q.rsh >= 0 ? qsun_hrow_dg2<T, V, true>(q, r, sG, G)
                    : qsun_hrow_dg2<T, V, false>(q, r, sG, G);
}

}  // namespace pf
