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

// ----------------------------------------------- block rows staged in LDS
// A block that owns the aligned row range [r0, r0 + 2^m) of one operator
// (row stride a power of two) copies that range of b2 to LDS once; the flips
// whose mask lies below 2^m (outer spins j < m - Lg: the flipped row stays in
// the range) and the grain block are then read from LDS, and only the other
// flips go through the vector memory pipeline.

constexpr int kLdsFlipBatch = 4;  // flips per batch in the LDS form (both lists)

// Local flips of an operator with Lg grain and nloc outer spins in a block of
// 2^m rows.
__host__ __device__ __forceinline__ int qsun_local_flips(int m, int Lg, int nloc) {
  const int nl = m - Lg;
  return nl < 0 ? 0 : (nl > nloc ? nloc : nl);
}

// Stage sF/sC/sG for the LDS form: the nl local flips first, padded to a
// multiple of kLdsFlipBatch with zero masks and weights (a read of the row
// itself that adds nothing), then the nloc - nl others, padded the same way.
// Masks are shifted by rsh (the row stride is 2^rsh bytes).  Call from all
// threads, then __syncthreads().
//
// Row-vector encoding (lw >= 1, the row stride one element): a lane holds 2^lw
// consecutive rows in one 16-byte vector, and the flip r ^ f maps the vector v
// of r to v ^ (f >> lw) with its components permuted by i -> i ^ (f mod 2^lw);
// sF then holds ((f >> lw) << 4) | (f mod 2^lw), the byte XOR of the vector
// with the permutation in its four low bits, and the grain matrix in sG has a
// zero diagonal (the diagonal is d(r)).
template <typename T>
__device__ __forceinline__ void qsun_stage_split(int nloc, int nl, int dg, const uint32_t* flip,
                                                 const double* coup, const double* G,
                                                 double scale, int rsh, uint32_t* sF, T* sC,
                                                 T* sG, int lw = 0) {
  constexpr int B = kLdsFlipBatch;
  const int nlp = (nl + B - 1) / B * B, nr = nloc - nl, nrp = (nr + B - 1) / B * B;
  for (int t = threadIdx.x; t < nlp + nrp; t += blockDim.x) {
    const bool loc = t < nlp;
    const int i = loc ? t : t - nlp;  // index inside its list
    const bool real = loc ? i < nl : i < nr;
    const int j = loc ? i : nl + i;
    const uint32_t f = real ? flip[j] : 0u;
    sF[t] = lw ? (((f >> lw) << 4) | (f & ((1u << lw) - 1u))) : (f << rsh);
    sC[t] = real ? (T)(coup[j] * scale) : (T)0;
  }
  for (int t = threadIdx.x; t < dg * dg; t += blockDim.x)
    sG[t] = (lw && t / dg == t % dg) ? (T)0 : (T)(G[t] * scale);
}

// Vector load at byte offset off from sB, known to be LDS.
template <typename T, int V>
__device__ __forceinline__ typename VT<T, V>::t lds_vload(const char* sB, uint32_t off) {
  typedef typename VT<T, V>::t vt;
  typedef const __attribute__((address_space(3))) vt* lp;
  typedef const __attribute__((address_space(3))) char* lc;
  return *(lp)((lc)sB + off);
}

// Rows a lane takes at once in the LDS form: about 16 bytes of each gather
// per lane and row group (4 FP32 rows, 2 FP64 rows, one row of 4+ columns).
template <typename T, int V>
constexpr int qsun_lds_rows() {
  return (16 / (V * (int)sizeof(T))) < 1 ? 1 : (16 / (V * (int)sizeof(T))) > 4 ? 4
                                                 : 16 / (V * (int)sizeof(T));
}

// H~ b2 at the R rows r[i] of this lane (its V columns) in the LDS form, into
// acc[i].  q.b2/q.cofs_b/q.rsh address the operator's b2 (row stride 2^rsh
// bytes); sB holds rows [r0, r0 + 2^m) with the same stride and rl_b[i] =
// ((r[i] - r0) << rsh) + cofs_b; sF/sC are laid out by qsun_stage_split with
// nlp local and nrp other entries (multiples of kLdsFlipBatch).  The R rows
// share their grain position a (they differ by multiples of 2^Lg).  Every
// batch issues its loads for all R rows before the first wait; the other
// flips go first (their gathers travel while the LDS reads run).
template <typename T, int V, int DG, int R>
__device__ __forceinline__ void qsun_hrows_lds(const QsunRow& q, const uint32_t* r,
                                               const uint32_t* rl_b, const char* sB,
                                               const T* sG, int nlp, int nrp,
                                               typename VT<T, V>::t* acc) {
  typedef typename VT<T, V>::t vt;
  constexpr int B = kLdsFlipBatch;
  const T* sC = reinterpret_cast<const T*>(q.sC);
  const int a = (int)(r[0] & (uint32_t)(DG - 1));
  double dl[R], dh[R];
  uint32_t rg_b[R];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    dl[i] = as_global(q.dlo)[r[i] & ((1u << q.lb) - 1u)];
    dh[i] = as_global(q.dhi)[r[i] >> q.lb];
    rg_b[i] = (r[i] << q.rsh) + q.cofs_b;  // the row's byte offset in b2
    acc[i] = (vt)(T)0;
  }
  for (int j0 = nlp; j0 < nlp + nrp; j0 += B) {
    // masks are uniform: a padding entry (mask 0) issues no gather
    uint32_t f[B];
#pragma unroll
    for (int u = 0; u < B; ++u) f[u] = (uint32_t)__builtin_amdgcn_readfirstlane((int)q.sF[j0 + u]);
    vt x[R][B];
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int u = 0; u < B; ++u) x[i][u] = f[u] ? gload<T, V>(q.b2, rg_b[i] ^ f[u]) : (vt)(T)0;
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int u = 0; u < B; ++u) acc[i] += sC[j0 + u] * x[i][u];
  }
  for (int j0 = 0; j0 < nlp; j0 += B) {
    uint32_t f[B];
#pragma unroll
    for (int u = 0; u < B; ++u) f[u] = q.sF[j0 + u];
    vt x[R][B];
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int u = 0; u < B; ++u) x[i][u] = lds_vload<T, V>(sB, rl_b[i] ^ f[u]);
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int u = 0; u < B; ++u) acc[i] += sC[j0 + u] * x[i][u];
  }
#pragma unroll
  for (int b = 1; b < DG; ++b) {
    const T w = sG[a * DG + (a ^ b)];
#pragma unroll
    for (int i = 0; i < R; ++i)
      acc[i] += w * lds_vload<T, V>(sB, rl_b[i] ^ ((uint32_t)b << q.rsh));
  }
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const T d = (T)((dl[i] + dh[i] - q.shift) * q.scale);
    acc[i] += d * lds_vload<T, V>(sB, rl_b[i]);
  }
}

template <typename T, int V, int R>
__device__ __forceinline__ void qsun_hrows_lds_dg(const QsunRow& q, const uint32_t* r,
                                                  const uint32_t* rl_b, const char* sB,
                                                  const T* sG, int nlp, int nrp,
                                                  typename VT<T, V>::t* acc) {
  switch (q.dg) {  // uniform
    case 2: qsun_hrows_lds<T, V, 2, R>(q, r, rl_b, sB, sG, nlp, nrp, acc); break;
    case 4: qsun_hrows_lds<T, V, 4, R>(q, r, rl_b, sB, sG, nlp, nrp, acc); break;
    case 16: qsun_hrows_lds<T, V, 16, R>(q, r, rl_b, sB, sG, nlp, nrp, acc); break;
    default: qsun_hrows_lds<T, V, 8, R>(q, r, rl_b, sB, sG, nlp, nrp, acc); break;
  }
}

}  // namespace pf
