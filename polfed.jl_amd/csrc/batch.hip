// Batched multi-instance Clenshaw filter for Quantum-Sun operators.
//
// Many independent instances (disorder realizations, uneven sizes) are
// filtered together: every Clenshaw step of every instance is one launch of
// qsun_batch_step_kernel over the concatenated rows of all instances.  The
// instances run in lockstep aligned at the END of their filters: at global
// step k (Kmax .. 0) instance i is active iff k <= K_i, its local step is
// g = K_i - k, and its buffer rotation, zero tracking and coefficient index
// all follow from (k, K_i) on the device.  Instances are laid out in
// decreasing K, so the active ones at step k are a prefix of the launch
// order and the stream-mode grid shrinks to them.
//
// Graph mode captures M consecutive step nodes, which read the current step
// from device memory (node j handles k = kstate - j), plus one node that
// lowers kstate by M; the graph is replayed ceil((Kmax+1)/M) times per
// filter application.  Nothing in the graph is a memset or a copy.
//
// Each instance has its own arithmetic (FP64 or FP32) per application, so a
// mixed-precision driver can switch instances independently.  Buffers are
// sized for FP64; an FP32 instance uses the first half of its region.
//
// Three kernel forms, chosen by the autotuner: rows from memory (each lane
// gathers its row's grain block and outer flips), block rows staged in LDS
// (the grain block and the flips that stay inside the block's aligned row
// range read from LDS), and row vectors for one column (a lane takes 16
// bytes of consecutive rows per load; blocks staged in LDS as well).

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "pf_device.h"
#include "pf_internal.h"
#include "pf_qsun.h"

namespace pf {

// Static per-instance record (device, launch order).
struct BatchInst {
  int32_t Lg, Lloc;
  int64_t D;
  int64_t row0;      // first row of this instance in the batch row space
  int64_t coef_off;  // first filter coefficient of this instance
  double scale, shift;
  const double* qdiag;  // the operator's diagonal tables (Dlo then Dhi)
  int32_t lb;           // low bits of the diagonal split
  uint32_t flip[kMaxLoc];
  double coup[kMaxLoc];
  double G[(1 << kMaxGrain) * (1 << kMaxGrain)];
};

// Per-application record (device, launch order).
struct BatchDyn {
  int32_t K;      // order of this application's filter; -1: not taking part
  int32_t dtype;  // PF_F64 / PF_F32
  double* Y;      // FP64 column-major output (device)
  int64_t ldy;
};

struct BatchArgs {
  const BatchInst* insts;
  const BatchDyn* dyn;
  const int64_t* blk0;  // n+1 prefix of blocks per instance, launch order
  const double* coefs;  // null: timing mode (every coefficient 0.5)
  const int* kstate;    // graph mode: kstate[0] is the step of node 0
  char* Xb;
  char* B0;
  char* B1;
  char* B2;
  int n;
  int polytype;
  int s;     // columns per instance
  int G;     // lanes per row
  int rpb;   // rows per block pass (256 / G)
  int rpt;   // block passes (rows per lane)
  int xcd;   // XCD-contiguous block mapping
  int shfl;  // 1: grain block by cross-lane shuffles, 0: by reloads
  int lgrows;  // LDS form: log2(rows per block)
};

// LDS form: the block's rows of b2 (dynamic shared memory, rows x row stride)
extern __shared__ __align__(16) char batch_rows[];

__device__ __forceinline__ void poly_ab_dev(int type, int k, double& alpha, double& beta_k) {
  switch (type) {
    case PF_LEGENDRE: alpha = (2.0 * k + 1) / (k + 1); beta_k = -(double)k / (k + 1); break;
    case PF_HERMITE: alpha = 2.0; beta_k = -2.0 * k; break;
    case PF_TAYLOR: alpha = 1.0; beta_k = 0.0; break;
    default: alpha = 2.0; beta_k = -1.0; break;  // Chebyshev
  }
}

// Buffer slot written at local step g (rotation (3,1,2) of the recurrence).
__device__ __forceinline__ int rot_slot(int g) { return (3 - g % 3) % 3; }

// One instance's step k: which terms it has, their coefficients and its
// buffers (the rotation follows the local step g = K - k).
template <typename T>
struct StepSetup {
  bool hasH, hasB3;
  T cx, ch, cb;
  const T* X;
  T* OUT;
  const char* B2;  // b2 (bytes)
  const T* B3;
};
template <typename T>
__device__ __forceinline__ StepSetup<T> step_setup(const BatchArgs& A, const BatchInst& I,
                                                   const BatchDyn& dy, int k) {
  StepSetup<T> r;
  const int g = dy.K - k;  // local step, >= 0
  double cxd, chd, cbd;
  const double ck = A.coefs ? A.coefs[I.coef_off + k] : 0.5;
  if (k >= 1) {
    double alpha, beta1, dummy;
    poly_ab_dev(A.polytype, k, alpha, dummy);
    poly_ab_dev(A.polytype, k + 1, dummy, beta1);
    cxd = ck;
    chd = alpha;
    cbd = beta1;
  } else {
    double dummy, beta1;
    poly_ab_dev(A.polytype, 1, dummy, beta1);
    cxd = ck;  // phi0 = 1 for every family
    chd = A.polytype == PF_HERMITE ? 2.0 : 1.0;
    cbd = beta1;
  }
  r.hasH = g >= 1;
  r.hasB3 = g >= 2 && cbd != 0.0;
  r.cx = (T)cxd;
  r.ch = (T)chd;
  r.cb = (T)cbd;
  const int64_t off = I.row0 * (int64_t)A.s * 8;  // byte offset of this instance's region
  char* bufs[3] = {A.B0, A.B1, A.B2};
  r.X = reinterpret_cast<const T*>(A.Xb + off);
  r.OUT = reinterpret_cast<T*>(bufs[rot_slot(g)] + off);
  r.B2 = bufs[rot_slot(g + 2)] + off;                         // g-1
  r.B3 = reinterpret_cast<const T*>(bufs[rot_slot(g + 1)] + off);  // g-2
  return r;
}

template <typename T, int V>
__device__ __forceinline__ void batch_body(const BatchArgs& A, const BatchInst& I,
                                           const BatchDyn& dy, int k, int64_t lb) {
  typedef typename VT<T, V>::t vt;
  __shared__ T sG[kMaxDg * kMaxDg];
  __shared__ T sC[kMaxLocPad];
  __shared__ uint32_t sF[kMaxLocPad];
  const int dg = 1 << I.Lg;
  const uint32_t rsb = (uint32_t)(A.s * sizeof(T));
  const int rsh = (rsb & (rsb - 1u)) == 0 ? __ffs(rsb) - 1 : -1;
  const int npad = qsun_pad(I.Lloc, qsun_flip_batch<T, V>());
  qsun_stage<T>(I.Lloc, npad, dg, I.flip, I.coup, I.G, I.scale, rsh, sF, sC, sG);
  __syncthreads();

  const StepSetup<T> st = step_setup<T>(A, I, dy, k);
  const bool hasH = st.hasH, hasB3 = st.hasB3;
  const T cx = st.cx, ch = st.ch, cb = st.cb;
  const T* __restrict__ X = st.X;
  T* __restrict__ OUT = st.OUT;
  const char* B2 = st.B2;
  const T* __restrict__ B3 = st.B3;

  const int rloc = threadIdx.x / A.G;
  const int gl = threadIdx.x - rloc * A.G;
  if (rloc >= A.rpb) return;
  const int64_t cofs = (int64_t)gl * V;
  QsunRow q;
  q.b2 = B2;
  q.cofs_b = (uint32_t)(cofs * sizeof(T));
  q.rsb = rsb;
  q.rsh = rsh;
  q.dg = dg;
  q.nloc = npad;
  q.lb = I.lb;
  q.dlo = I.qdiag;
  q.dhi = I.qdiag + ((int64_t)1 << I.lb);
  q.scale = I.scale;
  q.shift = I.shift;
  q.sF = sF;
  q.sC = sC;
  const int G_ = A.shfl ? A.G : 0;  // 0 selects the reload path
  for (int u = 0; u < A.rpt; ++u) {
    const int64_t row = (lb * A.rpt + u) * A.rpb + rloc;
    if (row >= I.D) break;
    const int64_t o = row * A.s + cofs;
    // (the rotating buffers are picked from a local table: global-space
    // accesses keep them off the flat path)
    const vt xv = *as_global(reinterpret_cast<const vt*>(X + o));
    const vt b3v = hasB3 ? *as_global(reinterpret_cast<const vt*>(B3 + o)) : (vt)(T)0;
    vt hv = (vt)(T)0;
    if (hasH) hv = qsun_hrow_dg<T, V>(q, (uint32_t)row, sG, G_);
    vt out = cx * xv;
    if (hasH) out += ch * hv;
    if (hasB3) out += cb * b3v;
    if (k >= 1) {
      *as_global(reinterpret_cast<vt*>(OUT + o)) = out;
    } else {
#pragma unroll
      for (int v = 0; v < V; ++v)
        as_global(dy.Y)[row + (int64_t)(gl * V + v) * dy.ldy] = (double)vget<T, V>(out, v);
    }
  }
}

// This block's instance (launch position lo), its block index lb inside the
// instance and the step k; false when the instance is idle at this step.
// kexpl >= 0: the step index; kexpl < 0: k = kstate[0] - j (graph nodes).
__device__ __forceinline__ bool batch_locate(const BatchArgs& A, int kexpl, int j, int& lo,
                                             BatchDyn& dy, int& k, int64_t& lb) {
  const int64_t nb = gridDim.x;
  int64_t b = blockIdx.x;
  if (A.xcd) b = xcd_block(b, nb);
  lo = 0;
  int hi = A.n;  // blk0[lo] <= b < blk0[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (A.blk0[mid] <= b) lo = mid;
    else hi = mid;
  }
  dy = A.dyn[lo];
  k = kexpl >= 0 ? kexpl : A.kstate[0] - j;
  lb = b - A.blk0[lo];
  return !(k < 0 || dy.K < 0 || k > dy.K);  // uniform over the block
}

// batch_body with the block's rows of b2 staged in LDS: the block owns the
// aligned range [r0, r0 + 2^lgrows) of its instance (G a power of two, so
// every thread has a row), copies it to LDS, and reads the grain block and
// the flips that stay in the range from there (qsun_hrows_lds).
template <typename T, int V>
__device__ __forceinline__ void batch_body_lds(const BatchArgs& A, const BatchInst& I,
                                               const BatchDyn& dy, int k, int64_t lb) {
  typedef typename VT<T, V>::t vt;
  __shared__ T sG[kMaxDg * kMaxDg];
  __shared__ T sC[kMaxLocPad];
  __shared__ uint32_t sF[kMaxLocPad];
  const int dg = 1 << I.Lg;
  const uint32_t rsb = (uint32_t)(A.s * sizeof(T));
  const int rsh = __ffs(rsb) - 1;  // the row stride is a power of two in this form
  const int nl = qsun_local_flips(A.lgrows, I.Lg, I.Lloc);
  constexpr int B = kLdsFlipBatch;
  const int nlp = (nl + B - 1) / B * B, nrp = (I.Lloc - nl + B - 1) / B * B;
  qsun_stage_split<T>(I.Lloc, nl, dg, I.flip, I.coup, I.G, I.scale, rsh, sF, sC, sG);

  const StepSetup<T> st = step_setup<T>(A, I, dy, k);
  const bool hasH = st.hasH, hasB3 = st.hasB3;
  const T cx = st.cx, ch = st.ch, cb = st.cb;
  const T* __restrict__ X = st.X;
  T* __restrict__ OUT = st.OUT;
  const char* B2 = st.B2;
  const T* __restrict__ B3 = st.B3;

  const int rloc = threadIdx.x / A.G;
  const int gl = threadIdx.x - rloc * A.G;
  const uint32_t cofs_b = (uint32_t)(gl * V * sizeof(T));
  const int64_t r0 = lb * A.rpt * A.rpb;  // first row of the block, a multiple of 2^lgrows
  // A block larger than its instance (D < 2^lgrows, so r0 = 0) reads its
  // rows >= D at row mod D (D is a power of two) and stores nothing for them:
  // every row group is whole and needs no branch.
  const uint32_t dmask = (uint32_t)(I.D - 1);
  constexpr int R = qsun_lds_rows<T, V>();  // A.rpt is a multiple of R
  typedef __attribute__((address_space(3))) vt lvt;
  typedef __attribute__((address_space(3))) char lch;
  if (hasH) {  // uniform over the block
    for (int u0 = 0; u0 < A.rpt; u0 += R) {
      vt xr[R];
#pragma unroll
      for (int i = 0; i < R; ++i) {
        const uint32_t row = (uint32_t)(r0 + (u0 + i) * A.rpb + rloc) & dmask;
        xr[i] = gload<T, V>(B2, (row << rsh) + cofs_b);
      }
#pragma unroll
      for (int i = 0; i < R; ++i)
        *reinterpret_cast<lvt*>((lch*)batch_rows +
                                ((uint32_t)((u0 + i) * A.rpb + rloc) << rsh) + cofs_b) = xr[i];
    }
  }
  __syncthreads();  // the LDS tables and rows

  QsunRow q;
  q.b2 = B2;
  q.cofs_b = cofs_b;
  q.rsb = rsb;
  q.rsh = rsh;
  q.dg = dg;
  q.nloc = nlp + nrp;
  q.lb = I.lb;
  q.dlo = I.qdiag;
  q.dhi = I.qdiag + ((int64_t)1 << I.lb);
  q.scale = I.scale;
  q.shift = I.shift;
  q.sF = sF;
  q.sC = sC;
  const int64_t cofs = (int64_t)gl * V;
  for (int u0 = 0; u0 < A.rpt; u0 += R) {
    uint32_t row[R], rl_b[R];
    vt xv[R], b3v[R], hv[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int rl = (u0 + i) * A.rpb + rloc;
      row[i] = (uint32_t)(r0 + rl) & dmask;
      rl_b[i] = ((uint32_t)rl << rsh) + cofs_b;
      const int64_t o = (int64_t)row[i] * A.s + cofs;
      xv[i] = *as_global(reinterpret_cast<const vt*>(X + o));
      b3v[i] = hasB3 ? *as_global(reinterpret_cast<const vt*>(B3 + o)) : (vt)(T)0;
      hv[i] = (vt)(T)0;
    }
    if (hasH) qsun_hrows_lds_dg<T, V, R>(q, row, rl_b, batch_rows, sG, nlp, nrp, hv);
#pragma unroll
    for (int i = 0; i < R; ++i) {
      if (r0 + (u0 + i) * A.rpb + rloc >= I.D) continue;
      vt out = cx * xv[i];
      if (hasH) out += ch * hv[i];
      if (hasB3) out += cb * b3v[i];
      const int64_t o = (int64_t)row[i] * A.s + cofs;
      if (k >= 1) {
        *as_global(reinterpret_cast<vt*>(OUT + o)) = out;
      } else {
#pragma unroll
        for (int v = 0; v < V; ++v)
          as_global(dy.Y)[(int64_t)row[i] + (int64_t)(gl * V + v) * dy.ldy] =
              (double)vget<T, V>(out, v);
      }
    }
  }
}

// ------------------------------------------------- row-vector form (ncols 1)
// Single-column instances: a lane takes W = 16 / sizeof(T) consecutive rows
// in one 16-byte vector (4 FP32 or 2 FP64 rows).  The flip r ^ f maps the
// aligned vector v of r to v ^ (f >> log2 W), components permuted by
// i -> i ^ (f mod W), so one 16-byte load serves W rows: a quarter (FP32) or
// half (FP64) of the vector-memory instructions and address arithmetic of a
// lane per row.  A block owns 1024 * rpt rows (rpt 2 or 4), staged in LDS as
// in the LDS form; each lane takes its vectors two at a time.

template <typename T>
struct RowVec {
  static constexpr int W = 16 / (int)sizeof(T);  // rows per vector
  static constexpr int LW = W == 4 ? 2 : 1;      // log2 W
  typedef typename VT<T, W>::t t;
};

// acc += c * x with the components of x permuted by i -> i ^ p (p uniform)
template <typename T>
__device__ __forceinline__ void rv_fma_perm(typename RowVec<T>::t& acc, T c,
                                            const typename RowVec<T>::t& x, uint32_t p) {
  if constexpr (RowVec<T>::W == 2) {
    if (p & 1u) {
      acc[0] += c * x[1];
      acc[1] += c * x[0];
    } else {
      acc += c * x;
    }
  } else {
    switch (p & 3u) {
      case 0: acc += c * x; break;
      case 1:
        acc[0] += c * x[1]; acc[1] += c * x[0]; acc[2] += c * x[3]; acc[3] += c * x[2];
        break;
      case 2:
        acc[0] += c * x[2]; acc[1] += c * x[3]; acc[2] += c * x[0]; acc[3] += c * x[1];
        break;
      default:
        acc[0] += c * x[3]; acc[1] += c * x[2]; acc[2] += c * x[1]; acc[3] += c * x[0];
        break;
    }
  }
}

// H~ b2 on the R vectors vg[i] (instance vector index; vl[i] the index in
// the block's LDS copy sB) of this lane, into acc[i].  sF/sC/sG as staged by
// qsun_stage_split in the row-vector encoding.
template <typename T, int DG, int R>
__device__ __forceinline__ void rv_hvec(const char* b2, const char* sB, const uint32_t* vg,
                                        const uint32_t* vl, const uint32_t* sF, const T* sC,
                                        const T* sG, int nlp, int nrp, const double* dlo,
                                        const double* dhi, int lb, double shift, double scale,
                                        typename RowVec<T>::t* acc) {
  typedef typename RowVec<T>::t wt;
  constexpr int W = RowVec<T>::W, LW = RowVec<T>::LW, B = kLdsFlipBatch;
  double dl[R][W], dh[R];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const uint32_t row = vg[i] << LW;
    const uint32_t lo = (row & ((1u << lb) - 1u)) * 8u;  // byte offset in Dlo, 16-byte aligned
#pragma unroll
    for (int h = 0; h < W / 2; ++h) {
      const typename VT<double, 2>::t d2 =
          gload<double, 2>(reinterpret_cast<const char*>(dlo), lo + 16u * h);
      dl[i][2 * h] = d2[0];
      dl[i][2 * h + 1] = d2[1];
    }
    dh[i] = as_global(dhi)[row >> lb];
    acc[i] = (wt)(T)0;
  }
  // flips whose rows lie outside the block: gathered from memory
  for (int j0 = nlp; j0 < nlp + nrp; j0 += B) {
    uint32_t F[B];
#pragma unroll
    for (int u = 0; u < B; ++u) F[u] = (uint32_t)__builtin_amdgcn_readfirstlane((int)sF[j0 + u]);
    wt x[R][B];
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int u = 0; u < B; ++u)  // a padding entry (F = 0) issues no gather
        x[i][u] = F[u] ? gload<T, W>(b2, (vg[i] << 4) ^ (F[u] & ~15u)) : (wt)(T)0;
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const T c = sC[j0 + u];
#pragma unroll
      for (int i = 0; i < R; ++i) rv_fma_perm<T>(acc[i], c, x[i][u], F[u] & 15u);
    }
  }
  // flips inside the block: from LDS
  for (int j0 = 0; j0 < nlp; j0 += B) {
    uint32_t F[B];
#pragma unroll
    for (int u = 0; u < B; ++u) F[u] = (uint32_t)__builtin_amdgcn_readfirstlane((int)sF[j0 + u]);
    wt x[R][B];
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int u = 0; u < B; ++u) x[i][u] = lds_vload<T, W>(sB, (vl[i] << 4) ^ (F[u] & ~15u));
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const T c = sC[j0 + u];
#pragma unroll
      for (int i = 0; i < R; ++i) rv_fma_perm<T>(acc[i], c, x[i][u], F[u] & 15u);
    }
  }
  // grain block (zero diagonal in sG) and the diagonal
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const wt own = lds_vload<T, W>(sB, vl[i] << 4);
    if constexpr (DG >= W) {
      constexpr int GV = DG / W;  // vectors per grain block
      const uint32_t vb = vl[i] & ~(uint32_t)(GV - 1);
      wt xg[GV];
#pragma unroll
      for (int t = 0; t < GV; ++t) xg[t] = lds_vload<T, W>(sB, (vb + t) << 4);
      const int a0 = (int)(vl[i] & (uint32_t)(GV - 1)) << LW;  // grain position of component 0
#pragma unroll
      for (int c = 0; c < W; ++c) {
        const char* grow = reinterpret_cast<const char*>(sG + (a0 + c) * DG);
#pragma unroll
        for (int t = 0; t < GV; ++t) {
          const wt w = lds_vload<T, W>(grow, 16u * t);
#pragma unroll
          for (int e = 0; e < W; ++e) acc[i][c] += w[e] * xg[t][e];
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < W; ++c) {
        const int a = c & (DG - 1), base = c & ~(DG - 1);
#pragma unroll
        for (int b = 0; b < DG; ++b) acc[i][c] += sG[a * DG + b] * own[base + b];
      }
    }
#pragma unroll
    for (int c = 0; c < W; ++c) acc[i][c] += (T)((dl[i][c] + dh[i] - shift) * scale) * own[c];
  }
}

template <typename T, int DG, int R>
__device__ __forceinline__ void batch_body_rv_dg(const BatchArgs& A, const BatchInst& I,
                                                 const BatchDyn& dy, int k, int64_t lb, T* sG,
                                                 T* sC, uint32_t* sF) {
  typedef typename RowVec<T>::t wt;
  constexpr int W = RowVec<T>::W, LW = RowVec<T>::LW;
  const int nl = qsun_local_flips(A.lgrows, I.Lg, I.Lloc);
  constexpr int B = kLdsFlipBatch;
  const int nlp = (nl + B - 1) / B * B, nrp = (I.Lloc - nl + B - 1) / B * B;
  qsun_stage_split<T>(I.Lloc, nl, DG, I.flip, I.coup, I.G, I.scale, 0, sF, sC, sG, LW);
  const StepSetup<T> st = step_setup<T>(A, I, dy, k);

  const uint32_t tid = threadIdx.x;
  const int64_t r0 = lb << A.lgrows;                    // first row of the block
  const uint32_t v0 = (uint32_t)(r0 >> LW);             // its first vector
  const int nvl = A.rpt * (4 / W);                      // vectors per lane, a multiple of R
  const uint32_t nvec = (uint32_t)(I.D >> LW);          // vectors of the instance
  const uint32_t vmask = nvec - 1;                      // (a power of two)
  typedef __attribute__((address_space(3))) wt lwt;
  typedef __attribute__((address_space(3))) char lch;
  if (st.hasH) {  // uniform over the block
    for (int u0 = 0; u0 < nvl; u0 += R) {
      wt xr[R];
#pragma unroll
      for (int i = 0; i < R; ++i)
        xr[i] = gload<T, W>(st.B2, ((v0 + (uint32_t)(u0 + i) * 256u + tid) & vmask) << 4);
#pragma unroll
      for (int i = 0; i < R; ++i)
        *reinterpret_cast<lwt*>((lch*)batch_rows + (((uint32_t)(u0 + i) * 256u + tid) << 4)) =
            xr[i];
    }
  }
  __syncthreads();  // the LDS tables and rows

  const char* Xc = reinterpret_cast<const char*>(st.X);
  const char* B3c = reinterpret_cast<const char*>(st.B3);
  for (int u0 = 0; u0 < nvl; u0 += R) {
    uint32_t vl[R], vg[R];
    wt xv[R], b3v[R], hv[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      vl[i] = (uint32_t)(u0 + i) * 256u + tid;
      vg[i] = (v0 + vl[i]) & vmask;  // a block larger than its instance reads row mod D
      xv[i] = gload<T, W>(Xc, vg[i] << 4);
      b3v[i] = st.hasB3 ? gload<T, W>(B3c, vg[i] << 4) : (wt)(T)0;
      hv[i] = (wt)(T)0;
    }
    if (st.hasH)
      rv_hvec<T, DG, R>(st.B2, batch_rows, vg, vl, sF, sC, sG, nlp, nrp, I.qdiag,
                     I.qdiag + ((int64_t)1 << I.lb), I.lb, I.shift, I.scale, hv);
#pragma unroll
    for (int i = 0; i < R; ++i) {
      if (v0 + vl[i] >= nvec) continue;  // past the end of a small instance: no store
      wt out = st.cx * xv[i];
      if (st.hasH) out += st.ch * hv[i];
      if (st.hasB3) out += st.cb * b3v[i];
      if (k >= 1) {
        *as_global(reinterpret_cast<wt*>(reinterpret_cast<char*>(st.OUT) + (vg[i] << 4))) = out;
      } else {
#pragma unroll
        for (int c = 0; c < W; ++c)
          as_global(dy.Y)[((int64_t)vg[i] << LW) + c] = (double)out[c];
      }
    }
  }
}

template <typename T, int R>
__device__ __forceinline__ void batch_body_rv(const BatchArgs& A, const BatchInst& I,
                                              const BatchDyn& dy, int k, int64_t lb) {
  __shared__ __align__(16) T sG[kMaxDg * kMaxDg];
  __shared__ T sC[kMaxLocPad];
  __shared__ uint32_t sF[kMaxLocPad];
  switch (I.Lg) {  // uniform
    case 1: batch_body_rv_dg<T, 2, R>(A, I, dy, k, lb, sG, sC, sF); break;
    case 2: batch_body_rv_dg<T, 4, R>(A, I, dy, k, lb, sG, sC, sF); break;
    case 4: batch_body_rv_dg<T, 16, R>(A, I, dy, k, lb, sG, sC, sF); break;
    default: batch_body_rv_dg<T, 8, R>(A, I, dy, k, lb, sG, sC, sF); break;
  }
}

// Kernel forms: rows per lane from memory (0), block rows in LDS (1), row
// vectors with block rows in LDS, two (2) or one (3) vectors at a time (one
// column, V = 1).
constexpr int kFormPlain = 0, kFormLds = 1, kFormRowVec = 2, kFormRowVec1 = 3;

template <typename T, int V, int FORM>
__device__ __forceinline__ void batch_body_any(const BatchArgs& A, const BatchInst& I,
                                               const BatchDyn& dy, int k, int64_t lb) {
  if constexpr (FORM == kFormRowVec) batch_body_rv<T, 2>(A, I, dy, k, lb);
  else if constexpr (FORM == kFormRowVec1) batch_body_rv<T, 1>(A, I, dy, k, lb);
  else if constexpr (FORM == kFormLds) batch_body_lds<T, V>(A, I, dy, k, lb);
  else batch_body<T, V>(A, I, dy, k, lb);
}

// One Clenshaw step of every active instance whose arithmetic is T.
template <typename T, int V, int FORM>
__global__ __launch_bounds__(256) void qsun_batch_step_kernel(const BatchArgs A, int kexpl,
                                                              int j) {
  int lo, k;
  BatchDyn dy;
  int64_t lb;
  constexpr int kDt = sizeof(T) == 8 ? PF_F64 : PF_F32;
  if (!batch_locate(A, kexpl, j, lo, dy, k, lb) || dy.dtype != kDt) return;
  batch_body_any<T, V, FORM>(A, A.insts[lo], dy, k, lb);
}

// One Clenshaw step of every active instance, FP64 and FP32 instances in the
// same launch (each block takes its instance's arithmetic): an application
// with both kinds costs one launch per step, not two over the whole grid.
template <int V, int FORM>
__global__ __launch_bounds__(256) void qsun_batch_step_kernel_mixed(const BatchArgs A, int kexpl,
                                                                    int j) {
  int lo, k;
  BatchDyn dy;
  int64_t lb;
  if (!batch_locate(A, kexpl, j, lo, dy, k, lb)) return;
  if (dy.dtype == PF_F64) batch_body_any<double, V, FORM>(A, A.insts[lo], dy, k, lb);
  else batch_body_any<float, V, FORM>(A, A.insts[lo], dy, k, lb);
}

// dtype bit set of an application: bit 0 FP64 instances, bit 1 FP32 ones
constexpr int kMaskF64 = 1, kMaskF32 = 2;

// LDS form: most bytes of b2 rows a block stages (FP64 rows)
constexpr size_t kBatchLdsBytes = 32 << 10;

__global__ void batch_kset_kernel(int* kstate, int value) {
  if (threadIdx.x == 0 && blockIdx.x == 0) kstate[0] = value;
}
__global__ void batch_kdec_kernel(int* kstate, int m) {
  if (threadIdx.x == 0 && blockIdx.x == 0) kstate[0] -= m;
}

static bool batch_vec_ok(int s, int V) {
  return (V == 1 || V == 2 || V == 4 || V == 8) && s % V == 0 && s / V <= 256;
}

}  // namespace pf

using namespace pf;

struct pf_batch {
  pf_handle* h = nullptr;
  int n = 0, s = 1;
  std::vector<pf_op*> ops;
  std::vector<int> order;      // launch position -> instance
  std::vector<int> pos;        // instance -> launch position
  std::vector<int64_t> row0;   // launch order
  std::vector<int64_t> blk0;   // launch order, n+1
  std::vector<int32_t> K;      // launch order (-1 before a filter is set)
  std::vector<int64_t> coef_off;  // launch order
  int64_t rows_total = 0;
  int polytype = PF_CHEBYSHEV;
  bool have_filter = false;
  DevBuf insts, dyn, blk0d, coefs, kstate, bufs;
  // tuning
  int V = 1, rpt = 1, xcd = 1, use_graph = 1, gsteps = 96, shfl = 1, lds = 0, rv = 0;
  // pinned staging of the per-application records; the event marks the end
  // of the last copy out of it
  BatchDyn* hdyn = nullptr;
  hipEvent_t dyn_done = nullptr;
  bool dyn_pending = false;
  // graph state
  hipStream_t cap = nullptr;
  hipGraph_t graph[4] = {nullptr, nullptr, nullptr, nullptr};      // by dtype mask
  hipGraphExec_t exec[4] = {nullptr, nullptr, nullptr, nullptr};

  int G() const { return s / V; }
  int rpb() const { return 256 / G(); }
  // The LDS form needs power-of-two rows per block and row strides, at least
  // kMaxDg rows per pass (the rows of a lane then share their grain
  // position), rows per lane in whole FP32 row groups, and the block's rows
  // within kBatchLdsBytes in FP64.
  bool lds_ok() const {
    const int rows = rpb() * rpt;
    return 256 % G() == 0 && (rows & (rows - 1)) == 0 && rpb() >= kMaxDg &&
           rpt % lds_rows_f32() == 0 && (s & (s - 1)) == 0 &&
           (size_t)rows * s * 8 <= kBatchLdsBytes;
  }
  bool use_lds() const { return lds && lds_ok(); }
  // The row-vector form (rv = vectors a lane takes at once, 1 or 2): one
  // column, one column per lane, 1024 * rpt rows per block with rpt 1, 2 or
  // 4 (whole vector pairs per lane in FP32 for rv = 2; FP64 rows within
  // kBatchLdsBytes).
  bool rv_ok() const {
    return s == 1 && V == 1 && (rpt == 1 || rpt == 2 || rpt == 4) && !(rv == 2 && rpt == 1);
  }
  bool use_rv() const { return rv && rv_ok(); }
  int rows_per_block() const { return use_rv() ? 1024 * rpt : rpb() * rpt; }
  int form() const {
    return use_rv() ? (rv == 2 ? kFormRowVec : kFormRowVec1) : use_lds() ? kFormLds : kFormPlain;
  }
  // rows a lane takes at once in the LDS form with FP32 (a multiple of the
  // FP64 figure)
  int lds_rows_f32() const {
    switch (V) {
      case 1: return qsun_lds_rows<float, 1>();
      case 2: return qsun_lds_rows<float, 2>();
      case 4: return qsun_lds_rows<float, 4>();
      default: return qsun_lds_rows<float, 8>();
    }
  }
  void drop_graph() {
    for (int m = 0; m < 4; ++m) {
      if (exec[m]) (void)hipGraphExecDestroy(exec[m]);
      if (graph[m]) (void)hipGraphDestroy(graph[m]);
      exec[m] = nullptr;
      graph[m] = nullptr;
    }
  }
  // Stage per-application records and copy them to the device on st.
  void upload_dyn(const std::vector<BatchDyn>& dy, hipStream_t st) {
    if (!hdyn) {
      PF_HIP(hipHostMalloc(reinterpret_cast<void**>(&hdyn), n * sizeof(BatchDyn)));
      PF_HIP(hipEventCreateWithFlags(&dyn_done, hipEventDisableTiming));
    }
    if (dyn_pending) PF_HIP(hipEventSynchronize(dyn_done));
    std::memcpy(hdyn, dy.data(), n * sizeof(BatchDyn));
    PF_HIP(hipMemcpyAsync(dyn.p, hdyn, n * sizeof(BatchDyn), hipMemcpyHostToDevice, st));
    PF_HIP(hipEventRecord(dyn_done, st));
    dyn_pending = true;
  }
  // block prefix in launch order for the current tiling (after the device is
  // idle: kernels of an earlier call may still read the old table)
  void layout_blocks() {
    PF_HIP(hipDeviceSynchronize());
    blk0.assign(n + 1, 0);
    const int64_t per = rows_per_block();
    for (int p = 0; p < n; ++p) {
      const int64_t D = ops[order[p]]->D;
      blk0[p + 1] = blk0[p] + (D + per - 1) / per;
    }
    PF_REQUIRE(blk0[n] < ((int64_t)1 << 31), "batch: grid too large");
    int64_t* d = reinterpret_cast<int64_t*>(blk0d.get((n + 1) * sizeof(int64_t)));
    PF_HIP(hipMemcpy(d, blk0.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    drop_graph();
  }
  // instance records in launch order
  void upload_insts() {
    PF_HIP(hipDeviceSynchronize());
    std::vector<BatchInst> recs(n);
    int64_t r0 = 0;
    row0.assign(n, 0);
    for (int p = 0; p < n; ++p) {
      const pf_op* op = ops[order[p]];
      BatchInst& I = recs[p];
      std::memset(&I, 0, sizeof(I));
      I.Lg = op->qsun.Lg;
      I.Lloc = op->qsun.Lloc;
      I.D = op->D;
      I.row0 = r0;
      I.coef_off = coef_off.empty() ? 0 : coef_off[p];
      const Affine af = op->affine(true);
      I.scale = af.scale;
      I.shift = af.shift;
      I.qdiag = op->qdiag;
      I.lb = op->qlb;
      for (int j = 0; j < kMaxLoc; ++j) {
        I.flip[j] = op->qsun.flip[j];
        I.coup[j] = op->qsun.coup[j];
      }
      std::memcpy(I.G, op->qsun.G, sizeof(I.G));
      row0[p] = r0;
      r0 += op->D;
    }
    rows_total = r0;
    void* d = insts.get(n * sizeof(BatchInst));
    PF_HIP(hipMemcpy(d, recs.data(), n * sizeof(BatchInst), hipMemcpyHostToDevice));
  }
  BatchArgs args() {
    BatchArgs A{};
    A.insts = reinterpret_cast<const BatchInst*>(insts.p);
    A.dyn = reinterpret_cast<const BatchDyn*>(dyn.p);
    A.blk0 = reinterpret_cast<const int64_t*>(blk0d.p);
    A.coefs = reinterpret_cast<const double*>(coefs.p);
    A.kstate = reinterpret_cast<const int*>(kstate.p);
    char* base = reinterpret_cast<char*>(bufs.p);
    const size_t reg = (size_t)rows_total * s * 8;
    A.Xb = base;
    A.B0 = base + reg;
    A.B1 = base + 2 * reg;
    A.B2 = base + 3 * reg;
    A.n = n;
    A.polytype = polytype;
    A.s = s;
    A.G = G();
    A.rpb = rpb();
    A.rpt = rpt;
    A.xcd = xcd;
    A.shfl = shfl;
    A.lgrows = __builtin_ctz((unsigned)rows_per_block());
    return A;
  }
  // dynamic LDS of a launch: the block's rows in the LDS forms
  size_t lds_bytes(size_t elem) const {
    return form() == kFormPlain ? 0 : (size_t)rows_per_block() * s * elem;
  }
  template <typename T, int F>
  void launch_t(const BatchArgs& A, int64_t grid, int kexpl, int j, hipStream_t st) {
    const size_t sh = lds_bytes(sizeof(T));
#define PF_BATCH_LAUNCH(VV)                                                                \
  hipLaunchKernelGGL((qsun_batch_step_kernel<T, VV, F>), dim3((unsigned)grid), dim3(256), sh, \
                     st, A, kexpl, j)
    if constexpr (F == kFormRowVec || F == kFormRowVec1) {
      PF_BATCH_LAUNCH(1);
    } else {
      switch (V) {
        case 2: PF_BATCH_LAUNCH(2); break;
        case 4: PF_BATCH_LAUNCH(4); break;
        case 8: PF_BATCH_LAUNCH(8); break;
        default: PF_BATCH_LAUNCH(1); break;
      }
    }
#undef PF_BATCH_LAUNCH
  }
  template <int F>
  void launch_mixed(const BatchArgs& A, int64_t grid, int kexpl, int j, hipStream_t st) {
    const size_t sh = lds_bytes(sizeof(double));
#define PF_BATCH_LAUNCH(VV)                                                       \
  hipLaunchKernelGGL((qsun_batch_step_kernel_mixed<VV, F>), dim3((unsigned)grid), \
                     dim3(256), sh, st, A, kexpl, j)
    if constexpr (F == kFormRowVec || F == kFormRowVec1) {
      PF_BATCH_LAUNCH(1);
    } else {
      switch (V) {
        case 2: PF_BATCH_LAUNCH(2); break;
        case 4: PF_BATCH_LAUNCH(4); break;
        case 8: PF_BATCH_LAUNCH(8); break;
        default: PF_BATCH_LAUNCH(1); break;
      }
    }
#undef PF_BATCH_LAUNCH
  }
  template <int F>
  void launch_f(const BatchArgs& A, int64_t grid, int kexpl, int j, hipStream_t st, int mask) {
    if (mask == (kMaskF64 | kMaskF32)) launch_mixed<F>(A, grid, kexpl, j, st);
    else if (mask & kMaskF64) launch_t<double, F>(A, grid, kexpl, j, st);
    else if (mask & kMaskF32) launch_t<float, F>(A, grid, kexpl, j, st);
  }
  // one step launch covering the arithmetics present in the application (mask)
  void launch(const BatchArgs& A, int64_t grid, int kexpl, int j, hipStream_t st, int mask) {
    if (grid <= 0) return;
    switch (form()) {
      case kFormRowVec: launch_f<kFormRowVec>(A, grid, kexpl, j, st, mask); break;
      case kFormRowVec1: launch_f<kFormRowVec1>(A, grid, kexpl, j, st, mask); break;
      case kFormLds: launch_f<kFormLds>(A, grid, kexpl, j, st, mask); break;
      default: launch_f<kFormPlain>(A, grid, kexpl, j, st, mask); break;
    }
    PF_HIP(hipGetLastError());
  }
  // timing: capture with constant coefficients (no coefficient reads), for
  // the autotuner's out-of-range step indices
  void ensure_graph(int mask, bool timing = false) {
    if (exec[mask]) return;
    if (!cap) PF_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    BatchArgs A = args();
    if (timing) A.coefs = nullptr;
    PF_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    try {
      for (int j = 0; j < gsteps; ++j) launch(A, blk0[n], -1, j, cap, mask);
      hipLaunchKernelGGL(batch_kdec_kernel, dim3(1), dim3(64), 0, cap,
                         reinterpret_cast<int*>(kstate.p), gsteps);
      PF_HIP(hipGetLastError());
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(cap, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    PF_HIP(hipStreamEndCapture(cap, &graph[mask]));
    PF_HIP(hipGraphInstantiate(&exec[mask], graph[mask], nullptr, nullptr, 0));
  }
  // Run steps kmax .. 0 of the current application on stream st.
  void run_steps(int kmax, hipStream_t st, int mask) {
    const BatchArgs A = args();
    if (use_graph) {
      ensure_graph(mask);
      hipLaunchKernelGGL(batch_kset_kernel, dim3(1), dim3(64), 0, st,
                         reinterpret_cast<int*>(kstate.p), kmax);
      PF_HIP(hipGetLastError());
      const int reps = (kmax + 1 + gsteps - 1) / gsteps;
      for (int r = 0; r < reps; ++r) PF_HIP(hipGraphLaunch(exec[mask], st));
    } else {
      // active at step k: launch positions with K >= k, a prefix that grows
      // as k falls (K is non-increasing in launch order)
      int nact = 0;
      for (int k = kmax; k >= 0; --k) {
        while (nact < n && K[nact] >= k) ++nact;
        launch(A, blk0[nact], k, 0, st, mask);
      }
    }
  }
  ~pf_batch() {
    (void)hipDeviceSynchronize();
    drop_graph();
    if (cap) (void)hipStreamDestroy(cap);
    if (hdyn) (void)hipHostFree(hdyn);
    if (dyn_done) (void)hipEventDestroy(dyn_done);
    insts.release();
    dyn.release();
    blk0d.release();
    coefs.release();
    kstate.release();
    bufs.release();
  }
};

extern "C" int pf_batch_create(pf_handle* h, int32_t n, pf_op* const* ops, int32_t ncols,
                               pf_batch** out) {
  return guarded([&] {
    PF_REQUIRE(h && ops && out, "batch: null argument");
    PF_REQUIRE(n >= 1 && n <= 4096, "batch: 1 <= n <= 4096");
    PF_REQUIRE(ncols >= 1 && ncols <= 32, "batch: 1 <= ncols <= 32");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
      PF_REQUIRE(ops[i], "batch: null operator");
      PF_REQUIRE(ops[i]->kind == kQsun, "batch: only Quantum-Sun operators can be batched");
      PF_REQUIRE(ops[i]->h == h, "batch: operators must belong to this handle");
    }
    PF_HIP(hipSetDevice(h->device));
    pf_batch* b = new pf_batch();
    try {
      b->h = h;
      b->n = n;
      b->s = ncols;
      b->ops.assign(ops, ops + n);
      b->order.resize(n);
      std::iota(b->order.begin(), b->order.end(), 0);
      b->pos = b->order;
      b->K.assign(n, -1);
      b->V = ncols % 2 == 0 ? 2 : 1;
      b->upload_insts();
      b->layout_blocks();
      b->kstate.get(256);
      b->dyn.get(n * sizeof(BatchDyn));
      PF_REQUIRE((uint64_t)b->ops[0]->D * ncols * 8 < ((uint64_t)1 << 32) &&
                     std::all_of(b->ops.begin(), b->ops.end(),
                                 [&](const pf_op* o) {
                                   return (uint64_t)o->D * ncols * 8 < ((uint64_t)1 << 32);
                                 }),
                 "batch: an instance block exceeds 32-bit byte offsets");
      b->bufs.get(4 * (size_t)b->rows_total * ncols * 8);
    } catch (...) {
      delete b;
      throw;
    }
    *out = b;
  });
}

extern "C" int pf_batch_destroy(pf_batch* b) {
  return guarded([&] { delete b; });
}

extern "C" int pf_batch_set_filter(pf_batch* b, int32_t polytype, const int32_t* orders,
                                   const double* coeffs) {
  return guarded([&] {
    PF_REQUIRE(b && orders && coeffs, "batch: null argument");
    PF_REQUIRE(polytype >= PF_CHEBYSHEV && polytype <= PF_TAYLOR, "batch: polynomial type");
    const int n = b->n;
    std::vector<int64_t> off_inst(n);
    int64_t tot = 0;
    for (int i = 0; i < n; ++i) {
      PF_REQUIRE(orders[i] >= 0 && orders[i] < (1 << 30), "batch: 0 <= order < 2^30");
      off_inst[i] = tot;
      tot += orders[i] + 1;
    }
    // launch order: decreasing order, ties by instance index
    std::vector<int> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return orders[x] > orders[y]; });
    b->order = ord;
    for (int p = 0; p < n; ++p) b->pos[ord[p]] = p;
    b->K.resize(n);
    b->coef_off.resize(n);
    for (int p = 0; p < n; ++p) {
      b->K[p] = orders[ord[p]];
      b->coef_off[p] = off_inst[ord[p]];
    }
    b->polytype = polytype;
    void* d = b->coefs.get(tot * sizeof(double));
    PF_HIP(hipMemcpy(d, coeffs, tot * sizeof(double), hipMemcpyHostToDevice));
    b->upload_insts();   // row0 and coefficient offsets follow the launch order
    b->layout_blocks();  // drops the graph
    b->have_filter = true;
  });
}

extern "C" int pf_batch_filter(pf_batch* b, const int32_t* active, const int32_t* dtypes,
                               const double* const* X, const int64_t* ldx, double* const* Y,
                               const int64_t* ldy, pf_stream stream) {
  return guarded([&] {
    PF_REQUIRE(b && active && dtypes && X && ldx && Y && ldy, "batch: null argument");
    PF_REQUIRE(b->have_filter, "batch: pf_batch_set_filter first");
    hipStream_t st = S(stream);
    const int n = b->n;
    std::vector<BatchDyn> dy(n);
    int kmax = -1;
    int mask = 0;
    for (int i = 0; i < n; ++i) {
      const int p = b->pos[i];
      BatchDyn& d = dy[p];
      d.K = -1;
      d.dtype = PF_F64;
      d.Y = nullptr;
      d.ldy = 0;
      if (!active[i]) continue;
      const int64_t D = b->ops[i]->D;
      PF_REQUIRE(X[i] && Y[i], "batch: null block of an active instance");
      PF_REQUIRE(ldx[i] >= D && ldy[i] >= D, "batch: leading dimension < D");
      PF_REQUIRE(dtypes[i] == PF_F64 || dtypes[i] == PF_F32, "batch: dtype");
      d.K = b->K[p];
      d.dtype = dtypes[i];
      d.Y = Y[i];
      d.ldy = ldy[i];
      kmax = std::max(kmax, d.K);
      mask |= dtypes[i] == PF_F64 ? kMaskF64 : kMaskF32;
    }
    if (kmax < 0) return;
    b->upload_dyn(dy, st);
    char* base = reinterpret_cast<char*>(b->bufs.p);
    for (int i = 0; i < n; ++i) {
      if (!active[i]) continue;
      const int p = b->pos[i];
      pack_block(dtypes[i], X[i], ldx[i], base + b->row0[p] * (int64_t)b->s * 8, b->ops[i]->D,
                 b->s, nullptr, st);
    }
    b->run_steps(kmax, st, mask);
  });
}

extern "C" int pf_batch_set_tuning(pf_batch* b, int32_t key, int32_t value) {
  return guarded([&] {
    PF_REQUIRE(b, "batch: null argument");
    switch (key) {
      case PF_BATCH_GRAPH:
        PF_REQUIRE(value == 0 || value == 1, "batch: graph 0/1");
        b->use_graph = value;
        break;
      case PF_BATCH_GRAPH_STEPS:
        PF_REQUIRE(value >= 1 && value <= 4096, "batch: 1 <= graph steps <= 4096");
        b->gsteps = value;
        b->drop_graph();
        break;
      case PF_BATCH_VEC:
        PF_REQUIRE(batch_vec_ok(b->s, value), "batch: vector width must divide ncols (1/2/4/8)");
        b->V = value;
        b->layout_blocks();
        break;
      case PF_BATCH_ROWS_PER_LANE:
        PF_REQUIRE(value >= 1 && value <= 64, "batch: 1 <= rows per lane <= 64");
        b->rpt = value;
        b->layout_blocks();
        break;
      case PF_BATCH_XCD:
        PF_REQUIRE(value == 0 || value == 1, "batch: xcd 0/1");
        b->xcd = value;
        b->drop_graph();
        break;
      case PF_BATCH_GRAIN_SHUFFLE:
        PF_REQUIRE(value == 0 || value == 1, "batch: grain shuffle 0/1");
        b->shfl = value;
        b->drop_graph();
        break;
      case PF_BATCH_LDS_ROWS:
        PF_REQUIRE(value == 0 || value == 1, "batch: LDS rows 0/1");
        b->lds = value;
        b->drop_graph();
        break;
      case PF_BATCH_ROW_VECTORS:
        PF_REQUIRE(value >= 0 && value <= 2, "batch: row vectors 0/1/2");
        b->rv = value;
        b->layout_blocks();
        break;
      default: throw Error{PF_ERR_INVALID, "batch: unknown tuning key"};
    }
  });
}

extern "C" int pf_batch_get_tuning(const pf_batch* b, int32_t* cfg) {
  return guarded([&] {
    PF_REQUIRE(b && cfg, "batch: null argument");
    const int32_t v[10] = {b->V,   b->rpt,       b->xcd,       b->shfl,   b->lds,
                           b->use_lds(), b->use_graph, b->gsteps, b->rv, b->use_rv()};
    for (int i = 0; i < 10; ++i) cfg[i] = v[i];
  });
}

extern "C" int pf_batch_autotune(pf_batch* b, int32_t dtype, int32_t nsteps, double* best_us,
                                 pf_stream stream) {
  return guarded([&] {
    PF_REQUIRE(b, "batch: null argument");
    PF_REQUIRE(dtype == PF_F64 || dtype == PF_F32, "batch: dtype");
    PF_REQUIRE(nsteps >= 3 && nsteps <= 100000, "batch: 3 <= nsteps <= 100000");
    hipStream_t st = S(stream);
    const int n = b->n;
    const int tmask = dtype == PF_F64 ? kMaskF64 : kMaskF32;
    // timing mode: every instance active at every step, constant coefficients
    std::vector<BatchDyn> dy(n);
    for (auto& d : dy) {
      d.K = 1 << 29;
      d.dtype = dtype;
      d.Y = nullptr;
      d.ldy = 0;
    }
    b->upload_dyn(dy, st);
    PF_HIP(hipMemsetAsync(b->bufs.p, 0, b->bufs.bytes, st));
    hipEvent_t e0, e1;
    PF_HIP(hipEventCreate(&e0));
    PF_HIP(hipEventCreate(&e1));
    struct Cand {
      int V, rpt, xcd, shfl, lds, rv;
      double us;
    };
    std::vector<Cand> cands;
    for (int V : {1, 2, 4, 8})
      if (batch_vec_ok(b->s, V) && !(dtype == PF_F64 && V > 8)) {
        for (int rpt : {1, 4, 16, 64})
          for (int x : {1, 0})
            for (int sh : {1, 0}) cands.push_back({V, rpt, x, sh, 0, 0, 1e300});
        // LDS form (the grain block comes from LDS: no shuffle choice)
        for (int rpt : {1, 2, 4, 8, 16}) {
          const int saveV0 = b->V, saveR0 = b->rpt;
          b->V = V;
          b->rpt = rpt;
          const bool ok = b->lds_ok();
          b->V = saveV0;
          b->rpt = saveR0;
          if (ok)
            for (int x : {1, 0}) cands.push_back({V, rpt, x, 1, 1, 0, 1e300});
        }
      }
    if (b->s == 1)  // row-vector forms
      for (int rv : {1, 2})
        for (int rpt : {1, 2, 4})
          if (!(rv == 2 && rpt == 1))
            for (int x : {1, 0}) cands.push_back({1, rpt, x, 1, 0, rv, 1e300});
    const int saveV = b->V, saveR = b->rpt, saveX = b->xcd, saveS = b->shfl, saveL = b->lds,
              saveRV = b->rv;
    const bool had_filter = b->have_filter;
    for (int round = 0; round < 3; ++round) {
      for (auto& c : cands) {
        b->V = c.V;
        b->rpt = c.rpt;
        b->xcd = c.xcd;
        b->shfl = c.shfl;
        b->lds = c.lds;
        b->rv = c.rv;
        b->layout_blocks();
        BatchArgs A = b->args();
        A.coefs = nullptr;
        const int64_t grid = b->blk0[n];
        const int kbig = 1 << 28;
        for (int w = 0; w < 2; ++w) b->launch(A, grid, kbig, 0, st, tmask);
        PF_HIP(hipEventRecord(e0, st));
        for (int k = 0; k < nsteps; ++k) b->launch(A, grid, kbig - k, 0, st, tmask);
        PF_HIP(hipEventRecord(e1, st));
        PF_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        PF_HIP(hipEventElapsedTime(&ms, e0, e1));
        c.us = std::min(c.us, 1e3 * ms / nsteps);
      }
    }
    double best = 1e300;
    Cand bc{saveV, saveR, saveX, saveS, saveL, saveRV, 1e300};
    for (const auto& c : cands)
      if (c.us < 0.99 * best) {
        best = c.us;
        bc = c;
      }
    b->V = bc.V;
    b->rpt = bc.rpt;
    b->xcd = bc.xcd;
    b->shfl = bc.shfl;
    b->lds = bc.lds;
    b->rv = bc.rv;
    b->layout_blocks();
    // graph replays against stream launches for the chosen tiling
    {
      const int saveG = b->gsteps;
      b->gsteps = std::min(saveG, nsteps);
      b->drop_graph();
      b->ensure_graph(tmask, true);
      double g_us = 1e300, s_us = 1e300;
      for (int round = 0; round < 3; ++round) {
        hipLaunchKernelGGL(batch_kset_kernel, dim3(1), dim3(64), 0, st,
                           reinterpret_cast<int*>(b->kstate.p), 1 << 28);
        PF_HIP(hipGetLastError());
        PF_HIP(hipGraphLaunch(b->exec[tmask], st));
        PF_HIP(hipEventRecord(e0, st));
        for (int r = 0; r < 2; ++r) PF_HIP(hipGraphLaunch(b->exec[tmask], st));
        PF_HIP(hipEventRecord(e1, st));
        PF_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        PF_HIP(hipEventElapsedTime(&ms, e0, e1));
        g_us = std::min(g_us, 1e3 * ms / (2 * b->gsteps));
        BatchArgs A = b->args();
        A.coefs = nullptr;
        PF_HIP(hipEventRecord(e0, st));
        for (int k = 0; k < 2 * b->gsteps; ++k)
          b->launch(A, b->blk0[n], (1 << 27) - k, 0, st, tmask);
        PF_HIP(hipEventRecord(e1, st));
        PF_HIP(hipEventSynchronize(e1));
        PF_HIP(hipEventElapsedTime(&ms, e0, e1));
        s_us = std::min(s_us, 1e3 * ms / (2 * b->gsteps));
      }
      b->gsteps = saveG;
      b->drop_graph();
      b->use_graph = g_us < s_us ? 1 : 0;
      best = std::min(g_us, s_us);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->have_filter = had_filter;
    if (best_us) *best_us = best;
  });
}
