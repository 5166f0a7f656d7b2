// Matrix-free Hamiltonian kernels, fused Clenshaw steps, filter driver and KPM.
//
// Layouts.  The C ABI takes column-major D x s blocks (Julia's layout).  The
// filter and KPM drivers repack their input once into a row-interleaved block
// (the s values of a row contiguous), run every recurrence step on that layout
// -- a gathered neighbour row is then one contiguous s*sizeof(T) segment loaded
// by G = s/V lanes with V-wide vector loads -- and write the result back
// column-major in the last (fused) step.  The hook entry points
// (pf_apply, pf_clenshaw_*) work directly on the caller's column-major blocks.
//
// Per row the spin-chain kernel reads the basis state (4 B) and the diagonal
// (8 B), then for every antiparallel bond gathers the neighbour row at the
// combinadic rank r + delta (pf_rank.h); no index table is stored.  The Quantum
// Sun kernel needs no table at all: its full-space basis makes every neighbour
// an XOR of the row index.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "pf_device.h"
#include "pf_internal.h"
#include "pf_qsun.h"
#include "pf_step.h"

namespace pf {

// ------------------------------------------------------ spin-chain kernel
// Where a lane finds neighbour ranks: the two rank tables staged in LDS
// (TBL) or the O(1) update from the binomial table; bond masks in LDS.
struct RankCtx {
  const uint32_t* C;   // binomials C(p, t) (!TBL)
  const uint16_t* TA;  // rank of the low m bits (TBL)
  const uint32_t* TB;  // segment offset of the high L-m bits (TBL)
  const uint32_t* M;   // bond masks (1 << lo) | (1 << hi)
  uint32_t maskA;
  int m;
};

// Rank of the neighbour across bond b, or r when its two spins are parallel
// (also for the zero padding bonds, mask 0).  With tables the lookups are made
// for every bond -- any L-bit value indexes both tables in bounds -- and the
// result selected afterwards, so a batch issues all its LDS reads before one
// wait and takes no branch.
template <bool TBL>
__device__ __forceinline__ uint32_t one_rank(const RankCtx& R, int b, uint32_t st, uint32_t r) {
  const uint32_t mk = R.M[b];
  if constexpr (TBL) {
    const uint32_t x = st ^ mk;
    const uint32_t nr = (uint32_t)R.TA[x & R.maskA] + R.TB[x >> R.m];
    return popc32(st & mk) == 1 ? nr : r;
  } else {
    if (popc32(st & mk) != 1) return r;
    return neighbor_rank(st, r, ctz32(mk), 31 - __clz(mk), R.C);
  }
}

// Ranks of the neighbours of bonds b0 + j, j < NBAT, as seen by every lane of
// the row group: with GM in {1,2,4,8} lane g computes the bonds j = g + GM*u
// and the group shares them; GM = 0 (any other group size) computes all NBAT.
// The bond list is padded with zero bonds to whole batches.
template <int GM, int NBAT, bool TBL>
__device__ __forceinline__ void bond_ranks(const RankCtx& R, int b0, uint32_t st, uint32_t r,
                                           int g, uint32_t (&out)[NBAT]) {
  if constexpr (GM == 0) {
#pragma unroll
    for (int j = 0; j < NBAT; ++j) out[j] = one_rank<TBL>(R, b0 + j, st, r);
  } else {
    constexpr int UB = NBAT / GM;
    uint32_t mine[UB];
#pragma unroll
    for (int u = 0; u < UB; ++u) mine[u] = one_rank<TBL>(R, b0 + g + GM * u, st, r);
    group_share<GM, UB>(mine, out);
  }
}

// Byte offset of row r (row stride rsb bytes, rsh = log2(rsb) or -1).  With
// rank tables D < 2^24 (L <= 26), so one full-rate 24-bit multiply serves
// every stride; otherwise a shift or a 32-bit multiply.
template <bool TBL>
__device__ __forceinline__ uint32_t row_bytes(uint32_t r, uint32_t rsb, int rsh) {
  if constexpr (TBL) return __umul24(r, rsb);
  else return rsh >= 0 ? (r << rsh) : r * rsb;
}

// out = cx*X + ch*((H - shift)*scale) b2 + cb*b3 on rows of the Lup sector.
// The G = s/V lanes of a row split the neighbour-rank work (GM = G) and each
// gathers its V columns of every flippable neighbour; NBAT neighbours are
// loaded before any is accumulated.  Block p (XCD-grouped) owns the
// contiguous row groups [p*n/nblk, (p+1)*n/nblk): with a persistent grid the
// rank tables are staged in LDS once per block, not once per row group.  A
// parallel-spin bond's gather is skipped (masked lanes); gathering a zero row
// instead removes the masking instructions but costs more memory-pipeline
// work than it saves.
template <typename T, int V, int GM, bool TBL, int NBW>
__device__ __forceinline__ void chain_step_body(
    const ChainParams& P, const uint32_t* __restrict__ states, const uint32_t* __restrict__ binom,
    const uint4* __restrict__ rtab, const int rtab_vec, const double* __restrict__ diag,
    const Affine af, const int flags, const T cx, const T ch, const T cb,
    const T* __restrict__ X, const T* __restrict__ B2, const T* __restrict__ B3,
    void* __restrict__ OUT, const Layout& lay, const int64_t out_ld,
    const uint32_t* __restrict__ operm, double* __restrict__ dot_part, const T* __restrict__ Pr,
    const int64_t ngroups) {
  typedef typename VT<T, V>::t vt;
  constexpr int GE = GM < 1 ? 1 : GM;
  constexpr int NB0 = ((V * (int)sizeof(T) <= 32) ? 8 : 4) * NBW;  // gathers in flight
  constexpr int NBAT = NB0 < GE ? GE : NB0;
  extern __shared__ uint4 tab_lds[];
  __shared__ uint32_t sM[kMaxBonds];
  __shared__ T sW[kMaxBonds];
  __shared__ double red[4];
  RankCtx R{};
  R.M = sM;
  if constexpr (TBL) {
    for (int i = threadIdx.x; i < rtab_vec; i += blockDim.x) tab_lds[i] = rtab[i];
    R.TA = reinterpret_cast<const uint16_t*>(tab_lds);
    R.TB = reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(tab_lds) +
                                             (2u << P.tsplit));
    R.maskA = (1u << P.tsplit) - 1u;
    R.m = P.tsplit;
  } else {
    __shared__ uint32_t sC[32 * kRankT];
    for (int i = threadIdx.x; i < 32 * kRankT; i += blockDim.x) sC[i] = binom[i];
    R.C = sC;
  }
  // bonds padded with zero bonds (mask 0, weight 0) to whole batches of NBAT
  const int nbp = (P.nb + NBAT - 1) / NBAT * NBAT;  // <= kMaxBonds (a multiple of NBAT)
  for (int i = threadIdx.x; i < nbp; i += blockDim.x) {
    const bool real = i < P.nb;
    sW[i] = real ? (T)(P.bval[i] * af.scale) : (T)0;
    sM[i] = real ? (1u << P.lo[i]) | (1u << P.hi[i]) : 0u;
  }
  __syncthreads();

  const int rloc = threadIdx.x / lay.G;
  const int g = threadIdx.x - rloc * lay.G;
  const bool lane_ok = rloc < lay.rows_per_block;
  const int64_t cofs = (int64_t)blockIdx.y * lay.cstride + (int64_t)g * V;
  const int out_col0 = (int)blockIdx.y + g * V;
  // block-uniform column offset in the 64-bit base, the lane's own in 32 bits
  const char* b2b = reinterpret_cast<const char*>(B2 + (int64_t)blockIdx.y * lay.cstride);
  const uint32_t cofs_b = (uint32_t)((int64_t)g * V * sizeof(T));
  const uint32_t rsb = (uint32_t)(lay.rs * sizeof(T));
  const int rsh = (rsb & (rsb - 1u)) == 0 ? __ffs(rsb) - 1 : -1;
  double dacc = 0.0;

  const int64_t nblk = gridDim.x;
  const int64_t p = lay.xcd ? xcd_block(blockIdx.x, nblk) : (int64_t)blockIdx.x;
  const int64_t g_end = (p + 1) * ngroups / nblk;  // balanced contiguous ranges
  const bool nt = flags & kStreamNT;
  const bool use_h = flags & kHasH;
  // the state and diagonal of the next row group are loaded one iteration
  // ahead, so their latency overlaps this group's gathers
  auto row_of = [&](int64_t grp) { return grp * lay.rows_per_block + rloc; };
  auto row_valid = [&](int64_t grp) { return lane_ok && grp < g_end && row_of(grp) < P.D; };
  uint32_t st_nx = 0;
  double d_nx = 0.0;
  int64_t grp = p * ngroups / nblk;
  if (use_h && row_valid(grp)) {
    const int64_t r0 = row_of(grp);
    st_nx = nt ? __builtin_nontemporal_load(states + r0) : states[r0];
    d_nx = nt ? __builtin_nontemporal_load(diag + r0) : diag[r0];
  }
  for (; grp < g_end; ++grp) {
    const int64_t row = row_of(grp);
    const uint32_t st = st_nx;
    const double drow = d_nx;
    if (use_h && row_valid(grp + 1)) {
      const int64_t r1 = row_of(grp + 1);
      st_nx = nt ? __builtin_nontemporal_load(states + r1) : states[r1];
      d_nx = nt ? __builtin_nontemporal_load(diag + r1) : diag[r1];
    }
    if (!lane_ok || row >= P.D) continue;  // whole row groups drop out together
    const int64_t off = row * lay.rs + cofs;
    vt xv, b3v;
    step_load<T, V>(flags, off, X, B3, xv, b3v);
    vt h = (vt)(T)0;
    if (use_h) {
      const uint32_t r32 = (uint32_t)row;
      vt tot = (vt)(T)0;
      for (int b0 = 0; b0 < nbp; b0 += NBAT) {
        uint32_t rk[NBAT];
        bond_ranks<GM, NBAT, TBL>(R, b0, st, r32, g, rk);
        vt x[NBAT];
#pragma unroll
        for (int j = 0; j < NBAT; ++j) {
          x[j] = (vt)(T)0;
          // 32-bit byte offset from the block base (D*rs*sizeof(T) < 2^32 is
          // checked on the host)
          const uint32_t o = row_bytes<TBL>(rk[j], rsb, rsh) + cofs_b;
          if (rk[j] != r32) x[j] = *reinterpret_cast<const vt*>(b2b + o);
        }
#pragma unroll
        for (int j = 0; j < NBAT; ++j) tot += sW[b0 + j] * x[j];
      }
      const T d = (T)((drow - af.shift) * af.scale);
      const uint32_t o = row_bytes<TBL>(r32, rsb, rsh) + cofs_b;
      h = d * *reinterpret_cast<const vt*>(b2b + o) + tot;
    }
    step_store<T, V>(flags, row, off, out_col0, cx, ch, cb, h, xv, b3v, OUT, out_ld, operm, Pr,
                     dacc);
  }
  if (flags & kDot) {
    const double t = block_sum_256(dacc, red);
    if (threadIdx.x == 0) dot_part[blockIdx.x] = t;
  }
}

#define PF_CHAIN_ARGS                                                                             \
  const ChainParams P, const uint32_t* __restrict__ states, const uint32_t* __restrict__ binom,  \
      const uint4* __restrict__ rtab, const int rtab_vec, const double* __restrict__ diag,       \
      const Affine af, const int flags, const T cx, const T ch, const T cb,                     \
      const T* __restrict__ X, const T* __restrict__ B2, const T* __restrict__ B3,              \
      void* __restrict__ OUT, const Layout lay, const int64_t out_ld,                           \
      const uint32_t* __restrict__ operm, double* __restrict__ dot_part,                        \
      const T* __restrict__ Pr, const int64_t ngroups
#define PF_CHAIN_PASS                                                                       \
  P, states, binom, rtab, rtab_vec, diag, af, flags, cx, ch, cb, X, B2, B3, OUT, lay, out_ld, \
      operm, dot_part, Pr, ngroups

template <typename T, int V, int GM, bool TBL, int NBW>
__global__ __launch_bounds__(256) void chain_step_kernel(PF_CHAIN_ARGS) {
  chain_step_body<T, V, GM, TBL, NBW>(PF_CHAIN_PASS);
}

// The same step held to 8 waves per SIMD (at most 64 VGPRs): more rows in
// flight per CU to hide the neighbour-gather latency, at the price of fewer
// gathers in flight per wave.
template <typename T, int V, int GM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void
chain_step_kernel_occ(PF_CHAIN_ARGS) {
  chain_step_body<T, V, GM, true, 1>(PF_CHAIN_PASS);
}
#undef PF_CHAIN_ARGS
#undef PF_CHAIN_PASS

// ------------------------------------------------------ Quantum Sun kernel
template <typename T, int V>
__global__ __launch_bounds__(256) void qsun_step_kernel(
    const QsunParams P, const double* __restrict__ qdiag, const int qlb, const Affine af,
    const int flags, const T cx, const T ch, const T cb, const T* __restrict__ X,
    const T* __restrict__ B2, const T* __restrict__ B3, void* __restrict__ OUT, const Layout lay,
    const int64_t out_ld, double* __restrict__ dot_part, const T* __restrict__ Pr,
    const int64_t ngroups) {
  typedef typename VT<T, V>::t vt;
  const int dg = 1 << P.Lg;
  __shared__ T sG[kMaxDg * kMaxDg];
  __shared__ T sC[kMaxLocPad];
  __shared__ uint32_t sF[kMaxLocPad];
  __shared__ double red[4];
  const uint32_t rsb = (uint32_t)(lay.rs * sizeof(T));
  const int rsh = (rsb & (rsb - 1u)) == 0 ? __ffs(rsb) - 1 : -1;
  const int npad = qsun_pad(P.Lloc, qsun_flip_batch<T, V>());
  qsun_stage<T>(P.Lloc, npad, dg, P.flip, P.coup, P.G, af.scale, rsh, sF, sC, sG);
  __syncthreads();

  const int rloc = threadIdx.x / lay.G;
  const int g = threadIdx.x - rloc * lay.G;
  const bool lane_ok = rloc < lay.rows_per_block;
  const int64_t cofs = (int64_t)blockIdx.y * lay.cstride + (int64_t)g * V;
  const int out_col0 = (int)blockIdx.y + g * V;
  QsunRow q;
  // block-uniform column offset in the 64-bit base, the lane's own in 32 bits
  q.b2 = reinterpret_cast<const char*>(B2 + (int64_t)blockIdx.y * lay.cstride);
  q.cofs_b = (uint32_t)((int64_t)g * V * sizeof(T));
  q.rsb = rsb;
  q.rsh = rsh;
  q.dg = dg;
  q.nloc = npad;
  q.lb = qlb;
  q.dlo = qdiag;
  q.dhi = qdiag + ((int64_t)1 << qlb);
  q.scale = af.scale;
  q.shift = af.shift;
  q.sF = sF;
  q.sC = sC;
  const int G_ = lay.shfl ? lay.G : 0;  // 0 selects the reload path
  double dacc = 0.0;

  const int64_t nblk = gridDim.x;
  const int64_t p = lay.xcd ? xcd_block(blockIdx.x, nblk) : (int64_t)blockIdx.x;
  const int64_t g_end = (p + 1) * ngroups / nblk;
  for (int64_t grp = p * ngroups / nblk; grp < g_end; ++grp) {
    const int64_t row = grp * lay.rows_per_block + rloc;
    if (!lane_ok || row >= P.D) continue;
    const int64_t off = row * lay.rs + cofs;
    vt xv, b3v;
    step_load<T, V>(flags, off, X, B3, xv, b3v);
    vt h = (vt)(T)0;
    if (flags & kHasH) h = qsun_hrow_dg<T, V>(q, (uint32_t)row, sG, G_);
    step_store<T, V>(flags, row, off, out_col0, cx, ch, cb, h, xv, b3v, OUT, out_ld, nullptr, Pr,
                     dacc);
  }
  if (flags & kDot) {
    const double t = block_sum_256(dacc, red);
    if (threadIdx.x == 0) dot_part[blockIdx.x] = t;
  }
}

// ------------------------------------------------------------ launching
static bool valid_vec(int dtype, int s, int V) {
  if (V < 1 || s % V != 0) return false;
  if (dtype == PF_F64 && !(V == 1 || V == 2 || V == 4 || V == 8 || V == 16)) return false;
  if (dtype == PF_F32 && !(V == 1 || V == 2 || V == 4 || V == 8 || V == 16 || V == 32))
    return false;
  return s / V <= 256;
}

Layout interleaved_layout(int dtype, int s, int V, int xcd, int nt, int pb, int wide, int shfl,
                          int occ, int part) {
  Layout l{};
  if (V == 0 || !valid_vec(dtype, s, V)) {
    if (dtype == PF_F64)
      V = (s % 2 == 0) ? 2 : 1;
    else
      V = (s % 4 == 0) ? 4 : (s % 2 == 0 ? 2 : 1);
  }
  l.V = V;
  l.G = s / V;
  l.rs = s;
  l.cstride = 0;
  l.rows_per_block = 256 / l.G;
  l.xcd = xcd;
  l.nt = nt;
  l.pblocks = pb;
  l.wide = wide;
  l.shfl = shfl;
  l.occ = occ;
  l.part = part;
  return l;
}

// Default persistent grid: spin chains with LDS rank tables stage them once
// per block, so they run 8 blocks per CU over all row groups.
static int default_pblocks(const pf_op* op) {
  return (op->kind == kChain && op->chain.tsplit > 0) ? 8 : 0;
}

// Configuration for (dtype, s): forced knobs, else the autotuned entry, else defaults.
Layout tuned_layout(pf_op* op, int dtype, int s) {
  TuneCfg c;
  auto it = op->tuned.find({dtype, s});
  if (it != op->tuned.end()) c = it->second;
  if (op->force.V) c.V = op->force.V;
  if (op->force.xcd >= 0) c.xcd = op->force.xcd;
  if (op->force.nt >= 0) c.nt = op->force.nt;
  if (op->force.pb >= 0) c.pb = op->force.pb;
  if (op->force.wide >= 0) c.wide = op->force.wide;
  if (op->force.shfl >= 0) c.shfl = op->force.shfl;
  if (op->force.occ >= 0) c.occ = op->force.occ;
  if (op->force.part >= 0) c.part = op->force.part;
  return interleaved_layout(dtype, s, c.V, c.xcd < 0 ? 1 : c.xcd, c.nt < 0 ? 1 : c.nt,
                            c.pb < 0 ? default_pblocks(op) : c.pb, c.wide < 0 ? 0 : c.wide,
                            c.shfl < 0 ? 1 : c.shfl, c.occ < 0 ? 0 : c.occ,
                            c.part < 0 ? 0 : c.part);
}

Layout colmajor_layout(int64_t ld) {
  Layout l{};
  l.V = 1;
  l.G = 1;
  l.rs = 1;
  l.cstride = ld;
  l.rows_per_block = 256;
  l.xcd = 1;
  l.nt = 0;
  return l;
}

// Blocks of one step launch: one per row group, capped by the persistent grid
// (pblocks per CU) and, for KPM dot launches, by the number of partial slots.
int64_t step_grid_x(const pf_op* op, const Layout& lay, int flags, int dot_blocks) {
  const int64_t ngroups = (op->D + lay.rows_per_block - 1) / lay.rows_per_block;
  int64_t gx = ngroups;
  if (lay.pblocks > 0)
    gx = std::min<int64_t>(gx, (int64_t)lay.pblocks * std::max(1, op->h->prop.multiProcessorCount));
  if (flags & kDot) gx = std::min<int64_t>(gx, dot_blocks);
  return gx;
}

template <typename T, int V>
static void launch_typed(pf_op* op, int flags, const Layout& lay, int ncols_grid, Affine af,
                         double cx, double ch, double cb, const void* X, const void* B2,
                         const void* B3, void* out, int64_t out_ld, hipStream_t st,
                         double* dot_part, const void* P, int dot_blocks) {
  PF_REQUIRE((uint64_t)op->D * (uint64_t)lay.rs * sizeof(T) < ((uint64_t)1 << 32),
             "block too large for 32-bit byte offsets");
  const int64_t ngroups = (op->D + lay.rows_per_block - 1) / lay.rows_per_block;
  const int64_t gx = step_grid_x(op, lay, flags, dot_blocks);
  PF_REQUIRE(gx < (int64_t)1 << 31, "grid too large");
  dim3 grid((unsigned)gx, (unsigned)ncols_grid), block(256);
  if (lay.nt) flags |= kStreamNT;
  const T* Xt = reinterpret_cast<const T*>(X);
  const T* B2t = reinterpret_cast<const T*>(B2);
  const T* B3t = reinterpret_cast<const T*>(B3);
  const T* Pt = reinterpret_cast<const T*>(P);
  if (op->kind == kChain) {
    const bool tbl = op->chain.tsplit > 0 && op->rtab;
    // drivers run in the partition order when it is built; the hooks never do
    const bool part = (flags & kRowPart) && op->pstates;
    const uint32_t* sts = part ? op->pstates : op->states;
    const double* dg = part ? op->pdiag : op->diag;
    const uint32_t* operm = part ? op->pperm : nullptr;
    const uint4* rt = reinterpret_cast<const uint4*>(part ? op->prtab : op->rtab);
    const int rvec = (int)(op->rtab_bytes / 16);
    const size_t lds = tbl ? op->rtab_bytes : 0;
#define PF_CHAIN_LAUNCH(GMV, TB, NW)                                                           \
  if (TB && NW == 1 && lay.occ)                                                                \
    hipLaunchKernelGGL((chain_step_kernel_occ<T, V, GMV>), grid, block, lds, st, op->chain,    \
                       sts, op->binom, rt, rvec, dg, af, flags, (T)cx, (T)ch, (T)cb, Xt, B2t,  \
                       B3t, out, lay, out_ld, operm, dot_part, Pt, ngroups);                   \
  else                                                                                         \
    hipLaunchKernelGGL((chain_step_kernel<T, V, GMV, TB, NW>), grid, block, lds, st, op->chain,  \
                       sts, op->binom, rt, rvec, dg, af, flags, (T)cx, (T)ch, (T)cb, Xt, B2t,  \
                       B3t, out, lay, out_ld, operm, dot_part, Pt, ngroups)
#define PF_CHAIN_G(TB, NW)                      \
  switch (lay.G) {                              \
    case 1: PF_CHAIN_LAUNCH(1, TB, NW); break;  \
    case 2: PF_CHAIN_LAUNCH(2, TB, NW); break;  \
    case 4: PF_CHAIN_LAUNCH(4, TB, NW); break;  \
    case 8: PF_CHAIN_LAUNCH(8, TB, NW); break;  \
    default: PF_CHAIN_LAUNCH(0, TB, NW); break; \
  }
    if (tbl && lay.wide) {
      PF_CHAIN_G(true, 2);
    } else if (tbl) {
      PF_CHAIN_G(true, 1);
    } else {
      PF_CHAIN_G(false, 1);
    }
#undef PF_CHAIN_G
#undef PF_CHAIN_LAUNCH
  } else {
    hipLaunchKernelGGL((qsun_step_kernel<T, V>), grid, block, 0, st, op->qsun, op->qdiag, op->qlb,
                       af, flags, (T)cx, (T)ch, (T)cb, Xt, B2t, B3t, out, lay, out_ld, dot_part, Pt,
                       ngroups);
  }
  PF_HIP(hipGetLastError());
}

void launch_step(pf_op* op, int dtype, int flags, const Layout& lay, int ncols, Affine af,
                 double cx, double ch, double cb, const void* X, const void* B2, const void* B3,
                 void* out, int64_t out_ld, hipStream_t st, double* dot_part, const void* P,
                 int dot_blocks) {
  const int ny = lay.cstride ? ncols : 1;
  if (op->kind == kPacked) {
    launch_packed(op, dtype, flags, lay, ny, af, cx, ch, cb, X, B2, B3, out, out_ld, st, dot_part,
                  P, dot_blocks);
    return;
  }
  if (dtype == PF_F64) {
    switch (lay.V) {
      case 1: launch_typed<double, 1>(op, flags, lay, ny, af, cx, ch, cb, X, B2, B3, out, out_ld, st, dot_part, P, dot_blocks); break;
      case 2: launch_typed<double, 2>(op, flags, lay, ny, af, cx, ch, cb, X, B2, B3, out, out_ld, st, dot_part, P, dot_blocks); break;
      case 4: launch_typed<double, 4>(op, flags, lay, ny, af, cx, ch, cb, X, B2, B3, out, out_ld, st, dot_part, P, dot_blocks); break;
      case 8: launch_typed<double, 8>(op, flags, lay, ny, af, cx, ch, cb, X, B2, B3, out, out_ld, st, dot_part, P, dot_blocks); break;
      case 16: launch_typed<double, 16>(op, flags, lay, ny, af, cx, ch, cb, X, B2, B3, out, out_ld, st, dot_part, P, dot_blocks); break;
      default: throw Error{PF_ERR_UNSUPPORTED, "unsupported vector width"};
    }
  } else if (dtype == PF_F32) {
    switch (lay.V) {
      case 1: launch_typed<float, 1>(op, flags, lay, ny, af, cx, ch, cb, X, B2, B3, out, out_ld, st, dot_part, P, dot_blocks); break;
      case 2: launch_typed<float, 2>(op, flags, lay, ny, af, cx, ch, cb, X, B2, B3, out, out_ld, st, dot_part, P, dot_blocks); break;
      case 4: launch_typed<float, 4>(op, flags, lay, ny, af, cx, ch, cb, X, B2, B3, out, out_ld, st, dot_part, P, dot_blocks); break;
      case 8: launch_typed<float, 8>(op, flags, lay, ny, af, cx, ch, cb, X, B2, B3, out, out_ld, st, dot_part, P, dot_blocks); break;
      case 16: launch_typed<float, 16>(op, flags, lay, ny, af, cx, ch, cb, X, B2, B3, out, out_ld, st, dot_part, P, dot_blocks); break;
      case 32: launch_typed<float, 32>(op, flags, lay, ny, af, cx, ch, cb, X, B2, B3, out, out_ld, st, dot_part, P, dot_blocks); break;
      default: throw Error{PF_ERR_UNSUPPORTED, "unsupported vector width"};
    }
  } else {
    throw Error{PF_ERR_INVALID, "dtype must be PF_F64 or PF_F32"};
  }
}

// ------------------------------------------------------------- pack kernel
// dst[r*s + c] = (T)(src[perm[r] + c*ld] * scale[c]) through a 64-row LDS
// tile (perm null: the identity).  A partition permutation keeps segments of
// consecutive reference ranks together, so the column reads stay in runs.
template <typename T>
__global__ __launch_bounds__(256) void pack_kernel(const double* __restrict__ src, int64_t ld,
                                                   T* __restrict__ dst, int64_t D, int s,
                                                   const double* __restrict__ scale,
                                                   const uint32_t* __restrict__ perm) {
  extern __shared__ double tile[];  // 64 x (s+1)
  const int64_t r0 = (int64_t)blockIdx.x * 64;
  const int n = 64 * s;
  for (int idx = threadIdx.x; idx < n; idx += blockDim.x) {
    const int c = idx / 64, rr = idx - c * 64;
    const int64_t r = r0 + rr;
    double v = 0.0;
    if (r < D) v = src[(perm ? (int64_t)perm[r] : r) + (int64_t)c * ld] * (scale ? scale[c] : 1.0);
    tile[rr * (s + 1) + c] = v;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < n; idx += blockDim.x) {
    const int rr = idx / s, c = idx - rr * s;
    const int64_t r = r0 + rr;
    if (r < D) dst[r * s + c] = (T)tile[rr * (s + 1) + c];
  }
}

void pack_block(int dtype, const double* src, int64_t ld, void* dst, int64_t D, int s,
                const double* scale, hipStream_t st, const uint32_t* perm) {
  PF_REQUIRE(s >= 1 && s <= 64, "pack: 1 <= s <= 64");
  const int64_t nb = (D + 63) / 64;
  const size_t lds = (size_t)64 * (s + 1) * sizeof(double);
  if (dtype == PF_F64)
    hipLaunchKernelGGL(pack_kernel<double>, dim3((unsigned)nb), dim3(256), lds, st, src, ld,
                       reinterpret_cast<double*>(dst), D, s, scale, perm);
  else
    hipLaunchKernelGGL(pack_kernel<float>, dim3((unsigned)nb), dim3(256), lds, st, src, ld,
                       reinterpret_cast<float*>(dst), D, s, scale, perm);
  PF_HIP(hipGetLastError());
}

// --------------------------------------------------------- setup kernels
// Diagonal of the spin-chain Hamiltonian in the reference summation order
// (fields for sites 1..L, then J*Delta*zi*zj per bond in input order), with
// FMA contraction disabled so the value is bit-identical to the reference's
// H[i, i] (src/Models/SpinHalfChains.jl:285-298).
__global__ void chain_diag_kernel(const uint32_t* __restrict__ states, int64_t D, int L, int nb,
                                  const double* __restrict__ fields,
                                  const int8_t* __restrict__ bi, const int8_t* __restrict__ bj,
                                  const double* __restrict__ jd, double* __restrict__ diag) {
#pragma clang fp contract(off)
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= D) return;
  const uint32_t st = states[r];
  double d = 0.0;
  for (int site = 0; site < L; ++site) {
    const double z = ((st >> site) & 1u) ? 0.5 : -0.5;
    const double t = fields[site] * z;
    d = d + t;
  }
  for (int b = 0; b < nb; ++b) {
    const double zi = ((st >> bi[b]) & 1u) ? 0.5 : -0.5;
    const double zj = ((st >> bj[b]) & 1u) ? 0.5 : -0.5;
    const double t = ((jd[b] * zi) * zj);
    d = d + t;
  }
  diag[r] = d;
}

// Connectivity export (parity surface).  Each row's neighbours come from the
// same rank function the step kernels call (one_rank<TBL>: the TA/TB tables
// when the operator has them, else the O(1) update), over the kernels' bond
// list (value-grouped, without the zero padding), and the diagonal from the
// table the step kernels stream.  Reference order: row = rank in the ascending basis.  In
// the XCD partition order (kRowPart drivers) the kernel runs over partition
// rows q with the partition tables and writes reference rank pperm[q]'s slot,
// its neighbours mapped back through pperm: that is the integer path those
// drivers apply.
template <bool TBL>
__global__ void chain_conn_kernel(const ChainParams P, const uint32_t* __restrict__ states,
                                  const uint32_t* __restrict__ binom, const void* __restrict__ rtab,
                                  const double* __restrict__ diag,
                                  const uint32_t* __restrict__ perm, int64_t row0, int64_t nrows,
                                  int64_t nscan, int maxper, int64_t* __restrict__ cols,
                                  double* __restrict__ vals, int32_t* __restrict__ counts,
                                  double* __restrict__ dout) {
  __shared__ uint32_t sM[kMaxBonds];
  for (int i = threadIdx.x; i < P.nb; i += blockDim.x) sM[i] = (1u << P.lo[i]) | (1u << P.hi[i]);
  __syncthreads();
  RankCtx R{};
  R.M = sM;
  if constexpr (TBL) {
    R.TA = reinterpret_cast<const uint16_t*>(rtab);
    R.TB = reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(rtab) +
                                             (2u << P.tsplit));
    R.maskA = (1u << P.tsplit) - 1u;
    R.m = P.tsplit;
  } else {
    R.C = binom;
  }
  const int64_t q = (perm ? 0 : row0) + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (perm ? nscan : row0 + nrows)) return;
  const int64_t ref = perm ? (int64_t)perm[q] : q;
  if (ref < row0 || ref >= row0 + nrows) return;
  const int64_t t = ref - row0;
  const uint32_t st = states[q];
  int cnt = 0;
  for (int b = 0; b < P.nb; ++b) {
    const uint32_t nr = one_rank<TBL>(R, b, st, (uint32_t)q);
    if (nr == (uint32_t)q) continue;  // parallel spins: no matrix element
    if (cnt < maxper) {
      cols[t * maxper + cnt] = perm ? (int64_t)perm[nr] : (int64_t)nr;
      vals[t * maxper + cnt] = P.bval[b];
    }
    ++cnt;
  }
  counts[t] = cnt;
  dout[t] = diag[q];
}

__global__ void qsun_conn_kernel(const QsunParams P, const double* __restrict__ qdiag, int qlb,
                                 int64_t row0, int64_t nrows, int maxper,
                                 int64_t* __restrict__ cols, double* __restrict__ vals,
                                 int32_t* __restrict__ counts, double* __restrict__ dout) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrows) return;
  const uint32_t r = (uint32_t)(row0 + t);
  const int dg = 1 << P.Lg;
  const int a = (int)(r & (uint32_t)(dg - 1));
  const uint32_t base = r & ~(uint32_t)(dg - 1);
  int cnt = 0;
  for (int j = 0; j < P.Lloc; ++j) {
    if (cnt < maxper) {
      cols[t * maxper + cnt] = (int64_t)(r ^ P.flip[j]);
      vals[t * maxper + cnt] = P.coup[j];
    }
    ++cnt;
  }
  for (int b = 0; b < dg; ++b) {
    if (b == a) continue;
    if (cnt < maxper) {
      cols[t * maxper + cnt] = (int64_t)(base | (uint32_t)b);
      vals[t * maxper + cnt] = P.G[a * dg + b];
    }
    ++cnt;
  }
  counts[t] = cnt;
  // the split diagonal the step kernels read: Dlo[r mod 2^lb] + Dhi[r >> lb]
  dout[t] = qdiag[r & ((1u << qlb) - 1u)] + qdiag[((size_t)1 << qlb) + (r >> qlb)];
}

// KPM / probe helpers ---------------------------------------------------
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__global__ void probes_kernel(int kind, uint64_t seed, int64_t D, int R, double* __restrict__ out,
                              int64_t ld) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= D) return;
  const uint64_t key = splitmix64(seed ^ splitmix64(((uint64_t)c << 40) ^ (uint64_t)r));
  double v;
  if (kind == PF_PROBE_PHASE) {
    v = (key >> 63) ? 1.0 : -1.0;
  } else {
    // Box-Muller on two 53-bit uniforms in (0, 1]
    const uint64_t k2 = splitmix64(key);
    const double u1 = ((double)(key >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    const double u2 = (double)(k2 >> 11) * (1.0 / 9007199254740992.0);
    v = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  }
  out[r + (int64_t)c * ld] = v;
}

// column sums of squares of a column-major block -> partials[c*nblk + blockIdx.x]
__global__ __launch_bounds__(256) void colnorm2_kernel(const double* __restrict__ A, int64_t ld,
                                                       int64_t D, double* __restrict__ part) {
  __shared__ double red[4];
  const int c = blockIdx.y;
  double acc = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < D;
       r += (int64_t)gridDim.x * blockDim.x) {
    const double v = A[r + (int64_t)c * ld];
    acc += v * v;
  }
  const double t = block_sum_256(acc, red);
  if (threadIdx.x == 0) part[(int64_t)c * gridDim.x + blockIdx.x] = t;
}

__global__ void inv_sqrt_sum_kernel(const double* __restrict__ part, int nblk, int ncols,
                                    double* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  double t = 0.0;
  for (int i = 0; i < nblk; ++i) t += part[(int64_t)c * nblk + i];
  out[c] = 1.0 / sqrt(t);
}

// sum over the blocks of each KPM step: mu[n] += sum_{i < count} part[n*stride + i]
__global__ void kpm_reduce_kernel(const double* __restrict__ part, int stride, int count, int N,
                                  double* __restrict__ mu) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  double t = 0.0;
  for (int i = 0; i < count; ++i) t += part[(int64_t)n * stride + i];
  mu[n] += t;
}

// Order of the high-bit segments (b = state >> m) for the XCD partition:
// reverse Cuthill-McKee over the graph that joins segment b to every segment
// its rows reach through one bond.  The drivers give each XCD a contiguous
// range of rows; in this order most neighbour rows of a range fall in the
// range itself (its own L2) instead of being spread over all XCDs as in the
// colex order.  Returns the valid segments (0 <= Lup - |b| <= m) in order.
static std::vector<uint32_t> rcm_segment_order(const ChainParams& P, int m) {
  const uint32_t nB = 1u << (P.L - m);
  std::vector<int> id(nB, -1);
  std::vector<uint32_t> seg;
  for (uint32_t b = 0; b < nB; ++b) {
    const int na = P.Lup - popc32(b);
    if (na >= 0 && na <= m) {
      id[b] = (int)seg.size();
      seg.push_back(b);
    }
  }
  const int n = (int)seg.size();
  std::vector<std::vector<int>> adj(n);
  for (int i = 0; i < n; ++i) {
    const uint32_t b = seg[i];
    for (int k = 0; k < P.nb; ++k) {
      const int lo = P.lo[k], hi = P.hi[k];
      if (hi < m) continue;  // both sites in the low bits: the segment stays
      uint32_t nbs;
      if (lo >= m) {
        const uint32_t mk = (1u << (lo - m)) | (1u << (hi - m));
        if ((b & mk) == 0 || (b & mk) == mk) continue;
        nbs = b ^ mk;
      } else {
        nbs = b ^ (1u << (hi - m));  // the low site flips with it
      }
      if (id[nbs] >= 0) adj[i].push_back(id[nbs]);
    }
    std::sort(adj[i].begin(), adj[i].end());
    adj[i].erase(std::unique(adj[i].begin(), adj[i].end()), adj[i].end());
  }
  std::vector<int> deg(n);
  for (int i = 0; i < n; ++i) deg[i] = (int)adj[i].size();
  for (auto& a : adj)  // Cuthill-McKee visits neighbours by ascending degree
    std::sort(a.begin(), a.end(),
              [&](int x, int y) { return deg[x] != deg[y] ? deg[x] < deg[y] : x < y; });
  std::vector<char> done(n, 0);
  std::vector<int> seen(n, 0), lvl(n, 0);
  int stamp = 0;
  // breadth-first order from s over the nodes not yet placed; returns the depth
  auto bfs = [&](int s, std::vector<int>& out) {
    ++stamp;
    out.clear();
    out.push_back(s);
    seen[s] = stamp;
    lvl[s] = 0;
    for (size_t h = 0; h < out.size(); ++h) {
      const int u = out[h];
      for (int v : adj[u])
        if (!done[v] && seen[v] != stamp) {
          seen[v] = stamp;
          lvl[v] = lvl[u] + 1;
          out.push_back(v);
        }
    }
    return lvl[out.back()];
  };
  std::vector<int> order, comp, trial;
  order.reserve(n);
  for (int s0 = 0; s0 < n; ++s0) {
    if (done[s0]) continue;
    // pseudo-peripheral start: move to a least-degree node of the deepest
    // level while that makes the breadth-first depth grow
    int s = s0;
    int depth = bfs(s, comp);
    for (int it = 0; it < 16; ++it) {
      const int last = lvl[comp.back()];
      int cand = comp.back();
      for (int i = (int)comp.size() - 1; i >= 0 && lvl[comp[i]] == last; --i)
        if (deg[comp[i]] < deg[cand]) cand = comp[i];
      const int d2 = bfs(cand, trial);
      if (d2 <= depth) break;
      depth = d2;
      s = cand;
      comp.swap(trial);
    }
    bfs(s, comp);
    for (int u : comp) {
      done[u] = 1;
      order.push_back(u);
    }
  }
  std::reverse(order.begin(), order.end());
  std::vector<uint32_t> out(n);
  for (int i = 0; i < n; ++i) out[i] = seg[order[i]];
  return out;
}

}  // namespace pf

using namespace pf;

// =========================================================== C ABI: ops
extern "C" int pf_op_create_spin_chain(pf_handle* h, int32_t L, int32_t Lup, int32_t nbonds,
                                       const int32_t* site_i, const int32_t* site_j,
                                       const double* J, const double* Delta, const double* fields,
                                       pf_op** out) {
  return guarded([&] {
    PF_REQUIRE(h && out, "null handle/output");
    PF_REQUIRE(L >= 2 && L <= 32, "spin chain: 2 <= L <= 32 required");
    PF_REQUIRE(Lup >= 0 && Lup <= L, "spin chain: 0 <= Lup <= L required");
    PF_REQUIRE(Lup + 2 <= kRankT, "spin chain: Lup <= 18 required");
    PF_REQUIRE(nbonds >= 0 && nbonds <= kMaxBonds, "spin chain: at most 128 bonds");
    PF_REQUIRE(nbonds == 0 || (site_i && site_j && J && Delta), "spin chain: null bond arrays");
    PF_REQUIRE(fields, "spin chain: null fields");
    PF_HIP(hipSetDevice(h->device));
    // basis size
    uint64_t D = 1;
    for (int i = 1; i <= Lup; ++i) D = D * (uint64_t)(L - Lup + i) / (uint64_t)i;
    PF_REQUIRE(D >= 1 && D < ((uint64_t)1 << 32), "spin chain: dimension too large");

    pf_op* op = new pf_op();
    op->h = h;
    op->kind = kChain;
    op->D = (int64_t)D;
    ChainParams& P = op->chain;
    P.L = L;
    P.Lup = Lup;
    P.D = (int64_t)D;
    P.nb = nbonds;
    // value classes: J/2 per bond, grouped in ascending value order
    std::vector<double> halfj(nbonds), jd(nbonds);
    std::vector<int8_t> oi(nbonds), oj(nbonds);
    std::vector<double> classes;
    for (int b = 0; b < nbonds; ++b) {
      PF_REQUIRE(site_i[b] >= 1 && site_i[b] <= L && site_j[b] >= 1 && site_j[b] <= L &&
                     site_i[b] != site_j[b],
                 "spin chain: bond sites must be distinct and in 1..L");
      oi[b] = (int8_t)(site_i[b] - 1);
      oj[b] = (int8_t)(site_j[b] - 1);
      halfj[b] = J[b] / 2;
      jd[b] = J[b] * Delta[b];
      if (halfj[b] != 0.0 &&
          std::find(classes.begin(), classes.end(), halfj[b]) == classes.end())
        classes.push_back(halfj[b]);
    }
    std::sort(classes.begin(), classes.end());
    int nb = 0;
    for (const double cv : classes)
      for (int b = 0; b < nbonds; ++b)
        if (halfj[b] == cv) {
          P.lo[nb] = std::min(oi[b], oj[b]);
          P.hi[nb] = std::max(oi[b], oj[b]);
          P.bval[nb] = cv;
          ++nb;
        }
    P.nb = nb;  // bonds with a nonzero off-diagonal value, in class order

    // basis (ascending, Gosper's hack) and binomial table
    std::vector<uint32_t> states(D);
    if (Lup == 0) {
      states[0] = 0;
    } else {
      uint64_t v = ((uint64_t)1 << Lup) - 1;
      for (uint64_t r = 0; r < D; ++r) {
        states[r] = (uint32_t)v;
        const uint64_t t = v | (v - 1);
        v = (t + 1) | (((~t & (t + 1)) - 1) >> (__builtin_ctzll(v) + 1));
      }
    }
    std::vector<uint32_t> C(32 * kRankT);
    fill_binomials(C.data());
    // Two-table rank: with a = low m bits, b = high L-m bits of a state,
    // rank = TA[a] + TB[b], TA[a] = sum_t C(p_t, t+1) over the set bits of a
    // and TB[b] = sum_t' C(m + q_t', nA + t' + 1) over the set bits q of b,
    // nA = Lup - |b| (the particles of a come first in the combinadic sum).
    std::vector<unsigned char> rtab;
    if (L <= kMaxTableL && D < ((uint64_t)1 << 24)) {  // ranks fit 24-bit offsets
      const int m = (L + 1) / 2;
      const uint32_t nA = 1u << m, nB = 1u << (L - m);
      const size_t bytes = ((size_t)nA * 2 + (size_t)nB * 4 + 15) & ~(size_t)15;
      rtab.assign(bytes, 0);
      uint16_t* TA = reinterpret_cast<uint16_t*>(rtab.data());
      uint32_t* TB = reinterpret_cast<uint32_t*>(rtab.data() + (size_t)nA * 2);
      for (uint32_t a = 0; a < nA; ++a) TA[a] = (uint16_t)full_rank(a, C.data());
      for (uint32_t b = 0; b < nB; ++b) {
        const int na = Lup - popc32(b);
        if (na < 0 || na > m) continue;
        uint64_t off = 0;
        int t = 0;
        for (uint32_t x = b; x; x &= x - 1u) {
          off += C[(m + ctz32(x)) * kRankT + na + t + 1];
          ++t;
        }
        TB[b] = (uint32_t)off;
      }
      for (uint64_t r = 0; r < D; ++r) {
        const uint32_t st = states[r];
        if ((uint64_t)TA[st & (nA - 1)] + TB[st >> m] != r) {
          delete op;
          throw Error{PF_ERR_INVALID, "spin chain: rank table check failed"};
        }
      }
      P.tsplit = m;
    } else {
      P.tsplit = 0;
    }
    // The XCD partition order: the same segments (rows of equal high bits, in
    // colex order inside) placed in rcm_segment_order; TA is unchanged.
    std::vector<unsigned char> prtab;
    std::vector<uint32_t> pstates, pperm;
    if (P.tsplit > 0 && P.nb > 0) {
      const int m = P.tsplit;
      const uint32_t nA = 1u << m, nB = 1u << (L - m);
      const uint16_t* TA = reinterpret_cast<const uint16_t*>(rtab.data());
      prtab = rtab;
      uint32_t* TBp = reinterpret_cast<uint32_t*>(prtab.data() + (size_t)nA * 2);
      std::fill(TBp, TBp + nB, 0u);
      uint64_t off = 0;
      for (const uint32_t b : rcm_segment_order(P, m)) {
        TBp[b] = (uint32_t)off;
        off += C[m * kRankT + (Lup - popc32(b))];  // C(m, set bits of the low part)
      }
      pstates.assign(D, 0u);
      pperm.assign(D, ~0u);
      bool ok = off == D;
      for (uint64_t r = 0; ok && r < D; ++r) {
        const uint32_t st = states[r];
        const uint64_t i = (uint64_t)TA[st & (nA - 1)] + TBp[st >> m];
        if (i >= D || pperm[i] != ~0u) {
          ok = false;
          break;
        }
        pstates[i] = st;
        pperm[i] = (uint32_t)r;
      }
      if (!ok) {
        delete op;
        throw Error{PF_ERR_INVALID, "spin chain: partition order check failed"};
      }
    }
    try {
      if (!rtab.empty()) {
        PF_HIP(hipMalloc(&op->rtab, rtab.size()));
        PF_HIP(hipMemcpy(op->rtab, rtab.data(), rtab.size(), hipMemcpyHostToDevice));
        op->rtab_bytes = rtab.size();
      }
      if (!pstates.empty()) {
        PF_HIP(hipMalloc(&op->prtab, prtab.size()));
        PF_HIP(hipMalloc(&op->pstates, D * sizeof(uint32_t)));
        PF_HIP(hipMalloc(&op->pperm, D * sizeof(uint32_t)));
        PF_HIP(hipMalloc(&op->pdiag, D * sizeof(double)));
        PF_HIP(hipMemcpy(op->prtab, prtab.data(), prtab.size(), hipMemcpyHostToDevice));
        PF_HIP(hipMemcpy(op->pstates, pstates.data(), D * sizeof(uint32_t), hipMemcpyHostToDevice));
        PF_HIP(hipMemcpy(op->pperm, pperm.data(), D * sizeof(uint32_t), hipMemcpyHostToDevice));
      }
      PF_HIP(hipMalloc(&op->states, D * sizeof(uint32_t)));
      PF_HIP(hipMalloc(&op->binom, C.size() * sizeof(uint32_t)));
      PF_HIP(hipMalloc(&op->diag, D * sizeof(double)));
      PF_HIP(hipMemcpy(op->states, states.data(), D * sizeof(uint32_t), hipMemcpyHostToDevice));
      PF_HIP(hipMemcpy(op->binom, C.data(), C.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
      // diagonal on the device (reference order)
      double* dfields;
      int8_t* dbi;
      int8_t* dbj;
      double* djd;
      const int nbm = std::max(nbonds, 1);
      PF_HIP(hipMalloc(&dfields, L * sizeof(double)));
      PF_HIP(hipMalloc(&dbi, nbm));
      PF_HIP(hipMalloc(&dbj, nbm));
      PF_HIP(hipMalloc(&djd, nbm * sizeof(double)));
      PF_HIP(hipMemcpy(dfields, fields, L * sizeof(double), hipMemcpyHostToDevice));
      if (nbonds) {
        PF_HIP(hipMemcpy(dbi, oi.data(), nbonds, hipMemcpyHostToDevice));
        PF_HIP(hipMemcpy(dbj, oj.data(), nbonds, hipMemcpyHostToDevice));
        PF_HIP(hipMemcpy(djd, jd.data(), nbonds * sizeof(double), hipMemcpyHostToDevice));
      }
      hipLaunchKernelGGL(chain_diag_kernel, dim3((unsigned)((D + 255) / 256)), dim3(256), 0, 0,
                         op->states, (int64_t)D, L, nbonds, dfields, dbi, dbj, djd, op->diag);
      PF_HIP(hipGetLastError());
      if (op->pstates) {  // the same per-state values, in partition order
        hipLaunchKernelGGL(chain_diag_kernel, dim3((unsigned)((D + 255) / 256)), dim3(256), 0, 0,
                           op->pstates, (int64_t)D, L, nbonds, dfields, dbi, dbj, djd, op->pdiag);
        PF_HIP(hipGetLastError());
      }
      PF_HIP(hipDeviceSynchronize());
      (void)hipFree(dfields);
      (void)hipFree(dbi);
      (void)hipFree(dbj);
      (void)hipFree(djd);
    } catch (...) {
      pf_op_destroy(op);
      throw;
    }
    *out = op;
  });
}

extern "C" int pf_op_create_qsun(pf_handle* h, int32_t L_grain, int32_t L_loc, double g0,
                                 const double* H_grain, const int32_t* neighbors,
                                 const double* couplings, const double* fields, pf_op** out) {
  return guarded([&] {
    PF_REQUIRE(h && out, "null handle/output");
    PF_REQUIRE(L_grain >= 1 && L_grain <= kMaxGrain, "qsun: 1 <= L_grain <= 4");
    PF_REQUIRE(L_loc >= 1 && L_loc <= kMaxLoc && L_grain + L_loc <= 31, "qsun: bad L_loc");
    PF_REQUIRE(H_grain && neighbors && couplings && fields, "qsun: null arrays");
    PF_HIP(hipSetDevice(h->device));
    pf_op* op = new pf_op();
    op->h = h;
    op->kind = kQsun;
    op->D = (int64_t)1 << (L_grain + L_loc);
    QsunParams& P = op->qsun;
    P.Lg = L_grain;
    P.Lloc = L_loc;
    P.D = op->D;
    const int dg = 1 << L_grain;
    for (int j = 0; j < L_loc; ++j) {
      if (neighbors[j] < 0 || neighbors[j] >= L_grain) {
        delete op;
        throw Error{PF_ERR_INVALID, "qsun: neighbors must be in [0, L_grain)"};
      }
      P.flip[j] = (1u << (L_grain + j)) | (1u << neighbors[j]);
      P.coup[j] = g0 * couplings[j] * 0.5 * 0.5;
      P.h[j] = fields[j];
    }
    // row-major copy of the (column-major) grain matrix: G[a*dg + b] = H[a, b]
    for (int a = 0; a < dg; ++a)
      for (int b = 0; b < dg; ++b) P.G[a * dg + b] = H_grain[a + b * dg];
    // Diagonal split: d(r) = Dlo[r mod 2^lb] + Dhi[r >> lb], Dlo carrying the
    // grain diagonal and the z-fields of the outer spins below bit lb.
    const int lb = L_grain + (L_loc + 1) / 2;
    const int hb = L_grain + L_loc - lb;
    std::vector<double> tab(((size_t)1 << lb) + ((size_t)1 << hb));
    for (uint32_t t = 0; t < (1u << lb); ++t) {
      double v = 0.0;
      for (int j = 0; L_grain + j < lb; ++j) v += fields[j] * (((t >> (L_grain + j)) & 1u) ? 0.5 : -0.5);
      tab[t] = v + P.G[(t & (dg - 1)) * (dg + 1)];
    }
    for (uint32_t u = 0; u < (1u << hb); ++u) {
      double v = 0.0;
      for (int j = lb - L_grain; j < L_loc; ++j)
        v += fields[j] * (((u >> (L_grain + j - lb)) & 1u) ? 0.5 : -0.5);
      tab[((size_t)1 << lb) + u] = v;
    }
    op->qlb = lb;
    if (hipMalloc(&op->qdiag, tab.size() * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      delete op;
      throw Error{PF_ERR_NOMEM, "qsun: diagonal table allocation failed"};
    }
    if (hipMemcpy(op->qdiag, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice) !=
        hipSuccess) {
      (void)hipFree(op->qdiag);
      delete op;
      throw Error{PF_ERR_HIP, "qsun: diagonal table upload failed"};
    }
    *out = op;
  });
}

extern "C" int pf_op_destroy(pf_op* op) {
  if (!op) return PF_OK;
  if (op->states) (void)hipFree(op->states);
  if (op->binom) (void)hipFree(op->binom);
  if (op->diag) (void)hipFree(op->diag);
  if (op->rtab) (void)hipFree(op->rtab);
  if (op->pstates) (void)hipFree(op->pstates);
  if (op->pdiag) (void)hipFree(op->pdiag);
  if (op->prtab) (void)hipFree(op->prtab);
  if (op->pperm) (void)hipFree(op->pperm);
  if (op->qdiag) (void)hipFree(op->qdiag);
  if (op->prow) (void)hipFree(op->prow);
  if (op->pent) (void)hipFree(op->pent);
  if (op->pvals) (void)hipFree(op->pvals);
  op->fbuf.release();
  op->kpart.release();
  delete op;
  return PF_OK;
}

extern "C" int64_t pf_op_dim(const pf_op* op) { return op ? op->D : 0; }

extern "C" int pf_op_set_rescale(pf_op* op, double Emin, double Emax) {
  return guarded([&] {
    PF_REQUIRE(op, "null operator");
    PF_REQUIRE(std::isfinite(Emin) && std::isfinite(Emax) && Emax > Emin,
               "rescale: need finite Emin < Emax");
    op->Emin = Emin;
    op->Emax = Emax;
  });
}

extern "C" int pf_op_connectivity_order(pf_op* op, int32_t order, int64_t row0, int64_t nrows,
                                        int32_t max_per_row, int64_t* cols, double* vals,
                                        int32_t* counts, double* diag, pf_stream stream) {
  return guarded([&] {
    PF_REQUIRE(op, "null operator");
    PF_REQUIRE(row0 >= 0 && nrows >= 0 && row0 + nrows <= op->D, "connectivity: rows out of range");
    PF_REQUIRE(max_per_row >= 1 && cols && vals && counts && diag, "connectivity: bad outputs");
    PF_REQUIRE(order == 0 || order == 1, "connectivity: order must be 0 (reference) or 1 (partition)");
    PF_REQUIRE(order == 0 || (op->kind == kChain && op->pstates),
               "connectivity: this operator has no partition row order");
    if (nrows == 0) return;
    hipStream_t st = S(stream);
    if (op->kind == kChain) {
      const ChainParams& P = op->chain;
      const bool part = order == 1;
      const int64_t nscan = part ? op->D : nrows;
      const unsigned nb = (unsigned)((nscan + 255) / 256);
      const uint32_t* states = part ? op->pstates : op->states;
      const void* tab = part ? op->prtab : op->rtab;
      const double* dg = part ? op->pdiag : op->diag;
      const uint32_t* perm = part ? op->pperm : nullptr;
      if (P.tsplit > 0)
        hipLaunchKernelGGL(chain_conn_kernel<true>, dim3(nb), dim3(256), 0, st, P, states,
                           op->binom, tab, dg, perm, row0, nrows, nscan, max_per_row, cols, vals,
                           counts, diag);
      else
        hipLaunchKernelGGL(chain_conn_kernel<false>, dim3(nb), dim3(256), 0, st, P, states,
                           op->binom, tab, dg, perm, row0, nrows, nscan, max_per_row, cols, vals,
                           counts, diag);
      PF_HIP(hipGetLastError());
      PF_HIP(hipStreamSynchronize(st));
    } else if (op->kind == kPacked) {
      packed_connectivity(op, row0, nrows, max_per_row, cols, vals, counts, diag, st);
    } else {
      const unsigned nb = (unsigned)((nrows + 255) / 256);
      hipLaunchKernelGGL(qsun_conn_kernel, dim3(nb), dim3(256), 0, st, op->qsun, op->qdiag,
                         op->qlb, row0, nrows, max_per_row, cols, vals, counts, diag);
      PF_HIP(hipGetLastError());
      PF_HIP(hipStreamSynchronize(st));
    }
  });
}

extern "C" int pf_op_connectivity(pf_op* op, int64_t row0, int64_t nrows, int32_t max_per_row,
                                  int64_t* cols, double* vals, int32_t* counts, double* diag,
                                  pf_stream stream) {
  return pf_op_connectivity_order(op, 0, row0, nrows, max_per_row, cols, vals, counts, diag,
                                  stream);
}

// =========================================================== C ABI: maps
extern "C" int pf_apply(pf_op* op, int32_t dtype, int32_t rescaled, const void* X, int64_t ldx,
                        void* Y, int64_t ldy, int32_t ncols, pf_stream stream) {
  return guarded([&] {
    PF_REQUIRE(op && X && Y, "apply: null argument");
    PF_REQUIRE(ncols >= 1 && ncols <= 65535, "apply: 1 <= ncols <= 65535");
    PF_REQUIRE(ldx == ldy && ldx >= op->D, "apply: X and Y need the same leading dimension >= D");
    PF_REQUIRE(X != Y, "apply: X and Y must not alias");
    Layout lay = colmajor_layout(ldx);
    lay.pblocks = default_pblocks(op);
    launch_step(op, dtype, kHasH, lay, ncols, op->affine(rescaled != 0), 0.0, 1.0, 0.0, nullptr, X,
                nullptr, Y, 0, S(stream));
  });
}

extern "C" int pf_clenshaw_recurrence(pf_op* op, int32_t dtype, void* b1, const void* b2,
                                      const void* b3, double c, double alpha, double beta,
                                      const void* X, int64_t ld, int32_t ncols, pf_stream stream) {
  return guarded([&] {
    PF_REQUIRE(op && b1 && b2 && X, "recurrence: null argument");
    PF_REQUIRE(ncols >= 1 && ncols <= 65535 && ld >= op->D, "recurrence: bad shape");
    PF_REQUIRE(b1 != b2, "recurrence: b1 must not alias b2");
    int flags = kHasX | kHasH;
    if (beta != 0.0) {
      PF_REQUIRE(b3 != nullptr, "recurrence: b3 is null but beta != 0");
      flags |= kHasB3;
    }
    Layout lay = colmajor_layout(ld);
    lay.pblocks = default_pblocks(op);
    launch_step(op, dtype, flags, lay, ncols, op->affine(true), c, alpha, beta, X,
                b2, b3, b1, 0, S(stream));
  });
}

extern "C" int pf_clenshaw_finalsum(pf_op* op, int32_t dtype, const void* b2, const void* b3,
                                    double c0, double beta1, double phi0, double phi1, void* Y,
                                    const void* X, int64_t ld, int32_t ncols, pf_stream stream) {
  return guarded([&] {
    PF_REQUIRE(op && b2 && Y && X, "finalsum: null argument");
    PF_REQUIRE(ncols >= 1 && ncols <= 65535 && ld >= op->D, "finalsum: bad shape");
    PF_REQUIRE(Y != b2, "finalsum: Y must not alias b2");
    int flags = kHasX | kHasH;
    if (beta1 * phi0 != 0.0) {
      PF_REQUIRE(b3 != nullptr, "finalsum: b3 is null");
      flags |= kHasB3;
    }
    Layout lay = colmajor_layout(ld);
    lay.pblocks = default_pblocks(op);
    launch_step(op, dtype, flags, lay, ncols, op->affine(true), c0 * phi0, phi1,
                beta1 * phi0, X, b2, b3, Y, 0, S(stream));
  });
}

// polynomial_properties (src/Clenshaw/ClenshawMapping.jl:7-12)
static void poly_props(int type, int k, double& alpha, double& beta_k) {
  switch (type) {
    case PF_CHEBYSHEV: alpha = 2.0; beta_k = -1.0; break;
    case PF_LEGENDRE: alpha = (2.0 * k + 1) / (k + 1); beta_k = -(double)k / (k + 1); break;
    case PF_HERMITE: alpha = 2.0; beta_k = -2.0 * k; break;
    case PF_TAYLOR: alpha = 1.0; beta_k = 0.0; break;
    default: throw Error{PF_ERR_INVALID, "unknown polynomial type"};
  }
}
static void poly_phi(int type, double& phi0, double& phi1) {
  phi0 = 1.0;
  phi1 = (type == PF_HERMITE) ? 2.0 : 1.0;
}

namespace pf {
// Chunk width for the row-interleaved drivers (columns are independent).
static int chunk_cols(int dtype) { return dtype == PF_F64 ? 32 : 64; }
// Flag of the driver launches: the partition row order when built and chosen.
static int driver_part(const pf_op* op, const Layout& lay) {
  return (op->pstates && lay.part) ? kRowPart : 0;
}
// Reference rank of each partition row, or null for the reference order.
static const uint32_t* driver_perm(const pf_op* op, const Layout& lay) {
  return driver_part(op, lay) ? op->pperm : nullptr;
}
}  // namespace pf

namespace pf {

// The filter driver shared by pf_filter and pf_filter_noalloc.  `work(bytes)`
// returns the device memory of one column chunk's four row-interleaved blocks
// (packed X, b1, b2, b3); nothing in them needs to start at zero, because the
// first two steps read neither b2 nor b3 (zero tracking below).
template <class Work>
static void filter_body(pf_op* op, int32_t dtype, int32_t polytype, const double* coeffs,
                        int32_t order, const double* X, int64_t ldx, double* Y, int64_t ldy,
                        int32_t ncols, hipStream_t st, Work&& work) {
  double phi0, phi1, a_tmp, b_tmp;
  poly_phi(polytype, phi0, phi1);
  poly_props(polytype, 1, a_tmp, b_tmp);
  const Affine af = op->affine(true);
  const size_t esz = dtype == PF_F64 ? 8 : 4;
  const int cmax = chunk_cols(dtype);
  for (int c0 = 0; c0 < ncols; c0 += cmax) {
    const int s = std::min(cmax, ncols - c0);
    const Layout lay = tuned_layout(op, dtype, s);
    const size_t blk = (size_t)op->D * s * esz;
    char* base = reinterpret_cast<char*>(work(4 * blk));
    void* Xp = base;
    void* b[3] = {base + blk, base + 2 * blk, base + 3 * blk};
    pack_block(dtype, X + (int64_t)c0 * ldx, ldx, Xp, op->D, s, nullptr, st,
               driver_perm(op, lay));
    double* Yc = Y + (int64_t)c0 * ldy;
    bool z2 = true, z3 = true;  // b2 / b3 known to be zero
    for (int k = order; k >= 1; --k) {
      double alpha, beta_k1, dummy;
      poly_props(polytype, k, alpha, dummy);
      poly_props(polytype, k + 1, dummy, beta_k1);
      int flags = kHasX | driver_part(op, lay);
      if (!z2) flags |= kHasH;
      if (!z3 && beta_k1 != 0.0) flags |= kHasB3;
      launch_step(op, dtype, flags, lay, s, af, coeffs[k], alpha, beta_k1, Xp, b[1], b[2], b[0],
                  0, st);
      void* t = b[0];
      b[0] = b[2];
      b[2] = b[1];
      b[1] = t;
      z3 = z2;
      z2 = false;
    }
    double beta1, dummy;
    poly_props(polytype, 1, dummy, beta1);
    int flags = kHasX | kOutCM | driver_part(op, lay);
    if (!z2) flags |= kHasH;
    if (!z3 && beta1 * phi0 != 0.0) flags |= kHasB3;
    launch_step(op, dtype, flags, lay, s, af, coeffs[0] * phi0, phi1, beta1 * phi0, Xp, b[1],
                b[2], Yc, ldy, st);
  }
}

static void filter_args(pf_op* op, int32_t dtype, const double* coeffs, int32_t order,
                        const double* X, int64_t ldx, double* Y, int64_t ldy, int32_t ncols) {
  PF_REQUIRE(op && coeffs && X && Y, "filter: null argument");
  PF_REQUIRE(order >= 0, "filter: order >= 0");
  PF_REQUIRE(ncols >= 1, "filter: ncols >= 1");
  PF_REQUIRE(ldx >= op->D && ldy >= op->D, "filter: leading dimension < D");
  PF_REQUIRE(dtype == PF_F64 || dtype == PF_F32, "filter: dtype");
}

// ------------------------------------------------ no-allocation filter work
// Work memory of pf_filter_noalloc: a fixed device array, so that the entry
// point never allocates.  Each GPU the library runs on holds its own copy
// (kNoAllocScratchBytes) until the process ends.  Host code must not use the
// array's name as an address (it is a host-side stand-in): the device address
// is looked up once per device by noalloc_scratch_init and kept in
// g_scratch_ptr.  All calls on one GPU share that GPU's array, so calls that
// overlap on one GPU (different streams or threads) race.
constexpr size_t kNoAllocScratchBytes = (size_t)128 << 20;
constexpr int kMaxScratchDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static void* g_scratch_ptr[kMaxScratchDevices] = {};
static std::mutex g_scratch_mu;

// Called by pf_create with `device` current.
void noalloc_scratch_init(int device) {
  PF_REQUIRE(device >= 0 && device < kMaxScratchDevices, "scratch: device index too large");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    throw Error{PF_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
  g_scratch_ptr[device] = ptr;
}

static void* noalloc_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  void* p = (device >= 0 && device < kMaxScratchDevices) ? g_scratch_ptr[device] : nullptr;
  PF_REQUIRE(p, "filter_noalloc: no scratch address for this device (pf_create not called?)");
  return p;
}

}  // namespace pf

extern "C" int pf_filter(pf_op* op, int32_t dtype, int32_t polytype, const double* coeffs,
                         int32_t order, const double* X, int64_t ldx, double* Y, int64_t ldy,
                         int32_t ncols, pf_stream stream) {
  return guarded([&] {
    filter_args(op, dtype, coeffs, order, X, ldx, Y, ldy, ncols);
    filter_body(op, dtype, polytype, coeffs, order, X, ldx, Y, ldy, ncols, S(stream),
                [&](size_t bytes) { return op->fbuf.get(bytes); });
  });
}

extern "C" int pf_filter_noalloc(pf_op* op, int32_t dtype, int32_t polytype,
                                 const double* coeffs, int32_t order, const double* X,
                                 int64_t ldx, double* Y, int64_t ldy, int32_t ncols,
                                 pf_stream stream) {
  return guarded([&] {
    filter_args(op, dtype, coeffs, order, X, ldx, Y, ldy, ncols);
    const size_t esz = dtype == PF_F64 ? 8 : 4;
    const size_t need = 4 * (size_t)op->D * (size_t)std::min(ncols, chunk_cols(dtype)) * esz;
    if (need > kNoAllocScratchBytes)
      throw Error{PF_ERR_INVALID,
                  "filter_noalloc: 4 x D x columns x element size = " + std::to_string(need) +
                      " bytes exceeds the " + std::to_string(kNoAllocScratchBytes >> 20) +
                      " MiB no-allocation scratch (use pf_filter)"};
    char* base = reinterpret_cast<char*>(noalloc_scratch(op->h->device));
    filter_body(op, dtype, polytype, coeffs, order, X, ldx, Y, ldy, ncols, S(stream),
                [&](size_t) { return (void*)base; });
  });
}

extern "C" int pf_fill_probes(int32_t kind, uint64_t seed, int64_t D, int32_t R, double* out,
                              int64_t ld, pf_stream stream) {
  return guarded([&] {
    PF_REQUIRE(out && D >= 1 && R >= 1 && R <= 65535 && ld >= D, "probes: bad shape");
    PF_REQUIRE(kind == PF_PROBE_GAUSSIAN || kind == PF_PROBE_PHASE, "probes: bad kind");
    hipLaunchKernelGGL(probes_kernel, dim3((unsigned)((D + 255) / 256), (unsigned)R), dim3(256), 0,
                       S(stream), kind, seed, D, R, out, ld);
    PF_HIP(hipGetLastError());
  });
}

extern "C" int pf_kpm_moments(pf_op* op, int32_t dtype, int32_t N, int32_t R, const double* probes,
                              int64_t ldp, double* mu_out, pf_stream stream) {
  return guarded([&] {
    PF_REQUIRE(op && probes && mu_out, "kpm: null argument");
    PF_REQUIRE(N >= 2 && R >= 1 && ldp >= op->D, "kpm: need N >= 2, R >= 1, ldp >= D");
    PF_REQUIRE(dtype == PF_F64 || dtype == PF_F32, "kpm: dtype");
    hipStream_t st = S(stream);
    const Affine af = op->affine(true);
    const size_t esz = dtype == PF_F64 ? 8 : 4;
    const int cmax = chunk_cols(dtype);
    const int dot_blocks = 1024;
    // partials: [N][dot_blocks] for the moments + column norms + mu accumulator
    const int nb_norm = 256;
    double* part = reinterpret_cast<double*>(
        op->kpart.get(((size_t)N * dot_blocks + (size_t)cmax * nb_norm + cmax + N) * sizeof(double)));
    double* npart = part + (size_t)N * dot_blocks;
    double* cscale = npart + (size_t)cmax * nb_norm;
    double* mu_dev = cscale + cmax;
    PF_HIP(hipMemsetAsync(mu_dev, 0, N * sizeof(double), st));
    for (int c0 = 0; c0 < R; c0 += cmax) {
      const int s = std::min(cmax, R - c0);
      const Layout lay = tuned_layout(op, dtype, s);
      const size_t blk = (size_t)op->D * s * esz;
      char* base = reinterpret_cast<char*>(op->fbuf.get(4 * blk));
      void* P = base;  // normalized probes = T_0
      void* t[3] = {base + blk, base + 2 * blk, base + 3 * blk};
      const double* src = probes + (int64_t)c0 * ldp;
      hipLaunchKernelGGL(colnorm2_kernel, dim3(nb_norm, s), dim3(256), 0, st, src, ldp, op->D, npart);
      hipLaunchKernelGGL(inv_sqrt_sum_kernel, dim3(1), dim3(64), 0, st, npart, nb_norm, s, cscale);
      PF_HIP(hipGetLastError());
      // (in the partition row order the dots are the same sums, reordered)
      const int rp = driver_part(op, lay);
      pack_block(dtype, src, ldp, P, op->D, s, cscale, st, driver_perm(op, lay));
      // mu_0: <P, P>  (an H-free step that copies P and dots it)
      launch_step(op, dtype, kHasX | kDot | rp, lay, s, af, 1.0, 0.0, 0.0, P, nullptr, nullptr,
                  t[0], 0, st, part + 0, P, dot_blocks);
      // mu_1: T_1 = H~ T_0
      launch_step(op, dtype, kHasH | kDot | rp, lay, s, af, 0.0, 1.0, 0.0, nullptr, P, nullptr,
                  t[1], 0, st, part + (size_t)1 * dot_blocks, P, dot_blocks);
      // T_{n} = 2 H~ T_{n-1} - T_{n-2}; t[0] holds T_0 (copy), t[1] T_1
      void* tm2 = t[0];
      void* tm1 = t[1];
      void* tn = t[2];
      for (int n = 2; n < N; ++n) {
        launch_step(op, dtype, kHasH | kHasB3 | kDot | rp, lay, s, af, 0.0, 2.0, -1.0, nullptr,
                    tm1, tm2, tn, 0, st, part + (size_t)n * dot_blocks, P, dot_blocks);
        void* tmp = tm2;
        tm2 = tm1;
        tm1 = tn;
        tn = tmp;
      }
      // every step wrote nblk block partials at stride dot_blocks
      const int nblk = (int)step_grid_x(op, lay, kDot, dot_blocks);
      hipLaunchKernelGGL(kpm_reduce_kernel, dim3((N + 255) / 256), dim3(256), 0, st, part,
                         dot_blocks, nblk, N, mu_dev);
      PF_HIP(hipGetLastError());
    }
    op->khost.resize(N);
    PF_HIP(hipMemcpyAsync(op->khost.data(), mu_dev, N * sizeof(double), hipMemcpyDeviceToHost, st));
    PF_HIP(hipStreamSynchronize(st));
    const double scale = (double)op->D / (double)R;
    for (int n = 0; n < N; ++n) mu_out[n] = op->khost[n] * scale;
  });
}

extern "C" int pf_op_set_tuning(pf_op* op, int32_t key, int32_t value) {
  return guarded([&] {
    PF_REQUIRE(op, "null operator");
    if (key == PF_TUNE_XCD_REMAP) {
      PF_REQUIRE(value >= -1 && value <= 1, "tuning: xcd remap is -1, 0 or 1");
      op->force.xcd = value;
    } else if (key == PF_TUNE_VEC) {
      PF_REQUIRE(value == 0 || value == 1 || value == 2 || value == 4 || value == 8 ||
                     value == 16 || value == 32,
                 "tuning: vector width 0, 1, 2, 4, 8, 16 or 32");
      op->force.V = value;
    } else if (key == PF_TUNE_NONTEMPORAL) {
      PF_REQUIRE(value >= -1 && value <= 1, "tuning: non-temporal is -1, 0 or 1");
      op->force.nt = value;
    } else if (key == PF_TUNE_OCCUPANCY) {
      PF_REQUIRE(value >= -1 && value <= 1, "tuning: occupancy variant is -1, 0 or 1");
      op->force.occ = value;
    } else if (key == PF_TUNE_GRAIN_SHUFFLE) {
      PF_REQUIRE(value >= -1 && value <= 1, "tuning: grain shuffle is -1, 0 or 1");
      op->force.shfl = value;
    } else if (key == PF_TUNE_WIDE) {
      PF_REQUIRE(value >= -1 && value <= 1, "tuning: wide gather batches is -1, 0 or 1");
      op->force.wide = value;
    } else if (key == PF_TUNE_PBLOCKS) {
      PF_REQUIRE(value >= -1 && value <= 64, "tuning: persistent blocks per CU is -1..64");
      op->force.pb = value;
    } else if (key == PF_TUNE_ROW_PARTITION) {
      PF_REQUIRE(value >= -1 && value <= 1, "tuning: row partition is -1, 0 or 1");
      op->force.part = value;
    } else if (key == PF_TUNE_RESET) {
      op->force = TuneCfg{};
      op->tuned.clear();
    } else {
      throw Error{PF_ERR_INVALID, "tuning: unknown key"};
    }
  });
}

// Times the fused recurrence step for every candidate (vector width, XCD
// mapping) on the operator's own shape and keeps the fastest for (dtype, ncols).
extern "C" int pf_op_get_tuning(pf_op* op, int32_t dtype, int32_t ncols, int32_t* cfg) {
  return guarded([&] {
    PF_REQUIRE(op && cfg, "get_tuning: null argument");
    PF_REQUIRE(dtype == PF_F64 || dtype == PF_F32, "get_tuning: dtype");
    PF_REQUIRE(ncols >= 1 && ncols <= chunk_cols(dtype), "get_tuning: 1 <= ncols <= chunk width");
    const Layout lay = tuned_layout(op, dtype, ncols);
    const int32_t v[8] = {lay.V, lay.xcd, lay.nt, lay.pblocks, lay.wide, lay.shfl, lay.occ,
                          lay.part};
    for (int i = 0; i < 8; ++i) cfg[i] = v[i];
  });
}

extern "C" int pf_autotune(pf_op* op, int32_t dtype, int32_t ncols, int32_t nsteps,
                           double* best_us, int32_t* best_vec, int32_t* best_xcd,
                           pf_stream stream) {
  return guarded([&] {
    PF_REQUIRE(op, "null operator");
    PF_REQUIRE(dtype == PF_F64 || dtype == PF_F32, "autotune: dtype");
    PF_REQUIRE(ncols >= 1 && ncols <= chunk_cols(dtype), "autotune: 1 <= ncols <= chunk width");
    PF_REQUIRE(nsteps >= 1 && nsteps <= 100000, "autotune: nsteps");
    hipStream_t st = S(stream);
    const size_t esz = dtype == PF_F64 ? 8 : 4;
    const size_t blk = (size_t)op->D * ncols * esz;
    char* base = reinterpret_cast<char*>(op->fbuf.get(4 * blk));
    PF_HIP(hipMemsetAsync(base, 0, 4 * blk, st));
    void* b[4] = {base, base + blk, base + 2 * blk, base + 3 * blk};
    const Affine af = op->affine(true);
    hipEvent_t e0, e1;
    PF_HIP(hipEventCreate(&e0));
    PF_HIP(hipEventCreate(&e1));
    // Candidates are timed in interleaved rounds (min per candidate) so that
    // clock ramp-up and transient interference do not favour one of them; a
    // later candidate must beat the best by more than 1% to replace it.
    struct Cand {
      TuneCfg c;
      double us;
    };
    std::vector<Cand> cands;
    const int vs[6] = {1, 2, 4, 8, 16, 32};
    // persistent blocks per CU: beyond the resident count (12, 16) the extra
    // blocks start as others finish, which balances the last row groups
    const int pbs[5] = {default_pblocks(op), 4, 0, 12, 16};
    for (int vi = 0; vi < 6; ++vi) {
      if (!valid_vec(dtype, ncols, vs[vi])) continue;
      // kernel variants: for spin chains with rank tables the plain step and
      // wide gather batches, each in the reference and the partition row
      // order; for the Quantum Sun grain shuffles and grain reloads (the
      // 8-waves-per-SIMD build is left to forced tuning: it never won)
      const bool ctab = op->kind == kChain && op->chain.tsplit > 0;
      const int nvar = ctab ? (op->pstates ? 4 : 2) : (op->kind == kQsun ? 2 : 1);
      for (int pi = 0; pi < 5; ++pi) {
        if (pi > 0 && pbs[pi] == pbs[0]) continue;
        for (int w = 0; w < nvar; ++w) {
          TuneCfg c;
          c.V = vs[vi];
          c.xcd = 1;
          c.nt = 1;
          c.pb = pbs[pi];
          c.wide = ctab && (w & 1) ? 1 : 0;
          c.occ = 0;
          c.part = ctab && (w & 2) ? 1 : 0;
          c.shfl = op->kind == kQsun ? 1 - w : 1;
          cands.push_back({c, 1e300});
        }
      }
    }
    for (int round = 0; round < 3; ++round) {
      for (auto& cd : cands) {
        const Layout lay = interleaved_layout(dtype, ncols, cd.c.V, cd.c.xcd, cd.c.nt, cd.c.pb,
                                              cd.c.wide, cd.c.shfl, cd.c.occ, cd.c.part);
        const int fl = kHasX | kHasH | kHasB3 | driver_part(op, lay);
        for (int w = 0; w < 2; ++w)
          launch_step(op, dtype, fl, lay, ncols, af, 0.5, 2.0, -1.0, b[0], b[1], b[2], b[3], 0, st);
        PF_HIP(hipEventRecord(e0, st));
        for (int k = 0; k < nsteps; ++k)  // rotate three distinct recurrence buffers
          launch_step(op, dtype, fl, lay, ncols, af, 0.5, 2.0, -1.0, b[0], b[1 + k % 3],
                      b[1 + (k + 1) % 3], b[1 + (k + 2) % 3], 0, st);
        PF_HIP(hipEventRecord(e1, st));
        PF_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        PF_HIP(hipEventElapsedTime(&ms, e0, e1));
        cd.us = std::min(cd.us, 1e3 * ms / nsteps);
      }
    }
    double best = 1e300;
    TuneCfg bestc;
    for (const auto& cd : cands) {
      if (cd.us < 0.99 * best) {
        best = cd.us;
        bestc = cd.c;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    op->tuned[{dtype, ncols}] = bestc;
    if (best_us) *best_us = best;
    if (best_vec) *best_vec = bestc.V;
    if (best_xcd) *best_xcd = bestc.xcd;
  });
}
