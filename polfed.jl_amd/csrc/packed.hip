// Packed sparse operator: any real symmetric matrix given in the reference's
// value-bucket form (get_diags_and_offdiagonals_by_value,
// src/Optimization/optimization.jl:134-182): the diagonal, the distinct
// off-diagonal values v (ascending) and, per value, the columns of each row
// holding v.  The reference's row kernel (mapping_state_i,
// src/Optimization/cpu/single_row_mappings.jl:66-79) computes
//   y_i = d_i x_i + sum_v v * sum_{j in N_v(i)} x_j.
//
// Layout in HBM: one row-major entry table instead of one CSR per value --
// row r's entries are pent[prow[r] .. prow[r+1]), ordered by value index then
// column, each a u32 column with the value index in bits 24..31 (a plain
// column when there is a single value).  A row is then one contiguous run of
// 4-byte entries whatever the number of values, and the values sit in LDS.
// The step kernel shares the row-interleaved blocks, the persistent
// XCD-grouped grid and the fused epilogue of the matrix-free kernels.

#include <algorithm>
#include <cstring>
#include <vector>

#include "pf_device.h"
#include "pf_internal.h"
#include "pf_step.h"

namespace pf {

// out = cx*X + ch*((H - shift)*scale) b2 + cb*b3 for a packed operator.  The
// G = s/V lanes of a row each gather their V columns of every entry; NB
// entries are loaded before any is accumulated.  With GM = G in {2, 4, 8} the
// lanes of a row load different entries of a batch and share them by
// cross-lane broadcasts (GM = 1 or 0: every lane loads all of them).  P2: the
// row stride in bytes is a power of two (offsets by shift).  Each entry adds
// (v*scale) * x_j, so the terms of one value are not summed before the
// multiply as in the reference (a rounding-level difference).
template <typename T, int V, bool P2, int GM>
__global__ __launch_bounds__(256) void packed_step_kernel(
    const int64_t D, const int64_t* __restrict__ prow, const uint32_t* __restrict__ pent,
    const double* __restrict__ diag, const double* __restrict__ pvals, const int nvals,
    const Affine af, const int flags, const T cx, const T ch, const T cb,
    const T* __restrict__ X, const T* __restrict__ B2, const T* __restrict__ B3,
    void* __restrict__ OUT, const Layout lay, const int64_t out_ld,
    double* __restrict__ dot_part, const T* __restrict__ Pr, const int64_t ngroups) {
  typedef typename VT<T, V>::t vt;
  constexpr int NB0 = (V * (int)sizeof(T) <= 32) ? 8 : 4;  // gathers in flight
  constexpr int NB = NB0 < GM ? GM : NB0;
  __shared__ T sV[kMaxPackedVals];
  __shared__ double red[4];
  for (int i = threadIdx.x; i < nvals; i += blockDim.x) sV[i] = (T)(pvals[i] * af.scale);
  __syncthreads();
  // entry -> (column, value index): one value keeps all 32 bits for the column
  const uint32_t cmask = nvals > 1 ? 0x00FFFFFFu : 0xFFFFFFFFu;
  const uint32_t vmask = nvals > 1 ? 0xFFu : 0u;

  const int rloc = threadIdx.x / lay.G;
  const int g = threadIdx.x - rloc * lay.G;
  const bool lane_ok = rloc < lay.rows_per_block;
  const int64_t cofs = (int64_t)blockIdx.y * lay.cstride + (int64_t)g * V;
  const int out_col0 = (int)blockIdx.y + g * V;
  const char* b2b = reinterpret_cast<const char*>(B2 + (int64_t)blockIdx.y * lay.cstride);
  const uint32_t cofs_b = (uint32_t)((int64_t)g * V * sizeof(T));
  const uint32_t rsb = (uint32_t)(lay.rs * sizeof(T));
  const int rsh = P2 ? __ffs(rsb) - 1 : 0;
  auto row_bytes = [&](uint32_t r) -> uint32_t { return P2 ? (r << rsh) : r * rsb; };
  double dacc = 0.0;
  const bool use_h = flags & kHasH;

  const int64_t nblk = gridDim.x;
  const int64_t p = lay.xcd ? xcd_block(blockIdx.x, nblk) : (int64_t)blockIdx.x;
  const int64_t g_end = (p + 1) * ngroups / nblk;  // balanced contiguous ranges
  for (int64_t grp = p * ngroups / nblk; grp < g_end; ++grp) {
    const int64_t row = grp * lay.rows_per_block + rloc;
    if (!lane_ok || row >= D) continue;
    const int64_t off = row * lay.rs + cofs;
    vt xv, b3v;
    step_load<T, V>(flags, off, X, B3, xv, b3v);
    vt h = (vt)(T)0;
    if (use_h) {
      const int64_t e0 = prow[row], e1 = prow[row + 1];
      const double drow = diag[row];
      vt tot = (vt)(T)0;
      for (int64_t e = e0; e < e1; e += NB) {
        uint32_t en[NB];
        if constexpr (GM >= 2) {
          constexpr int UB = NB / GM;
          uint32_t mine[UB];
#pragma unroll
          for (int u = 0; u < UB; ++u) {
            const int64_t ee = e + g + GM * u;
            mine[u] = ee < e1 ? pent[ee] : 0u;
          }
          group_share<GM, UB>(mine, en);
        } else {
#pragma unroll
          for (int u = 0; u < NB; ++u) en[u] = e + u < e1 ? pent[e + u] : 0u;
        }
        vt x[NB];
        T w[NB];
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          x[u] = (vt)(T)0;
          w[u] = (T)0;
          if (e + u < e1) {
            // 32-bit byte offset from the block base (D*rs*sizeof(T) < 2^32 is
            // checked on the host)
            x[u] = *reinterpret_cast<const vt*>(b2b + row_bytes(en[u] & cmask) + cofs_b);
            w[u] = sV[(en[u] >> 24) & vmask];
          }
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) tot += w[u] * x[u];
      }
      const T d = (T)((drow - af.shift) * af.scale);
      h = d * *reinterpret_cast<const vt*>(b2b + row_bytes((uint32_t)row) + cofs_b) + tot;
    }
    step_store<T, V>(flags, row, off, out_col0, cx, ch, cb, h, xv, b3v, OUT, out_ld, nullptr, Pr,
                     dacc);
  }
  if (flags & kDot) {
    const double t = block_sum_256(dacc, red);
    if (threadIdx.x == 0) dot_part[blockIdx.x] = t;
  }
}

template <typename T, int V>
static void launch_packed_typed(pf_op* op, int flags, const Layout& lay, int ncols_grid,
                                Affine af, double cx, double ch, double cb, const void* X,
                                const void* B2, const void* B3, void* out, int64_t out_ld,
                                hipStream_t st, double* dot_part, const void* P,
                                int dot_blocks) {
  PF_REQUIRE((uint64_t)op->D * (uint64_t)lay.rs * sizeof(T) < ((uint64_t)1 << 32),
             "block too large for 32-bit byte offsets");
  const int64_t ngroups = (op->D + lay.rows_per_block - 1) / lay.rows_per_block;
  const int64_t gx = step_grid_x(op, lay, flags, dot_blocks);
  PF_REQUIRE(gx < (int64_t)1 << 31, "grid too large");
  dim3 grid((unsigned)gx, (unsigned)ncols_grid), block(256);
  if (lay.nt) flags |= kStreamNT;
  const uint32_t rsb = (uint32_t)(lay.rs * sizeof(T));
  const bool p2 = (rsb & (rsb - 1u)) == 0;
#define PF_PACKED_LAUNCH(P2V, GMV)                                                               \
  hipLaunchKernelGGL((packed_step_kernel<T, V, P2V, GMV>), grid, block, 0, st, op->D, op->prow,  \
                     op->pent, op->diag, op->pvals, op->pnvals, af, flags, (T)cx, (T)ch, (T)cb,   \
                     reinterpret_cast<const T*>(X), reinterpret_cast<const T*>(B2),               \
                     reinterpret_cast<const T*>(B3), out, lay, out_ld, dot_part,                  \
                     reinterpret_cast<const T*>(P), ngroups)
#define PF_PACKED_G(P2V)                          \
  switch (lay.G) {                                \
    case 2: PF_PACKED_LAUNCH(P2V, 2); break;      \
    case 4: PF_PACKED_LAUNCH(P2V, 4); break;      \
    case 8: PF_PACKED_LAUNCH(P2V, 8); break;      \
    default: PF_PACKED_LAUNCH(P2V, 1); break;     \
  }
  if (p2) {
    PF_PACKED_G(true);
  } else {
    PF_PACKED_G(false);
  }
#undef PF_PACKED_G
#undef PF_PACKED_LAUNCH
  PF_HIP(hipGetLastError());
}

void launch_packed(pf_op* op, int dtype, int flags, const Layout& lay, int ncols_grid, Affine af,
                   double cx, double ch, double cb, const void* X, const void* B2, const void* B3,
                   void* out, int64_t out_ld, hipStream_t st, double* dot_part, const void* P,
                   int dot_blocks) {
#define PF_PK(TT, VV)                                                                       \
  launch_packed_typed<TT, VV>(op, flags, lay, ncols_grid, af, cx, ch, cb, X, B2, B3, out, \
                              out_ld, st, dot_part, P, dot_blocks)
  if (dtype == PF_F64) {
    switch (lay.V) {
      case 1: PF_PK(double, 1); break;
      case 2: PF_PK(double, 2); break;
      case 4: PF_PK(double, 4); break;
      case 8: PF_PK(double, 8); break;
      case 16: PF_PK(double, 16); break;
      default: throw Error{PF_ERR_UNSUPPORTED, "unsupported vector width"};
    }
  } else if (dtype == PF_F32) {
    switch (lay.V) {
      case 1: PF_PK(float, 1); break;
      case 2: PF_PK(float, 2); break;
      case 4: PF_PK(float, 4); break;
      case 8: PF_PK(float, 8); break;
      case 16: PF_PK(float, 16); break;
      case 32: PF_PK(float, 32); break;
      default: throw Error{PF_ERR_UNSUPPORTED, "unsupported vector width"};
    }
  } else {
    throw Error{PF_ERR_INVALID, "dtype must be PF_F64 or PF_F32"};
  }
#undef PF_PK
}

// Per-row entries (column, unscaled value) and diagonal of rows row0 + t.
__global__ void packed_conn_kernel(const int64_t* __restrict__ prow,
                                   const uint32_t* __restrict__ pent,
                                   const double* __restrict__ pvals, const int nvals,
                                   const double* __restrict__ dg, int64_t row0, int64_t nrows,
                                   int maxper, int64_t* __restrict__ cols,
                                   double* __restrict__ vals, int32_t* __restrict__ counts,
                                   double* __restrict__ dout) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrows) return;
  const int64_t r = row0 + t;
  const uint32_t cmask = nvals > 1 ? 0x00FFFFFFu : 0xFFFFFFFFu;
  const uint32_t vmask = nvals > 1 ? 0xFFu : 0u;
  const int64_t e0 = prow[r], e1 = prow[r + 1];
  int cnt = 0;
  for (int64_t e = e0; e < e1; ++e, ++cnt) {
    if (cnt < maxper) {
      cols[t * maxper + cnt] = (int64_t)(pent[e] & cmask);
      vals[t * maxper + cnt] = pvals[(pent[e] >> 24) & vmask];
    }
  }
  counts[t] = cnt;
  dout[t] = dg[r];
}

void packed_connectivity(pf_op* op, int64_t row0, int64_t nrows, int32_t max_per_row,
                         int64_t* cols, double* vals, int32_t* counts, double* diag,
                         hipStream_t st) {
  hipLaunchKernelGGL(packed_conn_kernel, dim3((unsigned)((nrows + 255) / 256)), dim3(256), 0, st,
                     op->prow, op->pent, op->pvals, op->pnvals, op->diag, row0, nrows,
                     max_per_row, cols, vals, counts, diag);
  PF_HIP(hipGetLastError());
  PF_HIP(hipStreamSynchronize(st));
}

}  // namespace pf

using namespace pf;

extern "C" int pf_op_create_packed(pf_handle* h, int64_t D, const double* diag, int32_t nvals,
                                   const double* vals, const int64_t* starts,
                                   const int32_t* flat, int64_t nflat, pf_op** out) {
  return guarded([&] {
    PF_REQUIRE(h && out && diag, "packed: null argument");
    PF_REQUIRE(D >= 1 && D <= (int64_t)0xFFFFFFFF, "packed: 1 <= D < 2^32");
    PF_REQUIRE(nvals >= 0 && nvals <= kMaxPackedVals, "packed: at most 256 distinct values");
    PF_REQUIRE(nvals == 0 || (vals && starts && (flat || nflat == 0)), "packed: null value arrays");
    PF_REQUIRE(nflat >= 0, "packed: nflat >= 0");
    PF_REQUIRE(nvals <= 1 || D <= ((int64_t)1 << 24),
               "packed: more than one distinct value needs D <= 2^24");
    *out = nullptr;
    // per row: its entries of value 0, then of value 1, ... (columns as given)
    std::vector<int64_t> rowptr(D + 1, 0);
    for (int v = 0; v < nvals; ++v) {
      const int64_t* sv = starts + (int64_t)v * (D + 1);
      PF_REQUIRE(sv[0] >= 0 && sv[D] <= nflat, "packed: starts outside flat[0, nflat)");
      for (int64_t r = 0; r < D; ++r) {
        PF_REQUIRE(sv[r + 1] >= sv[r], "packed: starts must be non-decreasing");
        rowptr[r + 1] += sv[r + 1] - sv[r];
      }
    }
    for (int64_t r = 0; r < D; ++r) rowptr[r + 1] += rowptr[r];
    const int64_t nnz = rowptr[D];
    std::vector<uint32_t> ent(std::max<int64_t>(nnz, 1));
    std::vector<int64_t> fill(rowptr.begin(), rowptr.end() - 1);
    for (int v = 0; v < nvals; ++v) {
      const int64_t* sv = starts + (int64_t)v * (D + 1);
      for (int64_t r = 0; r < D; ++r)
        for (int64_t e = sv[r]; e < sv[r + 1]; ++e) {
          const int64_t c = flat[e];
          PF_REQUIRE(c >= 0 && c < D, "packed: column out of range");
          ent[fill[r]++] = (uint32_t)c | (nvals > 1 ? (uint32_t)v << 24 : 0u);
        }
    }
    PF_HIP(hipSetDevice(h->device));
    pf_op* op = new pf_op();
    op->h = h;
    op->kind = kPacked;
    op->D = D;
    op->pnvals = nvals;
    op->pnnz = nnz;
    try {
      PF_HIP(hipMalloc(&op->diag, D * sizeof(double)));
      PF_HIP(hipMalloc(&op->prow, (D + 1) * sizeof(int64_t)));
      PF_HIP(hipMalloc(&op->pent, ent.size() * sizeof(uint32_t)));
      PF_HIP(hipMalloc(&op->pvals, std::max(nvals, 1) * sizeof(double)));
      PF_HIP(hipMemcpy(op->diag, diag, D * sizeof(double), hipMemcpyHostToDevice));
      PF_HIP(hipMemcpy(op->prow, rowptr.data(), (D + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
      PF_HIP(hipMemcpy(op->pent, ent.data(), ent.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
      if (nvals)
        PF_HIP(hipMemcpy(op->pvals, vals, nvals * sizeof(double), hipMemcpyHostToDevice));
    } catch (...) {
      pf_op_destroy(op);
      throw;
    }
    *out = op;
  });
}
