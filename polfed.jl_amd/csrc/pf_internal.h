// Internal declarations shared by the translation units of libpolfed_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/polfed_hip.h"
#include "pf_rank.h"

namespace pf {

// ------------------------------------------------------------------ errors
void set_error(const std::string& msg);
void clear_error();

struct Error {
  int code;
  std::string msg;
};

#define PF_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw ::pf::Error{PF_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
  } while (0)

#define PF_REQUIRE(cond, msg)                                         \
  do {                                                                \
    if (!(cond)) throw ::pf::Error{PF_ERR_INVALID, std::string(msg)}; \
  } while (0)

// Runs f, converting exceptions into a status code + thread-local message.
template <class F>
int guarded(F&& f) {
  try {
    f();
    clear_error();
    return PF_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return PF_ERR_NOMEM;
  }
}

inline hipStream_t S(pf_stream s) { return reinterpret_cast<hipStream_t>(s); }

// --------------------------------------------------- device scratch buffer
// Grow-only device allocation owned by a handle or operator.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void* get(size_t need) {
    if (need > bytes) {
      if (p) PF_HIP(hipFree(p));
      p = nullptr;
      bytes = 0;
      if (hipMalloc(&p, need) != hipSuccess) {
        p = nullptr;
        (void)hipGetLastError();
        throw Error{PF_ERR_NOMEM, "hipMalloc of " + std::to_string(need) + " bytes failed"};
      }
      bytes = need;
    }
    return p;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// ----------------------------------------------------------- operator data
constexpr int kMaxBonds = 128;
constexpr int kMaxLoc = 27;
constexpr int kMaxGrain = 4;

// Spin-chain operator parameters, passed by value (kernel argument segment).
struct ChainParams {
  int32_t L, Lup, nb;
  int64_t D;
  // bonds with a nonzero off-diagonal value, grouped by value (ascending)
  int8_t lo[kMaxBonds], hi[kMaxBonds];  // 0-based bit positions, lo < hi
  double bval[kMaxBonds];               // unscaled off-diagonal value J/2 of the bond
  // Two-table rank: rank(s) = TA[s & (2^m - 1)] + TB[s >> m] (m = tsplit,
  // 0 = tables not built, the kernels then use the O(1) neighbour update).
  int32_t tsplit;
};

// Largest L whose rank tables are staged in LDS (2^m u16 + 2^(L-m) u32).
constexpr int kMaxTableL = 26;

// Quantum-Sun operator parameters (spin 1/2, full space).
struct QsunParams {
  int32_t Lg, Lloc;
  int64_t D;
  uint32_t flip[kMaxLoc];  // XOR mask of outer spin j and its grain partner
  double coup[kMaxLoc];    // unscaled off-diagonal value g0 * couplings[j] / 4
  double h[kMaxLoc];       // z-fields of the outer spins
  double G[(1 << kMaxGrain) * (1 << kMaxGrain)];  // grain matrix, G[a*dg + b]
};

enum OpKind { kChain = 0, kQsun = 1, kPacked = 2 };

// Packed operator: at most this many distinct off-diagonal values when there
// is more than one (the value index sits in the top 8 bits of an entry).
constexpr int kMaxPackedVals = 256;

// Affine map of the operator: H' = (H - shift) * scale (scale = 1/a, shift = b
// for the rescaled operator; 1, 0 for the plain one).
struct Affine {
  double scale, shift;
};

// Kernel flag bits of the fused step out = cx*X + ch*(H' b2) + cb*b3.
enum : int {
  kHasX = 1,    // read X
  kHasH = 2,    // apply the operator to b2
  kHasB3 = 4,   // read b3
  kOutCM = 8,   // write the output column-major (FP64) instead of interleaved
  kDot = 16,    // accumulate <P, out> per block (KPM)
  kStreamNT = 32,  // non-temporal hints on the streamed vectors
  kRowPart = 64,   // spin chains: rows in the XCD partition order (drivers only)
};

struct Layout {
  int64_t rs;      // row stride (s for interleaved, 1 for column-major)
  int64_t cstride; // column stride for blockIdx.y (column-major), 0 for interleaved
  int G;           // lanes per row
  int rows_per_block;
  int V;           // elements per lane
  int xcd;         // 1: map consecutive row groups to the same XCD (L2 locality)
  int nt;          // 1: non-temporal hints on streamed loads/stores
  int pblocks;     // persistent grid: blocks per CU (0: one block per row group)
  int wide;        // 1: twice the neighbour gathers in flight per batch (spin chains)
  int shfl;        // 1: Quantum-Sun grain block by cross-lane shuffles (else reloads)
  int occ;         // 1: spin-chain kernel held to 8 waves per SIMD (64 VGPRs)
  int part;        // 1: spin-chain drivers run rows in the XCD partition order
};

// Launch configuration of the row-interleaved step kernel for one (dtype, s).
struct TuneCfg {
  int V = 0;    // elements per lane (0 = default)
  int xcd = -1; // XCD-contiguous row mapping (-1 = default)
  int nt = -1;  // non-temporal streamed accesses (-1 = default)
  int pb = -1;  // persistent blocks per CU (-1 = default)
  int wide = -1;  // wide gather batches (-1 = default)
  int shfl = -1;  // Quantum-Sun grain shuffles (-1 = default)
  int occ = -1;   // spin-chain 8-waves-per-SIMD variant (-1 = default)
  int part = -1;  // XCD partition row order in the drivers (-1 = default)
};

}  // namespace pf

struct pf_handle {
  int device = 0;
  hipDeviceProp_t prop{};
  pf::DevBuf partials;  // gram partial sums
  pf::DevBuf small;     // small matrices staged to the device
  pf::DevBuf vecs;      // residual / rotation scratch
};

struct pf_op {
  pf_handle* h = nullptr;
  pf::OpKind kind = pf::kChain;
  int64_t D = 0;
  double Emin = -1.0, Emax = 1.0;
  pf::ChainParams chain{};
  pf::QsunParams qsun{};
  uint32_t* states = nullptr;  // spin chain: ascending basis states
  uint32_t* binom = nullptr;   // spin chain: C(p, t), p < 32, t < kRankT
  double* diag = nullptr;      // spin chain: unscaled diagonal, reference summation order
  double* qdiag = nullptr;     // Quantum Sun: Dlo[2^qlb] then Dhi[2^(Lg+Lloc-qlb)] (diagonal)
  // Packed operator (diag in `diag`): row r's off-diagonal entries are
  // pent[prow[r] .. prow[r+1]), ordered by value index then column; an entry is
  // the column, with the value index in bits 24..31 when pnvals > 1.
  int64_t* prow = nullptr;
  uint32_t* pent = nullptr;
  double* pvals = nullptr;     // the distinct off-diagonal values, ascending (unscaled)
  int pnvals = 0;
  int64_t pnnz = 0;
  int qlb = 0;                 // Quantum Sun: low bits of the diagonal split
  void* rtab = nullptr;        // spin chain: u16 TA[2^m] then u32 TB[2^(L-m)] (rank tables)
  size_t rtab_bytes = 0;       // bytes of rtab (dynamic LDS of the step kernel)
  // Spin chain with rank tables: the same basis in the XCD partition order
  // (segments of equal high bits in reverse Cuthill-McKee order), used by the
  // row-interleaved drivers; the hook entry points keep the reference order.
  uint32_t* pstates = nullptr;  // basis states in partition order
  double* pdiag = nullptr;      // diagonal in partition order
  void* prtab = nullptr;        // TA, then TB of the partition order
  uint32_t* pperm = nullptr;    // partition row -> reference rank
  pf::DevBuf fbuf;             // filter / KPM row-interleaved work blocks
  pf::TuneCfg force;           // pf_op_set_tuning overrides
  std::map<std::pair<int, int>, pf::TuneCfg> tuned;  // autotuned (dtype, s) configs
  pf::DevBuf kpart;            // KPM partial dots
  std::vector<double> khost;   // KPM host staging
  pf::Affine affine(bool rescaled) const {
    if (!rescaled) return {1.0, 0.0};
    const double a = (Emax - Emin) / 2, b = (Emax + Emin) / 2;
    return {1.0 / a, b};
  }
};

namespace pf {

// Launch one fused operator step over `ncols` columns (ops.hip).
// out = cx*X + ch*(H' b2) + cb*b3 with H' = (H - shift)*scale.
// T is the arithmetic type of X, b2, b3 (and of out unless kOutCM, which
// writes FP64 column-major to `outp` with leading dimension out_ld).
void launch_step(pf_op* op, int dtype, int flags, const Layout& lay, int ncols, Affine af,
                 double cx, double ch, double cb, const void* X, const void* B2, const void* B3,
                 void* out, int64_t out_ld, hipStream_t st, double* dot_part = nullptr,
                 const void* P = nullptr, int dot_blocks = 0);

Layout interleaved_layout(int dtype, int s, int V = 0, int xcd = 1, int nt = 1, int pb = 0,
                          int wide = 0, int shfl = 1, int occ = 0, int part = 0);
Layout tuned_layout(pf_op* op, int dtype, int s);
Layout colmajor_layout(int64_t ld);

// Blocks of one step launch of the matrix-free or packed kernels (ops.hip):
// one per row group, capped by the persistent grid and the KPM partial slots.
int64_t step_grid_x(const pf_op* op, const Layout& lay, int flags, int dot_blocks);

// One fused step of a packed operator (packed.hip); arguments as launch_step.
void launch_packed(pf_op* op, int dtype, int flags, const Layout& lay, int ncols_grid, Affine af,
                   double cx, double ch, double cb, const void* X, const void* B2, const void* B3,
                   void* out, int64_t out_ld, hipStream_t st, double* dot_part, const void* P,
                   int dot_blocks);

// Looks up this device's address of pf_filter_noalloc's scratch array once per
// device (ops.hip); called by pf_create with the device current.
void noalloc_scratch_init(int device);

// Connectivity rows of a packed operator (packed.hip).
void packed_connectivity(pf_op* op, int64_t row0, int64_t nrows, int32_t max_per_row,
                         int64_t* cols, double* vals, int32_t* counts, double* diag,
                         hipStream_t st);

// Row-interleaved pack: dst[r*s + c] = (T)(src[perm[r] + c*ld] * colscale[c])
// (colscale may be null; perm null is the identity).
void pack_block(int dtype, const double* src, int64_t ld, void* dst, int64_t D, int s,
                const double* colscale_dev, hipStream_t st, const uint32_t* perm = nullptr);

}  // namespace pf
