/*
 * polfed_hip.h -- C ABI of libpolfed_hip, the MI355X (gfx950) hot path of the
 * POLFED eigensolver.
 *
 * Conventions
 *   - Vector blocks are column-major D x ncols with a leading dimension
 *     (element (i, j) at ptr[i + j*ld]), FP64 unless a dtype argument says
 *     otherwise, exactly Julia's Matrix{Float64} layout.  All such pointers are
 *     DEVICE pointers owned by the caller; the library only borrows them for the
 *     duration of the call.  Host pointers are marked "host".
 *   - Every call is ordered on the given stream (NULL = default stream).  Calls
 *     that return host values synchronize that stream before returning.
 *   - Return value: PF_OK (0) or a negative PF_ERR_* code; the message of the
 *     last failure on the calling thread is available from pf_last_error().
 *   - Handles are not thread-safe; use one handle per host thread.
 *
 * The reference (Julia, RockClimbingRocks/Polfed.jl) has no C ABI: each entry
 * point below names the reference interface it replaces.
 */
#ifndef POLFED_HIP_H
#define POLFED_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pf_handle pf_handle;  /* device context + scratch + autotune cache */
typedef struct pf_op pf_op;          /* matrix-free Hamiltonian on one device     */
typedef void* pf_stream;             /* hipStream_t                               */
typedef struct pf_batch pf_batch;    /* many operators filtered in one launch     */

enum {
  PF_OK = 0,
  PF_ERR_INVALID = -1,
  PF_ERR_NOMEM = -2,
  PF_ERR_HIP = -3,
  PF_ERR_UNSUPPORTED = -4,
  PF_ERR_NUMERIC = -5
};

enum { PF_F64 = 0, PF_F32 = 1 };

/* polynomial families of src/Clenshaw/ClenshawMapping.jl:7-12 */
enum { PF_CHEBYSHEV = 0, PF_LEGENDRE = 1, PF_HERMITE = 2, PF_TAYLOR = 3 };

/* KPM probe kinds: normalized Gaussian (src/DensetiesOfStates/Moments.jl:66),
 * or real random phase +-1 (normalized to 1/sqrt(D)). */
enum { PF_PROBE_GAUSSIAN = 0, PF_PROBE_PHASE = 1 };

/* ---------------------------------------------------------------- context */

/* Library version (major*10000 + minor*100 + patch). */
int pf_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* pf_last_error(void);

/* Replaces the CUDA backend registry fill of ext/PolfedCUDAExt.jl:19-40
 * (src/Polfed.jl:34-72): binds the library to HIP device `device`. */
int pf_create(pf_handle** out, int device);
int pf_destroy(pf_handle* h);

/* ---------------------------------------------------------------- operators */

/* Matrix-free spin-1/2 chain in the fixed-popcount (Lup) sector.
 * Replaces the SparseMatrixCSC built by _spin_half_hamiltonian
 * (src/Models/SpinHalfChains.jl:267-315) and its packed form
 * (src/Optimization/optimization.jl:134-182); covers xxz_hamiltonian (:74-99)
 * and j1j2_hamiltonian (:189-220).  Sites are 1-based like the reference's bond
 * tuples (site k <-> bit k-1); `fields` has L entries (host arrays).
 * Basis order = ascending integers (combinadic rank), bit-exact with the
 * reference's Dict index.  Requires 2 <= L <= 32, 0 <= Lup <= 18 (the
 * binomial table's width), nbonds <= 128 and C(L, Lup) < 2^32; any number of
 * distinct couplings (one value per bond). */
int pf_op_create_spin_chain(pf_handle* h, int32_t L, int32_t Lup, int32_t nbonds,
                            const int32_t* site_i, const int32_t* site_j,
                            const double* J, const double* Delta, const double* fields,
                            pf_op** out);

/* Matrix-free Quantum Sun Hamiltonian in the full 2^(L_grain+L_loc) space for
 * spin 1/2.  Replaces build_hamiltonian (src/QSun/QuantumSun.jl:1-102): grain
 * = low L_grain bits, H_grain is the dg x dg (dg = 2^L_grain) grain matrix
 * already scaled by gamma/sqrt(dg+1) (column-major, host), neighbors[j] in
 * [0, L_grain) (0-based grain site), couplings[j] = alpha^u_j, fields[j] the
 * z-field of outer spin j (host arrays of length L_loc).  L_grain <= 4,
 * L_grain + L_loc <= 31. */
int pf_op_create_qsun(pf_handle* h, int32_t L_grain, int32_t L_loc, double g0,
                      const double* H_grain, const int32_t* neighbors,
                      const double* couplings, const double* fields, pf_op** out);

/* Packed sparse operator: any real symmetric matrix in the reference's
 * value-bucket form, the output of get_diags_and_offdiagonals_by_value
 * (src/Optimization/optimization.jl:134-182) that optimize_spectral_transform
 * (:75-119) builds for polfed(mat, ...) and that optimized_mapping! applies
 * (src/Optimization/cpu/single_row_mappings.jl:66-79): diag[D]; nvals distinct
 * off-diagonal values vals[nvals] (ascending); for value v and row r the
 * columns flat[starts[v*(D+1)+r] .. starts[v*(D+1)+r+1]) (0-based columns,
 * absolute offsets into the concatenated flat array, one (D+1)-entry start
 * table per value; the reference keeps one 1-based flat/starts pair per
 * value); nflat = length of flat (every start must lie in [0, nflat]).  Host
 * arrays.  nvals <= 256; more than one value needs D <= 2^24. */
int pf_op_create_packed(pf_handle* h, int64_t D, const double* diag, int32_t nvals,
                        const double* vals, const int64_t* starts, const int32_t* flat,
                        int64_t nflat, pf_op** out);

int pf_op_destroy(pf_op* op);

/* Hilbert-space dimension D of the operator (0 on error). */
int64_t pf_op_dim(const pf_op* op);

/* Set the rescaling H~ = (H - b)/a with a = (Emax-Emin)/2, b = (Emax+Emin)/2
 * (build_mapping_plan, src/Polfed/Plan/Plan.jl:170-177; optimize_spectral_transform,
 * src/Optimization/optimization.jl:98-103).  Default: Emin = -1, Emax = 1. */
int pf_op_set_rescale(pf_op* op, double Emin, double Emax);

/* Export the connectivity the kernels compute on the device, for rows
 * [row0, row0+nrows): for each row, up to max_per_row off-diagonal (column,
 * value) pairs in the kernels' bond order and the diagonal entry (unscaled H).
 * The columns come from the step kernels' own rank function (the two rank
 * tables when L <= 26, else the O(1) combinadic update) and the diagonal from
 * the table they stream (Quantum Sun: the split Dlo + Dhi table).  Device
 * output buffers: cols/vals are nrows x max_per_row row-major, counts nrows,
 * diag nrows.  This is the parity surface for the integer rank/connectivity
 * work (reference: rows/cols/vals of SpinHalfChains.jl:284-311,
 * QuantumSun.jl:1-102). */
int pf_op_connectivity(pf_op* op, int64_t row0, int64_t nrows, int32_t max_per_row,
                       int64_t* cols, double* vals, int32_t* counts, double* diag,
                       pf_stream stream);

/* pf_op_connectivity with a row order: 0 = the reference order (what
 * pf_apply, the Clenshaw hooks and the drivers without the partition order
 * apply), 1 = the spin-chain XCD partition order the filter and KPM drivers
 * may run (its rank tables, states and diagonal; rows and columns are mapped
 * back to reference ranks, so the output is comparable row for row).  Order 1
 * needs a spin chain with rank tables.  Same reference rows as above. */
int pf_op_connectivity_order(pf_op* op, int32_t order, int64_t row0, int64_t nrows,
                             int32_t max_per_row, int64_t* cols, double* vals, int32_t* counts,
                             double* diag, pf_stream stream);

/* --------------------------------------------------------------- mappings */

/* Y = H~ X (rescaled != 0) or Y = H X (rescaled == 0) on ncols columns.
 * dtype PF_F64 or PF_F32 selects the element type of X and Y.
 * Replaces the f!/f!_rescaled callbacks (MappingConfig.jl:28-33,
 * optimized_mapping!, src/Optimization/cpu/mapping.jl:9-13,
 * src/Optimization/gpu/vector_dispatch.jl:105-111). */
int pf_apply(pf_op* op, int32_t dtype, int32_t rescaled, const void* X, int64_t ldx,
             void* Y, int64_t ldy, int32_t ncols, pf_stream stream);

/* b1 = c*X + alpha*H~b2 + beta*b3 (b3 may be NULL when beta == 0).
 * Replaces the clenshaw_recurrence hook (MappingConfig.jl:35-40;
 * optimized_clenshaw_recurrence_relation!, cpu/clenshaw_recurrence_relation.jl:87-95,
 * gpu/vector_dispatch.jl:113-125; generic form Clenshaw.jl:69-81).  All blocks
 * share leading dimension ld. */
int pf_clenshaw_recurrence(pf_op* op, int32_t dtype, void* b1, const void* b2, const void* b3,
                           double c, double alpha, double beta, const void* X, int64_t ld,
                           int32_t ncols, pf_stream stream);

/* Y = c0*phi0*X + phi1*H~b2 + beta1*phi0*b3.
 * Replaces the clenshaw_finalsum hook (MappingConfig.jl:42-45;
 * optimized_clenshaw_final_sum!, cpu/clenshaw_final_sum.jl:7-15,
 * gpu/vector_dispatch.jl:127-139; generic form Clenshaw.jl:104-117). */
int pf_clenshaw_finalsum(pf_op* op, int32_t dtype, const void* b2, const void* b3, double c0,
                         double beta1, double phi0, double phi1, void* Y, const void* X,
                         int64_t ld, int32_t ncols, pf_stream stream);

/* Whole polynomial filter Y = p(H~) X, p = sum_{n=0..order} coeffs[n] P_n for
 * the family `polytype`, evaluated by the Clenshaw recurrence of
 * clenshaw_algorithm! (src/Clenshaw/Algorithm/ClenshawAlgorithm.jl:12-29).
 * X, Y are FP64 (device); `dtype` is the arithmetic of the recurrence:
 * PF_F64 (the plain drop-in) or PF_F32 (FP32 storage and accumulation).
 * coeffs is a host array of order+1 values (already normalized).
 * Replaces the whole-block filter call clenshaw(transform, W, V, b, ::GPU,
 * ::NoParallel) (ParallelizationClenshaw.jl:41-51) reached from
 * f!_transformed (DefineTransformation.jl:42-45). */
int pf_filter(pf_op* op, int32_t dtype, int32_t polytype, const double* coeffs, int32_t order,
              const double* X, int64_t ldx, double* Y, int64_t ldy, int32_t ncols,
              pf_stream stream);

/* pf_filter without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * kernels and results as pf_filter; the four row-interleaved work blocks of a
 * column chunk live in a fixed per-GPU device array of 128 MiB (allocated with
 * the code object, held until the process ends), so 4 * D * min(ncols, 32 for
 * FP64 / 64 for FP32) * element size must be <= 128 MiB, else PF_ERR_INVALID.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (several streams or host threads) race.  Requires pf_create on the device. */
int pf_filter_noalloc(pf_op* op, int32_t dtype, int32_t polytype, const double* coeffs,
                      int32_t order, const double* X, int64_t ldx, double* Y, int64_t ldy,
                      int32_t ncols, pf_stream stream);

/* ------------------------------------------------------------- tuning */

enum {
  PF_TUNE_XCD_REMAP = 1,
  PF_TUNE_VEC = 2,
  PF_TUNE_RESET = 3,
  PF_TUNE_NONTEMPORAL = 4,
  PF_TUNE_PBLOCKS = 5,
  PF_TUNE_WIDE = 6,
  PF_TUNE_GRAIN_SHUFFLE = 7,
  PF_TUNE_OCCUPANCY = 8,
  PF_TUNE_ROW_PARTITION = 9
};

/* Force a launch knob of the row-interleaved step kernels for this operator:
 * PF_TUNE_XCD_REMAP (-1 default, 0 off, 1 on: contiguous row ranges per XCD),
 * PF_TUNE_VEC (0 default, or 1/2/4/8/16/32 elements per lane),
 * PF_TUNE_NONTEMPORAL (-1 default, 0/1: non-temporal hints on the streamed
 * vectors), PF_TUNE_PBLOCKS (-1 default, 0: one block per row group, n: a
 * persistent grid of n blocks per CU, each over a contiguous range of rows),
 * PF_TUNE_WIDE (-1 default, 0/1: twice the neighbour gathers in flight per
 * batch in the spin-chain kernel with rank tables), PF_TUNE_GRAIN_SHUFFLE
 * (-1 default, 0/1: Quantum-Sun grain block by cross-lane shuffles or by
 * reloads), PF_TUNE_OCCUPANCY (-1 default, 0/1: spin-chain kernel built for
 * 8 waves per SIMD), PF_TUNE_ROW_PARTITION (-1 default, 0/1: spin chains
 * with rank tables run pf_filter / pf_kpm_moments on the basis reordered so
 * that the row range each XCD owns gathers mostly its own rows; inputs and
 * outputs stay in the reference order and filter results are bit-identical;
 * the autotuner tries both orders),
 * PF_TUNE_RESET (drop forced and autotuned settings).  No reference
 * counterpart: launch tuning of this library. */
int pf_op_set_tuning(pf_op* op, int32_t key, int32_t value);

/* Autotuner: times `nsteps` fused recurrence steps for each candidate launch
 * configuration on this operator's shape (dtype, ncols columns) and keeps the
 * fastest for later pf_filter / pf_kpm_moments calls.  Outputs (host, may be
 * NULL): best time per step in microseconds, chosen vector width and XCD
 * mapping.  No reference counterpart. */
int pf_autotune(pf_op* op, int32_t dtype, int32_t ncols, int32_t nsteps, double* best_us,
                int32_t* best_vec, int32_t* best_xcd, pf_stream stream);

/* The launch configuration the filter and KPM drivers use for (dtype, ncols)
 * on this operator -- the autotuned one with any pf_op_set_tuning overrides --
 * as cfg[8] = {vector width V, XCD row mapping, non-temporal hints, persistent
 * blocks per CU, wide gather batches, grain shuffles, 8-waves-per-SIMD build,
 * partition row order}.  No counterpart in the reference (an observability
 * helper for profiling the exact kernel instance a solve runs). */
int pf_op_get_tuning(pf_op* op, int32_t dtype, int32_t ncols, int32_t* cfg);

/* ------------------------------------------------------------------- KPM */

/* Fill R probe columns (device, column-major, ld) with a counter-based
 * generator: PF_PROBE_GAUSSIAN standard normals, PF_PROBE_PHASE +-1.
 * Deterministic in (seed, row, column). Not normalized. */
int pf_fill_probes(int32_t kind, uint64_t seed, int64_t D, int32_t R, double* out, int64_t ld,
                   pf_stream stream);

/* KPM moments mu_n = (D/R) sum_r <r|T_n(H~)|r>, n = 0..N-1, each probe
 * normalized first; one operator application per moment per probe, as in
 * trace! (src/DensetiesOfStates/Moments.jl:124-140) and _dos_moments_impl
 * (:58-109).  probes: device D x R column-major FP64 (copied, not modified);
 * dtype selects the arithmetic of the recurrence (PF_F64 / PF_F32; dot
 * products accumulate in FP64).  mu_out: host array of N doubles.
 * Replaces dos_moments (Moments.jl:42-46). */
int pf_kpm_moments(pf_op* op, int32_t dtype, int32_t N, int32_t R, const double* probes,
                   int64_t ldp, double* mu_out, pf_stream stream);

/* --------------------------------------------------- dense FP64 block ops */

/* G = A^T B (m x n, device, column-major, leading dimension ldg); A is D x m,
 * B is D x n.  FP64 MFMA.  Replaces the projections V'*W of
 * reorthogonalization! (Factorization/FactorizationSteps/reorthogonalization.jl:164)
 * and calcoverlap! (BlockLanczosFactorization.jl:102-111). */
int pf_gram(pf_handle* h, const double* A, int64_t lda, int32_t m, const double* B, int64_t ldb,
            int32_t n, int64_t D, double* G, int32_t ldg, pf_stream stream);

/* B = beta*B + alpha*A*G, A D x m, G m x n (device, column-major ldg), B D x n.
 * B must not alias A.  FP64 MFMA.  Replaces W .-= V*(V'*W) (second half),
 * r .-= v1*alpha + v2*beta' (orthogonalization.jl:74-76) and the Ritz vectors
 * mul!(vecs, V, phi) (Convergence/getvecs.jl:51-53). */
int pf_gemm_update(pf_handle* h, const double* A, int64_t lda, int32_t m, const double* G,
                   int32_t ldg, double* B, int64_t ldb, int32_t n, int64_t D, double alpha,
                   double beta, pf_stream stream);

/* Full reorthogonalization W -= V (V^T W) applied `passes` times (2 = CGS2)
 * against the k columns of V.  Replaces reorthogonalization!(::FullRO)
 * (reorthogonalization.jl:12-28, 77-84), which applies one pass. */
int pf_reorth(pf_handle* h, const double* V, int64_t ldv, int32_t k, double* W, int64_t ldw,
              int32_t s, int64_t D, int32_t passes, pf_stream stream);

/* In-place thin QR of the D x s block W by Cholesky-QR applied twice in FP64:
 * W <- Q, R_host (host, s x s column-major, upper triangular) <- R.  Returns
 * PF_ERR_NUMERIC (W untouched) when the Gram matrix is not numerically
 * positive definite, so the caller can fall back to Householder QR.
 * Replaces qr(fact.r) (BlockLanczosFactorization.jl:115-131). */
int pf_block_qr(pf_handle* h, double* W, int64_t ldw, int64_t D, int32_t s, double* R_host,
                pf_stream stream);

/* W <- W * M for a small n x n matrix M (host, column-major), row by row in
 * place (used to apply R^{-1} or a basis rotation). */
int pf_rotate(pf_handle* h, double* W, int64_t ldw, int64_t D, int32_t n, const double* M_host,
              pf_stream stream);

/* Rayleigh quotients E_i = <psi_i|H|psi_i> and residuals ||H psi_i - E_i psi_i||
 * with the UNSCALED operator for nev columns (host outputs).  Replaces
 * calculate_eigenvalues! (Convergence/getvals.jl:10-36) and
 * calculate_eigenresiduals! (Convergence/ConvergenceInfo.jl:139-173). */
int pf_residuals(pf_handle* h, pf_op* op, const double* psi, int64_t ld, int32_t nev,
                 double* E_host, double* res_host, pf_stream stream);

/* Column dot products out[j] = <A_j, B_j> for j < n (host output), FP64. */
int pf_coldots(pf_handle* h, const double* A, int64_t lda, const double* B, int64_t ldb,
               int64_t D, int32_t n, double* out_host, pf_stream stream);

/* ------------------------------------------------------------- batching */

/* A batch of n Quantum-Sun operators (disorder realizations, possibly of
 * different sizes) whose filters run together: each Clenshaw step of every
 * instance is one multi-instance kernel launch, optionally replayed from a
 * captured HIP graph.  Every instance filters `ncols` columns (1..32).  The
 * operators must outlive the batch; their rescaling (pf_op_set_rescale) is
 * read by pf_batch_set_filter.  Calls on one batch must be ordered (one
 * stream, or synchronized).  The reference filters one problem per call
 * (ParallelizationClenshaw.jl:41-51, reached once per Lanczos iteration of
 * each polfed() call); this is the batched form of that call. */
int pf_batch_create(pf_handle* h, int32_t n, pf_op* const* ops, int32_t ncols, pf_batch** out);
int pf_batch_destroy(pf_batch* b);

/* Polynomial of each instance: orders[i] (>= 0) and its orders[i]+1
 * normalized coefficients, concatenated in instance order in `coeffs`
 * (host).  Same meaning as pf_filter's coeffs/order. */
int pf_batch_set_filter(pf_batch* b, int32_t polytype, const int32_t* orders,
                        const double* coeffs);

/* Y[i] = p_i(H~_i) X[i] for every instance with active[i] != 0, in the
 * arithmetic dtypes[i] (PF_F64 / PF_F32).  X[i], Y[i]: device FP64
 * column-major D_i x ncols blocks with leading dimensions ldx[i], ldy[i]
 * (host arrays of n pointers / values).  Replaces, for a set of problems,
 * the per-problem whole-block filter call clenshaw(transform, W, V, b, ::GPU,
 * ::NoParallel) (ParallelizationClenshaw.jl:41-51). */
int pf_batch_filter(pf_batch* b, const int32_t* active, const int32_t* dtypes,
                    const double* const* X, const int64_t* ldx, double* const* Y,
                    const int64_t* ldy, pf_stream stream);

enum {
  PF_BATCH_GRAPH = 1,          /* 0: stream launches, 1: graph replays (default) */
  PF_BATCH_GRAPH_STEPS = 2,    /* step nodes per captured graph (default 96)      */
  PF_BATCH_VEC = 3,            /* columns per lane: 1/2/4/8 dividing ncols        */
  PF_BATCH_ROWS_PER_LANE = 4,  /* rows per lane: 1..64                            */
  PF_BATCH_XCD = 5,            /* 1: XCD-contiguous block mapping (default)       */
  PF_BATCH_GRAIN_SHUFFLE = 6,  /* 1: grain block by cross-lane shuffles (default) */
  PF_BATCH_LDS_ROWS = 7,       /* 1: a block stages its rows in LDS and reads the
                                  grain block and the flips that stay among them
                                  from there; used when ncols is a power of two,
                                  ncols / vec <= 16, rows per lane a power of two
                                  and a multiple of 4 / vec (vec < 4), and the
                                  block's rows (256 / (ncols / vec) x rows per
                                  lane) hold <= 32 KiB in FP64; ignored otherwise
                                  (default 0)                                     */
  PF_BATCH_ROW_VECTORS = 8     /* 1 or 2: with ncols = 1, a lane takes 16 bytes
                                  of consecutive rows (4 FP32 or 2 FP64) per load,
                                  1 or 2 such vectors at a time, and a block
                                  1024 x rows per lane rows, staged in LDS; used
                                  when vec = 1 and rows per lane is 1, 2 or 4 (2
                                  or 4 for the value 2), ignored otherwise
                                  (default 0)                                     */
};
/* Launch settings of the batch.  No reference counterpart. */
int pf_batch_set_tuning(pf_batch* b, int32_t key, int32_t value);

/* The batch's launch settings: cfg[0..9] = columns per lane, rows per lane,
 * XCD mapping, grain shuffles, LDS rows requested, LDS rows in use (the
 * request and the tiling allow it), graph replays, step nodes per graph, row
 * vectors requested, row vectors in use.  No reference counterpart. */
int pf_batch_get_tuning(const pf_batch* b, int32_t* cfg);

/* Times every (columns per lane, rows per lane, XCD mapping, grain-block
 * shuffles or reloads, rows in LDS or not, row vectors or not) tiling of the
 * multi-instance step kernel on all
 * instances in arithmetic `dtype`, then graph replays against stream
 * launches, and keeps the fastest.  best_us
 * (host, may be NULL): time per step of the whole batch.  Overwrites the
 * batch's work buffers.  No reference counterpart. */
int pf_batch_autotune(pf_batch* b, int32_t dtype, int32_t nsteps, double* best_us,
                      pf_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* POLFED_HIP_H */
