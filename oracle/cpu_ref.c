/*
 * CPU restatement of the reference's packed-operator Clenshaw filter
 * (TEST INFRASTRUCTURE: parity checker and the bench's CPU baseline only).
 *
 * Follows src/Optimization/cpu/single_row_mappings.jl:9-79 (mapping_state_i:
 * y_i = d_i x_i + sum_v v * sum_{j in N_v(i)} x_j, neighbours in ascending
 * column order), src/Optimization/cpu/clenshaw_recurrence_relation.jl:104-145
 * (b1 = c X + 2 H b2 - b3), src/Optimization/cpu/clenshaw_final_sum.jl:25-60
 * (Y = c X + H b1 - b2) and src/Clenshaw/Algorithm/ClenshawAlgorithm.jl:12-29
 * (k = K..1, buffer rotation (3,1,2)).  Columns are filtered independently on
 * separate threads with per-column buffers, as MulColsParallel(1) does
 * (src/Polfed/Algorithm/SpectralTransformation/ParallelizationClenshaw.jl:77-95).
 * Indices are 0-based int64 (the reference's flat/starts are 1-based Int).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* y = H x for one column (packed buckets). */
static void packed_map(int64_t D, const double* diag, int nb, const double* vals,
                       const int64_t* const* flat, const int64_t* const* starts,
                       const int64_t* flat_len, const double* x, double* y) {
  for (int64_t i = 0; i < D; ++i) {
    double yi = diag[i] * x[i];
    for (int b = 0; b < nb; ++b) {
      const int64_t s0 = starts[b][i];
      const int64_t s1 = (i == D - 1) ? flat_len[b] : starts[b][i + 1];
      double acc = 0.0;
      for (int64_t p = s0; p < s1; ++p) acc += x[flat[b][p]];
      yi += vals[b] * acc;
    }
    y[i] = yi;
  }
}

/* one recurrence row-loop: b1 = c*X + 2*H b2 - b3 (single-bucket fused form,
 * multi-bucket via a mapped temporary like the reference). */
static void recurrence(int64_t D, const double* diag, int nb, const double* vals,
                       const int64_t* const* flat, const int64_t* const* starts,
                       const int64_t* flat_len, double c, const double* X, const double* b2,
                       const double* b3, double* b1) {
  if (nb == 1) {
    for (int64_t i = 0; i < D; ++i) {
      const int64_t s0 = starts[0][i];
      const int64_t s1 = (i == D - 1) ? flat_len[0] : starts[0][i + 1];
      double acc = 0.0;
      for (int64_t p = s0; p < s1; ++p) acc += b2[flat[0][p]];
      const double yi = diag[i] * b2[i] + vals[0] * acc;
      b1[i] = c * X[i] + 2.0 * yi - b3[i];
    }
  } else {
    packed_map(D, diag, nb, vals, flat, starts, flat_len, b2, b1);
    for (int64_t i = 0; i < D; ++i) b1[i] = c * X[i] + 2.0 * b1[i] - b3[i];
  }
}

/*
 * Y[:, j] = p(H) X[:, j] for j < ncols, p = sum_k coeffs[k] T_k (Chebyshev),
 * X, Y column-major with leading dimension D.  nthreads <= 0: all threads.
 * Returns 0 on success.
 */
int cpu_clenshaw_filter(int64_t D, const double* diag, int nb, const double* vals,
                        const int64_t* const* flat, const int64_t* const* starts,
                        const int64_t* flat_len, const double* coeffs, int K, const double* X,
                        double* Y, int ncols, int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  int err = 0;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ncols; ++j) {
    double* buf = (double*)malloc(sizeof(double) * 3 * (size_t)D);
    if (!buf) {
      err = 1;
      continue;
    }
    double* b[3] = {buf, buf + D, buf + 2 * D};
    memset(buf, 0, sizeof(double) * 3 * (size_t)D);
    const double* x = X + (int64_t)j * D;
    for (int k = K; k >= 1; --k) {
      recurrence(D, diag, nb, vals, flat, starts, flat_len, coeffs[k], x, b[1], b[2], b[0]);
      double* t = b[0];
      b[0] = b[2];
      b[2] = b[1];
      b[1] = t;
    }
    double* y = Y + (int64_t)j * D;
    packed_map(D, diag, nb, vals, flat, starts, flat_len, b[1], y);
    for (int64_t i = 0; i < D; ++i) y[i] = coeffs[0] * x[i] + y[i] - b[2][i];
    free(buf);
  }
  return err;
}

/* Row-chunk threaded variant for a single column (MulColsParallel(nt) with
 * nt > 1 splits each column into row chunks, make_loop.jl:45-63). */
int cpu_clenshaw_filter_rows(int64_t D, const double* diag, int nb, const double* vals,
                             const int64_t* const* flat, const int64_t* const* starts,
                             const int64_t* flat_len, const double* coeffs, int K,
                             const double* x, double* y, int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  double* buf = (double*)calloc(3 * (size_t)D, sizeof(double));
  if (!buf) return 1;
  double* b[3] = {buf, buf + D, buf + 2 * D};
  for (int k = K; k >= 1; --k) {
    const double c = coeffs[k];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < D; ++i) {
      double yi = diag[i] * b[1][i];
      for (int bb = 0; bb < nb; ++bb) {
        const int64_t s0 = starts[bb][i];
        const int64_t s1 = (i == D - 1) ? flat_len[bb] : starts[bb][i + 1];
        double acc = 0.0;
        for (int64_t p = s0; p < s1; ++p) acc += b[1][flat[bb][p]];
        yi += vals[bb] * acc;
      }
      b[0][i] = c * x[i] + 2.0 * yi - b[2][i];
    }
    double* t = b[0];
    b[0] = b[2];
    b[2] = b[1];
    b[1] = t;
  }
  packed_map(D, diag, nb, vals, flat, starts, flat_len, b[1], y);
  for (int64_t i = 0; i < D; ++i) y[i] = coeffs[0] * x[i] + y[i] - b[2][i];
  free(buf);
  return 0;
}
