/* A plain C caller of libpolfed_hip (no Python): the way a Julia ccall or any
 * C/C++ host uses the drop-in boundary.  On device 0 it builds a spin-1/2 XXZ
 * chain (L = 10, Lup = 5, periodic, fields 0.3 sin(i)) and checks
 *   - pf_apply on unit vectors against the pf_op_connectivity export,
 *   - pf_filter with coefficients (0, 0, 1) against 2 H~(H~ x) - x built from
 *     two rescaled pf_apply calls (T_2(x) = 2x^2 - 1),
 *   - pf_filter_noalloc giving pf_filter's output bit for bit, and
 *     pf_op_get_tuning reporting the launch configuration,
 *   - an invalid operator request returning an error code and a message.
 * Prints "c_abi_check: ok ..." and exits 0 on success. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "../../include/polfed_hip.h"

#define CHECK(call)                                                                   \
  do {                                                                                \
    int rc_ = (call);                                                                 \
    if (rc_ != 0) {                                                                   \
      fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, pf_last_error());          \
      return 1;                                                                       \
    }                                                                                 \
  } while (0)
#define HIPCHECK(call)                                                                \
  do {                                                                                \
    if ((call) != hipSuccess) {                                                       \
      fprintf(stderr, "%s failed\n", #call);                                          \
      return 1;                                                                       \
    }                                                                                 \
  } while (0)

int main(void) {
  const int L = 10, Lup = 5, nb = 10, maxper = 16;
  int32_t si[10], sj[10];
  double J[10], Dl[10], h[10];
  for (int i = 0; i < L; ++i) {
    si[i] = i + 1;             /* 1-based sites, bond (i, i+1 mod L) */
    sj[i] = (i + 1) % L + 1;
    J[i] = 1.0;
    Dl[i] = 0.8;
    h[i] = 0.3 * sin((double)i);
  }
  pf_handle* H = NULL;
  pf_op* op = NULL;
  CHECK(pf_create(&H, 0));
  CHECK(pf_op_create_spin_chain(H, L, Lup, nb, si, sj, J, Dl, h, &op));
  const int64_t D = pf_op_dim(op);
  if (D != 252) {
    fprintf(stderr, "dimension %lld, expected 252\n", (long long)D);
    return 1;
  }

  /* connectivity of every row (host copies) */
  int64_t* dcols;
  double *dvals, *ddiag;
  int32_t* dcnt;
  HIPCHECK(hipMalloc((void**)&dcols, D * maxper * sizeof(int64_t)));
  HIPCHECK(hipMalloc((void**)&dvals, D * maxper * sizeof(double)));
  HIPCHECK(hipMalloc((void**)&dcnt, D * sizeof(int32_t)));
  HIPCHECK(hipMalloc((void**)&ddiag, D * sizeof(double)));
  CHECK(pf_op_connectivity(op, 0, D, maxper, dcols, dvals, dcnt, ddiag, NULL));
  int64_t* cols = malloc(D * maxper * sizeof(int64_t));
  double* vals = malloc(D * maxper * sizeof(double));
  int32_t* cnt = malloc(D * sizeof(int32_t));
  double* dg = malloc(D * sizeof(double));
  HIPCHECK(hipMemcpy(cols, dcols, D * maxper * sizeof(int64_t), hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(vals, dvals, D * maxper * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(cnt, dcnt, D * sizeof(int32_t), hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(dg, ddiag, D * sizeof(double), hipMemcpyDeviceToHost));

  /* H e_r, three unit vectors at once (a D x 3 column-major block) */
  const int rs[3] = {0, 77, 251};
  double *dx, *dy, *dz;
  double* hx = calloc(D * 3, sizeof(double));
  double* hy = malloc(D * 3 * sizeof(double));
  for (int k = 0; k < 3; ++k) hx[rs[k] + k * D] = 1.0;
  HIPCHECK(hipMalloc((void**)&dx, D * 3 * sizeof(double)));
  HIPCHECK(hipMalloc((void**)&dy, D * 3 * sizeof(double)));
  HIPCHECK(hipMalloc((void**)&dz, D * 3 * sizeof(double)));
  HIPCHECK(hipMemcpy(dx, hx, D * 3 * sizeof(double), hipMemcpyHostToDevice));
  CHECK(pf_apply(op, PF_F64, 0, dx, D, dy, D, 3, NULL));
  HIPCHECK(hipMemcpy(hy, dy, D * 3 * sizeof(double), hipMemcpyDeviceToHost));
  double err_apply = 0.0;
  for (int k = 0; k < 3; ++k) {
    /* H symmetric: column r of H = row r of the connectivity */
    double* ref = calloc(D, sizeof(double));
    const int r = rs[k];
    ref[r] += dg[r];
    for (int t = 0; t < cnt[r] && t < maxper; ++t) ref[cols[r * maxper + t]] += vals[r * maxper + t];
    for (int64_t i = 0; i < D; ++i) {
      const double e = fabs(hy[i + k * D] - ref[i]);
      if (e > err_apply) err_apply = e;
    }
    free(ref);
  }

  /* filter T_2 on the rescaled operator against 2 H~(H~ x) - x */
  CHECK(pf_op_set_rescale(op, -8.0, 9.0));
  for (int64_t i = 0; i < D * 3; ++i) hx[i] = cos(0.37 * (double)i) + 0.1;
  HIPCHECK(hipMemcpy(dx, hx, D * 3 * sizeof(double), hipMemcpyHostToDevice));
  const double coeffs[3] = {0.0, 0.0, 1.0};
  CHECK(pf_filter(op, PF_F64, PF_CHEBYSHEV, coeffs, 2, dx, D, dy, D, 3, NULL));
  double* hf = malloc(D * 3 * sizeof(double));
  HIPCHECK(hipMemcpy(hf, dy, D * 3 * sizeof(double), hipMemcpyDeviceToHost));
  /* the same filter without allocation: identical bits */
  CHECK(pf_filter_noalloc(op, PF_F64, PF_CHEBYSHEV, coeffs, 2, dx, D, dz, D, 3, NULL));
  double* hn = malloc(D * 3 * sizeof(double));
  HIPCHECK(hipMemcpy(hn, dz, D * 3 * sizeof(double), hipMemcpyDeviceToHost));
  const int noalloc_ok = memcmp(hn, hf, D * 3 * sizeof(double)) == 0;
  free(hn);
  int32_t cfg[8];
  CHECK(pf_op_get_tuning(op, PF_F64, 3, cfg));
  const int tuning_ok = cfg[0] >= 1 && cfg[0] <= 32 && cfg[3] >= 0;
  CHECK(pf_apply(op, PF_F64, 1, dx, D, dz, D, 3, NULL));
  CHECK(pf_apply(op, PF_F64, 1, dz, D, dy, D, 3, NULL));
  HIPCHECK(hipMemcpy(hy, dy, D * 3 * sizeof(double), hipMemcpyDeviceToHost));
  double err_filter = 0.0, scale = 0.0;
  for (int64_t i = 0; i < D * 3; ++i) {
    const double ref = 2.0 * hy[i] - hx[i];
    if (fabs(ref) > scale) scale = fabs(ref);
    if (fabs(hf[i] - ref) > err_filter) err_filter = fabs(hf[i] - ref);
  }
  err_filter /= scale;

  /* the error path: L > 32 is rejected with a message */
  pf_op* bad = NULL;
  const int rc = pf_op_create_spin_chain(H, 40, 20, nb, si, sj, J, Dl, h, &bad);
  const char* msg = pf_last_error();
  const int err_ok = rc != 0 && bad == NULL && msg != NULL && strlen(msg) > 0;

  hipFree(dx);
  hipFree(dy);
  hipFree(dz);
  hipFree(dcols);
  hipFree(dvals);
  hipFree(dcnt);
  hipFree(ddiag);
  free(hx);
  free(hy);
  free(hf);
  free(cols);
  free(vals);
  free(cnt);
  free(dg);
  CHECK(pf_op_destroy(op));
  CHECK(pf_destroy(H));
  if (err_apply > 1e-14 || err_filter > 1e-14 || !err_ok || !noalloc_ok || !tuning_ok) {
    fprintf(stderr, "c_abi_check: FAILED apply %.3e filter %.3e error-path %d noalloc %d "
            "tuning %d\n", err_apply, err_filter, err_ok, noalloc_ok, tuning_ok);
    return 1;
  }
  printf("c_abi_check: ok D=%lld apply err %.1e filter err %.1e error path \"%s\"\n",
         (long long)D, err_apply, err_filter, msg);
  return 0;
}
