// Dense FP64 block kernels of the block-Lanczos step: tall-skinny Gram
// products A^T B and updates B = beta B + alpha A G on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64), Cholesky-QR, small rotations and residual norms.
//
// MFMA f64 16x16x4 operand layout used here (lane l of a 64-wide wave):
//   A fragment (16 x 4): lane l holds A[l % 16][l / 16]
//   B fragment (4 x 16): lane l holds B[l / 16][l % 16]
//   C/D (16 x 16):       lane l, register v holds D[l/16 + 4*v][l % 16]
// (output layout measured on gfx950 with an in-bounds probe of known inputs).
// The contraction index k only has to agree between the A and B fragments, so
// the Gram kernel assigns k-slot l/16 to rows 4*(l/16)+q of a 16-row slab for
// the q-th of four MFMAs: each lane then loads 4 consecutive rows (32 bytes) of
// one column and 4 lanes cover a full 128-byte line.

#include <algorithm>
#include <cmath>
#include <cstring>

#include "pf_internal.h"

namespace pf {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma16x16x4(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int kGramRowBlocks = 128;  // row partitions of the Gram reduction
constexpr int kMaxN = 64;            // widest B block of the Gram/update kernels

// Gram partials: part[(rb * m + i) * n + j] = sum over rb's slabs of A[:, i] B[:, j].
// Grid: (kGramRowBlocks, ceil(m/(64*AT))); block: 4 waves, wave w owns the AT
// tiles of 16 A columns from i0 = 64*AT*blockIdx.y + 16*AT*w and all n columns
// of B (NT <= 4 tiles), so each B load serves AT A tiles.  Slab sl (16 rows)
// belongs to row block sl mod gridDim.x; a block takes its full slabs two at
// a time with every load issued before the first MFMA.  Lanes past m or n
// read a clamped column (their results are not stored); only the rows past D
// of the last, partial slab are zeroed.
template <int NT, int AT>
__global__ __launch_bounds__(256) void gram_kernel(const double* __restrict__ A, int64_t lda, int m,
                                                   const double* __restrict__ B, int64_t ldb, int n,
                                                   int64_t D, double* __restrict__ part) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i0 = (blockIdx.y * 64 + w * 16) * AT;
  const int li = lane & 15, kq = lane >> 4;
  const double* pa[AT];
#pragma unroll
  for (int a = 0; a < AT; ++a) {
    const int ia = i0 + 16 * a + li;
    pa[a] = A + (int64_t)(ia < m ? ia : m - 1) * lda;
  }
  const double* pb[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int jb = t * 16 + li;
    pb[t] = B + (int64_t)(jb < n ? jb : n - 1) * ldb;
  }
  d4 acc[AT][NT];
#pragma unroll
  for (int a = 0; a < AT; ++a)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[a][t] = (d4){0.0, 0.0, 0.0, 0.0};
  // one slab's loads (rows r0..r0+3 of this lane) and MFMAs; zero rows >= D
  // only when `tail`
  auto load = [&](int64_t r0, bool tail, double (&av)[AT][4], double (&bv)[NT][4]) {
#pragma unroll
    for (int a = 0; a < AT; ++a)
#pragma unroll
      for (int q = 0; q < 4; ++q) av[a][q] = (!tail || r0 + q < D) ? pa[a][r0 + q] : 0.0;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int q = 0; q < 4; ++q) bv[t][q] = (!tail || r0 + q < D) ? pb[t][r0 + q] : 0.0;
  };
  auto fma = [&](const double (&av)[AT][4], const double (&bv)[NT][4]) {
#pragma unroll
    for (int a = 0; a < AT; ++a)
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[a][t] = mfma16x16x4(av[a][q], bv[t][q], acc[a][t]);
  };
  const int64_t gx = gridDim.x, nfull = D / 16;
  int64_t sl = blockIdx.x;
  for (; sl + gx < nfull; sl += 2 * gx) {
    double av0[AT][4], bv0[NT][4], av1[AT][4], bv1[NT][4];
    load(sl * 16 + 4 * kq, false, av0, bv0);
    load((sl + gx) * 16 + 4 * kq, false, av1, bv1);
    fma(av0, bv0);
    fma(av1, bv1);
  }
  if (sl < nfull) {  // one full slab left
    double av[AT][4], bv[NT][4];
    load(sl * 16 + 4 * kq, false, av, bv);
    fma(av, bv);
  }
  if (D % 16 != 0 && blockIdx.x == nfull % gx) {  // the partial last slab
    double av[AT][4], bv[NT][4];
    load(nfull * 16 + 4 * kq, true, av, bv);
    fma(av, bv);
  }
  // D-fragment: row (A column) i0 + 16a + kq + 4*v, column (B column) 16*t + li
  double* out = part + (int64_t)blockIdx.x * m * n;
#pragma unroll
  for (int a = 0; a < AT; ++a)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int j = t * 16 + li;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + 16 * a + kq + 4 * v;
        if (i < m && j < n) out[(int64_t)i * n + j] = acc[a][t][v];
      }
    }
}

// G[i*gsi + j*gsj] = sum_rb part[(rb*m + i)*n + j]  (fixed order:
// deterministic).  The partials are loaded 16 at a time before they are added
// in order, so one memory round trip serves 16 terms.
__global__ void gram_reduce_kernel(const double* __restrict__ part, int nrb, int m, int n,
                                   double* __restrict__ G, int64_t gsi, int64_t gsj) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)m * n) return;
  const int i = (int)(idx / n), j = (int)(idx % n);
  const int64_t stride = (int64_t)m * n;
  const double* p = part + (int64_t)i * n + j;
  constexpr int U = 16;
  double t = 0.0;
  int rb = 0;
  for (; rb + U <= nrb; rb += U, p += U * stride) {
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = p[u * stride];
#pragma unroll
    for (int u = 0; u < U; ++u) t += v[u];
  }
  for (; rb < nrb; ++rb, p += stride) t += *p;
  G[i * gsi + j * gsj] = t;
}

// B = beta*B + alpha*A*G.  Block: 4 waves, 64 rows (wave w: rows r0+16w..+15),
// all n output columns (NT tiles of 16).  Contraction over the m columns of A
// in steps of 4: A fragment lane l -> A[r0 + l%16, k + l/16] (16 consecutive
// rows of one column), G fragment lane l -> G(k + l/16, 16t + l%16), element
// (k, j) at G[k*gsk + j*gsj] (row-major G, gsj = 1, gives the 16 lanes of a
// k-slot one 128-byte line).  Four steps' loads are issued before their
// MFMAs.  Lanes past D or n read a clamped row or column (their results are
// not stored); only the contraction slots past m in the last steps are zeroed.
template <int NT>
__global__ __launch_bounds__(256) void update_kernel(const double* __restrict__ A, int64_t lda,
                                                     int m, const double* __restrict__ G,
                                                     int64_t gsk, int64_t gsj,
                                                     double* __restrict__ B, int64_t ldb, int n,
                                                     int64_t D, double alpha, double beta) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t r0 = (int64_t)blockIdx.x * 64 + w * 16;
  const int li = lane & 15, kq = lane >> 4;
  const int64_t ra = r0 + li;
  const double* pa = A + (ra < D ? ra : D - 1);
  const double* pg[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int j = t * 16 + li;
    pg[t] = G + (int64_t)(j < n ? j : n - 1) * gsj;
  }
  d4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = (d4){0.0, 0.0, 0.0, 0.0};
  constexpr int U = 4;  // steps of 4 columns whose loads travel together
  int k = 0;
  for (; k + 4 * U <= m; k += 4 * U) {
    double a[U], g[U][NT];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int ka = k + 4 * u + kq;
      a[u] = pa[(int64_t)ka * lda];
#pragma unroll
      for (int t = 0; t < NT; ++t) g[u][t] = pg[t][ka * gsk];
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[t] = mfma16x16x4(a[u], g[u][t], acc[t]);
  }
  for (; k < m; k += 4) {
    const int ka = k + kq;
    const bool kok = ka < m;
    const int kc = kok ? ka : m - 1;
    const double a = kok ? pa[(int64_t)kc * lda] : 0.0;
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = mfma16x16x4(a, pg[t][kc * gsk], acc[t]);
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int j = t * 16 + li;
    if (j >= n) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t r = r0 + kq + 4 * v;
      if (r < D) {
        double* pb = B + r + (int64_t)j * ldb;
        const double old = (beta == 0.0) ? 0.0 : beta * *pb;
        *pb = old + alpha * acc[t][v];
      }
    }
  }
}

// W[r, :] = W[r, :] * M for an n x n matrix M (column-major), one row per thread.
__global__ __launch_bounds__(256) void rotate_kernel(double* __restrict__ W, int64_t ldw, int64_t D,
                                                     int n, const double* __restrict__ M) {
  __shared__ double sM[kMaxN * kMaxN];
  for (int i = threadIdx.x; i < n * n; i += blockDim.x) sM[i] = M[i];
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= D) return;
  double row[kMaxN];
#pragma unroll 8
  for (int i = 0; i < n; ++i) row[i] = W[r + (int64_t)i * ldw];
  for (int j = 0; j < n; ++j) {
    double t = 0.0;
    for (int i = 0; i < n; ++i) t += row[i] * sM[i + j * n];
    W[r + (int64_t)j * ldw] = t;
  }
}

// Column partials of <A_j, B_j> (mode 0) or ||A_j - e_j B_j||^2 (mode 1).
__global__ __launch_bounds__(256) void coldot_kernel(const double* __restrict__ A, int64_t lda,
                                                     const double* __restrict__ B, int64_t ldb,
                                                     int64_t D, const double* __restrict__ e,
                                                     int mode, double* __restrict__ part) {
  __shared__ double red[4];
  const int c = blockIdx.y;
  double acc = 0.0;
  const double ec = mode ? e[c] : 0.0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < D;
       r += (int64_t)gridDim.x * blockDim.x) {
    const double a = A[r + (int64_t)c * lda], b = B[r + (int64_t)c * ldb];
    if (mode) {
      const double d = a - ec * b;
      acc += d * d;
    } else {
      acc += a * b;
    }
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[(int64_t)c * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void copy_block_kernel(const double* __restrict__ src, int64_t lds, double* __restrict__ dst,
                                  int64_t ldd, int64_t D) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r < D) dst[r + (int64_t)c * ldd] = src[r + (int64_t)c * lds];
}

// ---------------------------------------------------------------- drivers
// G(i, j) = <A_i, B_j> stored at G[i*gsi + j*gsj] (column-major: gsi = 1,
// gsj = ldg >= m; row-major: gsi = ldg >= n, gsj = 1).
void gram(pf_handle* h, const double* A, int64_t lda, int m, const double* B, int64_t ldb, int n,
          int64_t D, double* G, int64_t gsi, int64_t gsj, hipStream_t st) {
  PF_REQUIRE(m >= 1 && n >= 1 && n <= kMaxN, "gram: 1 <= n <= 64, m >= 1");
  PF_REQUIRE(D >= 1 && lda >= D && ldb >= D, "gram: bad leading dimensions");
  PF_REQUIRE((gsi == 1 && gsj >= m) || (gsj == 1 && gsi >= n), "gram: bad output layout");
  const int nrb = (int)std::min<int64_t>(kGramRowBlocks, (D + 15) / 16);
  double* part = reinterpret_cast<double*>(h->partials.get((size_t)nrb * m * n * sizeof(double)));
  // two A tiles per wave once A is wider than one block's 64 columns
  const int AT = m > 64 ? 2 : 1;
  dim3 grid(nrb, (m + 64 * AT - 1) / (64 * AT)), block(256);
  const int NT = (n + 15) / 16;
#define PF_GRAM(NTV, ATV) \
  hipLaunchKernelGGL((gram_kernel<NTV, ATV>), grid, block, 0, st, A, lda, m, B, ldb, n, D, part)
  if (AT == 2) {
    switch (NT) {
      case 1: PF_GRAM(1, 2); break;
      case 2: PF_GRAM(2, 2); break;
      case 3: PF_GRAM(3, 2); break;
      default: PF_GRAM(4, 2); break;
    }
  } else {
    switch (NT) {
      case 1: PF_GRAM(1, 1); break;
      case 2: PF_GRAM(2, 1); break;
      case 3: PF_GRAM(3, 1); break;
      default: PF_GRAM(4, 1); break;
    }
  }
#undef PF_GRAM
  PF_HIP(hipGetLastError());
  const int64_t mn = (int64_t)m * n;
  hipLaunchKernelGGL(gram_reduce_kernel, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, st, part,
                     nrb, m, n, G, gsi, gsj);
  PF_HIP(hipGetLastError());
}

// B = beta*B + alpha*A*G with G(k, j) at G[k*gsk + j*gsj] (column-major:
// gsk = 1, gsj = ldg >= m; row-major: gsk = ldg >= n, gsj = 1).
void update(const double* A, int64_t lda, int m, const double* G, int64_t gsk, int64_t gsj,
            double* B, int64_t ldb, int n, int64_t D, double alpha, double beta, hipStream_t st) {
  PF_REQUIRE(m >= 1 && n >= 1 && n <= kMaxN, "update: 1 <= n <= 64, m >= 1");
  PF_REQUIRE(D >= 1 && lda >= D && ldb >= D, "update: bad leading dimensions");
  PF_REQUIRE((gsk == 1 && gsj >= m) || (gsj == 1 && gsk >= n), "update: bad G layout");
  dim3 grid((unsigned)((D + 63) / 64)), block(256);
  const int NT = (n + 15) / 16;
#define PF_UPDATE(NTV)                                                                         \
  hipLaunchKernelGGL(update_kernel<NTV>, grid, block, 0, st, A, lda, m, G, gsk, gsj, B, ldb, n, \
                     D, alpha, beta)
  switch (NT) {
    case 1: PF_UPDATE(1); break;
    case 2: PF_UPDATE(2); break;
    case 3: PF_UPDATE(3); break;
    default: PF_UPDATE(4); break;
  }
#undef PF_UPDATE
  PF_HIP(hipGetLastError());
}

// per-column reductions returned to the host (fixed order)
static void coldots_host(pf_handle* h, const double* A, int64_t lda, const double* B, int64_t ldb,
                         int64_t D, int n, const double* e_dev, int mode, double* out,
                         hipStream_t st) {
  const int nblk = (int)std::min<int64_t>(256, (D + 255) / 256);
  double* part = reinterpret_cast<double*>(h->partials.get((size_t)nblk * n * sizeof(double)));
  hipLaunchKernelGGL(coldot_kernel, dim3(nblk, n), dim3(256), 0, st, A, lda, B, ldb, D, e_dev, mode,
                     part);
  PF_HIP(hipGetLastError());
  std::vector<double> hp((size_t)nblk * n);
  PF_HIP(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  PF_HIP(hipStreamSynchronize(st));
  for (int c = 0; c < n; ++c) {
    double t = 0.0;
    for (int b = 0; b < nblk; ++b) t += hp[(size_t)c * nblk + b];
    out[c] = t;
  }
}

// Cholesky G = R^T R of a small SPD matrix (column-major, n x n); returns false
// when a pivot is not positive.
static bool cholesky_upper(const std::vector<double>& G, int n, std::vector<double>& R) {
  R.assign((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j) {
    double d = G[j + (size_t)j * n];
    for (int k = 0; k < j; ++k) d -= R[k + (size_t)j * n] * R[k + (size_t)j * n];
    if (!(d > 0.0) || !std::isfinite(d)) return false;
    const double rjj = std::sqrt(d);
    R[j + (size_t)j * n] = rjj;
    for (int i = j + 1; i < n; ++i) {
      double t = G[j + (size_t)i * n];
      for (int k = 0; k < j; ++k) t -= R[k + (size_t)j * n] * R[k + (size_t)i * n];
      R[j + (size_t)i * n] = t / rjj;
    }
  }
  return true;
}

// inverse of an upper-triangular n x n matrix (column-major)
static void inv_upper(const std::vector<double>& R, int n, std::vector<double>& Ri) {
  Ri.assign((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j) {
    Ri[j + (size_t)j * n] = 1.0 / R[j + (size_t)j * n];
    for (int i = j - 1; i >= 0; --i) {
      double t = 0.0;
      for (int k = i + 1; k <= j; ++k) t += R[i + (size_t)k * n] * Ri[k + (size_t)j * n];
      Ri[i + (size_t)j * n] = -t / R[i + (size_t)i * n];
    }
  }
}

static void rotate(pf_handle* h, double* W, int64_t ldw, int64_t D, int n, const double* M_host,
                   hipStream_t st) {
  PF_REQUIRE(n >= 1 && n <= kMaxN, "rotate: 1 <= n <= 64");
  double* dM = reinterpret_cast<double*>(h->small.get((size_t)n * n * sizeof(double)));
  PF_HIP(hipMemcpyAsync(dM, M_host, (size_t)n * n * sizeof(double), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(rotate_kernel, dim3((unsigned)((D + 255) / 256)), dim3(256), 0, st, W, ldw, D,
                     n, dM);
  PF_HIP(hipGetLastError());
  // the host matrix may be freed by the caller after return
  PF_HIP(hipStreamSynchronize(st));
}

}  // namespace pf

using namespace pf;

extern "C" int pf_gram(pf_handle* h, const double* A, int64_t lda, int32_t m, const double* B,
                       int64_t ldb, int32_t n, int64_t D, double* G, int32_t ldg, pf_stream stream) {
  return guarded([&] {
    PF_REQUIRE(h && A && B && G, "gram: null argument");
    PF_REQUIRE(ldg >= m, "gram: ldg < m");
    gram(h, A, lda, m, B, ldb, n, D, G, 1, ldg, S(stream));
  });
}

extern "C" int pf_gemm_update(pf_handle* h, const double* A, int64_t lda, int32_t m,
                              const double* G, int32_t ldg, double* B, int64_t ldb, int32_t n,
                              int64_t D, double alpha, double beta, pf_stream stream) {
  return guarded([&] {
    PF_REQUIRE(h && A && G && B, "update: null argument");
    PF_REQUIRE((const void*)A != (const void*)B, "update: B must not alias A");
    PF_REQUIRE(ldg >= m, "update: ldg < m");
    update(A, lda, m, G, 1, ldg, B, ldb, n, D, alpha, beta, S(stream));
  });
}

extern "C" int pf_reorth(pf_handle* h, const double* V, int64_t ldv, int32_t k, double* W,
                         int64_t ldw, int32_t s, int64_t D, int32_t passes, pf_stream stream) {
  return guarded([&] {
    PF_REQUIRE(h && V && W, "reorth: null argument");
    PF_REQUIRE(passes >= 1 && passes <= 4, "reorth: 1 <= passes <= 4");
    if (k <= 0) return;
    hipStream_t st = S(stream);
    // G = V^T W held row-major (G(i, j) at G[i*s + j]): the update's 16 lanes
    // of one contraction index then read one contiguous line of G
    double* G = reinterpret_cast<double*>(h->vecs.get((size_t)k * s * sizeof(double)));
    for (int p = 0; p < passes; ++p) {
      gram(h, V, ldv, k, W, ldw, s, D, G, s, 1, st);
      update(V, ldv, k, G, s, 1, W, ldw, s, D, -1.0, 1.0, st);
    }
  });
}

extern "C" int pf_block_qr(pf_handle* h, double* W, int64_t ldw, int64_t D, int32_t s,
                           double* R_host, pf_stream stream) {
  return guarded([&] {
    PF_REQUIRE(h && W && R_host, "qr: null argument");
    PF_REQUIRE(s >= 1 && s <= kMaxN && D >= s && ldw >= D, "qr: bad shape");
    hipStream_t st = S(stream);
    // keep a copy to restore W when Cholesky breaks down
    double* save = reinterpret_cast<double*>(
        h->vecs.get((size_t)D * s * sizeof(double) + (size_t)s * s * sizeof(double)));
    double* dG = save + (size_t)D * s;
    hipLaunchKernelGGL(copy_block_kernel, dim3((unsigned)((D + 255) / 256), s), dim3(256), 0, st, W,
                       ldw, save, D, D);
    PF_HIP(hipGetLastError());
    std::vector<double> G((size_t)s * s), R1, R2, Ri, R((size_t)s * s, 0.0);
    for (int pass = 0; pass < 2; ++pass) {
      gram(h, W, ldw, s, W, ldw, s, D, dG, 1, s, st);
      PF_HIP(hipMemcpyAsync(G.data(), dG, G.size() * sizeof(double), hipMemcpyDeviceToHost, st));
      PF_HIP(hipStreamSynchronize(st));
      std::vector<double>& Rp = pass == 0 ? R1 : R2;
      if (!cholesky_upper(G, s, Rp)) {
        hipLaunchKernelGGL(copy_block_kernel, dim3((unsigned)((D + 255) / 256), s), dim3(256), 0,
                           st, save, D, W, ldw, D);
        PF_HIP(hipGetLastError());
        PF_HIP(hipStreamSynchronize(st));
        throw Error{PF_ERR_NUMERIC, "qr: Gram matrix not positive definite (rank-deficient block)"};
      }
      inv_upper(Rp, s, Ri);
      rotate(h, W, ldw, D, s, Ri.data(), st);
    }
    // R = R2 * R1 (both upper triangular)
    for (int j = 0; j < s; ++j)
      for (int i = 0; i <= j; ++i) {
        double t = 0.0;
        for (int k = i; k <= j; ++k) t += R2[i + (size_t)k * s] * R1[k + (size_t)j * s];
        R[i + (size_t)j * s] = t;
      }
    std::memcpy(R_host, R.data(), R.size() * sizeof(double));
  });
}

extern "C" int pf_rotate(pf_handle* h, double* W, int64_t ldw, int64_t D, int32_t n,
                         const double* M_host, pf_stream stream) {
  return guarded([&] {
    PF_REQUIRE(h && W && M_host && ldw >= D && D >= 1, "rotate: bad argument");
    rotate(h, W, ldw, D, n, M_host, S(stream));
  });
}

extern "C" int pf_coldots(pf_handle* h, const double* A, int64_t lda, const double* B, int64_t ldb,
                          int64_t D, int32_t n, double* out_host, pf_stream stream) {
  return guarded([&] {
    PF_REQUIRE(h && A && B && out_host && n >= 1 && n <= 65535 && D >= 1 && lda >= D && ldb >= D,
               "coldots: bad argument");
    coldots_host(h, A, lda, B, ldb, D, n, nullptr, 0, out_host, S(stream));
  });
}

extern "C" int pf_residuals(pf_handle* h, pf_op* op, const double* psi, int64_t ld, int32_t nev,
                            double* E_host, double* res_host, pf_stream stream) {
  return guarded([&] {
    PF_REQUIRE(h && op && psi && E_host && res_host, "residuals: null argument");
    PF_REQUIRE(nev >= 1 && nev <= 65535 && ld >= op->D, "residuals: bad shape");
    hipStream_t st = S(stream);
    const int64_t D = op->D;
    // H psi with the same leading dimension as psi (the apply kernels share offsets)
    double* Hpsi = reinterpret_cast<double*>(
        h->vecs.get((size_t)ld * nev * sizeof(double) + (size_t)nev * sizeof(double)));
    double* dE = Hpsi + (size_t)ld * nev;
    int rc = pf_apply(op, PF_F64, 0, psi, ld, Hpsi, ld, nev, stream);
    if (rc != PF_OK) throw Error{rc, pf_last_error()};
    coldots_host(h, psi, ld, Hpsi, ld, D, nev, nullptr, 0, E_host, st);
    PF_HIP(hipMemcpyAsync(dE, E_host, nev * sizeof(double), hipMemcpyHostToDevice, st));
    coldots_host(h, Hpsi, ld, psi, ld, D, nev, dE, 1, res_host, st);
    for (int i = 0; i < nev; ++i) res_host[i] = std::sqrt(res_host[i]);
  });
}
