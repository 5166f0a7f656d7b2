// Combinadic ranking of fixed-popcount spin states (host + device).
//
// The Sz-sector basis of the reference (_lup_basis, src/Models/SpinHalfChains.jl:259-262)
// lists the L-bit integers with Lup set bits in ascending order, and looks a
// flipped state up through a Dict (:278, :301).  The position of a state s with
// set bits p_0 < p_1 < ... is the combinadic rank  r(s) = sum_t C(p_t, t+1),
// so the matrix-free kernels compute the rank of a bond-flipped neighbour as
// r(s) + delta with an O(1) update instead of storing any index table.
#pragma once

#include <stdint.h>

#ifndef PF_HD
#if defined(__HIPCC__)
#define PF_HD __host__ __device__
#else
#define PF_HD
#endif
#endif

namespace pf {

constexpr int kRankT = 20;  // row width of the binomial table C[p*kRankT + t]

PF_HD inline int popc32(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __popc(x);
#else
  return __builtin_popcount(x);
#endif
}

PF_HD inline int ctz32(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __ffs(x) - 1;
#else
  return __builtin_ctz(x);
#endif
}

// Rank of the state obtained from `st` (rank r) by exchanging the different
// bits at positions lo < hi.  C is the binomial table C(p, t) for p < 32,
// t < kRankT.  Moving the set bit from lo up to hi, with k set bits below lo
// and set bits q_1 < ... < q_m strictly between, changes the rank by
//   C(hi, k+m+1) - C(lo, k+1) + sum_t [C(q_t, k+t) - C(q_t, k+t+1)],
// which for adjacent sites (hi = lo+1) is C(lo, k) (Pascal's rule).  Moving the
// set bit down from hi to lo is the inverse move, with the same k and q_t.
PF_HD inline uint32_t neighbor_rank(uint32_t st, uint32_t r, int lo, int hi, const uint32_t* C) {
  const int k = popc32(st & ((1u << lo) - 1u));
  int32_t d;
  if (hi == lo + 1) {
    d = (int32_t)C[lo * kRankT + k];
  } else {
    uint32_t between = st & ((1u << hi) - 1u) & ~((2u << lo) - 1u);
    const int m = popc32(between);
    d = (int32_t)C[hi * kRankT + k + m + 1] - (int32_t)C[lo * kRankT + k + 1];
    int t = 0;
    while (between) {
      const int q = ctz32(between);
      between &= between - 1u;
      ++t;
      d += (int32_t)C[q * kRankT + k + t] - (int32_t)C[q * kRankT + k + t + 1];
    }
  }
  return ((st >> lo) & 1u) ? r + (uint32_t)d : r - (uint32_t)d;
}

// Full rank of a state (used on the host to build and check tables).
inline uint64_t full_rank(uint32_t st, const uint32_t* C) {
  uint64_t r = 0;
  int t = 0;
  while (st) {
    const int p = ctz32(st);
    st &= st - 1u;
    r += C[p * kRankT + t + 1];
    ++t;
  }
  return r;
}

// Fill the binomial table C(p, t) for p < 32, t < kRankT.
inline void fill_binomials(uint32_t* C) {
  for (int p = 0; p < 32; ++p)
    for (int t = 0; t < kRankT; ++t) {
      uint64_t c;
      if (t > p) {
        c = 0;
      } else {
        c = 1;
        for (int i = 1; i <= t; ++i) c = c * (uint64_t)(p - t + i) / (uint64_t)i;
      }
      C[p * kRankT + t] = (uint32_t)c;
    }
}

}  // namespace pf
