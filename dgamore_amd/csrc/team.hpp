// Wavefront "team" algebra on small complex blocks: a team of TS lanes
// (power of two, TS >= T, inside one wavefront) owns one T x T block and lane
// r holds row r in registers. Rows move between lanes by shuffles only.
#pragma once

#include "dgami_internal.hpp"

namespace dgami {

__device__ __forceinline__ double2 shfl2(double2 v, int src, int width) {
  return make_double2(__shfl(v.x, src, width), __shfl(v.y, src, width));
}

// In-place inverse of the team's block. Lane r (< T) holds row r of `m`;
// on return it holds row r of m^-1. Returns true if a pivot was exactly zero.
template <int T, int TS>
__device__ __forceinline__ bool team_invert(double2 (&m)[T], int r) {
  double2 e[T];
#pragma unroll
  for (int j = 0; j < T; ++j) e[j] = make_double2(r == j ? 1.0 : 0.0, 0.0);
  bool singular = false;
#pragma unroll
  for (int k = 0; k < T; ++k) {
    double val = (r >= k && r < T) ? cabs1(m[k]) : -1.0;
    int idx = r;
#pragma unroll
    for (int off = TS / 2; off > 0; off >>= 1) {
      const double ov = __shfl_xor(val, off, TS);
      const int oi = __shfl_xor(idx, off, TS);
      if (ov > val || (ov == val && oi < idx)) {
        val = ov;
        idx = oi;
      }
    }
    const int p = idx;
    if (!(val > 0.0)) singular = true;
    const int src = (r == k) ? p : ((r == p) ? k : r);
#pragma unroll
    for (int j = 0; j < T; ++j) {
      m[j] = shfl2(m[j], src, TS);
      e[j] = shfl2(e[j], src, TS);
    }
    const double2 piv = shfl2(m[k], k, TS);
    const double2 pinv = (val > 0.0) ? cinv(piv) : make_double2(0.0, 0.0);
    if (r == k) {
#pragma unroll
      for (int j = 0; j < T; ++j) {
        m[j] = cmul(m[j], pinv);
        e[j] = cmul(e[j], pinv);
      }
    }
    const double2 f = m[k];
#pragma unroll
    for (int j = 0; j < T; ++j) {
      const double2 mk = shfl2(m[j], k, TS);
      const double2 ek = shfl2(e[j], k, TS);
      if (r != k) {
        m[j] = cfms(m[j], f, mk);
        e[j] = cfms(e[j], f, ek);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < T; ++j) m[j] = e[j];
  return singular;
}

// Row r of the team product A.B, lane r holding row r of both operands.
template <int T, int TS>
__device__ __forceinline__ void team_matmul(const double2 (&a)[T], const double2 (&b)[T], double2 (&out)[T]) {
#pragma unroll
  for (int j = 0; j < T; ++j) out[j] = make_double2(0.0, 0.0);
#pragma unroll
  for (int c = 0; c < T; ++c) {
#pragma unroll
    for (int j = 0; j < T; ++j) out[j] = cfma(out[j], a[c], shfl2(b[j], c, TS));
  }
}

}  // namespace dgami
