// Plain-FP64 ladder solves for small compound sizes, register resident.
//
// One workgroup solves one (chan, w, q) problem A X = B, A of size n <= TR*RPT
// and [A | B] of width n + nb^2 <= TC*CPT. Thread (tr, tc) holds the RPT x CPT
// elements at rows tr + TR*ii and columns tc + TC*jj of the augmented matrix
// in registers for the whole factorisation, so A is read from memory once
// (coalesced along the columns of the shared vertex Gamma(w)) and never
// written back.
//
// Elimination is right-looking Gaussian elimination with partial pivoting
// (pivot = largest |Re|+|Im| of the column, LAPACK's izamax criterion), with
// physical row swaps through LDS so that the active trailing block stays the
// contiguous range i, j > k and finished row/column groups are skipped with
// wave-uniform branches. Two barriers per elimination step.
//
// The upper-triangular back substitution runs in chunks of 32 columns:
// wavefront 0 solves the chunk's triangle with readlane broadcasts (no LDS
// round trip on the critical path), then all threads update the remaining
// right-hand sides with the chunk (rectangle update).
#include "dgami.h"
#include "dgami_internal.hpp"
#include "dgami_launch.hpp"
#include "ladder_common.hpp"
#include "ladder_tune.hpp"

namespace dgami {

constexpr int LADDER_CB = 32;   // back-substitution chunk width

__device__ __forceinline__ double readlane_d(double v, int lane) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), lane);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// Problem `item` = (chan, w, q) of model `a`, solved by the calling workgroup.
template <int TR, int TC, int RPT, int CPT>
__device__ __forceinline__ void ladder_reg_solve(const LadderArgs& a, int item) {
  constexpr int NT = TR * TC;
  constexpr int NWAVE = NT / 64;
  static_assert(64 % TC == 0, "TC must divide the wavefront");
  const int tid = threadIdx.x;
  const int tc = tid % TC, tr = tid / TC;
  const int lane = tid & 63, wave = tid >> 6;

  const int q = item % a.nq;
  const int rest = item / a.nq;
  const int w = rest % a.nw;
  const int chan = rest / a.nw;
  const int n = a.n, T = a.T, ncol = n + T;

  __shared__ double2 s_u[2][TC * CPT];
  __shared__ double2 s_w[2][TC * CPT];
  __shared__ double2 s_l[2][TR * RPT];
  __shared__ double s_val[2][NWAVE];
  __shared__ int s_idx[2][NWAVE];
  __shared__ double2 s_pv[2][NWAVE];
  __shared__ double2 s_dinv[TR * RPT];
  __shared__ int s_info;
  extern __shared__ double2 s_dyn[];  // back substitution: Y[n][T] then UC[n][CB]
  double2* s_y = s_dyn;
  double2* s_uc = s_dyn + n * T;

  const double2* G = ladder_matrix_ptr(a, chan, w, q);
  const double2* U = a.dense ? nullptr : a.u_r + chan * T * T;
  const double2* V = (a.v_r && !a.dense) ? a.v_r + (static_cast<size_t>(chan) * a.nq + q) * T * T : nullptr;

  double2 A[RPT][CPT];
#pragma unroll
  for (int ii = 0; ii < RPT; ++ii) {
    const int i = tr + TR * ii;
#pragma unroll
    for (int jj = 0; jj < CPT; ++jj) {
      const int j = tc + TC * jj;
      double2 val = make_double2(0.0, 0.0);
      if (i < n && j < ncol) val = (j < n) ? ladder_elem(a, G, U, V, w, q, i, j) : ladder_rhs(a, i, j - n);
      A[ii][jj] = val;
    }
  }
  if (tid == 0) s_info = 0;

  for (int k = 0; k < n; ++k) {
    const int buf = k & 1;
    const int kc = k % TC, kj = k / TC;
    const int kr = k % TR, ki = k / TR;
    // ---- 1. pivot search in column k over rows i >= k (column owners) ----
    double best = -1.0;
    int bidx = 0x7fffffff;
    double2 bval = make_double2(0.0, 0.0);
    if (tc == kc) {
#pragma unroll
      for (int ii = 0; ii < RPT; ++ii) {
        const int i = tr + TR * ii;
        double2 x = make_double2(0.0, 0.0);
#pragma unroll
        for (int jj = 0; jj < CPT; ++jj)
          if (jj == kj) x = A[ii][jj];
        const double m = cabs1(x);
        if (i >= k && i < n && m > best) {
          best = m;
          bidx = i;
          bval = x;
        }
      }
    }
#pragma unroll
    for (int off = TC; off < 64; off <<= 1) {
      const double ov = __shfl_xor(best, off);
      const int oi = __shfl_xor(bidx, off);
      const double ox = __shfl_xor(bval.x, off);
      const double oy = __shfl_xor(bval.y, off);
      if (ov > best || (ov == best && oi < bidx)) {
        best = ov;
        bidx = oi;
        bval = make_double2(ox, oy);
      }
    }
    if (lane == kc) {
      s_val[buf][wave] = best;
      s_idx[buf][wave] = bidx;
      s_pv[buf][wave] = bval;
    }
    __syncthreads();
    // ---- 2. every thread reduces the per-wave candidates (identical result) ----
    best = s_val[buf][0];
    bidx = s_idx[buf][0];
    bval = s_pv[buf][0];
#pragma unroll
    for (int wv = 1; wv < NWAVE; ++wv) {
      const double ov = s_val[buf][wv];
      const int oi = s_idx[buf][wv];
      if (ov > best || (ov == best && oi < bidx)) {
        best = ov;
        bidx = oi;
        bval = s_pv[buf][wv];
      }
    }
    const int p = bidx == 0x7fffffff ? k : bidx;  // no finite candidate (NaN column): keep row k
    const bool ok = best > 0.0;
    const double2 pinv = ok ? cinv(bval) : make_double2(0.0, 0.0);
    if (tid == 0) {
      if (!ok && s_info == 0) s_info = k + 1;
      s_dinv[k] = pinv;
    }
    const int pr = p % TR, pi = p / TR;
    // pivot row -> s_u (columns > k)
    if (tr == pr) {
#pragma unroll
      for (int ii = 0; ii < RPT; ++ii)
        if (ii == pi) {
#pragma unroll
          for (int jj = 0; jj < CPT; ++jj) {
            const int j = tc + TC * jj;
            if (j > k && j < ncol) s_u[buf][j] = A[ii][jj];
          }
        }
    }
    // old row k -> s_w (it moves to position p)
    if (p != k && tr == kr) {
#pragma unroll
      for (int ii = 0; ii < RPT; ++ii)
        if (ii == ki) {
#pragma unroll
          for (int jj = 0; jj < CPT; ++jj) {
            const int j = tc + TC * jj;
            if (j > k && j < ncol) s_w[buf][j] = A[ii][jj];
          }
        }
    }
    // multipliers of rows > k after the swap
    if (tc == kc) {
#pragma unroll
      for (int ii = 0; ii < RPT; ++ii) {
        const int i = tr + TR * ii;
        double2 x = make_double2(0.0, 0.0);
#pragma unroll
        for (int jj = 0; jj < CPT; ++jj)
          if (jj == kj) x = A[ii][jj];
        if (i > k && i < n && i != p) s_l[buf][i] = cmul(x, pinv);
        if (i == k && p != k) s_l[buf][p] = cmul(x, pinv);
      }
    }
    __syncthreads();
    // ---- 3. swap in registers, then rank-1 update of rows/cols > k ----
    double2 uj[CPT];
#pragma unroll
    for (int jj = 0; jj < CPT; ++jj) {
      const int j = tc + TC * jj;
      uj[jj] = (j > k && j < ncol) ? s_u[buf][j] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int ii = 0; ii < RPT; ++ii) {
      if (TR * ii + TR - 1 < k) continue;  // whole row group finished (uniform)
      const int i = tr + TR * ii;
      if (i == k) {
#pragma unroll
        for (int jj = 0; jj < CPT; ++jj) {
          const int j = tc + TC * jj;
          if (j > k && j < ncol) A[ii][jj] = uj[jj];
          if (j == k) A[ii][jj] = bval;
        }
      } else if (i > k && i < n) {
        if (i == p) {
#pragma unroll
          for (int jj = 0; jj < CPT; ++jj) {
            const int j = tc + TC * jj;
            if (j > k && j < ncol) A[ii][jj] = s_w[buf][j];
          }
        }
        const double2 l = s_l[buf][i];
#pragma unroll
        for (int jj = 0; jj < CPT; ++jj) {
          if (TC * jj + TC - 1 <= k) continue;  // whole column group finished (uniform)
          A[ii][jj] = cfms(A[ii][jj], l, uj[jj]);
        }
      }
    }
  }

  // ---- back substitution: U x = y, chunks of LADDER_CB columns from the right ----
  const int nr = T;
#pragma unroll
  for (int ii = 0; ii < RPT; ++ii) {
    const int i = tr + TR * ii;
#pragma unroll
    for (int jj = 0; jj < CPT; ++jj) {
      const int j = tc + TC * jj;
      if (i < n && j >= n && j < ncol) s_y[i * nr + (j - n)] = A[ii][jj];
    }
  }
  for (int c1 = n; c1 > 0; c1 -= LADDER_CB) {
    const int c0 = c1 > LADDER_CB ? c1 - LADDER_CB : 0;
#pragma unroll
    for (int ii = 0; ii < RPT; ++ii) {
      const int i = tr + TR * ii;
#pragma unroll
      for (int jj = 0; jj < CPT; ++jj) {
        const int j = tc + TC * jj;
        if (i < c1 && j >= c0 && j < c1 && j >= i) s_uc[i * LADDER_CB + (j - c0)] = A[ii][jj];
      }
    }
    __syncthreads();
    if (wave == 0) {
      // lane t <-> row c0 + t of the chunk; one right-hand side at a time
      const int row = c0 + lane;
      const bool live = lane < (c1 - c0);
      const double2 dinv = live ? s_dinv[row] : make_double2(0.0, 0.0);
      for (int r = 0; r < nr; ++r) {
        double2 y = live ? s_y[row * nr + r] : make_double2(0.0, 0.0);
#pragma unroll
        for (int t = LADDER_CB - 1; t >= 0; --t) {
          if (c0 + t >= c1) continue;  // uniform
          // lane t finalises x_t = y_t / U_tt, then every lane above eliminates it
          if (lane == t) y = cmul(y, dinv);
          const double2 x = make_double2(readlane_d(y.x, t), readlane_d(y.y, t));
          if (lane < t) y = cfms(y, s_uc[row * LADDER_CB + t], x);
        }
        if (live) s_y[row * nr + r] = y;
      }
    }
    __syncthreads();
    // rectangle: rows < c0 eliminate the chunk's unknowns
    for (int idx = tid; idx < c0 * nr; idx += NT) {
      const int i = idx / nr, r = idx - i * nr;
      double2 acc = s_y[idx];
      for (int t = 0; t < c1 - c0; ++t) acc = cfms(acc, s_uc[i * LADDER_CB + t], s_y[(c0 + t) * nr + r]);
      s_y[idx] = acc;
    }
    __syncthreads();
  }
  for (int idx = tid; idx < n * nr; idx += NT) {
    const int i = idx / nr, r = idx - i * nr;
    a.out[ladder_out_index(a, chan, w, q, i, r)] = cscale(s_y[idx], a.inv_beta);
  }
  if (tid == 0 && a.info) a.info[(static_cast<size_t>(chan) * a.nw + w) * a.nq + q] = s_info;
}

template <int TR, int TC, int RPT, int CPT>
__global__ void __launch_bounds__(TR* TC) ladder_reg_kernel(LadderArgs a) {
  ladder_reg_solve<TR, TC, RPT, CPT>(a, blockIdx.x);
}

// Several models in one launch: workgroup b solves item b - start[m] of the model m with
// start[m] <= b < start[m + 1] (the table travels in the kernel arguments, so a captured
// graph keeps it).
template <int TR, int TC, int RPT, int CPT>
__global__ void __launch_bounds__(TR* TC) ladder_reg_multi_kernel(LadderMulti mm) {
  const int b = blockIdx.x;
  int m = 0;
  while (m + 1 < mm.nmodels && b >= mm.start[m + 1]) ++m;
  ladder_reg_solve<TR, TC, RPT, CPT>(mm.model[m], b - mm.start[m]);
}

struct RegVariant {
  int max_rows, max_cols, threads;
  void (*kernel)(LadderArgs);
  void (*multi)(LadderMulti);
};

#define DGAMI_REG_VARIANT(TR, TC, RPT, CPT)                                            \
  RegVariant {                                                                         \
    TR * RPT, TC * CPT, TR * TC, ladder_reg_kernel<TR, TC, RPT, CPT>, ladder_reg_multi_kernel<TR, TC, RPT, CPT> \
  }

static const RegVariant kRegVariants[] = {
    DGAMI_REG_VARIANT(16, 16, 3, 4),   // n <= 48,  n + nb^2 <= 64
    DGAMI_REG_VARIANT(16, 16, 4, 5),   // n <= 64,  n + nb^2 <= 80
    DGAMI_REG_VARIANT(32, 16, 3, 6),   // n <= 96,  n + nb^2 <= 96
    DGAMI_REG_VARIANT(32, 16, 4, 8),   // n <= 128, n + nb^2 <= 128
};

int ladder_reg_variant_count() { return static_cast<int>(sizeof(kRegVariants) / sizeof(kRegVariants[0])); }

bool ladder_reg_variant_fits(int v, int n, int nrhs) {
  if (v < 0 || v >= ladder_reg_variant_count()) return false;
  const size_t lds = static_cast<size_t>(n) * (nrhs + LADDER_CB) * sizeof(double2);
  return n <= kRegVariants[v].max_rows && n + nrhs <= kRegVariants[v].max_cols && lds <= 160 * 1024;
}

// Index of the register variant for (n, nrhs), or -1 when n is too large.
int ladder_reg_variant(int n, int nrhs) {
  for (int v = 0; v < static_cast<int>(sizeof(kRegVariants) / sizeof(kRegVariants[0])); ++v)
    if (n <= kRegVariants[v].max_rows && n + nrhs <= kRegVariants[v].max_cols) return v;
  return -1;
}

int ladder_reg_launch_multi(int variant, const LadderMulti& mm, hipStream_t stream) {
  const RegVariant& rv = kRegVariants[variant];
  const int items = mm.start[mm.nmodels];
  if (items == 0) return 0;
  size_t lds = 0;
  for (int m = 0; m < mm.nmodels; ++m) {
    const size_t l = static_cast<size_t>(mm.model[m].n) * (mm.model[m].T + LADDER_CB) * sizeof(double2);
    if (l > lds) lds = l;
  }
  allow_lds(rv.multi, lds);
  hipLaunchKernelGGL(rv.multi, dim3(static_cast<unsigned>(items)), dim3(rv.threads), lds, stream, mm);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}

int ladder_reg_launch(int variant, const LadderArgs& a, hipStream_t stream) {
  const RegVariant& rv = kRegVariants[variant];
  const size_t items = static_cast<size_t>(a.nchan) * a.nw * a.nq;
  DGAMI_ARG_CHECK(items < (size_t(1) << 31), "ladder: too many problems in one call");
  const size_t lds = static_cast<size_t>(a.n) * (a.T + LADDER_CB) * sizeof(double2);
  allow_lds(rv.kernel, lds);
  hipLaunchKernelGGL(rv.kernel, dim3(static_cast<unsigned>(items)), dim3(rv.threads), lds, stream, a);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace dgami
