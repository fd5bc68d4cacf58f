// Plain-FP64 ladder solves for n <= 128, column-cyclic register layout.
//
// One workgroup of NW wavefronts solves one (chan, w, q) problem A X = B.
// Wavefront w owns the columns j = w + NW*jj of the augmented matrix [A | B]
// and lane l owns the rows i = l + 64*m, all in registers (RS*CS complex per
// lane). Gaussian elimination with partial pivoting (largest |Re|+|Im| by the
// upper word of its FP64 value, lowest lane on ties) then needs only one
// workgroup barrier per step:
//   * the wavefront owning column k finds the pivot with a wavefront
//     reduction and publishes the pivot row index and the multipliers of
//     column k through LDS;            -- barrier --
//   * every wavefront swaps rows k and p of its own columns and broadcasts the
//     pivot row to its lanes through a wavefront-private LDS slot, then applies
//     the rank-1 update to rows > k of its columns > k.
// The multiplier buffer is double-buffered, so step k+1's owner can publish
// while slower wavefronts still read step k's values.
// Back substitution: the upper triangle is packed column-major into LDS and
// wavefront 0 solves U x = y with readlane broadcasts (no further barriers).
#include "dgami.h"
#include "dgami_internal.hpp"
#include "dgami_launch.hpp"
#include "ladder_common.hpp"
#include "ladder_tune.hpp"

namespace dgami {

__device__ __forceinline__ double rl_d(double v, int lane) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), lane);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int NW, int RS, int CS>
__global__ void __launch_bounds__(NW * 64) ladder_col_kernel(LadderArgs a) {
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int q = blockIdx.x % a.nq;
  const int rest = blockIdx.x / a.nq;
  const int w = rest % a.nw;
  const int chan = rest / a.nw;
  const int n = a.n, T = a.T, ncol = n + T;

  __shared__ double2 s_l[2][64 * RS];   // multipliers of column k (by row)
  __shared__ double2 s_prow[NW][CS];    // pivot row, per wavefront
  __shared__ double2 s_krow[NW][CS];    // old row k, per wavefront
  __shared__ double2 s_dinv[64 * RS];   // inverse pivots
  __shared__ int s_p[2];
  __shared__ double2 s_pv[2];
  __shared__ int s_info;
  extern __shared__ double2 s_dyn[];    // packed U (column major) then Y[n][T]

  const double2* G = ladder_matrix_ptr(a, chan, w, q);
  const double2* U = a.dense ? nullptr : a.u_r + chan * T * T;
  const double2* V = (a.v_r && !a.dense) ? a.v_r + (static_cast<size_t>(chan) * a.nq + q) * T * T : nullptr;

  double2 A[RS][CS];
#pragma unroll
  for (int m = 0; m < RS; ++m) {
    const int i = lane + 64 * m;
#pragma unroll
    for (int jj = 0; jj < CS; ++jj) {
      const int j = wave + NW * jj;
      double2 val = make_double2(0.0, 0.0);
      if (i < n && j < ncol) val = (j < n) ? ladder_elem(a, G, U, V, w, q, i, j) : ladder_rhs(a, i, j - n);
      A[m][jj] = val;
    }
  }
  if (tid == 0) s_info = 0;

  for (int k = 0; k < n; ++k) {
    const int buf = k & 1;
    const int ow = k % NW, kj = k / NW;
    const int kl = k & 63, km = k >> 6;
    if (wave == ow) {
      // ---- pivot search in column k (owned by this wavefront) ----
      double best = -1.0;
      int bidx = 0x7fffffff;
      double2 col[RS];
#pragma unroll
      for (int m = 0; m < RS; ++m) {
        double2 x = make_double2(0.0, 0.0);
#pragma unroll
        for (int jj = 0; jj < CS; ++jj)
          if (jj == kj) x = A[m][jj];
        col[m] = x;
        const int i = lane + 64 * m;
        const double v = cabs1(x);
        if (i >= k && i < n && v > best) {
          best = v;
          bidx = i;
        }
      }
      {
        const unsigned key = pivot_key(best, best >= 0.0);
        const unsigned wkey = wave_max_u32(key);
        const unsigned long long hit = __ballot(key == wkey);
        bidx = __builtin_amdgcn_readlane(bidx, __ffsll(static_cast<long long>(hit)) - 1);
        best = wkey > 1u ? 1.0 : 0.0;  // nonzero pivot magnitude
        if (wkey == 0u) bidx = k;      // no finite candidate (NaN column): keep row k, flagged singular below
      }
      const int p = bidx;
      const int pl = p & 63, pm = p >> 6;
      // pivot value and the old diagonal value, broadcast from their lanes
      double2 mine_p = make_double2(0.0, 0.0), mine_k = make_double2(0.0, 0.0);
#pragma unroll
      for (int m = 0; m < RS; ++m) {
        if (m == pm) mine_p = col[m];
        if (m == km) mine_k = col[m];
      }
      const double2 pv = make_double2(rl_d(mine_p.x, pl), rl_d(mine_p.y, pl));
      const double2 akk = make_double2(rl_d(mine_k.x, kl), rl_d(mine_k.y, kl));
      const bool ok = best > 0.0;
      const double2 pinv = ok ? cinv(pv) : make_double2(0.0, 0.0);
#pragma unroll
      for (int m = 0; m < RS; ++m) {
        const int i = lane + 64 * m;
        if (i > k && i < n && i != p) s_l[buf][i] = cmul(col[m], pinv);
      }
      if (lane == 0) {
        if (p != k) s_l[buf][p] = cmul(akk, pinv);
        s_p[buf] = p;
        s_pv[buf] = pv;
        s_dinv[k] = pinv;
        if (!ok && s_info == 0) s_info = k + 1;
      }
    }
    __syncthreads();
    const int p = s_p[buf];
    const double2 pv = s_pv[buf];
    const int pl = p & 63, pm = p >> 6;
    // ---- swap rows k and p in this wavefront's columns; broadcast the pivot row ----
    if (lane == pl) {
#pragma unroll
      for (int m = 0; m < RS; ++m)
        if (m == pm) {
#pragma unroll
          for (int jj = 0; jj < CS; ++jj) s_prow[wave][jj] = A[m][jj];
        }
    }
    if (p != k && lane == kl) {
#pragma unroll
      for (int m = 0; m < RS; ++m)
        if (m == km) {
#pragma unroll
          for (int jj = 0; jj < CS; ++jj) s_krow[wave][jj] = A[m][jj];
        }
    }
    wave_sync_lds();
    double2 u[CS];
#pragma unroll
    for (int jj = 0; jj < CS; ++jj) u[jj] = s_prow[wave][jj];
    if (p != k && lane == pl) {
#pragma unroll
      for (int m = 0; m < RS; ++m)
        if (m == pm) {
#pragma unroll
          for (int jj = 0; jj < CS; ++jj) A[m][jj] = s_krow[wave][jj];
        }
    }
    if (lane == kl) {
#pragma unroll
      for (int m = 0; m < RS; ++m)
        if (m == km) {
#pragma unroll
          for (int jj = 0; jj < CS; ++jj) {
            const int j = wave + NW * jj;
            A[m][jj] = (j == k) ? pv : u[jj];
          }
        }
    }
    // ---- rank-1 update of rows > k, columns > k ----
#pragma unroll
    for (int m = 0; m < RS; ++m) {
      if (64 * m + 63 <= k) continue;  // whole row slot finished (uniform)
      const int i = lane + 64 * m;
      const bool live = i > k && i < n;
      const double2 l = live ? s_l[buf][i] : make_double2(0.0, 0.0);
#pragma unroll
      for (int jj = 0; jj < CS; ++jj) {
        if (wave + NW * jj <= k) continue;  // column finished (uniform per wavefront)
        A[m][jj] = cfms(A[m][jj], l, u[jj]);
      }
    }
  }

  // ---- back substitution: packed U (column-major) and Y in LDS, solved by wavefront 0 ----
  double2* s_u = s_dyn;
  double2* s_y = s_dyn + (static_cast<size_t>(n) * (n + 1)) / 2;
#pragma unroll
  for (int m = 0; m < RS; ++m) {
    const int i = lane + 64 * m;
#pragma unroll
    for (int jj = 0; jj < CS; ++jj) {
      const int j = wave + NW * jj;
      if (i < n && j < n && i <= j) s_u[(j * (j + 1)) / 2 + i] = A[m][jj];
      if (i < n && j >= n && j < ncol) s_y[i * T + (j - n)] = A[m][jj];
    }
  }
  __syncthreads();
  if (wave != 0) return;
  for (int r = 0; r < T; ++r) {
    double2 y[RS];
#pragma unroll
    for (int m = 0; m < RS; ++m) {
      const int i = lane + 64 * m;
      y[m] = i < n ? s_y[i * T + r] : make_double2(0.0, 0.0);
    }
    for (int j = n - 1; j >= 0; --j) {
      const int jl = j & 63, jm = j >> 6;
      const double2 dj = s_dinv[j];
      double2 mine = make_double2(0.0, 0.0);
#pragma unroll
      for (int m = 0; m < RS; ++m)
        if (m == jm) mine = y[m];
      const double2 yj = make_double2(rl_d(mine.x, jl), rl_d(mine.y, jl));
      const double2 x = cmul(yj, dj);
      const double2* ucol = s_u + (j * (j + 1)) / 2;
#pragma unroll
      for (int m = 0; m < RS; ++m) {
        const int i = lane + 64 * m;
        if (i == j) y[m] = x;
        else if (i < j) y[m] = cfms(y[m], ucol[i], x);
      }
    }
#pragma unroll
    for (int m = 0; m < RS; ++m) {
      const int i = lane + 64 * m;
      if (i < n) a.out[ladder_out_index(a, chan, w, q, i, r)] = cscale(y[m], a.inv_beta);
    }
  }
  if (lane == 0 && a.info) a.info[(static_cast<size_t>(chan) * a.nw + w) * a.nq + q] = s_info;
}

struct ColVariant {
  int max_rows, max_cols, threads;
  void (*kernel)(LadderArgs);
};

#define DGAMI_COL_VARIANT(NW, RS, CS) \
  ColVariant { 64 * RS, NW * CS, NW * 64, ladder_col_kernel<NW, RS, CS> }

static const ColVariant kColVariants[] = {
    DGAMI_COL_VARIANT(4, 1, 16),  // n <= 64,  n + nb^2 <= 64
    DGAMI_COL_VARIANT(8, 1, 10),  // n <= 64,  n + nb^2 <= 80
    DGAMI_COL_VARIANT(8, 2, 12),  // n <= 128, n + nb^2 <= 96
    DGAMI_COL_VARIANT(8, 2, 16),  // n <= 128, n + nb^2 <= 128
};

int ladder_col_variant_count() { return static_cast<int>(sizeof(kColVariants) / sizeof(kColVariants[0])); }

bool ladder_col_variant_fits(int v, int n, int nrhs) {
  if (v < 0 || v >= ladder_col_variant_count()) return false;
  const size_t lds = (static_cast<size_t>(n) * (n + 1) / 2 + static_cast<size_t>(n) * nrhs) * sizeof(double2);
  return n <= kColVariants[v].max_rows && n + nrhs <= kColVariants[v].max_cols && lds <= 140 * 1024;
}

int ladder_col_variant(int n, int nrhs) {
  for (int v = 0; v < static_cast<int>(sizeof(kColVariants) / sizeof(kColVariants[0])); ++v) {
    const size_t lds = (static_cast<size_t>(n) * (n + 1) / 2 + static_cast<size_t>(n) * nrhs) * sizeof(double2);
    if (n <= kColVariants[v].max_rows && n + nrhs <= kColVariants[v].max_cols && lds <= 140 * 1024) return v;
  }
  return -1;
}

int ladder_col_launch(int variant, const LadderArgs& a, hipStream_t stream) {
  const ColVariant& cv = kColVariants[variant];
  const size_t items = static_cast<size_t>(a.nchan) * a.nw * a.nq;
  DGAMI_ARG_CHECK(items < (size_t(1) << 31), "ladder: too many problems in one call");
  const size_t lds = (static_cast<size_t>(a.n) * (a.n + 1) / 2 + static_cast<size_t>(a.n) * a.T) * sizeof(double2);
  allow_lds(cv.kernel, lds);
  hipLaunchKernelGGL(cv.kernel, dim3(static_cast<unsigned>(items)), dim3(cv.threads), lds, stream, a);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace dgami
