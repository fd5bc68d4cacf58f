// Plain-FP64 ladder solves for compound sizes beyond the register kernels.
//
// Batched, blocked right-looking LU with partial pivoting on the augmented
// matrices [A | B] held in a device workspace ([batch][n][ld], row major):
//   assemble -> for each panel of NB columns: panel factorisation (one
//   workgroup per problem, full-row swaps), TRSM of the block row, GEMM update
//   of the trailing block (all problems of the batch in one launch)
//   -> back substitution and output.
// The right-hand-side columns are part of every row, so the forward
// elimination of B happens inside the TRSM/GEMM updates.
#include "dgami.h"
#include "dgami_internal.hpp"
#include "dgami_launch.hpp"
#include "ladder_common.hpp"

namespace dgami {

constexpr int BLK_NB = 32;      // panel width
constexpr int GEMM_TILE = 64;   // trailing-update tile (rows and columns)
constexpr int BS_CB = 32;       // back-substitution chunk

struct BlockedState {
  double2* W;     // [batch][n][ld]
  double2* dinv;  // [batch][n] inverse pivots
  int* info;      // [batch]
  int n, ld, ncol;
};

// ---------------------------------------------------------------------------
__global__ void blk_assemble_kernel(LadderArgs a, BlockedState st, int item0, int batch) {
  const size_t per = static_cast<size_t>(st.n) * st.ld;
  const size_t total = per * batch;
  for (size_t idx = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; idx < total;
       idx += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int b = static_cast<int>(idx / per);
    const int rem = static_cast<int>(idx - static_cast<size_t>(b) * per);
    const int i = rem / st.ld, j = rem - (rem / st.ld) * st.ld;
    const int item = item0 + b;
    const int q = item % a.nq, w = (item / a.nq) % a.nw, chan = item / (a.nq * a.nw);
    double2 val = make_double2(0.0, 0.0);
    if (j < st.n) {
      const double2* G = ladder_matrix_ptr(a, chan, w, q);
      const double2* U = a.dense ? nullptr : a.u_r + chan * a.T * a.T;
      const double2* V = (a.v_r && !a.dense) ? a.v_r + (static_cast<size_t>(chan) * a.nq + q) * a.T * a.T : nullptr;
      val = ladder_elem(a, G, U, V, w, q, i, j);
    } else if (j < st.ncol) {
      val = ladder_rhs(a, i, j - st.n);
    }
    st.W[idx] = val;
  }
  if (blockIdx.x == 0)
    for (int b = threadIdx.x; b < batch; b += blockDim.x) st.info[b] = 0;
}

// ---------------------------------------------------------------------------
// panel factorisation of columns [p0, p0+NB) (one workgroup of 256 per problem)
__global__ void __launch_bounds__(256) blk_panel_kernel(BlockedState st, int p0) {
  const int b = blockIdx.x;
  const int n = st.n, ld = st.ld, ncol = st.ncol;
  double2* W = st.W + static_cast<size_t>(b) * n * ld;
  const int pend = p0 + BLK_NB < n ? p0 + BLK_NB : n;
  __shared__ double s_val[4];
  __shared__ int s_idx[4];
  __shared__ double2 s_row[BLK_NB];
  __shared__ double2 s_piv;
  __shared__ int s_p;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int jc = p0; jc < pend; ++jc) {
    double best = -1.0;
    int bi = 0x7fffffff;
    for (int i = jc + tid; i < n; i += 256) {
      const double m = cabs1(W[static_cast<size_t>(i) * ld + jc]);
      if (m > best) {
        best = m;
        bi = i;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(best, off);
      const int oi = __shfl_xor(bi, off);
      if (ov > best || (ov == best && oi < bi)) {
        best = ov;
        bi = oi;
      }
    }
    if (lane == 0) {
      s_val[wave] = best;
      s_idx[wave] = bi;
    }
    __syncthreads();
    if (tid == 0) {
      double bv = s_val[0];
      int bx = s_idx[0];
      for (int wv = 1; wv < 4; ++wv)
        if (s_val[wv] > bv || (s_val[wv] == bv && s_idx[wv] < bx)) {
          bv = s_val[wv];
          bx = s_idx[wv];
        }
      s_p = bx;
      const double2 pv = W[static_cast<size_t>(bx) * ld + jc];
      if (!(bv > 0.0)) {
        if (st.info[b] == 0) st.info[b] = jc + 1;
        s_piv = make_double2(0.0, 0.0);
      } else {
        s_piv = cinv(pv);
      }
      st.dinv[static_cast<size_t>(b) * n + jc] = s_piv;
    }
    __syncthreads();
    const int p = s_p;
    if (p != jc) {
      for (int c = p0 + tid; c < ncol; c += 256) {
        const double2 t = W[static_cast<size_t>(jc) * ld + c];
        W[static_cast<size_t>(jc) * ld + c] = W[static_cast<size_t>(p) * ld + c];
        W[static_cast<size_t>(p) * ld + c] = t;
      }
    }
    __syncthreads();
    const double2 dinv = s_piv;
    // pivot-row segment of the panel into LDS
    for (int c = jc + 1 + tid; c < pend; c += 256) s_row[c - p0] = W[static_cast<size_t>(jc) * ld + c];
    __syncthreads();
    const int width = pend - jc - 1;
    for (int i = jc + 1 + tid / 4; i < n; i += 64) {
      double2* row = W + static_cast<size_t>(i) * ld;
      const double2 l = cmul(row[jc], dinv);
      for (int c = jc + 1 + (tid & 3); c < jc + 1 + width; c += 4) row[c] = cfms(row[c], l, s_row[c - p0]);
    }
    __syncthreads();
    // store the multipliers after the update used the unscaled column
    for (int i = jc + 1 + tid; i < n; i += 256) {
      double2* e = W + static_cast<size_t>(i) * ld + jc;
      *e = cmul(*e, dinv);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// U12 = L11^-1 A12 for block row [p0, p0+NB), columns [p0+NB, ncol)
__global__ void __launch_bounds__(64) blk_trsm_kernel(BlockedState st, int p0, int ntile) {
  const int b = blockIdx.x / ntile;
  const int tile = blockIdx.x - b * ntile;
  const int n = st.n, ld = st.ld;
  double2* W = st.W + static_cast<size_t>(b) * n * ld;
  const int kb = (p0 + BLK_NB < n ? BLK_NB : n - p0);
  __shared__ double2 s_l[BLK_NB][BLK_NB];
  for (int e = threadIdx.x; e < kb * kb; e += 64) {
    const int r = e / kb, c = e - (e / kb) * kb;
    s_l[r][c] = W[static_cast<size_t>(p0 + r) * ld + p0 + c];
  }
  __syncthreads();
  const int col = p0 + kb + tile * 64 + threadIdx.x;
  if (col >= st.ncol) return;
  double2 x[BLK_NB];
#pragma unroll
  for (int r = 0; r < BLK_NB; ++r) {
    x[r] = make_double2(0.0, 0.0);
    if (r < kb) {
      double2 acc = W[static_cast<size_t>(p0 + r) * ld + col];
#pragma unroll
      for (int s = 0; s < r; ++s) acc = cfms(acc, s_l[r][s], x[s]);
      x[r] = acc;
      W[static_cast<size_t>(p0 + r) * ld + col] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// A22 -= L21 U12 (rows [p0+NB, n), columns [p0+NB, ncol)), 64x64 tiles, 256 threads
__global__ void __launch_bounds__(256) blk_gemm_kernel(BlockedState st, int p0, int tiles_r, int tiles_c) {
  const int per = tiles_r * tiles_c;
  const int b = blockIdx.x / per;
  const int t = blockIdx.x - b * per;
  const int tr = t / tiles_c, tcl = t - (t / tiles_c) * tiles_c;
  const int n = st.n, ld = st.ld, ncol = st.ncol;
  const int k0 = p0 + BLK_NB;
  double2* W = st.W + static_cast<size_t>(b) * n * ld;
  const int r0 = k0 + tr * GEMM_TILE, c0 = k0 + tcl * GEMM_TILE;
  __shared__ double2 s_a[GEMM_TILE][BLK_NB + 1];
  __shared__ double2 s_b[BLK_NB][GEMM_TILE + 1];
  const int tid = threadIdx.x;
  for (int e = tid; e < GEMM_TILE * BLK_NB; e += 256) {
    const int r = e / BLK_NB, s = e - (e / BLK_NB) * BLK_NB;
    s_a[r][s] = (r0 + r < n) ? W[static_cast<size_t>(r0 + r) * ld + p0 + s] : make_double2(0.0, 0.0);
  }
  for (int e = tid; e < GEMM_TILE * BLK_NB; e += 256) {
    const int s = e / GEMM_TILE, c = e - (e / GEMM_TILE) * GEMM_TILE;
    s_b[s][c] = (c0 + c < ncol) ? W[static_cast<size_t>(p0 + s) * ld + c0 + c] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  const int ty = tid / 16, tx = tid % 16;
  double2 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = make_double2(0.0, 0.0);
#pragma unroll 4
  for (int s = 0; s < BLK_NB; ++s) {
    double2 av[4], bv[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) av[i] = s_a[ty + 16 * i][s];
#pragma unroll
    for (int j = 0; j < 4; ++j) bv[j] = s_b[s][tx + 16 * j];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = cfma(acc[i][j], av[i], bv[j]);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = r0 + ty + 16 * i;
    if (r >= n) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = c0 + tx + 16 * j;
      if (c < ncol) {
        double2* e = W + static_cast<size_t>(r) * ld + c;
        *e = csub(*e, acc[i][j]);
      }
    }
  }
}

__device__ __forceinline__ double readlane_dbl(double v, int lane) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), lane);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// ---------------------------------------------------------------------------
// back substitution U x = y on the right-hand-side columns, then output / beta
__global__ void __launch_bounds__(256) blk_backsub_kernel(LadderArgs a, BlockedState st, int item0) {
  const int b = blockIdx.x;
  const int n = st.n, ld = st.ld, T = a.T;
  double2* W = st.W + static_cast<size_t>(b) * n * ld;
  const double2* dinv = st.dinv + static_cast<size_t>(b) * n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int c1 = n; c1 > 0; c1 -= BS_CB) {
    const int c0 = c1 > BS_CB ? c1 - BS_CB : 0;
    if (wave == 0) {
      const int row = c0 + lane;
      const bool live = lane < c1 - c0;
      const double2 di = live ? dinv[row] : make_double2(0.0, 0.0);
      for (int r = 0; r < T; ++r) {
        double2 y = live ? W[static_cast<size_t>(row) * ld + n + r] : make_double2(0.0, 0.0);
#pragma unroll
        for (int t = BS_CB - 1; t >= 0; --t) {
          if (c0 + t >= c1) continue;
          if (lane == t) y = cmul(y, di);
          const double2 x = make_double2(readlane_dbl(y.x, t), readlane_dbl(y.y, t));
          if (lane < t) y = cfms(y, W[static_cast<size_t>(row) * ld + c0 + t], x);
        }
        if (live) W[static_cast<size_t>(row) * ld + n + r] = y;
      }
    }
    __syncthreads();
    for (int idx = tid; idx < c0 * T; idx += 256) {
      const int i = idx / T, r = idx - (idx / T) * T;
      const double2* urow = W + static_cast<size_t>(i) * ld;
      double2 acc = urow[n + r];
      for (int t = c0; t < c1; ++t) acc = cfms(acc, urow[t], W[static_cast<size_t>(t) * ld + n + r]);
      W[static_cast<size_t>(i) * ld + n + r] = acc;
    }
    __syncthreads();
  }
  const int item = item0 + b;
  const int q = item % a.nq, w = (item / a.nq) % a.nw, chan = item / (a.nq * a.nw);
  for (int idx = tid; idx < n * T; idx += 256) {
    const int i = idx / T, r = idx - (idx / T) * T;
    a.out[ladder_out_index(a, chan, w, q, i, r)] = cscale(W[static_cast<size_t>(i) * ld + n + r], a.inv_beta);
  }
  if (tid == 0 && a.info) a.info[item] = st.info[b];
}

int ladder_blocked_fp64(const LadderArgs& a, hipStream_t stream) {
  const int n = a.n, ncol = n + a.T;
  const int ld = (ncol + 3) & ~3;
  const size_t items = static_cast<size_t>(a.nchan) * a.nw * a.nq;
  const size_t per = static_cast<size_t>(n) * ld * sizeof(double2);
  // batch so that the workspace stays below ~4 GiB
  size_t batch = (size_t(4) << 30) / per;
  if (batch < 1) batch = 1;
  if (batch > items) batch = items;
  if (batch > 65535) batch = 65535;
  BlockedState st;
  st.W = static_cast<double2*>(workspace(WS_LADDER, batch * per));
  st.dinv = static_cast<double2*>(workspace(WS_LADDER_AUX, batch * n * sizeof(double2) + batch * sizeof(int) + 256));
  if (!st.W || !st.dinv) return -2;
  st.info = reinterpret_cast<int*>(st.dinv + batch * n);
  st.n = n;
  st.ld = ld;
  st.ncol = ncol;
  for (size_t item0 = 0; item0 < items; item0 += batch) {
    const int nb = static_cast<int>(items - item0 < batch ? items - item0 : batch);
    const size_t tot = static_cast<size_t>(nb) * n * ld;
    const unsigned ablocks = ceil_div(tot, 256) > 262144u ? 262144u : ceil_div(tot, 256);
    hipLaunchKernelGGL(blk_assemble_kernel, dim3(ablocks), dim3(256), 0, stream, a, st, static_cast<int>(item0), nb);
    DGAMI_HIP_CHECK(hipGetLastError());
    for (int p0 = 0; p0 < n; p0 += BLK_NB) {
      hipLaunchKernelGGL(blk_panel_kernel, dim3(nb), dim3(256), 0, stream, st, p0);
      DGAMI_HIP_CHECK(hipGetLastError());
      const int k1 = p0 + (n - p0 < BLK_NB ? n - p0 : BLK_NB);
      if (k1 < ncol) {
        const int ntile = ceil_div(ncol - k1, 64);
        hipLaunchKernelGGL(blk_trsm_kernel, dim3(nb * ntile), dim3(64), 0, stream, st, p0, ntile);
        DGAMI_HIP_CHECK(hipGetLastError());
      }
      if (k1 < n) {  // then k1 == p0 + BLK_NB
        const int tr = ceil_div(n - k1, GEMM_TILE), tc = ceil_div(ncol - k1, GEMM_TILE);
        const size_t g = static_cast<size_t>(nb) * tr * tc;
        DGAMI_ARG_CHECK(g < (size_t(1) << 31), "ladder: trailing update grid too large");
        hipLaunchKernelGGL(blk_gemm_kernel, dim3(static_cast<unsigned>(g)), dim3(256), 0, stream, st, p0, tr, tc);
        DGAMI_HIP_CHECK(hipGetLastError());
      }
    }
    hipLaunchKernelGGL(blk_backsub_kernel, dim3(nb), dim3(256), 0, stream, a, st, static_cast<int>(item0));
    DGAMI_HIP_CHECK(hipGetLastError());
  }
  return 0;
}

}  // namespace dgami
