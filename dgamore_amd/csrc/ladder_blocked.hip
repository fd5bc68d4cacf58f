// Ladder solves for compound sizes beyond the register kernels (n > 128).
//
// Batched, blocked right-looking LU with partial pivoting on a device
// workspace W[batch][n][ld] (row major), templated on the storage type:
// complex128 for the plain FP64 form, complex64 for the mixed form (whose
// factors are refined in FP64 by ladder_mixed.hip). Per panel of NB columns:
//   1. panel kernel: one workgroup per problem holds the panel rows in
//      registers (one thread per row or per few rows), factors the NB columns
//      with partial pivoting (two barriers per column, no global round trips),
//      and writes L/U of the panel, the pivots and the row map of the panel's
//      interchanges;
//   2. rows kernel: for every column outside the panel, applies the panel's
//      interchanges as one gather through the row map (independent loads, not
//      a dependent chain of swaps) and, on trailing columns, the unit-lower
//      triangular solve U12 = L11^-1 A12;
//   3. trailing update A22 -= L21 U12 for all problems of the batch in one launch.
// With the right-hand sides carried as extra columns (plain form), back
// substitution on them finishes the solve.
#include "dgami.h"
#include "dgami_internal.hpp"
#include "dgami_launch.hpp"
#include "ladder_blocked.hpp"
#include "ladder_common.hpp"
#include "tri_solve.hpp"

#include <cstdlib>
#include <string>
#include <type_traits>
#include <utility>

namespace dgami {

constexpr int MAP_STRIDE = BLK_MAP_STRIDE;  // interchange map per problem: count, dst[NB], src[NB], dest[NB]

// ---------------------------------------------------------------------------
// grid (ceil(n / ASM_ROWS), batch), 256 threads: ASM_ROWS rows of one problem per workgroup.
constexpr int ASM_ROWS = 4;

template <typename S>
__global__ void __launch_bounds__(256) blk_assemble_kernel(LadderArgs a, BlockedState<S> st, int item0) {
  const int b = blockIdx.y;
  const int item = item0 + b;
  const int q = item % a.nq, w = (item / a.nq) % a.nw, chan = item / (a.nq * a.nw);
  const double2* G = ladder_matrix_ptr(a, chan, w, q);
  const double2* U = a.dense ? nullptr : a.u_r + chan * a.T * a.T;
  const double2* V = (a.v_r && !a.dense) ? a.v_r + (static_cast<size_t>(chan) * a.nq + q) * a.T * a.T : nullptr;
  S* Wb = st.W + static_cast<size_t>(b) * st.n * st.ld;
  const int r1 = min(st.n, (blockIdx.x + 1) * ASM_ROWS);
  for (int i = blockIdx.x * ASM_ROWS; i < r1; ++i) {
    for (int j = threadIdx.x; j < st.ld; j += 256) {
      double2 val = make_double2(0.0, 0.0);
      if (j < st.n)
        val = ladder_elem(a, G, U, V, w, q, i, j);
      else if (j < st.ncol)
        val = ladder_rhs(a, i, j - st.n);
      Wb[static_cast<size_t>(i) * st.ld + j] = from_c128<S>(val);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) st.info[b] = 0;
}

// ---------------------------------------------------------------------------
// Panel factorisation: thread t owns rows p0 + t + blockDim*m (m < RPT) of columns [p0, p0+NB).
// Each column is one instance of panel_column<C>, expanded at compile time so that every
// register index is a constant.
// Pivot search per column: every thread takes the largest |Re|+|Im| of its rows, a
// wavefront reduces the keys with DPP row shifts/broadcasts (the key is the upper word of
// the FP64 magnitude plus one, so any valid row beats the invalid ones; candidates equal
// in that word are interchangeable pivots), the wavefront leaders meet in LDS, and the
// owner of the winning row publishes it for the rank-1 update (two barriers per column).
__device__ __forceinline__ unsigned wave_max_u32(unsigned v) {
  v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x111, 0xf, 0xf, false)));
  v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x112, 0xf, 0xf, false)));
  v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x114, 0xf, 0xf, false)));
  v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x118, 0xf, 0xf, false)));
  v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x142, 0xa, 0xf, false)));
  v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x143, 0xc, 0xf, false)));
  return static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}

template <typename S, int NB>
struct PanelShared {
  unsigned key[2][16];
  int idx[2][16];
  S prow[2][NB];
  S krow[2][NB];
  int piv[NB];
  int mdst[2 * NB], msrc[2 * NB];
};

template <typename S, int NB, int RPT, int C>
__device__ __forceinline__ void panel_column(S (&P)[RPT][NB], PanelShared<S, NB>& sh, BlockedState<S>& st, int b,
                                             int p0, int nbw) {
  if (C >= nbw) return;
  const int n = st.n, TH = blockDim.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwave = TH >> 6;
  const int col = p0 + C;
  constexpr int buf = C & 1;
  double best = -1.0;
  int bidx = 0x7fffffff;
#pragma unroll
  for (int m = 0; m < RPT; ++m) {
    const int r = p0 + tid + TH * m;
    const double v = static_cast<double>(cabs1(P[m][C]));
    if (r >= col && r < n && v > best) {
      best = v;
      bidx = r;
    }
  }
  {
    const unsigned key = best >= 0.0 ? static_cast<unsigned>(__double_as_longlong(best) >> 32) + 1u : 0u;
    const unsigned wkey = wave_max_u32(key);
    const unsigned long long hit = __ballot(key == wkey);
    const int wlane = __ffsll(static_cast<long long>(hit)) - 1;
    const int widx = __builtin_amdgcn_readlane(bidx, wlane);
    if (lane == 0) {
      sh.key[buf][wave] = wkey;
      sh.idx[buf][wave] = widx;
    }
  }
  __syncthreads();
  unsigned gkey = sh.key[buf][0];
  bidx = sh.idx[buf][0];
  for (int wv = 1; wv < nwave; ++wv) {
    const unsigned k2 = sh.key[buf][wv];
    const int oi = sh.idx[buf][wv];
    if (k2 > gkey || (k2 == gkey && oi < bidx)) {
      gkey = k2;
      bidx = oi;
    }
  }
  best = gkey > 1u ? 1.0 : 0.0;  // nonzero pivot magnitude
  const int p = bidx;
  constexpr int mk = 0;  // column C's diagonal row p0 + C lies in the first register row (C < NB <= TH)
  if constexpr (RPT == 1) {
    if (tid == p - p0) {
#pragma unroll
      for (int cc = 0; cc < NB; ++cc) sh.prow[buf][cc] = P[0][cc];
    }
  } else {
#pragma unroll
    for (int m = 0; m < RPT; ++m)
      if (p0 + tid + TH * m == p) {
#pragma unroll
        for (int cc = 0; cc < NB; ++cc) sh.prow[buf][cc] = P[m][cc];
      }
  }
  if (p != col && tid == C) {
#pragma unroll
    for (int cc = 0; cc < NB; ++cc) sh.krow[buf][cc] = P[mk][cc];
  }
  if (tid == 0) sh.piv[C] = p;
  __syncthreads();
  const bool ok = best > 0.0;
  const S pinv = ok ? cinv(sh.prow[buf][C]) : czero<S>();
  if (tid == 0) {
    st.dinv[static_cast<size_t>(b) * n + col] = to_c128(pinv);
    if (!ok && st.info[b] == 0) st.info[b] = col + 1;
  }
#pragma unroll
  for (int m = 0; m < RPT; ++m) {
    const int r = p0 + tid + TH * m;
    if (r == col) {
#pragma unroll
      for (int cc = 0; cc < NB; ++cc) P[m][cc] = sh.prow[buf][cc];
    } else if (r > col && r < n) {
      if (r == p) {
#pragma unroll
        for (int cc = 0; cc < NB; ++cc) P[m][cc] = sh.krow[buf][cc];
      }
      const S l = cmul(P[m][C], pinv);
      P[m][C] = l;
#pragma unroll
      for (int cc = C + 1; cc < NB; ++cc) P[m][cc] = cfms(P[m][cc], l, sh.prow[buf][cc]);
    }
  }
}

template <typename S, int NB, int RPT, int... Cs>
__device__ __forceinline__ void panel_columns(S (&P)[RPT][NB], PanelShared<S, NB>& sh, BlockedState<S>& st, int b,
                                              int p0, int nbw, std::integer_sequence<int, Cs...>) {
  (panel_column<S, NB, RPT, Cs>(P, sh, st, b, p0, nbw), ...);
}

template <typename S, int NB, int RPT, int MAXTH>
__global__ void __launch_bounds__(MAXTH) blk_panel_kernel(BlockedState<S> st, int p0) {
  const int b = blockIdx.x;
  const int n = st.n, ld = st.ld;
  const int TH = blockDim.x;
  const int tid = threadIdx.x;
  S* W = st.W + static_cast<size_t>(b) * n * ld;
  const int nbw = (n - p0 < NB) ? n - p0 : NB;
  __shared__ PanelShared<S, NB> sh;

  S P[RPT][NB];
#pragma unroll
  for (int m = 0; m < RPT; ++m) {
    const int r = p0 + tid + TH * m;
#pragma unroll
    for (int c = 0; c < NB; ++c) P[m][c] = (r < n && c < nbw) ? W[static_cast<size_t>(r) * ld + p0 + c] : czero<S>();
  }
  panel_columns<S, NB, RPT>(P, sh, st, b, p0, nbw, std::make_integer_sequence<int, NB>{});
#pragma unroll
  for (int m = 0; m < RPT; ++m) {
    const int r = p0 + tid + TH * m;
#pragma unroll
    for (int c = 0; c < NB; ++c)
      if (r < n && c < nbw) W[static_cast<size_t>(r) * ld + p0 + c] = P[m][c];
  }
  // Interchange map of the panel. A row below the panel only ever exchanges its value with a
  // panel row, so after the NB interchanges every row below the panel holds a former panel
  // row's value, and the former values of rows below the panel end in panel rows:
  //   outer[e] = (dst, src): panel row dst receives row src (src below the panel), e < count;
  //   dest[c]:  panel row p0 + c's former value goes to row dest[c] (-1: stays).
  if (tid == 0) {
    int* dst = sh.mdst;
    int* src = sh.msrc;
    int cnt = 0;
    for (int c = 0; c < nbw; ++c) {
      const int j = p0 + c, p = sh.piv[c];
      st.piv[static_cast<size_t>(b) * n + j] = p;
      if (p == j) continue;
      int ij = -1, ip = -1;
      for (int e = 0; e < cnt; ++e) {
        if (dst[e] == j) ij = e;
        if (dst[e] == p) ip = e;
      }
      if (ij < 0) {
        ij = cnt++;
        dst[ij] = j;
        src[ij] = j;
      }
      if (ip < 0) {
        ip = cnt++;
        dst[ip] = p;
        src[ip] = p;
      }
      const int t = src[ij];
      src[ij] = src[ip];
      src[ip] = t;
    }
    int* map = st.rowmap + static_cast<size_t>(b) * MAP_STRIDE;
    for (int c = 0; c < NB; ++c) map[1 + 2 * NB + c] = -1;
    int outer = 0;
    for (int e = 0; e < cnt; ++e) {
      if (src[e] == dst[e]) continue;
      if (src[e] < p0 + nbw) {
        map[1 + 2 * NB + (src[e] - p0)] = dst[e];
      } else {
        map[1 + outer] = dst[e];
        map[1 + NB + outer] = src[e];
        ++outer;
      }
    }
    map[0] = outer;
  }
}

// ---------------------------------------------------------------------------
// Interchanges for every column outside the panel, then U12 = L11^-1 A12 on trailing columns.
// grid: xcd_grid(batch, ntile), 64 threads; tile t covers columns t*64 .. t*64+63 of [0, ncol).
template <typename S, int NB>
__global__ void __launch_bounds__(64) blk_rows_kernel(BlockedState<S> st, int p0, int ntile, int nprob) {
  int b, tile;
  if (!xcd_tile(blockIdx.x, nprob, ntile, b, tile)) return;
  const int n = st.n, ld = st.ld;
  S* W = st.W + static_cast<size_t>(b) * n * ld;
  const int nbw = (n - p0 < NB) ? n - p0 : NB;
  __shared__ S s_l[NB][NB + 1];
  __shared__ int s_dst[NB], s_src[NB], s_dest[NB];
  const int* map = st.rowmap + static_cast<size_t>(b) * MAP_STRIDE;
  const int outer = map[0];
  for (int e = threadIdx.x; e < NB; e += 64) {
    s_dst[e] = e < outer ? map[1 + e] : 0;
    s_src[e] = e < outer ? map[1 + NB + e] : 0;
    s_dest[e] = map[1 + 2 * NB + e];
  }
  for (int e = threadIdx.x; e < nbw * nbw; e += 64) {
    const int r = e / nbw, c = e - (e / nbw) * nbw;
    s_l[r][c] = W[static_cast<size_t>(p0 + r) * ld + p0 + c];
  }
  __syncthreads();
  const int col = tile * 64 + threadIdx.x;
  if (col >= st.ncol || (col >= p0 && col < p0 + nbw)) return;
  if (col < p0 && !st.swap_left) return;
  // former panel values first (registers), then rows from below the panel into panel rows,
  // then the former panel values to their destinations
  S pv[NB];
#pragma unroll
  for (int c = 0; c < NB; ++c)
    pv[c] = (c < nbw && s_dest[c] >= 0) ? W[static_cast<size_t>(p0 + c) * ld + col] : czero<S>();
  for (int e = 0; e < outer; ++e) W[static_cast<size_t>(s_dst[e]) * ld + col] = W[static_cast<size_t>(s_src[e]) * ld + col];
#pragma unroll
  for (int c = 0; c < NB; ++c)
    if (c < nbw && s_dest[c] >= 0) W[static_cast<size_t>(s_dest[c]) * ld + col] = pv[c];
  if (col < p0) return;  // left of the panel: interchanges only
  S x[NB];
#pragma unroll
  for (int r = 0; r < NB; ++r) {
    x[r] = czero<S>();
    if (r < nbw) {
      S acc = W[static_cast<size_t>(p0 + r) * ld + col];
#pragma unroll
      for (int s = 0; s < r; ++s) acc = cfms(acc, s_l[r][s], x[s]);
      x[r] = acc;
      W[static_cast<size_t>(p0 + r) * ld + col] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// A22 -= L21 U12 (rows [p0+NB, n), columns [p0+NB, ncol)), 64x64 tiles, 256 threads.
template <typename S, int NB>
__global__ void __launch_bounds__(256) blk_gemm_kernel(BlockedState<S> st, int p0, int tiles_r, int tiles_c,
                                                       int nprob) {
  constexpr int TILE = 64;
  const int per = tiles_r * tiles_c;
  int b, t;
  if (!xcd_tile(blockIdx.x, nprob, per, b, t)) return;
  const int tr = t / tiles_c, tcl = t - (t / tiles_c) * tiles_c;
  const int n = st.n, ld = st.ld, ncol = st.ncol;
  const int k0 = p0 + NB;
  S* W = st.W + static_cast<size_t>(b) * n * ld;
  const int r0 = k0 + tr * TILE, c0 = k0 + tcl * TILE;
  __shared__ S s_a[TILE][NB + 1];
  __shared__ S s_b[NB][TILE + 1];
  const int tid = threadIdx.x;
  const int ty = tid / 16, tx = tid % 16;
  // the C tile is loaded first so that its latency overlaps the operand staging
  S acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = r0 + ty + 16 * i, c = c0 + tx + 16 * j;
      acc[i][j] = (r < n && c < ncol) ? W[static_cast<size_t>(r) * ld + c] : czero<S>();
    }
  for (int e = tid; e < TILE * NB; e += 256) {
    const int r = e / NB, s = e - (e / NB) * NB;
    s_a[r][s] = (r0 + r < n) ? W[static_cast<size_t>(r0 + r) * ld + p0 + s] : czero<S>();
  }
  for (int e = tid; e < TILE * NB; e += 256) {
    const int s = e / TILE, c = e - (e / TILE) * TILE;
    s_b[s][c] = (c0 + c < ncol) ? W[static_cast<size_t>(p0 + s) * ld + c0 + c] : czero<S>();
  }
  __syncthreads();
#pragma unroll 4
  for (int s = 0; s < NB; ++s) {
    S av[4], bv[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) av[i] = s_a[ty + 16 * i][s];
#pragma unroll
    for (int j = 0; j < 4; ++j) bv[j] = s_b[s][tx + 16 * j];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = cfms(acc[i][j], av[i], bv[j]);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = r0 + ty + 16 * i;
    if (r >= n) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = c0 + tx + 16 * j;
      if (c < ncol) W[static_cast<size_t>(r) * ld + c] = acc[i][j];
    }
  }
}

// Same update on the matrix cores: each wavefront owns a 32x32 quarter of the tile as 2x2
// blocks of 16x16, and a complex product becomes two real MFMA chains over the interleaved
// k index (re/im): C_re -= [a_re, -a_im] . [b_re; b_im], C_im -= [a_im, a_re] . [b_re; b_im].
// Operand layout of the 16x16x4 MFMA (lane l): a = A[l%16][l/16], b = B[l/16][l%16],
// d[e] = D[4*(l/16)+e][l%16].
template <typename S>
struct Mfma16x16x4;
template <>
struct Mfma16x16x4<float2> {
  using R = float;
  typedef float V4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ V4 op(R a, R b, V4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};

template <typename S, int NB>
__global__ void __launch_bounds__(256) blk_gemm_mfma_kernel(BlockedState<S> st, int p0, int tiles_r, int tiles_c,
                                                            int nprob) {
  using M = Mfma16x16x4<S>;
  using R = typename M::R;
  using V4 = typename M::V4;
  constexpr int TILE = 64;
  const int per = tiles_r * tiles_c;
  int b, t;
  if (!xcd_tile(blockIdx.x, nprob, per, b, t)) return;
  const int tr = t / tiles_c, tcl = t - (t / tiles_c) * tiles_c;
  const int n = st.n, ld = st.ld, ncol = st.ncol;
  const int k0 = p0 + NB;
  S* W = st.W + static_cast<size_t>(b) * n * ld;
  const int r0 = k0 + tr * TILE, c0 = k0 + tcl * TILE;
  __shared__ S s_a[TILE][NB + 1];
  __shared__ S s_b[NB][TILE + 1];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int l16 = lane & 15, kq = lane >> 4, kc = kq >> 1;
  const bool odd = (kq & 1) != 0;
  // accumulators start from the C tile (loaded first, overlapping the operand staging);
  // the A operand enters negated, so the chains compute C - A B
  V4 accr[2][2], acci[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int c = c0 + wc + 16 * j + l16;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = r0 + wr + 16 * i + 4 * kq + e;
        const S v = (r < n && c < ncol) ? W[static_cast<size_t>(r) * ld + c] : czero<S>();
        accr[i][j][e] = v.x;
        acci[i][j][e] = v.y;
      }
    }
  for (int e = tid; e < TILE * NB; e += 256) {
    const int r = e / NB, s = e - (e / NB) * NB;
    s_a[r][s] = (r0 + r < n) ? W[static_cast<size_t>(r0 + r) * ld + p0 + s] : czero<S>();
  }
  for (int e = tid; e < TILE * NB; e += 256) {
    const int s = e / TILE, c = e - (e / TILE) * TILE;
    s_b[s][c] = (c0 + c < ncol) ? W[static_cast<size_t>(p0 + s) * ld + c0 + c] : czero<S>();
  }
  __syncthreads();
#pragma unroll
  for (int kk = 0; kk < NB; kk += 2) {
    const int k = kk + kc;
    R ar[2], ai[2], bv[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const S v = s_a[wr + 16 * i + l16][k];
      ar[i] = odd ? v.y : -v.x;
      ai[i] = odd ? -v.x : -v.y;
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const S v = s_b[k][wc + 16 * j + l16];
      bv[j] = odd ? v.y : v.x;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        accr[i][j] = M::op(ar[i], bv[j], accr[i][j]);
        acci[i][j] = M::op(ai[i], bv[j], acci[i][j]);
      }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int c = c0 + wc + 16 * j + l16;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = r0 + wr + 16 * i + 4 * kq + e;
        if (r < n && c < ncol) {
          S v;
          v.x = accr[i][j][e];
          v.y = acci[i][j][e];
          W[static_cast<size_t>(r) * ld + c] = v;
        }
      }
    }
}

__global__ void blk_info_kernel(const int* __restrict__ st_info, int32_t* __restrict__ info, int item0, int nitems) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nitems) info[item0 + b] = st_info[b];
}

int blocked_copy_info(const int* st_info, int32_t* info, int item0, int nitems, hipStream_t stream) {
  if (!info) return 0;
  hipLaunchKernelGGL(blk_info_kernel, dim3(ceil_div(nitems, 256)), dim3(256), 0, stream, st_info, info, item0, nitems);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// Complex64 trailing updates run on the FP32 matrix cores; DGAMI_BLOCKED_GEMM=vector selects
// the vector-FMA tile kernel instead (for comparisons). Complex128 updates always use the
// vector-FMA kernel: the FP64 MFMA (v_mfma_f64_16x16x4_f64) returned wrong products on
// gfx950 when the compiler rewrote its source registers right after issue, which source
// code cannot prevent.
static bool use_vector_gemm() {
  const char* v = getenv("DGAMI_BLOCKED_GEMM");
  return v && std::string(v) == "vector";
}

template <typename S, int NB, int RPT, int MAXTH>
static int blocked_factor_nb(const LadderArgs& a, BlockedState<S>& st, int item0, int nb, hipStream_t stream) {
  const int n = st.n, ld = st.ld, ncol = st.ncol;
  DGAMI_ARG_CHECK(nb <= 65535, "ladder: batch too large");
  hipLaunchKernelGGL(blk_assemble_kernel<S>, dim3(ceil_div(n, ASM_ROWS), nb), dim3(256), 0, stream, a, st, item0);
  DGAMI_HIP_CHECK(hipGetLastError());
  for (int p0 = 0; p0 < n; p0 += NB) {
    const int rows = n - p0;
    int th = ((rows + RPT - 1) / RPT + 63) / 64 * 64;
    DGAMI_ARG_CHECK(th <= MAXTH, "ladder: compound size too large for the panel kernel");
    hipLaunchKernelGGL((blk_panel_kernel<S, NB, RPT, MAXTH>), dim3(nb), dim3(th), 0, stream, st, p0);
    DGAMI_HIP_CHECK(hipGetLastError());
    const int ntile = ceil_div(ncol, 64);
    hipLaunchKernelGGL((blk_rows_kernel<S, NB>), dim3(static_cast<unsigned>(xcd_grid(nb, ntile))), dim3(64), 0, stream,
                       st, p0, ntile, nb);
    DGAMI_HIP_CHECK(hipGetLastError());
    const int k1 = p0 + NB;
    if (k1 < n) {
      const int tr = ceil_div(n - k1, 64), tc = ceil_div(ncol - k1, 64);
      const size_t g = xcd_grid(nb, static_cast<size_t>(tr) * tc);
      DGAMI_ARG_CHECK(g < (size_t(1) << 31), "ladder: trailing update grid too large");
      bool vector = true;
      if constexpr (std::is_same<S, float2>::value) vector = use_vector_gemm();
      if (vector) {
        hipLaunchKernelGGL((blk_gemm_kernel<S, NB>), dim3(static_cast<unsigned>(g)), dim3(256), 0, stream, st, p0, tr,
                           tc, nb);
      } else if constexpr (std::is_same<S, float2>::value) {
        hipLaunchKernelGGL((blk_gemm_mfma_kernel<S, NB>), dim3(static_cast<unsigned>(g)), dim3(256), 0, stream, st, p0,
                           tr, tc, nb);
      }
      DGAMI_HIP_CHECK(hipGetLastError());
    }
  }
  return 0;
}

template <>
int blocked_factor<double2>(const LadderArgs& a, BlockedState<double2>& st, int item0, int nb, hipStream_t stream) {
  if (st.n <= 768) return blocked_factor_nb<double2, 32, 1, 768>(a, st, item0, nb, stream);
  return blocked_factor_nb<double2, 16, 3, 512>(a, st, item0, nb, stream);
}

template <>
int blocked_factor<float2>(const LadderArgs& a, BlockedState<float2>& st, int item0, int nb, hipStream_t stream) {
  if (st.n <= 1024) return blocked_factor_nb<float2, 32, 1, 1024>(a, st, item0, nb, stream);
  return blocked_factor_nb<float2, 16, 3, 512>(a, st, item0, nb, stream);
}

int ladder_blocked_fp64(const LadderArgs& a, hipStream_t stream) {
  const int n = a.n, ncol = n + a.T;
  const int ld = (ncol + 3) & ~3;
  DGAMI_ARG_CHECK(n <= 1536, "ladder: compound size above 1536 is not supported");
  const size_t items = static_cast<size_t>(a.nchan) * a.nw * a.nq;
  const size_t per = static_cast<size_t>(n) * ld * sizeof(double2);
  // MI355X holds 288 GB of HBM: factor up to 48 GiB of problems per pass
  size_t batch = (size_t(48) << 30) / per;
  if (batch < 1) batch = 1;
  if (batch > items) batch = items;
  if (batch > 65535) batch = 65535;
  BlockedState<double2> st;
  if (!blocked_alloc(st, n, ld, ncol, batch, WS_LADDER)) return -2;
  const int nblk = static_cast<int>(ceil_div(n, TRI_B));
  double2* tri = static_cast<double2*>(workspace(WS_TRI, batch * nblk * 2 * TRI_B * TRI_B * sizeof(double2)));
  if (!tri) return -2;
  for (size_t item0 = 0; item0 < items; item0 += batch) {
    const int nb = static_cast<int>(items - item0 < batch ? items - item0 : batch);
    int rc = blocked_factor<double2>(a, st, static_cast<int>(item0), nb, stream);
    if (rc) return rc;
    // back substitution on the carried right-hand sides, x / beta straight to the output
    TriFactors<double2> f{st.W, tri, st.dinv, n, ld, nblk};
    rc = tri_diag_inverse(f, nb, false, stream);
    if (rc) return rc;
    TriVec<double2> y{st.W + n, static_cast<size_t>(n) * ld, ld};
    rc = tri_backward(f, nb, a.T, y, TriVec<double2>{nullptr, 0, 0}, &a, static_cast<int>(item0), stream);
    if (rc) return rc;
    rc = blocked_copy_info(st.info, a.info, static_cast<int>(item0), nb, stream);
    if (rc) return rc;
  }
  return 0;
}

}  // namespace dgami
