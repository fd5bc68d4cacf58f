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
#include "ladder_tune.hpp"
#include "tri_solve.hpp"

#include <algorithm>
#include <cstdlib>
#include <string>
#include <type_traits>
#include <utility>

namespace dgami {

constexpr int MAP_STRIDE = BLK_MAP_STRIDE;  // interchange map per problem: count, dst[NB], src[NB], dest[NB]

// ---------------------------------------------------------------------------
// grid (ceil(n / ASM_ROWS), ceil(batch / ASM_QG)), 256 threads: ASM_ROWS rows of ASM_QG
// consecutive problems per workgroup. Consecutive problems are consecutive q-points of one
// (channel, w) and share Gamma(channel, w): each Gamma element is loaded once for the group and
// reused for every problem that shares it (the load is repeated only where the group crosses
// to another (channel, w), and for every problem of the dense form). Each thread owns column
// pairs (j, j + 1); their orbital/frequency split is computed once and reused for every row, and
// each pair is written as one 16- or 32-byte store.
constexpr int ASM_ROWS = 8;
constexpr int ASM_QG = 8;

template <typename S, bool NT>
__device__ __forceinline__ void store_pair(S* dst, double2 x, double2 y) {
  typedef float F4 __attribute__((ext_vector_type(4)));
  typedef double D2 __attribute__((ext_vector_type(2)));
  if constexpr (std::is_same<S, float2>::value) {
    const F4 v = F4{static_cast<float>(x.x), static_cast<float>(x.y), static_cast<float>(y.x), static_cast<float>(y.y)};
    if constexpr (NT)
      __builtin_nontemporal_store(v, reinterpret_cast<F4*>(dst));
    else
      *reinterpret_cast<F4*>(dst) = v;
  } else {
    if constexpr (NT) {
      __builtin_nontemporal_store(D2{x.x, x.y}, reinterpret_cast<D2*>(dst));
      __builtin_nontemporal_store(D2{y.x, y.y}, reinterpret_cast<D2*>(dst + 1));
    } else {
      dst[0] = x;
      dst[1] = y;
    }
  }
}

template <typename S>
__device__ __forceinline__ S to_storage(double2 x) {
  if constexpr (std::is_same<S, float2>::value)
    return make_float2(static_cast<float>(x.x), static_cast<float>(x.y));
  else
    return x;
}

// With a fused first block (st.fuse_e = E > 0) only columns [0, E) of every row and the rows
// [0, E) are written, and the rows' flags are cleared (E is a multiple of ASM_ROWS).
template <typename S, bool NT>
__global__ void __launch_bounds__(256) blk_assemble_kernel(LadderArgs a, BlockedState<S> st, int item0, int nbatch) {
  const int b0 = blockIdx.y * ASM_QG;
  const int ng = min(ASM_QG, nbatch - b0);
  const int n = st.n, ncol = st.ncol, nv = a.nv, T = a.T;
  const int i0 = blockIdx.x * ASM_ROWS, r1 = min(n, i0 + ASM_ROWS);
  const int E = st.fuse_e;
  const int ld = (E > 0 && i0 >= E) ? E : st.ld;  // columns written in these rows (ld and E are even)
  if (E > 0)
    for (int e = threadIdx.x; e < ng * (r1 - i0); e += blockDim.x)
      st.rowflag[static_cast<size_t>(b0 + e / (r1 - i0)) * n + i0 + e % (r1 - i0)] = 0;
  const double2 zero = make_double2(0.0, 0.0);
  // the group's problems: matrix, U_r + V_r(q) and chi0^-1 bases, output matrix
  const double2* Gk[ASM_QG];
  const double2* Uk[ASM_QG];
  const double2* Vk[ASM_QG];
  const double2* Ck[ASM_QG];
  int wk[ASM_QG], qk[ASM_QG];
#pragma unroll
  for (int k = 0; k < ASM_QG; ++k) {
    const int item = item0 + b0 + min(k, ng - 1);
    const int q = item % a.nq, w = (item / a.nq) % a.nw, chan = item / (a.nq * a.nw);
    Gk[k] = ladder_matrix_ptr(a, chan, w, q);
    Uk[k] = a.dense ? nullptr : a.u_r + chan * T * T;
    Vk[k] = (a.v_r && !a.dense) ? a.v_r + (static_cast<size_t>(chan) * a.nq + q) * T * T : nullptr;
    Ck[k] = a.dense ? nullptr : a.chi0_inv + static_cast<size_t>(w) * T * T * nv * a.nq + q;
    wk[k] = w;
    qk[k] = q;
  }
  for (int jb = 2 * threadIdx.x; jb < ld; jb += 2 * blockDim.x) {  // ld is even
    int cj[2], vj[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int j = jb + h;
      cj[h] = j < n ? j / nv : 0;
      vj[h] = j < n ? j - cj[h] * nv : -1;
    }
    for (int i = i0; i < r1; ++i) {
      const int ri = i / nv, vi = i - ri * nv;
      double2 g[2] = {zero, zero};
#pragma unroll
      for (int k = 0; k < ASM_QG; ++k) {
        if (k >= ng) break;
        if (k == 0 || Gk[k] != Gk[k - 1]) {
#pragma unroll
          for (int h = 0; h < 2; ++h) g[h] = jb + h < n ? Gk[k][static_cast<size_t>(i) * n + jb + h] : zero;
        }
        double2 val[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int j = jb + h;
          val[h] = zero;
          if (j < n) {
            if (a.dense) {
              val[h] = g[h];
            } else {
              double2 uv = Uk[k][ri * T + cj[h]];
              if (Vk[k]) uv = cadd(uv, Vk[k][ri * T + cj[h]]);
              val[h] = cscale(csub(g[h], uv), a.inv_beta2);
              if (vi == vj[h]) val[h] = cadd(val[h], Ck[k][(static_cast<size_t>(ri * T + cj[h]) * nv + vi) * a.nq]);
            }
          } else if (j < ncol) {
            val[h] = a.rhs_mode == LADDER_RHS_VERTEX ? ladder_vertex_part(a, Gk[k], Uk[k], Vk[k], i, ladder_rhs_col(a, j - n))
                                                     : ladder_rhs_at(a, wk[k], qk[k], i, j - n);
          }
        }
        store_pair<S, NT>(st.W + static_cast<size_t>(b0 + k) * n * st.ld + static_cast<size_t>(i) * st.ld + jb, val[0],
                          val[1]);
      }
    }
  }
  if (E > 0 && i0 >= E && st.g32 != nullptr) {
    // complex64 fused block: the trailing elements of these rows that the complex64 vertex does
    // not hold, the v = v' diagonal of every orbital pair and the right-hand sides, are written here
    const int nr = ncol - n, per = T + nr, rows = r1 - i0;
    for (int e = threadIdx.x; e < rows * ng * per; e += blockDim.x) {
      const int s2 = e % per, k = (e / per) % ng, i = i0 + e / (per * ng);
      const int c = s2 < T ? s2 * nv + (i % nv) : n + (s2 - T);
      if (c < E) continue;
      const LadderSrc src = ladder_src(a, item0 + b0 + k);
      st.W[static_cast<size_t>(b0 + k) * n * st.ld + static_cast<size_t>(i) * st.ld + c] =
          to_storage<S>(ladder_aug_at(a, src, i, c));
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < ng) st.info[b0 + threadIdx.x] = 0;
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

template <typename S, int NB>
struct PanelShared {
  unsigned key[2][16];
  int idx[2][16];
  S prow[2][NB];
  int piv[NB];
  int mdst[2 * NB], msrc[2 * NB];
  int mdest[NB];
  int ocount;
  int rowat[BLK_MAX_N];  // rowat[i]: the original row now at row p0 + i (after the interchanges so far)
  int pos[BLK_MAX_N];    // pos[r - p0]: where original row r is now, relative to p0
};

// Rows stay in the registers of their original owner; the interchanges are only booked (one
// thread keeps rowat/pos in LDS) and applied when the panel is written back. A row is
// eliminated ("done") once it has been a pivot: the candidates of column C are the rows not
// yet done, exactly the rows at positions >= p0 + C in the interchanged order.
template <typename S, int NB, int RPT, int C>
__device__ __forceinline__ void panel_column(S (&P)[RPT][NB], unsigned& done, PanelShared<S, NB>& sh,
                                             BlockedState<S>& st, int b, int p0, int nbw) {
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
    if (!((done >> m) & 1u) && r < n && v > best) {
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
  // no finite candidate (every magnitude NaN): pivot on the row already at p0 + C, which is not
  // yet done; the zero inverse pivot below marks the problem singular in info
  const int p = gkey == 0u ? sh.rowat[C] : bidx;  // original index of the pivot row
#pragma unroll
  for (int m = 0; m < RPT; ++m)
    if (p0 + tid + TH * m == p) {
#pragma unroll
      for (int cc = C; cc < NB; ++cc) sh.prow[buf][cc] = P[m][cc];
      done |= 1u << m;
    }
  if (tid == 0) {
    // the interchange of this column: the pivot row moves to p0 + C, the row there to its place
    const int o = p - p0, pc = sh.pos[o], o2 = sh.rowat[C] - p0;
    sh.piv[C] = p0 + pc;
    sh.rowat[pc] = o2 + p0;
    sh.pos[o2] = pc;
    sh.rowat[C] = p;
    sh.pos[o] = C;
  }
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
    if (!((done >> m) & 1u) && r < n) {
      const S l = cmul(P[m][C], pinv);
      P[m][C] = l;
#pragma unroll
      for (int cc = C + 1; cc < NB; ++cc) P[m][cc] = cfms(P[m][cc], l, sh.prow[buf][cc]);
    }
  }
}

template <typename S, int NB, int RPT, int... Cs>
__device__ __forceinline__ void panel_columns(S (&P)[RPT][NB], unsigned& done, PanelShared<S, NB>& sh,
                                              BlockedState<S>& st, int b, int p0, int nbw,
                                              std::integer_sequence<int, Cs...>) {
  (panel_column<S, NB, RPT, Cs>(P, done, sh, st, b, p0, nbw), ...);
}

template <typename S, int NB, int RPT, int MAXTH>
__global__ void __launch_bounds__(MAXTH) blk_panel_kernel(BlockedState<S> st, int p0, int slot) {
  const int b = blockIdx.x;
  const int n = st.n, ld = st.ld;
  const int TH = blockDim.x;
  const int tid = threadIdx.x;
  S* W = st.W + static_cast<size_t>(b) * n * ld;
  const int nbw = (n - p0 < NB) ? n - p0 : NB;
  __shared__ PanelShared<S, NB> sh;

  // complex64 rows of a full panel move as 16-byte pairs (row starts are 32-byte aligned)
  constexpr bool pairs = std::is_same<S, float2>::value;
  S P[RPT][NB];
#pragma unroll
  for (int m = 0; m < RPT; ++m) {
    const int r = p0 + tid + TH * m;
    const S* src = W + static_cast<size_t>(r) * ld + p0;
    if (r < n) {
      sh.rowat[r - p0] = r;
      sh.pos[r - p0] = r - p0;
    }
    if (pairs && r < n && nbw == NB) {
#pragma unroll
      for (int c = 0; c < NB; c += 2) {
        const float4 x = *reinterpret_cast<const float4*>(src + c);
        P[m][c] = S{x.x, x.y};
        P[m][c + 1] = S{x.z, x.w};
      }
    } else {
#pragma unroll
      for (int c = 0; c < NB; ++c) P[m][c] = (r < n && c < nbw) ? src[c] : czero<S>();
    }
  }
  // (the first column's barrier orders these LDS writes before thread 0's bookkeeping)
  unsigned done = 0;
  panel_columns<S, NB, RPT>(P, done, sh, st, b, p0, nbw, std::make_integer_sequence<int, NB>{});
  __syncthreads();
  // Interchange map of the panel. A row below the panel only ever exchanges its value with a
  // panel row, so after the NB interchanges every row below the panel holds a former panel
  // row's value, and the former values of rows below the panel end in panel rows:
  //   outer[e] = (dst, src): panel row dst receives row src (src below the panel), e < count;
  //   dest[c]:  panel row p0 + c's former value goes to row dest[c] (-1: stays).
  // rowat holds the interchanged order; every panel row and every pivot row below the panel is
  // classified in parallel.
  const int* rv = sh.rowat;
  const int pbase = p0 + nbw;
  if (tid < nbw) {
    sh.mdest[tid] = -1;
    st.piv[static_cast<size_t>(b) * n + p0 + tid] = sh.piv[tid];
  }
  if (tid == 0) sh.ocount = 0;
  __syncthreads();
  if (tid < nbw) {
    const int r = p0 + tid, src = rv[tid];
    if (src != r) {
      if (src < pbase) {
        sh.mdest[src - p0] = r;
      } else {
        const int e = atomicAdd(&sh.ocount, 1);
        sh.mdst[e] = r;
        sh.msrc[e] = src;
      }
    }
    const int p = sh.piv[tid];
    if (p >= pbase) {  // a pivot row below the panel, classified by the first column that chose it
      bool first = true;
      for (int c = 0; c < tid; ++c) first = first && sh.piv[c] != p;
      const int sp = rv[p - p0];
      if (first && sp != p) sh.mdest[sp - p0] = p;
    }
  }
  // rows are written to their interchanged places
#pragma unroll
  for (int m = 0; m < RPT; ++m) {
    const int r = p0 + tid + TH * m;
    if (r >= n) continue;
    S* dst = W + static_cast<size_t>(p0 + sh.pos[r - p0]) * ld + p0;
    if (pairs && nbw == NB) {
#pragma unroll
      for (int c = 0; c < NB; c += 2)
        *reinterpret_cast<float4*>(dst + c) = make_float4(P[m][c].x, P[m][c].y, P[m][c + 1].x, P[m][c + 1].y);
    } else {
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c < nbw) dst[c] = P[m][c];
    }
  }
  __syncthreads();
  int* map = st.rowmap + (static_cast<size_t>(b) * BLK_PANELS + slot) * MAP_STRIDE;
  const int outer = sh.ocount;
  if (tid < NB) map[1 + 2 * NB + tid] = tid < nbw ? sh.mdest[tid] : -1;
  if (tid < outer) {
    map[1 + tid] = sh.mdst[tid];
    map[1 + NB + tid] = sh.msrc[tid];
  }
  if (tid == 0) map[0] = outer;
}

// ---------------------------------------------------------------------------
// Interchanges of panel p0 (its map in `slot`) and/or U12 = L11^-1 A12 for the columns of
// [c_lo, c_hi) and [0, l_hi), the panel's own columns excluded; the triangular solve only
// touches columns right of the panel. grid: xcd_grid(batch, ntile), 64 threads; the first
// ntile_a tiles of 64 columns cover [c_lo, c_hi), the rest [0, l_hi).
constexpr int ROWS_SWAP = 1, ROWS_TRSM = 2;

// Row r of column col of problem b as the trailing update of a fused first block sees it: W
// once the row has been written (flag set by an earlier slot than `slot`), else the source
// matrix. Rows above fuse_e are always in W.
template <typename S>
__device__ __forceinline__ S fused_read(const BlockedState<S>& st, const S* W, int b, int r, int col, int slot,
                                        const LadderSrc& src) {
  if (r >= st.fuse_e) {
    const int f = st.rowflag[static_cast<size_t>(b) * st.n + r];
    if (f == 0 || f > slot) return to_storage<S>(ladder_aug_at(st.src, src, r, col));
  }
  return W[static_cast<size_t>(r) * st.ld + col];
}

template <typename S, int NB>
__global__ void __launch_bounds__(64) blk_rows_kernel(BlockedState<S> st, int p0, int slot, int c_lo, int c_hi,
                                                      int l_hi, int flags, int ntile_a, int ntile, int nprob) {
  int b, tile;
  if (!xcd_tile(blockIdx.x, nprob, ntile, b, tile)) return;
  const int n = st.n, ld = st.ld;
  S* W = st.W + static_cast<size_t>(b) * n * ld;
  const int nbw = (n - p0 < NB) ? n - p0 : NB;
  __shared__ S s_l[NB][NB + 1];
  __shared__ int s_dst[NB], s_src[NB], s_dest[NB];
  const int* map = st.rowmap + (static_cast<size_t>(b) * BLK_PANELS + slot) * MAP_STRIDE;
  const int outer = map[0];
  for (int e = threadIdx.x; e < NB; e += 64) {
    s_dst[e] = e < outer ? map[1 + e] : 0;
    s_src[e] = e < outer ? map[1 + NB + e] : 0;
    s_dest[e] = map[1 + 2 * NB + e];
  }
  if (flags & ROWS_TRSM) {
    for (int e = threadIdx.x; e < nbw * nbw; e += 64) {
      const int r = e / nbw, c = e - (e / nbw) * nbw;
      s_l[r][c] = W[static_cast<size_t>(p0 + r) * ld + p0 + c];
    }
  }
  __syncthreads();
  const int col = tile < ntile_a ? c_lo + tile * 64 + threadIdx.x : (tile - ntile_a) * 64 + threadIdx.x;
  if (col >= (tile < ntile_a ? c_hi : l_hi) || (col >= p0 && col < p0 + nbw)) return;
  if (flags & ROWS_SWAP) {
    // former panel values first (registers), then rows from below the panel into panel rows,
    // then the former panel values to their destinations
    S pv[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c)
      pv[c] = (c < nbw && s_dest[c] >= 0) ? W[static_cast<size_t>(p0 + c) * ld + col] : czero<S>();
    if (st.fuse_e > 0 && p0 < st.fuse_e && col >= st.fuse_e) {
      // trailing columns of a fused first block: rows below it that were not written yet are the
      // source rows, and a row written here is flagged with this panel's slot
      const LadderSrc src = ladder_src(st.src, st.item0 + b);
      for (int e = 0; e < outer; ++e)
        W[static_cast<size_t>(s_dst[e]) * ld + col] = fused_read<S>(st, W, b, s_src[e], col, slot, src);
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c < nbw && s_dest[c] >= 0) {
          const int d = s_dest[c];
          W[static_cast<size_t>(d) * ld + col] = pv[c];
          if (d >= st.fuse_e) {
            int* f = st.rowflag + static_cast<size_t>(b) * n + d;
            if (*f == 0) *f = slot + 1;
          }
        }
    } else {
      for (int e = 0; e < outer; ++e)
        W[static_cast<size_t>(s_dst[e]) * ld + col] = W[static_cast<size_t>(s_src[e]) * ld + col];
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c < nbw && s_dest[c] >= 0) W[static_cast<size_t>(s_dest[c]) * ld + col] = pv[c];
    }
  }
  if (!(flags & ROWS_TRSM) || col < p0) return;
  S x[NB];
#pragma unroll
  for (int r = 0; r < NB; ++r) {
    x[r] = czero<S>();
    if (r < nbw) {
      S acc = W[static_cast<size_t>(p0 + r) * ld + col];
#pragma unroll
      for (int s2 = 0; s2 < r; ++s2) acc = cfms(acc, s_l[r][s2], x[s2]);
      x[r] = acc;
      W[static_cast<size_t>(p0 + r) * ld + col] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// W[r0:r1, c0:c1] -= W[r0:r1, k0:k0+kw] W[k0:k0+kw, c0:c1] for every problem: 64x64 tiles,
// 256 threads, the k range in LDS chunks of KC columns with the C tile held in registers.
struct GemmRange {
  int r0, r1, c0, c1, k0, kw;
  int tiles_r, tiles_c;
  int fused;  // the trailing update of a fused first block: C rows not yet written come from the source
};

inline GemmRange gemm_range(int r0, int r1, int c0, int c1, int k0, int kw, int tm = 64, int fused = 0) {
  return GemmRange{r0, r1, c0, c1, k0, kw, static_cast<int>(ceil_div(r1 - r0, tm)), static_cast<int>(ceil_div(c1 - c0, 64)),
                   fused};
}

// The complex64 vertex of problem b's (chan, w) group for a fused update, or nullptr.
template <typename S>
__device__ __forceinline__ const float2* fused_vertex32(const BlockedState<S>& st, const GemmRange& g, int b) {
  if (st.g32 == nullptr) return nullptr;
  const int grp = (st.item0 + b) / st.src.nq;
  return st.g32 + static_cast<size_t>(grp - st.g32_first) * st.n * st.n;
}

// C element (r, c) of an update: W, or for a fused update (all its rows lie below the fused
// block) the source matrix where the block's interchanges did not write the row.
template <typename S>
__device__ __forceinline__ S gemm_c_at(const BlockedState<S>& st, const S* W, const GemmRange& g, int b, int r, int c,
                                       const LadderSrc& src, const float2* g32) {
  if (st.rowflag[static_cast<size_t>(b) * st.n + r] == 0) {
    if constexpr (std::is_same<S, float2>::value) {
      // off the v = v' diagonal the element is the stored complex64 (gamma - u_r) / beta^2; the
      // assembly wrote the diagonal and the right-hand sides of these rows
      if (g32 != nullptr)
        return (c < st.n && r % st.src.nv != c % st.src.nv) ? g32[static_cast<size_t>(r) * st.n + c]
                                                            : W[static_cast<size_t>(r) * st.ld + c];
    }
    return to_storage<S>(ladder_aug_at(st.src, src, r, c));
  }
  return W[static_cast<size_t>(r) * st.ld + c];
}

// The complex64 vertex of a fused first block: (gamma - u_r) / beta^2 of the (chan, w) groups
// [g0, g0 + ng) in complex64, each element as the assembly forms it.
__global__ void __launch_bounds__(256) blk_vertex32_kernel(LadderArgs a, float2* __restrict__ g32, int g0, int ng) {
  const size_t nn = static_cast<size_t>(a.n) * a.n;
  const size_t total = nn * ng;
  for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < total;
       e += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int gl = static_cast<int>(e / nn);
    const size_t ij = e - static_cast<size_t>(gl) * nn;
    const int grp = g0 + gl, chan = grp / a.nw;
    const int i = static_cast<int>(ij / a.n), j = static_cast<int>(ij - static_cast<size_t>(i) * a.n);
    const int ri = i / a.nv, cj = j / a.nv;
    const double2 x = cscale(csub(a.gamma[static_cast<size_t>(grp) * nn + ij], a.u_r[chan * a.T * a.T + ri * a.T + cj]),
                             a.inv_beta2);
    g32[e] = make_float2(static_cast<float>(x.x), static_cast<float>(x.y));
  }
}

template <typename S, int KC, int TM = 64>
__device__ __forceinline__ void gemm_stage(const S* W, int ld, const GemmRange& g, int rt, int ct, int kc, int kn,
                                           S (&s_a)[TM][KC + 1], S (&s_b)[KC][65]) {
  for (int e = threadIdx.x; e < TM * KC; e += 256) {
    const int r = e / KC, s = e - (e / KC) * KC;
    s_a[r][s] = (rt + r < g.r1 && s < kn) ? W[static_cast<size_t>(rt + r) * ld + g.k0 + kc + s] : czero<S>();
  }
  for (int e = threadIdx.x; e < 64 * KC; e += 256) {
    const int s = e / 64, c = e - (e / 64) * 64;
    s_b[s][c] = (ct + c < g.c1 && s < kn) ? W[static_cast<size_t>(g.k0 + kc + s) * ld + ct + c] : czero<S>();
  }
}

// Thread (ty, tx) of the 16x16 layout owns rows ty + 16 i (i < RI) and columns tx + 16 j
// (j < 4) of a (16 RI) x 64 tile; RI = 8 gives 12 LDS reads per 32 complex FMAs, enough
// to keep the FP64 pipes busy.
template <typename S, int KC, int RI, bool FUSED = false>
__global__ void __launch_bounds__(256) blk_gemm_kernel(BlockedState<S> st, GemmRange g, int nprob) {
  constexpr int TM = 16 * RI;
  int b, t;
  if (!xcd_tile(blockIdx.x, nprob, g.tiles_r * g.tiles_c, b, t)) return;
  const int tr = t / g.tiles_c, tcl = t - (t / g.tiles_c) * g.tiles_c;
  const int ld = st.ld;
  S* W = st.W + static_cast<size_t>(b) * st.n * ld;
  const int rt = g.r0 + tr * TM, ct = g.c0 + tcl * 64;
  __shared__ S s_a[TM][KC + 1];
  __shared__ S s_b[KC][65];
  const int tid = threadIdx.x;
  const int ty = tid / 16, tx = tid % 16;
  // the C tile is loaded first so that its latency overlaps the operand staging
  S acc[RI][4];
  if constexpr (FUSED) {
    const LadderSrc src = ladder_src(st.src, st.item0 + b);
    const float2* g32 = fused_vertex32(st, g, b);
#pragma unroll
    for (int i = 0; i < RI; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = rt + ty + 16 * i, c = ct + tx + 16 * j;
        acc[i][j] = (r < g.r1 && c < g.c1) ? gemm_c_at<S>(st, W, g, b, r, c, src, g32) : czero<S>();
      }
  } else {
#pragma unroll
    for (int i = 0; i < RI; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = rt + ty + 16 * i, c = ct + tx + 16 * j;
        acc[i][j] = (r < g.r1 && c < g.c1) ? W[static_cast<size_t>(r) * ld + c] : czero<S>();
      }
  }
  for (int kc = 0; kc < g.kw; kc += KC) {
    const int kn = min(KC, g.kw - kc);
    if (kc > 0) __syncthreads();
    gemm_stage<S, KC, TM>(W, ld, g, rt, ct, kc, kn, s_a, s_b);
    __syncthreads();
#pragma unroll 2
    for (int s2 = 0; s2 < KC; ++s2) {
      S av[RI], bv[4];
#pragma unroll
      for (int i = 0; i < RI; ++i) av[i] = s_a[ty + 16 * i][s2];
#pragma unroll
      for (int j = 0; j < 4; ++j) bv[j] = s_b[s2][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < RI; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = cfms(acc[i][j], av[i], bv[j]);
    }
  }
#pragma unroll
  for (int i = 0; i < RI; ++i) {
    const int r = rt + ty + 16 * i;
    if (r >= g.r1) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = ct + tx + 16 * j;
      if (c < g.c1) W[static_cast<size_t>(r) * ld + c] = acc[i][j];
    }
  }
}

// Same update on the FP32 matrix cores: each wavefront owns a 32x32 quarter of the tile as
// 2x2 blocks of 16x16, and a complex product becomes two real MFMA chains over the
// interleaved k index (re/im): C_re -= [a_re, -a_im] . [b_re; b_im],
// C_im -= [a_im, a_re] . [b_re; b_im]. Operand layout of the 16x16x4 MFMA (lane l):
// a = A[l%16][l/16], b = B[l/16][l%16], d[e] = D[4*(l/16)+e][l%16]. The accumulators start
// from the C tile and the A operand enters negated.
template <int KC, bool FUSED = false>
__global__ void __launch_bounds__(256) blk_gemm_mfma_kernel(BlockedState<float2> st, GemmRange g, int nprob) {
  typedef float V4 __attribute__((ext_vector_type(4)));
  int b, t;
  if (!xcd_tile(blockIdx.x, nprob, g.tiles_r * g.tiles_c, b, t)) return;
  const int tr = t / g.tiles_c, tcl = t - (t / g.tiles_c) * g.tiles_c;
  const int ld = st.ld;
  float2* W = st.W + static_cast<size_t>(b) * st.n * ld;
  const int rt = g.r0 + tr * 64, ct = g.c0 + tcl * 64;
  __shared__ float2 s_a[64][KC + 1];
  __shared__ float2 s_b[KC][65];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int l16 = lane & 15, kq = lane >> 4, kcl = kq >> 1;
  const bool odd = (kq & 1) != 0;
  const LadderSrc src = FUSED ? ladder_src(st.src, st.item0 + b) : LadderSrc{};
  const float2* g32 = FUSED ? fused_vertex32(st, g, b) : nullptr;
  V4 accr[2][2], acci[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int c = ct + wc + 16 * j + l16;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = rt + wr + 16 * i + 4 * kq + e;
        float2 v = make_float2(0.f, 0.f);
        if (r < g.r1 && c < g.c1) {
          if constexpr (FUSED)
            v = gemm_c_at<float2>(st, W, g, b, r, c, src, g32);
          else
            v = W[static_cast<size_t>(r) * ld + c];
        }
        accr[i][j][e] = v.x;
        acci[i][j][e] = v.y;
      }
    }
  for (int kc = 0; kc < g.kw; kc += KC) {
    const int kn = min(KC, g.kw - kc);
    if (kc > 0) __syncthreads();
    gemm_stage<float2, KC>(W, ld, g, rt, ct, kc, kn, s_a, s_b);
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KC; kk += 2) {
      const int k = kk + kcl;
      float ar[2], ai[2], bv[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const float2 v = s_a[wr + 16 * i + l16][k];
        ar[i] = odd ? v.y : -v.x;
        ai[i] = odd ? -v.x : -v.y;
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const float2 v = s_b[k][wc + 16 * j + l16];
        bv[j] = odd ? v.y : v.x;
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          accr[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[i], bv[j], accr[i][j], 0, 0, 0);
          acci[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[i], bv[j], acci[i][j], 0, 0, 0);
        }
    }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int c = ct + wc + 16 * j + l16;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = rt + wr + 16 * i + 4 * kq + e;
        if (r < g.r1 && c < g.c1) W[static_cast<size_t>(r) * ld + c] = make_float2(accr[i][j][e], acci[i][j][e]);
      }
    }
}

// Split-BF16 form of the complex64 update (three MFMA products per chain): every FP32
// operand x is split as x = hi + lo with hi = bf16(x), lo = bf16(x - hi) (round to
// nearest even), and a b ~ hi_a hi_b + hi_a lo_b + lo_a hi_b with FP32 accumulation
// (relative error ~2^-16 per product; the FP64 refinement of the mixed solve removes it).
// The interleaved real operands are staged in LDS as bf16: A' = [-a_re, a_im] and
// A'' = [-a_im, -a_re] (negated for the subtraction) by rows, B' = [b_re; b_im] by columns.
// Operand layout of the 16x16x32 bf16 MFMA (lane l): a[i] = A[l%16][8*(l/16)+i],
// b[i] = B[8*(l/16)+i][l%16], d[e] = D[4*(l/16)+e][l%16].
struct Bf16Split {
  __bf16 hi, lo;
};

__device__ __forceinline__ Bf16Split bf16_split(float x) {
  const __bf16 h = static_cast<__bf16>(x);
  return Bf16Split{h, static_cast<__bf16>(x - static_cast<float>(h))};
}

// (x0, x1) -> (-x1, x0) in every 32-bit pair of bf16 values
__device__ __forceinline__ __bf16 __attribute__((ext_vector_type(8))) swap_negate_pairs(
    __bf16 __attribute__((ext_vector_type(8))) x) {
  typedef unsigned U4 __attribute__((ext_vector_type(4)));
  U4 u = __builtin_bit_cast(U4, x);
#pragma unroll
  for (int e = 0; e < 4; ++e) u[e] = __builtin_amdgcn_alignbit(u[e], u[e], 16) ^ 0x8000u;
  return __builtin_bit_cast(__bf16 __attribute__((ext_vector_type(8))), u);
}

template <int KC, bool FUSED = false>
__global__ void __launch_bounds__(256) blk_gemm_bf16x3_kernel(BlockedState<float2> st, GemmRange g, int nprob) {
  typedef float V4 __attribute__((ext_vector_type(4)));
  typedef __bf16 B8 __attribute__((ext_vector_type(8)));
  constexpr int KR = 2 * KC + 8;  // bf16 per staged row (interleaved re/im, padded)
  int b, t;
  if (!xcd_tile(blockIdx.x, nprob, g.tiles_r * g.tiles_c, b, t)) return;
  const int tr = t / g.tiles_c, tcl = t - (t / g.tiles_c) * g.tiles_c;
  const int ld = st.ld;
  float2* W = st.W + static_cast<size_t>(b) * st.n * ld;
  const int rt = g.r0 + tr * 64, ct = g.c0 + tcl * 64;
  // two LDS buffers: the next chunk is split and staged while the current one feeds the
  // matrix cores, one barrier per chunk
  __shared__ __attribute__((aligned(16))) __bf16 s_a[2][2][64][KR];  // [buffer] A' hi, A' lo
  __shared__ __attribute__((aligned(16))) __bf16 s_b[2][2][64][KR];  // [buffer] B' hi, B' lo, [column][k]
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int l16 = lane & 15, kg = lane >> 4;
  // tiles that lie wholly inside the update range take unpredicated 16-byte loads; 32-bit
  // element offsets (n * ld < 2^31) keep the address arithmetic in single registers
  const bool full = rt + 64 <= g.r1 && ct + 64 <= g.c1 && ((g.c0 | g.k0 | ld) & 1) == 0;
  // the C tile is held apart from the accumulators, so its loads overlap staging and compute
  // (reading it after the products instead frees 32 registers for a third wave per SIMD, but
  // measured 7% slower)
  float2 cin[2][2][4];
  if constexpr (FUSED) {
    // rows not yet written come from the complex64 vertex (off the v = v' diagonal) or the source;
    // each row's flag and frequency, and each column's, are looked up once
    const LadderSrc src = ladder_src(st.src, st.item0 + b);
    const float2* g32 = fused_vertex32(st, g, b);
    const int n = st.n, nv = st.src.nv;
    int rflag[2][4], rv[2][4], cv[2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = rt + wr + 16 * i + 4 * kg + e;
        rflag[i][e] = r < g.r1 ? st.rowflag[static_cast<size_t>(b) * n + r] : 1;
        rv[i][e] = r % nv;
      }
#pragma unroll
    for (int j = 0; j < 2; ++j) cv[j] = (ct + wc + 16 * j + l16) % nv;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int c = ct + wc + 16 * j + l16;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int r = rt + wr + 16 * i + 4 * kg + e;
          float2 x = make_float2(0.f, 0.f);
          if (r < g.r1 && c < g.c1) {
            if (g32 == nullptr)
              x = rflag[i][e] != 0 ? W[static_cast<unsigned>(r * ld + c)]
                                   : to_storage<float2>(ladder_aug_at(st.src, src, r, c));
            else  // the assembly wrote the diagonal and the right-hand sides of these rows
              x = (rflag[i][e] != 0 || c >= n || rv[i][e] == cv[j]) ? W[static_cast<unsigned>(r * ld + c)]
                                                                     : g32[static_cast<unsigned>(r * n + c)];
          }
          cin[i][j][e] = x;
        }
      }
  } else {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int c = ct + wc + 16 * j + l16;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int r = rt + wr + 16 * i + 4 * kg + e;
          cin[i][j][e] = (full || (r < g.r1 && c < g.c1)) ? W[static_cast<unsigned>(r * ld + c)] : make_float2(0.f, 0.f);
        }
      }
  }
  V4 accr[2][2], acci[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      accr[i][j] = V4{0.f, 0.f, 0.f, 0.f};
      acci[i][j] = V4{0.f, 0.f, 0.f, 0.f};
    }
  // operand chunks of KC columns of k: raw values are loaded to registers (the next chunk's
  // loads are issued before the current chunk's products), then split and packed into LDS.
  // A: thread (row tid/16 + 16p, k pair 2 (tid % 16)); B: thread (k pair tid/32 + 8q, column
  // pair 2 (tid % 32)), each a 16-byte load per row.
  static_assert(KC == 32, "the chunk mapping assumes 32 complex columns of k");
  float4 ra[4], rb[2][2];
  const int a_r = tid >> 4, a_s = 2 * (tid & 15);
  const int b_p = tid >> 5, b_c = 2 * (tid & 31);
  auto load_chunk = [&](int kc) {
    const int kn = min(KC, g.kw - kc);
    if (full && kn == KC) {
#pragma unroll
      for (int p = 0; p < 4; ++p)
        ra[p] = *reinterpret_cast<const float4*>(W + static_cast<unsigned>((rt + a_r + 16 * p) * ld + g.k0 + kc + a_s));
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int h = 0; h < 2; ++h)
          rb[q][h] = *reinterpret_cast<const float4*>(
              W + static_cast<unsigned>((g.k0 + kc + 2 * (b_p + 8 * q) + h) * ld + ct + b_c));
    } else {
      auto ld1 = [&](bool in, int r, int c) { return in ? W[static_cast<unsigned>(r * ld + c)] : make_float2(0.f, 0.f); };
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        const int r = rt + a_r + 16 * p;
        const float2 x0 = ld1(r < g.r1 && a_s < kn, r, g.k0 + kc + a_s);
        const float2 x1 = ld1(r < g.r1 && a_s + 1 < kn, r, g.k0 + kc + a_s + 1);
        ra[p] = make_float4(x0.x, x0.y, x1.x, x1.y);
      }
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int k = 2 * (b_p + 8 * q) + h, c = ct + b_c;
          const float2 x0 = ld1(k < kn && c < g.c1, g.k0 + kc + k, c);
          const float2 x1 = ld1(k < kn && c + 1 < g.c1, g.k0 + kc + k, c + 1);
          rb[q][h] = make_float4(x0.x, x0.y, x1.x, x1.y);
        }
    }
  };
  typedef __bf16 B4 __attribute__((ext_vector_type(4)));
  auto stage = [&](int buf) {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int r = a_r + 16 * p;
      const Bf16Split r0 = bf16_split(ra[p].x), i0 = bf16_split(ra[p].y), r1 = bf16_split(ra[p].z),
                      i1 = bf16_split(ra[p].w);
      *reinterpret_cast<B4*>(&s_a[buf][0][r][2 * a_s]) = B4{-r0.hi, i0.hi, -r1.hi, i1.hi};
      *reinterpret_cast<B4*>(&s_a[buf][1][r][2 * a_s]) = B4{-r0.lo, i0.lo, -r1.lo, i1.lo};
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int k2 = 2 * (2 * (b_p + 8 * q));
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const float br0 = h ? rb[q][0].z : rb[q][0].x, bi0 = h ? rb[q][0].w : rb[q][0].y;
        const float br1 = h ? rb[q][1].z : rb[q][1].x, bi1 = h ? rb[q][1].w : rb[q][1].y;
        const Bf16Split b0 = bf16_split(br0), j0 = bf16_split(bi0), b1 = bf16_split(br1), j1 = bf16_split(bi1);
        *reinterpret_cast<B4*>(&s_b[buf][0][b_c + h][k2]) = B4{b0.hi, j0.hi, b1.hi, j1.hi};
        *reinterpret_cast<B4*>(&s_b[buf][1][b_c + h][k2]) = B4{b0.lo, j0.lo, b1.lo, j1.lo};
      }
    }
  };
  auto products = [&](int buf) {
#pragma unroll
    for (int ks = 0; ks < 2 * KC; ks += 32) {
      const int k = ks + 8 * kg;
      B8 ah[2], al[2], a2h[2], a2l[2], bh[2], bl[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int row = wr + 16 * i + l16;
        ah[i] = *reinterpret_cast<const B8*>(&s_a[buf][0][row][k]);
        al[i] = *reinterpret_cast<const B8*>(&s_a[buf][1][row][k]);
        // A'' = (-im, -re) pairs from A' = (-re, im): swap the halves of each word, negate the low one
        a2h[i] = swap_negate_pairs(ah[i]);
        a2l[i] = swap_negate_pairs(al[i]);
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int col = wc + 16 * j + l16;
        bh[j] = *reinterpret_cast<const B8*>(&s_b[buf][0][col][k]);
        bl[j] = *reinterpret_cast<const B8*>(&s_b[buf][1][col][k]);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          accr[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], accr[i][j], 0, 0, 0);
          accr[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], accr[i][j], 0, 0, 0);
          accr[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], accr[i][j], 0, 0, 0);
          acci[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2h[i], bh[j], acci[i][j], 0, 0, 0);
          acci[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2h[i], bl[j], acci[i][j], 0, 0, 0);
          acci[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2l[i], bh[j], acci[i][j], 0, 0, 0);
        }
    }
  };
  load_chunk(0);
  stage(0);
  __syncthreads();
  int cur = 0;
  for (int kc = 0; kc < g.kw; kc += KC, cur ^= 1) {
    const bool more = kc + KC < g.kw;
    if (more) load_chunk(kc + KC);
    products(cur);
    if (more) stage(cur ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int c = ct + wc + 16 * j + l16;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = rt + wr + 16 * i + 4 * kg + e;
        if (full || (r < g.r1 && c < g.c1))
          W[static_cast<unsigned>(r * ld + c)] = make_float2(cin[i][j][e].x + accr[i][j][e], cin[i][j][e].y + acci[i][j][e]);
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
// The first outer block takes over the assembly of its trailing rows unless DGAMI_LADDER_FUSE=0.
static bool ladder_fuse_enabled() {
  const char* v = getenv("DGAMI_LADDER_FUSE");
  return !(v && std::string(v) == "0");
}

static bool use_vector_gemm() {
  const char* v = getenv("DGAMI_BLOCKED_GEMM");
  return v && std::string(v) == "vector";
}

template <typename S, int NB, int RPT, int MAXTH>
struct BlockedSteps {
  const LadderArgs& a;
  BlockedState<S>& st;
  int nb;
  hipStream_t stream;

  int panel(int p0, int slot) {
    const int rows = st.n - p0;
    const int th = ((rows + RPT - 1) / RPT + 63) / 64 * 64;
    DGAMI_ARG_CHECK(th <= MAXTH, "ladder: compound size too large for the panel kernel");
    hipLaunchKernelGGL((blk_panel_kernel<S, NB, RPT, MAXTH>), dim3(nb), dim3(th), 0, stream, st, p0, slot);
    DGAMI_HIP_CHECK(hipGetLastError());
    return 0;
  }
  // interchanges / triangular solve of panel p0 on [c_lo, c_hi) and, with swap_left, [0, l_hi)
  int rows(int p0, int slot, int c_lo, int c_hi, int l_hi, int flags) {
    if (!st.swap_left) l_hi = 0;
    const int ta = c_hi > c_lo ? static_cast<int>(ceil_div(c_hi - c_lo, 64)) : 0;
    const int ntile = ta + (l_hi > 0 ? static_cast<int>(ceil_div(l_hi, 64)) : 0);
    if (ntile == 0) return 0;
    hipLaunchKernelGGL((blk_rows_kernel<S, NB>), dim3(static_cast<unsigned>(xcd_grid(nb, ntile))), dim3(64), 0, stream,
                       st, p0, slot, c_lo, c_hi, l_hi, flags, ta, ntile, nb);
    DGAMI_HIP_CHECK(hipGetLastError());
    return 0;
  }
  int gemm(int r0, int r1, int c0, int c1, int k0, int kw, int fused = 0) {
    if (r1 <= r0 || c1 <= c0 || kw <= 0) return 0;
    int mode = BLK_GEMM_VECTOR;
    if constexpr (std::is_same<S, float2>::value) mode = use_vector_gemm() ? BLK_GEMM_VECTOR : st.gemm_mode;
    // update tiles are 64 x 64; complex128 variants with tall_tiles use 128 x 64 (8 rows per
    // thread) for updates of more than 96 rows. Measured at n = 720 (5642 problems) the 64-row
    // tiles are 8% faster over the whole factorisation (more workgroups per CU), and at n = 120
    // a 128-row tile over the short updates is mostly empty.
    constexpr bool c128 = std::is_same<S, double2>::value;
    const bool tall = c128 && mode == BLK_GEMM_VECTOR && r1 - r0 > 96 && st.tall_tiles;
    const GemmRange g = gemm_range(r0, r1, c0, c1, k0, kw, tall ? 128 : 64, fused);
    const size_t grid = xcd_grid(nb, static_cast<size_t>(g.tiles_r) * g.tiles_c);
    DGAMI_ARG_CHECK(grid < (size_t(1) << 31), "ladder: trailing update grid too large");
    hipEvent_t start = nullptr;
    const bool prof = profile_on();
    if (prof && profile_begin(stream, &start)) return -2;
    const dim3 gd(static_cast<unsigned>(grid));
    if (mode == BLK_GEMM_VECTOR) {
      if constexpr (c128) {
        if (tall) {
          if (fused)
            hipLaunchKernelGGL((blk_gemm_kernel<S, 16, 8, true>), gd, dim3(256), 0, stream, st, g, nb);
          else
            hipLaunchKernelGGL((blk_gemm_kernel<S, 16, 8>), gd, dim3(256), 0, stream, st, g, nb);
        } else {
          if (fused)
            hipLaunchKernelGGL((blk_gemm_kernel<S, 16, 4, true>), gd, dim3(256), 0, stream, st, g, nb);
          else
            hipLaunchKernelGGL((blk_gemm_kernel<S, 16, 4>), gd, dim3(256), 0, stream, st, g, nb);
        }
      } else {
        if (fused)
          hipLaunchKernelGGL((blk_gemm_kernel<S, 32, 4, true>), gd, dim3(256), 0, stream, st, g, nb);
        else
          hipLaunchKernelGGL((blk_gemm_kernel<S, 32, 4>), gd, dim3(256), 0, stream, st, g, nb);
      }
    } else if constexpr (std::is_same<S, float2>::value) {
      if (mode == BLK_GEMM_BF16X3) {
        if (fused)
          hipLaunchKernelGGL((blk_gemm_bf16x3_kernel<32, true>), gd, dim3(256), 0, stream, st, g, nb);
        else
          hipLaunchKernelGGL((blk_gemm_bf16x3_kernel<32>), gd, dim3(256), 0, stream, st, g, nb);
      } else {
        if (fused)
          hipLaunchKernelGGL((blk_gemm_mfma_kernel<32, true>), gd, dim3(256), 0, stream, st, g, nb);
        else
          hipLaunchKernelGGL((blk_gemm_mfma_kernel<32>), gd, dim3(256), 0, stream, st, g, nb);
      }
    }
    DGAMI_HIP_CHECK(hipGetLastError());
    if (prof) {
      const double m = r1 - r0, c = c1 - c0, es = sizeof(S);
      const double bytes = static_cast<double>(nb) * es * (2.0 * m * c + m * kw + static_cast<double>(kw) * c);
      if (profile_end(stream, start, bytes, static_cast<double>(nb) * 8.0 * m * c * kw)) return -2;
    }
    return 0;
  }
};

// Right-looking LU in outer blocks of NPAN * NB columns, each factored as NPAN register
// panels of NB columns, so that the trailing matrix is read and written once per
// outer block. Inside the block the panels are brought up to date left-looking: before panel
// j, for every earlier panel i, its interchanges, the update of its rows by the panels before
// it and its U block are applied to panel j's columns, then panel j's rows and the rows below
// are updated by all earlier panels at once; after panel j its interchanges are applied to the
// earlier panels' columns. The trailing columns then receive, panel by panel, the
// interchanges, the row-block update and the U blocks, and the rows below the block one
// trailing update of depth NPAN * NB. Complex128 uses two panels per block (its vector-FMA
// trailing update is compute-bound, and the extra in-block updates of deeper blocks cost more
// than they save); complex64 uses four, halving the trailing matrix's HBM traffic.
template <typename S, int NB, int RPT, int MAXTH, int NPAN>
static int blocked_factor_nb(const LadderArgs& a, BlockedState<S>& st, int item0, int nb, hipStream_t stream) {
  static_assert(NPAN >= 1 && NPAN <= BLK_PANELS, "panels per outer block");
  const int n = st.n, ncol = st.ncol;
  DGAMI_ARG_CHECK(nb <= 65535, "ladder: batch too large");
  // the first outer block takes the assembly of its trailing rows (when there are trailing columns)
  // complex64 fuses only with a shared vertex and no V_r, whose elements it reads from a complex64
  // copy (the complex128 vertex would double the bytes of the C tile it replaces)
  constexpr bool c64 = std::is_same<S, float2>::value;
  const bool fuse = ladder_fuse_enabled() && n > NPAN * NB && (!c64 || (!a.dense && a.v_r == nullptr));
  st.fuse_e = fuse ? NPAN * NB : 0;
  st.item0 = item0;
  st.src = a;
  st.g32 = nullptr;
  if (fuse && c64) {
    const int g0 = item0 / a.nq, g1 = (item0 + nb - 1) / a.nq;
    float2* g32 = static_cast<float2*>(
        workspace(st.ws_g32, static_cast<size_t>(g1 - g0 + 1) * n * n * sizeof(float2)));
    if (!g32) return -2;
    hipLaunchKernelGGL(blk_vertex32_kernel, dim3(8192), dim3(256), 0, stream, a, g32, g0, g1 - g0 + 1);
    DGAMI_HIP_CHECK(hipGetLastError());
    st.g32 = g32;
    st.g32_first = g0;
  }
  // the matrices are written once and read later from HBM: streaming (non-temporal) stores
  hipLaunchKernelGGL((blk_assemble_kernel<S, true>), dim3(ceil_div(n, ASM_ROWS), ceil_div(nb, ASM_QG)), dim3(256), 0,
                     stream, a, st, item0, nb);
  DGAMI_HIP_CHECK(hipGetLastError());
  BlockedSteps<S, NB, RPT, MAXTH> k{a, st, nb, stream};
  const int BOTH = ROWS_SWAP | ROWS_TRSM;
  int rc = 0;
#define DGAMI_STEP(x) \
  if ((rc = (x)) != 0) return rc
  for (int P = 0; P < n; P += NPAN * NB) {
    const int np = min(NPAN, static_cast<int>(ceil_div(n - P, NB)));
    const int E = min(n, P + NPAN * NB);  // end of the outer block
    for (int j = 0; j < np; ++j) {
      const int Pj = P + j * NB, Ej = min(n, Pj + NB);
      if (j > 0) {
        for (int i = 0; i < j; ++i) {
          const int Pi = P + i * NB;
          if (i == 0) {
            DGAMI_STEP(k.rows(Pi, i, Pj, Ej, 0, BOTH));
          } else {
            DGAMI_STEP(k.rows(Pi, i, Pj, Ej, 0, ROWS_SWAP));
            DGAMI_STEP(k.gemm(Pi, Pi + NB, Pj, Ej, P, i * NB));
            DGAMI_STEP(k.rows(Pi, i, Pj, Ej, 0, ROWS_TRSM));
          }
        }
        DGAMI_STEP(k.gemm(Pj, n, Pj, Ej, P, j * NB));
      }
      DGAMI_STEP(k.panel(Pj, j));
      if (j > 0) DGAMI_STEP(k.rows(Pj, j, P, Pj, 0, ROWS_SWAP));
    }
    for (int j = 0; j < np; ++j) {
      const int Pj = P + j * NB, Ej = min(n, Pj + NB);
      if (j == 0) {
        DGAMI_STEP(k.rows(Pj, j, E, ncol, P, BOTH));
      } else {
        DGAMI_STEP(k.rows(Pj, j, E, ncol, P, ROWS_SWAP));
        DGAMI_STEP(k.gemm(Pj, Ej, E, ncol, P, j * NB));
        DGAMI_STEP(k.rows(Pj, j, E, ncol, 0, ROWS_TRSM));
      }
    }
    DGAMI_STEP(k.gemm(E, n, E, ncol, P, E - P, P == 0 && st.fuse_e > 0));
  }
#undef DGAMI_STEP
  return 0;
}

template <>
int blocked_factor<double2>(const LadderArgs& a, BlockedState<double2>& st, int item0, int nb, hipStream_t stream) {
  // by size: 32-column panels of one row per thread up to n = 768; above, 16-column panels
  // of three rows per thread in four-panel outer blocks (7% faster than two panels at n = 1080)
  const int v = blocked_variant_fits(false, st.variant, st.n) ? st.variant : (st.n <= 768 ? 0 : 4);
  st.tall_tiles = v == 2 || v == 3;
  if (v == 0 || v == 2) return blocked_factor_nb<double2, 32, 1, 768, 2>(a, st, item0, nb, stream);
  if (v == 4) return blocked_factor_nb<double2, 16, 3, 512, 4>(a, st, item0, nb, stream);
  return blocked_factor_nb<double2, 16, 3, 512, 2>(a, st, item0, nb, stream);
}

template <>
int blocked_factor<float2>(const LadderArgs& a, BlockedState<float2>& st, int item0, int nb, hipStream_t stream) {
  // three rows per thread: 256-thread panels for n <= 768, two problems per CU
  const int v = blocked_variant_fits(true, st.variant, st.n) ? st.variant : (st.n <= 1152 ? 0 : 1);
  switch (v) {
    case 0: return blocked_factor_nb<float2, 32, 3, 384, 4>(a, st, item0, nb, stream);
    case 2: return blocked_factor_nb<float2, 32, 3, 384, 8>(a, st, item0, nb, stream);
    case 3: return blocked_factor_nb<float2, 16, 3, 512, 8>(a, st, item0, nb, stream);
    case 4: return blocked_factor_nb<float2, 32, 2, 512, 4>(a, st, item0, nb, stream);
    case 5: return blocked_factor_nb<float2, 32, 3, 384, 6>(a, st, item0, nb, stream);
    default: return blocked_factor_nb<float2, 16, 3, 512, 4>(a, st, item0, nb, stream);
  }
}

int blocked_gemm_fp64(BlockedState<double2>& st, int nb, int r0, int r1, int c0, int c1, int k0, int kw,
                      hipStream_t stream) {
  const LadderArgs unused{};
  BlockedSteps<double2, 32, 1, 768> k{unused, st, nb, stream};
  return k.gemm(r0, r1, c0, c1, k0, kw);
}

int blocked_gemm_c64(BlockedState<float2>& st, int nb, int r0, int r1, int c0, int c1, int k0, int kw,
                     hipStream_t stream) {
  const LadderArgs unused{};
  BlockedSteps<float2, 32, 3, 384> k{unused, st, nb, stream};
  return k.gemm(r0, r1, c0, c1, k0, kw);
}

int ladder_lanes() {
  const char* v = getenv("DGAMI_LADDER_STREAMS");
  const int s = v ? atoi(v) : 0;
  return s >= 1 && s <= 1 + MAX_EXTRA_STREAMS ? s : 2;
}

int ladder_blocked_fp64(const LadderArgs& a, int variant, hipStream_t stream) {
  const int n = a.n, ncol = n + a.T;
  // rows start on 128-byte boundaries (8 complex128 values)
  const int ld = (ncol + 7) & ~7;
  DGAMI_ARG_CHECK(n <= BLK_MAX_N, "ladder: compound size above 1536 is not supported");
  const size_t items = static_cast<size_t>(a.nchan) * a.nw * a.nq;
  const size_t per = static_cast<size_t>(n) * ld * sizeof(double2);
  const int nl = static_cast<int>(std::min<size_t>(ladder_lanes(), items));
  // MI355X holds 288 GB of HBM: factor up to 48 GiB of problems per pass (over all lanes)
  size_t batch = (size_t(48) << 30) / (per * nl);
  if (batch < 1) batch = 1;
  if (batch > (items + nl - 1) / nl) batch = (items + nl - 1) / nl;
  if (batch > 65535) batch = 65535;
  const int nblk = static_cast<int>(ceil_div(n, TRI_B));
  struct Lane {
    BlockedState<double2> st;
    double2* tri;
    hipStream_t s;
  };
  Lane lanes[1 + MAX_EXTRA_STREAMS];
  for (int l = 0; l < nl; ++l) {
    Lane& L = lanes[l];
    if (!blocked_alloc(L.st, n, ld, ncol, batch, WS_LADDER + l * WS_STREAM_STRIDE)) return -2;
    L.st.variant = variant;
    L.tri = static_cast<double2*>(
        workspace(WS_TRI + l * WS_STREAM_STRIDE, batch * nblk * 2 * TRI_B * TRI_B * sizeof(double2)));
    if (!L.tri) return -2;
    L.s = stream;
    if (l > 0) {
      const int rc = extra_stream_fork(stream, l, &L.s);
      if (rc) return rc;
    }
  }
  int k = 0;
  for (size_t item0 = 0; item0 < items; item0 += batch, k = (k + 1) % nl) {
    Lane& L = lanes[k];
    const int nb = static_cast<int>(items - item0 < batch ? items - item0 : batch);
    int rc = blocked_factor<double2>(a, L.st, static_cast<int>(item0), nb, L.s);
    if (rc) return rc;
    // back substitution on the carried right-hand sides, x / beta straight to the output
    TriFactors<double2> f{L.st.W, L.tri, L.st.dinv, n, ld, nblk};
    rc = tri_diag_inverse(f, nb, false, L.s);
    if (rc) return rc;
    TriVec<double2> y{L.st.W + n, static_cast<size_t>(n) * ld, ld};
    rc = tri_backward(f, nb, a.T, y, TriVec<double2>{nullptr, 0, 0}, &a, static_cast<int>(item0), L.s);
    if (rc) return rc;
    rc = blocked_copy_info(L.st.info, a.info, static_cast<int>(item0), nb, L.s);
    if (rc) return rc;
  }
  for (int l = 1; l < nl; ++l) {
    const int rc = extra_stream_join(stream, l);
    if (rc) return rc;
  }
  return 0;
}

}  // namespace dgami
