// Full inverse of the ladder matrices, the full vertex built from it, and the
// particle-particle pairing vertices of the Eliashberg equation (complex128).
//
//   X(q,w)   = A^-1,  A = chi0^-1 delta_vv' + (Gamma_r - U_r - V_r^q) / beta^2
//   F_r(q,w) = beta^2 (chi0^-1 - chi0^-1 X chi0^-1) + (u - u chi_r u) vrg(v) vrg(v'),  u = U_r + V_r^q
//
// Every (q, w) matrix is factored by the batched blocked LU of the ladder path, which carries
// n right-hand-side columns: the identity for X, the block-diagonal chi0^-1 for F (so that
// the columns end as X chi0^-1 and the two GEMMs of chi0^-1 X chi0^-1 reduce to one
// nb^2-wide product per row). The back substitution runs in steps of 64 rows from the bottom:
// the two inverse 32x32 diagonal blocks of U are applied in LDS, then the rows above are
// updated by the trailing-update GEMM of the factorisation.
#include "dgami.h"
#include "dgami_internal.hpp"
#include "dgami_launch.hpp"
#include "ladder_blocked.hpp"
#include "tri_solve.hpp"

#include <algorithm>
#include <type_traits>

namespace dgami {

constexpr int VTX_STEP = 2 * TRI_B;  // rows per back-substitution step

// s_y[h*TRI_B + r][lane] = (D_h s_y_h)[r][lane] for the 32 rows of half h; D_h = inverse of the
// diagonal block of U (row-major TRI_B x TRI_B). Thread (grp, lane) forms rows grp*8 .. grp*8+7.
template <typename S>
__device__ __forceinline__ void vtx_apply_inverse(S (&s_y)[VTX_STEP][65], S (&s_m)[TRI_B][TRI_B + 1],
                                                  const S* __restrict__ D, int h, int grp, int lane) {
  __syncthreads();
  for (int e = threadIdx.x; e < TRI_B * TRI_B; e += 256) s_m[e / TRI_B][e % TRI_B] = D[e];
  __syncthreads();
  S acc[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) acc[i] = czero<S>();
#pragma unroll 4
  for (int k = 0; k < TRI_B; ++k) {
    const S y = s_y[h * TRI_B + k][lane];
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[i] = cfma(acc[i], s_m[grp * 8 + i][k], y);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 8; ++i) s_y[h * TRI_B + grp * 8 + i][lane] = acc[i];
}

// One back-substitution step on rows [c0, c0 + cb) (c0 a multiple of 64, cb <= 64) of the
// right-hand-side columns [n, n + nr): rows below c0 + cb are final and already eliminated
// from these rows, so with the halves 0 = [c0, c0+32) and 1 = [c0+32, c0+cb)
//   X_1 = U_11^-1 Y_1,  Y_0 -= U_01 X_1,  X_0 = U_00^-1 Y_0.
// grid: xcd_grid(nprob, ntile) workgroups of 256 threads, one 64-column tile each.
template <typename S>
__global__ void __launch_bounds__(256) vtx_backsolve_kernel(BlockedState<S> st, const S* __restrict__ dinvblk,
                                                            int nblk, int c0, int cb, int nr, int ntile, int nprob) {
  int b, tile;
  if (!xcd_tile(blockIdx.x, nprob, ntile, b, tile)) return;
  const int n = st.n, ld = st.ld;
  S* W = st.W + static_cast<size_t>(b) * n * ld;
  const int cc = n + tile * 64;
  __shared__ S s_y[VTX_STEP][65];
  __shared__ S s_m[TRI_B][TRI_B + 1];
  const int lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
  const bool colok = tile * 64 + lane < nr;
  for (int r = grp; r < VTX_STEP; r += 4)
    s_y[r][lane] = (r < cb && colok) ? W[static_cast<size_t>(c0 + r) * ld + cc + lane] : czero<S>();
  // inverse diagonal blocks of U: [nblk][2][TRI_B][TRI_B], U at offset TRI_B*TRI_B
  const S* Dinv = dinvblk + static_cast<size_t>(b) * nblk * 2 * TRI_B * TRI_B + TRI_B * TRI_B;
  const int kb0 = c0 / TRI_B;
  if (cb > TRI_B) {
    vtx_apply_inverse(s_y, s_m, Dinv + static_cast<size_t>(kb0 + 1) * 2 * TRI_B * TRI_B, 1, grp, lane);
    __syncthreads();
    const int kw = cb - TRI_B;
    for (int e = threadIdx.x; e < TRI_B * TRI_B; e += 256) {
      const int r = e / TRI_B, k = e % TRI_B;
      s_m[r][k] = k < kw ? W[static_cast<size_t>(c0 + r) * ld + c0 + TRI_B + k] : czero<S>();
    }
    __syncthreads();
    S acc[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[i] = s_y[grp * 8 + i][lane];
#pragma unroll 4
    for (int k = 0; k < TRI_B; ++k) {
      const S x = s_y[TRI_B + k][lane];
#pragma unroll
      for (int i = 0; i < 8; ++i) acc[i] = cfms(acc[i], s_m[grp * 8 + i][k], x);
    }
    // each thread rewrites only the elements it read
#pragma unroll
    for (int i = 0; i < 8; ++i) s_y[grp * 8 + i][lane] = acc[i];
  }
  vtx_apply_inverse(s_y, s_m, Dinv + static_cast<size_t>(kb0) * 2 * TRI_B * TRI_B, 0, grp, lane);
  __syncthreads();
  if (colok)
    for (int r = grp; r < cb; r += 4) W[static_cast<size_t>(c0 + r) * ld + cc + lane] = s_y[r][lane];
}

// out[item][i][j] = X[i][j] (the right-hand-side columns after the back substitution).
// grid (ceil(n/4), nprob), 256 threads.
template <typename S>
__global__ void __launch_bounds__(256) vtx_copy_kernel(BlockedState<S> st, double2* __restrict__ out, int item0) {
  const int b = blockIdx.y, n = st.n;
  const S* W = st.W + static_cast<size_t>(b) * n * st.ld + n;
  double2* O = out + static_cast<size_t>(item0 + b) * n * n;
  const int r1 = min(n, (blockIdx.x + 1) * 4);
  for (int i = blockIdx.x * 4; i < r1; ++i)
    for (int j = threadIdx.x; j < n; j += 256)
      O[static_cast<size_t>(i) * n + j] = to_c128(W[static_cast<size_t>(i) * st.ld + j]);
}

struct VertexArgs {
  const double2* vrg;       // QF [nw][T*T][nv][nq] or nullptr (first term only)
  const double2* chi_phys;  // QW [nw][T*T][nq]
  double2* out;             // [nw][nq][n][n]
  double beta2;
  int z_form;               // 1: the right-hand sides hold Z = X (A - chi0^-1), 0: Y = X chi0^-1
  int vlo, vw;              // vw > 0: only rows and columns with v in [vlo, vlo + vw) are formed
};

// Full vertex of one problem, rows (ri, v) for every ri < T at one v, every column j = (cj, v')
// (with a window vw > 0: v = vlo + blockIdx.x and the columns with v' in the window, whose
// right-hand sides are r = cj vw + v' - vlo):
//   F = beta^2 (D(v) delta_vv' - sum_s D(v)[ri][s] Y[(s, v)][j]) + sum_s P(v)[ri][s] V(v')[s][cj]
// with Y = X chi0^-1 (right-hand-side columns), D(v) = chi0^-1(q,w,v), V(v) = vrg(q,w,v) and
// P(v) = (u - u chi_r u) V(v), all nb^2 x nb^2 in M order. With z_form the columns hold
// Z = X (A - chi0^-1) and the first term is beta^2 sum_s D(v)[ri][s] Z[(s, v)][j], the same
// value (chi0^-1 - chi0^-1 X chi0^-1 = chi0^-1 X (A - chi0^-1)) without the cancellation of two
// terms that grow as v^2. grid (nv, nprob), 256 threads.
template <int T, typename S>
__global__ void __launch_bounds__(256) vtx_full_kernel(LadderArgs a, BlockedState<S> st, VertexArgs o, int item0) {
  constexpr int TT = T * T;
  const int v = o.vlo + blockIdx.x, b = blockIdx.y;
  const int item = item0 + b;
  const int nq = a.nq, nv = a.nv, n = a.n;
  const int q = item % nq, w = item / nq;
  __shared__ double2 s_d[T][T + 1], s_u[T][T + 1], s_t[T][T + 1], s_p[T][T + 1];
  const int tid = threadIdx.x;
  const bool f2 = o.vrg != nullptr;
  for (int e = tid; e < TT; e += 256) {
    const int r = e / T, c = e % T;
    s_d[r][c] = a.chi0_inv[((static_cast<size_t>(w) * TT + e) * nv + v) * nq + q];
    if (f2) {
      double2 u = a.u_r[e];
      if (a.v_r) u = cadd(u, a.v_r[static_cast<size_t>(q) * TT + e]);
      s_u[r][c] = u;
    }
  }
  __syncthreads();
  if (f2) {
    for (int e = tid; e < TT; e += 256) {  // s_t = chi_r u
      const int r = e / T, c = e % T;
      double2 acc = make_double2(0.0, 0.0);
      for (int k = 0; k < T; ++k)
        acc = cfma(acc, o.chi_phys[(static_cast<size_t>(w) * TT + r * T + k) * nq + q], s_u[k][c]);
      s_t[r][c] = acc;
    }
    __syncthreads();
    for (int e = tid; e < TT; e += 256) {  // s_p = u - u chi_r u
      const int r = e / T, c = e % T;
      double2 acc = s_u[r][c];
      for (int k = 0; k < T; ++k) acc = cfms(acc, s_u[r][k], s_t[k][c]);
      s_p[r][c] = acc;
    }
    __syncthreads();
    for (int e = tid; e < TT; e += 256)  // s_t = V(v)
      s_t[e / T][e % T] = o.vrg[((static_cast<size_t>(w) * TT + e) * nv + v) * nq + q];
    __syncthreads();
    for (int e = tid; e < TT; e += 256) {  // s_u = P(v) = (u - u chi_r u) V(v)
      const int r = e / T, c = e % T;
      double2 acc = make_double2(0.0, 0.0);
      for (int k = 0; k < T; ++k) acc = cfma(acc, s_p[r][k], s_t[k][c]);
      s_u[r][c] = acc;
    }
    __syncthreads();
  }
  const S* Y = st.W + static_cast<size_t>(b) * n * st.ld + n;
  double2* O = o.out + static_cast<size_t>(item) * n * n;
  const int ncols = o.vw > 0 ? T * o.vw : n;
  for (int rc = tid; rc < ncols; rc += 256) {
    const int cj = o.vw > 0 ? rc / o.vw : rc / nv;
    const int vj = o.vw > 0 ? o.vlo + (rc - cj * o.vw) : rc - cj * nv;
    const int j = cj * nv + vj;
    double2 y[T], vc[T];
#pragma unroll
    for (int s = 0; s < T; ++s) y[s] = to_c128(Y[static_cast<size_t>(s * nv + v) * st.ld + rc]);
    if (f2) {
#pragma unroll
      for (int s = 0; s < T; ++s) vc[s] = o.vrg[((static_cast<size_t>(w) * TT + s * T + cj) * nv + vj) * nq + q];
    }
#pragma unroll 1
    for (int ri = 0; ri < T; ++ri) {
      double2 dy = make_double2(0.0, 0.0);
#pragma unroll
      for (int s = 0; s < T; ++s) dy = cfma(dy, s_d[ri][s], y[s]);
      double2 f = o.z_form ? dy : csub(vj == v ? s_d[ri][cj] : make_double2(0.0, 0.0), dy);
      f = cscale(f, o.beta2);
      if (f2) {
#pragma unroll
        for (int s = 0; s < T; ++s) f = cfma(f, s_u[ri][s], vc[s]);
      }
      O[static_cast<size_t>(ri * nv + v) * n + j] = f;
    }
  }
}

// ---------------------------------------------------------------------------
// Pairing vertices at w_pp = 0 from F_dens and F_magn (half bosonic range [nw][nq][n][n]):
//   F_pp[q][o1][o2][o3][o4][i][j] = -F(w = i - j)[(o1,o2, i + off)][(o4,o3, 2np-1-j + off)]   (w >= 0)
//                                 = -conj F(-w)[(o1,o2, 2np-1-i + off)][(o4,o3, j + off)]      (w < 0)
// with off = niv - np, zero where |w| >= nw; gamma_sing = 0.5 F_pp,d - 1.5 F_pp,m and
// gamma_trip = 0.5 F_pp,d + 0.5 F_pp,m; without fm, gs = F_pp of fd alone. One thread per
// output element, j fastest.
// f_dens / f_magn hold the bosonic frequencies [w0, w0 + nw); entries whose |w| lies outside that
// range are zeroed (fill) or left as they are (a later call of another range writes them).
__global__ void __launch_bounds__(256) vtx_pp_kernel(const double2* __restrict__ fd, const double2* __restrict__ fm,
                                                     double2* __restrict__ gs, double2* __restrict__ gt, int nq, int nb,
                                                     int nw, int niv, int np, int w0, int fill) {
  const int T = nb * nb, TT = T * T, P2 = 2 * np, nv = 2 * niv, n = T * nv, off = niv - np;
  const size_t total = static_cast<size_t>(nq) * TT * P2 * P2;
  for (size_t idx = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
       idx += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int j = static_cast<int>(idx % P2);
    const int i = static_cast<int>((idx / P2) % P2);
    const size_t rest = idx / (static_cast<size_t>(P2) * P2);
    const int o = static_cast<int>(rest % TT), q = static_cast<int>(rest / TT);
    const int o4 = o % nb, o3 = (o / nb) % nb, o2 = (o / T) % nb, o1 = o / (T * nb);
    const int ri = o1 * nb + o2, cj = o4 * nb + o3;
    const int w = i - j;
    const int lw = (w >= 0 ? w : -w) - w0;  // position of |w| in the given range
    if (!fill && (lw < 0 || lw >= nw)) continue;
    double2 vd = make_double2(0.0, 0.0), vm = vd;
    if (w >= 0 && lw >= 0 && lw < nw) {
      const size_t e = ((static_cast<size_t>(lw) * nq + q) * n + ri * nv + i + off) * n + cj * nv + (P2 - 1 - j) + off;
      vd = fd[e];
      if (fm) vm = fm[e];
    } else if (w < 0 && lw >= 0 && lw < nw) {
      const size_t e = ((static_cast<size_t>(lw) * nq + q) * n + ri * nv + (P2 - 1 - i) + off) * n + cj * nv + j + off;
      vd = cconj(fd[e]);
      if (fm) vm = cconj(fm[e]);
    }
    if (!fm) {  // one channel: the transformed vertex itself
      gs[idx] = make_double2(-vd.x, -vd.y);
      continue;
    }
    gs[idx] = make_double2(-(0.5 * vd.x - 1.5 * vm.x), -(0.5 * vd.y - 1.5 * vm.y));
    if (gt) gt[idx] = make_double2(-(0.5 * vd.x + 0.5 * vm.x), -(0.5 * vd.y + 0.5 * vm.y));
  }
}

// ---------------------------------------------------------------------------
enum VertexOut { VTX_OUT_INVERSE = 0, VTX_OUT_FULL = 1 };

template <typename S>
static int vertex_solve_t(LadderArgs a, int what, const VertexArgs& o_in, int gemm_mode, hipStream_t stream) {
  DGAMI_ARG_CHECK(o_in.vw == 0 || what == VTX_OUT_FULL, "vertex: a frequency window needs the full vertex");
  const int n = a.n, nr = o_in.vw > 0 ? a.T * o_in.vw : n, ncol = n + nr;
  // complex64 factorisations carry the vertex part as right-hand sides (no cancellation in F)
  constexpr bool c64 = std::is_same<S, float2>::value;
  // rows start on 128-byte boundaries
  constexpr int align = c64 ? 16 : 8;
  const int ld = (ncol + align - 1) & ~(align - 1);
  DGAMI_ARG_CHECK(n <= BLK_MAX_N, "vertex: compound size above 1536 is not supported");
  a.rhs_mode = what == VTX_OUT_FULL ? (c64 ? LADDER_RHS_VERTEX : LADDER_RHS_CHI0INV) : LADDER_RHS_IDENTITY;
  a.rhs_vlo = o_in.vlo;
  a.rhs_w = o_in.vw;
  VertexArgs o = o_in;
  o.z_form = c64 && what == VTX_OUT_FULL;
  const size_t items = static_cast<size_t>(a.nw) * a.nq;
  const size_t per = static_cast<size_t>(n) * ld * sizeof(S);
  const int nl = static_cast<int>(std::min<size_t>(ladder_lanes(), items));
  // up to 32 GiB of problems (factors and n right-hand sides) per pass over all lanes
  size_t batch = (size_t(32) << 30) / (per * nl);
  if (batch < 1) batch = 1;
  if (batch > (items + nl - 1) / nl) batch = (items + nl - 1) / nl;
  if (batch > 65535) batch = 65535;
  const int nblk = static_cast<int>(ceil_div(n, TRI_B));
  // lanes (ladder_lanes): parts of the problems on the caller's and the library's extra streams
  struct Lane {
    BlockedState<S> st;
    S* tri;
    hipStream_t s;
  };
  Lane lanes[1 + MAX_EXTRA_STREAMS];
  for (int l = 0; l < nl; ++l) {
    Lane& L = lanes[l];
    if (!blocked_alloc(L.st, n, ld, ncol, batch, WS_LADDER + l * WS_STREAM_STRIDE)) return -2;
    if constexpr (std::is_same<S, float2>::value) L.st.gemm_mode = gemm_mode;
    L.tri = static_cast<S*>(workspace(WS_TRI + l * WS_STREAM_STRIDE, batch * nblk * 2 * TRI_B * TRI_B * sizeof(S)));
    if (!L.tri) return -2;
    L.s = stream;
    if (l > 0) {
      const int rc = extra_stream_fork(stream, l, &L.s);
      if (rc) return rc;
    }
  }
  const int ntile = static_cast<int>(ceil_div(nr, 64));
  int lane = 0;
  for (size_t item0 = 0; item0 < items; item0 += batch, lane = (lane + 1) % nl) {
    BlockedState<S>& st = lanes[lane].st;
    S* tri = lanes[lane].tri;
    hipStream_t stream = lanes[lane].s;  // this part's stream
    const int nb = static_cast<int>(items - item0 < batch ? items - item0 : batch);
    int rc = blocked_factor<S>(a, st, static_cast<int>(item0), nb, stream);
    if (rc) return rc;
    TriFactors<S> f{st.W, tri, st.dinv, n, ld, nblk};
    rc = tri_diag_inverse(f, nb, false, stream);
    if (rc) return rc;
    for (int c0 = (n - 1) / VTX_STEP * VTX_STEP; c0 >= 0; c0 -= VTX_STEP) {
      const int cb = min(VTX_STEP, n - c0);
      hipLaunchKernelGGL(vtx_backsolve_kernel<S>, dim3(static_cast<unsigned>(xcd_grid(nb, ntile))), dim3(256), 0,
                         stream, st, tri, nblk, c0, cb, nr, ntile, nb);
      DGAMI_HIP_CHECK(hipGetLastError());
      if (c0 > 0) {
        if constexpr (std::is_same<S, float2>::value)
          rc = blocked_gemm_c64(st, nb, 0, c0, n, n + nr, c0, cb, stream);
        else
          rc = blocked_gemm_fp64(st, nb, 0, c0, n, n + nr, c0, cb, stream);
        if (rc) return rc;
      }
    }
    if (what == VTX_OUT_INVERSE) {
      hipLaunchKernelGGL(vtx_copy_kernel<S>, dim3(static_cast<unsigned>(ceil_div(n, 4)), nb), dim3(256), 0, stream, st,
                         o.out, static_cast<int>(item0));
    } else {
#define DGAMI_VTX_FULL(TV)                                                                                       \
  hipLaunchKernelGGL((vtx_full_kernel<TV, S>), dim3(o.vw > 0 ? o.vw : a.nv, nb), dim3(256), 0, stream, a, st, o, \
                     static_cast<int>(item0))
      switch (a.T) {
        case 1: DGAMI_VTX_FULL(1); break;
        case 4: DGAMI_VTX_FULL(4); break;
        case 9: DGAMI_VTX_FULL(9); break;
        case 16: DGAMI_VTX_FULL(16); break;
        default: DGAMI_ARG_CHECK(false, "vertex: nb must be 1..4");
      }
#undef DGAMI_VTX_FULL
    }
    DGAMI_HIP_CHECK(hipGetLastError());
    rc = blocked_copy_info(st.info, a.info, static_cast<int>(item0), nb, stream);
    if (rc) return rc;
  }
  for (int l = 1; l < nl; ++l) {
    const int rc = extra_stream_join(stream, l);
    if (rc) return rc;
  }
  return 0;
}

// precision: DGAMI_PLAIN_FP64 (complex128 LU), or a complex64 LU with FP32 (DGAMI_MIXED_FP32_IR) or
// split-BF16 (DGAMI_MIXED_BF16X3_IR) matrix-core updates and no FP64 refinement.
static int vertex_solve(LadderArgs a, int what, const VertexArgs& o, hipStream_t stream, int precision = 0) {
  if (precision == 0) return vertex_solve_t<double2>(a, what, o, 0, stream);
  DGAMI_ARG_CHECK(precision == 1 || precision == 2, "vertex: precision must be 0, 1 or 2");
  return vertex_solve_t<float2>(a, what, o, precision == 2 ? BLK_GEMM_BF16X3 : BLK_GEMM_F32_MFMA, stream);
}

static LadderArgs vertex_ladder_args(int nq, int nb, int nw, int niv, double beta, const dgami_c128* chi0_inv,
                                     const dgami_c128* gamma, const dgami_c128* u_r, const dgami_c128* v_r,
                                     int32_t* info) {
  LadderArgs a{};
  a.chi0_inv = reinterpret_cast<const double2*>(chi0_inv);
  a.gamma = reinterpret_cast<const double2*>(gamma);
  a.u_r = reinterpret_cast<const double2*>(u_r);
  a.v_r = reinterpret_cast<const double2*>(v_r);
  a.out = nullptr;
  a.info = info;
  a.nchan = 1;
  a.nq = nq;
  a.nw = nw;
  a.T = nb * nb;
  a.nv = 2 * niv;
  a.n = a.T * a.nv;
  a.inv_beta2 = 1.0 / (beta * beta);
  a.inv_beta = 1.0 / beta;
  a.dense = 0;
  return a;
}

}  // namespace dgami

extern "C" int dgami_ladder_invert(int nq, int nb, int nw, int niv, double beta, const dgami_c128* chi0_inv,
                                   const dgami_c128* gamma, const dgami_c128* u_r, const dgami_c128* v_r,
                                   dgami_c128* out, int32_t* info, void* stream) {
  using namespace dgami;
  DGAMI_ARG_CHECK(chi0_inv && gamma && u_r && out, "ladder_invert: null pointer");
  DGAMI_ARG_CHECK(nq >= 1 && nb >= 1 && nb <= 4 && nw >= 1 && niv >= 1 && beta > 0, "ladder_invert: invalid sizes");
  const LadderArgs a = vertex_ladder_args(nq, nb, nw, niv, beta, chi0_inv, gamma, u_r, v_r, info);
  VertexArgs o{nullptr, nullptr, reinterpret_cast<double2*>(out), 0.0, 0, 0, 0};
  return vertex_solve(a, VTX_OUT_INVERSE, o, static_cast<hipStream_t>(stream));
}

extern "C" int dgami_invert_dense(int nmat, int n, const dgami_c128* a, dgami_c128* a_inv, int32_t* info,
                                  void* stream) {
  using namespace dgami;
  DGAMI_ARG_CHECK(a && a_inv, "invert_dense: null pointer");
  DGAMI_ARG_CHECK(nmat >= 1 && n >= 1 && n <= BLK_MAX_N, "invert_dense: invalid sizes (1 <= n <= 1536)");
  LadderArgs la = vertex_ladder_args(nmat, 1, 1, 1, 1.0, nullptr, a, nullptr, nullptr, info);
  la.dense = 1;  // item q reads its matrix from a + q n^2
  la.T = 1;
  la.nv = n;
  la.n = n;
  la.inv_beta2 = 1.0;
  VertexArgs o{nullptr, nullptr, reinterpret_cast<double2*>(a_inv), 0.0, 0, 0, 0};
  return vertex_solve(la, VTX_OUT_INVERSE, o, static_cast<hipStream_t>(stream));
}

extern "C" int dgami_full_vertex(int nq, int nb, int nw, int niv, double beta, const dgami_c128* chi0_inv,
                                 const dgami_c128* gamma, const dgami_c128* u_r, const dgami_c128* v_r,
                                 const dgami_c128* vrg, const dgami_c128* chi_phys, dgami_c128* out, int32_t* info,
                                 void* stream) {
  using namespace dgami;
  DGAMI_ARG_CHECK(chi0_inv && gamma && u_r && out, "full_vertex: null pointer");
  DGAMI_ARG_CHECK((vrg == nullptr) == (chi_phys == nullptr), "full_vertex: vrg and chi_phys go together");
  DGAMI_ARG_CHECK(nq >= 1 && nb >= 1 && nb <= 4 && nw >= 1 && niv >= 1 && beta > 0, "full_vertex: invalid sizes");
  const LadderArgs a = vertex_ladder_args(nq, nb, nw, niv, beta, chi0_inv, gamma, u_r, v_r, info);
  VertexArgs o{reinterpret_cast<const double2*>(vrg), reinterpret_cast<const double2*>(chi_phys),
               reinterpret_cast<double2*>(out), beta * beta, 0, 0, 0};
  return vertex_solve(a, VTX_OUT_FULL, o, static_cast<hipStream_t>(stream));
}

extern "C" int dgami_full_vertex_prec(int nq, int nb, int nw, int niv, double beta, int precision, int v_lo,
                                      int v_hi, const dgami_c128* chi0_inv, const dgami_c128* gamma,
                                      const dgami_c128* u_r, const dgami_c128* v_r, const dgami_c128* vrg,
                                      const dgami_c128* chi_phys, dgami_c128* out, int32_t* info, void* stream) {
  using namespace dgami;
  DGAMI_ARG_CHECK(chi0_inv && gamma && u_r && out, "full_vertex: null pointer");
  DGAMI_ARG_CHECK((vrg == nullptr) == (chi_phys == nullptr), "full_vertex: vrg and chi_phys go together");
  DGAMI_ARG_CHECK(nq >= 1 && nb >= 1 && nb <= 4 && nw >= 1 && niv >= 1 && beta > 0, "full_vertex: invalid sizes");
  DGAMI_ARG_CHECK(precision >= 0 && precision <= 2, "full_vertex: precision must be 0, 1 or 2");
  DGAMI_ARG_CHECK((v_lo == 0 && v_hi == 0) || (0 <= v_lo && v_lo < v_hi && v_hi <= 2 * niv),
                  "full_vertex: the window needs 0 <= v_lo < v_hi <= 2 niv (or both 0)");
  const LadderArgs a = vertex_ladder_args(nq, nb, nw, niv, beta, chi0_inv, gamma, u_r, v_r, info);
  VertexArgs o{reinterpret_cast<const double2*>(vrg), reinterpret_cast<const double2*>(chi_phys),
               reinterpret_cast<double2*>(out), beta * beta, 0, v_lo, v_hi - v_lo};
  return vertex_solve(a, VTX_OUT_FULL, o, static_cast<hipStream_t>(stream), precision);
}

namespace dgami {
static int pairing_vertex_pp(int nq, int nb, int nw, int w0, int fill, int niv, int np, const dgami_c128* f_dens,
                             const dgami_c128* f_magn, dgami_c128* gamma_sing, dgami_c128* gamma_trip,
                             hipStream_t stream) {
  DGAMI_ARG_CHECK(f_dens && gamma_sing, "pairing_vertex_pp: null pointer");
  DGAMI_ARG_CHECK(f_magn || !gamma_trip, "pairing_vertex_pp: the triplet vertex needs f_magn");
  DGAMI_ARG_CHECK(nq >= 1 && nb >= 1 && nb <= 4 && nw >= 1 && w0 >= 0 && niv >= 1 && np >= 1 && np <= niv,
                  "pairing_vertex_pp: invalid sizes (need 1 <= np <= niv, w0 >= 0)");
  const size_t total = static_cast<size_t>(nq) * nb * nb * nb * nb * 4 * np * np;
  const size_t blocks = ceil_div(total, 256);
  hipLaunchKernelGGL(vtx_pp_kernel, dim3(static_cast<unsigned>(blocks < 65536 ? blocks : 65536)), dim3(256), 0, stream,
                     reinterpret_cast<const double2*>(f_dens), reinterpret_cast<const double2*>(f_magn),
                     reinterpret_cast<double2*>(gamma_sing), reinterpret_cast<double2*>(gamma_trip), nq, nb, nw, niv,
                     np, w0, fill);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}
}  // namespace dgami

extern "C" int dgami_pairing_vertex_pp(int nq, int nb, int nw, int niv, int np, const dgami_c128* f_dens,
                                       const dgami_c128* f_magn, dgami_c128* gamma_sing, dgami_c128* gamma_trip,
                                       void* stream) {
  return dgami::pairing_vertex_pp(nq, nb, nw, 0, 1, niv, np, f_dens, f_magn, gamma_sing, gamma_trip,
                                  static_cast<hipStream_t>(stream));
}

extern "C" int dgami_pairing_vertex_pp_range(int nq, int nb, int nw, int w0, int niv, int np, const dgami_c128* f_dens,
                                             const dgami_c128* f_magn, dgami_c128* gamma_sing, dgami_c128* gamma_trip,
                                             void* stream) {
  return dgami::pairing_vertex_pp(nq, nb, nw, w0, 0, niv, np, f_dens, f_magn, gamma_sing, gamma_trip,
                                  static_cast<hipStream_t>(stream));
}
