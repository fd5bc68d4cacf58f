// Shared definitions of the ladder solves (plain FP64 and mixed precision).
#pragma once

#include "dgami_internal.hpp"

namespace dgami {

// One batch of ladder problems: item = (chan, w, q), q fastest.
struct LadderArgs {
  const double2* chi0_inv;  // QF [nw][T*T][nv][nq]
  const double2* gamma;     // [nchan][nw][n][n]
  const double2* u_r;       // [nchan][T*T]
  const double2* v_r;       // [nchan][nq][T*T] or nullptr
  double2* out;             // [nchan][nw][T*T][nv][nq]
  int32_t* info;            // [nchan][nw][nq]
  int nchan, nq, nw;
  int dense;  // 1: gamma holds the assembled matrix of every item ([item][n][n]), out is [item][n][T]
  int T;     // nb^2 (number of right-hand sides)
  int nv;    // 2 * niv
  int n;     // T * nv
  double inv_beta2, inv_beta;
  int rhs_mode = 0;  // LadderRhs: right-hand sides carried by the blocked LU
  // CHI0INV / VERTEX right-hand sides restricted to the columns (c, v') with v' in
  // [rhs_vlo, rhs_vlo + rhs_w) for every orbital pair c (rhs_w = 0: every column)
  int rhs_vlo = 0, rhs_w = 0;
};

// Compound column of right-hand side r (CHI0INV / VERTEX modes).
__device__ __forceinline__ int ladder_rhs_col(const LadderArgs& a, int r) {
  if (a.rhs_w == 0) return r;
  const int c = r / a.rhs_w;
  return c * a.nv + a.rhs_vlo + (r - c * a.rhs_w);
}

// Right-hand sides of the blocked LU: the nb^2 indicator columns of the summed ladder, the
// identity (full inverse), the block-diagonal chi0^-1 (X chi0^-1 for the full vertex) or the
// vertex part A - chi0^-1 = (gamma - u_r - v_r)/beta^2 (X (A - chi0^-1) for the full vertex
// without the cancellation of chi0^-1 - chi0^-1 X chi0^-1).
enum LadderRhs { LADDER_RHS_INDICATOR = 0, LADDER_RHS_IDENTITY = 1, LADDER_RHS_CHI0INV = 2, LADDER_RHS_VERTEX = 3 };

// (gamma - u_r - v_r)[i][j] / beta^2 in compound indices (non-dense items).
__device__ __forceinline__ double2 ladder_vertex_part(const LadderArgs& a, const double2* __restrict__ G,
                                                      const double2* __restrict__ U, const double2* __restrict__ V,
                                                      int i, int j) {
  const int T = a.T, ri = i / a.nv, cj = j / a.nv;
  double2 uv = U[ri * T + cj];
  if (V) uv = cadd(uv, V[ri * T + cj]);
  return cscale(csub(G[static_cast<size_t>(i) * a.n + j], uv), a.inv_beta2);
}

// Element (i, j) of the ladder matrix A(chan, w, q) in compound indices
// (nonlocal_sde.py:93-95): chi0_inv delta_vv' + (gamma - u_r - v_r)/beta^2.
__device__ __forceinline__ double2 ladder_elem(const LadderArgs& a, const double2* __restrict__ G,
                                               const double2* __restrict__ U, const double2* __restrict__ V, int w,
                                               int q, int i, int j) {
  if (a.dense) return G[static_cast<size_t>(i) * a.n + j];
  const int nv = a.nv, T = a.T;
  const int ri = i / nv, vi = i - ri * nv;
  const int cj = j / nv, vj = j - cj * nv;
  double2 val = ladder_vertex_part(a, G, U, V, i, j);
  if (vi == vj) val = cadd(val, a.chi0_inv[((static_cast<size_t>(w) * T * T + ri * T + cj) * nv + vi) * a.nq + q]);
  return val;
}

// Indicator right-hand side B[i][r] = 1 iff the orbital pair of row i is r (four_point.py:661-668).
__device__ __forceinline__ double2 ladder_rhs(const LadderArgs& a, int i, int r) {
  return make_double2((i / a.nv) == r ? 1.0 : 0.0, 0.0);
}

// Right-hand side B[i][r] of item (w, q) for every LadderRhs mode.
__device__ __forceinline__ double2 ladder_rhs_at(const LadderArgs& a, int w, int q, int i, int r) {
  if (a.rhs_mode == LADDER_RHS_IDENTITY) return make_double2(i == r ? 1.0 : 0.0, 0.0);
  if (a.rhs_mode == LADDER_RHS_CHI0INV) {
    const int nv = a.nv, T = a.T;
    const int ri = i / nv, vi = i - ri * nv;
    const int col = ladder_rhs_col(a, r);
    const int cj = col / nv, vj = col - cj * nv;
    if (vi != vj) return make_double2(0.0, 0.0);
    return a.chi0_inv[((static_cast<size_t>(w) * T * T + ri * T + cj) * nv + vi) * a.nq + q];
  }
  return ladder_rhs(a, i, r);
}

// Output element chi_aux_sum[chan][w][(ri, r)][vi][q] = x[i][r] / beta.
__device__ __forceinline__ size_t ladder_out_index(const LadderArgs& a, int chan, int w, int q, int i, int r) {
  if (a.dense) return ((static_cast<size_t>(chan) * a.nw + w) * a.nq + q) * a.n * a.T + static_cast<size_t>(i) * a.T + r;
  const int ri = i / a.nv, vi = i - ri * a.nv;
  return ((((static_cast<size_t>(chan) * a.nw + w) * a.T * a.T + ri * a.T + r) * a.nv) + vi) * a.nq + q;
}

// Matrix of one item: the shared vertex Gamma(chan, w), or the item's own matrix in dense mode.
__device__ __forceinline__ const double2* ladder_matrix_ptr(const LadderArgs& a, int chan, int w, int q) {
  const size_t nn = static_cast<size_t>(a.n) * a.n;
  if (a.dense) return a.gamma + ((static_cast<size_t>(chan) * a.nw + w) * a.nq + q) * nn;
  return a.gamma + (static_cast<size_t>(chan) * a.nw + w) * nn;
}

// Operands of one item for its augmented matrix [A | B]: the matrix (shared vertex or dense),
// U_r and V_r(q) (non-dense only), and its (w, q).
struct LadderSrc {
  const double2* G;
  const double2* U;
  const double2* V;
  int w, q;
};

__device__ __forceinline__ LadderSrc ladder_src(const LadderArgs& a, int item) {
  const int q = item % a.nq, w = (item / a.nq) % a.nw, chan = item / (a.nq * a.nw);
  LadderSrc s;
  s.G = ladder_matrix_ptr(a, chan, w, q);
  s.U = a.dense ? nullptr : a.u_r + chan * a.T * a.T;
  s.V = (a.v_r && !a.dense) ? a.v_r + (static_cast<size_t>(chan) * a.nq + q) * a.T * a.T : nullptr;
  s.w = w;
  s.q = q;
  return s;
}

// Element (i, j) of the augmented matrix [A | B] of an item, j < n + (right-hand sides).
__device__ __forceinline__ double2 ladder_aug_at(const LadderArgs& a, const LadderSrc& s, int i, int j) {
  if (j < a.n) return ladder_elem(a, s.G, s.U, s.V, s.w, s.q, i, j);
  if (a.rhs_mode == LADDER_RHS_VERTEX) return ladder_vertex_part(a, s.G, s.U, s.V, i, ladder_rhs_col(a, j - a.n));
  return ladder_rhs_at(a, s.w, s.q, i, j - a.n);
}

// Up to LADDER_MULTI_MAX models solved by one launch of the register-resident kernel.
constexpr int LADDER_MULTI_MAX = 16;
struct LadderMulti {
  LadderArgs model[LADDER_MULTI_MAX];
  int start[LADDER_MULTI_MAX + 1];  // first workgroup of each model; start[nmodels] = total
  int nmodels;
};

int ladder_solve_fp64(const LadderArgs& a, hipStream_t stream);
int ladder_solve_mixed(const LadderArgs& a, int precision, hipStream_t stream);

}  // namespace dgami
