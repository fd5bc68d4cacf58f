// Mixed-precision ladder solves: complex64 LU factorisation, FP64 iterative refinement.
//
// Every item's matrix A is factored once in FP32 (PA = LU, batched blocked LU of
// ladder_blocked.hip with the interchanges also applied left of each panel; trailing
// updates on the FP32 matrix cores, or as three split-BF16 MFMA products). The
// solution X of A X = B (B = indicator right-hand sides, the nu' sum of A^-1)
// is then refined in FP64:
//   R = B - A X      in FP64 from the FP64 operands, with the vertex term Gamma(chan,w) X
//                    formed for all q of one (chan, w) as a single complex128 GEMM
//                    (Gamma is shared by the q's; the chi0^-1 blocks and the rank-nb^2
//                    U_r + V_r term are applied per item);
//   D = (LU)^-1 P R  in FP32 (blocked triangular sweeps with inverse diagonal blocks);
//   X += D           in FP64.
// Starting from X = 0 (R = B), every item is refined until its predicted relative error is
// below 1e-10 (see MIXED_TOL). Each refinement step multiplies the error by about
// cond(A) * 2^-24 (FP32 factors) or cond(A) * 2^-16 (split-BF16 updates), so the number of
// solves differs between items: the host issues a fixed maximum number of refinement steps,
// each item's status lives on the device, and the kernels of a step skip items that have
// converged (and the vertex GEMM skips (channel, w) groups with no active item). No value
// is read back by the host, so a call can be captured in a HIP graph. An item that has not
// converged after the last step, or whose correction became non-finite, is reported in
// info (DGAMI_INFO_REFINE_STALLED, DGAMI_INFO_REFINE_NONFINITE).
#include <algorithm>
#include <cstdlib>

#include "dgami.h"
#include "dgami_internal.hpp"
#include "dgami_launch.hpp"
#include "ladder_blocked.hpp"
#include "ladder_common.hpp"
#include "ladder_tune.hpp"
#include "tri_solve.hpp"
#include "zgemm.hpp"

namespace dgami {

// Refinement stops when the predicted relative error of X, |D_s| (|D_s| / |D_s-1|) / |X|,
// falls below MIXED_TOL (the contraction |D_s| / |D_s-1| estimates the error factor of one
// refinement step), or after the maximum number of solves.
constexpr int MIXED_MAX_SOLVES_FP32 = 4;
constexpr int MIXED_MAX_SOLVES_BF16 = 8;
constexpr double MIXED_TOL = 1e-10;

// Refinement status of an item (device array, one int per item of the batch).
enum MixStatus { MIX_ACTIVE = 0, MIX_CONVERGED = 1, MIX_STALLED = 2, MIX_NONFINITE = 3 };

// Tolerance on the predicted relative error; DGAMI_MIXED_TOL overrides it (for measurements).
static double mixed_tol() {
  const char* v = getenv("DGAMI_MIXED_TOL");
  const double t = v ? atof(v) : 0.0;
  return t > 0.0 ? t : MIXED_TOL;
}

// Item layout of the refinement vectors: group g = (chan, w) (or the item itself in
// dense mode) holds gq items; X and R are [group][n][gq][T] so that Gamma(g) X(g) is
// one GEMM with N = gq * T.
struct MixedArgs {
  LadderArgs a;
  const float2* W;      // factors [item][n][ld]
  const int* perm;      // [item][n] row i of P r is r[perm[i]]
  const int* pinv;      // [item][n] row j of r is row pinv[j] of P r
  const double2* dinv;  // [item][n] inverse pivots
  double2* X;           // [group][n][gq][T]
  double2* R;           // same
  int* status;          // [item of the batch] MixStatus
  float* prev;          // [item of the batch] relative correction of the previous refinement step
  int* gactive;         // [group of the batch] 1 while the group has an active item
  int ld, gq, group0, item0;
};

// ---------------------------------------------------------------------------
// Row permutation of each factorisation from its sequence of interchanges, and its inverse.
// Also starts the item's refinement (status active, no previous correction).
__global__ void __launch_bounds__(64) mix_perm_kernel(const int* __restrict__ piv, int* __restrict__ perm,
                                                      int* __restrict__ pinv, int n, int* __restrict__ status,
                                                      float* __restrict__ prev) {
  extern __shared__ int s_perm[];
  const int b = blockIdx.x;
  if (threadIdx.x == 0) {
    status[b] = MIX_ACTIVE;
    prev[b] = 0.f;
  }
  for (int i = threadIdx.x; i < n; i += 64) s_perm[i] = i;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int* pv = piv + static_cast<size_t>(b) * n;
    for (int j = 0; j < n; ++j) {
      const int p = pv[j];
      const int t = s_perm[j];
      s_perm[j] = s_perm[p];
      s_perm[p] = t;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += 64) {
    perm[static_cast<size_t>(b) * n + i] = s_perm[i];
    pinv[static_cast<size_t>(b) * n + s_perm[i]] = i;
  }
}

// ---------------------------------------------------------------------------
// R = B - A X given R = Gamma X on entry (vertex term of the GEMM). One workgroup per item.
__global__ void __launch_bounds__(256) mix_residual_kernel(MixedArgs m) {
  const LadderArgs& a = m.a;
  const int lb = blockIdx.x;                  // item within the batch
  if (m.status[lb] != MIX_ACTIVE) return;
  const int gl = lb / m.gq, q = lb - gl * m.gq;  // local group, q within the group
  const int item = m.item0 + lb;
  const int n = a.n, T = a.T, nv = a.nv, gq = m.gq;
  const size_t gbase = static_cast<size_t>(gl) * n * gq * T;
  const double2* X = m.X + gbase;
  double2* R = m.R + gbase;
  __shared__ double2 s_uv[16][16];
  __shared__ double2 s_s[16][16];
  const int tid = threadIdx.x;
  if (a.dense) {
    for (int idx = tid; idx < n * T; idx += 256) {
      const int i = idx / T, t = idx - (idx / T) * T;
      const size_t o = (static_cast<size_t>(i) * gq + q) * T + t;
      R[o] = csub(ladder_rhs(a, i, t), R[o]);
    }
    return;
  }
  const int w = (item / a.nq) % a.nw, chan = item / (a.nq * a.nw);
  // S[c][t] = sum_nu X[(c, nu)][t]; UV = U_r + V_r(q)
  for (int idx = tid; idx < T * T; idx += 256) {
    const int c = idx / T, t = idx - (idx / T) * T;
    double2 acc = make_double2(0.0, 0.0);
    for (int v = 0; v < nv; ++v) acc = cadd(acc, X[(static_cast<size_t>(c * nv + v) * gq + q) * T + t]);
    s_s[c][t] = acc;
    double2 uv = a.u_r[static_cast<size_t>(chan) * T * T + idx];
    if (a.v_r) uv = cadd(uv, a.v_r[(static_cast<size_t>(chan) * a.nq + q) * T * T + idx]);
    s_uv[c][t] = uv;
  }
  __syncthreads();
  const double2* ci = a.chi0_inv + static_cast<size_t>(w) * T * T * nv * a.nq;
  for (int idx = tid; idx < n * T; idx += 256) {
    const int i = idx / T, t = idx - (idx / T) * T;
    const int ri = i / nv, vi = i - ri * nv;
    double2 uvs = make_double2(0.0, 0.0);
    double2 cx = make_double2(0.0, 0.0);
    for (int c = 0; c < T; ++c) {
      uvs = cfma(uvs, s_uv[ri][c], s_s[c][t]);
      cx = cfma(cx, ci[(static_cast<size_t>(ri * T + c) * nv + vi) * a.nq + q],
                X[(static_cast<size_t>(c * nv + vi) * gq + q) * T + t]);
    }
    const size_t o = (static_cast<size_t>(i) * gq + q) * T + t;
    const double2 ax = cadd(cx, cscale(csub(R[o], uvs), a.inv_beta2));
    R[o] = csub(ladder_rhs(a, i, t), ax);
  }
}

// ---------------------------------------------------------------------------
// Y = P R in complex64 (R = B on the first solve). grid (ceil(n*T/256), batch).
__global__ void __launch_bounds__(256) mix_gather_kernel(MixedArgs m, float2* __restrict__ Y, int first) {
  const LadderArgs& a = m.a;
  const int lb = blockIdx.y;
  if (m.status[lb] != MIX_ACTIVE) return;
  const int gl = lb / m.gq, q = lb - gl * m.gq;
  const int n = a.n, T = a.T, gq = m.gq;
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= n * T) return;
  const int i = idx / T, t = idx - (idx / T) * T;
  const int src = m.perm[static_cast<size_t>(lb) * n + i];
  const double2* R = m.R + static_cast<size_t>(gl) * n * gq * T;
  Y[static_cast<size_t>(lb) * n * T + idx] =
      first ? from_c128<float2>(ladder_rhs(a, src, t)) : from_c128<float2>(R[(static_cast<size_t>(src) * gq + q) * T + t]);
}

// The same residual with the q-points of a (channel, w) group as the fastest index, so that
// X, R and chi0^-1 (whose q is also fastest) are read in whole cache lines. It runs in two
// kernels: mix_vsum_kernel forms the U_r/V_r term, and mix_residual_rows_kernel forms one row i
// of every item of the group and writes it, in complex64, to its row of Y = P R (the
// sweeps' input), so R is not written back and no gather runs. The sums run in the same order
// as in mix_residual_kernel, so the results are the same bit for bit.
//
// UVS[g][ri][q][t] = sum_c (U_r + V_r(q))[ri][c] S[c][q][t], S[c][q][t] = sum_nu X[(c, nu)][q][t].
// grid (ceil(gq*T/128), groups of the batch), one thread per (q, t).
template <int T>
__global__ void __launch_bounds__(128) mix_vsum_kernel(MixedArgs m, double2* __restrict__ uvs) {
  const LadderArgs& a = m.a;
  const int gl = blockIdx.y;
  if (!m.gactive[gl]) return;
  const int nv = a.nv, gq = m.gq, n = a.n;
  const int e = blockIdx.x * 128 + threadIdx.x;
  if (e >= gq * T) return;
  const int q = e / T;
  const int chan = (m.group0 + gl) / a.nw;
  const size_t rstride = static_cast<size_t>(gq) * T;  // one row of X
  const double2* X = m.X + static_cast<size_t>(gl) * n * rstride + e;
  double2 s[T];
#pragma unroll
  for (int c = 0; c < T; ++c) s[c] = make_double2(0.0, 0.0);
  for (int v = 0; v < nv; ++v) {
#pragma unroll
    for (int c = 0; c < T; ++c) s[c] = cadd(s[c], X[static_cast<size_t>(c * nv + v) * rstride]);
  }
  const double2* u = a.u_r + static_cast<size_t>(chan) * T * T;
  const double2* vr = a.v_r ? a.v_r + (static_cast<size_t>(chan) * a.nq + q) * T * T : nullptr;
  double2* out = uvs + static_cast<size_t>(gl) * T * rstride + e;
  for (int ri = 0; ri < T; ++ri) {
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int c = 0; c < T; ++c) {
      double2 uv = u[ri * T + c];
      if (vr) uv = cadd(uv, vr[ri * T + c]);
      acc = cfma(acc, uv, s[c]);
    }
    out[static_cast<size_t>(ri) * rstride] = acc;
  }
}

// Y[item][pinv[i]][t] = B - A X for row i = (ri, vi) of every active item of a group, given
// R = Gamma X and UVS. One thread per (q, t); the workgroups of one (group, vi), whose T rows
// read the same rows of X, run on one XCD (xcd_tile over (group, vi) x (ri, chunk of (q, t))).
template <int T>
__global__ void __launch_bounds__(256) mix_residual_rows_kernel(MixedArgs m, const double2* __restrict__ uvs,
                                                                float2* __restrict__ Y, int ng, int xb) {
  const LadderArgs& a = m.a;
  const int nv = a.nv, gq = m.gq, n = a.n;
  int unit, tile;
  if (!xcd_tile(blockIdx.x, ng * nv, T * xb, unit, tile)) return;
  const int gl = unit / nv, vi = unit - gl * nv;
  const int ri = tile / xb, bx = tile - ri * xb;
  if (!m.gactive[gl]) return;
  const int e = bx * 256 + threadIdx.x;
  if (e >= gq * T) return;
  const int q = e / T, t = e - q * T;
  const int lb = gl * gq + q;
  if (m.status[lb] != MIX_ACTIVE) return;
  const int i = ri * nv + vi;
  const int w = (m.group0 + gl) % a.nw;
  const size_t rstride = static_cast<size_t>(gq) * T;
  const double2* X = m.X + static_cast<size_t>(gl) * n * rstride + e;
  const double2* ci = a.chi0_inv + ((static_cast<size_t>(w) * T + ri) * T * nv + vi) * a.nq + q;
  double2 cx = make_double2(0.0, 0.0);
#pragma unroll
  for (int c = 0; c < T; ++c)
    cx = cfma(cx, ci[static_cast<size_t>(c) * nv * a.nq], X[static_cast<size_t>(c * nv + vi) * rstride]);
  const size_t o = (static_cast<size_t>(gl) * n + i) * rstride + e;
  const double2 us = uvs[(static_cast<size_t>(gl) * T + ri) * rstride + e];
  const double2 ax = cadd(cx, cscale(csub(m.R[o], us), a.inv_beta2));
  const int row = m.pinv[static_cast<size_t>(lb) * n + i];
  Y[(static_cast<size_t>(lb) * n + row) * T + t] = from_c128<float2>(csub(ladder_rhs(a, i, t), ax));
}

// The residual of the items of a batch and Y = P R in complex64, in yv. uvs (the forward
// sweep's output buffer, free until that sweep) holds ng * T * gq * T complex128 values.
static int mix_residual(const MixedArgs& m, int ng, int nit, double2* uvs, float2* yv, hipStream_t s) {
  const LadderArgs& a = m.a;
  const int T = a.T, gq = m.gq;
  auto rows = [&](auto kv, auto kr) -> int {
    hipLaunchKernelGGL(kv, dim3(ceil_div(gq * T, 128), ng), dim3(128), 0, s, m, uvs);
    DGAMI_HIP_CHECK(hipGetLastError());
    const int xb = static_cast<int>(ceil_div(gq * T, 256));
    const size_t grid = xcd_grid(static_cast<size_t>(ng) * a.nv, static_cast<size_t>(T) * xb);
    DGAMI_ARG_CHECK(grid < (size_t(1) << 31), "ladder: residual grid too large");
    hipLaunchKernelGGL(kr, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, m, uvs, yv, ng, xb);
    DGAMI_HIP_CHECK(hipGetLastError());
    return 0;
  };
  // uvs fits the forward sweep's buffer (bitems * n * T complex64) since n = T * nv with nv >= 2
  if (!a.dense && a.nv >= 2) {
    switch (T) {
      case 1: return rows(mix_vsum_kernel<1>, mix_residual_rows_kernel<1>);
      case 4: return rows(mix_vsum_kernel<4>, mix_residual_rows_kernel<4>);
      case 9: return rows(mix_vsum_kernel<9>, mix_residual_rows_kernel<9>);
      case 16: return rows(mix_vsum_kernel<16>, mix_residual_rows_kernel<16>);
      default: break;
    }
  }
  hipLaunchKernelGGL(mix_residual_kernel, dim3(nit), dim3(256), 0, s, m);
  DGAMI_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(mix_gather_kernel, dim3(ceil_div(static_cast<size_t>(a.n) * T, 256), nit), dim3(256), 0, s, m, yv, 0);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}

// X += D (FP64) for the active items, then the item's convergence test
// from its largest |D| and |X| (|Re|+|Im|): r = |D|/|X| estimates the relative error before this
// correction and r / r_prev the contraction of one refinement step, so the error after it is
// predicted as r * contraction; the item converges when that is below tol. A non-finite D or X
// ends the item's refinement as MIX_NONFINITE; an item still active after the last step is
// MIX_STALLED. One workgroup per item; step 0 (the first solve, X = D) only records r.
__global__ void __launch_bounds__(256) mix_accumulate_kernel(MixedArgs m, const float2* __restrict__ D, int step,
                                                             int last, double tol) {
  const LadderArgs& a = m.a;
  const int lb = blockIdx.x;
  if (m.status[lb] != MIX_ACTIVE) return;
  const int gl = lb / m.gq, q = lb - gl * m.gq;
  const int n = a.n, T = a.T, gq = m.gq;
  double2* X = m.X + static_cast<size_t>(gl) * n * gq * T;
  const float2* Db = D + static_cast<size_t>(lb) * n * T;
  float dmax = 0.f, xmax = 0.f;
  unsigned bad = 0;
  for (int idx = threadIdx.x; idx < n * T; idx += 256) {
    const int i = idx / T, t = idx - (idx / T) * T;
    const size_t o = (static_cast<size_t>(i) * gq + q) * T + t;
    const double2 d = to_c128(Db[idx]);
    const double2 x = step == 0 ? d : cadd(X[o], d);
    X[o] = x;
    const float dv = static_cast<float>(cabs1(d)), xv = static_cast<float>(cabs1(x));
    bad |= !(dv <= 3.0e38f) || !(xv <= 3.0e38f);  // NaN or infinite (fmaxf would drop a NaN)
    dmax = fmaxf(dmax, dv);
    xmax = fmaxf(xmax, xv);
  }
  __shared__ float s_d[4], s_x[4];
  __shared__ unsigned s_bad[4];
  const unsigned dk = wave_max_u32(__float_as_uint(dmax)), xk = wave_max_u32(__float_as_uint(xmax));
  const unsigned bk = wave_max_u32(bad);
  if ((threadIdx.x & 63) == 0) {
    s_d[threadIdx.x >> 6] = __uint_as_float(dk);
    s_x[threadIdx.x >> 6] = __uint_as_float(xk);
    s_bad[threadIdx.x >> 6] = bk;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const float d = fmaxf(fmaxf(s_d[0], s_d[1]), fmaxf(s_d[2], s_d[3]));
    const float x = fmaxf(fmaxf(s_x[0], s_x[1]), fmaxf(s_x[2], s_x[3]));
    if (s_bad[0] | s_bad[1] | s_bad[2] | s_bad[3]) {
      m.status[lb] = MIX_NONFINITE;
      return;
    }
    const double r = x > 0.f ? static_cast<double>(d) / x : (d > 0.f ? 1.0 : 0.0);
    if (step > 0) {
      const double pr = m.prev[lb];
      const double contraction = (step >= 2 && pr > 0.0) ? r / pr : r;
      if (r * contraction < tol)
        m.status[lb] = MIX_CONVERGED;
      else if (last)
        m.status[lb] = MIX_STALLED;
    }
    m.prev[lb] = static_cast<float>(r);
  }
}

// out = X / beta for every item of the batch, after the last refinement step (an item's X
// stops changing when it leaves the active state). grid (n, groups of the batch): one row i of
// every item of a group, written with q fastest as ladder_out_index orders it.
__global__ void __launch_bounds__(256) mix_output_kernel(MixedArgs m) {
  const LadderArgs& a = m.a;
  const int i = blockIdx.x, gl = blockIdx.y;
  const int n = a.n, T = a.T, gq = m.gq;
  const double2* X = m.X + (static_cast<size_t>(gl) * n + i) * gq * T;
  const int g = m.group0 + gl, w = g % a.nw, chan = g / a.nw;
  for (int e = threadIdx.x; e < gq * T; e += 256) {
    const int t = e / gq, q = e - t * gq;
    a.out[ladder_out_index(a, chan, w, q, i, t)] = cscale(X[q * T + t], a.inv_beta);
  }
}

// The same in dense mode, where X of the batch and out share the [item][n][T] layout.
__global__ void __launch_bounds__(256) mix_output_dense_kernel(MixedArgs m, size_t count) {
  const size_t e = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  if (e >= count) return;
  const LadderArgs& a = m.a;
  a.out[static_cast<size_t>(m.item0) * a.n * a.T + e] = cscale(m.X[e], a.inv_beta);
}

// gactive[g] = 1 iff an item of group g of the batch is still active (one thread per group).
__global__ void __launch_bounds__(256) mix_group_active_kernel(const int* __restrict__ status, int* __restrict__ gactive,
                                                               int gq, int ng) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= ng) return;
  int any = 0;
  for (int i = 0; i < gq; ++i) any |= status[static_cast<size_t>(g) * gq + i] == MIX_ACTIVE;
  gactive[g] = any;
}

// info of the batch's items: a zero pivot of the factorisation (k + 1) first, else the outcome
// of the refinement (0, DGAMI_INFO_REFINE_STALLED or DGAMI_INFO_REFINE_NONFINITE).
__global__ void __launch_bounds__(256) mix_info_kernel(const int* __restrict__ st_info, const int* __restrict__ status,
                                                       int32_t* __restrict__ info, int item0, int nitems) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= nitems) return;
  const int st = status[b];
  info[item0 + b] = st_info[b] != 0 ? st_info[b]
                    : st == MIX_STALLED ? DGAMI_INFO_REFINE_STALLED
                    : st == MIX_NONFINITE ? DGAMI_INFO_REFINE_NONFINITE
                                          : 0;
}

// ---------------------------------------------------------------------------
int ladder_solve_mixed(const LadderArgs& a, int precision, hipStream_t stream) {
  DGAMI_ARG_CHECK(precision == DGAMI_MIXED_FP32_IR || precision == DGAMI_MIXED_BF16X3_IR,
                  "ladder: unknown mixed-precision mode");
  // the split-BF16 factors carry ~2^-16 relative error per product instead of 2^-24, so
  // each refinement step gains fewer digits and more steps may be needed
  const int max_solves = precision == DGAMI_MIXED_BF16X3_IR ? MIXED_MAX_SOLVES_BF16 : MIXED_MAX_SOLVES_FP32;
  const double tol = mixed_tol();
  const int n = a.n, T = a.T;
  DGAMI_ARG_CHECK(T <= 16, "ladder: at most 4 orbitals in the mixed-precision form");
  DGAMI_ARG_CHECK(n <= BLK_MAX_N, "ladder: compound size above 1536 is not supported");
  // the factorisation carries the nb^2 indicator columns, so the first solve needs only the
  // backward sweep
  // rows start on 128-byte boundaries (16 complex64 values), so a row's 32-column block segment
  // spans two cache lines, not three
  const int ncol = n + T, ld = (ncol + 15) & ~15;
  const size_t items = static_cast<size_t>(a.nchan) * a.nw * a.nq;
  const int gq = a.dense ? 1 : a.nq;
  const size_t groups = items / gq;
  const int nblk = static_cast<int>(ceil_div(n, TRI_B));
  // per item: FP32 factors, their inverse diagonal blocks, FP64 X and R, FP32 sweep vectors, permutation
  const size_t per_item = static_cast<size_t>(n) * ld * sizeof(float2) +
                          static_cast<size_t>(nblk) * 2 * TRI_B * TRI_B * sizeof(float2) +
                          2 * static_cast<size_t>(n) * T * sizeof(double2) + 2 * static_cast<size_t>(n) * T * sizeof(float2) +
                          static_cast<size_t>(n) * sizeof(int);
  // lanes (ladder_lanes): the (channel, w) groups split over the caller's and the extra streams
  const int nl = static_cast<int>(std::min<size_t>(ladder_lanes(), groups));
  size_t gb = (size_t(48) << 30) / (per_item * gq * nl);
  gb = std::max<size_t>(1, std::min(gb, (groups + nl - 1) / nl));
  gb = std::min<size_t>(gb, 65535 / static_cast<size_t>(gq) > 0 ? 65535 / gq : 1);
  const size_t bitems = gb * gq;
  DGAMI_ARG_CHECK(bitems <= 65535, "ladder: too many q-points in one (channel, w) group");
  const int tuned = ladder_tuned_kernel(n, T, precision);

  // one lane per stream: factors, refinement vectors and sweep buffers in its own workspace slots
  struct Lane {
    BlockedState<float2> st;
    MixedArgs m;
    float2* tri;
    float2* yv;
    float2* xv;
    int* perm;  // the permutation the perm kernel writes (m.perm reads it)
    int* pinv;  // and its inverse (m.pinv)
    hipStream_t s;
  };
  auto setup = [&](Lane& L, int off, hipStream_t s) -> int {
    if (!blocked_alloc(L.st, n, ld, ncol, bitems, WS_MIXED + off)) return -2;
    L.st.swap_left = 1;
    L.st.gemm_mode = precision == DGAMI_MIXED_BF16X3_IR ? BLK_GEMM_BF16X3 : BLK_GEMM_F32_MFMA;
    if (tuned >= 0 && ladder_kernel_family(tuned) == LADDER_FAM_BLOCKED) L.st.variant = ladder_kernel_variant(tuned);
    char* vec = static_cast<char*>(workspace(WS_MIXED_VEC + off, 2 * bitems * n * T * sizeof(double2) +
                                                                    2 * bitems * n * sizeof(int) + 2 * bitems * sizeof(int) +
                                                                    bitems * sizeof(float) + gb * sizeof(int)));
    if (!vec) return -2;
    MixedArgs& m = L.m;
    m.a = a;
    m.W = L.st.W;
    m.X = reinterpret_cast<double2*>(vec);
    m.R = m.X + bitems * n * T;
    int* perm = reinterpret_cast<int*>(m.R + bitems * n * T);
    m.perm = perm;
    L.perm = perm;
    L.pinv = perm + bitems * n;
    m.pinv = L.pinv;
    m.status = L.pinv + bitems * n;
    m.gactive = m.status + bitems;
    m.prev = reinterpret_cast<float*>(m.gactive + gb);
    m.dinv = L.st.dinv;
    m.ld = ld;
    m.gq = gq;
    L.tri = static_cast<float2*>(workspace(WS_TRI + off, bitems * nblk * 2 * TRI_B * TRI_B * sizeof(float2)));
    L.yv = static_cast<float2*>(workspace(WS_TRI_VEC + off, 2 * bitems * n * T * sizeof(float2)));
    if (!L.tri || !L.yv) return -2;
    L.xv = L.yv + bitems * n * T;
    L.s = s;
    return 0;
  };
  auto run = [&](Lane& L, size_t g0) -> int {
    hipStream_t s = L.s;
    MixedArgs& m = L.m;
    BlockedState<float2>& st = L.st;
    const TriVec<float2> Yv{L.yv, static_cast<size_t>(n) * T, T}, Xv{L.xv, static_cast<size_t>(n) * T, T};
    const int ng = static_cast<int>(std::min(gb, groups - g0));
    const int nit = ng * gq;
    const int item0 = static_cast<int>(g0 * gq);
    m.group0 = static_cast<int>(g0);
    m.item0 = item0;
    int rc = blocked_factor<float2>(a, st, item0, nit, s);
    if (rc) return rc;
    hipLaunchKernelGGL(mix_perm_kernel, dim3(nit), dim3(64), n * sizeof(int), s, st.piv, L.perm, L.pinv, n, m.status,
                       m.prev);
    DGAMI_HIP_CHECK(hipGetLastError());
    TriFactors<float2> f{st.W, L.tri, st.dinv, n, ld, nblk};
    rc = tri_diag_inverse(f, nit, true, s);
    if (rc) return rc;
    f.skip = m.status;  // the sweeps leave converged items alone
    for (int sv = 0; sv < max_solves; ++sv) {
      if (sv > 0) {
        hipLaunchKernelGGL(mix_group_active_kernel, dim3(ceil_div(ng, 256)), dim3(256), 0, s, m.status, m.gactive, gq,
                           ng);
        DGAMI_HIP_CHECK(hipGetLastError());
        ZgemmArgs g;
        g.A = a.gamma + static_cast<size_t>(g0) * n * n;
        g.B = m.X;
        g.C = m.R;
        g.M = n;
        g.N = gq * T;
        g.K = n;
        g.lda = n;
        g.ldb = gq * T;
        g.ldc = gq * T;
        g.sa = static_cast<long long>(n) * n;
        g.sb = static_cast<long long>(n) * gq * T;
        g.sc = g.sb;
        g.alpha = make_double2(1.0, 0.0);
        g.beta = make_double2(0.0, 0.0);
        g.active = m.gactive;
        rc = launch_zgemm(g, ng, s);
        if (rc) return rc;
        // Y = P R into yv; xv, free until the forward sweep, holds the U_r/V_r term
        rc = mix_residual(m, ng, nit, reinterpret_cast<double2*>(L.xv), L.yv, s);
        if (rc) return rc;
      }
      // D = U^-1 L^-1 P R in complex64, X += D; the first R is B, whose L^-1 P B the
      // factorisation left in columns [n, n + T)
      if (sv == 0) {
        rc = tri_backward(f, nit, T, TriVec<float2>{st.W + n, static_cast<size_t>(n) * ld, ld}, Yv, nullptr, 0, s);
        if (rc) return rc;
      } else {
        rc = tri_forward(f, nit, T, Yv, Xv, s);
        if (rc) return rc;
        rc = tri_backward(f, nit, T, Xv, Yv, nullptr, 0, s);
        if (rc) return rc;
      }
      hipLaunchKernelGGL(mix_accumulate_kernel, dim3(nit), dim3(256), 0, s, m, L.yv, sv, sv == max_solves - 1 ? 1 : 0,
                         tol);
      DGAMI_HIP_CHECK(hipGetLastError());
    }
    if (a.dense) {
      const size_t count = static_cast<size_t>(nit) * n * T;
      hipLaunchKernelGGL(mix_output_dense_kernel, dim3(ceil_div(count, 256)), dim3(256), 0, s, m, count);
    } else {
      hipLaunchKernelGGL(mix_output_kernel, dim3(n, ng), dim3(256), 0, s, m);
    }
    DGAMI_HIP_CHECK(hipGetLastError());
    if (a.info) {
      hipLaunchKernelGGL(mix_info_kernel, dim3(ceil_div(nit, 256)), dim3(256), 0, s, st.info, m.status, a.info, item0,
                         nit);
      DGAMI_HIP_CHECK(hipGetLastError());
    }
    return 0;
  };

  Lane lanes[1 + MAX_EXTRA_STREAMS];
  int rc = setup(lanes[0], 0, stream);
  if (rc) return rc;
  for (int l = 1; l < nl; ++l) {
    hipStream_t sx = nullptr;
    if ((rc = extra_stream_fork(stream, l, &sx)) != 0) return rc;
    if ((rc = setup(lanes[l], l * WS_STREAM_STRIDE, sx)) != 0) return rc;
  }
  int k = 0;
  for (size_t g0 = 0; g0 < groups; g0 += gb, k = (k + 1) % nl)
    if ((rc = run(lanes[k], g0)) != 0) return rc;
  for (int l = 1; l < nl; ++l)
    if ((rc = extra_stream_join(stream, l)) != 0) return rc;
  return 0;
}

}  // namespace dgami
