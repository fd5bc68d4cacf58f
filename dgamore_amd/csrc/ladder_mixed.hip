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
// Starting from X = 0 (R = B), solves continue until the predicted relative error is below
// 1e-10 (see MIXED_TOL). Each refinement step multiplies the error by about cond(A) * 2^-24
// (FP32 factors) or cond(A) * 2^-16 (split-BF16 updates); the number of solves is data
// dependent, so the host reads one float per solve.
#include <algorithm>

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

// Item layout of the refinement vectors: group g = (chan, w) (or the item itself in
// dense mode) holds gq items; X and R are [group][n][gq][T] so that Gamma(g) X(g) is
// one GEMM with N = gq * T.
struct MixedArgs {
  LadderArgs a;
  const float2* W;      // factors [item][n][ld]
  const int* perm;      // [item][n] row i of P r is r[perm[i]]
  const double2* dinv;  // [item][n] inverse pivots
  double2* X;           // [group][n][gq][T]
  double2* R;           // same
  int ld, gq, group0, item0;
};

// ---------------------------------------------------------------------------
// Row permutation of each factorisation from its sequence of interchanges.
__global__ void __launch_bounds__(64) mix_perm_kernel(const int* __restrict__ piv, int* __restrict__ perm, int n) {
  extern __shared__ int s_perm[];
  const int b = blockIdx.x;
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
  for (int i = threadIdx.x; i < n; i += 64) perm[static_cast<size_t>(b) * n + i] = s_perm[i];
}

// ---------------------------------------------------------------------------
// R = B - A X given R = Gamma X on entry (vertex term of the GEMM). One workgroup per item.
__global__ void __launch_bounds__(256) mix_residual_kernel(MixedArgs m) {
  const LadderArgs& a = m.a;
  const int lb = blockIdx.x;                  // item within the batch
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

// X += D (FP64) and out = X / beta; per item, the largest |D| and |X| (|Re|+|Im|) go to
// norms[2*item], norms[2*item+1]. One workgroup per item.
__global__ void __launch_bounds__(256) mix_accumulate_kernel(MixedArgs m, const float2* __restrict__ D, int first,
                                                             float* __restrict__ norms) {
  const LadderArgs& a = m.a;
  const int lb = blockIdx.x;
  const int gl = lb / m.gq, q = lb - gl * m.gq;
  const int n = a.n, T = a.T, gq = m.gq;
  const int item = m.item0 + lb;
  const int qt = item % a.nq, w = (item / a.nq) % a.nw, chan = item / (a.nq * a.nw);
  double2* X = m.X + static_cast<size_t>(gl) * n * gq * T;
  const float2* Db = D + static_cast<size_t>(lb) * n * T;
  float dmax = 0.f, xmax = 0.f;
  for (int idx = threadIdx.x; idx < n * T; idx += 256) {
    const int i = idx / T, t = idx - (idx / T) * T;
    const size_t o = (static_cast<size_t>(i) * gq + q) * T + t;
    const double2 d = to_c128(Db[idx]);
    const double2 x = first ? d : cadd(X[o], d);
    X[o] = x;
    a.out[ladder_out_index(a, chan, w, qt, i, t)] = cscale(x, a.inv_beta);
    dmax = fmaxf(dmax, static_cast<float>(cabs1(d)));
    xmax = fmaxf(xmax, static_cast<float>(cabs1(x)));
  }
  __shared__ float s_d[4], s_x[4];
  const unsigned dk = wave_max_u32(__float_as_uint(dmax)), xk = wave_max_u32(__float_as_uint(xmax));
  if ((threadIdx.x & 63) == 0) {
    s_d[threadIdx.x >> 6] = __uint_as_float(dk);
    s_x[threadIdx.x >> 6] = __uint_as_float(xk);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    norms[2 * static_cast<size_t>(lb)] = fmaxf(fmaxf(s_d[0], s_d[1]), fmaxf(s_d[2], s_d[3]));
    norms[2 * static_cast<size_t>(lb) + 1] = fmaxf(fmaxf(s_x[0], s_x[1]), fmaxf(s_x[2], s_x[3]));
  }
}

// Largest relative correction max_item(|D|_max / |X|_max) of the batch into *ratio.
__global__ void __launch_bounds__(256) mix_ratio_kernel(const float* __restrict__ norms, int nitems,
                                                        float* __restrict__ ratio) {
  float r = 0.f;
  for (int b = threadIdx.x; b < nitems; b += 256) {
    const float d = norms[2 * b], x = norms[2 * b + 1];
    r = fmaxf(r, x > 0.f ? d / x : (d > 0.f ? 1.f : 0.f));
  }
  __shared__ float s_r[4];
  const unsigned k = wave_max_u32(__float_as_uint(r));
  if ((threadIdx.x & 63) == 0) s_r[threadIdx.x >> 6] = __uint_as_float(k);
  __syncthreads();
  if (threadIdx.x == 0) *ratio = fmaxf(fmaxf(s_r[0], s_r[1]), fmaxf(s_r[2], s_r[3]));
}

// ---------------------------------------------------------------------------
int ladder_solve_mixed(const LadderArgs& a, int precision, hipStream_t stream) {
  DGAMI_ARG_CHECK(precision == DGAMI_MIXED_FP32_IR || precision == DGAMI_MIXED_BF16X3_IR,
                  "ladder: unknown mixed-precision mode");
  // the split-BF16 factors carry ~2^-16 relative error per product instead of 2^-24, so
  // each refinement step gains fewer digits and more steps may be needed
  const int max_solves = precision == DGAMI_MIXED_BF16X3_IR ? MIXED_MAX_SOLVES_BF16 : MIXED_MAX_SOLVES_FP32;
  const int n = a.n, T = a.T;
  DGAMI_ARG_CHECK(T <= 16, "ladder: at most 4 orbitals in the mixed-precision form");
  DGAMI_ARG_CHECK(n <= BLK_MAX_N, "ladder: compound size above 1536 is not supported");
  // the factorisation carries the nb^2 indicator columns, so the first solve needs only the
  // backward sweep
  const int ncol = n + T, ld = (ncol + 7) & ~7;
  const size_t items = static_cast<size_t>(a.nchan) * a.nw * a.nq;
  const int gq = a.dense ? 1 : a.nq;
  const size_t groups = items / gq;
  const int nblk = static_cast<int>(ceil_div(n, TRI_B));
  // per item: FP32 factors, their inverse diagonal blocks, FP64 X and R, FP32 sweep vectors, permutation
  const size_t per_item = static_cast<size_t>(n) * ld * sizeof(float2) +
                          static_cast<size_t>(nblk) * 2 * TRI_B * TRI_B * sizeof(float2) +
                          2 * static_cast<size_t>(n) * T * sizeof(double2) + 2 * static_cast<size_t>(n) * T * sizeof(float2) +
                          static_cast<size_t>(n) * sizeof(int);
  size_t gb = (size_t(48) << 30) / (per_item * gq);
  gb = std::max<size_t>(1, std::min(gb, groups));
  gb = std::min<size_t>(gb, 65535 / static_cast<size_t>(gq) > 0 ? 65535 / gq : 1);
  const size_t bitems = gb * gq;
  DGAMI_ARG_CHECK(bitems <= 65535, "ladder: too many q-points in one (channel, w) group");

  BlockedState<float2> st;
  if (!blocked_alloc(st, n, ld, ncol, bitems, WS_MIXED)) return -2;
  st.swap_left = 1;
  st.gemm_mode = precision == DGAMI_MIXED_BF16X3_IR ? BLK_GEMM_BF16X3 : BLK_GEMM_F32_MFMA;
  const int tuned = ladder_tuned_kernel(n, T, precision);
  if (tuned >= 0 && ladder_kernel_family(tuned) == LADDER_FAM_BLOCKED) st.variant = ladder_kernel_variant(tuned);
  char* vec = static_cast<char*>(workspace(WS_MIXED_VEC, 2 * bitems * n * T * sizeof(double2) + bitems * n * sizeof(int) +
                                                             (2 * bitems + 4) * sizeof(float)));
  if (!vec) return -2;
  MixedArgs m;
  m.a = a;
  m.W = st.W;
  m.X = reinterpret_cast<double2*>(vec);
  m.R = m.X + bitems * n * T;
  int* perm = reinterpret_cast<int*>(m.R + bitems * n * T);
  m.perm = perm;
  float* norms = reinterpret_cast<float*>(perm + bitems * n);  // [bitems][2] + the batch ratio
  m.dinv = st.dinv;
  m.ld = ld;
  m.gq = gq;
  float2* tri = static_cast<float2*>(workspace(WS_TRI, bitems * nblk * 2 * TRI_B * TRI_B * sizeof(float2)));
  float2* yv = static_cast<float2*>(workspace(WS_TRI_VEC, 2 * bitems * n * T * sizeof(float2)));
  if (!tri || !yv) return -2;
  float2* xv = yv + bitems * n * T;
  const TriVec<float2> Yv{yv, static_cast<size_t>(n) * T, T}, Xv{xv, static_cast<size_t>(n) * T, T};

  for (size_t g0 = 0; g0 < groups; g0 += gb) {
    const int ng = static_cast<int>(std::min(gb, groups - g0));
    const int nit = ng * gq;
    const int item0 = static_cast<int>(g0 * gq);
    m.group0 = static_cast<int>(g0);
    m.item0 = item0;
    int rc = blocked_factor<float2>(a, st, item0, nit, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(mix_perm_kernel, dim3(nit), dim3(64), n * sizeof(int), stream, st.piv, perm, n);
    DGAMI_HIP_CHECK(hipGetLastError());
    const TriFactors<float2> f{st.W, tri, st.dinv, n, ld, nblk};
    rc = tri_diag_inverse(f, nit, true, stream);
    if (rc) return rc;
    const dim3 vgrid(ceil_div(static_cast<size_t>(n) * T, 256), nit);
    double prev = 0.0;
    for (int s = 0; s < max_solves; ++s) {
      if (s > 0) {
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
        rc = launch_zgemm(g, ng, stream);
        if (rc) return rc;
        hipLaunchKernelGGL(mix_residual_kernel, dim3(nit), dim3(256), 0, stream, m);
        DGAMI_HIP_CHECK(hipGetLastError());
      }
      // D = U^-1 L^-1 P R in complex64, X += D; the first R is B, whose L^-1 P B the
      // factorisation left in columns [n, n + T)
      if (s == 0) {
        rc = tri_backward(f, nit, T, TriVec<float2>{st.W + n, static_cast<size_t>(n) * ld, ld}, Yv, nullptr, 0, stream);
        if (rc) return rc;
      } else {
        hipLaunchKernelGGL(mix_gather_kernel, vgrid, dim3(256), 0, stream, m, yv, 0);
        DGAMI_HIP_CHECK(hipGetLastError());
        rc = tri_forward(f, nit, T, Yv, Xv, stream);
        if (rc) return rc;
        rc = tri_backward(f, nit, T, Xv, Yv, nullptr, 0, stream);
        if (rc) return rc;
      }
      hipLaunchKernelGGL(mix_accumulate_kernel, dim3(nit), dim3(256), 0, stream, m, yv, s == 0 ? 1 : 0, norms);
      DGAMI_HIP_CHECK(hipGetLastError());
      if (s == 0) continue;
      hipLaunchKernelGGL(mix_ratio_kernel, dim3(1), dim3(256), 0, stream, norms, nit, norms + 2 * bitems);
      DGAMI_HIP_CHECK(hipGetLastError());
      float ratio = 0.f;
      DGAMI_HIP_CHECK(hipMemcpyAsync(&ratio, norms + 2 * bitems, sizeof(float), hipMemcpyDeviceToHost, stream));
      DGAMI_HIP_CHECK(hipStreamSynchronize(stream));
      const double r = ratio;
      const double contraction = (s >= 2 && prev > 0.0) ? r / prev : r;
      prev = r;
      if (r * contraction < MIXED_TOL || !(r == r)) break;
    }
    rc = blocked_copy_info(st.info, a.info, item0, nit, stream);
    if (rc) return rc;
  }
  return 0;
}

}  // namespace dgami
