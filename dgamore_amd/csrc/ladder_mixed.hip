// Mixed-precision ladder solves: complex64 LU factorisation, FP64 iterative refinement.
//
// Every item's matrix A is factored once in FP32 (PA = LU, batched blocked LU of
// ladder_blocked.hip with the interchanges also applied left of each panel). The
// solution X of A X = B (B = indicator right-hand sides, the nu' sum of A^-1)
// is then refined in FP64:
//   R = B - A X      in FP64 from the FP64 operands, with the vertex term Gamma(chan,w) X
//                    formed for all q of one (chan, w) as a single complex128 GEMM
//                    (Gamma is shared by the q's; the chi0^-1 blocks and the rank-nb^2
//                    U_r + V_r term are applied per item);
//   D = (LU)^-1 P R  in FP32 (triangular solves on LDS-resident right-hand sides);
//   X += D           in FP64.
// Starting from X = 0 (R = B), MIXED_SOLVES solves are made. Each refinement step
// multiplies the error by about cond(A) * 2^-24, so two refinements bring the result
// to FP64 accuracy for the condition numbers of ladder matrices.
#include <algorithm>

#include "dgami.h"
#include "dgami_internal.hpp"
#include "dgami_launch.hpp"
#include "ladder_blocked.hpp"
#include "ladder_common.hpp"
#include "zgemm.hpp"

namespace dgami {

constexpr int MIXED_SOLVES = 3;
constexpr int TRI_CB = 32;        // rows per triangular block of the substitutions
constexpr int SOLVE_THREADS = 256;

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
__device__ __forceinline__ float readlane_f(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// D = U^-1 L^-1 P R in FP32 with R[item] staged in LDS; X += D. On the last solve the
// result is also written to the output (divided by beta). One workgroup per item.
template <int T>
__global__ void __launch_bounds__(SOLVE_THREADS) mix_solve_kernel(MixedArgs m, int first, int last) {
  extern __shared__ float2 s_y[];  // [n][T]
  const LadderArgs& a = m.a;
  const int lb = blockIdx.x;
  const int gl = lb / m.gq, q = lb - gl * m.gq;
  const int item = m.item0 + lb;
  const int n = a.n, gq = m.gq, ld = m.ld;
  const size_t gbase = static_cast<size_t>(gl) * n * gq * T;
  double2* X = m.X + gbase;
  const double2* R = m.R + gbase;
  const float2* W = m.W + static_cast<size_t>(lb) * n * ld;
  const int* perm = m.perm + static_cast<size_t>(lb) * n;
  const double2* dinv = m.dinv + static_cast<size_t>(lb) * n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  for (int idx = tid; idx < n * T; idx += SOLVE_THREADS) {
    const int i = idx / T, t = idx - (idx / T) * T;
    const int src = perm[i];
    s_y[idx] = first ? from_c128<float2>(ladder_rhs(a, src, t))
                     : from_c128<float2>(R[(static_cast<size_t>(src) * gq + q) * T + t]);
  }
  __syncthreads();

  // forward substitution with the unit lower factor
  for (int c0 = 0; c0 < n; c0 += TRI_CB) {
    const int cb = min(TRI_CB, n - c0);
    if (wave == 0) {
      const int row = c0 + lane;
      const bool live = lane < cb;
      float2 lrow[TRI_CB];
#pragma unroll
      for (int s = 0; s < TRI_CB; ++s)
        lrow[s] = (live && s < lane) ? W[static_cast<size_t>(row) * ld + c0 + s] : make_float2(0.f, 0.f);
#pragma unroll
      for (int t = 0; t < T; ++t) {
        float2 y = live ? s_y[row * T + t] : make_float2(0.f, 0.f);
#pragma unroll
        for (int s = 0; s < TRI_CB; ++s) {
          const float2 xs = make_float2(readlane_f(y.x, s), readlane_f(y.y, s));
          if (lane > s) y = cfms(y, lrow[s], xs);
        }
        if (live) s_y[row * T + t] = y;
      }
    }
    __syncthreads();
    for (int i = c0 + cb + tid; i < n; i += SOLVE_THREADS) {
      float2 acc[T];
#pragma unroll
      for (int t = 0; t < T; ++t) acc[t] = s_y[i * T + t];
      const float2* lr = W + static_cast<size_t>(i) * ld + c0;
      for (int s = 0; s < cb; ++s) {
        const float2 l = lr[s];
#pragma unroll
        for (int t = 0; t < T; ++t) acc[t] = cfms(acc[t], l, s_y[(c0 + s) * T + t]);
      }
#pragma unroll
      for (int t = 0; t < T; ++t) s_y[i * T + t] = acc[t];
    }
    __syncthreads();
  }

  // back substitution with the upper factor
  for (int c1 = n; c1 > 0; c1 -= TRI_CB) {
    const int c0 = c1 > TRI_CB ? c1 - TRI_CB : 0;
    const int cb = c1 - c0;
    if (wave == 0) {
      const int row = c0 + lane;
      const bool live = lane < cb;
      float2 urow[TRI_CB];
#pragma unroll
      for (int s = 0; s < TRI_CB; ++s)
        urow[s] = (live && s > lane && s < cb) ? W[static_cast<size_t>(row) * ld + c0 + s] : make_float2(0.f, 0.f);
      const float2 di = live ? from_c128<float2>(dinv[row]) : make_float2(0.f, 0.f);
#pragma unroll
      for (int t = 0; t < T; ++t) {
        float2 y = live ? s_y[row * T + t] : make_float2(0.f, 0.f);
#pragma unroll
        for (int s = TRI_CB - 1; s >= 0; --s) {
          if (lane == s) y = cmul(y, di);
          const float2 xs = make_float2(readlane_f(y.x, s), readlane_f(y.y, s));
          if (lane < s) y = cfms(y, urow[s], xs);
        }
        if (live) s_y[row * T + t] = y;
      }
    }
    __syncthreads();
    for (int i = tid; i < c0; i += SOLVE_THREADS) {
      float2 acc[T];
#pragma unroll
      for (int t = 0; t < T; ++t) acc[t] = s_y[i * T + t];
      const float2* ur = W + static_cast<size_t>(i) * ld + c0;
      for (int s = 0; s < cb; ++s) {
        const float2 u = ur[s];
#pragma unroll
        for (int t = 0; t < T; ++t) acc[t] = cfms(acc[t], u, s_y[(c0 + s) * T + t]);
      }
#pragma unroll
      for (int t = 0; t < T; ++t) s_y[i * T + t] = acc[t];
    }
    __syncthreads();
  }

  const int qt = item % a.nq, w = (item / a.nq) % a.nw, chan = item / (a.nq * a.nw);
  for (int idx = tid; idx < n * T; idx += SOLVE_THREADS) {
    const int i = idx / T, t = idx - (idx / T) * T;
    const size_t o = (static_cast<size_t>(i) * gq + q) * T + t;
    const double2 x = first ? to_c128(s_y[idx]) : cadd(X[o], to_c128(s_y[idx]));
    X[o] = x;
    if (last) a.out[ladder_out_index(a, chan, w, qt, i, t)] = cscale(x, a.inv_beta);
  }
}

__global__ void mix_info_kernel(const int* __restrict__ st_info, int32_t* __restrict__ info, int item0, int nitems) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nitems) info[item0 + b] = st_info[b];
}

// ---------------------------------------------------------------------------
int ladder_solve_mixed(const LadderArgs& a, int precision, hipStream_t stream) {
  DGAMI_ARG_CHECK(precision == DGAMI_MIXED_FP32_IR,
                  "ladder: the split-BF16 mixed mode is not available in this build; use DGAMI_MIXED_FP32_IR");
  const int n = a.n, T = a.T;
  DGAMI_ARG_CHECK(T <= 16, "ladder: at most 4 orbitals in the mixed-precision form");
  DGAMI_ARG_CHECK(n <= 1536, "ladder: compound size above 1536 is not supported");
  const size_t lds = static_cast<size_t>(n) * T * sizeof(float2);
  DGAMI_ARG_CHECK(lds <= 128 * 1024, "ladder: compound size times right-hand sides too large for the LDS solve");
  const int ld = (n + 7) & ~7;
  const size_t items = static_cast<size_t>(a.nchan) * a.nw * a.nq;
  const int gq = a.dense ? 1 : a.nq;
  const size_t groups = items / gq;
  // per item: FP32 factors + FP64 X and R + permutation
  const size_t per_item = static_cast<size_t>(n) * ld * sizeof(float2) + 2 * static_cast<size_t>(n) * T * sizeof(double2) +
                          static_cast<size_t>(n) * sizeof(int);
  size_t gb = (size_t(48) << 30) / (per_item * gq);
  gb = std::max<size_t>(1, std::min(gb, groups));
  gb = std::min<size_t>(gb, 65535 / static_cast<size_t>(gq) > 0 ? 65535 / gq : 1);
  const size_t bitems = gb * gq;
  DGAMI_ARG_CHECK(bitems <= 65535, "ladder: too many q-points in one (channel, w) group");

  BlockedState<float2> st;
  if (!blocked_alloc(st, n, ld, n, bitems, WS_MIXED)) return -2;
  st.swap_left = 1;
  char* vec = static_cast<char*>(workspace(WS_MIXED_VEC, 2 * bitems * n * T * sizeof(double2) + bitems * n * sizeof(int)));
  if (!vec) return -2;
  MixedArgs m;
  m.a = a;
  m.W = st.W;
  m.X = reinterpret_cast<double2*>(vec);
  m.R = m.X + bitems * n * T;
  int* perm = reinterpret_cast<int*>(m.R + bitems * n * T);
  m.perm = perm;
  m.dinv = st.dinv;
  m.ld = ld;
  m.gq = gq;
  void (*solve)(MixedArgs, int, int) = nullptr;
  switch (T) {
    case 1: solve = mix_solve_kernel<1>; break;
    case 4: solve = mix_solve_kernel<4>; break;
    case 9: solve = mix_solve_kernel<9>; break;
    case 16: solve = mix_solve_kernel<16>; break;
    default: DGAMI_ARG_CHECK(false, "ladder: nb^2 must be 1, 4, 9 or 16");
  }
  allow_lds(solve, lds);

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
    for (int s = 0; s < MIXED_SOLVES; ++s) {
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
      hipLaunchKernelGGL(solve, dim3(nit), dim3(SOLVE_THREADS), lds, stream, m, s == 0 ? 1 : 0,
                         s == MIXED_SOLVES - 1 ? 1 : 0);
      DGAMI_HIP_CHECK(hipGetLastError());
    }
    if (a.info) {
      hipLaunchKernelGGL(mix_info_kernel, dim3(ceil_div(nit, 256)), dim3(256), 0, stream, st.info, a.info, item0, nit);
      DGAMI_HIP_CHECK(hipGetLastError());
    }
  }
  return 0;
}

}  // namespace dgami
