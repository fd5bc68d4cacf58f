// Batched blocked triangular solves (see tri_solve.hpp for the scheme).
#include <type_traits>

#include "dgami.h"
#include "dgami_internal.hpp"
#include "dgami_launch.hpp"
#include "tri_solve.hpp"

namespace dgami {

template <typename S>
__device__ __forceinline__ S cone();
template <>
__device__ __forceinline__ double2 cone<double2>() {
  return make_double2(1.0, 0.0);
}
template <>
__device__ __forceinline__ float2 cone<float2>() {
  return make_float2(1.0f, 0.0f);
}

// ---------------------------------------------------------------------------
// grid (nblk, batch), one wavefront: lanes 0-31 form column j of L_bb^-1, lanes 32-63
// column j of U_bb^-1 (padding rows of the last block act as identity).
template <typename S>
__global__ void __launch_bounds__(64) tri_diag_inverse_kernel(TriFactors<S> f, int need_lower) {
  const int kb = blockIdx.x, b = blockIdx.y;
  const int c0 = kb * TRI_B;
  const int cb = min(TRI_B, f.n - c0);
  const int tid = threadIdx.x;
  const S* W = f.W + static_cast<size_t>(b) * f.n * f.ld;
  __shared__ S s[TRI_B][TRI_B + 1];
  __shared__ S s_d[TRI_B];
  for (int e = tid; e < TRI_B * TRI_B; e += 64) {
    const int r = e / TRI_B, c = e - (e / TRI_B) * TRI_B;
    s[r][c] = (r < cb && c < cb) ? W[static_cast<size_t>(c0 + r) * f.ld + c0 + c] : (r == c ? cone<S>() : czero<S>());
  }
  for (int e = tid; e < TRI_B; e += 64)
    s_d[e] = e < cb ? from_c128<S>(f.dinv[static_cast<size_t>(b) * f.n + c0 + e]) : cone<S>();
  __syncthreads();
  S* out = f.dinvblk + (static_cast<size_t>(b) * f.nblk + kb) * 2 * TRI_B * TRI_B;
  const int j = tid & 31;
  S x[TRI_B];
  if (tid < 32) {
    if (!need_lower) return;
    // L x = e_j, L unit lower
#pragma unroll
    for (int i = 0; i < TRI_B; ++i) {
      S acc = czero<S>();
#pragma unroll
      for (int k = 0; k < i; ++k) acc = cfms(acc, s[i][k], x[k]);
      x[i] = (i == j) ? cone<S>() : (i > j ? acc : czero<S>());
    }
#pragma unroll
    for (int i = 0; i < TRI_B; ++i) out[i * TRI_B + j] = x[i];
  } else {
    // U x = e_j
#pragma unroll
    for (int i = TRI_B - 1; i >= 0; --i) {
      S acc = czero<S>();
#pragma unroll
      for (int k = i + 1; k < TRI_B; ++k) acc = cfma(acc, s[i][k], x[k]);
      x[i] = (i > j) ? czero<S>() : cmul(csub(i == j ? cone<S>() : czero<S>(), acc), s_d[i]);
    }
#pragma unroll
    for (int i = 0; i < TRI_B; ++i) out[TRI_B * TRI_B + i * TRI_B + j] = x[i];
  }
}

// Two consecutive factor elements (the first at an even column of a block column: 16-byte
// aligned for complex64, since ld is a multiple of 4).
template <typename S>
__device__ __forceinline__ void load_pair(const S* src, S& a, S& b);
template <>
__device__ __forceinline__ void load_pair<float2>(const float2* src, float2& a, float2& b) {
  const float4 x = *reinterpret_cast<const float4*>(src);
  a = make_float2(x.x, x.y);
  b = make_float2(x.z, x.w);
}
template <>
__device__ __forceinline__ void load_pair<double2>(const double2* src, double2& a, double2& b) {
  a = src[0];
  b = src[1];
}

// ---------------------------------------------------------------------------
// One whole sweep per problem: grid (batch), 256 threads. Thread t keeps rows t + 256 m
// (m < RPT) of the right-hand sides in registers. For every block b (forward: first to
// last, backward: last to first): the owners stage y_b in LDS, x_b = D_b^-1 y_b is formed
// in LDS, the owners of block b keep x_b, and every thread updates its rows below (L) or
// above (U) block b with its own row of the factor's block column.
template <typename S, int T, bool LOWER, bool TO_LADDER, int RPT>
__device__ __forceinline__ void tri_sweep_body(const TriFactors<S>& f, const TriVec<S>& y, const TriVec<S>& out,
                                               const LadderArgs& la, int item0) {
  const int b = blockIdx.x;
  if (f.skip && f.skip[b] != 0) return;
  const int tid = threadIdx.x;
  const int n = f.n;
  __shared__ S s_d[TRI_B][TRI_B + 1];
  __shared__ S s_y[TRI_B][T];
  __shared__ S s_x[TRI_B][T];
  const S* F = f.W + static_cast<size_t>(b) * n * f.ld;
  const S* Y = y.p + static_cast<size_t>(b) * y.item_stride;
  S yr[RPT][T];
#pragma unroll
  for (int m = 0; m < RPT; ++m) {
    const int r = tid + 256 * m;
#pragma unroll
    for (int t = 0; t < T; ++t) yr[m][t] = r < n ? Y[static_cast<size_t>(r) * y.row_stride + t] : czero<S>();
  }
  for (int s = 0; s < f.nblk; ++s) {
    const int kb = LOWER ? s : f.nblk - 1 - s;
    const int c0 = kb * TRI_B;
    const int cb = min(TRI_B, n - c0);
    const S* Db = f.dinvblk + (static_cast<size_t>(b) * f.nblk + kb) * 2 * TRI_B * TRI_B + (LOWER ? 0 : TRI_B * TRI_B);
    for (int e = tid; e < TRI_B * TRI_B; e += 256) s_d[e / TRI_B][e % TRI_B] = Db[e];
#pragma unroll
    for (int m = 0; m < RPT; ++m) {
      const int r = tid + 256 * m;
      if (r >= c0 && r < c0 + cb) {
#pragma unroll
        for (int t = 0; t < T; ++t) s_y[r - c0][t] = yr[m][t];
      }
    }
    if (tid < TRI_B && tid >= cb) {
#pragma unroll
      for (int t = 0; t < T; ++t) s_y[tid][t] = czero<S>();
    }
    __syncthreads();
    for (int e = tid; e < TRI_B * T; e += 256) {
      const int i = e / T, t = e - (e / T) * T;
      S acc = czero<S>();
#pragma unroll 8
      for (int k = 0; k < TRI_B; ++k) acc = cfma(acc, s_d[i][k], s_y[k][t]);
      s_x[i][t] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < RPT; ++m) {
      const int r = tid + 256 * m;
      if (r >= c0 && r < c0 + cb) {
#pragma unroll
        for (int t = 0; t < T; ++t) yr[m][t] = s_x[r - c0][t];
      } else if (r < n && (LOWER ? r >= c0 + cb : r < c0)) {
        const S* frow = F + static_cast<size_t>(r) * f.ld + c0;
        if (cb == TRI_B) {
          // full block: unrolled in pairs of 16-byte-aligned elements so that loads run ahead
#pragma unroll 4
          for (int k = 0; k < TRI_B; k += 2) {
            S l0, l1;
            load_pair<S>(frow + k, l0, l1);
#pragma unroll
            for (int t = 0; t < T; ++t) {
              yr[m][t] = cfms(yr[m][t], l0, s_x[k][t]);
              yr[m][t] = cfms(yr[m][t], l1, s_x[k + 1][t]);
            }
          }
        } else {
          for (int k = 0; k < cb; ++k) {
            const S l = frow[k];
#pragma unroll
            for (int t = 0; t < T; ++t) yr[m][t] = cfms(yr[m][t], l, s_x[k][t]);
          }
        }
      }
    }
  }
#pragma unroll
  for (int m = 0; m < RPT; ++m) {
    const int r = tid + 256 * m;
    if (r >= n) continue;
#pragma unroll
    for (int t = 0; t < T; ++t) {
      if constexpr (TO_LADDER) {
        const int item = item0 + b;
        const int q = item % la.nq, w = (item / la.nq) % la.nw, chan = item / (la.nq * la.nw);
        la.out[ladder_out_index(la, chan, w, q, r, t)] = cscale(to_c128(yr[m][t]), la.inv_beta);
      } else {
        out.p[static_cast<size_t>(b) * out.item_stride + static_cast<size_t>(r) * out.row_stride + t] = yr[m][t];
      }
    }
  }
}

template <typename S, int T, bool LOWER, bool TO_LADDER, int RPT>
__global__ void __launch_bounds__(256) tri_sweep_kernel(TriFactors<S> f, TriVec<S> y, TriVec<S> out, LadderArgs la,
                                                        int item0) {
  tri_sweep_body<S, T, LOWER, TO_LADDER, RPT>(f, y, out, la, item0);
}

// Complex64 sweeps of three rows of nine right-hand sides per thread (n <= 768, nb = 3) fit
// 128 VGPRs without spilling: four wavefronts per SIMD instead of three (the forward sweep took
// 152). Other shapes keep tri_sweep_kernel: under the same cap some of them spill.
template <int T, bool LOWER, int RPT>
__global__ void __launch_bounds__(256, 4) tri_sweep_c64_kernel(TriFactors<float2> f, TriVec<float2> y,
                                                               TriVec<float2> out, LadderArgs la, int item0) {
  tri_sweep_body<float2, T, LOWER, false, RPT>(f, y, out, la, item0);
}

// ---------------------------------------------------------------------------
template <typename S>
int tri_diag_inverse(const TriFactors<S>& f, int batch, bool need_lower, hipStream_t stream) {
  DGAMI_ARG_CHECK(batch >= 1 && batch <= 65535, "triangular solve: batch out of range");
  hipLaunchKernelGGL(tri_diag_inverse_kernel<S>, dim3(f.nblk, batch), dim3(64), 0, stream, f, need_lower ? 1 : 0);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}

template <typename S, int T, bool LOWER, bool TO_LADDER>
static int tri_sweep_t(const TriFactors<S>& f, int batch, TriVec<S> y, TriVec<S> out, const LadderArgs* la,
                       int item0, hipStream_t stream) {
  LadderArgs args{};
  if (la) args = *la;
  const int rpt = static_cast<int>(ceil_div(f.n, 256));
#define DGAMI_TRI_LAUNCH(R)                                                                                      \
  if constexpr (std::is_same<S, float2>::value && !TO_LADDER && R == 3 && T == 9)                                      \
    hipLaunchKernelGGL((tri_sweep_c64_kernel<T, LOWER, R>), dim3(batch), dim3(256), 0, stream, f, y, out, args,  \
                       item0);                                                                                   \
  else                                                                                                           \
    hipLaunchKernelGGL((tri_sweep_kernel<S, T, LOWER, TO_LADDER, R>), dim3(batch), dim3(256), 0, stream, f, y, out, \
                       args, item0)
  switch (rpt) {
    case 1: DGAMI_TRI_LAUNCH(1); break;
    case 2: DGAMI_TRI_LAUNCH(2); break;
    case 3: DGAMI_TRI_LAUNCH(3); break;
    case 4: DGAMI_TRI_LAUNCH(4); break;
    case 5:
    case 6: DGAMI_TRI_LAUNCH(6); break;
    default: DGAMI_ARG_CHECK(false, "triangular solve: compound size above 1536");
  }
#undef DGAMI_TRI_LAUNCH
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}

template <typename S, bool LOWER, bool TO_LADDER>
static int tri_sweep(const TriFactors<S>& f, int batch, int T, TriVec<S> y, TriVec<S> out, const LadderArgs* la,
                     int item0, hipStream_t stream) {
  DGAMI_ARG_CHECK(batch >= 1 && batch <= 65535, "triangular solve: batch out of range");
  switch (T) {
    case 1: return tri_sweep_t<S, 1, LOWER, TO_LADDER>(f, batch, y, out, la, item0, stream);
    case 4: return tri_sweep_t<S, 4, LOWER, TO_LADDER>(f, batch, y, out, la, item0, stream);
    case 9: return tri_sweep_t<S, 9, LOWER, TO_LADDER>(f, batch, y, out, la, item0, stream);
    case 16: return tri_sweep_t<S, 16, LOWER, TO_LADDER>(f, batch, y, out, la, item0, stream);
    default: DGAMI_ARG_CHECK(false, "triangular solve: nb^2 must be 1, 4, 9 or 16");
  }
  return 0;
}

template <typename S>
int tri_forward(const TriFactors<S>& f, int batch, int T, TriVec<S> y, TriVec<S> out, hipStream_t stream) {
  return tri_sweep<S, true, false>(f, batch, T, y, out, nullptr, 0, stream);
}

template <>
int tri_backward<double2>(const TriFactors<double2>& f, int batch, int T, TriVec<double2> y, TriVec<double2> out,
                          const LadderArgs* ladder, int item0, hipStream_t stream) {
  if (ladder) return tri_sweep<double2, false, true>(f, batch, T, y, out, ladder, item0, stream);
  return tri_sweep<double2, false, false>(f, batch, T, y, out, nullptr, 0, stream);
}

template <>
int tri_backward<float2>(const TriFactors<float2>& f, int batch, int T, TriVec<float2> y, TriVec<float2> out,
                         const LadderArgs* ladder, int item0, hipStream_t stream) {
  DGAMI_ARG_CHECK(ladder == nullptr, "triangular solve: ladder output needs complex128 factors");
  return tri_sweep<float2, false, false>(f, batch, T, y, out, nullptr, 0, stream);
}

template int tri_diag_inverse<double2>(const TriFactors<double2>&, int, bool, hipStream_t);
template int tri_diag_inverse<float2>(const TriFactors<float2>&, int, bool, hipStream_t);
template int tri_forward<double2>(const TriFactors<double2>&, int, int, TriVec<double2>, TriVec<double2>, hipStream_t);
template int tri_forward<float2>(const TriFactors<float2>&, int, int, TriVec<float2>, TriVec<float2>, hipStream_t);

}  // namespace dgami
