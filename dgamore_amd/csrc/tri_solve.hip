// Batched blocked triangular solves (see tri_solve.hpp for the scheme).
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

// ---------------------------------------------------------------------------
// One block step. grid (1 + ntiles, batch), 256 threads. Workgroup 0 writes x_b; workgroup
// t >= 1 updates row tile t-1 of the rows below (LOWER) or above (upper) block kb.
template <typename S, int T, bool LOWER, bool TO_LADDER>
__global__ void __launch_bounds__(256) tri_step_kernel(TriFactors<S> f, TriVec<S> y, TriVec<S> out, LadderArgs la,
                                                       int item0, int kb) {
  const int b = blockIdx.y;
  const int c0 = kb * TRI_B;
  const int cb = min(TRI_B, f.n - c0);
  const int tid = threadIdx.x;
  __shared__ S s_dinv[TRI_B][TRI_B + 1];
  __shared__ S s_y[TRI_B][T];
  __shared__ S s_x[TRI_B][T];
  __shared__ S s_f[TRI_TILE][TRI_B + 1];
  const S* Db = f.dinvblk + (static_cast<size_t>(b) * f.nblk + kb) * 2 * TRI_B * TRI_B + (LOWER ? 0 : TRI_B * TRI_B);
  S* Y = y.p + static_cast<size_t>(b) * y.item_stride;
  for (int e = tid; e < TRI_B * TRI_B; e += 256) {
    const int r = e / TRI_B, c = e - (e / TRI_B) * TRI_B;
    s_dinv[r][c] = Db[e];
  }
  for (int e = tid; e < TRI_B * T; e += 256) {
    const int i = e / T, t = e - (e / T) * T;
    s_y[i][t] = i < cb ? Y[static_cast<size_t>(c0 + i) * y.row_stride + t] : czero<S>();
  }
  __syncthreads();
  for (int e = tid; e < TRI_B * T; e += 256) {
    const int i = e / T, t = e - (e / T) * T;
    S acc = czero<S>();
#pragma unroll 8
    for (int k = 0; k < TRI_B; ++k) acc = cfma(acc, s_dinv[i][k], s_y[k][t]);
    s_x[i][t] = acc;
  }
  __syncthreads();
  if (blockIdx.x == 0) {
    for (int e = tid; e < cb * T; e += 256) {
      const int i = e / T, t = e - (e / T) * T;
      if constexpr (TO_LADDER) {
        const int item = item0 + b;
        const int q = item % la.nq, w = (item / la.nq) % la.nw, chan = item / (la.nq * la.nw);
        la.out[ladder_out_index(la, chan, w, q, c0 + i, t)] = cscale(to_c128(s_x[i][t]), la.inv_beta);
      } else {
        out.p[static_cast<size_t>(b) * out.item_stride + static_cast<size_t>(c0 + i) * out.row_stride + t] = s_x[i][t];
      }
    }
    return;
  }
  const int r0 = LOWER ? c0 + cb + (blockIdx.x - 1) * TRI_TILE : (blockIdx.x - 1) * TRI_TILE;
  const int r1 = LOWER ? f.n : c0;
  const S* F = f.W + static_cast<size_t>(b) * f.n * f.ld;
  for (int e = tid; e < TRI_TILE * TRI_B; e += 256) {
    const int r = e / TRI_B, c = e - (e / TRI_B) * TRI_B;
    s_f[r][c] = (r0 + r < r1 && c < cb) ? F[static_cast<size_t>(r0 + r) * f.ld + c0 + c] : czero<S>();
  }
  __syncthreads();
  const int rr = tid >> 2, tq = tid & 3;
  const int r = r0 + rr;
  if (r >= r1) return;
#pragma unroll
  for (int t = tq; t < T; t += 4) {
    S* yp = Y + static_cast<size_t>(r) * y.row_stride + t;
    S acc = *yp;
#pragma unroll 8
    for (int k = 0; k < TRI_B; ++k) acc = cfms(acc, s_f[rr][k], s_x[k][t]);
    *yp = acc;
  }
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
  for (int s = 0; s < f.nblk; ++s) {
    const int kb = LOWER ? s : f.nblk - 1 - s;
    const int c0 = kb * TRI_B;
    const int cb = min(TRI_B, f.n - c0);
    const int rows = LOWER ? f.n - c0 - cb : c0;
    const int tiles = 1 + static_cast<int>(ceil_div(rows, TRI_TILE));
    hipLaunchKernelGGL((tri_step_kernel<S, T, LOWER, TO_LADDER>), dim3(tiles, batch), dim3(256), 0, stream, f, y,
                       out, args, item0, kb);
    DGAMI_HIP_CHECK(hipGetLastError());
  }
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
