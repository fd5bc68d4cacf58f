// Batched complex128 GEMM on LDS tiles and the FP64 matrix cores, C = alpha A B + beta C.
// Row-major operands with leading dimensions and batch strides (elements).
#pragma once

#include "dgami_internal.hpp"

namespace dgami {

struct ZgemmArgs {
  const double2* A;
  const double2* B;
  double2* C;
  int M, N, K;
  int lda, ldb, ldc;
  long long sa, sb, sc;  // batch strides
  double2 alpha, beta;
};

constexpr int ZG_TM = 64, ZG_TN = 64, ZG_TK = 16;

// grid: (ceil(N/64), ceil(M/64), batch); block 256. Each wavefront owns a 32x32 quarter of the
// tile as 2x2 blocks of 16x16 on the FP64 matrix cores: a complex product is two real MFMA
// chains over the interleaved k index, C_re += [a_re, -a_im].[b_re; b_im] and
// C_im += [a_im, a_re].[b_re; b_im] (16x16x4 f64 layout, lane l: a = A[l%16][l/16],
// b = B[l/16][l%16], d[e] = D[4*(l/16)+e][l%16]).
static __global__ void __launch_bounds__(256) zgemm_kernel(ZgemmArgs g) {
  typedef double v4d __attribute__((ext_vector_type(4)));
  const int b = blockIdx.z;
  const double2* A = g.A + b * g.sa;
  const double2* B = g.B + b * g.sb;
  double2* C = g.C + b * g.sc;
  const int m0 = blockIdx.y * ZG_TM, n0 = blockIdx.x * ZG_TN;
  __shared__ double2 s_a[ZG_TM][ZG_TK + 1];
  __shared__ double2 s_b[ZG_TK][ZG_TN + 1];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int l16 = lane & 15, kq = lane >> 4, kc = kq >> 1;
  const bool odd = (kq & 1) != 0;
  v4d accr[2][2], acci[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      accr[i][j] = v4d{0, 0, 0, 0};
      acci[i][j] = v4d{0, 0, 0, 0};
    }
  for (int k0 = 0; k0 < g.K; k0 += ZG_TK) {
    for (int e = tid; e < ZG_TM * ZG_TK; e += 256) {
      const int r = e / ZG_TK, s = e % ZG_TK;
      s_a[r][s] = (m0 + r < g.M && k0 + s < g.K) ? A[static_cast<size_t>(m0 + r) * g.lda + k0 + s]
                                                  : make_double2(0.0, 0.0);
    }
    for (int e = tid; e < ZG_TK * ZG_TN; e += 256) {
      const int s = e / ZG_TN, c = e % ZG_TN;
      s_b[s][c] = (k0 + s < g.K && n0 + c < g.N) ? B[static_cast<size_t>(k0 + s) * g.ldb + n0 + c]
                                                  : make_double2(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < ZG_TK; kk += 2) {
      const int k = kk + kc;
      double ar[2], ai[2], bv[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const double2 v = s_a[wr + 16 * i + l16][k];
        ar[i] = odd ? -v.y : v.x;
        ai[i] = odd ? v.x : v.y;
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const double2 v = s_b[k][wc + 16 * j + l16];
        bv[j] = odd ? v.y : v.x;
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          accr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bv[j], accr[i][j], 0, 0, 0);
          acci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], bv[j], acci[i][j], 0, 0, 0);
        }
    }
    __syncthreads();
  }
  const bool has_beta = g.beta.x != 0.0 || g.beta.y != 0.0;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int c = n0 + wc + 16 * j + l16;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = m0 + wr + 16 * i + 4 * kq + e;
        if (r < g.M && c < g.N) {
          double2* p = C + static_cast<size_t>(r) * g.ldc + c;
          double2 val = cmul(g.alpha, make_double2(accr[i][j][e], acci[i][j][e]));
          if (has_beta) val = cfma(val, g.beta, *p);
          *p = val;
        }
      }
    }
}

inline int launch_zgemm(const ZgemmArgs& g, int batch, hipStream_t stream) {
  if (g.M <= 0 || g.N <= 0 || batch <= 0) return 0;
  dim3 grid(ceil_div(g.N, ZG_TN), ceil_div(g.M, ZG_TM), batch);
  hipLaunchKernelGGL(zgemm_kernel, grid, dim3(256), 0, stream, g);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace dgami
