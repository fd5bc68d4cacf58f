// Linearised Eliashberg operator (eliashberg_solver.py:376-428).
//
// Set-up: Gamma_pp(q) is mapped irr -> full zone, scaled by the channel sign
// s, Fourier transformed over q, and stored per lattice vector R as two dense
// matrices with rows (a,b,v) and columns (c,d,v'):
//   g_hat[R][(a,b,v)][(c,d,p)]      = Ghat(R)_{a c b d}(v, p)
//   g_hat_flip[R][(a,b,v)][(c,d,p)] = Ghat(-R)_{a d b c}(v, np2-1-p)
// (the reference's einsum "xyzacbdvp,xyzcdp->xyzabv" and its flipped copy).
// Matvec for nvec vectors: gg = fft_k(chi0pp . gap); per R one pass over
// both matrices (memory bound: the matrices are read once per application,
// shared by all nvec vectors); then ifft_k with the 0.5/(N beta) norm.
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "dgami.h"
#include "dgami_internal.hpp"
#include "dgami_launch.hpp"

namespace dgami {

constexpr int PAIR_MAXVEC = 8;

__global__ void pair_gather_kernel(const double2* __restrict__ girr, const int32_t* __restrict__ irrk_inv,
                                   double2* __restrict__ out, int N, int nq, int nelem, double sign) {
  const size_t total = static_cast<size_t>(nelem) * N;
  for (size_t idx = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; idx < total;
       idx += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int k = static_cast<int>(idx % N);
    const size_t e = idx / N;
    const int qi = irrk_inv[k];
    out[idx] = (qi >= 0 && qi < nq) ? cscale(girr[static_cast<size_t>(qi) * nelem + e], sign)
                                    : make_double2(__builtin_nan(""), __builtin_nan(""));
  }
}

// hat: [nb][nb][nb][nb][np2][np2][N] (reference element order, R fastest); M = double2 or float2
template <typename M>
__global__ void pair_scatter_kernel(const double2* __restrict__ hat, M* __restrict__ g_hat, M* __restrict__ g_flip,
                                    FftPlan3D plan, int nb, int np2) {
  const int N = plan.ntot;
  const int n = nb * nb * np2;
  const size_t total = static_cast<size_t>(N) * n * n;
  for (size_t idx = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; idx < total;
       idx += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int col = static_cast<int>(idx % n);
    const int row = static_cast<int>((idx / n) % n);
    const int R = static_cast<int>(idx / (static_cast<size_t>(n) * n));
    const int a = row / (nb * np2), b = (row / np2) % nb, v = row % np2;
    const int c = col / (nb * np2), d = (col / np2) % nb, p = col % np2;
    const int z = R % plan.n[2], y = (R / plan.n[2]) % plan.n[1], x = R / (plan.n[2] * plan.n[1]);
    const int Rm = (((plan.n[0] - x) % plan.n[0]) * plan.n[1] + (plan.n[1] - y) % plan.n[1]) * plan.n[2] +
                   (plan.n[2] - z) % plan.n[2];
    const size_t e1 = (((((static_cast<size_t>(a) * nb + c) * nb + b) * nb + d) * np2 + v) * np2 + p);
    const size_t e2 = (((((static_cast<size_t>(a) * nb + d) * nb + b) * nb + c) * np2 + v) * np2 + (np2 - 1 - p));
    g_hat[idx] = from_c128<M>(hat[e1 * N + R]);
    g_flip[idx] = from_c128<M>(hat[e2 * N + Rm]);
  }
}

// t[vec][(a,b,v)][k] = sum_cd chi0pp[k][abcd][v] gap[vec][k][c][d][v]
__global__ void pair_bubble_kernel(const double2* __restrict__ chi0pp, const double2* __restrict__ gap,
                                   double2* __restrict__ t, int N, int nb, int np2, int nvec) {
  const int T = nb * nb;
  const size_t total = static_cast<size_t>(nvec) * T * np2 * N;
  for (size_t idx = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; idx < total;
       idx += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int k = static_cast<int>(idx % N);
    size_t r = idx / N;
    const int v = static_cast<int>(r % np2);
    r /= np2;
    const int ab = static_cast<int>(r % T);
    const int vec = static_cast<int>(r / T);
    double2 acc = make_double2(0.0, 0.0);
    for (int cd = 0; cd < T; ++cd)
      acc = cfma(acc, chi0pp[(static_cast<size_t>(k) * T * T + ab * T + cd) * np2 + v],
                 gap[((static_cast<size_t>(vec) * N + k) * T + cd) * np2 + v]);
    t[idx] = acc;
  }
}

// y[vec][row][R] = sum_col g_hat[R][row][col] x[vec][col][R] + s g_flip[R][row][col] x[vec][col][-R]
// The matrices are complex128 or complex64 (M); products and sums are complex128.
template <typename M>
__global__ void __launch_bounds__(256) pair_gemv_kernel(const M* __restrict__ g_hat, const M* __restrict__ g_flip,
                                                       const double2* __restrict__ x, double2* __restrict__ y,
                                                       FftPlan3D plan, int n, int nvec, double sign) {
  extern __shared__ double2 s_x[];  // [2][nvec][n]: x(R) and x(-R)
  const int N = plan.ntot;
  const int R = blockIdx.x;
  const int z = R % plan.n[2], yy = (R / plan.n[2]) % plan.n[1], xx = R / (plan.n[2] * plan.n[1]);
  const int Rm = (((plan.n[0] - xx) % plan.n[0]) * plan.n[1] + (plan.n[1] - yy) % plan.n[1]) * plan.n[2] +
                 (plan.n[2] - z) % plan.n[2];
  for (int e = threadIdx.x; e < nvec * n; e += blockDim.x) {
    s_x[e] = x[static_cast<size_t>(e) * N + R];
    s_x[nvec * n + e] = x[static_cast<size_t>(e) * N + Rm];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const M* M1 = g_hat + static_cast<size_t>(R) * n * n;
  const M* M2 = g_flip + static_cast<size_t>(R) * n * n;
  for (int row = wave; row < n; row += 4) {
    double2 acc[PAIR_MAXVEC];
#pragma unroll
    for (int vv = 0; vv < PAIR_MAXVEC; ++vv) acc[vv] = make_double2(0.0, 0.0);
    for (int col = lane; col < n; col += 64) {
      const double2 m1 = to_c128(M1[static_cast<size_t>(row) * n + col]);
      const double2 m2 = cscale(to_c128(M2[static_cast<size_t>(row) * n + col]), sign);
#pragma unroll
      for (int vv = 0; vv < PAIR_MAXVEC; ++vv) {
        if (vv < nvec) {
          acc[vv] = cfma(acc[vv], m1, s_x[vv * n + col]);
          acc[vv] = cfma(acc[vv], m2, s_x[nvec * n + vv * n + col]);
        }
      }
    }
#pragma unroll
    for (int vv = 0; vv < PAIR_MAXVEC; ++vv) {
      if (vv >= nvec) break;
      double re = acc[vv].x, im = acc[vv].y;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        re += __shfl_xor(re, off);
        im += __shfl_xor(im, off);
      }
      if (lane == 0) y[(static_cast<size_t>(vv) * n + row) * N + R] = make_double2(re, im);
    }
  }
}

// The same product with RB rows per wave at a time, so that every x value read from LDS feeds RB
// rows (the one-row form spends two LDS reads per two complex FMAs), and, for complex64 matrices
// (CPL = 2), two adjacent columns per lane (16-byte loads). The channel sign is applied to x(-R)
// when it is staged. NV >= nvec vectors are held in registers. Requires n even for CPL = 2.
template <typename M, int NV, int RB, int CPL>
__global__ void __launch_bounds__(256) pair_gemv_rb_kernel(const M* __restrict__ g_hat, const M* __restrict__ g_flip,
                                                          const double2* __restrict__ x, double2* __restrict__ y,
                                                          FftPlan3D plan, int n, int nvec, double sign) {
  extern __shared__ double2 s_x[];  // [2][nvec][n]: x(R) and sign * x(-R)
  const int N = plan.ntot;
  const int R = blockIdx.x;
  const int z = R % plan.n[2], yy = (R / plan.n[2]) % plan.n[1], xx = R / (plan.n[2] * plan.n[1]);
  const int Rm = (((plan.n[0] - xx) % plan.n[0]) * plan.n[1] + (plan.n[1] - yy) % plan.n[1]) * plan.n[2] +
                 (plan.n[2] - z) % plan.n[2];
  for (int e = threadIdx.x; e < nvec * n; e += blockDim.x) {
    s_x[e] = x[static_cast<size_t>(e) * N + R];
    s_x[nvec * n + e] = cscale(x[static_cast<size_t>(e) * N + Rm], sign);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const M* M1 = g_hat + static_cast<size_t>(R) * n * n;
  const M* M2 = g_flip + static_cast<size_t>(R) * n * n;
  for (int r0 = wave * RB; r0 < n; r0 += 4 * RB) {
    double2 acc[RB][NV];
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int vv = 0; vv < NV; ++vv) acc[r][vv] = make_double2(0.0, 0.0);
    for (int c0 = lane * CPL; c0 < n; c0 += 64 * CPL) {
      double2 xa[CPL][NV], xb[CPL][NV];
#pragma unroll
      for (int h = 0; h < CPL; ++h)
#pragma unroll
        for (int vv = 0; vv < NV; ++vv) {
          const bool ok = vv < nvec && c0 + h < n;
          xa[h][vv] = ok ? s_x[vv * n + c0 + h] : make_double2(0.0, 0.0);
          xb[h][vv] = ok ? s_x[(nvec + vv) * n + c0 + h] : make_double2(0.0, 0.0);
        }
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const int row = r0 + r;
        if (row >= n) break;
        double2 m1[CPL], m2[CPL];
        if constexpr (CPL == 2) {
          const float4 a = *reinterpret_cast<const float4*>(M1 + static_cast<size_t>(row) * n + c0);
          const float4 b = *reinterpret_cast<const float4*>(M2 + static_cast<size_t>(row) * n + c0);
          m1[0] = make_double2(a.x, a.y);
          m1[1] = make_double2(a.z, a.w);
          m2[0] = make_double2(b.x, b.y);
          m2[1] = make_double2(b.z, b.w);
        } else {
          m1[0] = to_c128(M1[static_cast<size_t>(row) * n + c0]);
          m2[0] = to_c128(M2[static_cast<size_t>(row) * n + c0]);
        }
#pragma unroll
        for (int h = 0; h < CPL; ++h)
#pragma unroll
          for (int vv = 0; vv < NV; ++vv) {
            acc[r][vv] = cfma(acc[r][vv], m1[h], xa[h][vv]);
            acc[r][vv] = cfma(acc[r][vv], m2[h], xb[h][vv]);
          }
      }
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int row = r0 + r;
#pragma unroll
      for (int vv = 0; vv < NV; ++vv) {
        double re = acc[r][vv].x, im = acc[r][vv].y;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
          re += __shfl_xor(re, off);
          im += __shfl_xor(im, off);
        }
        if (lane == 0 && row < n && vv < nvec) y[(static_cast<size_t>(vv) * n + row) * N + R] = make_double2(re, im);
      }
    }
  }
}

// The row-blocked product for nvec vectors: NV = 1, 2, 4 or 8 registers per row, RB rows per
// wave (RB * NV <= 16 accumulators), two columns per lane for complex64 matrices with up to four
// vectors and even n.
template <typename M, int NV, int RB>
static void pair_gemv_rb_launch(const M* g_hat, const M* g_flip, const double2* t, double2* yv, const FftPlan3D& plan,
                                int n, int nvec, double sign, size_t lds, hipStream_t stream) {
  if constexpr (std::is_same<M, float2>::value && NV <= 4) {
    if (n % 2 == 0) {
      allow_lds(pair_gemv_rb_kernel<M, NV, RB, 2>, lds);
      hipLaunchKernelGGL((pair_gemv_rb_kernel<M, NV, RB, 2>), dim3(plan.ntot), dim3(256), lds, stream, g_hat, g_flip,
                         t, yv, plan, n, nvec, sign);
      return;
    }
  }
  allow_lds(pair_gemv_rb_kernel<M, NV, RB, 1>, lds);
  hipLaunchKernelGGL((pair_gemv_rb_kernel<M, NV, RB, 1>), dim3(plan.ntot), dim3(256), lds, stream, g_hat, g_flip, t, yv,
                     plan, n, nvec, sign);
}

// Rows per wave: measured on the 2-orbital 48x48 operator (n = 512, 2304 lattice vectors), whole
// application with one / four vectors, complex128 and complex64 matrices (ms):
//   one-row form 3.83 / 6.65, 2.50 / 6.36;  RB = 2: 3.65 / 5.05, 2.13 / 3.69;
//   RB = 4: 3.62 / 4.89, 2.11 / 3.34;       RB = 8: 3.62 / 5.01, 2.10 / 3.91.
template <typename M, int NV>
static void pair_gemv_nv_launch(const M* g_hat, const M* g_flip, const double2* t, double2* yv, const FftPlan3D& plan,
                                int n, int nvec, double sign, size_t lds, hipStream_t stream) {
  constexpr int RB = NV >= 8 ? 2 : 4;
  pair_gemv_rb_launch<M, NV, RB>(g_hat, g_flip, t, yv, plan, n, nvec, sign, lds, stream);
}

// out[vec][k][a][b][v] <- t[vec][(a,b,v)][k]
__global__ void pair_out_kernel(const double2* __restrict__ t, double2* __restrict__ out, int N, int n, int nvec) {
  const size_t total = static_cast<size_t>(nvec) * N * n;
  for (size_t idx = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; idx < total;
       idx += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int row = static_cast<int>(idx % n);
    const int k = static_cast<int>((idx / n) % N);
    const int vec = static_cast<int>(idx / (static_cast<size_t>(n) * N));
    out[idx] = t[(static_cast<size_t>(vec) * n + row) * N + k];
  }
}

// chi0pp[k][abcd][v] = G_ad(k, v) conj(G_cb(k, v)) over the centred 2*np of 2*niv_g frequencies,
// |Re|, |Im| < 1e-12 -> 0 (bubble_gen.py:268-279, n_point_base.py:205-211).
__global__ void pair_chi0pp_kernel(const double2* __restrict__ g, double2* __restrict__ out, int nk, int nb, int niv_g,
                                   int np) {
  const int T = nb * nb, P2 = 2 * np;
  const size_t total = static_cast<size_t>(nk) * T * T * P2;
  for (size_t idx = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; idx < total;
       idx += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int v = static_cast<int>(idx % P2);
    const int o = static_cast<int>((idx / P2) % (T * T));
    const int k = static_cast<int>(idx / (static_cast<size_t>(P2) * T * T));
    const int d = o % nb, c = (o / nb) % nb, b = (o / T) % nb, a = o / (T * nb);
    const size_t base = static_cast<size_t>(k) * T * 2 * niv_g + (niv_g - np) + v;
    const double2 x = cmul(g[base + static_cast<size_t>(a * nb + d) * 2 * niv_g],
                           cconj(g[base + static_cast<size_t>(c * nb + b) * 2 * niv_g]));
    out[idx] = make_double2(fabs(x.x) < 1e-12 && fabs(x.y) < 1e-12 ? 0.0 : x.x,
                            fabs(x.x) < 1e-12 && fabs(x.y) < 1e-12 ? 0.0 : x.y);
  }
}

inline unsigned grid_for(size_t total) { return ceil_div(total, 256) > 262144u ? 262144u : ceil_div(total, 256); }

template <typename M>
static int pairing_setup(int nx, int ny, int nz, int nb, int np2, int channel, const dgami_c128* gamma_pp_irr,
                         const int32_t* irrk_inv, int nq_irr, M* g_hat, M* g_hat_flip, hipStream_t stream) {
  DGAMI_ARG_CHECK(gamma_pp_irr && irrk_inv && g_hat && g_hat_flip, "pairing_setup: null pointer");
  DGAMI_ARG_CHECK(nb >= 1 && np2 >= 1 && nq_irr >= 1, "pairing_setup: invalid sizes");
  DGAMI_ARG_CHECK(channel == 0 || channel == 1, "pairing_setup: channel must be 0 (singlet) or 1 (triplet)");
  FftPlan3D plan;
  DGAMI_ARG_CHECK(nx > 0 && ny > 0 && nz > 0 && fft_plan_3d(nx, ny, nz, &plan),
                  "pairing_setup: grid lengths must factor into 2, 3 and 5");
  const int N = plan.ntot;
  const int nelem = nb * nb * nb * nb * np2 * np2;
  double2* hat = static_cast<double2*>(workspace(WS_PAIR, static_cast<size_t>(nelem) * N * sizeof(double2)));
  if (!hat) return -2;
  const double sign = channel == 0 ? 1.0 : -1.0;
  hipLaunchKernelGGL(pair_gather_kernel, dim3(grid_for(static_cast<size_t>(nelem) * N)), dim3(256), 0, stream,
                     reinterpret_cast<const double2*>(gamma_pp_irr), irrk_inv, hat, N, nq_irr, nelem, sign);
  DGAMI_HIP_CHECK(hipGetLastError());
  int rc = launch_fft_grid(hat, hat, nelem, nx, ny, nz, -1, 1.0, stream);
  if (rc) return rc;
  const int n = nb * nb * np2;
  hipLaunchKernelGGL(pair_scatter_kernel<M>, dim3(grid_for(static_cast<size_t>(N) * n * n)), dim3(256), 0, stream,
                     hat, g_hat, g_hat_flip, plan, nb, np2);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}

// DGAMI_PAIR_GEMV=row selects the one-row product (for comparisons).
static bool pair_gemv_one_row() {
  const char* v = getenv("DGAMI_PAIR_GEMV");
  return v && strcmp(v, "row") == 0;
}

template <typename M>
static int pairing_matvec(int nx, int ny, int nz, int nb, int np2, int channel, double beta, const M* g_hat,
                          const M* g_hat_flip, const dgami_c128* chi0pp, int nvec, const dgami_c128* gap,
                          dgami_c128* out, hipStream_t stream) {
  DGAMI_ARG_CHECK(g_hat && g_hat_flip && chi0pp && gap && out, "pairing_matvec: null pointer");
  DGAMI_ARG_CHECK(nvec >= 1 && nvec <= PAIR_MAXVEC, "pairing_matvec: nvec must be 1..8");
  DGAMI_ARG_CHECK(channel == 0 || channel == 1, "pairing_matvec: channel must be 0 or 1");
  DGAMI_ARG_CHECK(nb >= 1 && np2 >= 1 && beta > 0, "pairing_matvec: invalid sizes");
  FftPlan3D plan;
  DGAMI_ARG_CHECK(nx > 0 && ny > 0 && nz > 0 && fft_plan_3d(nx, ny, nz, &plan),
                  "pairing_matvec: grid lengths must factor into 2, 3 and 5");
  const int N = plan.ntot;
  const int n = nb * nb * np2;
  const size_t lds = static_cast<size_t>(2) * nvec * n * sizeof(double2);
  DGAMI_ARG_CHECK(lds <= 160 * 1024, "pairing_matvec: nvec * nb^2 * np2 too large for LDS");
  const size_t vsz = static_cast<size_t>(nvec) * n * N;
  double2* t = static_cast<double2*>(workspace(WS_MISC, 2 * vsz * sizeof(double2)));
  if (!t) return -2;
  double2* yv = t + vsz;
  hipLaunchKernelGGL(pair_bubble_kernel, dim3(grid_for(vsz)), dim3(256), 0, stream,
                     reinterpret_cast<const double2*>(chi0pp), reinterpret_cast<const double2*>(gap), t, N, nb, np2,
                     nvec);
  DGAMI_HIP_CHECK(hipGetLastError());
  int rc = launch_fft_grid(t, t, static_cast<int64_t>(nvec) * n, nx, ny, nz, -1, 1.0, stream);
  if (rc) return rc;
  const double sign = channel == 0 ? 1.0 : -1.0;
  if (pair_gemv_one_row()) {
    allow_lds(pair_gemv_kernel<M>, lds);
    hipLaunchKernelGGL(pair_gemv_kernel<M>, dim3(N), dim3(256), lds, stream, g_hat, g_hat_flip, t, yv, plan, n, nvec,
                       sign);
  } else if (nvec == 1) {
    pair_gemv_nv_launch<M, 1>(g_hat, g_hat_flip, t, yv, plan, n, nvec, sign, lds, stream);
  } else if (nvec == 2) {
    pair_gemv_nv_launch<M, 2>(g_hat, g_hat_flip, t, yv, plan, n, nvec, sign, lds, stream);
  } else if (nvec <= 4) {
    pair_gemv_nv_launch<M, 4>(g_hat, g_hat_flip, t, yv, plan, n, nvec, sign, lds, stream);
  } else {
    pair_gemv_nv_launch<M, 8>(g_hat, g_hat_flip, t, yv, plan, n, nvec, sign, lds, stream);
  }
  DGAMI_HIP_CHECK(hipGetLastError());
  rc = launch_fft_grid(yv, yv, static_cast<int64_t>(nvec) * n, nx, ny, nz, 1, 0.5 / (static_cast<double>(N) * beta) / N,
                       stream);
  if (rc) return rc;
  hipLaunchKernelGGL(pair_out_kernel, dim3(grid_for(vsz)), dim3(256), 0, stream, yv, reinterpret_cast<double2*>(out), N,
                     n, nvec);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace dgami

extern "C" int dgami_pairing_setup(int nx, int ny, int nz, int nb, int np2, int channel,
                                   const dgami_c128* gamma_pp_irr, const int32_t* irrk_inv, int nq_irr,
                                   dgami_c128* g_hat, dgami_c128* g_hat_flip, void* stream) {
  using namespace dgami;
  return pairing_setup(nx, ny, nz, nb, np2, channel, gamma_pp_irr, irrk_inv, nq_irr, reinterpret_cast<double2*>(g_hat),
                       reinterpret_cast<double2*>(g_hat_flip), static_cast<hipStream_t>(stream));
}

extern "C" int dgami_pairing_setup_c64(int nx, int ny, int nz, int nb, int np2, int channel,
                                       const dgami_c128* gamma_pp_irr, const int32_t* irrk_inv, int nq_irr,
                                       dgami_c64* g_hat, dgami_c64* g_hat_flip, void* stream) {
  using namespace dgami;
  return pairing_setup(nx, ny, nz, nb, np2, channel, gamma_pp_irr, irrk_inv, nq_irr, reinterpret_cast<float2*>(g_hat),
                       reinterpret_cast<float2*>(g_hat_flip), static_cast<hipStream_t>(stream));
}

extern "C" int dgami_pairing_matvec(int nx, int ny, int nz, int nb, int np2, int channel, double beta,
                                    const dgami_c128* g_hat, const dgami_c128* g_hat_flip, const dgami_c128* chi0pp,
                                    int nvec, const dgami_c128* gap, dgami_c128* out, void* stream) {
  using namespace dgami;
  return pairing_matvec(nx, ny, nz, nb, np2, channel, beta, reinterpret_cast<const double2*>(g_hat),
                        reinterpret_cast<const double2*>(g_hat_flip), chi0pp, nvec, gap, out,
                        static_cast<hipStream_t>(stream));
}

extern "C" int dgami_pairing_matvec_c64(int nx, int ny, int nz, int nb, int np2, int channel, double beta,
                                        const dgami_c64* g_hat, const dgami_c64* g_hat_flip, const dgami_c128* chi0pp,
                                        int nvec, const dgami_c128* gap, dgami_c128* out, void* stream) {
  using namespace dgami;
  return pairing_matvec(nx, ny, nz, nb, np2, channel, beta, reinterpret_cast<const float2*>(g_hat),
                        reinterpret_cast<const float2*>(g_hat_flip), chi0pp, nvec, gap, out,
                        static_cast<hipStream_t>(stream));
}

extern "C" int dgami_chi0_pp(int nk, int nb, int niv_g, int np, const dgami_c128* g_k, dgami_c128* chi0pp,
                             void* stream) {
  using namespace dgami;
  DGAMI_ARG_CHECK(g_k && chi0pp, "chi0_pp: null pointer");
  DGAMI_ARG_CHECK(nk >= 1 && nb >= 1 && nb <= 4 && np >= 1 && np <= niv_g, "chi0_pp: invalid sizes (np <= niv_g)");
  const size_t total = static_cast<size_t>(nk) * nb * nb * nb * nb * 2 * np;
  hipLaunchKernelGGL(pair_chi0pp_kernel, dim3(grid_for(total)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     reinterpret_cast<const double2*>(g_k), reinterpret_cast<double2*>(chi0pp), nk, nb, niv_g, np);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}
