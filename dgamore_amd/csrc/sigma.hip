// Non-local self-energy from the ladder kernel, R-space form
// (nonlocal_sde.py:544-594 with the caller's ifft and to_full_niv_range at
// :874-879, and the irr->FBZ map of mpi_utils.py:123-131).
//
//   Sigma_ij(R, v) = -1/(2 beta N) sum_w sum_ad G_ad(R, v-w) Khat_aijd(R, w, v),
//   Khat(R) = sum_q K(q) e^{+i q R},  Sigma(k) = (1/N) sum_R e^{+i k R} Sigma(R).
//
// Pass 1 (one workgroup per (i, j, v, w-chunk)): for each w of the chunk and
// each (a, d), gather K(q) for every q of the full zone from the irreducible
// kernel (K(q) = K_irr(irrk_inv[q]); w < 0 from w > 0 by conjugation and
// v -> -v, the reference's to_full_niw_range), transform it in LDS, multiply
// by G_ad(R, v - w) and accumulate in registers. Partial sums per w-chunk go
// to a workspace; pass 2 adds the chunks in a fixed order (deterministic),
// applies the prefactor, transforms back to k in LDS and writes both halves
// of the fermionic range.
#include <algorithm>

#include "dgami.h"
#include "dgami_internal.hpp"
#include "dgami_launch.hpp"

namespace dgami {

constexpr int SIG_ACC = 16;  // accumulator elements per thread (N <= 16 * threads)

__global__ void __launch_bounds__(256) sigma_partial_kernel(const double2* __restrict__ kirr,
                                                            const int32_t* __restrict__ irrk_inv,
                                                            const double2* __restrict__ g_r, double2* __restrict__ part,
                                                            FftPlan3D plan, int nb, int niw, int niv_core, int nvg,
                                                            int nq, int nchunk, int wper) {
  extern __shared__ double2 smem[];
  const int N = plan.ntot;
  double2* buf = smem;
  double2* tw = smem + N;
  int* s_inv = reinterpret_cast<int*>(tw + plan.n[0] + plan.n[1] + plan.n[2]);
  const int T = nb * nb;
  int idx = blockIdx.x;
  const int chunk = idx % nchunk;
  idx /= nchunk;
  const int v = idx % niv_core;
  idx /= niv_core;
  const int j = idx % nb;
  const int i = idx / nb;
  const int nvk = 2 * niv_core;

  fft_twiddles_3d(tw, plan, 1.0);
  for (int r = threadIdx.x; r < N; r += blockDim.x) s_inv[r] = irrk_inv[r];
  double2 acc[SIG_ACC];
#pragma unroll
  for (int e = 0; e < SIG_ACC; ++e) acc[e] = make_double2(0.0, 0.0);
  __syncthreads();

  const int w_begin = -niw + chunk * wper;
  const int w_end = (w_begin + wper - 1 < niw) ? w_begin + wper - 1 : niw;
  for (int w = w_begin; w <= w_end; ++w) {
    const int aw = w < 0 ? -w : w;
    const int vk = w < 0 ? niv_core - 1 - v : niv_core + v;  // frequency index in the kernel's 2*niv_core box
    for (int a = 0; a < nb; ++a) {
      for (int d = 0; d < nb; ++d) {
        const int blk = ((a * nb + i) * nb + d) * nb + j;  // M order of (o1,o2,o3,o4) = (a,i,j,d)
        const double2* ks = kirr + ((static_cast<size_t>(aw) * T * T + blk) * nvk + vk) * nq;
        for (int r = threadIdx.x; r < N; r += blockDim.x) {
          const int qi = s_inv[r];
          double2 x = (qi >= 0 && qi < nq) ? ks[qi] : make_double2(__builtin_nan(""), __builtin_nan(""));
          buf[r] = w < 0 ? cconj(x) : x;
        }
        __syncthreads();
        fft_3d_lds(buf, plan, tw, 1.0);
        const double2* gs = g_r + (static_cast<size_t>(a * nb + d) * nvg + (v - w + niv_core + niw)) * N;
#pragma unroll
        for (int e = 0; e < SIG_ACC; ++e) {
          const int r = threadIdx.x + e * blockDim.x;
          if (r < N) acc[e] = cfma(acc[e], gs[r], buf[r]);
        }
        __syncthreads();
      }
    }
  }
  double2* dst = part + ((static_cast<size_t>(chunk) * T + i * nb + j) * niv_core + v) * N;
#pragma unroll
  for (int e = 0; e < SIG_ACC; ++e) {
    const int r = threadIdx.x + e * blockDim.x;
    if (r < N) dst[r] = acc[e];
  }
}

__global__ void __launch_bounds__(256) sigma_final_kernel(const double2* __restrict__ part, double2* __restrict__ sigma,
                                                          FftPlan3D plan, int nb, int niv_core, int nchunk,
                                                          double prefactor) {
  extern __shared__ double2 smem[];
  const int N = plan.ntot;
  double2* buf = smem;
  double2* tw = smem + N;
  const int T = nb * nb;
  const int v = blockIdx.x % niv_core;
  const int ij = blockIdx.x / niv_core;
  fft_twiddles_3d(tw, plan, 1.0);
  for (int r = threadIdx.x; r < N; r += blockDim.x) {
    double2 s = make_double2(0.0, 0.0);
    for (int c = 0; c < nchunk; ++c) s = cadd(s, part[((static_cast<size_t>(c) * T + ij) * niv_core + v) * N + r]);
    buf[r] = cscale(s, prefactor);
  }
  __syncthreads();
  fft_3d_lds(buf, plan, tw, 1.0);
  const double inv_n = 1.0 / N;
  const int nvs = 2 * niv_core;
  for (int k = threadIdx.x; k < N; k += blockDim.x) {
    const double2 s = cscale(buf[k], inv_n);
    double2* out = sigma + (static_cast<size_t>(k) * T + ij) * nvs;
    out[niv_core + v] = s;
    out[niv_core - 1 - v] = cconj(s);
  }
}

// Irreducible q -> list of (irreducible q, orbital transform) entries of the full zone:
//   out[w][(o1,o2),(o4,o3)][v][j] = sum_p U[o1,p1] U*[o2,p2] U*[o4,p4] U[o3,p3] x[w][(p1,p2),(p4,p3)][v][src[j]]
// with x = conj(kernel) for entries flagged conj. One thread per output element, j fastest;
// zero entries of U are skipped, so a signed permutation costs one product per element.
__global__ void __launch_bounds__(256) kernel_orbital_expand_kernel(const double2* __restrict__ kin,
                                                                   const int* __restrict__ src,
                                                                   const double2* __restrict__ u,
                                                                   const int* __restrict__ conj, double2* __restrict__ out,
                                                                   int nq, int next, int nb, int nw, int nv) {
  const int M = nb * nb * nb * nb;
  const size_t total = static_cast<size_t>(nw) * M * nv * next;
  for (size_t idx = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; idx < total;
       idx += static_cast<size_t>(gridDim.x) * 256) {
    const int j = static_cast<int>(idx % next);
    size_t rest = idx / next;
    const int v = static_cast<int>(rest % nv);
    rest /= nv;
    const int m = static_cast<int>(rest % M);
    const int w = static_cast<int>(rest / M);
    const int o3 = m % nb, o4 = (m / nb) % nb, o2 = (m / (nb * nb)) % nb, o1 = m / (nb * nb * nb);
    const int s = src[j];
    if (static_cast<unsigned>(s) >= static_cast<unsigned>(nq)) {  // invalid entry: NaN, no access
      out[idx] = make_double2(__builtin_nan(""), __builtin_nan(""));
      continue;
    }
    const bool cj = conj && conj[j];
    const double2* U = u + static_cast<size_t>(j) * nb * nb;
    double2 acc = make_double2(0.0, 0.0);
    for (int p1 = 0; p1 < nb; ++p1) {
      const double2 a = U[o1 * nb + p1];
      if (a.x == 0.0 && a.y == 0.0) continue;
      for (int p2 = 0; p2 < nb; ++p2) {
        const double2 b = U[o2 * nb + p2];
        if (b.x == 0.0 && b.y == 0.0) continue;
        const double2 ab = cmul(a, cconj(b));
        for (int p4 = 0; p4 < nb; ++p4) {
          const double2 d = U[o4 * nb + p4];
          if (d.x == 0.0 && d.y == 0.0) continue;
          const double2 abd = cmul(ab, cconj(d));
          for (int p3 = 0; p3 < nb; ++p3) {
            const double2 c = U[o3 * nb + p3];
            if (c.x == 0.0 && c.y == 0.0) continue;
            const int mi = ((p1 * nb + p2) * nb + p4) * nb + p3;
            double2 x = kin[((static_cast<size_t>(w) * M + mi) * nv + v) * nq + s];
            if (cj) x = cconj(x);
            acc = cadd(acc, cmul(cmul(abd, c), x));
          }
        }
      }
    }
    out[idx] = acc;
  }
}

}  // namespace dgami

extern "C" int dgami_kernel_orbital_expand(int nq, int next, int nb, int nw, int nv, const dgami_c128* kernel_irr,
                                           const int32_t* src, const dgami_c128* u, const int32_t* conj,
                                           dgami_c128* kernel_ext, void* stream) {
  using namespace dgami;
  DGAMI_ARG_CHECK(kernel_irr && src && u && kernel_ext, "kernel_orbital_expand: null pointer");
  DGAMI_ARG_CHECK(nq >= 1 && next >= 1 && nb >= 1 && nb <= 4 && nw >= 1 && nv >= 1,
                  "kernel_orbital_expand: invalid sizes");
  const size_t total = static_cast<size_t>(nw) * nb * nb * nb * nb * nv * next;
  const size_t blocks = std::min<size_t>(ceil_div(total, 256), size_t(1) << 20);
  hipLaunchKernelGGL(kernel_orbital_expand_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), reinterpret_cast<const double2*>(kernel_irr), src,
                     reinterpret_cast<const double2*>(u), conj, reinterpret_cast<double2*>(kernel_ext), nq, next, nb,
                     nw, nv);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}

extern "C" int dgami_sigma_from_kernel(int nx, int ny, int nz, int nb, int niw, int niv_core, int niv_g, double beta,
                                       const dgami_c128* kernel_irr, const int32_t* irrk_inv, int nq_irr,
                                       const dgami_c128* g_k, dgami_c128* sigma, void* stream_) {
  using namespace dgami;
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  DGAMI_ARG_CHECK(kernel_irr && irrk_inv && g_k && sigma, "sigma: null pointer");
  DGAMI_ARG_CHECK(nb >= 1 && nb <= 4 && niw >= 0 && niv_core >= 1 && nq_irr >= 1 && beta > 0, "sigma: invalid sizes");
  DGAMI_ARG_CHECK(niv_g >= niv_core + niw, "sigma: G must hold at least niw + niv_core fermionic frequencies");
  FftPlan3D plan;
  DGAMI_ARG_CHECK(nx > 0 && ny > 0 && nz > 0 && fft_plan_3d(nx, ny, nz, &plan),
                  "sigma: grid lengths must factor into 2, 3 and 5");
  const int N = plan.ntot;
  int threads = fft_block_threads(plan);
  while (threads * SIG_ACC < N) threads += 64;
  DGAMI_ARG_CHECK(threads <= 256, "sigma: momentum grid too large (max 4096 points)");
  const int nvg = 2 * (niv_core + niw);
  const size_t gbytes = static_cast<size_t>(nb) * nb * nvg * N * sizeof(double2);
  double2* gk = static_cast<double2*>(workspace(WS_GK, gbytes));
  double2* gr = static_cast<double2*>(workspace(WS_GR, gbytes));
  if (!gk || !gr) return -2;
  int rc = launch_ref2_to_gk(reinterpret_cast<const double2*>(g_k), gk, N, nb, 2 * niv_g, niv_g - niv_core - niw, nvg,
                             stream);
  if (rc) return rc;
  rc = launch_fft_grid(gk, gr, static_cast<int64_t>(nb) * nb * nvg, nx, ny, nz, -1, 1.0, stream);
  if (rc) return rc;
  const int nw_full = 2 * niw + 1;
  const int base = nb * nb * niv_core;
  // enough workgroups (about 16k) that the gathers of the kernel are hidden: measured at 24x24, 3
  // orbitals, niw = 30, niv = 40: 5.4 ms with 2k workgroups, 3.7 ms with 16k, 3.9 ms with 64k
  int nchunk = (16384 + base - 1) / base;
  if (nchunk > nw_full) nchunk = nw_full;
  if (nchunk < 1) nchunk = 1;
  const int wper = (nw_full + nchunk - 1) / nchunk;
  nchunk = (nw_full + wper - 1) / wper;
  double2* part = static_cast<double2*>(workspace(WS_SIGMA_PART, static_cast<size_t>(nchunk) * base * N * sizeof(double2)));
  if (!part) return -2;
  const size_t lds1 = (N + nx + ny + nz) * sizeof(double2) + N * sizeof(int);
  DGAMI_ARG_CHECK(lds1 <= 160 * 1024, "sigma: momentum grid too large for LDS");
  allow_lds(sigma_partial_kernel, lds1);
  hipLaunchKernelGGL(sigma_partial_kernel, dim3(base * nchunk), dim3(threads), lds1, stream,
                     reinterpret_cast<const double2*>(kernel_irr), irrk_inv, gr, part, plan, nb, niw, niv_core, nvg,
                     nq_irr, nchunk, wper);
  DGAMI_HIP_CHECK(hipGetLastError());
  const size_t lds2 = (N + nx + ny + nz) * sizeof(double2);
  allow_lds(sigma_final_kernel, lds2);
  hipLaunchKernelGGL(sigma_final_kernel, dim3(base), dim3(threads), lds2, stream, part,
                     reinterpret_cast<double2*>(sigma), plan, nb, niv_core, nchunk, -0.5 / beta / N);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}
