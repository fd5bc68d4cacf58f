// Small-block algebra of the ladder: nb^2 x nb^2 complex blocks per (q, w[, v]).
//
// A "team" of TS lanes (TS >= T = nb^2, power of two, inside one wavefront)
// owns one block; lane r holds row r in registers. Inversion is Gauss-Jordan
// with partial pivoting (largest |Re|+|Im|, first index on ties, as LAPACK's
// izamax); rows move between lanes with wavefront shuffles, so no LDS and no
// barriers are needed. Blocks are in M order: element (r, c) of the
// compound block [(o1,o2),(o4,o3)] sits at r*T + c.
#include "dgami.h"
#include "dgami_internal.hpp"
#include "dgami_launch.hpp"

namespace dgami {

__device__ __forceinline__ double2 shfl2(double2 v, int src, int width) {
  return make_double2(__shfl(v.x, src, width), __shfl(v.y, src, width));
}

// In-place inverse of the team's block. Lane r (< T) holds row r of `m`;
// on return it holds row r of m^-1. Returns true if a pivot was exactly zero.
template <int T, int TS>
__device__ __forceinline__ bool team_invert(double2 (&m)[T], int r) {
  double2 e[T];
#pragma unroll
  for (int j = 0; j < T; ++j) e[j] = make_double2(r == j ? 1.0 : 0.0, 0.0);
  bool singular = false;
#pragma unroll
  for (int k = 0; k < T; ++k) {
    double val = (r >= k && r < T) ? cabs1(m[k]) : -1.0;
    int idx = r;
#pragma unroll
    for (int off = TS / 2; off > 0; off >>= 1) {
      const double ov = __shfl_xor(val, off, TS);
      const int oi = __shfl_xor(idx, off, TS);
      if (ov > val || (ov == val && oi < idx)) {
        val = ov;
        idx = oi;
      }
    }
    const int p = idx;
    if (!(val > 0.0)) singular = true;
    const int src = (r == k) ? p : ((r == p) ? k : r);
#pragma unroll
    for (int j = 0; j < T; ++j) {
      m[j] = shfl2(m[j], src, TS);
      e[j] = shfl2(e[j], src, TS);
    }
    const double2 piv = shfl2(m[k], k, TS);
    const double2 pinv = (val > 0.0) ? cinv(piv) : make_double2(0.0, 0.0);
    if (r == k) {
#pragma unroll
      for (int j = 0; j < T; ++j) {
        m[j] = cmul(m[j], pinv);
        e[j] = cmul(e[j], pinv);
      }
    }
    const double2 f = m[k];
#pragma unroll
    for (int j = 0; j < T; ++j) {
      const double2 mk = shfl2(m[j], k, TS);
      const double2 ek = shfl2(e[j], k, TS);
      if (r != k) {
        m[j] = cfms(m[j], f, mk);
        e[j] = cfms(e[j], f, ek);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < T; ++j) m[j] = e[j];
  return singular;
}

// Row r of the team product A.B, lane r holding row r of both operands.
template <int T, int TS>
__device__ __forceinline__ void team_matmul(const double2 (&a)[T], const double2 (&b)[T], double2 (&out)[T]) {
#pragma unroll
  for (int j = 0; j < T; ++j) out[j] = make_double2(0.0, 0.0);
#pragma unroll
  for (int c = 0; c < T; ++c) {
#pragma unroll
    for (int j = 0; j < T; ++j) out[j] = cfma(out[j], a[c], shfl2(b[j], c, TS));
  }
}

// ---------------------------------------------------------------------------
// chi0: full/core frequency sums  (thread per (w, blk, q))
// ---------------------------------------------------------------------------
__global__ void chi0_sums_kernel(const double2* __restrict__ chi0, double2* full_sum, double2* core_sum, int nq,
                                 int nb4, int nw, int nvf, int v0, int nvc, double inv_beta2) {
  const size_t total = static_cast<size_t>(nw) * nb4 * nq;
  for (size_t idx = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; idx < total;
       idx += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const size_t q = idx % nq;
    const size_t wb = idx / nq;  // w * nb4 + blk
    const double2* src = chi0 + wb * nvf * nq + q;
    double2 sf = make_double2(0.0, 0.0), sc = make_double2(0.0, 0.0);
    for (int v = 0; v < nvf; ++v) {
      const double2 x = src[static_cast<size_t>(v) * nq];
      sf = cadd(sf, x);
      if (v >= v0 && v < v0 + nvc) sc = cadd(sc, x);
    }
    if (full_sum) full_sum[idx] = cscale(sf, inv_beta2);
    if (core_sum) core_sum[idx] = cscale(sc, inv_beta2);
  }
}

// ---------------------------------------------------------------------------
// chi0 core block inverse  (team per (w, v, q))
// ---------------------------------------------------------------------------
template <int T, int TS>
__global__ void __launch_bounds__(256) chi0_inv_kernel(const double2* __restrict__ chi0, double2* __restrict__ inv,
                                                      int nq, int nw, int nvf, int v0, int nvc) {
  const int lane = threadIdx.x % TS;
  const size_t team = (blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x) / TS;
  const size_t nteam = static_cast<size_t>(nw) * nvc * nq;
  const bool active = team < nteam;
  const size_t tt = active ? team : 0;
  const int q = static_cast<int>(tt % nq);
  const int v = static_cast<int>((tt / nq) % nvc);
  const int w = static_cast<int>(tt / (static_cast<size_t>(nq) * nvc));
  const int r = lane < T ? lane : 0;
  double2 m[T];
#pragma unroll
  for (int c = 0; c < T; ++c)
    m[c] = chi0[((static_cast<size_t>(w) * T * T + r * T + c) * nvf + v0 + v) * nq + q];
  team_invert<T, TS>(m, lane);
  if (active && lane < T) {
#pragma unroll
    for (int c = 0; c < T; ++c) inv[((static_cast<size_t>(w) * T * T + r * T + c) * nvc + v) * nq + q] = m[c];
  }
}

// ---------------------------------------------------------------------------
// vrg = beta chi0_inv . chi_aux_sum  (team per (w, v, q))
// ---------------------------------------------------------------------------
template <int T, int TS>
__global__ void __launch_bounds__(256) vrg_kernel(const double2* __restrict__ chi0_inv, const double2* __restrict__ aux,
                                                 double2* __restrict__ vrg, int nq, int nw, int nv, double beta) {
  const int lane = threadIdx.x % TS;
  const size_t team = (blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x) / TS;
  const size_t nteam = static_cast<size_t>(nw) * nv * nq;
  const bool active = team < nteam;
  const size_t tt = active ? team : 0;
  const int q = static_cast<int>(tt % nq);
  const int v = static_cast<int>((tt / nq) % nv);
  const int w = static_cast<int>(tt / (static_cast<size_t>(nq) * nv));
  const int r = lane < T ? lane : 0;
  double2 a[T], b[T], o[T];
#pragma unroll
  for (int c = 0; c < T; ++c) {
    const size_t off = ((static_cast<size_t>(w) * T * T + r * T + c) * nv + v) * nq + q;
    a[c] = chi0_inv[off];
    b[c] = aux[off];
  }
  team_matmul<T, TS>(a, b, o);
  if (active && lane < T) {
#pragma unroll
    for (int c = 0; c < T; ++c)
      vrg[((static_cast<size_t>(w) * T * T + r * T + c) * nv + v) * nq + q] = cscale(o[c], beta);
  }
}

// ---------------------------------------------------------------------------
// chi_phys = ((sum_v aux / beta + full - core)^-1 + u_r + v_r)^-1  (team per (w, q))
// ---------------------------------------------------------------------------
template <int T, int TS>
__global__ void __launch_bounds__(256) chi_phys_kernel(const double2* __restrict__ aux, const double2* __restrict__ full,
                                                      const double2* __restrict__ core, const double2* __restrict__ u_r,
                                                      const double2* __restrict__ v_r, double2* __restrict__ chi_phys,
                                                      int32_t* info, int nq, int nw, int nv, double inv_beta) {
  const int lane = threadIdx.x % TS;
  const size_t team = (blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x) / TS;
  const size_t nteam = static_cast<size_t>(nw) * nq;
  const bool active = team < nteam;
  const size_t tt = active ? team : 0;
  const int q = static_cast<int>(tt % nq);
  const int w = static_cast<int>(tt / nq);
  const int r = lane < T ? lane : 0;
  double2 m[T];
#pragma unroll
  for (int c = 0; c < T; ++c) {
    const size_t blk = static_cast<size_t>(w) * T * T + r * T + c;
    double2 s = make_double2(0.0, 0.0);
    const double2* src = aux + blk * nv * nq + q;
    for (int v = 0; v < nv; ++v) s = cadd(s, src[static_cast<size_t>(v) * nq]);
    s = cscale(s, inv_beta);
    m[c] = csub(cadd(s, full[blk * nq + q]), core[blk * nq + q]);
  }
  bool sing = team_invert<T, TS>(m, lane);
#pragma unroll
  for (int c = 0; c < T; ++c) {
    double2 uu = u_r[r * T + c];
    if (v_r) uu = cadd(uu, v_r[static_cast<size_t>(q) * T * T + r * T + c]);
    m[c] = cadd(m[c], uu);
  }
  sing |= team_invert<T, TS>(m, lane);
  if (active && lane < T) {
#pragma unroll
    for (int c = 0; c < T; ++c) chi_phys[(static_cast<size_t>(w) * T * T + r * T + c) * nq + q] = m[c];
    if (lane == 0 && info) info[static_cast<size_t>(w) * nq + q] = sing ? 1 : 0;
  }
}

// ---------------------------------------------------------------------------
// K_r = u . (vrg - vrg . u . chi_phys - [magn] 2/3 1)  (team per (w, v, q))
// ---------------------------------------------------------------------------
template <int T, int TS>
__global__ void __launch_bounds__(256) kernel_r_kernel(const double2* __restrict__ vrg, const double2* __restrict__ chi_phys,
                                                      const double2* __restrict__ u_r, const double2* __restrict__ v_r,
                                                      double2* kernel, int nq, int nw, int nv, int magn, double weight,
                                                      int accumulate) {
  const int lane = threadIdx.x % TS;
  const size_t team = (blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x) / TS;
  const size_t nteam = static_cast<size_t>(nw) * nv * nq;
  const bool active = team < nteam;
  const size_t tt = active ? team : 0;
  const int q = static_cast<int>(tt % nq);
  const int v = static_cast<int>((tt / nq) % nv);
  const int w = static_cast<int>(tt / (static_cast<size_t>(nq) * nv));
  const int r = lane < T ? lane : 0;
  double2 g[T], u[T], x[T], t1[T], t2[T];
#pragma unroll
  for (int c = 0; c < T; ++c) {
    const size_t blk = static_cast<size_t>(w) * T * T + r * T + c;
    g[c] = vrg[(blk * nv + v) * nq + q];
    x[c] = chi_phys[blk * nq + q];
    double2 uu = u_r[r * T + c];
    if (v_r) uu = cadd(uu, v_r[static_cast<size_t>(q) * T * T + r * T + c]);
    u[c] = uu;
  }
  team_matmul<T, TS>(g, u, t1);   // vrg . u
  team_matmul<T, TS>(t1, x, t2);  // vrg . u . chi
#pragma unroll
  for (int c = 0; c < T; ++c) {
    t2[c] = csub(g[c], t2[c]);
    if (magn && c == r) t2[c].x -= 2.0 / 3.0;
  }
  team_matmul<T, TS>(u, t2, t1);  // u . kern
  if (active && lane < T) {
#pragma unroll
    for (int c = 0; c < T; ++c) {
      const size_t off = ((static_cast<size_t>(w) * T * T + r * T + c) * nv + v) * nq + q;
      const double2 val = cscale(t1[c], weight);
      kernel[off] = accumulate ? cadd(kernel[off], val) : val;
    }
  }
}

}  // namespace dgami

// ---------------------------------------------------------------------------
// entry points
// ---------------------------------------------------------------------------
namespace {

using namespace dgami;

template <int T, int TS>
unsigned team_blocks(size_t nteam) {
  return ceil_div(nteam * TS, 256);
}

#define DGAMI_DISPATCH_NB(nb, MACRO) \
  switch (nb) {                      \
    case 1:                          \
      MACRO(1, 1);                   \
      break;                         \
    case 2:                          \
      MACRO(4, 4);                   \
      break;                         \
    case 3:                          \
      MACRO(9, 16);                  \
      break;                         \
    case 4:                          \
      MACRO(16, 16);                 \
      break;                         \
    default:                         \
      set_error("nb must be 1..4");  \
      return -1;                     \
  }

}  // namespace

extern "C" int dgami_chi0_prepare(int nq, int nb, int nw, int niv_full, int niv_core, double beta,
                                  const dgami_c128* chi0_, dgami_c128* chi0_inv_, dgami_c128* full_,
                                  dgami_c128* core_, void* stream_) {
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  DGAMI_ARG_CHECK(chi0_ != nullptr, "chi0_prepare: null chi0");
  DGAMI_ARG_CHECK(nq >= 1 && nw >= 1 && niv_core >= 1 && niv_full >= niv_core && beta > 0, "chi0_prepare: sizes");
  const double2* chi0 = reinterpret_cast<const double2*>(chi0_);
  const int nb4 = nb * nb * nb * nb;
  const int nvf = 2 * niv_full, nvc = 2 * niv_core, v0 = niv_full - niv_core;
  if (full_ || core_) {
    const size_t total = static_cast<size_t>(nw) * nb4 * nq;
    const unsigned blocks = ceil_div(total, 256) > 65536u ? 65536u : ceil_div(total, 256);
    hipLaunchKernelGGL(chi0_sums_kernel, dim3(blocks), dim3(256), 0, stream, chi0, reinterpret_cast<double2*>(full_),
                       reinterpret_cast<double2*>(core_), nq, nb4, nw, nvf, v0, nvc, 1.0 / (beta * beta));
    DGAMI_HIP_CHECK(hipGetLastError());
  }
  if (chi0_inv_) {
    double2* inv = reinterpret_cast<double2*>(chi0_inv_);
    const size_t nteam = static_cast<size_t>(nw) * nvc * nq;
#define LAUNCH(T, TS)                                                                                              \
  hipLaunchKernelGGL((chi0_inv_kernel<T, TS>), dim3(team_blocks<T, TS>(nteam)), dim3(256), 0, stream, chi0, inv, nq, \
                     nw, nvf, v0, nvc)
    DGAMI_DISPATCH_NB(nb, LAUNCH)
#undef LAUNCH
    DGAMI_HIP_CHECK(hipGetLastError());
  }
  return 0;
}

extern "C" int dgami_vrg_chi_phys(int nq, int nb, int nw, int niv, double beta, const dgami_c128* chi0_inv_,
                                  const dgami_c128* aux_, const dgami_c128* full_, const dgami_c128* core_,
                                  const dgami_c128* u_r_, const dgami_c128* v_r_, dgami_c128* vrg_,
                                  dgami_c128* chi_phys_, int32_t* info, void* stream_) {
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  DGAMI_ARG_CHECK(aux_ != nullptr, "vrg_chi_phys: null chi_aux_sum");
  DGAMI_ARG_CHECK(nq >= 1 && nw >= 1 && niv >= 1 && beta > 0, "vrg_chi_phys: sizes");
  const int nv = 2 * niv;
  const double2* aux = reinterpret_cast<const double2*>(aux_);
  if (vrg_) {
    DGAMI_ARG_CHECK(chi0_inv_ != nullptr, "vrg_chi_phys: vrg needs chi0_inv");
    const double2* inv = reinterpret_cast<const double2*>(chi0_inv_);
    double2* vrg = reinterpret_cast<double2*>(vrg_);
    const size_t nteam = static_cast<size_t>(nw) * nv * nq;
#define LAUNCH(T, TS) \
  hipLaunchKernelGGL((vrg_kernel<T, TS>), dim3(team_blocks<T, TS>(nteam)), dim3(256), 0, stream, inv, aux, vrg, nq, nw, nv, beta)
    DGAMI_DISPATCH_NB(nb, LAUNCH)
#undef LAUNCH
    DGAMI_HIP_CHECK(hipGetLastError());
  }
  if (chi_phys_) {
    DGAMI_ARG_CHECK(full_ != nullptr && core_ != nullptr && u_r_ != nullptr, "vrg_chi_phys: chi_phys inputs");
    const size_t nteam = static_cast<size_t>(nw) * nq;
    const double2* full = reinterpret_cast<const double2*>(full_);
    const double2* core = reinterpret_cast<const double2*>(core_);
    const double2* ur = reinterpret_cast<const double2*>(u_r_);
    const double2* vr = reinterpret_cast<const double2*>(v_r_);
    double2* out = reinterpret_cast<double2*>(chi_phys_);
#define LAUNCH(T, TS)                                                                                                 \
  hipLaunchKernelGGL((chi_phys_kernel<T, TS>), dim3(team_blocks<T, TS>(nteam)), dim3(256), 0, stream, aux, full, core, \
                     ur, vr, out, info, nq, nw, nv, 1.0 / beta)
    DGAMI_DISPATCH_NB(nb, LAUNCH)
#undef LAUNCH
    DGAMI_HIP_CHECK(hipGetLastError());
  }
  return 0;
}

extern "C" int dgami_kernel_r(int nq, int nb, int nw, int niv, int channel, double weight, const dgami_c128* vrg_,
                              const dgami_c128* chi_phys_, const dgami_c128* u_r_, const dgami_c128* v_r_,
                              int accumulate, dgami_c128* kernel_, void* stream_) {
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  DGAMI_ARG_CHECK(vrg_ && chi_phys_ && u_r_ && kernel_, "kernel_r: null pointer");
  DGAMI_ARG_CHECK(channel == 0 || channel == 1, "kernel_r: channel must be 0 (dens) or 1 (magn)");
  DGAMI_ARG_CHECK(nq >= 1 && nw >= 1 && niv >= 1, "kernel_r: sizes");
  const int nv = 2 * niv;
  const size_t nteam = static_cast<size_t>(nw) * nv * nq;
  const double2* vrg = reinterpret_cast<const double2*>(vrg_);
  const double2* chi = reinterpret_cast<const double2*>(chi_phys_);
  const double2* ur = reinterpret_cast<const double2*>(u_r_);
  const double2* vr = reinterpret_cast<const double2*>(v_r_);
  double2* ker = reinterpret_cast<double2*>(kernel_);
#define LAUNCH(T, TS)                                                                                               \
  hipLaunchKernelGGL((kernel_r_kernel<T, TS>), dim3(team_blocks<T, TS>(nteam)), dim3(256), 0, stream, vrg, chi, ur, \
                     vr, ker, nq, nw, nv, channel, weight, accumulate)
  DGAMI_DISPATCH_NB(nb, LAUNCH)
#undef LAUNCH
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}
