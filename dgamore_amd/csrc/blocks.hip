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
#include "team.hpp"

namespace dgami {

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

// ---------------------------------------------------------------------------
// G(k, v) = [(i v + mu) 1 - e_k - Sigma(k, v)]^-1  (team per (k, v), T = nb)
// ---------------------------------------------------------------------------
namespace dgami {
template <int T, int TS>
__global__ void __launch_bounds__(256) g_full_kernel(const double2* __restrict__ ek, const double2* __restrict__ sigma,
                                                    double2* __restrict__ g, int nk, int nv, int sigma_nk, int s_nv,
                                                    int s_v0, double mu_v, const double* __restrict__ mu_p,
                                                    double beta) {
  const double mu = mu_p != nullptr ? *mu_p : mu_v;
  const int lane = threadIdx.x % TS;
  const size_t team = (blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x) / TS;
  const size_t nteam = static_cast<size_t>(nk) * nv;
  const bool active = team < nteam;
  const size_t tt = active ? team : 0;
  const int v = static_cast<int>(tt % nv);
  const int k = static_cast<int>(tt / nv);
  const int r = lane < T ? lane : 0;
  const double iv = M_PI / beta * (2 * (v - nv / 2) + 1);
  const int ks = sigma_nk == 1 ? 0 : k;
  double2 m[T];
#pragma unroll
  for (int c = 0; c < T; ++c) {
    const double2 e = ek[(static_cast<size_t>(k) * T + r) * T + c];
    const double2 sg = sigma[((static_cast<size_t>(ks) * T + r) * T + c) * s_nv + s_v0 + v];
    double2 x = make_double2(-e.x - sg.x, -e.y - sg.y);
    if (c == r) x = make_double2(x.x + mu, x.y + iv);
    m[c] = x;
  }
  team_invert<T, TS>(m, lane);
  if (active && lane < T) {
#pragma unroll
    for (int c = 0; c < T; ++c) g[((static_cast<size_t>(k) * T + r) * T + c) * nv + v] = m[c];
  }
}

// generic block inverse: in/out [T*T][nmat] (element-major, matrix index fastest)
template <int T, int TS>
__global__ void __launch_bounds__(256) block_invert_kernel(const double2* __restrict__ in, double2* __restrict__ out,
                                                          int32_t* info, int nmat) {
  const int lane = threadIdx.x % TS;
  const size_t team = (blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x) / TS;
  const bool active = team < static_cast<size_t>(nmat);
  const int q = active ? static_cast<int>(team) : 0;
  const int r = lane < T ? lane : 0;
  double2 m[T];
#pragma unroll
  for (int c = 0; c < T; ++c) m[c] = in[static_cast<size_t>(r * T + c) * nmat + q];
  const bool sing = team_invert<T, TS>(m, lane);
  if (active && lane < T) {
#pragma unroll
    for (int c = 0; c < T; ++c) out[static_cast<size_t>(r * T + c) * nmat + q] = m[c];
    if (lane == 0 && info) info[q] = sing ? 1 : 0;
  }
}
}  // namespace dgami

static int g_full_launch(int nk, int nb, int niv, double beta, double mu, const double* mu_p, const dgami_c128* ek_,
                         const dgami_c128* sigma_, int sigma_nk, int sigma_niv, dgami_c128* g_, void* stream_) {
  using namespace dgami;
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  DGAMI_ARG_CHECK(ek_ && sigma_ && g_, "g_full: null pointer");
  DGAMI_ARG_CHECK(nk >= 1 && nb >= 1 && nb <= 4 && niv >= 1 && beta > 0, "g_full: invalid sizes");
  DGAMI_ARG_CHECK(sigma_nk == 1 || sigma_nk == nk, "g_full: sigma must be local (1 point) or on the k-grid");
  DGAMI_ARG_CHECK(sigma_niv >= niv, "g_full: sigma box smaller than the requested G box");
  const double2* ek = reinterpret_cast<const double2*>(ek_);
  const double2* sg = reinterpret_cast<const double2*>(sigma_);
  double2* g = reinterpret_cast<double2*>(g_);
  const int nv = 2 * niv;
  const size_t nteam = static_cast<size_t>(nk) * nv;
#define LAUNCH(T, TS)                                                                                             \
  hipLaunchKernelGGL((g_full_kernel<T, TS>), dim3(team_blocks<T, TS>(nteam)), dim3(256), 0, stream, ek, sg, g, nk, \
                     nv, sigma_nk, 2 * sigma_niv, sigma_niv - niv, mu, mu_p, beta)
  switch (nb) {
    case 1: LAUNCH(1, 1); break;
    case 2: LAUNCH(2, 2); break;
    case 3: LAUNCH(3, 4); break;
    default: LAUNCH(4, 4); break;
  }
#undef LAUNCH
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}

extern "C" int dgami_g_full(int nk, int nb, int niv, double beta, double mu, const dgami_c128* ek,
                            const dgami_c128* sigma, int sigma_nk, int sigma_niv, dgami_c128* g, void* stream) {
  return g_full_launch(nk, nb, niv, beta, mu, nullptr, ek, sigma, sigma_nk, sigma_niv, g, stream);
}

extern "C" int dgami_g_full_dmu(int nk, int nb, int niv, double beta, const double* mu, const dgami_c128* ek,
                                const dgami_c128* sigma, int sigma_nk, int sigma_niv, dgami_c128* g, void* stream) {
  DGAMI_ARG_CHECK(mu != nullptr, "g_full_dmu: null mu");
  return g_full_launch(nk, nb, niv, beta, 0.0, mu, ek, sigma, sigma_nk, sigma_niv, g, stream);
}

extern "C" int dgami_block_invert(int nmat, int nb, const dgami_c128* in_, dgami_c128* out_, int32_t* info,
                                  void* stream_) {
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  DGAMI_ARG_CHECK(in_ && out_, "block_invert: null pointer");
  DGAMI_ARG_CHECK(nmat >= 1, "block_invert: nmat must be positive");
  const double2* in = reinterpret_cast<const double2*>(in_);
  double2* out = reinterpret_cast<double2*>(out_);
#define LAUNCH(T, TS)                                                                                        \
  hipLaunchKernelGGL((block_invert_kernel<T, TS>), dim3(team_blocks<T, TS>(nmat)), dim3(256), 0, stream, in, \
                     out, info, nmat)
  DGAMI_DISPATCH_NB(nb, LAUNCH)
#undef LAUNCH
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}
