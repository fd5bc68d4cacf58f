// Local Schwinger-Dyson self-energy from the local three-leg vertices and susceptibilities of the
// density and magnetic channels (local_sde.py:150-172):
//   inner_r(w, v) = vrg_r - vrg_r . U_r . chi_r                    (products over orbital pairs)
//   inner(w, v)   = (inner_dens - inner_magn) / 2, extended to w < 0 by inner(-w, v) = conj inner(w, -v)
//   Sigma_ij(v)   = -(1/beta) sum_{w = -niw..niw} sum_klop U_kjop inner_ilpo(w, v) G_lk(v - w)
// plus, when the occupation n is given, the Hartree-Fock constant sum_cd U_dens,ijcd n_dc. One
// workgroup per fermionic frequency; per bosonic frequency the nb^4 entries of inner are formed in
// LDS, then nb^2 threads accumulate.
#include <algorithm>

#include "dgami.h"
#include "dgami_internal.hpp"

namespace dgami {

constexpr int LSIG_MAX_T = 16;  // nb^2 <= 16 (up to four orbitals)

// inner_r(w, v) of one channel into s_out[nb^4] (M order), w >= 0, frequency index vi of 2 niv.
// s_m: scratch [T][T] for U_r . chi_r(w).
__device__ void local_inner(const double2* __restrict__ vrg, const double2* __restrict__ chi,
                            const double2* __restrict__ u_m, int T, int nv, int w, int vi, double2 (*s_m)[LSIG_MAX_T],
                            double2* s_out, double sign) {
  const int TT = T * T;
  // s_m = U_r . chi_r(w): (U_r)_{(a b),(d c)} (chi_r)_{(d c),(f e)}, both in M order
  for (int e = threadIdx.x; e < TT; e += blockDim.x) {
    const int r = e / T, c = e - (e / T) * T;
    double2 acc = make_double2(0.0, 0.0);
    for (int k = 0; k < T; ++k) acc = cfma(acc, u_m[r * T + k], chi[static_cast<size_t>(w) * TT + k * T + c]);
    s_m[r][c] = acc;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < TT; e += blockDim.x) {
    const int r = e / T, c = e - (e / T) * T;
    const double2* vw = vrg + static_cast<size_t>(w) * TT * nv;  // QF with nq = 1: [w][TT][nv]
    double2 acc = vw[static_cast<size_t>(e) * nv + vi];
    for (int k = 0; k < T; ++k) acc = csub(acc, cmul(vw[static_cast<size_t>(r * T + k) * nv + vi], s_m[k][c]));
    s_out[e] = cadd(s_out[e], cscale(acc, sign));
  }
  __syncthreads();
}

__global__ void __launch_bounds__(256) local_sigma_kernel(const double2* __restrict__ vrg_d,
                                                          const double2* __restrict__ vrg_m,
                                                          const double2* __restrict__ chi_d,
                                                          const double2* __restrict__ chi_m,
                                                          const double2* __restrict__ ud_m,
                                                          const double2* __restrict__ um_m,
                                                          const double2* __restrict__ u, const double2* __restrict__ g,
                                                          const double2* __restrict__ occ,
                                                          double2* __restrict__ sigma, int nb, int niw, int niv,
                                                          int niv_g, double beta) {
  const int vi = blockIdx.x;  // fermionic index of 2 niv: v = vi - niv
  const int T = nb * nb, TT = T * T, nv = 2 * niv;
  __shared__ double2 s_m[LSIG_MAX_T][LSIG_MAX_T];
  __shared__ double2 s_inner[LSIG_MAX_T * LSIG_MAX_T];
  double2 acc = make_double2(0.0, 0.0);
  const int t = threadIdx.x;
  const int i = t / nb, j = t - (t / nb) * nb;
  for (int w = -niw; w <= niw; ++w) {
    // inner at |w|; for w < 0 at the mirrored fermionic index, conjugated below
    const int aw = w < 0 ? -w : w, vv = w < 0 ? nv - 1 - vi : vi;
    for (int e = t; e < TT; e += blockDim.x) s_inner[e] = make_double2(0.0, 0.0);
    __syncthreads();
    local_inner(vrg_d, chi_d, ud_m, T, nv, aw, vv, s_m, s_inner, 0.5);
    local_inner(vrg_m, chi_m, um_m, T, nv, aw, vv, s_m, s_inner, -0.5);
    if (t < T) {
      const int gi = niv_g + (vi - niv) - w;  // G at v - w
      for (int k = 0; k < nb; ++k)
        for (int l = 0; l < nb; ++l) {
          const double2 gl = g[(static_cast<size_t>(l) * nb + k) * 2 * niv_g + gi];
          for (int o = 0; o < nb; ++o)
            for (int p = 0; p < nb; ++p) {
              double2 in = s_inner[((i * nb + l) * nb + o) * nb + p];  // inner_ilpo in M order
              if (w < 0) in = cconj(in);
              acc = cfma(acc, cmul(u[((k * nb + j) * nb + o) * nb + p], in), gl);
            }
        }
    }
    __syncthreads();
  }
  if (t < T) {
    acc = cscale(acc, -1.0 / beta);
    if (occ != nullptr)  // U_dens,ijcd in M order sits at ((i, j), (d, c))
      for (int d = 0; d < nb; ++d)
        for (int c = 0; c < nb; ++c) acc = cfma(acc, ud_m[(t * nb + d) * nb + c], occ[d * nb + c]);
    sigma[static_cast<size_t>(t) * nv + vi] = acc;
  }
}

// chi_r = beta g2 (w >= 0); density channel: chi_r(w = 0) -= 2 beta G_ab(v) G_cd(v') (local_sde.py:17-29).
// Both in CMP layout [nw][(a, b, v)][(d, c, v')]; g: [nb][nb][2 niv_g].
__global__ void __launch_bounds__(256) generalized_chi_kernel(const double2* __restrict__ g2,
                                                              const double2* __restrict__ g,
                                                              double2* __restrict__ chi, int nb, int nw, int niv,
                                                              int niv_g, double beta, int dens) {
  const int nv = 2 * niv, n = nb * nb * nv;
  const size_t total = static_cast<size_t>(nw) * n * n;
  for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < total;
       e += static_cast<size_t>(gridDim.x) * blockDim.x) {
    double2 x = cscale(g2[e], beta);
    if (dens && e < static_cast<size_t>(n) * n) {  // w = 0
      const int r = static_cast<int>(e / n), c = static_cast<int>(e - static_cast<size_t>(r) * n);
      const int ab = r / nv, v = r - ab * nv, dc = c / nv, vp = c - dc * nv;
      const int a = ab / nb, b = ab - a * nb, d = dc / nb, cc = dc - d * nb;
      const double2 gab = g[(static_cast<size_t>(a) * nb + b) * 2 * niv_g + niv_g - niv + v];
      const double2 gcd = g[(static_cast<size_t>(cc) * nb + d) * 2 * niv_g + niv_g - niv + vp];
      x = csub(x, cscale(cmul(gab, gcd), 2.0 * beta));
    }
    chi[e] = x;
  }
}

// Gamma_r = beta^2 (chi_inv - chit_inv) + U_r on the core box [nw][nc][nc] (chit_inv NULL: chi_inv is
// Gamma_r itself); optionally the vertex padded with U_r to the full box [nw][nf][nf]
// (local_sde.py:40-51, local_four_point.py:652-668).
__global__ void __launch_bounds__(256) local_gamma_kernel(const double2* __restrict__ chi_inv,
                                                          const double2* __restrict__ chit_inv,
                                                          const double2* __restrict__ u_m,
                                                          double2* __restrict__ gamma, double2* __restrict__ pad,
                                                          int nb, int nw, int nvc, int nvf, double beta) {
  const int T = nb * nb, nc = T * nvc, nf = T * nvf, off = (nvf - nvc) / 2;
  const size_t total = static_cast<size_t>(nw) * nf * nf;
  const double b2 = beta * beta;
  for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < total;
       e += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int w = static_cast<int>(e / (static_cast<size_t>(nf) * nf));
    const size_t rc = e - static_cast<size_t>(w) * nf * nf;
    const int r = static_cast<int>(rc / nf), c = static_cast<int>(rc - static_cast<size_t>(r) * nf);
    const int ab = r / nvf, v = r - ab * nvf - off, dc = c / nvf, vp = c - dc * nvf - off;
    double2 x = u_m[ab * T + dc];
    if (v >= 0 && v < nvc && vp >= 0 && vp < nvc) {
      const size_t ec = (static_cast<size_t>(w) * nc + ab * nvc + v) * nc + dc * nvc + vp;
      x = chit_inv != nullptr ? cadd(x, cscale(csub(chi_inv[ec], chit_inv[ec]), b2)) : chi_inv[ec];
      if (gamma != nullptr) gamma[ec] = x;
    }
    if (pad != nullptr) pad[e] = x;
  }
}

}  // namespace dgami

extern "C" int dgami_generalized_chi(int nb, int nw, int niv, int niv_g, double beta, int dens,
                                     const dgami_c128* g2, const dgami_c128* g, dgami_c128* chi, void* stream) {
  using namespace dgami;
  DGAMI_ARG_CHECK(g2 && chi && (g || !dens), "generalized_chi: null pointer");
  DGAMI_ARG_CHECK(nb >= 1 && nw >= 1 && niv >= 1 && beta > 0, "generalized_chi: invalid sizes");
  DGAMI_ARG_CHECK(!dens || niv_g >= niv, "generalized_chi: G must hold the core frequencies");
  const size_t total = static_cast<size_t>(nw) * nb * nb * 2 * niv * nb * nb * 2 * niv;
  const int blocks = static_cast<int>(std::min<size_t>((total + 255) / 256, 8192));
  hipLaunchKernelGGL(generalized_chi_kernel, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                     reinterpret_cast<const double2*>(g2), reinterpret_cast<const double2*>(g),
                     reinterpret_cast<double2*>(chi), nb, nw, niv, niv_g, beta, dens);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}

extern "C" int dgami_local_gamma(int nb, int nw, int niv_core, int niv_full, double beta, const dgami_c128* chi_inv,
                                 const dgami_c128* chit_inv, const dgami_c128* u_r, dgami_c128* gamma,
                                 dgami_c128* gamma_pad, void* stream) {
  using namespace dgami;
  DGAMI_ARG_CHECK(chi_inv && u_r && (gamma || gamma_pad), "local_gamma: null pointer");
  DGAMI_ARG_CHECK(gamma == nullptr || (gamma != chi_inv && gamma != chit_inv),
                  "local_gamma: gamma may not alias its inputs");
  DGAMI_ARG_CHECK(nb >= 1 && nw >= 1 && niv_core >= 1 && niv_full >= niv_core && beta > 0,
                  "local_gamma: invalid sizes");
  const int nvf = gamma_pad ? 2 * niv_full : 2 * niv_core;
  const size_t nf = static_cast<size_t>(nb) * nb * nvf;
  const size_t total = static_cast<size_t>(nw) * nf * nf;
  const int blocks = static_cast<int>(std::min<size_t>((total + 255) / 256, 8192));
  hipLaunchKernelGGL(local_gamma_kernel, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                     reinterpret_cast<const double2*>(chi_inv), reinterpret_cast<const double2*>(chit_inv),
                     reinterpret_cast<const double2*>(u_r), reinterpret_cast<double2*>(gamma),
                     reinterpret_cast<double2*>(gamma_pad), nb, nw, 2 * niv_core, nvf, beta);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}

extern "C" int dgami_local_sigma(int nb, int niw, int niv, int niv_g, double beta, const dgami_c128* vrg_dens,
                                 const dgami_c128* vrg_magn, const dgami_c128* chi_dens, const dgami_c128* chi_magn,
                                 const dgami_c128* u_dens, const dgami_c128* u_magn, const dgami_c128* u,
                                 const dgami_c128* g, const dgami_c128* occ, dgami_c128* sigma, void* stream) {
  using namespace dgami;
  DGAMI_ARG_CHECK(vrg_dens && vrg_magn && chi_dens && chi_magn && u_dens && u_magn && u && g && sigma,
                  "local_sigma: null pointer");
  DGAMI_ARG_CHECK(nb >= 1 && nb <= 4 && niw >= 0 && niv >= 1 && beta > 0, "local_sigma: invalid sizes");
  DGAMI_ARG_CHECK(niv_g >= niv + niw, "local_sigma: G must hold at least niv + niw fermionic frequencies");
  hipLaunchKernelGGL(local_sigma_kernel, dim3(2 * niv), dim3(64), 0, static_cast<hipStream_t>(stream),
                     reinterpret_cast<const double2*>(vrg_dens), reinterpret_cast<const double2*>(vrg_magn),
                     reinterpret_cast<const double2*>(chi_dens), reinterpret_cast<const double2*>(chi_magn),
                     reinterpret_cast<const double2*>(u_dens), reinterpret_cast<const double2*>(u_magn),
                     reinterpret_cast<const double2*>(u), reinterpret_cast<const double2*>(g),
                     reinterpret_cast<const double2*>(occ), reinterpret_cast<double2*>(sigma), nb, niw, niv, niv_g,
                     beta);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}
