// Self-consistency pieces around the Schwinger-Dyson sum (complex128):
//   * Hartree and Fock terms of the interaction (the Fock term as an FFT cross-correlation
//     over the momentum grid instead of a roll per q);
//   * assembly of the new self-energy (non-local part + Hartree + Fock + the local
//     correction, DMFT self-energy outside the core window) and its linear mixing;
//   * the Matsubara sums of the occupations that fix the chemical potential.
#include "dgami.h"
#include "dgami_internal.hpp"
#include "dgami_launch.hpp"

namespace dgami {

// ---------------------------------------------------------------------------
// H_ab = 2 sum_cd (U + V^{q=0})_abcd n_dc. u, v: reference orbital order abcd; v: [nk][nb^4]
// (the q = 0 entry is the first). One thread per (a, b).
__global__ void hf_hartree_kernel(const double2* __restrict__ u, const double2* __restrict__ v,
                                  const double2* __restrict__ occ, double2* __restrict__ h, int nb) {
  const int e = threadIdx.x;
  if (e >= nb * nb) return;
  const int a = e / nb, b = e % nb;
  double2 acc = make_double2(0.0, 0.0);
  for (int c = 0; c < nb; ++c)
    for (int d = 0; d < nb; ++d) {
      const int idx = ((a * nb + b) * nb + c) * nb + d;
      double2 w = u[idx];
      if (v) w = cadd(w, v[idx]);
      acc = cfma(acc, w, occ[d * nb + c]);
    }
  h[e] = cscale(acc, 2.0);
}

// wq[abcd][k] = U_abcd + V_abcd(k); nq[dc][k] = n_dc(k) (orbital-major rows for the batched FFTs).
__global__ void hf_stage_kernel(const double2* __restrict__ u, const double2* __restrict__ v,
                                const double2* __restrict__ occ_k, double2* __restrict__ wq, double2* __restrict__ nq,
                                int nk, int nb) {
  const int nb2 = nb * nb, nb4 = nb2 * nb2;
  const size_t total = static_cast<size_t>(nk) * (nb4 + nb2);
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int k = static_cast<int>(i % nk);
    const int r = static_cast<int>(i / nk);
    if (r < nb4) {
      double2 w = u[r];
      if (v) w = cadd(w, v[static_cast<size_t>(k) * nb4 + r]);
      wq[static_cast<size_t>(r) * nk + k] = w;
    } else {
      const int dc = r - nb4;
      nq[static_cast<size_t>(dc) * nk + k] = occ_k[static_cast<size_t>(k) * nb2 + dc];
    }
  }
}

// p[ab][R] = sum_cd wb[abcd][R] nf[dc][R]
__global__ void hf_product_kernel(const double2* __restrict__ wb, const double2* __restrict__ nf,
                                  double2* __restrict__ p, int nk, int nb) {
  const int nb2 = nb * nb;
  const size_t total = static_cast<size_t>(nk) * nb2;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int R = static_cast<int>(i % nk);
    const int ab = static_cast<int>(i / nk);
    double2 acc = make_double2(0.0, 0.0);
    for (int c = 0; c < nb; ++c)
      for (int d = 0; d < nb; ++d)
        acc = cfma(acc, wb[(static_cast<size_t>(ab) * nb2 + c * nb + d) * nk + R],
                   nf[(static_cast<size_t>(d) * nb + c) * nk + R]);
    p[i] = acc;
  }
}

// fock[k][ab] = -p[ab][k]
__global__ void hf_fock_kernel(const double2* __restrict__ p, double2* __restrict__ fock, int nk, int nb2) {
  const size_t total = static_cast<size_t>(nk) * nb2;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int k = static_cast<int>(i / nb2), ab = static_cast<int>(i % nb2);
    const double2 x = p[static_cast<size_t>(ab) * nk + k];
    fock[i] = make_double2(-x.x, -x.y);
  }
}

// ---------------------------------------------------------------------------
// out[k][a][b][v], v over 2*niv_out: inside the core window (off = niv_out - niv_core)
//   val = sigma_nl[k][a][b][v-off] + H_ab + F_ab(k) + dsig[a][b][v-off]
// outside it the DMFT self-energy (centred cut of 2*niv_dmft frequencies); then
//   out = alpha val + (1 - alpha) old  (old given) or val.
__global__ void sigma_assemble_kernel(int nk, int nb, int niv_core, int niv_out, const double2* __restrict__ snl,
                                      const double2* __restrict__ hart, const double2* __restrict__ fock,
                                      const double2* __restrict__ dsig, const double2* __restrict__ sdmft,
                                      int niv_dmft, const double2* __restrict__ old, double alpha,
                                      double2* __restrict__ out) {
  const int nb2 = nb * nb, nvo = 2 * niv_out, nvc = 2 * niv_core, off = niv_out - niv_core;
  const size_t total = static_cast<size_t>(nk) * nb2 * nvo;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int v = static_cast<int>(i % nvo);
    const size_t kab = i / nvo;
    const int ab = static_cast<int>(kab % nb2);
    const int k = static_cast<int>(kab / nb2);
    double2 val;
    const int vc = v - off;
    if (vc >= 0 && vc < nvc) {
      val = snl[kab * nvc + vc];
      if (hart) val = cadd(val, hart[ab]);
      if (fock) val = cadd(val, fock[kab]);
      if (dsig) val = cadd(val, dsig[static_cast<size_t>(ab) * nvc + vc]);
    } else {
      val = sdmft[static_cast<size_t>(ab) * 2 * niv_dmft + (niv_dmft - niv_out) + v];
    }
    if (old) val = cadd(cscale(val, alpha), cscale(old[i], 1.0 - alpha));
    out[i] = val;
  }
}

// ---------------------------------------------------------------------------
// In-register inverse of an NB x NB complex matrix (Gauss-Jordan, partial pivoting).
template <int NB>
__device__ __forceinline__ void small_inverse(double2 (&m)[NB][NB], double2 (&x)[NB][NB]) {
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) x[i][j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    // pivot row: largest |Re| + |Im| in column c at or below the diagonal, swapped in place
#pragma unroll
    for (int r = c + 1; r < NB; ++r) {
      if (cabs1(m[r][c]) > cabs1(m[c][c])) {
#pragma unroll
        for (int j = 0; j < NB; ++j) {
          const double2 t = m[c][j];
          m[c][j] = m[r][j];
          m[r][j] = t;
          const double2 s = x[c][j];
          x[c][j] = x[r][j];
          x[r][j] = s;
        }
      }
    }
    const double2 p = cinv(m[c][c]);
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      m[c][j] = cmul(m[c][j], p);
      x[c][j] = cmul(x[c][j], p);
    }
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      if (r == c) continue;
      const double2 f = m[r][c];
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        m[r][j] = cfms(m[r][j], f, m[c][j]);
        x[r][j] = cfms(x[r][j], f, x[c][j]);
      }
    }
  }
}

// occ[k][a][b] = (1/beta) sum_v Re[G(k,v) - G_model(k,v)]_ab over v in [-niv, niv), with
//   G = (i v + mu - e_k - Sigma(k,v))^-1,  G_model = (i v + mu - h - smom0)^-1,
// h = e_k (hloc == nullptr) or hloc. One workgroup per k, 256 threads over the frequencies.
// mu_p (or NULL): mu is read from device memory; skip (or NULL): nonzero makes the launch a no-op.
template <int NB>
__global__ void __launch_bounds__(256) fill_kernel(int niv, double beta, double mu_v, const double2* __restrict__ ek,
                                                   const double2* __restrict__ sigma, int sigma_nk,
                                                   const double2* __restrict__ smom0, const double2* __restrict__ hloc,
                                                   double* __restrict__ occ, const double* __restrict__ mu_p,
                                                   const int* __restrict__ skip) {
  if (skip != nullptr && *skip != 0) return;
  const double mu = mu_p != nullptr ? *mu_p : mu_v;
  const int k = blockIdx.x;
  const int nv = 2 * niv;
  const double2* E = ek + static_cast<size_t>(k) * NB * NB;
  const double2* S = sigma + static_cast<size_t>(sigma_nk == 1 ? 0 : k) * NB * NB * nv;
  const double2* H = hloc ? hloc : E;
  double acc[NB][NB];
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = 0.0;
  for (int v = threadIdx.x; v < nv; v += blockDim.x) {
    const double iv = M_PI / beta * (2 * (v - niv) + 1);
    double2 m[NB][NB], g[NB][NB];
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const double2 diag = make_double2(a == b ? mu : 0.0, a == b ? iv : 0.0);
        m[a][b] = csub(csub(diag, E[a * NB + b]), S[(static_cast<size_t>(a) * NB + b) * nv + v]);
      }
    small_inverse<NB>(m, g);
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        acc[a][b] += g[a][b].x;
        const double2 diag = make_double2(a == b ? mu : 0.0, a == b ? iv : 0.0);
        m[a][b] = csub(csub(diag, H[a * NB + b]), smom0[a * NB + b]);
      }
    small_inverse<NB>(m, g);
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b) acc[a][b] -= g[a][b].x;
  }
  __shared__ double s_red[NB * NB][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      double x = acc[a][b];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
      if (lane == 0) s_red[a * NB + b][wave] = x;
    }
  __syncthreads();
  if (threadIdx.x < NB * NB) {
    const int nw = blockDim.x >> 6;
    double x = 0.0;
    for (int w = 0; w < nw; ++w) x += s_red[threadIdx.x][w];
    occ[static_cast<size_t>(k) * NB * NB + threadIdx.x] = x / beta;
  }
}

// out[e] = (1/nk) sum_k x[k][e] for e < m (real or complex data as `width` doubles per element).
__global__ void kmean_kernel(int nk, size_t m, const double* __restrict__ x, double* __restrict__ out) {
  for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < m;
       e += static_cast<size_t>(gridDim.x) * blockDim.x) {
    double acc = 0.0;
    for (int k = 0; k < nk; ++k) acc += x[static_cast<size_t>(k) * m + e];
    out[e] = acc / nk;
  }
}

// ---------------------------------------------------------------------------
// Chemical potential on the device.
//
// Fermi function of a real symmetric nb x nb matrix through its eigen-decomposition by cyclic
// Jacobi rotations, with the overflow-safe form for positive eigenvalues: f = V diag(rho) V^T,
// rho = 1 / (1 + exp(lambda)) (greens_function.py:38-45 takes numpy's eig of the same matrix).
__device__ __forceinline__ double fermi_of(double x) {
  if (x > 0.0) {
    const double e = exp(-x);
    return e / (1.0 + e);
  }
  return 1.0 / (1.0 + exp(x));
}

template <int NB>
__device__ void fermi_symmetric(double (&a)[NB][NB], double (&f)[NB][NB]) {
  double v[NB][NB];
  double fro = 0.0;
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      v[i][j] = i == j ? 1.0 : 0.0;
      fro += a[i][j] * a[i][j];
    }
  for (int sweep = 0; sweep < 16; ++sweep) {
    double off = 0.0;
#pragma unroll
    for (int p = 0; p < NB; ++p)
#pragma unroll
      for (int q = p + 1; q < NB; ++q) off += a[p][q] * a[p][q];
    if (off <= 1e-34 * fro) break;
#pragma unroll
    for (int p = 0; p < NB; ++p)
#pragma unroll
      for (int q = p + 1; q < NB; ++q) {
        if (a[p][q] == 0.0) continue;
        // rotation in the (p, q) plane that annihilates a[p][q] (Golub & Van Loan, sec. 8.5)
        const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        const double t = fabs(theta) > 1e150 ? 0.5 / theta
                                             : (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
        for (int k = 0; k < NB; ++k) {
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
#pragma unroll
        for (int k = 0; k < NB; ++k) {
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
#pragma unroll
        for (int k = 0; k < NB; ++k) {
          const double vkp = v[k][p], vkq = v[k][q];
          v[k][p] = c * vkp - s * vkq;
          v[k][q] = s * vkp + c * vkq;
        }
      }
  }
  double rho[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) rho[i] = fermi_of(a[i][i]);
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      double x = 0.0;
#pragma unroll
      for (int k = 0; k < NB; ++k) x += v[i][k] * rho[k] * v[j][k];
      f[i][j] = x;
    }
}

// beta (Re h + (smom0 + smom0^T)/2 - mu), the argument of the model's density matrix. smom0 is
// symmetrised: the moment of a symmetric Sigma(v) is symmetric, and the asymmetry is reported.
template <int NB>
__device__ double model_matrix(const double2* __restrict__ h, const double2* __restrict__ smom0, double mu,
                               double beta, double (&a)[NB][NB]) {
  double asym = 0.0;
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const double s = 0.5 * (smom0[i * NB + j].x + smom0[j * NB + i].x);
      asym = fmax(asym, fabs(smom0[i * NB + j].x - smom0[j * NB + i].x));
      a[i][j] = beta * (0.5 * (h[i * NB + j].x + h[j * NB + i].x) + s - (i == j ? mu : 0.0));
    }
  return asym;
}

// Secant iteration on mu (the derivative-free branch of scipy.optimize.newton that
// update_mu calls, greens_function.py:63-80: tol 1e-6 on |p - p1|, 50 iterations), kept on the
// device: every evaluation is a fill_kernel launch at x followed by one mu_secant_kernel launch.
struct MuSecant {
  double p0, p1, q0, q1, x, result, asym;
  int stage, itr, done, status;  // status: 0 running, 1 converged, 2 failed
};

constexpr int MU_MAXITER = 50;

__global__ void mu_secant_init_kernel(MuSecant* st, const double* mu) {
  const double x0 = *mu;
  st->p0 = x0;
  double p1 = x0 * (1.0 + 1e-4);
  p1 += p1 >= 0.0 ? 1e-4 : -1e-4;
  st->p1 = p1;
  st->q0 = st->q1 = 0.0;
  st->x = x0;
  st->result = x0;
  st->asym = 0.0;
  st->stage = st->itr = st->done = st->status = 0;
}

// One WG: the k-mean of the Matsubara occupations at st->x plus the model's density matrix of
// the k-mean Hamiltonian hloc, q = 2 Tr occ - target, then one step of the secant.
template <int NB>
__global__ void __launch_bounds__(256) mu_secant_kernel(MuSecant* st, const double* __restrict__ occ_k, int nk,
                                                        const double2* __restrict__ hloc,
                                                        const double2* __restrict__ smom0, double beta,
                                                        double target) {
  if (st->done) return;
  __shared__ double s_red[256];
  double tr = 0.0;
  for (int e = threadIdx.x; e < nk * NB; e += blockDim.x) {  // diagonal entries of every k
    const int k = e / NB, a = e - k * NB;
    tr += occ_k[static_cast<size_t>(k) * NB * NB + a * NB + a];
  }
  s_red[threadIdx.x] = tr;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) s_red[threadIdx.x] += s_red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  double a[NB][NB], f[NB][NB];
  st->asym = fmax(st->asym, model_matrix<NB>(hloc, smom0, st->x, beta, a));
  fermi_symmetric<NB>(a, f);
  double ftr = 0.0;
#pragma unroll
  for (int i = 0; i < NB; ++i) ftr += f[i][i];
  const double q = 2.0 * (ftr + s_red[0] / nk) - target;
  if (st->stage == 0) {  // q0 at p0; next the first value at p1
    st->q0 = q;
    st->stage = 1;
    st->x = st->p1;
    return;
  }
  double p0 = st->p0, p1 = st->p1, q0 = st->q0, q1 = q;
  if (st->stage == 1) {
    if (fabs(q1) < fabs(q0)) {
      const double tp = p0, tq = q0;
      p0 = p1;
      q0 = q1;
      p1 = tp;
      q1 = tq;
    }
    st->stage = 2;
  }
  double p;
  if (q1 == q0) {
    if (p1 != p0) {  // the reference raises here, and update_mu returns nan
      st->status = 2;
      st->done = 1;
      return;
    }
    p = 0.5 * (p1 + p0);
    st->result = p;
    st->status = 1;
    st->done = 1;
    return;
  }
  if (fabs(q1) > fabs(q0))
    p = (-q0 / q1 * p1 + p0) / (1.0 - q0 / q1);
  else
    p = (-q1 / q0 * p0 + p1) / (1.0 - q1 / q0);
  if (fabs(p - p1) <= 1e-6) {
    st->result = p;
    st->status = 1;
    st->done = 1;
    return;
  }
  if (++st->itr >= MU_MAXITER) {  // no convergence in the reference's 50 iterations
    st->status = 2;
    st->done = 1;
    return;
  }
  st->p0 = p1;
  st->q0 = q1;
  st->p1 = p;
  st->x = p;
}

// mu <- mixing mu_new + (1 - mixing) mu (kept when the secant failed); info: [0] mu_new (nan on
// failure), [1] status, [2] the largest asymmetry |smom0 - smom0^T| met.
__global__ void mu_secant_finish_kernel(const MuSecant* st, double* mu, double mixing, double* info) {
  const double old = *mu;
  const bool ok = st->status == 1;
  if (ok) *mu = mixing * st->result + (1.0 - mixing) * old;
  if (info != nullptr) {
    info[0] = ok ? st->result : __builtin_nan("");
    info[1] = st->status;
    info[2] = st->asym;
  }
}

// occ_k (complex [nk][nb][nb]) = Fermi(beta (Re e_k + smom0 - mu)) + Matsubara part, entries below
// 1e-12 set to zero (greens_function.py:212-240); one thread per k.
template <int NB>
__global__ void occ_k_kernel(const double* __restrict__ occ_sum, const double2* __restrict__ ek,
                             const double2* __restrict__ smom0, const double* __restrict__ mu, double beta, int nk,
                             double2* __restrict__ occ_k) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nk) return;
  double a[NB][NB], f[NB][NB];
  model_matrix<NB>(ek + static_cast<size_t>(k) * NB * NB, smom0, *mu, beta, a);
  fermi_symmetric<NB>(a, f);
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      double x = f[i][j] + occ_sum[static_cast<size_t>(k) * NB * NB + i * NB + j];
      if (fabs(x) < 1e-12) x = 0.0;
      occ_k[static_cast<size_t>(k) * NB * NB + i * NB + j] = make_double2(x, 0.0);
    }
}

// occ = k-mean of occ_k with entries below 1e-12 set to zero; one thread per entry.
__global__ void occ_mean_kernel(const double2* __restrict__ occ_k, int nk, int nb2, double2* __restrict__ occ) {
  const int e = threadIdx.x;
  if (e >= nb2) return;
  double x = 0.0;
  for (int k = 0; k < nk; ++k) x += occ_k[static_cast<size_t>(k) * nb2 + e].x;
  x /= nk;
  if (fabs(x) < 1e-12) x = 0.0;
  occ[e] = make_double2(x, 0.0);
}

// Local moments of Sigma from its k-mean: mom0 = mean Re Sigma_loc, mom1 = mean Im Sigma_loc * v
// over the top n_fit positive frequencies (self_energy.py:51-75). One WG per (a, b).
__global__ void __launch_bounds__(256) fit_smom_kernel(const double2* __restrict__ sigma, int nk, int nb2, int nv,
                                                       int n_fit, double beta, double2* __restrict__ mom0,
                                                       double2* __restrict__ mom1) {
  const int e = blockIdx.x;
  const int niv = nv / 2;
  __shared__ double s0[256], s1[256];
  double a0 = 0.0, a1 = 0.0;
  for (int t = threadIdx.x; t < nk * n_fit; t += blockDim.x) {
    const int k = t / n_fit, j = t - k * n_fit;
    const int v = nv - n_fit + j;
    const double2 x = sigma[(static_cast<size_t>(k) * nb2 + e) * nv + v];
    const double iv = M_PI / beta * (2 * (v - niv) + 1);
    a0 += x.x;
    a1 += x.y * iv;
  }
  s0[threadIdx.x] = a0;
  s1[threadIdx.x] = a1;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s0[threadIdx.x] += s0[threadIdx.x + o];
      s1[threadIdx.x] += s1[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double d = static_cast<double>(nk) * n_fit;
    mom0[e] = make_double2(s0[0] / d, 0.0);
    mom1[e] = make_double2(s1[0] / d, 0.0);
  }
}

static unsigned grid_for(size_t total) {
  const size_t b = (total + 255) / 256;
  return static_cast<unsigned>(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

static int launch_kmean(int nk, size_t m, const double* x, double* out, hipStream_t stream) {
  hipLaunchKernelGGL(kmean_kernel, dim3(grid_for(m)), dim3(256), 0, stream, nk, m, x, out);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace dgami

extern "C" int dgami_hartree_fock(int nx, int ny, int nz, int nb, const dgami_c128* u, const dgami_c128* v,
                                  const dgami_c128* occ, const dgami_c128* occ_k, dgami_c128* hartree,
                                  dgami_c128* fock, void* stream) {
  using namespace dgami;
  DGAMI_ARG_CHECK(u && occ && occ_k && hartree && fock, "hartree_fock: null pointer");
  DGAMI_ARG_CHECK(nx > 0 && ny > 0 && nz > 0 && nb >= 1 && nb <= 4, "hartree_fock: invalid sizes");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nk = nx * ny * nz, nb2 = nb * nb, nb4 = nb2 * nb2;
  const double2* U = reinterpret_cast<const double2*>(u);
  const double2* V = reinterpret_cast<const double2*>(v);
  hipLaunchKernelGGL(hf_hartree_kernel, dim3(1), dim3(64), 0, s, U, V, reinterpret_cast<const double2*>(occ),
                     reinterpret_cast<double2*>(hartree), nb);
  DGAMI_HIP_CHECK(hipGetLastError());
  double2* ws = static_cast<double2*>(workspace(WS_SELFCONS, static_cast<size_t>(nk) * (nb4 + 2 * nb2) * sizeof(double2)));
  if (!ws) return -2;
  double2* wq = ws;
  double2* nq = wq + static_cast<size_t>(nk) * nb4;
  double2* p = nq + static_cast<size_t>(nk) * nb2;
  hipLaunchKernelGGL(hf_stage_kernel, dim3(grid_for(static_cast<size_t>(nk) * (nb4 + nb2))), dim3(256), 0, s, U, V,
                     reinterpret_cast<const double2*>(occ_k), wq, nq, nk, nb);
  DGAMI_HIP_CHECK(hipGetLastError());
  // sum_q W(q) n(k+q) = (1/N) sum_R W(-R) n(R) e^{ikR}, W(-R) = sum_q W(q) e^{iqR}, n(R) = sum_k n(k) e^{-ikR}
  int rc = launch_fft_grid(wq, wq, nb4, nx, ny, nz, +1, 1.0, s);
  if (rc) return rc;
  if ((rc = launch_fft_grid(nq, nq, nb2, nx, ny, nz, -1, 1.0, s)) != 0) return rc;
  hipLaunchKernelGGL(hf_product_kernel, dim3(grid_for(static_cast<size_t>(nk) * nb2)), dim3(256), 0, s, wq, nq, p, nk,
                     nb);
  DGAMI_HIP_CHECK(hipGetLastError());
  // Fock = -(1/N_q) sum_q W(q) n(k+q): one 1/N from the inverse transform, one from the q average
  const double inv_n2 = 1.0 / (static_cast<double>(nk) * nk);
  if ((rc = launch_fft_grid(p, p, nb2, nx, ny, nz, +1, inv_n2, s)) != 0) return rc;
  hipLaunchKernelGGL(hf_fock_kernel, dim3(grid_for(static_cast<size_t>(nk) * nb2)), dim3(256), 0, s, p,
                     reinterpret_cast<double2*>(fock), nk, nb2);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}

extern "C" int dgami_sigma_assemble(int nk, int nb, int niv_core, int niv_out, const dgami_c128* sigma_nl,
                                    const dgami_c128* hartree, const dgami_c128* fock, const dgami_c128* delta_sigma,
                                    const dgami_c128* sigma_dmft, int niv_dmft, const dgami_c128* sigma_old,
                                    double alpha, dgami_c128* out, void* stream) {
  using namespace dgami;
  DGAMI_ARG_CHECK(sigma_nl && sigma_dmft && out, "sigma_assemble: null pointer");
  DGAMI_ARG_CHECK(nk >= 1 && nb >= 1 && nb <= 4 && niv_core >= 1 && niv_out >= niv_core && niv_dmft >= niv_out,
                  "sigma_assemble: need 1 <= niv_core <= niv_out <= niv_dmft");
  const size_t total = static_cast<size_t>(nk) * nb * nb * 2 * niv_out;
  hipLaunchKernelGGL(sigma_assemble_kernel, dim3(grid_for(total)), dim3(256), 0, static_cast<hipStream_t>(stream), nk,
                     nb, niv_core, niv_out, reinterpret_cast<const double2*>(sigma_nl),
                     reinterpret_cast<const double2*>(hartree), reinterpret_cast<const double2*>(fock),
                     reinterpret_cast<const double2*>(delta_sigma), reinterpret_cast<const double2*>(sigma_dmft),
                     niv_dmft, reinterpret_cast<const double2*>(sigma_old), alpha, reinterpret_cast<double2*>(out));
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}

extern "C" int dgami_fill_sums(int nk, int nb, int niv, double beta, double mu, const dgami_c128* ek,
                               const dgami_c128* sigma, int sigma_nk, const dgami_c128* smom0, const dgami_c128* hloc,
                               double* occ_k, double* occ_mean, void* stream) {
  using namespace dgami;
  DGAMI_ARG_CHECK(ek && sigma && smom0 && occ_k, "fill_sums: null pointer");
  DGAMI_ARG_CHECK(nk >= 1 && nb >= 1 && nb <= 4 && niv >= 1 && beta > 0 && (sigma_nk == 1 || sigma_nk == nk),
                  "fill_sums: invalid sizes");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const double2* E = reinterpret_cast<const double2*>(ek);
  const double2* S = reinterpret_cast<const double2*>(sigma);
  const double2* M0 = reinterpret_cast<const double2*>(smom0);
  const double2* H = reinterpret_cast<const double2*>(hloc);
#define DGAMI_FILL(NBV)                                                                                           \
  hipLaunchKernelGGL(fill_kernel<NBV>, dim3(nk), dim3(256), 0, s, niv, beta, mu, E, S, sigma_nk, M0, H, occ_k, nullptr, \
                     nullptr)
  switch (nb) {
    case 1: DGAMI_FILL(1); break;
    case 2: DGAMI_FILL(2); break;
    case 3: DGAMI_FILL(3); break;
    default: DGAMI_FILL(4); break;
  }
#undef DGAMI_FILL
  DGAMI_HIP_CHECK(hipGetLastError());
  if (occ_mean) return launch_kmean(nk, static_cast<size_t>(nb) * nb, occ_k, occ_mean, s);
  return 0;
}

extern "C" int dgami_kmean(int nk, int64_t m, const dgami_c128* x, dgami_c128* out, void* stream) {
  using namespace dgami;
  DGAMI_ARG_CHECK(x && out && nk >= 1 && m >= 1, "kmean: invalid arguments");
  return launch_kmean(nk, static_cast<size_t>(m) * 2, reinterpret_cast<const double*>(x), reinterpret_cast<double*>(out),
                      static_cast<hipStream_t>(stream));
}

extern "C" int dgami_fit_smom(int nk, int nb, int niv, double beta, const dgami_c128* sigma, dgami_c128* mom0,
                              dgami_c128* mom1, void* stream) {
  using namespace dgami;
  DGAMI_ARG_CHECK(sigma && mom0 && mom1, "fit_smom: null pointer");
  DGAMI_ARG_CHECK(nk >= 1 && nb >= 1 && nb <= 4 && niv >= 1 && beta > 0, "fit_smom: invalid sizes");
  int n_fit = static_cast<int>(0.2 * niv);
  if (n_fit < 4) n_fit = 4;
  DGAMI_ARG_CHECK(n_fit <= 2 * niv, "fit_smom: too few frequencies");
  hipLaunchKernelGGL(fit_smom_kernel, dim3(nb * nb), dim3(256), 0, static_cast<hipStream_t>(stream),
                     reinterpret_cast<const double2*>(sigma), nk, nb * nb, 2 * niv, n_fit, beta,
                     reinterpret_cast<double2*>(mom0), reinterpret_cast<double2*>(mom1));
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}

namespace dgami {
template <int NB>
static int update_mu_nb(int nk, int niv, double beta, double target, double mixing, const double2* ek,
                        const double2* sigma, int sigma_nk, const double2* hloc, const double2* smom0, double* mu,
                        double* info, hipStream_t s) {
  // workspace: the secant state, then the Matsubara occupations of every k
  const size_t state_bytes = 256;
  char* ws = static_cast<char*>(workspace(WS_MU, state_bytes + static_cast<size_t>(nk) * NB * NB * sizeof(double)));
  if (!ws) return -2;
  MuSecant* st = reinterpret_cast<MuSecant*>(ws);
  double* occ = reinterpret_cast<double*>(ws + state_bytes);
  hipLaunchKernelGGL(mu_secant_init_kernel, dim3(1), dim3(1), 0, s, st, mu);
  DGAMI_HIP_CHECK(hipGetLastError());
  for (int e = 0; e < MU_MAXITER + 2; ++e) {  // two starting values, then one value per iteration
    hipLaunchKernelGGL(fill_kernel<NB>, dim3(nk), dim3(256), 0, s, niv, beta, 0.0, ek, sigma, sigma_nk, smom0, hloc, occ,
                       &st->x, &st->done);
    DGAMI_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(mu_secant_kernel<NB>, dim3(1), dim3(256), 0, s, st, occ, nk, hloc, smom0, beta, target);
    DGAMI_HIP_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(mu_secant_finish_kernel, dim3(1), dim3(1), 0, s, st, mu, mixing, info);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}

template <int NB>
static int occupations_nb(int nk, int niv, double beta, const double2* ek, const double2* sigma, int sigma_nk,
                          const double2* smom0, const double* mu, double2* occ_k, double2* occ, hipStream_t s) {
  double* occ_sum = static_cast<double*>(workspace(WS_MU, 256 + static_cast<size_t>(nk) * NB * NB * sizeof(double)));
  if (!occ_sum) return -2;
  occ_sum = reinterpret_cast<double*>(reinterpret_cast<char*>(occ_sum) + 256);
  hipLaunchKernelGGL(fill_kernel<NB>, dim3(nk), dim3(256), 0, s, niv, beta, 0.0, ek, sigma, sigma_nk, smom0, nullptr,
                     occ_sum, mu, nullptr);
  DGAMI_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(occ_k_kernel<NB>, dim3((nk + 127) / 128), dim3(128), 0, s, occ_sum, ek, smom0, mu, beta, nk,
                     occ_k);
  DGAMI_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(occ_mean_kernel, dim3(1), dim3(64), 0, s, occ_k, nk, NB * NB, occ);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}
}  // namespace dgami

extern "C" int dgami_update_mu(int nk, int nb, int niv, double beta, double target, double mixing,
                               const dgami_c128* ek, const dgami_c128* sigma, int sigma_nk, const dgami_c128* hloc,
                               const dgami_c128* smom0, double* mu, double* info, void* stream) {
  using namespace dgami;
  DGAMI_ARG_CHECK(ek && sigma && hloc && smom0 && mu, "update_mu: null pointer");
  DGAMI_ARG_CHECK(nk >= 1 && nb >= 1 && nb <= 4 && niv >= 1 && beta > 0 && (sigma_nk == 1 || sigma_nk == nk),
                  "update_mu: invalid sizes");
  const double2* E = reinterpret_cast<const double2*>(ek);
  const double2* S = reinterpret_cast<const double2*>(sigma);
  const double2* H = reinterpret_cast<const double2*>(hloc);
  const double2* M0 = reinterpret_cast<const double2*>(smom0);
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (nb) {
    case 1: return update_mu_nb<1>(nk, niv, beta, target, mixing, E, S, sigma_nk, H, M0, mu, info, s);
    case 2: return update_mu_nb<2>(nk, niv, beta, target, mixing, E, S, sigma_nk, H, M0, mu, info, s);
    case 3: return update_mu_nb<3>(nk, niv, beta, target, mixing, E, S, sigma_nk, H, M0, mu, info, s);
    default: return update_mu_nb<4>(nk, niv, beta, target, mixing, E, S, sigma_nk, H, M0, mu, info, s);
  }
}

extern "C" int dgami_occupations(int nk, int nb, int niv, double beta, const dgami_c128* ek, const dgami_c128* sigma,
                                 int sigma_nk, const dgami_c128* smom0, const double* mu, dgami_c128* occ_k,
                                 dgami_c128* occ, void* stream) {
  using namespace dgami;
  DGAMI_ARG_CHECK(ek && sigma && smom0 && mu && occ_k && occ, "occupations: null pointer");
  DGAMI_ARG_CHECK(nk >= 1 && nb >= 1 && nb <= 4 && niv >= 1 && beta > 0 && (sigma_nk == 1 || sigma_nk == nk),
                  "occupations: invalid sizes");
  const double2* E = reinterpret_cast<const double2*>(ek);
  const double2* S = reinterpret_cast<const double2*>(sigma);
  const double2* M0 = reinterpret_cast<const double2*>(smom0);
  double2* OK = reinterpret_cast<double2*>(occ_k);
  double2* O = reinterpret_cast<double2*>(occ);
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (nb) {
    case 1: return occupations_nb<1>(nk, niv, beta, E, S, sigma_nk, M0, mu, OK, O, s);
    case 2: return occupations_nb<2>(nk, niv, beta, E, S, sigma_nk, M0, mu, OK, O, s);
    case 3: return occupations_nb<3>(nk, niv, beta, E, S, sigma_nk, M0, mu, OK, O, s);
    default: return occupations_nb<4>(nk, niv, beta, E, S, sigma_nk, M0, mu, OK, O, s);
  }
}
