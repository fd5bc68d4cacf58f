// Double-counting kernel of the self-energy (nonlocal_sde.py:185-198, with the
// caller's prefactors at :817-820): for every bosonic frequency w one complex
// GEMM with rows (q, ab), contraction ((dc), v) and columns ((fe), v'):
//   L[w][(q,ab)][((dc),v)]   = (1/beta^2) (M(u') M(chi0(q,w,v)))[(ab),(dc)],  u'_abij = u_ajib
//   R[w][((dc),v)][((fe),v')] = f_dc_dcef(w,v,v') = 2 F_magn,loc[e][c][d][f][w][v][v']
//   kernel[w][(ab),(fe)][v'][q] = -(L R)[(q,ab)][((fe),v')],  v' cut to the core box.
#include "dgami.h"
#include "dgami_internal.hpp"
#include "dgami_launch.hpp"
#include "zgemm.hpp"

namespace dgami {

// L operand: thread per (w, q, ab, (dc), v)
__global__ void dc_left_kernel(const double2* __restrict__ chi0, const double2* __restrict__ u, double2* __restrict__ L,
                               int nq, int nb, int nw, int nvf, double inv_beta2) {
  const int T = nb * nb;
  const size_t total = static_cast<size_t>(nw) * nq * T * T * nvf;
  for (size_t idx = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; idx < total;
       idx += static_cast<size_t>(gridDim.x) * blockDim.x) {
    size_t r = idx;
    const int v = static_cast<int>(r % nvf);
    r /= nvf;
    const int dc = static_cast<int>(r % T);
    r /= T;
    const int ab = static_cast<int>(r % T);
    r /= T;
    const int q = static_cast<int>(r % nq);
    const int w = static_cast<int>(r / nq);
    const int a = ab / nb, b = ab % nb;
    double2 acc = make_double2(0.0, 0.0);
    for (int ji = 0; ji < T; ++ji) {
      const int j = ji / nb, i = ji % nb;
      // M(u')[(ab),(ji)] = u'_{a b i j} = u_{a j i b} (reference order)
      const double2 uu = u[((a * nb + j) * nb + i) * nb + b];
      const double2 x = chi0[((static_cast<size_t>(w) * T * T + ji * T + dc) * nvf + v) * nq + q];
      acc = cfma(acc, uu, x);
    }
    L[idx] = cscale(acc, inv_beta2);
  }
}

// R operand: thread per (w, (dc), v, (fe), v')
__global__ void dc_right_kernel(const double2* __restrict__ f, double2* __restrict__ R, int nb, int nw, int nvf,
                                int nvc, int p0) {
  const int T = nb * nb;
  const size_t total = static_cast<size_t>(nw) * T * nvf * T * nvc;
  for (size_t idx = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; idx < total;
       idx += static_cast<size_t>(gridDim.x) * blockDim.x) {
    size_t r = idx;
    const int p = static_cast<int>(r % nvc);
    r /= nvc;
    const int fe = static_cast<int>(r % T);
    r /= T;
    const int v = static_cast<int>(r % nvf);
    r /= nvf;
    const int dc = static_cast<int>(r % T);
    const int w = static_cast<int>(r / T);
    const int d = dc / nb, c = dc % nb, ff = fe / nb, e = fe % nb;
    const size_t src = ((((static_cast<size_t>(e) * nb + c) * nb + d) * nb + ff) * nw + w) * nvf * nvf +
                       static_cast<size_t>(v) * nvf + p0 + p;
    R[idx] = cscale(f[src], 2.0);
  }
}

// scatter C[w][(q,ab)][((fe),p)] -> kernel[w][(ab),(fe)][p][q] with a minus sign
__global__ void dc_scatter_kernel(const double2* __restrict__ C, double2* __restrict__ kernel, int nq, int nb, int nw,
                                  int nvc) {
  const int T = nb * nb;
  const size_t total = static_cast<size_t>(nw) * T * T * nvc * nq;
  for (size_t idx = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; idx < total;
       idx += static_cast<size_t>(gridDim.x) * blockDim.x) {
    size_t r = idx;
    const int q = static_cast<int>(r % nq);
    r /= nq;
    const int p = static_cast<int>(r % nvc);
    r /= nvc;
    const int blk = static_cast<int>(r % (T * T));
    const int w = static_cast<int>(r / (T * T));
    const int ab = blk / T, fe = blk % T;
    const double2 x = C[((static_cast<size_t>(w) * nq + q) * T + ab) * (static_cast<size_t>(T) * nvc) + fe * nvc + p];
    kernel[idx] = make_double2(-x.x, -x.y);
  }
}

}  // namespace dgami

extern "C" int dgami_dc_kernel(int nq, int nb, int nw, int niv_full, int niv_core, double beta, const dgami_c128* chi0,
                               const dgami_c128* u, const dgami_c128* f_magn_loc, dgami_c128* kernel, void* stream_) {
  using namespace dgami;
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  DGAMI_ARG_CHECK(chi0 && u && f_magn_loc && kernel, "dc_kernel: null pointer");
  DGAMI_ARG_CHECK(nq >= 1 && nb >= 1 && nw >= 1 && niv_core >= 1 && niv_full >= niv_core && beta > 0,
                  "dc_kernel: invalid sizes");
  const int T = nb * nb, nvf = 2 * niv_full, nvc = 2 * niv_core;
  const size_t lsz = static_cast<size_t>(nw) * nq * T * T * nvf;
  const size_t rsz = static_cast<size_t>(nw) * T * nvf * T * nvc;
  const size_t csz = static_cast<size_t>(nw) * nq * T * T * nvc;
  double2* ws = static_cast<double2*>(workspace(WS_DC, (lsz + rsz + csz) * sizeof(double2)));
  if (!ws) return -2;
  double2 *L = ws, *R = ws + lsz, *C = ws + lsz + rsz;
  auto blocks = [](size_t total) { return ceil_div(total, 256) > 262144u ? 262144u : ceil_div(total, 256); };
  hipLaunchKernelGGL(dc_left_kernel, dim3(blocks(lsz)), dim3(256), 0, stream, reinterpret_cast<const double2*>(chi0),
                     reinterpret_cast<const double2*>(u), L, nq, nb, nw, nvf, 1.0 / (beta * beta));
  DGAMI_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(dc_right_kernel, dim3(blocks(rsz)), dim3(256), 0, stream,
                     reinterpret_cast<const double2*>(f_magn_loc), R, nb, nw, nvf, nvc, niv_full - niv_core);
  DGAMI_HIP_CHECK(hipGetLastError());
  ZgemmArgs g;
  g.A = L;
  g.B = R;
  g.C = C;
  g.M = nq * T;
  g.K = T * nvf;
  g.N = T * nvc;
  g.lda = g.K;
  g.ldb = g.N;
  g.ldc = g.N;
  g.sa = static_cast<long long>(g.M) * g.K;
  g.sb = static_cast<long long>(g.K) * g.N;
  g.sc = static_cast<long long>(g.M) * g.N;
  g.alpha = make_double2(1.0, 0.0);
  g.beta = make_double2(0.0, 0.0);
  int rc = launch_zgemm(g, nw, stream);
  if (rc) return rc;
  hipLaunchKernelGGL(dc_scatter_kernel, dim3(blocks(csz)), dim3(256), 0, stream, C, reinterpret_cast<double2*>(kernel),
                     nq, nb, nw, nvc);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}
