// Bare bubble chi0(q, w, v) on the irreducible q-points (bubble_gen.py:33-69).
//
// G(k) -> G(R) by one batched forward FFT; then one workgroup per
// (w, a, b, c, d, v) forms G_ad(R, v) G_cb(-R, v - w) in LDS, runs the
// backward 3D FFT there and writes only the irreducible q-points, scaled by
// -beta/N^2 (ifft 1/N times -beta/N) and filtered at 1e-12. Output is q
// fastest (QF layout), so every workgroup writes one contiguous run.
#include "dgami.h"
#include "dgami_internal.hpp"
#include "dgami_launch.hpp"

namespace dgami {

__global__ void __launch_bounds__(256) bubble_kernel(const double2* __restrict__ g_r, const int32_t* __restrict__ irrk_ind,
                                                     double2* __restrict__ chi0, FftPlan3D plan, int nb, int niw, int niv,
                                                     int nvg, int nq, double scale) {
  extern __shared__ double2 smem[];
  const int N = plan.ntot;
  double2* buf = smem;
  double2* tw = smem + N;
  const int nb4 = nb * nb * nb * nb;
  const int nv = 2 * niv;
  // blockIdx.x -> (w, blk, v), v fastest; blk in M order ((a*nb+b)*nb+d)*nb+c
  int idx = blockIdx.x;
  const int v = idx % nv;
  idx /= nv;
  const int blk = idx % nb4;
  const int w = idx / nb4;
  const int c = blk % nb;
  const int d = (blk / nb) % nb;
  const int b = (blk / (nb * nb)) % nb;
  const int a = blk / (nb * nb * nb);

  fft_twiddles_3d(tw, plan, 1.0);
  // G(R) layout [nb][nb][nvg][N]; frequency index of v in the cut G is niw + v
  const double2* gl = g_r + (static_cast<size_t>(a * nb + d) * nvg + (niw + v)) * N;
  const double2* gr = g_r + (static_cast<size_t>(c * nb + b) * nvg + (niw + v - w)) * N;
  const int nx = plan.n[0], ny = plan.n[1], nz = plan.n[2];
  for (int r = threadIdx.x; r < N; r += blockDim.x) {
    const int z = r % nz, y = (r / nz) % ny, x = r / (nz * ny);
    const int rm = (((nx - x) % nx) * ny + (ny - y) % ny) * nz + (nz - z) % nz;  // index of -R
    buf[r] = cmul(gl[r], gr[rm]);
  }
  __syncthreads();
  fft_3d_lds(buf, plan, tw, 1.0);
  double2* dst = chi0 + ((static_cast<size_t>(w) * nb4 + blk) * nv + v) * nq;
  for (int iq = threadIdx.x; iq < nq; iq += blockDim.x) {
    const int k = irrk_ind[iq];
    double2 val = (k >= 0 && k < N) ? cscale(buf[k], scale) : make_double2(__builtin_nan(""), __builtin_nan(""));
    if (fabs(val.x) < 1e-12 && fabs(val.y) < 1e-12) val = make_double2(0.0, 0.0);
    dst[iq] = val;
  }
}

}  // namespace dgami

extern "C" int dgami_bubble_irrq(int nx, int ny, int nz, int nb, int niv_g, int niw, int niv, double beta,
                                 const dgami_c128* g_k, const int32_t* irrk_ind, int nq_irr, dgami_c128* chi0,
                                 void* stream_) {
  using namespace dgami;
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  DGAMI_ARG_CHECK(g_k != nullptr && irrk_ind != nullptr && chi0 != nullptr, "bubble: null pointer");
  DGAMI_ARG_CHECK(nb >= 1 && niw >= 0 && niv >= 1 && nq_irr >= 1, "bubble: invalid sizes");
  DGAMI_ARG_CHECK(niv_g >= niv + niw, "bubble: G must hold at least niv + niw fermionic frequencies");
  FftPlan3D plan;
  DGAMI_ARG_CHECK(nx > 0 && ny > 0 && nz > 0 && fft_plan_3d(nx, ny, nz, &plan),
                  "bubble: grid lengths must factor into 2, 3 and 5");
  DGAMI_ARG_CHECK(nq_irr <= plan.ntot, "bubble: more irreducible points than grid points");
  const int nk = plan.ntot;
  const int nvg = 2 * (niv + niw);
  const size_t gbytes = static_cast<size_t>(nb) * nb * nvg * nk * sizeof(double2);
  double2* gk = static_cast<double2*>(workspace(WS_GK, gbytes));
  double2* gr = static_cast<double2*>(workspace(WS_GR, gbytes));
  if (gk == nullptr || gr == nullptr) return -2;
  int rc = launch_ref2_to_gk(reinterpret_cast<const double2*>(g_k), gk, nk, nb, 2 * niv_g, niv_g - niv - niw, nvg,
                             stream);
  if (rc) return rc;
  rc = launch_fft_grid(gk, gr, static_cast<int64_t>(nb) * nb * nvg, nx, ny, nz, -1, 1.0, stream);
  if (rc) return rc;
  const size_t lds = (plan.ntot + nx + ny + nz) * sizeof(double2);
  const int threads = fft_block_threads(plan);
  const size_t nblocks = static_cast<size_t>(niw + 1) * nb * nb * nb * nb * 2 * niv;
  DGAMI_ARG_CHECK(nblocks < (size_t(1) << 31), "bubble: too many slices");
  const double scale = -beta / (static_cast<double>(nk) * nk);
  allow_lds(bubble_kernel, lds);
  hipLaunchKernelGGL(bubble_kernel, dim3(static_cast<unsigned>(nblocks)), dim3(threads), lds, stream, gr, irrk_ind,
                     reinterpret_cast<double2*>(chi0), plan, nb, niw, niv, nvg, nq_irr, scale);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}
