// Bare bubble chi0(q, w, v) on the irreducible q-points (bubble_gen.py:33-69).
//
// G(k) -> G(R) by one batched forward FFT; then one workgroup per
// (w, a, b, c, d, v) forms G_ad(R, v) G_cb(-R, v - w) in LDS, runs the
// backward 3D FFT there and writes only the irreducible q-points, scaled by
// -beta/N^2 (ifft 1/N times -beta/N) and filtered at 1e-12. Output is q
// fastest (QF layout), so every workgroup writes one contiguous run.
#include <mutex>
#include <string>

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
  // (x, y, z) of r = threadIdx.x + m blockDim.x, advanced by blockDim.x with carries instead of
  // divided out for every element
  const int st = blockDim.x;
  const int dz = st % nz, dy = (st / nz) % ny, dx = st / (nz * ny);
  int z = threadIdx.x % nz, y = (threadIdx.x / nz) % ny, x = threadIdx.x / (nz * ny);
  for (int r = threadIdx.x; r < N; r += st) {
    const int xm = x == 0 ? 0 : nx - x, ym = y == 0 ? 0 : ny - y, zm = z == 0 ? 0 : nz - z;
    buf[r] = cmul(gl[r], gr[(xm * ny + ym) * nz + zm]);  // G_cb at -R
    z += dz;
    int cy = dy, cx = dx;
    if (z >= nz) {
      z -= nz;
      ++cy;
    }
    y += cy;
    if (y >= ny) {
      y -= ny;
      ++cx;
    }
    x += cx;
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

namespace dgami {
namespace {

// Both entry points: validate, then G(k) -> [nb][nb][nvg][nk] (gk), its momentum FFT (gr) and the
// bubble kernel. gk and gr hold nb^2 * 2 (niv + niw) * nk complex128 values each.
struct BubbleShape {
  FftPlan3D plan;
  size_t gbytes;
};

int bubble_check(int nx, int ny, int nz, int nb, int niv_g, int niw, int niv, const dgami_c128* g_k,
                 const int32_t* irrk_ind, int nq_irr, const dgami_c128* chi0, BubbleShape* sh) {
  DGAMI_ARG_CHECK(g_k != nullptr && irrk_ind != nullptr && chi0 != nullptr, "bubble: null pointer");
  DGAMI_ARG_CHECK(nb >= 1 && niw >= 0 && niv >= 1 && nq_irr >= 1, "bubble: invalid sizes");
  DGAMI_ARG_CHECK(niv_g >= niv + niw, "bubble: G must hold at least niv + niw fermionic frequencies");
  DGAMI_ARG_CHECK(nx > 0 && ny > 0 && nz > 0 && fft_plan_3d(nx, ny, nz, &sh->plan),
                  "bubble: grid lengths must factor into 2, 3 and 5");
  DGAMI_ARG_CHECK(nq_irr <= sh->plan.ntot, "bubble: more irreducible points than grid points");
  sh->gbytes = static_cast<size_t>(nb) * nb * 2 * (niv + niw) * sh->plan.ntot * sizeof(double2);
  return 0;
}

int bubble_run(const BubbleShape& sh, int nx, int ny, int nz, int nb, int niv_g, int niw, int niv, double beta,
               const dgami_c128* g_k, const int32_t* irrk_ind, int nq_irr, dgami_c128* chi0, double2* gk, double2* gr,
               hipStream_t stream) {
  const FftPlan3D& plan = sh.plan;
  const int nk = plan.ntot;
  const int nvg = 2 * (niv + niw);
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

// Scratch of dgami_bubble_irrq_noalloc: a file-scope device array, one copy per GPU, that holds
// gk and gr. Host code never uses the array's name as an address: the address on each device is
// obtained once with hipGetSymbolAddress (with that device current) and kept in scratch_ptr.
constexpr size_t BUBBLE_SCRATCH_BYTES = DGAMI_BUBBLE_NOALLOC_MAX_BYTES;
__device__ __align__(256) double2 scratch[BUBBLE_SCRATCH_BYTES / sizeof(double2)];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[MAX_DEVICES] = {};

// Address of this device's scratch (looked up on the first call on the device).
int scratch_address(void** out) {
  int dev = 0;
  DGAMI_HIP_CHECK(hipGetDevice(&dev));
  DGAMI_ARG_CHECK(dev >= 0 && dev < MAX_DEVICES, "bubble_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    DGAMI_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace
}  // namespace dgami

extern "C" int dgami_bubble_irrq(int nx, int ny, int nz, int nb, int niv_g, int niw, int niv, double beta,
                                 const dgami_c128* g_k, const int32_t* irrk_ind, int nq_irr, dgami_c128* chi0,
                                 void* stream_) {
  using namespace dgami;
  BubbleShape sh;
  int rc = bubble_check(nx, ny, nz, nb, niv_g, niw, niv, g_k, irrk_ind, nq_irr, chi0, &sh);
  if (rc) return rc;
  double2* gk = static_cast<double2*>(workspace(WS_GK, sh.gbytes));
  double2* gr = static_cast<double2*>(workspace(WS_GR, sh.gbytes));
  if (gk == nullptr || gr == nullptr) return -2;
  return bubble_run(sh, nx, ny, nz, nb, niv_g, niw, niv, beta, g_k, irrk_ind, nq_irr, chi0, gk, gr,
                    static_cast<hipStream_t>(stream_));
}

extern "C" int dgami_bubble_irrq_noalloc(int nx, int ny, int nz, int nb, int niv_g, int niw, int niv, double beta,
                                         const dgami_c128* g_k, const int32_t* irrk_ind, int nq_irr, dgami_c128* chi0,
                                         void* stream_) {
  using namespace dgami;
  BubbleShape sh;
  int rc = bubble_check(nx, ny, nz, nb, niv_g, niw, niv, g_k, irrk_ind, nq_irr, chi0, &sh);
  if (rc) return rc;
  const size_t half = (sh.gbytes + 255) & ~size_t(255);
  DGAMI_ARG_CHECK(2 * half <= BUBBLE_SCRATCH_BYTES,
                  "bubble_noalloc: G needs " + std::to_string(2 * half) + " bytes of scratch, more than the " +
                      std::to_string(BUBBLE_SCRATCH_BYTES) + " bytes the no-allocation entry point holds");
  void* base = nullptr;
  rc = scratch_address(&base);
  if (rc) return rc;
  double2* gk = static_cast<double2*>(base);
  double2* gr = reinterpret_cast<double2*>(static_cast<char*>(base) + half);
  return bubble_run(sh, nx, ny, nz, nb, niv_g, niw, niv, beta, g_k, irrk_ind, nq_irr, chi0, gk, gr,
                    static_cast<hipStream_t>(stream_));
}
