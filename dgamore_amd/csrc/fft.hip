// Batched momentum-grid FFTs and the two-point layout transform.
//
// One workgroup transforms one [nx][ny][nz] slab entirely in LDS (mixed-radix
// Stockham, radices 2/3/4/5). Grids of the hot path are at most 64x64 points
// (64 KiB of complex128), so the slab never leaves LDS between passes.
#include "dgami.h"
#include "dgami_internal.hpp"
#include "dgami_launch.hpp"

namespace dgami {

__global__ void __launch_bounds__(256) fft_grid_kernel(const double2* __restrict__ in, double2* out, FftPlan3D plan,
                                                       double sign, double scale) {
  extern __shared__ double2 smem[];
  double2* buf = smem;
  double2* tw = smem + plan.ntot;
  const size_t off = static_cast<size_t>(blockIdx.x) * plan.ntot;
  fft_twiddles_3d(tw, plan, sign);
  for (int i = threadIdx.x; i < plan.ntot; i += blockDim.x) buf[i] = in[off + i];
  __syncthreads();
  fft_3d_lds(buf, plan, tw, sign);
  for (int i = threadIdx.x; i < plan.ntot; i += blockDim.x) out[off + i] = cscale(buf[i], scale);
}

__global__ void ref2_to_gk_kernel(const double2* __restrict__ in, double2* __restrict__ out, int nk, int nb,
                                  int nv_in, int v0, int nv) {
  // in: [nk][nb][nb][nv_in]; out: [nb][nb][nv][nk] with frequencies v0 .. v0+nv-1 of the input
  const size_t total = static_cast<size_t>(nb) * nb * nv * nk;
  for (size_t idx = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; idx < total;
       idx += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int k = static_cast<int>(idx % nk);
    size_t r = idx / nk;
    const int v = static_cast<int>(r % nv);
    r /= nv;  // (a, b)
    out[idx] = in[(static_cast<size_t>(k) * nb * nb + r) * nv_in + v0 + v];
  }
}

int launch_fft_grid(const double2* in, double2* out, int64_t batch, int nx, int ny, int nz, int sign, double scale,
                    hipStream_t stream) {
  FftPlan3D plan;
  DGAMI_ARG_CHECK(nx > 0 && ny > 0 && nz > 0 && batch >= 0, "fft: grid sizes must be positive");
  DGAMI_ARG_CHECK(fft_plan_3d(nx, ny, nz, &plan), "fft: grid lengths must factor into 2, 3 and 5");
  DGAMI_ARG_CHECK(sign == 1 || sign == -1, "fft: sign must be +1 or -1");
  const size_t lds = (plan.ntot + nx + ny + nz) * sizeof(double2);
  DGAMI_ARG_CHECK(lds <= 160 * 1024, "fft: grid too large for one LDS slab (max ~10000 points)");
  const int threads = fft_block_threads(plan);
  DGAMI_ARG_CHECK(threads <= 256, "fft: grid too large for one workgroup (max 4096 points)");
  if (batch == 0) return 0;
  DGAMI_ARG_CHECK(batch < (int64_t(1) << 31), "fft: batch too large");
  allow_lds(fft_grid_kernel, lds);
  hipLaunchKernelGGL(fft_grid_kernel, dim3(static_cast<unsigned>(batch)), dim3(threads), lds, stream, in, out, plan,
                     static_cast<double>(sign), scale);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_ref2_to_gk(const double2* in, double2* out, int nk, int nb, int nv_in, int v0, int nv, hipStream_t stream) {
  const size_t total = static_cast<size_t>(nb) * nb * nv * nk;
  if (total == 0) return 0;
  const unsigned blocks = ceil_div(total, 256) > 65536u ? 65536u : ceil_div(total, 256);
  hipLaunchKernelGGL(ref2_to_gk_kernel, dim3(blocks), dim3(256), 0, stream, in, out, nk, nb, nv_in, v0, nv);
  DGAMI_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace dgami

extern "C" int dgami_fft_grid(const dgami_c128* in, dgami_c128* out, int64_t batch, int nx, int ny, int nz, int sign,
                              double scale, void* stream) {
  DGAMI_ARG_CHECK(in != nullptr && out != nullptr, "fft: null pointer");
  return dgami::launch_fft_grid(reinterpret_cast<const double2*>(in), reinterpret_cast<double2*>(out), batch, nx, ny,
                                nz, sign, scale, static_cast<hipStream_t>(stream));
}
