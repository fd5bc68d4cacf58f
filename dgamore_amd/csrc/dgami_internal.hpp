// Shared device/host helpers for libdgami (gfx950, CDNA4).
//
// Complex numbers are interleaved (re, im) doubles, i.e. the memory layout of
// numpy.complex128; on the device they are handled as HIP's double2.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace dgami {

// ----------------------------------------------------------------------------
// complex double helpers
// ----------------------------------------------------------------------------
__host__ __device__ __forceinline__ double2 cmk(double r, double i) { return make_double2(r, i); }
__host__ __device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// a - b*c
__host__ __device__ __forceinline__ double2 cfms(double2 a, double2 b, double2 c) {
  double re = fma(-b.x, c.x, a.x);
  re = fma(b.y, c.y, re);
  double im = fma(-b.x, c.y, a.y);
  im = fma(-b.y, c.x, im);
  return make_double2(re, im);
}
// a + b*c
__host__ __device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 c) {
  double re = fma(b.x, c.x, a.x);
  re = fma(-b.y, c.y, re);
  double im = fma(b.x, c.y, a.y);
  im = fma(b.y, c.x, im);
  return make_double2(re, im);
}
// 1/b with Smith's scaling (robust against over/underflow of |b|^2)
__host__ __device__ __forceinline__ double2 cinv(double2 b) {
  if (fabs(b.x) >= fabs(b.y)) {
    double r = b.y / b.x;
    double d = b.x + b.y * r;
    return make_double2(1.0 / d, -r / d);
  }
  double r = b.x / b.y;
  double d = b.x * r + b.y;
  return make_double2(r / d, -1.0 / d);
}
__host__ __device__ __forceinline__ double cabs1(double2 a) { return fabs(a.x) + fabs(a.y); }

// complex float (interleaved, numpy.complex64 layout)
__host__ __device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ float2 cfms(float2 a, float2 b, float2 c) {
  float re = fmaf(-b.x, c.x, a.x);
  re = fmaf(b.y, c.y, re);
  float im = fmaf(-b.x, c.y, a.y);
  im = fmaf(-b.y, c.x, im);
  return make_float2(re, im);
}
__host__ __device__ __forceinline__ float2 cfma(float2 a, float2 b, float2 c) {
  float re = fmaf(b.x, c.x, a.x);
  re = fmaf(-b.y, c.y, re);
  float im = fmaf(b.x, c.y, a.y);
  im = fmaf(b.y, c.x, im);
  return make_float2(re, im);
}
__host__ __device__ __forceinline__ float2 cinv(float2 b) {
  if (fabsf(b.x) >= fabsf(b.y)) {
    float r = b.y / b.x;
    float d = b.x + b.y * r;
    return make_float2(1.0f / d, -r / d);
  }
  float r = b.x / b.y;
  float d = b.x * r + b.y;
  return make_float2(r / d, -1.0f / d);
}
__host__ __device__ __forceinline__ float cabs1(float2 a) { return fabsf(a.x) + fabsf(a.y); }

// conversions and zero of the two storage types
template <typename S>
__host__ __device__ __forceinline__ S czero();
template <>
__host__ __device__ __forceinline__ double2 czero<double2>() { return make_double2(0.0, 0.0); }
template <>
__host__ __device__ __forceinline__ float2 czero<float2>() { return make_float2(0.0f, 0.0f); }
template <typename S>
__host__ __device__ __forceinline__ S from_c128(double2 v);
template <>
__host__ __device__ __forceinline__ double2 from_c128<double2>(double2 v) { return v; }
template <>
__host__ __device__ __forceinline__ float2 from_c128<float2>(double2 v) {
  return make_float2(static_cast<float>(v.x), static_cast<float>(v.y));
}
__host__ __device__ __forceinline__ double2 to_c128(double2 v) { return v; }
__host__ __device__ __forceinline__ double2 to_c128(float2 v) {
  return make_double2(static_cast<double>(v.x), static_cast<double>(v.y));
}

// ----------------------------------------------------------------------------
// XCD-aware tile order. Workgroups are dispatched round-robin over MI355X's 8 XCDs, each
// with its own L2. With a grid of xcd_grid(nprob, per) workgroups, xcd_tile maps the
// linear id so that all `per` tiles of one problem run on the same XCD and share its L2
// copy of the problem's operand panels. Returns false for the padding ids.
// ----------------------------------------------------------------------------
constexpr int NUM_XCD = 8;

inline size_t xcd_grid(size_t nprob, size_t per) { return (nprob + NUM_XCD - 1) / NUM_XCD * NUM_XCD * per; }

__device__ __forceinline__ bool xcd_tile(unsigned id, int nprob, int per, int& prob, int& tile) {
  const unsigned x = id % NUM_XCD, slot = id / NUM_XCD;
  const unsigned j = slot / per;
  tile = static_cast<int>(slot - j * per);
  prob = static_cast<int>(j * NUM_XCD + x);
  return prob < nprob;
}

// ----------------------------------------------------------------------------
// Wavefront maximum of a 32-bit key with DPP row shifts and row broadcasts (the result
// collects in lane 63 and is returned uniform). Pivot searches reduce the key
// (upper word of the FP64 magnitude) + 1, then pick the lowest lane holding the maximum.
// ----------------------------------------------------------------------------
__device__ __forceinline__ unsigned wave_max_u32(unsigned v) {
  v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x111, 0xf, 0xf, false)));
  v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x112, 0xf, 0xf, false)));
  v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x114, 0xf, 0xf, false)));
  v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x118, 0xf, 0xf, false)));
  v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x142, 0xa, 0xf, false)));
  v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x143, 0xc, 0xf, false)));
  return static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}

__device__ __forceinline__ unsigned pivot_key(double mag, bool valid) {
  return valid ? static_cast<unsigned>(__double_as_longlong(mag) >> 32) + 1u : 0u;
}

// ----------------------------------------------------------------------------
// error reporting (thread-local last error; host side)
// ----------------------------------------------------------------------------
void set_error(const std::string& msg);
const char* last_error();

#define DGAMI_ARG_CHECK(cond, msg)      \
  do {                                  \
    if (!(cond)) {                      \
      ::dgami::set_error(msg);          \
      return -1;                        \
    }                                   \
  } while (0)

#define DGAMI_HIP_CHECK(expr)                                                              \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess) {                                                                \
      ::dgami::set_error(std::string("HIP error in " #expr ": ") + hipGetErrorString(_e)); \
      return -2;                                                                           \
    }                                                                                      \
  } while (0)

// Kernel timing (dgami_profile_enable): profile_begin records a start event on the stream,
// profile_end the stop event together with the launch's algorithmic bytes and flops.
bool profile_on();
int profile_begin(hipStream_t stream, hipEvent_t* start);
int profile_end(hipStream_t stream, hipEvent_t start, double bytes, double flops);

// Grow-only device scratch owned by the library, one slot per purpose id.
// Returns nullptr (and sets the error) when the allocation fails.
void* workspace(int slot, size_t bytes);

inline unsigned ceil_div(size_t a, size_t b) { return static_cast<unsigned>((a + b - 1) / b); }

// Allow a kernel to use more than the default 64 KiB of dynamic LDS (gfx950 has 160 KiB per CU).
template <typename K>
inline void allow_lds(K kernel, size_t bytes) {
  if (bytes > 48 * 1024)
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(bytes));
}

// ----------------------------------------------------------------------------
// in-LDS mixed-radix FFT (radices 2, 3, 4, 5), Stockham autosort
// ----------------------------------------------------------------------------
// Factorisation of a transform length into radix stages (largest radices first).
struct FftPlan1D {
  int len;
  int nstage;
  int radix[8];
};

__host__ __device__ inline bool fft_plan(int len, FftPlan1D* p) {
  p->len = len;
  p->nstage = 0;
  int m = len;
  while (m > 1) {
    int r = (m % 4 == 0) ? 4 : (m % 2 == 0) ? 2 : (m % 3 == 0) ? 3 : (m % 5 == 0) ? 5 : 0;
    if (r == 0 || p->nstage == 8) return false;
    p->radix[p->nstage++] = r;
    m /= r;
  }
  return true;
}

// Small DFT kernels on registers; sign = -1 forward, +1 backward.
template <int R>
__device__ __forceinline__ void dft_small(double2* v, double sign);

template <>
__device__ __forceinline__ void dft_small<2>(double2* v, double) {
  double2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <>
__device__ __forceinline__ void dft_small<4>(double2* v, double sign) {
  double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
  double2 b0 = cadd(v[1], v[3]), b1 = csub(v[1], v[3]);
  // multiply b1 by (sign * i)
  double2 b1r = make_double2(-sign * b1.y, sign * b1.x);
  v[0] = cadd(a0, b0);
  v[2] = csub(a0, b0);
  v[1] = cadd(a1, b1r);
  v[3] = csub(a1, b1r);
}

template <>
__device__ __forceinline__ void dft_small<3>(double2* v, double sign) {
  // w = exp(sign * 2 pi i / 3) = -1/2 + sign * i sqrt(3)/2
  const double s3 = 0.86602540378443864676 * sign;
  double2 t1 = cadd(v[1], v[2]);
  double2 t2 = csub(v[1], v[2]);
  double2 m = make_double2(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
  double2 r = make_double2(-s3 * t2.y, s3 * t2.x);
  v[0] = cadd(v[0], t1);
  v[1] = cadd(m, r);
  v[2] = csub(m, r);
}

template <>
__device__ __forceinline__ void dft_small<5>(double2* v, double sign) {
  // w = exp(sign 2 pi i / 5); with t1 = x1 + x4, t2 = x2 + x3, t3 = x1 - x4, t4 = x2 - x3:
  //   X0 = x0 + t1 + t2,  X1,4 = a1 +- i b1,  X2,3 = a2 +- i b2,
  //   a1 = x0 + c1 t1 + c2 t2, a2 = x0 + c2 t1 + c1 t2 (c_k = cos(2 pi k / 5)),
  //   b1 = sign (s1 t3 + s2 t4), b2 = sign (s2 t3 - s1 t4) (s_k = sin(2 pi k / 5))
  const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
  const double s1 = 0.95105651629515357212 * sign, s2 = 0.58778525229247312917 * sign;
  const double2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  const double2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  const double2 x0 = v[0];
  const double2 a1 = make_double2(x0.x + c1 * t1.x + c2 * t2.x, x0.y + c1 * t1.y + c2 * t2.y);
  const double2 a2 = make_double2(x0.x + c2 * t1.x + c1 * t2.x, x0.y + c2 * t1.y + c1 * t2.y);
  const double2 b1 = make_double2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
  const double2 b2 = make_double2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
  v[0] = cadd(x0, cadd(t1, t2));
  v[1] = make_double2(a1.x - b1.y, a1.y + b1.x);
  v[4] = make_double2(a1.x + b1.y, a1.y - b1.x);
  v[2] = make_double2(a2.x - b2.y, a2.y + b2.x);
  v[3] = make_double2(a2.x + b2.y, a2.y - b2.x);
}

// One Stockham stage of radix R over `nfft` transforms of length L held in
// LDS. Transform f starts at (f / nf2) * s1 + (f % nf2) * s2 and its elements
// are `xstride` apart. The stage reads all inputs into registers,
// synchronises, and writes back in place. `tw` holds exp(sign 2 pi i j / L),
// j < L. Must be called by all threads of the block; the launch guarantees
// nfft * L / R <= fft_max_bfly(R) * blockDim.x (see fft_block_threads).
__host__ __device__ constexpr int fft_max_bfly(int R) { return R == 2 ? 8 : R == 3 ? 5 : R == 4 ? 4 : 3; }

template <int R>
__device__ __forceinline__ void fft_stage(double2* buf, int L, int nfft, int nf2, int s1, int s2, int xstride, int Ns,
                                          const double2* tw, double sign) {
  constexpr int MB = fft_max_bfly(R);
  const int nbf = L / R;  // butterflies per transform
  const int total = nbf * nfft;
  const int step = L / (Ns * R);
  double2 v[MB][R];
  int boff[MB], dst[MB];  // each butterfly's transform base and output index, kept for the write
  // butterfly b -> (transform f, butterfly j): along a strided axis consecutive threads take
  // consecutive transforms, whose elements are adjacent in LDS (the butterfly-major order puts a
  // whole wavefront on one bank there); along a contiguous axis consecutive butterflies
  const bool fmajor = xstride > 1;
#pragma unroll
  for (int c = 0; c < MB; ++c) {
    const int b = threadIdx.x + c * blockDim.x;
    if (b < total) {
      const int f = fmajor ? b % nfft : b / nbf;
      const int j = fmajor ? b / nfft : b - f * nbf;
      const int jq = j / Ns, k = j - jq * Ns;
      const int fq = f / nf2;
      boff[c] = fq * s1 + (f - fq * nf2) * s2;
      dst[c] = jq * Ns * R + k;
      const double2* base = buf + boff[c];
#pragma unroll
      for (int r = 0; r < R; ++r) v[c][r] = base[(j + r * nbf) * xstride];
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[c][r] = cmul(v[c][r], tw[k * r * step]);  // k r step < L
      }
      dft_small<R>(v[c], sign);
    }
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < MB; ++c) {
    const int b = threadIdx.x + c * blockDim.x;
    if (b < total) {
      double2* base = buf + boff[c];
#pragma unroll
      for (int r = 0; r < R; ++r) base[(dst[c] + r * Ns) * xstride] = v[c][r];
    }
  }
  __syncthreads();
}

// Full 1D FFT of `nfft` transforms of length plan.len (layout as in fft_stage).
__device__ inline void fft_1d_lds(double2* buf, const FftPlan1D& plan, int nfft, int nf2, int s1, int s2, int xstride,
                                  const double2* tw, double sign) {
  int Ns = 1;
  for (int s = 0; s < plan.nstage; ++s) {
    const int R = plan.radix[s];
    if (R == 4)
      fft_stage<4>(buf, plan.len, nfft, nf2, s1, s2, xstride, Ns, tw, sign);
    else if (R == 2)
      fft_stage<2>(buf, plan.len, nfft, nf2, s1, s2, xstride, Ns, tw, sign);
    else if (R == 3)
      fft_stage<3>(buf, plan.len, nfft, nf2, s1, s2, xstride, Ns, tw, sign);
    else
      fft_stage<5>(buf, plan.len, nfft, nf2, s1, s2, xstride, Ns, tw, sign);
    Ns *= R;
  }
}

// Twiddle table exp(sign 2 pi i j / L), j < L, filled cooperatively (no barrier).
__device__ inline void fft_twiddles(double2* tw, int L, double sign) {
  for (int j = threadIdx.x; j < L; j += blockDim.x) {
    double s, c;
    sincospi(sign * 2.0 * j / L, &s, &c);
    tw[j] = make_double2(c, s);
  }
}

// Momentum-grid FFT plan: lengths and radix stages of the three axes.
struct FftPlan3D {
  FftPlan1D p[3];
  int n[3];
  int ntot;
};

inline bool fft_plan_3d(int nx, int ny, int nz, FftPlan3D* p) {
  p->n[0] = nx;
  p->n[1] = ny;
  p->n[2] = nz;
  p->ntot = nx * ny * nz;
  return fft_plan(nx, &p->p[0]) && fft_plan(ny, &p->p[1]) && fft_plan(nz, &p->p[2]);
}

// Threads per block so that every stage of a 3D transform of one grid fits the
// per-thread butterfly budget (a multiple of 64, at least 64).
inline int fft_block_threads(const FftPlan3D& p, int min_threads = 64) {
  int need = min_threads;
  for (int d = 0; d < 3; ++d)
    for (int s = 0; s < p.p[d].nstage; ++s) {
      const int R = p.p[d].radix[s];
      const int bfly = p.ntot / R;
      const int t = (bfly + fft_max_bfly(R) - 1) / fft_max_bfly(R);
      if (t > need) need = t;
    }
  return ((need + 63) / 64) * 64;
}

// LDS twiddle storage: exp(sign 2 pi i j / n_d) for the three axes, concatenated.
__device__ inline void fft_twiddles_3d(double2* tw, const FftPlan3D& p, double sign) {
  fft_twiddles(tw, p.n[0], sign);
  fft_twiddles(tw + p.n[0], p.n[1], sign);
  fft_twiddles(tw + p.n[0] + p.n[1], p.n[2], sign);
}

// 3D FFT of one [nx][ny][nz] slab in LDS (z fastest), unnormalised, sign -1
// forward (e^{-i k R}) / +1 backward. Axes of length 1 are skipped.
__device__ inline void fft_3d_lds(double2* buf, const FftPlan3D& p, const double2* tw, double sign) {
  const int nx = p.n[0], ny = p.n[1], nz = p.n[2];
  if (nz > 1) fft_1d_lds(buf, p.p[2], nx * ny, 1, nz, 0, 1, tw + nx + ny, sign);
  if (ny > 1) fft_1d_lds(buf, p.p[1], nx * nz, nz, ny * nz, 1, nz, tw + nx, sign);
  if (nx > 1) fft_1d_lds(buf, p.p[0], ny * nz, ny * nz, 0, 1, ny * nz, tw, sign);
}

}  // namespace dgami
