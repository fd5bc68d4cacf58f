// Host-side launchers shared between translation units of libdgami.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace dgami {

// Workspace slots (see workspace() in dgami_internal.hpp); one purpose each.
enum WorkspaceSlot {
  WS_GK = 0,         // two-point object in [nb][nb][nv][nk] layout
  WS_GR = 1,         // its momentum-space FFT
  WS_SIGMA_PART = 2, // per-frequency-chunk partial self-energies
  WS_LADDER = 3,     // ladder matrices / factors (large-n path)
  WS_LADDER_AUX = 4, // pivots, inverse pivots and interchange maps of WS_LADDER
  WS_DC = 5,         // double-counting kernel operands
  WS_PAIR = 6,       // Eliashberg matvec intermediates
  WS_MISC = 7,
  WS_MIXED = 8,      // mixed-precision factors (slots 8 and 9)
  WS_MIXED_VEC = 10, // refinement vectors
  WS_TRI = 11,       // inverse diagonal blocks of the triangular factors
  WS_TRI_VEC = 12,   // complex64 right-hand sides of the triangular sweeps
  WS_SELFCONS = 13,  // Hartree-Fock transforms
  WS_MU = 14,        // chemical-potential secant: per-k occupations and the secant state
  WS_LADDER_G32 = 15, // complex64 scaled vertex of the fused first block of the blocked LU
  // work issued on the library's extra stream k (1..MAX_EXTRA_STREAMS) uses its own slots:
  // slot + k * WS_STREAM_STRIDE
  WS_STREAM_STRIDE = 16,
};
constexpr int MAX_EXTRA_STREAMS = 3;

// The library's extra stream k of the current device (created on first use, default priority,
// not blocking the null stream): fork makes it wait for the work queued so far on `caller`,
// join makes `caller` wait for the work queued on it, so that a captured caller stream takes
// the forked work into its graph.
int extra_stream_fork(hipStream_t caller, int k, hipStream_t* out);
int extra_stream_join(hipStream_t caller, int k);

int launch_fft_grid(const double2* in, double2* out, int64_t batch, int nx, int ny, int nz, int sign, double scale,
                    hipStream_t stream);
int launch_ref2_to_gk(const double2* in, double2* out, int nk, int nb, int nv_in, int v0, int nv, hipStream_t stream);

}  // namespace dgami
