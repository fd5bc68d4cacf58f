// Workspace state of the batched blocked LU (shared by the plain and mixed forms).
#pragma once

#include "dgami_internal.hpp"
#include "dgami_launch.hpp"
#include "ladder_common.hpp"

namespace dgami {

constexpr int BLK_MAX_NB = 64;                      // widest panel
constexpr int BLK_MAX_N = 1536;                     // largest compound size of the blocked LU
constexpr int BLK_MAP_STRIDE = 3 * BLK_MAX_NB + 1;  // ints of one problem's interchange map
constexpr int BLK_PANELS = 8;                       // most register panels per outer block (trailing-update depth)

// Trailing-update kernels of the complex64 factorisation (complex128 always uses vector FMA).
enum BlockedGemm { BLK_GEMM_VECTOR = 0, BLK_GEMM_F32_MFMA = 1, BLK_GEMM_BF16X3 = 2 };

template <typename S>
struct BlockedState {
  S* W;            // [batch][n][ld] matrices (row major), factored in place
  double2* dinv;   // [batch][n] inverse pivots (FP64 copy)
  int* piv;        // [batch][n] pivot row of every elimination step
  int* rowmap;     // [batch][BLK_PANELS][BLK_MAP_STRIDE] interchange maps of the panels of an outer block
  int* info;       // [batch] first zero pivot (1-based) or 0
  int n, ld, ncol;
  int swap_left;   // also apply interchanges left of the panel (needed when L is reused)
  int gemm_mode;   // BlockedGemm (complex64 only)
  int variant;     // blocked-LU variant (panel width and rows per thread), -1: chosen by n
  int tall_tiles;  // complex128: 128-row update tiles for updates of more than 96 rows (0: 64-row tiles)
  // First outer block fused with the assembly (fuse_e > 0): only the block column [0, fuse_e) and
  // the block row [0, fuse_e) are assembled; every other element is formed from the source
  // matrix where it is first read (the interchanges of the block's panels and its trailing update).
  int fuse_e;       // width of the fused first block (0: the whole matrix is assembled)
  int* rowflag;     // [batch][n]: 0 while row r >= fuse_e of the trailing columns is the source row r
                    // (not in W), else 1 + the panel slot whose interchanges first wrote it
  int item0;        // first item of the batch
  LadderArgs src;   // the items' matrices
  // complex64 with a shared vertex and no V_r: (gamma - u_r) / beta^2 of the batch's (chan, w)
  // groups in complex64 ([group - g32_first][n][n]), the fused update's source off the v = v' diagonal
  const float2* g32;
  int g32_first;
  int ws_g32;  // workspace slot of the complex64 vertex copy
};

// Allocate the workspace of `batch` problems in slots ws and ws+1; returns false on failure.
template <typename S>
inline bool blocked_alloc(BlockedState<S>& st, int n, int ld, int ncol, size_t batch, int ws) {
  st.W = static_cast<S*>(workspace(ws, batch * n * ld * sizeof(S)));
  const size_t aux = batch * n * sizeof(double2) + batch * n * sizeof(int) + batch * BLK_PANELS * BLK_MAP_STRIDE * sizeof(int) +
                     batch * sizeof(int) + batch * n * sizeof(int) + 1024;
  char* p = static_cast<char*>(workspace(ws + 1, aux));
  if (!st.W || !p) return false;
  st.dinv = reinterpret_cast<double2*>(p);
  st.piv = reinterpret_cast<int*>(st.dinv + batch * n);
  st.rowmap = st.piv + batch * n;
  st.info = st.rowmap + batch * BLK_PANELS * BLK_MAP_STRIDE;
  st.rowflag = st.info + batch;
  st.fuse_e = 0;
  st.item0 = 0;
  st.src = LadderArgs{};
  st.g32 = nullptr;
  st.g32_first = 0;
  st.ws_g32 = WS_LADDER_G32 + ws / WS_STREAM_STRIDE * WS_STREAM_STRIDE;  // the lane's slot range
  st.n = n;
  st.ld = ld;
  st.ncol = ncol;
  st.swap_left = 0;
  st.gemm_mode = BLK_GEMM_F32_MFMA;
  st.variant = -1;
  st.tall_tiles = 0;
  return true;
}

template <typename S>
int blocked_factor(const LadderArgs& a, BlockedState<S>& st, int item0, int nb, hipStream_t stream);

// Lanes of a batched blocked-LU solve: with L > 1 the problems are split into L parts solved on
// the caller's stream and on the library's extra streams, each with its own workspace slots, so
// that the latency-bound steps of one part (panels, interchanges, sweeps) overlap the
// matrix-core or vector updates of another. Measured on the 3-orbital 24x24 workload (5642
// problems of n = 720, split-BF16): 1 lane 100.9 ms, 2 lanes 97.5, 3 and 4 lanes 97.5-97.8, with
// bit-identical results. DGAMI_LADDER_STREAMS (1..4) overrides the default of 2.
int ladder_lanes();

// W[r0:r1, c0:c1] -= W[r0:r1, k0:k0+kw] W[k0:k0+kw, c0:c1] on every complex128 problem of the
// batch (the trailing-update kernel of the factorisation, reused by the block back-substitution).
int blocked_gemm_fp64(BlockedState<double2>& st, int nb, int r0, int r1, int c0, int c1, int k0, int kw,
                      hipStream_t stream);

// The same update on complex64 problems with the state's matrix-core kernel (st.gemm_mode).
int blocked_gemm_c64(BlockedState<float2>& st, int nb, int r0, int r1, int c0, int c1, int k0, int kw,
                     hipStream_t stream);

// info[item0 + b] = first zero pivot of problem b of the batch (no-op for info == nullptr).
int blocked_copy_info(const int* st_info, int32_t* info, int item0, int nitems, hipStream_t stream);

}  // namespace dgami
