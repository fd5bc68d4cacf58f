// Batched blocked triangular solves on the LU factors of the blocked ladder path.
//
// The factors of every problem are stored row major in W[item][n][ld] (unit lower L
// below the diagonal, U on and above it). Solves proceed in blocks of TRI_B = 32 rows:
// the inverses of the 32x32 diagonal blocks of L and U are formed once per
// factorisation (tri_diag_inverse), after which a sweep is one launch with one workgroup
// per problem, the right-hand sides in registers, and per block b
//   x_b = D_b^-1 y_b                (D_b^-1 the stored diagonal-block inverse)
//   y_rows -= F[rows, b] x_b        (rows below b for L, above b for U)
// so that every factor element is read once per sweep.
#pragma once

#include "dgami_internal.hpp"
#include "ladder_common.hpp"

namespace dgami {

constexpr int TRI_B = 32;      // rows per diagonal block

template <typename S>
struct TriFactors {
  const S* W;          // [batch][n][ld] factors
  S* dinvblk;          // [batch][nblk][2][TRI_B][TRI_B]: inverse of L_bb, inverse of U_bb
  const double2* dinv; // [batch][n] inverse pivots (diagonal of U^-1)
  int n, ld, nblk;
  const int* skip = nullptr;  // [batch] or null: the sweeps leave problems with skip[b] != 0 untouched
};

// Right-hand-side block [batch][n][T] addressed with a row stride (elements) and an item stride.
template <typename S>
struct TriVec {
  S* p;
  size_t item_stride;
  int row_stride;
};

// Inverse diagonal blocks of every problem (one wavefront per (problem, block)).
template <typename S>
int tri_diag_inverse(const TriFactors<S>& f, int batch, bool need_lower, hipStream_t stream);

// x = L^-1 y (forward sweep): y is overwritten, x written to out.
template <typename S>
int tri_forward(const TriFactors<S>& f, int batch, int T, TriVec<S> y, TriVec<S> out, hipStream_t stream);

// x = U^-1 y (backward sweep): y is overwritten, x written to out; with `ladder` set, x / beta
// goes to the ladder output of items item0.. instead (out unused).
template <typename S>
int tri_backward(const TriFactors<S>& f, int batch, int T, TriVec<S> y, TriVec<S> out, const LadderArgs* ladder,
                 int item0, hipStream_t stream);

}  // namespace dgami
