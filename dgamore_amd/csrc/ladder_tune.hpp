// Kernel choice of the batched ladder solve per (compound size n, right-hand sides, precision).
//
// A kernel id is family * 256 + variant. The host-side autotuner times the candidates of a
// shape and stores the fastest in a process-wide table; the solve entry points read it on
// every call and fall back to the built-in size rule when the shape has no entry.
#pragma once

namespace dgami {

enum LadderFamily { LADDER_FAM_REG = 1, LADDER_FAM_COL = 2, LADDER_FAM_BLOCKED = 3 };

constexpr int ladder_kernel_id(int family, int variant) { return family * 256 + variant; }
constexpr int ladder_kernel_family(int id) { return id / 256; }
constexpr int ladder_kernel_variant(int id) { return id % 256; }

// Blocked-LU variants (panel width NB, rows per panel thread, panels per outer block):
//   complex128: 0 = NB 32, one row per thread, 2 panels (n <= 768); 1 = NB 16, three rows, 2 panels
//               2, 3 = 0, 1 with 128-row update tiles for long updates (64-row tiles otherwise)
//               4 = NB 16, three rows, 4 panels (64-column outer blocks for n > 768)
//   complex64:  0 = NB 32, three rows, 4 panels (n <= 1152);        1 = NB 16, three rows, 4 panels
//               2 = NB 32, three rows, 8 panels (n <= 1152);        3 = NB 16, three rows, 8 panels
//               4 = NB 32, two rows, 4 panels (n <= 1024);      5 = NB 32, three rows, 6 panels (n <= 1152)
constexpr int BLK_VARIANTS = 6;
bool blocked_variant_fits(bool complex64, int variant, int n);

// Tuned kernel id of (n, nrhs, precision), or -1 when the autotuner has not set one.
int ladder_tuned_kernel(int n, int nrhs, int precision);

// Register and column-cyclic variant limits (defined with their kernels).
int ladder_reg_variant_count();
bool ladder_reg_variant_fits(int variant, int n, int nrhs);
int ladder_col_variant_count();
bool ladder_col_variant_fits(int variant, int n, int nrhs);

}  // namespace dgami
