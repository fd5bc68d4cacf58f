/*
 * libdgami -- MI355X (gfx950) kernels for the nonlocal ladder-DGA hot path of
 * DGAmore (Julpe/DGAmore). C ABI: plain pointers, sizes and a stream handle.
 *
 * Conventions
 * -----------
 * - Complex numbers are dgami_c128: interleaved (re, im) doubles, the memory
 *   layout of numpy.complex128.
 * - Every array pointer is a DEVICE pointer owned by the caller (hipMalloc or
 *   a torch tensor on cuda:0); the library never frees caller memory. Scratch
 *   memory is a grow-only per-process workspace owned by the library.
 * - `stream` is a hipStream_t passed as void* (NULL: the default stream).
 *   Calls are asynchronous with respect to the host.
 * - Return value: 0 on success; -1 invalid argument (text in
 *   dgami_last_error()); -2 HIP runtime error; >0 numerical problem (the
 *   number of singular matrices; per-matrix codes are in the `info` array).
 *   The Python layer maps -1 to ValueError, -2 to RuntimeError and >0 to
 *   numpy.linalg.LinAlgError, the errors the reference raises.
 *
 * Layout names
 * ------------
 * REF2  two-point object in the reference layout [nx][ny][nz][nb][nb][2*niv]
 *       (GreensFunction.mat / SelfEnergy.mat, C order).
 * QF    4-orbital object per (q, w, v): [nw][nb^4][nv][nq]  (q fastest).
 * QW    4-orbital object per (q, w):    [nw][nb^4][nq].
 * QO    4-orbital object per q:         [nq][nb^4].
 * nb^4 is in "M order": orbital tuple (o1,o2,o3,o4) is stored at
 *       ((o1*nb + o2)*nb + o4)*nb + o3, i.e. the reference's compound block
 *       [(o1,o2),(o4,o3)] (four_point.py:184-221) row-major, so that the
 *       reference's contraction X_abcd Y_dcef is the matrix product M(X)M(Y).
 * CMP   compound ladder matrix per (w): [nw][n][n], n = nb^2 * 2*niv, rows
 *       (o1,o2,v) and cols (o4,o3,v') with v fastest (four_point.py:216-219).
 * Momentum index k = (x*ny + y)*nz + z everywhere.
 */
#ifndef DGAMI_H
#define DGAMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double re, im;
} dgami_c128;

/* Complex64: interleaved (re, im) floats, the layout of numpy.complex64. */
typedef struct {
  float re, im;
} dgami_c64;

/* Ladder precision modes (dgami_ladder_solve_sum). */
enum {
  DGAMI_PLAIN_FP64 = 0,    /* complex128 LU, the reference's algorithm in FP64 */
  DGAMI_MIXED_FP32_IR = 1, /* FP32 LU + FP64 residual iterative refinement */
  DGAMI_MIXED_BF16X3_IR = 2 /* split-BF16 (3-term) MFMA trailing updates + FP64 IR */
};

/* Negative per-matrix codes in the `info` arrays of the mixed-precision ladder forms: the FP64
 * refinement of the item did not reach its tolerance within the maximum number of refinement
 * steps (STALLED), or its correction became NaN or infinite (NONFINITE). Positive codes k + 1
 * still report an exactly singular pivot at elimination step k. The reference raises
 * numpy.linalg.LinAlgError for a singular ladder matrix (four_point.py:636-638); the Python layer
 * re-solves STALLED / NONFINITE items with the plain FP64 LU and raises for singular ones. */
#define DGAMI_INFO_REFINE_STALLED (-1)
#define DGAMI_INFO_REFINE_NONFINITE (-2)

/* Text of the last error on the calling host thread ("" if none). */
const char* dgami_last_error(void);

/* ABI version of the library (incremented on incompatible changes). */
int dgami_version(void);

/* Workspace holds for captured HIP graphs. A captured graph names the library's workspace
 * buffers by address; a later call that needs a larger buffer would normally free the old one.
 * While at least one hold is taken, outgrown buffers are kept instead and freed only when the
 * last hold is released. hold() returns the number of holds after taking one; release()
 * returns the number left (-1 if none was held). No counterpart in the reference, which
 * captures nothing; the Python layer takes a hold per captured graph and releases it when the
 * graph is dropped. */
int dgami_workspace_hold(void);
int dgami_workspace_release(void);

/* Kernel timing of the trailing-matrix updates of the blocked ladder LU (the dominant
 * kernel for compound sizes above 128). No counterpart in the reference: it is the
 * measurement facility behind the roofline figures. While enabled, every trailing-update
 * launch is bracketed by HIP events on its launch stream and its algorithmic work is
 * accumulated: bytes = C tile read + written + the A and B panels read once, flops =
 * 8 m n k. dgami_profile_enable(1) resets the counters; dgami_profile_read waits for the
 * recorded events and returns the totals (time in ms). */
int dgami_profile_enable(int on);
int dgami_profile_read(double* ms, double* bytes, double* flops, int64_t* launches);

/* Batched FFT over momentum grids. in/out: [batch][nx*ny*nz] (may alias).
 * sign -1: forward, unnormalised, exp(-i k.R); sign +1: backward exp(+i k.R);
 * the result is multiplied by `scale`.
 * Replaces scipy.fft.fftn / ifftn over the momentum axes
 * (n_point_base.py:509-551, mpi_utils.py:499-542, eliashberg_solver.py:421-428). */
int dgami_fft_grid(const dgami_c128* in, dgami_c128* out, int64_t batch, int nx, int ny, int nz, int sign,
                   double scale, void* stream);

/* Bare bubble on the irreducible q-points, FFT form:
 *   chi0[w][abcd][v][q] = -beta/N sum_k G_ad(k, v) G_cb(k - q, v - w),
 *   w in [0, niw], v in [-niv, niv), then |Re|,|Im| < 1e-12 -> 0.
 * g_k: REF2 with 2*niv_g frequencies, niv_g >= niv + niw. chi0: QF with nv = 2*niv, nq = nq_irr.
 * irrk_ind: [nq_irr] flat momentum index of each irreducible q (KGrid.irrk_ind).
 * Replaces BubbleGenerator.create_generalized_chi0_q_fft_{cpu,gpu,auto} (bubble_gen.py:33-144). */
int dgami_bubble_irrq(int nx, int ny, int nz, int nb, int niv_g, int niw, int niv, double beta,
                      const dgami_c128* g_k, const int32_t* irrk_ind, int nq_irr, dgami_c128* chi0, void* stream);

/* dgami_bubble_irrq without any device allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The G(k) re-layout and its momentum FFT live in a file-scope
 * device array of DGAMI_BUBBLE_NOALLOC_MAX_BYTES per GPU (held until the process ends; its
 * address is looked up once per device on the first call there). Same arguments and results as
 * dgami_bubble_irrq; sizes needing more than the maximum (2 * nb^2 * 2(niv + niw) * nx*ny*nz * 16
 * bytes) are refused with -1. Every call on a GPU shares that GPU's array, so calls that overlap
 * on one GPU (on different streams) race; nothing in it needs to start at zero. */
#define DGAMI_BUBBLE_NOALLOC_MAX_BYTES ((size_t)64 << 20)
int dgami_bubble_irrq_noalloc(int nx, int ny, int nz, int nb, int niv_g, int niw, int niv, double beta,
                              const dgami_c128* g_k, const int32_t* irrk_ind, int nq_irr, dgami_c128* chi0,
                              void* stream);

/* chi0 post-processing for one SDE iteration:
 *   chi0_full_sum = (1/beta^2) sum_v chi0           (all 2*niv_full frequencies)
 *   chi0_core_sum = (1/beta^2) sum_{|v|<niv_core} chi0
 *   chi0_inv      = per-(q,w,v) inverse of the core blocks [(o1,o2),(o4,o3)]
 * chi0: QF with nv = 2*niv_full; chi0_inv: QF with nv = 2*niv_core; sums: QW. Any output may be NULL.
 * Replaces cut_niv + FourPoint.invert(copy=False) for num_vn=1 + sum_over_all_vn
 * (nonlocal_sde.py:824-835, four_point.py:601-613, local_four_point.py:164-177). */
int dgami_chi0_prepare(int nq, int nb, int nw, int niv_full, int niv_core, double beta, const dgami_c128* chi0,
                       dgami_c128* chi0_inv, dgami_c128* chi0_full_sum, dgami_c128* chi0_core_sum, void* stream);

/* Batched ladder solves with the nu' sum, for nchan channels at once:
 *   A(c,q,w) = chi0_inv(q,w,v) delta_vv' + (gamma[c][w] - u_r[c] - v_r[c][q]) / beta^2   (CMP)
 *   chi_aux_sum[c][w][(o1,o2),(o4,o3)][v][q] = (1/beta) sum_v' [A^-1]_{(o1,o2,v),(o4,o3,v')}
 * computed as an LU solve with the nb^2 indicator right-hand sides.
 * chi0_inv: QF (nv = 2*niv); gamma: [nchan] x CMP; u_r: [nchan][nb^4] (M order);
 * v_r: [nchan] x QO or NULL (V = 0); chi_aux_sum: [nchan] x QF; info: [nchan][nw][nq] int32,
 * 0 or k+1 for an exactly singular pivot at elimination step k; the mixed forms also report
 * DGAMI_INFO_REFINE_STALLED / DGAMI_INFO_REFINE_NONFINITE per item. Nothing is read back to the
 * host in any precision, so every form can be captured in a HIP graph.
 * precision: DGAMI_PLAIN_FP64 / DGAMI_MIXED_FP32_IR / DGAMI_MIXED_BF16X3_IR.
 * Replaces create_auxiliary_chi_r_q_sum_v1/_v3 (nonlocal_sde.py:85-156) with
 * FourPoint.invert_and_sum_over_last_vn/_v2 (four_point.py:620-684). */
int dgami_ladder_solve_sum(int nchan, int nq, int nb, int nw, int niv, double beta, const dgami_c128* chi0_inv,
                           const dgami_c128* gamma, const dgami_c128* u_r, const dgami_c128* v_r, int precision,
                           dgami_c128* chi_aux_sum, int32_t* info, void* stream);

/* One model of dgami_ladder_solve_sum_multi: the arguments of dgami_ladder_solve_sum (plain FP64). */
typedef struct {
  int nchan, nq, nb, nw, niv;
  double beta;
  const dgami_c128* chi0_inv;
  const dgami_c128* gamma;
  const dgami_c128* u_r;
  const dgami_c128* v_r;
  dgami_c128* chi_aux_sum;
  int32_t* info;
} dgami_ladder_model;

/* Plain-FP64 ladder solves of several independent models (different orbital counts, frequency
 * boxes and momentum grids) in one call. `models` is a HOST array of nmodels descriptors whose
 * device pointers follow dgami_ladder_solve_sum. The models whose compound size takes the
 * register-resident kernel are solved together, up to 16 models per launch: one workgroup per
 * (model, chan, w, q) of the concatenated work list, the model table in the kernel arguments
 * (so a captured HIP graph keeps it). The larger models go through the blocked LU, one call
 * each. Outputs and info per model as in dgami_ladder_solve_sum. No counterpart in the
 * reference, which loops over its models one run at a time; it replaces that loop for a
 * batch of models. */
int dgami_ladder_solve_sum_multi(int nmodels, const dgami_ladder_model* models, void* stream);

/* dgami_ladder_solve_sum_multi with the blocked-LU models solved in `precision`
 * (DGAMI_PLAIN_FP64 / DGAMI_MIXED_FP32_IR / DGAMI_MIXED_BF16X3_IR, as in dgami_ladder_solve_sum,
 * info included); the register-resident models (compound size <= 64 by default) stay FP64.
 * No counterpart in the reference, as dgami_ladder_solve_sum_multi. */
int dgami_ladder_solve_sum_multi_prec(int nmodels, const dgami_ladder_model* models, int precision, void* stream);

/* Autotuner table of the ladder solves. A kernel id is family * 256 + variant, family 1 = register
 * tiles, 2 = column-cyclic, 3 = blocked LU (variant: panel shape). candidates() writes up to `cap`
 * ids that can run a compound size n with nrhs right-hand sides in `precision` and returns how many
 * there are; set() stores the choice for (n, nrhs, precision) in a process-wide table that every
 * later ladder call reads (kernel_id -1 removes it); get() returns the stored id or -1; clear()
 * empties the table. The environment variable DGAMI_LADDER_KERNEL still overrides the table. No
 * counterpart in the reference (NumPy picks its LAPACK kernels itself); the host-side autotuner
 * times the candidates and fills the table. */
int dgami_ladder_tune_candidates(int n, int nrhs, int precision, int* ids, int cap);
int dgami_ladder_tune_set(int n, int nrhs, int precision, int kernel_id);
int dgami_ladder_tune_get(int n, int nrhs, int precision);
int dgami_ladder_tune_clear(void);

/* Dense-matrix variant of the ladder solve: a_cmp holds nmat assembled compound matrices
 * [nmat][n][n] (n = nb^2 * 2*niv); out[nmat][n][nb^2] = (1/beta) sum_v' [A^-1]_{(o1,o2,v),(o4,o3,v')}
 * at row (o1,o2,v), column (o4,o3). info: [nmat].
 * Replaces FourPoint.invert_and_sum_over_last_vn / _v2 applied to an assembled object
 * (four_point.py:620-684). */
int dgami_solve_sum_dense(int nmat, int nb, int niv, double beta, const dgami_c128* a_cmp, int precision,
                          dgami_c128* out, int32_t* info, void* stream);

/* Full inverse of the ladder matrices of one channel (complex128 blocked LU carrying the
 * identity as n right-hand sides, then block back substitution):
 *   X(q,w) = A^-1,  A = chi0^-1 delta_vv' + (Gamma_r - U_r - V_r^q)/beta^2   (as in dgami_ladder_solve_sum)
 * chi0_inv: QF; gamma: CMP [nw][n][n]; u_r: [nb^4] M order; v_r: QO or NULL;
 * out: [nw][nq][n][n] compound (rows (o1,o2,v), cols (o4,o3,v')); info: [nw][nq] or NULL.
 * Replaces create_auxiliary_chi_r_q with FourPoint.invert (nonlocal_sde.py:72-82, four_point.py:590-600). */
int dgami_ladder_invert(int nq, int nb, int nw, int niv, double beta, const dgami_c128* chi0_inv,
                        const dgami_c128* gamma, const dgami_c128* u_r, const dgami_c128* v_r, dgami_c128* out,
                        int32_t* info, void* stream);

/* Full vertex of one channel in particle-hole notation, products in compound indices:
 *   F_r(q,w) = beta^2 (chi0^-1 - chi0^-1 X chi0^-1) + (u - u chi_r u) . (vrg * vrg),  u = U_r + V_r^q
 * with X as in dgami_ladder_invert, (vrg * vrg)[(a,b,v),(f,e,v')] = sum_dc vrg_abcd(v) vrg_dcef(v'),
 * vrg: QF and chi_r: QW (the outputs of dgami_vrg_chi_phys). vrg = chi_phys = NULL: first term only.
 * out: [nw][nq][n][n] compound; info: [nw][nq] or NULL.
 * Replaces create_full_vertex_q_r / create_full_vertex_q_r_v2 (eliashberg_solver.py:78-147, 175-209). */
int dgami_full_vertex(int nq, int nb, int nw, int niv, double beta, const dgami_c128* chi0_inv,
                      const dgami_c128* gamma, const dgami_c128* u_r, const dgami_c128* v_r, const dgami_c128* vrg,
                      const dgami_c128* chi_phys, dgami_c128* out, int32_t* info, void* stream);

/* dgami_full_vertex with the LU in a chosen precision and an optional frequency window:
 * DGAMI_PLAIN_FP64 is dgami_full_vertex;
 * DGAMI_MIXED_FP32_IR / DGAMI_MIXED_BF16X3_IR factor in complex64 with FP32 / split-BF16
 * matrix-core trailing updates, without refinement (refining n right-hand sides would cost
 * more than the complex128 LU). Their n right-hand sides are the vertex part
 * A - chi0^-1 = (gamma - u_r - v_r)/beta^2, so that the first term is formed as
 * beta^2 chi0^-1 X (A - chi0^-1), which equals beta^2 (chi0^-1 - chi0^-1 X chi0^-1) without its
 * cancellation; the output sums run in FP64.
 * v_lo < v_hi (0 <= v_lo, v_hi <= 2 niv) restricts the output to the rows and columns whose
 * fermionic index (0..2 niv-1) lies in [v_lo, v_hi), for every orbital pair: only those right-hand
 * sides are carried through the LU, and the other entries of `out` are left as they were (v_lo =
 * v_hi = 0: the whole vertex). The pairing transform of a range of bosonic frequencies
 * [w0, w1) reads only the window [w0 + niv - np, niv + np).
 * Replaces the same reference code as dgami_full_vertex, for the pairing vertices of the
 * Eliashberg equation (precision 1 or 2 where the leading eigenvalue tolerates ~1e-6 relative
 * error in F). */
int dgami_full_vertex_prec(int nq, int nb, int nw, int niv, double beta, int precision, int v_lo, int v_hi,
                           const dgami_c128* chi0_inv, const dgami_c128* gamma, const dgami_c128* u_r,
                           const dgami_c128* v_r, const dgami_c128* vrg, const dgami_c128* chi_phys,
                           dgami_c128* out, int32_t* info, void* stream);

/* Singlet and triplet pairing vertices at w_pp = 0 from the particle-hole full vertices of the
 * density and magnetic channels (f_dens, f_magn: [nw][nq][n][n], half bosonic range, n = nb^2*2niv):
 *   F_pp[q][o1][o2][o3][o4][i][j] = -F(w=i-j)[(o1,o2,i+off)][(o4,o3,2np-1-j+off)]        (w >= 0)
 *                                 = -conj F(j-i)[(o1,o2,2np-1-i+off)][(o4,o3,j+off)]     (w < 0)
 *   off = niv - np, zero where |w| >= nw;
 *   gamma_sing = 0.5 F_pp,dens - 1.5 F_pp,magn,  gamma_trip = 0.5 F_pp,dens + 0.5 F_pp,magn (or NULL).
 * f_magn = gamma_trip = NULL: gamma_sing receives F_pp of f_dens alone.
 * Output layout [nq][nb][nb][nb][nb][2np][2np], the input of dgami_pairing_setup. Needs np <= niv.
 * Replaces transform_vertex_q_frequencies_w0 (eliashberg_solver.py:44-58, 69-75) and the channel
 * combination of solve (eliashberg_solver.py:620-629). */
int dgami_pairing_vertex_pp(int nq, int nb, int nw, int niv, int np, const dgami_c128* f_dens,
                            const dgami_c128* f_magn, dgami_c128* gamma_sing, dgami_c128* gamma_trip, void* stream);

/* The same transform for one range of bosonic frequencies: f_dens / f_magn hold w in [w0, w0 + nw)
 * ([nw][nq][n][n]); the entries of gamma_sing / gamma_trip whose |w| = |i - j| lies in that range
 * are written, all others are left untouched. Called over consecutive ranges on zeroed outputs it
 * gives dgami_pairing_vertex_pp of all w, while the full vertex of only one range is held at a time
 * (at 48x48, 2 orbitals and 2np = 128 the full vertex of all w would take 176 GB per channel).
 * Replaces the same reference code as dgami_pairing_vertex_pp, which transforms each rank's
 * complete F (eliashberg_solver.py:44-58, 69-75, 620-629). */
int dgami_pairing_vertex_pp_range(int nq, int nb, int nw, int w0, int niv, int np, const dgami_c128* f_dens,
                                  const dgami_c128* f_magn, dgami_c128* gamma_sing, dgami_c128* gamma_trip,
                                  void* stream);

/* Batched inverse of nmat dense complex128 matrices a [nmat][n][n] (row major) into a_inv (same
 * layout; may not alias a), n <= 1536: the blocked LU with partial pivoting and n right-hand
 * sides of dgami_ladder_invert on given matrices. info: [nmat] (k + 1 for an exactly singular
 * pivot at step k) or NULL. Replaces the dense FourPoint/LocalFourPoint.invert of the local
 * vertices (local_sde.py:32-51, 62-66; four_point.py:590-600), which calls np.linalg.inv. */
int dgami_invert_dense(int nmat, int n, const dgami_c128* a, dgami_c128* a_inv, int32_t* info, void* stream);

/* Local generalized susceptibility of one channel from the DMFT two-particle Green's function:
 *   chi = beta g2,  density (dens != 0): chi(w = 0) -= 2 beta G_ab(v) G_cd(v')
 * g2, chi: CMP [nw][n][n] (w >= 0, n = nb^2 2 niv); g: [nb][nb][2 niv_g] (centred; NULL when
 * dens = 0). Replaces create_generalized_chi (local_sde.py:17-29). */
int dgami_generalized_chi(int nb, int nw, int niv, int niv_g, double beta, int dens, const dgami_c128* g2,
                          const dgami_c128* g, dgami_c128* chi, void* stream);

/* Local irreducible vertex from the inverted susceptibilities:
 *   gamma = beta^2 (chi_inv - chit_inv) + u_r                  CMP [nw][nc][nc], nc = nb^2 2 niv_core
 * chit_inv is the inverse of the shell-corrected chi~ cut to the core (or of chi0 without a shell);
 * chit_inv NULL: chi_inv holds gamma itself (only the padding is done). u_r: [nb^4] M order.
 * gamma_pad (or NULL): the same vertex on the full box CMP [nw][nf][nf], nf = nb^2 2 niv_full,
 * equal to u_r outside the core; gamma may be NULL when gamma_pad is given. Replaces the combination step of
 * create_gamma_r / create_gamma_r_with_shell_correction (local_sde.py:32-51) and pad_with_u
 * (local_four_point.py:652-668). */
int dgami_local_gamma(int nb, int nw, int niv_core, int niv_full, double beta, const dgami_c128* chi_inv,
                      const dgami_c128* chit_inv, const dgami_c128* u_r, dgami_c128* gamma, dgami_c128* gamma_pad,
                      void* stream);

/* Local self-energy from the Schwinger-Dyson equation with the local three-leg vertices:
 *   inner_r = vrg_r - vrg_r . U_r . chi_r,  inner = (inner_dens - inner_magn) / 2 (w < 0 by symmetry)
 *   Sigma_ij(v) = -(1/beta) sum_{w = -niw..niw} U_kjop inner_ilpo(w, v) G_lk(v - w)
 *               + sum_cd U_dens,ijcd n_dc                                  (when occ != NULL)
 * vrg_*: QF with nq = 1 ([niw+1][nb^4][2 niv]); chi_*: QW with nq = 1; u_dens / u_magn: [nb^4] M
 * order; u: [nb^4] reference order; g: [nb][nb][2 niv_g], niv_g >= niv + niw; occ: [nb][nb] or
 * NULL; sigma: [nb][nb][2 niv]. nb <= 4. Replaces get_loc_self_energy_vrg (local_sde.py:150-172). */
int dgami_local_sigma(int nb, int niw, int niv, int niv_g, double beta, const dgami_c128* vrg_dens,
                      const dgami_c128* vrg_magn, const dgami_c128* chi_dens, const dgami_c128* chi_magn,
                      const dgami_c128* u_dens, const dgami_c128* u_magn, const dgami_c128* u, const dgami_c128* g,
                      const dgami_c128* occ, dgami_c128* sigma, void* stream);

/* dgami_g_full with mu read from device memory (a double), so that a captured iteration can
 * carry the chemical potential that dgami_update_mu changes on the device. */
int dgami_g_full_dmu(int nk, int nb, int niv, double beta, const double* mu, const dgami_c128* ek,
                     const dgami_c128* sigma, int sigma_nk, int sigma_niv, dgami_c128* g, void* stream);

/* Batched inverse of nmat nb^2 x nb^2 blocks stored element-major: in/out [nb^4 (M order)][nmat].
 * info: [nmat] (1 where a pivot was exactly zero) or NULL.
 * Replaces FourPoint.invert for objects without fermionic axes (four_point.py:207-209, 615-618)
 * and LocalFourPoint.invert of such objects. */
int dgami_block_invert(int nmat, int nb, const dgami_c128* in, dgami_c128* out, int32_t* info, void* stream);

/* Lattice Green's function G(k, v) = [(i v + mu) 1 - e_k - Sigma(k, v)]^-1 on the box
 * v in [-niv, niv). ek: [nk][nb][nb]; sigma: REF2 with sigma_nk = 1 (local) or nk points
 * and 2*sigma_niv >= 2*niv frequencies (centred cut); g: REF2 [nk][nb][nb][2*niv].
 * Replaces GreensFunction.get_g_full (greens_function.py:125-142). */
int dgami_g_full(int nk, int nb, int niv, double beta, double mu, const dgami_c128* ek, const dgami_c128* sigma,
                 int sigma_nk, int sigma_niv, dgami_c128* g, void* stream);

/* Three-leg vertex and shell-corrected physical susceptibility of one channel:
 *   vrg      = beta chi0_inv . chi_aux_sum                              (QF)
 *   chi_phys = ((sum_v chi_aux_sum / beta + chi0_full_sum - chi0_core_sum)^-1 + u_r + v_r)^-1  (QW)
 * u_r: [nb^4] M order; v_r: QO or NULL. info: [nw][nq], nonzero where a block was singular.
 * Replaces create_vrg_r_q + sum_over_all_vn + create_generalized_chi_q_with_shell_correction
 * (nonlocal_sde.py:159-182, 289-303). */
int dgami_vrg_chi_phys(int nq, int nb, int nw, int niv, double beta, const dgami_c128* chi0_inv,
                       const dgami_c128* chi_aux_sum, const dgami_c128* chi0_full_sum,
                       const dgami_c128* chi0_core_sum, const dgami_c128* u_r, const dgami_c128* v_r,
                       dgami_c128* vrg, dgami_c128* chi_phys, int32_t* info, void* stream);

/* Self-energy kernel of one channel, accumulated:
 *   K_r = u . (vrg - vrg . u . chi_phys - [magn] 2/3 1),  u = u_r + v_r(q)
 *   kernel = accumulate ? kernel + weight K_r : weight K_r                   (QF)
 * channel: 0 density, 1 magnetic.
 * Replaces calculate_kernel_r_q (nonlocal_sde.py:201-215) and the channel sum
 * kernel += K_dens + 3 K_magn (nonlocal_sde.py:838-857). */
int dgami_kernel_r(int nq, int nb, int nw, int niv, int channel, double weight, const dgami_c128* vrg,
                   const dgami_c128* chi_phys, const dgami_c128* u_r, const dgami_c128* v_r, int accumulate,
                   dgami_c128* kernel, void* stream);

/* Double-counting kernel, written with the reference's sign (kernel = -K_dc):
 *   K_dc[q][abef][w][p] = sum_{cd,v} (1/beta^2) (U_adcb-permuted . chi0)_{abcd}(q,w,v) f_dc_{dcef}(w,v,p),
 *   f_dc = 2 F_magn,loc[abcd->cbad], then p cut to [niv_full-niv_core, niv_full+niv_core).
 * chi0: QF (nv = 2*niv_full); u: [nb^4] in the reference order abcd;
 * f_magn_loc: reference layout [nb][nb][nb][nb][nw][2*niv_full][2*niv_full];
 * kernel: QF with nv = 2*niv_core (overwritten).
 * Replaces calculate_sigma_dc_kernel and its caller's prefactors (nonlocal_sde.py:185-198, 817-820). */
int dgami_dc_kernel(int nq, int nb, int nw, int niv_full, int niv_core, double beta, const dgami_c128* chi0,
                    const dgami_c128* u, const dgami_c128* f_magn_loc, dgami_c128* kernel, void* stream);

/* Kernel entries of the full zone with the per-k orbital transformation of the reference's
 * auto symmetry mode: entry j is the irreducible kernel src[j] transformed by U_j (and
 * conjugated when conj[j] != 0),
 *   ext[w][(o1,o2),(o4,o3)][v][j] = U_o1p1 U*_o2p2 U*_o4p4 U_o3p3 K[w][(p1,p2),(p4,p3)][v][src[j]].
 * kernel_irr: QF with nq entries; src: [next] int32; u: [next][nb][nb]; conj: [next] int32 or NULL;
 * kernel_ext: QF with next entries. dgami_sigma_from_kernel then takes kernel_ext with an
 * irrk_inv that maps every k to its entry. Replaces the orbital step of _map_to_full_bz and
 * exchange_and_map_irrbz_fullbz (n_point_base.py:462-507, mpi_utils.py:270-290) with
 * apply_auto_orbital_transform (symmetry_reduction.py:791-903). */
int dgami_kernel_orbital_expand(int nq, int next, int nb, int nw, int nv, const dgami_c128* kernel_irr,
                                const int32_t* src, const dgami_c128* u, const int32_t* conj, dgami_c128* kernel_ext,
                                void* stream);

/* Non-local self-energy from the kernel, R-space FFT form, in the full fermionic range:
 *   Sigma_ij(k, v) = -1/(2 beta N) sum_{q, w in [-niw,niw]} sum_{ad} K_aijd(q, w, v) G_ad(k+q, v-w),
 *   v >= 0, and Sigma(k, -v) = conj(Sigma(k, v)).
 * kernel_irr: QF on the irreducible q (nv = 2*niv_core); irrk_inv: [nx*ny*nz] (KGrid.irrk_inv);
 * g_k: REF2 with 2*niv_g frequencies (niv_g >= niw + niv_core); sigma: REF2 with 2*niv_core.
 * Replaces map_irrbz_fullbz + calculate_sigma_from_kernel_fft_{cpu,gpu,auto} and the caller's
 * ifft / to_full_niv_range (mpi_utils.py:123-131, nonlocal_sde.py:544-674, 874-879). */
int dgami_sigma_from_kernel(int nx, int ny, int nz, int nb, int niw, int niv_core, int niv_g, double beta,
                            const dgami_c128* kernel_irr, const int32_t* irrk_inv, int nq_irr, const dgami_c128* g_k,
                            dgami_c128* sigma, void* stream);

/* Hartree and Fock terms of the self-energy (complex128):
 *   H_ab    = 2 sum_cd (U + V^{q=0})_abcd n_dc
 *   F_ab(k) = -(1/N) sum_q sum_cd (U + V^q)_abcd n_dc(k + q)   (as an FFT cross-correlation)
 * u: [nb^4] and v: [nk][nb^4] (full momentum grid, q = 0 first, or NULL) in the reference orbital
 * order abcd; occ: [nb][nb] k-mean occupation; occ_k: [nk][nb][nb]; hartree: [nb][nb]; fock: [nk][nb][nb].
 * Replaces get_hartree_fock (nonlocal_sde.py:31-69, 782-783). */
int dgami_hartree_fock(int nx, int ny, int nz, int nb, const dgami_c128* u, const dgami_c128* v,
                       const dgami_c128* occ, const dgami_c128* occ_k, dgami_c128* hartree, dgami_c128* fock,
                       void* stream);

/* Local moments of a self-energy: mom0 = mean Re Sigma_loc and mom1 = mean Im Sigma_loc * v over the
 * top n_fit = max(int(0.2 niv), 4) positive frequencies of the k-mean Sigma_loc.
 * sigma: [nk][nb][nb][2 niv]; mom0, mom1: [nb][nb] (imaginary parts zero).
 * Replaces SelfEnergy.fit_smom (self_energy.py:51-75). */
int dgami_fit_smom(int nk, int nb, int niv, double beta, const dgami_c128* sigma, dgami_c128* mom0, dgami_c128* mom1,
                   void* stream);

/* Chemical-potential update on the device: the secant iteration of update_mu on
 *   f(mu) = 2 Tr[ Fermi(beta (Re hloc + smom0 - mu)) + (1/nk) sum_k occ_k(mu) ] - target
 * (occ_k: the Matsubara sums of dgami_fill_sums with the model hloc) from mu_old = *mu, with the
 * derivative-free steps, tolerance (|p - p1| <= 1e-6) and 50-iteration limit of
 * scipy.optimize.newton, then *mu = mixing mu_new + (1 - mixing) mu_old, or *mu unchanged when the
 * iteration fails (the reference's update_mu returns nan then and its caller keeps mu).
 * mu: one double in device memory (in/out); info: three doubles or NULL: mu_new (nan on failure),
 * status (1 converged, 2 failed), the largest |smom0 - smom0^T| met (the model's density matrix
 * is formed from the symmetric part of smom0, exact for a symmetric Sigma). The number of kernel
 * launches is fixed (evaluations after convergence return at once), so the call can be captured.
 * hloc, smom0: [nb][nb] (real parts used); ek, sigma as in dgami_fill_sums.
 * Replaces update_mu / root_fun / get_total_fill (greens_function.py:20-80) and the mu mixing of
 * the self-consistency loop (nonlocal_sde.py:895-911). */
int dgami_update_mu(int nk, int nb, int niv, double beta, double target, double mixing, const dgami_c128* ek,
                    const dgami_c128* sigma, int sigma_nk, const dgami_c128* hloc, const dgami_c128* smom0, double* mu,
                    double* info, void* stream);

/* Occupations of the k-dependent model at the device mu:
 *   occ_k = Fermi(beta (Re e_k + smom0 - mu)) + Matsubara part of dgami_fill_sums (e_k model),
 *   occ = k-mean of occ_k, entries below 1e-12 set to zero (complex outputs, imaginary parts zero).
 * occ_k: [nk][nb][nb]; occ: [nb][nb]. Replaces GreensFunction.get_fill_nonlocal
 * (greens_function.py:212-240) as called by the self-consistency loop (nonlocal_sde.py:911-914). */
int dgami_occupations(int nk, int nb, int niv, double beta, const dgami_c128* ek, const dgami_c128* sigma,
                      int sigma_nk, const dgami_c128* smom0, const double* mu, dgami_c128* occ_k, dgami_c128* occ,
                      void* stream);

/* New self-energy of one iteration and its linear mixing:
 *   val[k][a][b][v] = sigma_nl + H_ab + F_ab(k) + delta_sigma_ab(v)   for |v| < niv_core (centred)
 *                   = sigma_dmft_ab(v)                                otherwise (2*niv_out frequencies)
 *   out = alpha val + (1 - alpha) sigma_old   (sigma_old NULL: out = val)
 * sigma_nl: [nk][nb][nb][2*niv_core]; delta_sigma: [nb][nb][2*niv_core] or NULL; hartree/fock or NULL;
 * sigma_dmft: [nb][nb][2*niv_dmft] (centred cut to niv_out); sigma_old, out: [nk][nb][nb][2*niv_out].
 * Replaces sigma + hartree + fock + delta_sigma, concatenate_self_energies and linear mixing
 * (nonlocal_sde.py:884-893, 1145-1151; self_energy.py:153-168). */
int dgami_sigma_assemble(int nk, int nb, int niv_core, int niv_out, const dgami_c128* sigma_nl,
                         const dgami_c128* hartree, const dgami_c128* fock, const dgami_c128* delta_sigma,
                         const dgami_c128* sigma_dmft, int niv_dmft, const dgami_c128* sigma_old, double alpha,
                         dgami_c128* out, void* stream);

/* Matsubara part of the occupations:
 *   occ_k[k][a][b] = (1/beta) sum_{v in [-niv, niv)} Re[G(k,v) - G_model(k,v)]_ab
 *   G = (i v + mu - e_k - Sigma(k,v))^-1,  G_model = (i v + mu - h - smom0)^-1
 * h = e_k (hloc NULL, get_fill_nonlocal) or hloc (get_total_fill); occ_mean (or NULL) = k-mean of occ_k.
 * ek: [nk][nb][nb]; sigma: [sigma_nk][nb][nb][2*niv], sigma_nk = 1 or nk; smom0, hloc: [nb][nb];
 * occ_k: [nk][nb][nb] doubles; occ_mean: [nb][nb] doubles. The density-matrix term of the model
 * (a Fermi function of an nb x nb matrix per k) is added by the caller.
 * Replaces the frequency sums of get_total_fill and GreensFunction.get_fill_nonlocal
 * (greens_function.py:20-51, 212-240). */
int dgami_fill_sums(int nk, int nb, int niv, double beta, double mu, const dgami_c128* ek, const dgami_c128* sigma,
                    int sigma_nk, const dgami_c128* smom0, const dgami_c128* hloc, double* occ_k, double* occ_mean,
                    void* stream);

/* out[e] = (1/nk) sum_k x[k][e], e < m (the k-mean of SelfEnergy.fit_smom, self_energy.py:60). */
int dgami_kmean(int nk, int64_t m, const dgami_c128* x, dgami_c128* out, void* stream);

/* Particle-particle bubble at w = 0: chi0pp[k][abcd][v] = G_ad(k,v) conj(G_cb(k,v)) on the centred 2*np of
 * the 2*niv_g frequencies of g_k (REF2 [nk][nb][nb][2*niv_g]), |Re|,|Im| < 1e-12 -> 0; chi0pp: [nk][nb^4][2*np]
 * (reference orbital order abcd), the input of dgami_pairing_matvec.
 * Replaces BubbleGenerator.create_generalized_chi0_q_pp_w0 (bubble_gen.py:268-279). */
int dgami_chi0_pp(int nk, int nb, int niv_g, int np, const dgami_c128* g_k, dgami_c128* chi0pp, void* stream);

/* Pairing-vertex preparation for the linearised Eliashberg operator:
 *   g_hat = s fft_q Gamma_pp (q mapped irr -> FBZ), g_hat_flip = flip_k flip_v' perm(abcd->adcb) g_hat,
 * stored for the matvec as [nk][(a,b,v)][(c,d,v')] (row (a,b,v) = out, col (c,d,v') = in), n = nb^2*2np.
 * gamma_pp_irr: [nq_irr][nb][nb][nb][nb][2np][2np] (reference layout, pp orbital order acbd).
 * channel: 0 singlet (s=+1), 1 triplet (s=-1).
 * Replaces the set-up of solve_eliashberg_lanczos (eliashberg_solver.py:390-398). */
int dgami_pairing_setup(int nx, int ny, int nz, int nb, int np2, int channel, const dgami_c128* gamma_pp_irr,
                        const int32_t* irrk_inv, int nq_irr, dgami_c128* g_hat, dgami_c128* g_hat_flip, void* stream);

/* One application of the linearised gap operator to nvec vectors:
 *   gg = fft_k(chi0pp . gap); y = ifft_k( (0.5/(N beta)) [g_hat gg + s g_hat_flip flip_k(gg)] ).
 * chi0pp: [nk][nb^4 (reference order abcd)][np2]; gap, out: [nvec][nk][nb][nb][np2] (reference layout).
 * Replaces mv() of solve_eliashberg_lanczos (eliashberg_solver.py:420-428). */
int dgami_pairing_matvec(int nx, int ny, int nz, int nb, int np2, int channel, double beta, const dgami_c128* g_hat,
                         const dgami_c128* g_hat_flip, const dgami_c128* chi0pp, int nvec, const dgami_c128* gap,
                         dgami_c128* out, void* stream);

/* The same operator with the transformed pairing vertex stored in complex64 (half the bytes
 * of the HBM-bound matvec); vectors, products and sums stay complex128. The vertex is
 * transformed in complex128 and rounded once when stored.
 * Replaces the same reference code as dgami_pairing_setup / dgami_pairing_matvec. */
int dgami_pairing_setup_c64(int nx, int ny, int nz, int nb, int np2, int channel, const dgami_c128* gamma_pp_irr,
                            const int32_t* irrk_inv, int nq_irr, dgami_c64* g_hat, dgami_c64* g_hat_flip,
                            void* stream);
int dgami_pairing_matvec_c64(int nx, int ny, int nz, int nb, int np2, int channel, double beta, const dgami_c64* g_hat,
                             const dgami_c64* g_hat_flip, const dgami_c128* chi0pp, int nvec, const dgami_c128* gap,
                             dgami_c128* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DGAMI_H */
