// Entry point of the batched ladder solves: picks the register-resident
// kernel for small compound sizes and the blocked workspace LU otherwise.
#include <algorithm>
#include <cstdlib>
#include <string>
#include <vector>

#include "dgami.h"
#include "dgami_internal.hpp"
#include "dgami_launch.hpp"
#include "ladder_common.hpp"
#include "ladder_tune.hpp"

namespace dgami {

int ladder_reg_variant(int n, int nrhs);
int ladder_reg_launch(int variant, const LadderArgs& a, hipStream_t stream);
int ladder_col_variant(int n, int nrhs);
int ladder_col_launch(int variant, const LadderArgs& a, hipStream_t stream);
int ladder_blocked_fp64(const LadderArgs& a, int variant, hipStream_t stream);
int ladder_reg_launch_multi(int variant, const LadderMulti& mm, hipStream_t stream);

// Kernel of the plain FP64 solve. The environment variable DGAMI_LADDER_KERNEL (col, reg or
// blocked) forces one family for comparisons; otherwise an autotuned choice for (n, nrhs)
// wins; otherwise the size rule: "reg" (2D register tiles) up to compound size 64, the
// blocked workspace LU above it. Measured on 8192 matrices per call, the register kernel
// leads at n = 40 (0.96 vs 1.85 ms) and ties at n = 64, while the blocked LU is 1.7-2.7x
// faster from n = 72 on. The column-cyclic kernel ("col", n <= 128) is the alternative.
constexpr int LADDER_REG_MAX_N = 64;

// Register variant a plain FP64 solve of (n, nrhs) runs in, or -1 when it takes another family.
static int fp64_reg_choice(int n, int nrhs) {
  const char* force = getenv("DGAMI_LADDER_KERNEL");
  if (force && *force) return std::string(force) == "reg" ? ladder_reg_variant(n, nrhs) : -1;
  const int tuned = ladder_tuned_kernel(n, nrhs, DGAMI_PLAIN_FP64);
  if (tuned >= 0) return ladder_kernel_family(tuned) == LADDER_FAM_REG ? ladder_kernel_variant(tuned) : -1;
  return n <= LADDER_REG_MAX_N ? ladder_reg_variant(n, nrhs) : -1;
}

int ladder_solve_fp64(const LadderArgs& a, hipStream_t stream) {
  const char* force = getenv("DGAMI_LADDER_KERNEL");
  const std::string f = force ? force : "";
  if (f == "col") {
    const int v = ladder_col_variant(a.n, a.T);
    if (v >= 0) return ladder_col_launch(v, a, stream);
  }
  if (f.empty()) {
    const int tuned = ladder_tuned_kernel(a.n, a.T, DGAMI_PLAIN_FP64);
    if (tuned >= 0) {
      const int fam = ladder_kernel_family(tuned), v = ladder_kernel_variant(tuned);
      if (fam == LADDER_FAM_COL && ladder_col_variant_fits(v, a.n, a.T)) return ladder_col_launch(v, a, stream);
      if (fam == LADDER_FAM_BLOCKED) return ladder_blocked_fp64(a, v, stream);
    }
  }
  const int v = fp64_reg_choice(a.n, a.T);
  if (v >= 0) return ladder_reg_launch(v, a, stream);
  return ladder_blocked_fp64(a, -1, stream);
}

}  // namespace dgami

namespace dgami {

// Models of a batch: the register-resident ones share launches in FP64 (at n <= 64 a
// complex64 factorisation would save nothing); the others take the blocked LU in `precision`.
static int ladder_multi(int nmodels, const dgami_ladder_model* models, int precision, hipStream_t s) {
  DGAMI_ARG_CHECK(nmodels >= 0 && (nmodels == 0 || models), "ladder_multi: null model table");
  DGAMI_ARG_CHECK(precision == DGAMI_PLAIN_FP64 || precision == DGAMI_MIXED_FP32_IR ||
                      precision == DGAMI_MIXED_BF16X3_IR,
                  "ladder_multi: unknown precision mode");
  std::vector<LadderArgs> small;
  int rc = 0;
  for (int m = 0; m < nmodels; ++m) {
    const dgami_ladder_model& md = models[m];
    DGAMI_ARG_CHECK(md.chi0_inv && md.gamma && md.u_r && md.chi_aux_sum, "ladder_multi: null pointer in a model");
    DGAMI_ARG_CHECK(md.nchan >= 1 && md.nq >= 1 && md.nb >= 1 && md.nb <= 4 && md.nw >= 1 && md.niv >= 1 &&
                        md.beta > 0,
                    "ladder_multi: invalid sizes in a model");
    LadderArgs a{};
    a.chi0_inv = reinterpret_cast<const double2*>(md.chi0_inv);
    a.gamma = reinterpret_cast<const double2*>(md.gamma);
    a.u_r = reinterpret_cast<const double2*>(md.u_r);
    a.v_r = reinterpret_cast<const double2*>(md.v_r);
    a.out = reinterpret_cast<double2*>(md.chi_aux_sum);
    a.info = md.info;
    a.nchan = md.nchan;
    a.nq = md.nq;
    a.nw = md.nw;
    a.T = md.nb * md.nb;
    a.nv = 2 * md.niv;
    a.n = a.T * a.nv;
    a.inv_beta2 = 1.0 / (md.beta * md.beta);
    a.inv_beta = 1.0 / md.beta;
    a.dense = 0;
    const size_t items = static_cast<size_t>(a.nchan) * a.nw * a.nq;
    DGAMI_ARG_CHECK(items < (size_t(1) << 30), "ladder_multi: too many problems in a model");
    if (fp64_reg_choice(a.n, a.T) >= 0)
      small.push_back(a);
    else if ((rc = precision == DGAMI_PLAIN_FP64 ? ladder_solve_fp64(a, s) : ladder_solve_mixed(a, precision, s)) != 0)
      return rc;
  }
  // the register-resident models, LADDER_MULTI_MAX per launch, in the largest variant of the group
  for (size_t g0 = 0; g0 < small.size(); g0 += LADDER_MULTI_MAX) {
    LadderMulti mm{};
    mm.nmodels = static_cast<int>(std::min(small.size() - g0, static_cast<size_t>(LADDER_MULTI_MAX)));
    int variant = 0;
    size_t total = 0;
    for (int m = 0; m < mm.nmodels; ++m) {
      const LadderArgs& a = small[g0 + m];
      mm.model[m] = a;
      mm.start[m] = static_cast<int>(total);
      total += static_cast<size_t>(a.nchan) * a.nw * a.nq;
      variant = std::max(variant, fp64_reg_choice(a.n, a.T));
    }
    DGAMI_ARG_CHECK(total < (size_t(1) << 31), "ladder_multi: too many problems in one launch");
    mm.start[mm.nmodels] = static_cast<int>(total);
    if ((rc = ladder_reg_launch_multi(variant, mm, s)) != 0) return rc;
  }
  return 0;
}

}  // namespace dgami

extern "C" int dgami_ladder_solve_sum_multi(int nmodels, const dgami_ladder_model* models, void* stream) {
  return dgami::ladder_multi(nmodels, models, DGAMI_PLAIN_FP64, static_cast<hipStream_t>(stream));
}

extern "C" int dgami_ladder_solve_sum_multi_prec(int nmodels, const dgami_ladder_model* models, int precision,
                                                 void* stream) {
  return dgami::ladder_multi(nmodels, models, precision, static_cast<hipStream_t>(stream));
}

extern "C" int dgami_ladder_solve_sum(int nchan, int nq, int nb, int nw, int niv, double beta,
                                      const dgami_c128* chi0_inv, const dgami_c128* gamma, const dgami_c128* u_r,
                                      const dgami_c128* v_r, int precision, dgami_c128* chi_aux_sum, int32_t* info,
                                      void* stream) {
  using namespace dgami;
  DGAMI_ARG_CHECK(chi0_inv && gamma && u_r && chi_aux_sum, "ladder: null pointer");
  DGAMI_ARG_CHECK(nchan >= 1 && nq >= 1 && nb >= 1 && nb <= 4 && nw >= 1 && niv >= 1 && beta > 0,
                  "ladder: invalid sizes");
  DGAMI_ARG_CHECK(precision == DGAMI_PLAIN_FP64 || precision == DGAMI_MIXED_FP32_IR ||
                      precision == DGAMI_MIXED_BF16X3_IR,
                  "ladder: unknown precision mode");
  LadderArgs a;
  a.chi0_inv = reinterpret_cast<const double2*>(chi0_inv);
  a.gamma = reinterpret_cast<const double2*>(gamma);
  a.u_r = reinterpret_cast<const double2*>(u_r);
  a.v_r = reinterpret_cast<const double2*>(v_r);
  a.out = reinterpret_cast<double2*>(chi_aux_sum);
  a.info = info;
  a.nchan = nchan;
  a.nq = nq;
  a.nw = nw;
  a.T = nb * nb;
  a.nv = 2 * niv;
  a.n = a.T * a.nv;
  a.inv_beta2 = 1.0 / (beta * beta);
  a.inv_beta = 1.0 / beta;
  a.dense = 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (precision == DGAMI_PLAIN_FP64) return ladder_solve_fp64(a, s);
  return ladder_solve_mixed(a, precision, s);
}

extern "C" int dgami_solve_sum_dense(int nmat, int nb, int niv, double beta, const dgami_c128* a_cmp, int precision,
                                     dgami_c128* out, int32_t* info, void* stream) {
  using namespace dgami;
  DGAMI_ARG_CHECK(a_cmp && out, "solve_sum_dense: null pointer");
  DGAMI_ARG_CHECK(nmat >= 1 && nb >= 1 && nb <= 4 && niv >= 1 && beta > 0, "solve_sum_dense: invalid sizes");
  DGAMI_ARG_CHECK(precision == DGAMI_PLAIN_FP64 || precision == DGAMI_MIXED_FP32_IR ||
                      precision == DGAMI_MIXED_BF16X3_IR,
                  "solve_sum_dense: unknown precision mode");
  LadderArgs a;
  a.chi0_inv = nullptr;
  a.gamma = reinterpret_cast<const double2*>(a_cmp);
  a.u_r = nullptr;
  a.v_r = nullptr;
  a.out = reinterpret_cast<double2*>(out);
  a.info = info;
  a.nchan = 1;
  a.nq = nmat;
  a.nw = 1;
  a.T = nb * nb;
  a.nv = 2 * niv;
  a.n = a.T * a.nv;
  a.inv_beta2 = 1.0;
  a.inv_beta = 1.0 / beta;
  a.dense = 1;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (precision == DGAMI_PLAIN_FP64) return ladder_solve_fp64(a, s);
  return ladder_solve_mixed(a, precision, s);
}
