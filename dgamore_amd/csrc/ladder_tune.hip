// Process-wide table of autotuned ladder kernels and its C-ABI (candidates, set, get, clear).
#include <mutex>
#include <vector>

#include "dgami.h"
#include "dgami_internal.hpp"
#include "ladder_blocked.hpp"
#include "ladder_tune.hpp"

namespace dgami {

namespace {
struct TuneEntry {
  int n, nrhs, precision, id;
};
std::mutex g_tune_mutex;
std::vector<TuneEntry> g_tune;  // a few dozen entries at most: one per distinct shape of a run

int candidates(int n, int nrhs, int precision, int* ids, int cap) {
  int k = 0;
  auto add = [&](int id) {
    if (k < cap && ids) ids[k] = id;
    ++k;
  };
  if (precision == DGAMI_PLAIN_FP64) {
    for (int v = 0; v < ladder_reg_variant_count(); ++v)
      if (ladder_reg_variant_fits(v, n, nrhs)) add(ladder_kernel_id(LADDER_FAM_REG, v));
    for (int v = 0; v < ladder_col_variant_count(); ++v)
      if (ladder_col_variant_fits(v, n, nrhs)) add(ladder_kernel_id(LADDER_FAM_COL, v));
  }
  // the blocked LU runs any n, but below one panel width the register kernels are the ones to use
  const bool c64 = precision != DGAMI_PLAIN_FP64;
  for (int v = 0; v < BLK_VARIANTS; ++v)
    if ((n >= 32 || c64) && blocked_variant_fits(c64, v, n)) add(ladder_kernel_id(LADDER_FAM_BLOCKED, v));
  return k;
}
}  // namespace

bool blocked_variant_fits(bool complex64, int variant, int n) {
  if (n < 1 || n > BLK_MAX_N) return false;
  if (variant == 0) return n <= (complex64 ? 1152 : 768);
  if (variant == 1) return true;
  if (!complex64) return variant == 3 || variant == 4 || (variant == 2 && n <= 768);
  return variant == 3 || ((variant == 2 || variant == 5) && n <= 1152) || (variant == 4 && n <= 1024);
}

int ladder_tuned_kernel(int n, int nrhs, int precision) {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  for (const TuneEntry& e : g_tune)
    if (e.n == n && e.nrhs == nrhs && e.precision == precision) return e.id;
  return -1;
}

}  // namespace dgami

using namespace dgami;

static bool valid_precision(int p) {
  return p == DGAMI_PLAIN_FP64 || p == DGAMI_MIXED_FP32_IR || p == DGAMI_MIXED_BF16X3_IR;
}

extern "C" int dgami_ladder_tune_candidates(int n, int nrhs, int precision, int* ids, int cap) {
  DGAMI_ARG_CHECK(n >= 1 && nrhs >= 1 && cap >= 0 && valid_precision(precision), "ladder_tune: invalid key");
  return candidates(n, nrhs, precision, ids, cap);
}

extern "C" int dgami_ladder_tune_set(int n, int nrhs, int precision, int kernel_id) {
  DGAMI_ARG_CHECK(n >= 1 && nrhs >= 1 && valid_precision(precision), "ladder_tune: invalid key");
  if (kernel_id >= 0) {
    int ids[64];
    const int k = candidates(n, nrhs, precision, ids, 64);
    bool ok = false;
    for (int i = 0; i < k && i < 64; ++i) ok = ok || ids[i] == kernel_id;
    DGAMI_ARG_CHECK(ok, "ladder_tune: kernel id is not a candidate for this shape");
  }
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  for (size_t i = 0; i < g_tune.size(); ++i) {
    TuneEntry& e = g_tune[i];
    if (e.n == n && e.nrhs == nrhs && e.precision == precision) {
      if (kernel_id >= 0)
        e.id = kernel_id;
      else
        g_tune.erase(g_tune.begin() + static_cast<long>(i));
      return 0;
    }
  }
  if (kernel_id >= 0) g_tune.push_back(TuneEntry{n, nrhs, precision, kernel_id});
  return 0;
}

extern "C" int dgami_ladder_tune_get(int n, int nrhs, int precision) {
  DGAMI_ARG_CHECK(n >= 1 && nrhs >= 1 && valid_precision(precision), "ladder_tune: invalid key");
  return ladder_tuned_kernel(n, nrhs, precision);
}

extern "C" int dgami_ladder_tune_clear(void) {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  g_tune.clear();
  return 0;
}
