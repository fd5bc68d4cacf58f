// Mixed-precision ladder solves (FP32 / split-BF16 factorisation with FP64
// iterative refinement).
#include "dgami.h"
#include "dgami_internal.hpp"
#include "dgami_launch.hpp"
#include "ladder_common.hpp"

namespace dgami {

int ladder_solve_mixed(const LadderArgs& a, int precision, hipStream_t stream) {
  (void)a;
  (void)precision;
  (void)stream;
  set_error("ladder: mixed-precision modes are not built yet");
  return -1;
}

}  // namespace dgami
