// FP32 inner CG inside FP64 iterative refinement (filled in below).
#include "tfg_internal.h"

namespace tfg {
int solve_ir(tfg_matrix*, int, const double*, const double*, int, int*, double, double*,
             const double*, const tfg_options&, tfg_report*) {
  return TFG_ENOTSUP;
}
}  // namespace tfg
