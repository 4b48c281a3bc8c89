// Gradient kernels (filled in below).
#include "qj_internal.h"

namespace qj {

qj_status ap_grads(qj_workspace*, const MapArgs&, const qj_rational*, const qj_grid*, const double*, double, double,
                   double*, double*, double*, qj_ap_grads*, hipStream_t) {
  set_error("ap gradients not built");
  return QJ_ERR_INVALID;
}
qj_status points_adjoint(qj_workspace*, const MapArgs&, const qj_rational*, qj_precision, const double*, int64_t,
                         const double*, double*, double*, hipStream_t) {
  set_error("adjoint not built");
  return QJ_ERR_INVALID;
}

}  // namespace qj
