// Adjoint products J^T.psi and (dR/dbeta)^T.psi.
#include <hip/hip_runtime.h>

#include "sg_context.hpp"

extern "C" {
int sg_jtv(sg_ctx* ctx, const double*, const double*, double*, double*) {
  return sg_fail(ctx, SG_ERR_STATE, "adjoint not available in this build");
}
int sg_jtv_dev(sg_ctx* ctx, const double*, const double*, double*, double*) {
  return sg_fail(ctx, SG_ERR_STATE, "adjoint not available in this build");
}
}
