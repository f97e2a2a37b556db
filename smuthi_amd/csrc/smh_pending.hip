// Entry points of include/smh.h whose kernels are not in this build yet.
#include "smh_op.h"

extern "C" int smh_volume_create(smh_ctx* ctx, int64_t, const double*, const int*, const int*, int, int, double,
                                 double, double, double, int64_t, int64_t, int, smh_op** out) {
  if (out) *out = nullptr;
  return smh_fail(ctx, SMH_ERR_UNSUPPORTED, "volume lookup operator not built");
}
