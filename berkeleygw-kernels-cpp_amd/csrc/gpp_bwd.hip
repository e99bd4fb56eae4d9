// gpp_bwd.hip -- adjoint of the GPP forward (placeholder until the kernel lands).
#include "bgw_internal.h"

namespace bgw {
int gpp_backward_device(bgw_ctx* ctx, size_t, size_t, size_t, size_t, const int*, const int*, const double*,
                        const double*, const double*, const double*, const double*, const double*, double,
                        double*, double*, double*) {
  return fail(ctx, BGW_EINVAL, "gpp backward not built");
}
}  // namespace bgw
