// Split-BF16 MFMA cost (placeholder until the MFMA tile lands).
#include "common.h"

namespace wotk {
size_t cost_bf16x3_scratch(int64_t, int64_t, int64_t) { return 0; }
int launch_cost_bf16x3(const double *, const double *, int64_t, int64_t, int64_t, double *, int64_t,
                       void *, hipStream_t) {
    return fail(WOTK_EINVAL, "bf16x3 cost mode is not available in this build");
}
}  // namespace wotk
