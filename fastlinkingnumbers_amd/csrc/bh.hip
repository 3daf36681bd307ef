// Two-tree Barnes–Hut on the device (placeholder until the traversal lands).
#include "flk_internal.h"

namespace flk {
void bh_run(flk_ctx *, const int32_t *, int64_t, const flk_params &, std::vector<double> &,
            std::vector<double> &, std::vector<uint32_t> &) {
  throw Error(FLK_ERR_STATE, "Barnes-Hut is not available in this build");
}
}  // namespace flk
