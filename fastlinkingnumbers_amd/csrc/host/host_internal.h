// Internal helpers of the drop-in host library.
#pragma once
#include <vector>

#include "curve.h"
#include "flk.h"

namespace VerifyCurves {

// Process-wide engine context (created on first use, one per device).
flk_ctx *engine(int device);
void engine_check(flk_ctx *ctx, int32_t rc, const char *what);

}  // namespace VerifyCurves
