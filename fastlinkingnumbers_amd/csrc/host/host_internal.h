// Internal helpers of the drop-in host library.
#pragma once
#include <vector>

#include "curve.h"
#include "flk.h"

namespace VerifyCurves {

// Bisects spline (sub)segments until no piece of a curve overlaps a piece of a
// potentially linked curve, then returns the closed polylines through the
// pieces' start points (reference src/discretize.cpp:102-281).  links[i] holds
// every curve whose box overlaps curve i (both directions).
std::vector<Curve> Discretize(const std::vector<Curve> &curves,
                              const std::vector<std::vector<int>> &links, bool verbose);

// Process-wide engine context (created on first use, one per device).
flk_ctx *engine(int device);
void engine_check(flk_ctx *ctx, int32_t rc, const char *what);

}  // namespace VerifyCurves
