// Frechet::Continuous::distance_cuda: the reference's host driver
// (include/frechet.hpp:14, src/frechet.cpp:18-75), here running the gfx950
// sweeps through the C ABI in rfd.h.
#pragma once

#include "curve.hpp"

namespace Frechet {
namespace Continuous {
// Bracket [VE-FD, DFD], bisection while (ub - lb > eps), result (ub + lb) / 2.
// Prints the reference's progress lines.  The decisions follow the
// reference's path choice (.cu:87-96): reachable<D>'s interval form when
// forceOversize is set or its interval arrays would not fit in free device
// memory, else the two-phase interval form.  Returns NaN on error.
double distance_cuda(const Curve &curve1, const Curve &curve2, const double eps,
                     bool forceOversize);
}  // namespace Continuous
}  // namespace Frechet
