// The drop-in entry point (reference include/wrapper.hpp:4).
#pragma once

#include "curve.hpp"
#include "frechet.hpp"

double compute_distance_parallel(const Curve &, const Curve &, bool forceOversize = false);
