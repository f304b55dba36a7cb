// Shared by the fused multi-operator line sweeps (momentum.hip,
// projection.hip).  The building blocks themselves -- the segment solver of
// the y/z sweeps (seg_solve, YzShape), the wavefront solver of the x sweeps
// (halo_window, wave_solve) -- live in lines_impl.h beside the single-operator
// kernels; chain_depth and same_set are host functions of schemes.cpp.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "lines_impl.h"

namespace x3d {

}  // namespace x3d
