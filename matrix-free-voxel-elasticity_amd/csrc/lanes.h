// Element-per-lanes layout control of the parity product (lanes.hip).
#pragma once
#include "common.h"

namespace vh {

// y = K(E) u (unconstrained), y zeroed and the element contributions added with atomics.
template <typename T>
int lanes24_launch(const double* blocks72, const T* E, const T* u, T* y, int nx, int ny, int nz, hipStream_t st);

}  // namespace vh
