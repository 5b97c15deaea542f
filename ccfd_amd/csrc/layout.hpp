// layout.hpp - constants of the device data layout, shared by the kernels
// and the host code that builds their arrays.
#pragma once

#include <limits.h>

namespace ccfd {

constexpr int kBlock = 256;       // elements per block (one per thread)
constexpr int kMaxBC = 32;        // boundary conditions of one mesh
constexpr int kMaxSlots = 4;      // sides per element
constexpr int kNoSide = INT_MIN;  // link of a slot the element does not have

}  // namespace ccfd
