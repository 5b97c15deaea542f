// layout.hpp - constants of the device data layout, shared by the kernels
// and the host code that builds their arrays.
#pragma once

#include <limits.h>

namespace ccfd {

constexpr int kBlock = 256;       // elements per block (one per thread)
constexpr int kMaxBC = 32;        // boundary conditions of one mesh
constexpr int kMaxSlots = 4;      // sides per element
constexpr int kNoSide = INT_MIN;  // link of a slot the element does not have
// elements x sides a mesh or arena may hold: the kernels index [4][S][Np]
// slot arrays with 32-bit ints, so 4 * S * Np must stay below 2^31
constexpr long kMaxIndexedElems = (INT_MAX - 4 * kBlock) / 4;

}  // namespace ccfd
