// Hook-mode launcher and the reference offset lists it walks.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.hpp"
#include "host_setup.hpp"

namespace misa {

// The reference's full lists (even x, odd x) as linear offsets into the
// ghost-extended doubled-x array.
struct FullOffsets {
  int n[2];
  int32_t off[2][kMaxFullOffsets];
};

// which: 0 rho, 1 df, 2 force
void launchHook(int which, const Box& b, double rc, const FullOffsets& o, const Tables<double>& tb, AtomElement* a,
                hipStream_t st);

}  // namespace misa
