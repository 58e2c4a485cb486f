// gStar4D for MI355X — the static device scratch array behind the allocation-free entry point.
#pragma once

#include <hip/hip_runtime.h>

namespace g4d {

// Looks up the device address of the scratch array on `device`, which must be the current
// device; once per device (later calls return at once).
hipError_t noalloc_setup(int device);
// The address noalloc_setup found on `device`, or nullptr before it ran there.
void *noalloc_scratch(int device);

} // namespace g4d
