// gStar4D for MI355X — diagnostic kernel launchers.
#pragma once

#include "g4d_prims.h"

namespace g4d {

hipError_t launch_diag_orient4(const double *pts5, const int *idx5, int64_t T, int mode, int *sign, int *depth,
                               unsigned long long *cnt, hipStream_t s);
hipError_t launch_diag_orient3d(const double *pts4, int64_t T, int *sign, unsigned long long *cnt, hipStream_t s);
// Plain device copy of n16 16-byte words (reads and writes each byte once).
hipError_t launch_diag_copy(const int4 *in, int4 *out, int64_t n16, hipStream_t s);

} // namespace g4d
