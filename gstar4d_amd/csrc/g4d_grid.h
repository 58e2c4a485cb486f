// gStar4D for MI355X — grid-stage launchers (Morton, PBA, pairs, worksets).
#pragma once

#include "g4d_prims.h"

namespace g4d {

hipError_t launch_bounds(const double *xyz, int n, double *part /* >= 768 doubles */, hipStream_t s);
hipError_t launch_morton_keys(const double *xyz, int n, int g, const double *bounds, uint64_t *keys, hipStream_t s);
hipError_t launch_gather_sorted(const double *xyz, const uint64_t *keys, int n, int g, const double *bounds,
                                double *pts, int *perm, int *vox, hipStream_t s);
hipError_t launch_dup_check(const double *pts, const int *vox, int n, unsigned long long *cnt, hipStream_t s);
hipError_t launch_fill_i32(int *a, int64_t n, int v, hipStream_t s);
hipError_t launch_scatter_sites(const int *vox, int n, int g, int *grid, int *idgrid, hipStream_t s);
// Ints of the PBA stack-spill workspace (per-column overflow of the in-LDS envelope stack).
int64_t pba_spill_ints(int g);
// One axis pass (2 = z, 1 = y, 0 = x); with idgrid the output is point ids.
hipError_t pba_pass(int axis, int g, const int *in, int *out, const int *idgrid, int *spill, hipStream_t s);
// Three axis passes; result (point ids, or site codes when idgrid is null) ends in `tmp`.
hipError_t pba_run(int g, int *grid, int *tmp, const int *idgrid, int *spill, hipStream_t s);
// Ordered pairs (min<<32 | max) of each non-leader point with its voxel's leader.
hipError_t launch_missing(const int *vox, int n, int g, const int *grid, uint64_t *out, unsigned long long *counter,
                          int64_t cap, hipStream_t s);
// Voronoi-neighbour pairs (min<<32 | max): pass write=false fills count[thread], the caller
// scans it to offsets (in place), pass write=true writes the pairs.
int64_t grid_pair_threads(int g, int zc);
hipError_t launch_grid_pairs(const int *grid, int g, int zc, int64_t *count, uint64_t *out, bool write, hipStream_t s);
hipError_t launch_csr_offsets(const uint64_t *keys, int64_t m, int n, int64_t *off, hipStream_t s);
hipError_t launch_pump(const uint64_t *keys, const int64_t *off, int n, int min_ws, uint64_t *out,
                       unsigned long long *counter, int64_t cap, hipStream_t s);

} // namespace g4d
