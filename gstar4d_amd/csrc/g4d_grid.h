// gStar4D for MI355X — grid-stage launchers (Morton, PBA, pairs, worksets).
#pragma once

#include "g4d_prims.h"

namespace g4d {

// The point sets of one call (batched mode; a single triangulation is one set).  Set k holds
// points [pt_off[k], pt_off[k+1]) in input order and, after the Morton sort, the same range of
// sorted positions; it is scaled by its own bounds and has its own G^3 grid at vox_off[k].
struct GridSets {
    int nsets;
    const int *pt_set;      // per point: its set; null when nsets == 1
    const int64_t *pt_off;  // nsets + 1
    const int *g;           // nsets
    const int64_t *vox_off; // nsets + 1
    const double *bounds;   // 2 per set: min and max over the set's coordinates
    // work decompositions (prefix sums over the sets)
    const int64_t *pba_blk; // nsets + 1: PBA work-groups
    const int64_t *spill;   // nsets + 1: PBA spill ints
    const int64_t *pair_thr;// nsets + 1: grid-pair threads
    int key_shift;          // one set: Morton keys are code << key_shift | index
};
// Batched Morton keys: set (16 bits) | code (24 bits, G <= 256) | index within the set (24 bits).
static constexpr int BATCH_MAX_GRID = 256;
static constexpr int64_t BATCH_MAX_SET_POINTS = (int64_t)1 << 24;
static constexpr int BATCH_MAX_SETS = 1 << 16;

// Bounds: one set -> part[0], part[1] (part >= 768 doubles of scratch); several sets ->
// bounds[2k], bounds[2k+1] per set.  A non-finite coordinate makes the minimum NaN.
hipError_t launch_bounds(const double *xyz, int n, double *part /* >= 768 doubles */, hipStream_t s);
hipError_t launch_set_bounds(const double *xyz, const GridSets &S, double *bounds, hipStream_t s);
// pt_set[k] = set of point k.
hipError_t launch_set_ids(const int64_t *pt_off, int nsets, int *pt_set, hipStream_t s);
// Morton keys: one set: code << key_shift | i, code 3 bits per level of the grid (sort bits
// key_shift..64); several sets: the batched layout (sort bits 24..64).
int morton_key_shift(int g);
hipError_t launch_morton_keys(const double *xyz, int n, const GridSets &S, uint64_t *keys, hipStream_t s);
hipError_t launch_gather_sorted(const double *xyz, const uint64_t *keys, int n, const GridSets &S, double *pts,
                                int *perm, uint64_t *vox, hipStream_t s);
hipError_t launch_dup_check(const double *pts, const uint64_t *vox, int n, const GridSets &S, unsigned long long *cnt,
                            hipStream_t s);
hipError_t launch_fill_i32(int *a, int64_t n, int v, hipStream_t s);
// Leaders into the grids: site codes into grid and point ids into idgrid, or, with idgrid null
// (required above G4D_CODE_GRID), point ids into grid.
hipError_t launch_scatter_sites(const uint64_t *vox, int n, const GridSets &S, int *grid, int *idgrid, hipStream_t s);
// PBA work-groups and stack-spill ints (per-column overflow of the in-LDS envelope stack) of a
// G^3 grid.
int64_t pba_blocks(int g);
int64_t pba_spill_ints(int g);
// One axis pass (2 = z, 1 = y, 0 = x) over every set's grid; with idgrid the output is point
// ids.  nblocks = pba_blk[nsets].  With vox (one set wider than G4D_CODE_GRID) the cells hold
// point ids throughout and the sites' coordinates are read from vox.
hipError_t pba_pass(int axis, const GridSets &S, int64_t nblocks, const int *in, int *out, const int *idgrid,
                    int *spill, const uint64_t *vox, hipStream_t s);
// The y (1) or x (0) pass in `bands` bands (2..PBA_MAX_BANDS): band envelopes, their merge, and
// the colouring of each band (same output as pba_pass).  benv has the spill area's size; bcnt
// holds 64 * bands ints per work-group of pba_pass, fcnt 64.
static constexpr int PBA_MAX_BANDS = 16;
hipError_t pba_pass_banded(int axis, const GridSets &S, int64_t nblocks, int bands, const int *in, int *out,
                           const int *idgrid, int *spill, int *benv, int *bcnt, int *fcnt, const uint64_t *vox,
                           hipStream_t s);
// Ordered pairs (min<<32 | max) of each non-leader point with its voxel's leader.
hipError_t launch_missing(const uint64_t *vox, int n, const GridSets &S, const int *grid, uint64_t *out,
                          unsigned long long *counter, int64_t cap, hipStream_t s);
// Voronoi-neighbour pairs (min<<32 | max): pass write=false fills count[thread], the caller
// scans it to offsets (in place), pass write=true writes the pairs.  nthreads = pair_thr[nsets].
int grid_pair_chunk(int g);
int64_t grid_pair_threads(int g, int zc);
hipError_t launch_grid_pairs(const int *grid, const GridSets &S, int64_t nthreads, int64_t *count, uint64_t *out,
                             bool write, hipStream_t s);
// Pass 1 with each 256-thread tile's pairs deduplicated on chip: tile b writes into its raw
// region [off[256 b], off[256 (b + 1)]) and reports (distinct, overflow) counts in cnt[b];
// launch_tile_totals + an exclusive scan give the dense offsets for launch_gather_tiles.
hipError_t launch_grid_pairs_dedup(const int *grid, const GridSets &S, int64_t nthreads, const int64_t *off,
                                   uint64_t *out, uint2 *cnt, hipStream_t s);
hipError_t launch_tile_totals(const uint2 *cnt, int64_t ntiles, int64_t *tot, hipStream_t s);
hipError_t launch_gather_tiles(const uint64_t *src, const int64_t *off, const uint2 *cnt, int64_t ntiles,
                               const int64_t *doff, uint64_t *dst, hipStream_t s);
hipError_t launch_csr_offsets(const uint64_t *keys, int64_t m, int n, int64_t *off, hipStream_t s);
// Pumping; a star short of candidates after its neighbours' lists takes the lowest ids of its
// own set.
hipError_t launch_pump(const uint64_t *keys, const int64_t *off, int n, const GridSets &S, int min_ws, uint64_t *out,
                       unsigned long long *counter, int64_t cap, hipStream_t s);

} // namespace g4d
