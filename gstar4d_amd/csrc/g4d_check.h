// gStar4D for MI355X — mesh validation launchers.
#pragma once

#include "g4d_prims.h"

namespace g4d {

// Accumulator words of a validation run.
enum check_acc : int {
    CHECK_ORIENT = 0,     // orient3d > 0
    CHECK_FLAT,           // orient3d == 0
    CHECK_INSPHERE,       // adjacent pairs with the opposite vertex strictly inside
    CHECK_ADJ,            // adjacency disagreements / invalid vertex ids
    CHECK_HULL,           // faces without a neighbour
    CHECK_SHARED,         // face slots with a neighbour (each shared face counted twice)
    CHECK_VERTS,          // distinct vertices used
    CHECK_EMPTY,          // (point, tetrahedron) pairs violating the empty sphere
    CHECK_EMPTY_TESTS,    // exact in-sphere tests of the empty-sphere check
    CHECK_NUM = 16
};

hipError_t launch_check_tets(const double *pts, const int *tets, const int *opp, const int *ref_opp, int64_t nt, int n,
                             unsigned char *used, unsigned long long *acc, unsigned long long *cnt, hipStream_t s);
hipError_t launch_check_edges(const int *tets, int64_t nt, uint64_t *keys, hipStream_t s);
hipError_t launch_check_count_used(const unsigned char *used, int n, unsigned long long *acc, hipStream_t s);
hipError_t launch_check_cells(const double *pts, int n, const double *box, int R, int *cid,
                              unsigned long long *cellcnt, hipStream_t s);
hipError_t launch_check_cell_fill(const int *cid, int n, const int64_t *start, int *fill, int *items, hipStream_t s);
hipError_t launch_check_empty(const double *pts, const int *tets, int64_t nt, const double *box, int R,
                              const int64_t *start, const int *items, unsigned long long *acc, unsigned long long *cnt,
                              hipStream_t s);

} // namespace g4d
