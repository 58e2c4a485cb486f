/*
 * gStar4D CPU oracle — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference gStar4D path (points in -> exact 3D Delaunay
 * tetrahedra with face adjacency out).  It is the parity checker for the HIP library and
 * the CPU baseline leg of bench.py.  Nothing in the product library links, loads or calls
 * it; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may.
 *
 * Each function cites the reference file:line whose behaviour it restates (paths are
 * relative to the reference repository root; GD/ = GDelaunay/).
 *
 * Where the reference's intermediate choice depends on storage order or on a write race
 * (PBA tie-breaks GD/PBA/pba3DKernel.h:80-123, racy leader GD/PBA/pba3DKernel.h:518-542,
 * atomic pump placement GD/GDelaunay/GDelKernels.cu:326-345, proof vertex choice
 * GD/GDelaunay/GDelPredKernels.cu:440-520) the oracle uses a canonical min-index rule.
 * These choices change only intermediate work, never the Delaunay output, which is
 * fixed by the exact predicates, the SoS table, the Morton ranks and the flat-tetra drop.
 *
 * Parity pins (see tests/): orient3d / insphere signs vs golden vectors from the
 * reference's own predicates.c and vs rational arithmetic; SoS vs a brute-force
 * epsilon-polynomial expansion; PBA vs brute-force nearest site; the full pipeline vs
 * brute-force empty-sphere enumeration (small N), scipy/Qhull (non-degenerate N) and an
 * empty-sphere validator (large N).
 */
#ifndef GSTAR4D_ORC_H
#define GSTAR4D_ORC_H

/* Widest PBA grid: 11-bit voxel coordinates. */
#define ORC_MAX_GRID 2048

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- predicates (orc_predicates.c) ---- */
int orc_orient3d_exact(const double *pa, const double *pb, const double *pc, const double *pd);
int orc_insphere_exact(const double *pa, const double *pb, const double *pc, const double *pd,
                       const double *pe);
int orc_orient3d(const double *pa, const double *pb, const double *pc, const double *pd);
int orc_insphere(const double *pa, const double *pb, const double *pc, const double *pd,
                 const double *pe);
/* Exact sign of det[[c_0 .. c_{k-2} 1]] over k rows; cols[j] in {0,1,2,3}, 3 = lifted. */
int orc_sign_minor(int k, const double *const *rows, const int *cols);
/* SoS orientation of 5 lifted points with indices (GD/GDelaunay/GDelPredDevice.h:142-698).
   Returns +1/-1, never 0.  depth_out (optional) receives the SoS depth used (0 = exact). */
int orc_orient4_sos(const double *pts, int i0, int i1, int i2, int i3, int i4, int *depth_out);
/* Same but with explicit coordinates (for predicate unit tests). */
int orc_orient4_sos_xyz(const double *p0, const double *p1, const double *p2, const double *p3,
                        const double *p4, int i0, int i1, int i2, int i3, int i4, int *depth_out);
int64_t orc_pred_counter(int which); /* 0 orient4 calls, 1 exact calls, 2 SoS calls */
void orc_pred_counter_reset(void);

/* ---- pipeline (orc_pipeline.c) ---- */
typedef struct {
    int32_t grid_size;      /* PBA grid width G (<= ORC_MAX_GRID) */
    int32_t min_workset;    /* MinWorksetSize, 8 in the reference */
    int32_t max_rounds;     /* splaying round limit (guard), e.g. 1000 */
    int32_t verbose;
    int32_t insert_order;   /* 0: each star's list in ascending id (the reference's order);
                               1: nearest to the star point first (the GPU grouped kernel's) */
} orc_options;

typedef struct {
    int32_t num_points;
    int32_t num_missing;
    int64_t num_grid_pairs;     /* after sort-unique */
    int64_t num_workset;        /* final workset entries (both directions) */
    int32_t max_workset;
    int32_t num_pumped;
    int32_t splay_rounds;
    int64_t splay_insertions;
    int32_t num_lowup_tetra;
    int32_t num_tetra;
    int64_t orient4_calls, exact_calls, sos_calls;
    double t_morton, t_pba, t_star, t_splay, t_out; /* seconds */
} orc_stats;

/* Morton order of points (GD/GDelaunay/GDelData.cu:51-131): perm[k] = input index of the
   k-th point in Morton order (stable: ties keep input order).  vox receives the voxel
   coordinates (int, 3 per input point) from the grid scaling. */
int orc_morton_order(const double *xyz, int32_t n, int32_t grid, int32_t *perm, int32_t *vox);
/* Grid scaling: voxel = (int)((G-3)*(x-mn)/(mx-mn) + 1); mn/mx over all coordinates. */
void orc_scale_to_voxels(const double *xyz, int32_t n, int32_t grid, int32_t *vox);

/* PBA digital Voronoi (GD/PBA/pba3DHost.cu:200-295): sites given as voxel coords of the
   Morton-sorted points (leader = lowest index in a voxel).  out_grid (G^3 int32) receives the
   nearest site's point index per voxel. */
int orc_pba(const int32_t *vox_sorted, int32_t n, int32_t grid, int32_t *out_grid);
/* Site-code grid before id conversion (for kernel parity tests); codes x<<20|y<<10|z. */
int orc_pba_codes(const int32_t *vox_sorted, int32_t n, int32_t grid, int32_t *out_codes);

/* Pairs (min<<32 | max, Marker = -1) of one 2x2x2 cube of labels v[8], corner c at offset
   (c & 1, c >> 1 & 1, c >> 2) (GD/GDelaunay/GDelKernels.cu:196-270); returns the count (<= 31),
   duplicates included, in the reference's emission order. */
int orc_cube_pairs(const int32_t *v, uint64_t *out);
/* Voronoi-neighbour pairs from the id grid (GD/GDelaunay/GDelKernels.cu:137-299), sorted and
   unique, each as (min<<32 | max).  Returns count, or -needed if cap too small. */
int64_t orc_grid_pairs(const int32_t *grid, int32_t g, uint64_t *out, int64_t cap);

/* Threads of the parallel stages (PBA columns, grid-pair slabs, star creation, drowned
   checks, proofs, insertions); < 1 selects every processor.  Returns the count in use.  The
   result does not depend on it: every parallel loop owns whole stars or columns and its
   outputs are placed in input order. */
int orc_set_threads(int nthreads);

/* Full triangulation.  xyz: input points (N x 3 FP64).  Output tetrahedra in MORTON index
   space, orient3d < 0, vertex order (owner-star triangle v0 v1 v2, owner star).
   tets_out (4*cap) may be NULL to query the count.  Returns tetra count or <0 on error. */
int64_t orc_triangulate(const double *xyz, int32_t n, const orc_options *opt,
                        int32_t *perm_out, int32_t *tets_out, int64_t cap, orc_stats *stats);

/* The same with extra candidate pairs (a << 32 | b, Morton index space) added to the
   worksets in both directions before pumping (the learned proposal's plug-in point,
   GD/GDelaunay/GDelHost.cu:352-529); pairs with a == b or ids out of range are ignored. */
int64_t orc_triangulate_ex(const double *xyz, int32_t n, const orc_options *opt, const uint64_t *extra,
                           int64_t n_extra, int32_t *perm_out, int32_t *tets_out, int64_t cap, orc_stats *stats);

/* Canonical form: map tets through perm into input index space, sort each quadruple
   ascending, sort the list lexicographically.  In place. */
void orc_canonicalize(int32_t *tets, int64_t nt, const int32_t *perm);
/* Face adjacency by face matching: opp[4*t+i] = tetra across the face opposite tets[4*t+i],
   or -1.  Works on any tetra list. */
int orc_adjacency(const int32_t *tets, int64_t nt, int32_t n, int32_t *opp);

/* Validator (GD/Common/Geometry.cu:200-393 plus a full empty-sphere test).
   Returns 0 if all checks pass; bit flags otherwise.  counts (optional, 4):
   [euler V-E+F-T, bad orientation, bad adjacent-insphere, empty-sphere violations]. */
int orc_validate(const double *xyz, int32_t n, const int32_t *tets, int64_t nt, int full_empty,
                 int64_t *counts);

#ifdef __cplusplus
}
#endif
#endif
