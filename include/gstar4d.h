/*
 * gStar4D for MI355X — C ABI.
 *
 * Drop-in boundary for the reference's triangulation call (GDelaunay/GDelaunay/GDelaunay.h:39-42):
 *   gdelInit(const Config&, Point3HVec& points, Point3HVec& scaled)   -> g4d_create + input
 *   gdelCompute(double& tPba, double& tStar, double& tSplay, double& tOut) -> g4d_triangulate
 *   getTetraFromGpu(TetraMesh&)                                        -> g4d_result / g4d_copy_result
 *   gdelDeInit()                                                       -> g4d_destroy
 * The C++ shim with those four functions is include/gdelaunay_shim.hpp.
 *
 * Points are FP64 xyz (N x 3, array of structs).  Output tetrahedra are int32 quadruples with
 * orient3d(v0, v1, v2, v3) < 0 (the reference's convention, GD/GDelaunay/GDelPredKernels.cu:648-661;
 * GD/Common/Geometry.cu:319-344) and opp[i] = tetrahedron across the face opposite v[i] or -1
 * (GD/Common/Geometry.h:147-151).  The tetrahedron set equals the reference's: the exact
 * Delaunay tetrahedralisation of the points lifted to the paraboloid, degeneracies resolved by
 * Simulation of Simplicity over Morton-order point ranks, flat tetrahedra dropped.
 *
 * Errors: every entry point returns a g4d_status; g4d_last_error() gives a message.  Nothing
 * calls exit().  One context per host thread; calls on one context are serialised.
 */
#ifndef GSTAR4D_H
#define GSTAR4D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct g4d_ctx g4d_ctx;

typedef enum {
    G4D_OK = 0,
    G4D_ERR_INVALID_INPUT = 1,
    G4D_ERR_DUPLICATE_POINTS = 2,
    G4D_ERR_GRID_LIMIT = 3,
    G4D_ERR_OOM = 4,
    G4D_ERR_DEVICE = 5,
    G4D_ERR_NOT_CONVERGED = 6,
    G4D_ERR_CAPACITY = 7,
    G4D_ERR_PREDICATE_RANGE = 8,
    G4D_ERR_INTERNAL = 9,
    G4D_ERR_IO = 10
} g4d_status;

typedef enum {
    G4D_INDEX_INPUT = 0,  /* vertex ids refer to the caller's input order */
    G4D_INDEX_MORTON = 1  /* vertex ids refer to Morton order, as the reference's gdelInit
                             leaves them (GD/GDelaunay/GDelData.cu:125-130) */
} g4d_index_space;

typedef struct {
    int32_t grid_size;          /* PBA grid width G, 4..2048 (Config::_gridSize) */
    int32_t min_workset;        /* MinWorksetSize (GD/GDelaunay/GDelCommon.h:49), 4..64 */
    int32_t max_rounds;         /* splaying round guard */
    int32_t index_space;        /* g4d_index_space */
    int32_t device;             /* HIP device ordinal */
    int32_t compute_adjacency;  /* 1: fill opp */
    int32_t lds_star_limit;     /* 0: default.  Stars whose link plus insertions exceed this
                                   (at most 511) run in the global-memory star kernel; lower
                                   values route more stars there (testing). */
} g4d_options;

typedef struct {
    int64_t num_tets;
    const int32_t *tets; /* 4*num_tets, owned by the context until the next call */
    const int32_t *opp;  /* 4*num_tets or NULL */
    const int32_t *perm; /* num_points: perm[k] = input index of the k-th point in Morton order */
    int32_t num_points;
} g4d_result;

typedef struct {
    double t_morton_ms, t_pba_ms, t_pairs_ms, t_star_ms, t_splay_ms, t_out_ms, t_adj_ms, t_total_ms;
    int32_t num_points;
    int32_t grid_size;
    int32_t splay_rounds;      /* LogLoopNum in the reference */
    int32_t max_workset;
    int32_t big_stars;         /* star passes routed to the large-star LDS kernel */
    int32_t huge_stars;        /* star passes routed to the global-memory star kernel */
    int64_t num_missing;
    int64_t num_grid_pairs;    /* sort-unique Voronoi-neighbour pairs (one direction) */
    int64_t num_workset;       /* workset entries, both directions, after pumping */
    int64_t num_pumped;
    int64_t splay_insertions;  /* reciprocated insertions over all rounds */
    int64_t num_tets;
    int64_t exact_calls;       /* orient4 filter failures resolved exactly */
    int64_t sos_calls;         /* exact zeros resolved by SoS */
    int64_t exact_stars;       /* star passes handed to the exact-capable star kernel */
    int64_t pba_bytes;         /* algorithmic PBA bytes: 24 G^3 */
    double t_pba_pass_ms[3];   /* per axis pass (z, y, x) */
    int64_t star_store_bytes;  /* star arenas in use at the end (vertices, triangles, history) */
    int32_t store_repacks;     /* times the star arenas were compacted into larger ones */
    int32_t star_passes;       /* star kernel passes: creation plus one per splaying round */
    /* The 32-vertex LDS star kernel (the dominant kernel): launches, summed time (hipEvents on
       the context's stream), cofactor-filter orient4 evaluations and cofactor computations. */
    int32_t star_t0_launches;
    int32_t pad1;
    double t_star_t0_ms;
    int64_t star_t0_filter_evals;
    int64_t star_t0_cofactors;
    /* Learned proposal: enabled for the call, candidate pairs it added (one direction, before
       removing those already in a workset), time of the proposal stage (candidates, features,
       network, merge). */
    int32_t prop_enabled;
    int32_t pba_bands;         /* bands of the y and x PBA passes (1: whole columns) */
    int64_t prop_pairs;
    double t_prop_ms;
    /* Grid stage (site scatter and PBA passes) run as a captured hipGraph: launches of a cached
       graph in this call, and graphs captured for it. */
    int32_t graph_launches;
    int32_t graph_captures;
} g4d_stats;

/* Mesh validation report (GDelaunay/Common/Geometry.cu:200-393 plus adjacency and a full
   empty-sphere test).  A Delaunay tetrahedralisation of points in general position has
   euler == 1 and every failure count 0. */
typedef struct {
    int64_t num_tets, num_vertices, num_edges, num_faces, hull_faces;
    int64_t euler;                 /* V - E + F - T over distinct vertices, edges, faces */
    int64_t orientation_failures;  /* exact orient3d(v0, v1, v2, v3) > 0 */
    int64_t flat_tets;             /* exact orient3d == 0 */
    int64_t insphere_failures;     /* adjacent pairs: the opposite vertex strictly inside */
    int64_t adjacency_failures;    /* opp entries that differ from face matching, bad ids,
                                      faces shared by more than two tetrahedra */
    int64_t empty_sphere_failures; /* (point, tetrahedron) pairs strictly inside; -1: not run */
    int64_t empty_sphere_tests;    /* exact in-sphere tests of that check */
} g4d_check_report;

/* Size limits.  Intermediate key lists (sorted and compacted with 32-bit counts and offsets)
   hold at most G4D_MAX_SORT_KEYS entries, and a result at most G4D_MAX_TETS tetrahedra
   (face keys (t << 2 | i) + 1 are 32-bit).  The largest list is the reciprocated worksets,
   about 20 keys per point for uniform inputs, and a result has about 6.8 tetrahedra per
   point, so uniform inputs reach these limits near 150 M points; a call that would pass them
   returns G4D_ERR_CAPACITY. */
#define G4D_MAX_SORT_KEYS 4294967295LL
#define G4D_MAX_TETS 1073741823LL
/* G4D_OK when lists of sort_keys keys and num_tets tetrahedra are within the limits,
   G4D_ERR_CAPACITY otherwise (host only; no device work). */
int g4d_check_sizes(int64_t sort_keys, int64_t num_tets);

int g4d_default_options(g4d_options *opt);
int g4d_create(g4d_ctx **ctx, const g4d_options *opt);
void g4d_destroy(g4d_ctx *ctx);
const char *g4d_last_error(const g4d_ctx *ctx);
int g4d_device_count(void);

/* Host points in, host tetrahedra out (valid until the next call on ctx). */
int g4d_triangulate(g4d_ctx *ctx, const double *xyz, int32_t n, g4d_result *out);
/* Device-resident points in, device-resident result kept in ctx (for timing on HBM-resident
   input); g4d_copy_result copies the last result to host buffers (any may be NULL). */
int g4d_triangulate_device(g4d_ctx *ctx, const double *d_xyz, int32_t n, int64_t *num_tets);
int g4d_copy_result(g4d_ctx *ctx, int32_t *tets, int32_t *opp, int32_t *perm);
/* Allocation-free triangulation.  No counterpart in the reference: it is for callers that cannot
   allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc).  It runs the same pipeline as g4d_triangulate_device, but takes every
   intermediate buffer from a static device array of G4D_NOALLOC_SCRATCH_BYTES bytes that each
   GPU the library runs on holds until the process ends, and makes no hipMalloc or hipFree.
   Limits: n <= G4D_NOALLOC_MAX_POINTS and the context's grid_size <= G4D_NOALLOC_MAX_GRID
   (G4D_ERR_CAPACITY otherwise); an input that needs more scratch than the array holds returns
   G4D_ERR_OOM.  Outputs are device buffers, in the context's index space, identical to what
   g4d_copy_result gives after g4d_triangulate_device on the same points: d_tets (max_tets x 4
   int32), d_opp (max_tets x 4 int32; required when the context computes adjacency, ignored
   otherwise), d_perm (n int32, Morton index -> input index; may be NULL).  *num_tets (may be
   NULL) receives the count, also when it exceeds max_tets (then G4D_ERR_CAPACITY and nothing
   is written); *scratch_used (may be NULL) the scratch bytes the call took.  All calls on one
   GPU share that GPU's array, so calls that overlap on one GPU race.  The context's own last
   result (g4d_copy_result, the checks) is left as it was. */
#define G4D_NOALLOC_SCRATCH_BYTES (256LL << 20)
#define G4D_NOALLOC_MAX_POINTS 16384
#define G4D_NOALLOC_MAX_GRID 128
int g4d_triangulate_device_noalloc(g4d_ctx *ctx, const double *d_xyz, int32_t n, int32_t *d_tets, int32_t *d_opp,
                                   int64_t max_tets, int32_t *d_perm, int64_t *num_tets, int64_t *scratch_used);
/* Tuning (no reference counterpart; gstar4d_amd/autotune.py measures the choices).
   g4d_set_star_kernel: variant of the 32-vertex star kernel, -1 for the built-in default,
   0 one star per wave, 1 four stars per wave (FP64 permanents), 2 four stars per wave (FP32
   largest permanent), 3 the same at four waves per SIMD, 4 two stars per wave; the output never
   depends on it.  g4d_set_grid_table: PBA grid of a batched set without an explicit grid =
   the smallest multiple of 8 in [32, 256] with at least voxels_per_point[k] voxels per point,
   k the first bucket with n <= max_points[k] (ascending; the last bucket takes every larger
   set); nbuckets = 0 restores g4d_batch_grid_size.  g4d_grid_for_set gives the grid a set of n
   points gets under the context's table. */
int g4d_set_star_kernel(g4d_ctx *ctx, int32_t variant);
int g4d_set_grid_table(g4d_ctx *ctx, int32_t nbuckets, const int64_t *max_points, const float *voxels_per_point);
int g4d_grid_for_set(const g4d_ctx *ctx, int64_t n);
/* Bands of the y and x PBA passes (the reference's m2/m3 bands): 0 picks them by grid width,
   1 walks whole columns, 2..16 split each column (same diagram). */
int g4d_set_pba_bands(g4d_ctx *ctx, int32_t bands);
/* Grid stage (grid fill, site scatter, PBA passes) as a hipGraph captured once per shape and
   buffer set and replayed (up to 16 cached per context): 1 always, 0 never (eager launches,
   with per-pass PBA times in the stats), -1 (default) for batched calls only. */
int g4d_set_stage_graphs(g4d_ctx *ctx, int32_t mode);
/* Upload host points into a device buffer owned by ctx and return its device address. */
int g4d_upload_points(g4d_ctx *ctx, const double *xyz, int32_t n, const double **d_xyz);
int g4d_get_stats(const g4d_ctx *ctx, g4d_stats *st);
int g4d_set_options(g4d_ctx *ctx, const g4d_options *opt);

/* ---- Batched mode (no reference counterpart): many independent point sets, one launch per
   stage for all of them ---- */
typedef struct {
    int32_t num_sets;
    int32_t pad;
    int64_t num_points;          /* total over the sets */
    int64_t num_tets;            /* total over the sets */
    const int64_t *tet_offsets;  /* num_sets + 1: set k's tetrahedra are [tet_offsets[k], tet_offsets[k+1]) */
    const int32_t *tets;         /* 4 * num_tets; vertex ids index the set's own points (input order) */
    const int32_t *opp;          /* 4 * num_tets; tetrahedron ids within the set's range, -1 on the hull;
                                    NULL without adjacency */
    const int32_t *grid_sizes;   /* num_sets: the PBA grid width used for each set */
} g4d_batch_result;

/* Triangulates num_sets independent point sets: set k is xyz[set_offsets[k] .. set_offsets[k+1])
   (set_offsets[0] = 0, every set at least 5 points, at most 2^24 with several sets).  Each set's
   tetrahedra are those g4d_triangulate returns for that set alone with grid_size = grids[k];
   grids (NULL, or num_sets entries) NULL picks them with g4d_batch_grid_size (several sets:
   at most 256).  All sets go
   through every stage together (one set table, one launch per stage).  Arrays belong to ctx
   until its next call. */
int g4d_triangulate_batch(g4d_ctx *ctx, const double *xyz, const int64_t *set_offsets, int32_t num_sets,
                          const int32_t *grids, g4d_batch_result *out);
/* Device-resident points; the result stays in ctx. */
int g4d_triangulate_batch_device(g4d_ctx *ctx, const double *d_xyz, const int64_t *set_offsets, int32_t num_sets,
                                 const int32_t *grids, int64_t *num_tets);
/* The grid width chosen for a set of n points: about 24 voxels per point, in steps of 16, 32..256. */
int g4d_batch_grid_size(int64_t n);

/* ---- Learned proposal (no reference counterpart; it plugs in where the reference builds
   worksets, GD/GDelaunay/GDelHost.cu:352-529, and sizes star storage, :1139-1186).  For every
   point the 8 nearest points among its neighbours in two fine Morton orders are candidates;
   12 features of the digital Voronoi diagram feed an MLP 12 -> 64 -> 64 -> 16 (ReLU; BF16
   weights, FP32 accumulation on the MFMA) whose first 8 outputs score the candidates and
   last 8 the final link-size bin.  Candidates scoring above the threshold join the worksets
   (both directions); predicted link sizes pre-size the star arenas.  Splaying still runs to
   full consistency, so the tetrahedra are those of the plain pipeline. ---- */
#define G4D_PROP_FEATURES 12
#define G4D_PROP_SLOTS 8
#define G4D_PROP_BINS 8
/* Parameters (FP32): W1[32][64] (rows 12..31 zero), b1[64], W2[64][64], b2[64], W3[64][16],
   b3[16]; W[i][o] maps input i to output o. */
#define G4D_MLP_PARAMS 7312
/* enable != 0 with params (G4D_MLP_PARAMS floats, copied) turns the proposal on for the
   following calls on ctx; a candidate is proposed when its logit exceeds `threshold`. */
int g4d_set_proposal(g4d_ctx *ctx, const float *params, int32_t nparams, float threshold, int32_t enable);
/* enable != 0 keeps each call's features, candidates and labels (which candidates are in the
   final link, and the link-size bin) for g4d_get_samples. */
int g4d_set_sample_export(g4d_ctx *ctx, int32_t enable);
/* Copies samples of the last call: points offset, offset + stride, ... (at most cap):
   features (12 floats), label words (bit j: candidate j in the final link; bit 8 + j:
   candidate j exists; bits 16..18: link-size bin) and candidates (8 ids, -1 when absent; may
   be NULL).  *count receives the number copied. */
int g4d_get_samples(g4d_ctx *ctx, int64_t offset, int64_t stride, float *features, uint32_t *labels,
                    int32_t *candidates, int64_t cap, int64_t *count);
/* Network outputs (16 logits per sample) of n samples (host arrays). */
int g4d_mlp_forward(g4d_ctx *ctx, const float *params, const float *features, int64_t n, float *logits);
/* Gradient of the summed loss over n samples (binary cross-entropy with logits over the
   existing candidates plus the link-size-bin cross-entropy) with respect to the parameters,
   and the summed loss. */
int g4d_mlp_grad(g4d_ctx *ctx, const float *params, const float *features, const uint32_t *labels, int64_t n,
                 float *grad, double *loss);
/* Trains params in place on the GPU: Adam (betas 0.9, 0.999), minibatches of `batch` samples
   in a seeded shuffled order, `epochs` passes; epoch_loss (may be NULL) receives the mean loss
   per sample of each epoch. */
int g4d_mlp_train(g4d_ctx *ctx, float *params, const float *features, const uint32_t *labels, int64_t n,
                  int32_t epochs, int32_t batch, float lr, uint64_t seed, double *epoch_loss);

/* Validate the last result of ctx on the GPU (TetraMesh::check, GD/Common/Geometry.cu:200-393);
   full_empty_sphere != 0 adds the empty-sphere test of every point near each tetrahedron. */
int g4d_check_result(g4d_ctx *ctx, int32_t full_empty_sphere, g4d_check_report *rep);
/* Validate a caller's mesh: points (n x 3), tets and opp (T x 4; opp may be NULL). */
int g4d_check_mesh(g4d_ctx *ctx, const double *xyz, int32_t n, const int32_t *tets, const int32_t *opp,
                   int64_t num_tets, int32_t full_empty_sphere, g4d_check_report *rep);

/* Star consistency of the last result's stars (areStarsConsistent -> kerCheckStarConsistency,
   GD/GDelaunay/GDelHost.cu:1362-1387, GD/GDelaunay/GDelKernels.cu:1193-1271): link triangle
   (a, b, c) of star s stands for tetrahedron {s, a, b, c}, so stars a, b and c must each hold the
   link triangle made of s and the other two (vertex sets compared without orientation).
   *inconsistent receives the number of link triangles that fail (0 after splaying ends),
   *triangles (may be NULL) the number checked. */
int g4d_check_stars(g4d_ctx *ctx, int64_t *inconsistent, int64_t *triangles);

/* ---- Point and mesh files (host only; no context, no device work) ---- */

/* Points of an .asc file (GD/Main/Application.cpp:475-566): whitespace-separated tokens read as
   floats, three per point; exact duplicates dropped (first occurrence kept); `scaled` is
   Application::_scalePoint (GD/Main/Application.cpp:568-585) into [1, G - 2] with the float
   min/max over all coordinates.  Arrays are malloc'ed; release with g4d_free_point_file.
   Returns G4D_ERR_IO when the file cannot be opened, G4D_ERR_DUPLICATE_POINTS when scaling
   merges points (the reference asserts it does not). */
typedef struct {
    int32_t num_points;     /* unique points */
    int32_t pad;
    int64_t num_read;       /* triples in the file */
    int64_t num_duplicates; /* num_read - num_points */
    float min_val, max_val; /* over every coordinate; 999 / -999 for an empty file */
    double *points;         /* num_points x 3, as read */
    double *scaled;         /* num_points x 3 */
} g4d_point_file;
int g4d_read_asc(const char *path, int32_t grid_size, g4d_point_file *out);
void g4d_free_point_file(g4d_point_file *f);
/* "x y z" per point, as the reference's Input-Points.asc dump (GD/Main/Application.cpp:231-243). */
int g4d_write_asc(const char *path, const double *xyz, int32_t n);
/* ASCII PLY of the points and three faces per tetrahedron (TetraMesh::writeToFile,
   GD/Common/Geometry.cu:395-466). */
int g4d_write_ply(const char *path, const double *xyz, int32_t n, const int32_t *tets, int64_t num_tets);

/* sizeof of g4d_options, g4d_result, g4d_stats, g4d_check_report, g4d_point_file and
   g4d_batch_result, in that order (for bindings that mirror the structs); returns the count. */
int g4d_struct_sizes(int64_t *out, int32_t max_out);

/* ---- Kernel-level diagnostics (used by parity tests; no reference counterpart) ---- */
/* orient4 of T tuples (15 doubles, 5 indices each).  mode 0: filter + exact + SoS;
   1: exact + SoS only; 2: the star kernel's cached-cofactor filter with exact fallback,
   tuples ordered (a, b, c, s, v).  depth_out (may be NULL): 0 filter/exact, else SoS depth. */
int g4d_diag_orient4(g4d_ctx *ctx, const double *pts5, const int32_t *idx5, int64_t T, int32_t mode,
                     int32_t *sign_out, int32_t *depth_out);
int g4d_diag_orient3d(g4d_ctx *ctx, const double *pts4, int64_t T, int32_t *sign_out);
/* Morton order and voxel coordinates (3 per input point). */
int g4d_diag_morton(g4d_ctx *ctx, const double *xyz, int32_t n, int32_t grid, int32_t *perm_out, int32_t *vox_out);
/* PBA of sites given as Morton-sorted voxel coordinates (3n); out: G^3 nearest-site point ids,
   index (z*G + y)*G + x.  codes: 0 ids, 1 site codes x<<20|y<<10|z (grids up to 1024), 2 ids
   through the point-id grids that wider grids use (any grid). */
int g4d_diag_pba(g4d_ctx *ctx, const int32_t *vox_sorted, int32_t n, int32_t grid, int32_t codes, int32_t *out);
/* Plain device-to-device copy of `bytes` (rounded down to 16) between two buffers the context
   allocates; *ms receives the kernel time.  Calibrates HBM traffic counters and bandwidth. */
int g4d_diag_copy(g4d_ctx *ctx, int64_t bytes, double *ms);
/* Overwrites link vertex `local` of star `star` (Morton index space) of the last result with
   point id v: plants a defect for the star-consistency tests. */
int g4d_diag_set_link_vertex(g4d_ctx *ctx, int32_t star, int32_t local, int32_t v);
/* Sorts m directed pair keys a<<32|b (a, b < 2^nb) by a, then b: method 0 as the pipeline does
   (the bucket sort for long lists: scatter into buckets of consecutive a, one LDS sort per
   bucket; the LSD radix sort otherwise), 1 the LSD radix sort, 3 the bucket sort at any length;
   methods 2 and 4 are 0 and 3 for drowned-insertion keys a<<33|b (b < 2^(nb+1)).  *ms (may be
   NULL): device time of the sort; *shift (may be NULL): the bucket width log2 chosen, -1 where
   the LSD sort runs. */
int g4d_diag_sort_pairs(g4d_ctx *ctx, const uint64_t *keys, int64_t m, int32_t nb, int32_t method, uint64_t *out,
                        double *ms, int32_t *shift);
/* Sorted unique Voronoi-neighbour pairs (min<<32|max) from an id grid. */
int g4d_diag_grid_pairs(g4d_ctx *ctx, const int32_t *id_grid, int32_t grid, uint64_t *out, int64_t cap,
                        int64_t *count);

#ifdef __cplusplus
}
#endif
#endif
