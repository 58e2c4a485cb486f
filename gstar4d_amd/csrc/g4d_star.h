// gStar4D for MI355X — star stage launchers.
#pragma once

#include "g4d_prims.h"

namespace g4d {

// LDS star kernel tiers by link vertices (plus this pass's insertions).
static constexpr int STAR_VMAX_T0 = 32;
static constexpr int STAR_VMAX_T1 = 64;
static constexpr int STAR_VMAX_T2 = 256;
// Stars taken per work-queue atomic by the filter-only LDS kernels.
static constexpr int STAR_CHUNK = 8;
// Star triangles in HBM.  A star with fewer than STAR_WIDE_MIN link vertices stores each
// triangle in one word: local vertex indices (9 bits each), adjacency slots (10 bits each),
// checked flag (bit 57).  A larger star stores two words per triangle: local vertices
// (21 bits each, checked flag in bit 63), then adjacency slots (21 bits each).
static constexpr int STAR_WIDE_MIN = 512;
// Stars whose link (plus this pass's insertions) exceeds the LDS kernels' limit run in the
// global-memory star kernel, which has no fixed capacity.
static constexpr int STAR_LDS_LIMIT = STAR_WIDE_MIN - 1;

// Triangle words reserved for a star of v link vertices (including this pass's insertions):
// 2v - 4 triangles, two words each in the wide format.
__host__ __device__ inline int64_t star_tcap_words(int64_t v)
{
    const int64_t t = v >= 4 ? 2 * v - 4 : 4;
    return v >= STAR_WIDE_MIN ? 2 * t : t;
}
static constexpr uint64_t STAR_PAIR_NONE = ~0ull;

// Star storage: per-star records (vertex count/offset, triangle count/offset, history
// count/offset) over three append-only arenas (link vertices, triangle words, history
// vertices).  A pass rebuilds only the stars that receive insertions: each is written to a
// fresh region reserved for it (nvoff/ntoff/nhoff) and its record is switched over when it
// is complete; untouched stars stay where they are.
struct StarParams {
    const double *pts;
    int n;
    int mode; // 0: create from worksets, 1: splaying insertions
    // insertion lists: directed keys (star<<32 | vertex), sorted, CSR offsets (n+1)
    const uint64_t *ins_keys;
    const int64_t *ins_off;
    // per-star records (read for the current star, rewritten when it is rebuilt)
    int *vnum;
    int64_t *voff;
    int *tnum;
    int64_t *toff;
    int *hnum;
    int64_t *hoff;
    // arenas
    int *vpool;
    uint64_t *tpool;
    int *hpool; // per star: sorted vertex ids inserted so far (workset and splaying)
    // regions reserved for this pass's stars, indexed by star id
    const int64_t *nvoff;
    const int64_t *ntoff;
    const int64_t *nhoff;
    // outputs without atomics (compacted by the host code after the pass)
    uint8_t *dflag;       // per insertion, indexed like ins_keys: 1 when the vertex drowned
    uint64_t *pairbuf;    // per star, from 2 * (ntoff[s] - tbase): its insertion pairs (a<<32 | b,
    int64_t tbase;        // a < b) of unchecked triangles, dense, in triangle order
    int *pcount;          // per star: the number of those pairs
    int *defer_list;      // stars that exceed this kernel's capacity, counter `defer_counter`
    int64_t defer_cap;
    int defer_counter;
    int *exact_list;      // stars whose filters were inconclusive (filter-only kernels),
    int64_t exact_cap;    // counter `exact_counter`
    int exact_counter;
    int work_counter; // counter word of this launch's work queue (zero at launch)
    int lds_limit;        // largest vn + m handled in LDS (<= STAR_LDS_LIMIT)
    // work list (nullptr = all stars); with nlist_dev the list length is read on the device
    // (the counter a previous kernel appended with), nlist being its capacity
    const int *star_list;
    int nlist;
    const unsigned long long *nlist_dev;
    unsigned long long *cnt;
    int t0_variant; // 32-vertex filter-only kernel: STAR_T0_* (STAR_T0_DEFAULT: the tuned default)
};

// Variants of the 32-vertex filter-only star kernel.
enum {
    STAR_T0_DEFAULT = -1,
    STAR_T0_WAVE = 0,     // one star per wave (k_star32)
    STAR_T0_G16 = 1,      // four stars per wave, FP64 permanents, 3 waves per SIMD (k_star32g16)
    STAR_T0_G16F = 2,     // four stars per wave, FP32 largest permanent, 3 waves per SIMD (k_star32g16f)
    STAR_T0_G16F4 = 3,    // the same at 4 waves per SIMD (k_star32g16f4)
    STAR_T0_G32F = 4,     // two stars per wave, FP32 largest permanent, 4 waves per SIMD (k_star32g32f)
    STAR_T0_NUM = 5
};

// LDS star kernel for stars of up to vmax (STAR_VMAX_T0, T1 or T2) link vertices plus
// insertions; larger ones go to p.defer_list.  exact = false: filters only (no multiprecision
// code, no private scratch), a star with an inconclusive orientation goes to p.exact_list
// untouched.  exact = true: exact predicates and SoS, for the stars of star_list.
// Zeroes the counter words whose bits are set in mask (one launch instead of one fill per word).
hipError_t launch_zero_counters(unsigned long long *cnt, uint32_t mask, hipStream_t s);
hipError_t launch_star_kernel(const StarParams &p, int vmax, bool exact, int nblocks, hipStream_t s);
// Global-memory star kernel for the stars in `list` (any size up to vcap link vertices);
// each block owns a workspace of star_global_bytes(vcap) bytes at ws + blockIdx * stride.
int64_t star_global_bytes(int vcap);
hipError_t launch_star_global(const StarParams &p, const int *list, int nitems, char *ws, int64_t ws_stride, int vcap,
                              int nblocks, hipStream_t s);
// Largest vn + m over the stars in `list`, atomically maxed into *out.
hipError_t launch_need_max(const StarParams &p, const int *list, int nitems, unsigned long long *out, hipStream_t s);

struct ProofParams {
    const double *pts;
    const uint64_t *tasks; // (star<<32 | vertex) drowned, non-mutual
    int64_t ntasks;
    int64_t *xlist;       // filter-only pass: tasks needing exact predicates, counter G4D_CNT_APPEND5;
    int64_t nx;           // exact pass: the task indices to redo (< 0: their count from the counter)
    int64_t xcap;
    const int *vnum;
    const int64_t *voff;
    const int *vpool;
    const int *tnum;
    const int64_t *toff;
    const uint64_t *tpool;
    uint64_t *out;        // 4 pairs per task (a<<32|b, a<b)
    unsigned long long *cnt;
};
// exact = false: filter-only pass over all tasks; exact = true: the tasks in xlist[0, nx).
hipError_t launch_proofs(const ProofParams &p, bool exact, hipStream_t s);

// Drowned pairs -> keys (min<<33 | max<<1 | dir) for the mutual test.
hipError_t launch_drown_keys(const uint64_t *drown, int64_t m, uint64_t *keys, hipStream_t s);
// Sorted drowned keys -> tasks[i] = proof task (star<<32 | vertex), or STAR_TASK_NONE for a
// mutual pair; compact the NONE entries away afterwards.
static constexpr uint64_t STAR_TASK_NONE = ~0ull;
hipError_t launch_nonmutual(const uint64_t *keys, int64_t m, uint64_t *tasks, hipStream_t s);
// keep[i] = pair (a<<32|b) with a < b and b not in the history of star a.
hipError_t launch_hist_filter(const uint64_t *pairs, int64_t m, const int *hpool, const int64_t *hoff, const int *hnum,
                              uint8_t *keep, hipStream_t s);
// Reciprocate m ordered pairs in place into 2m directed keys.
hipError_t launch_reciprocate(uint64_t *keys, int64_t m, hipStream_t s);
// Stars with at least one insertion (ins_off[s+1] > ins_off[s]): flags[s] in {0,1}; after
// an exclusive scan of flags into pos, list[pos[s]] = s for the flagged stars (ascending).
hipError_t launch_active_flags(const int64_t *ins_off, int n, uint32_t *flags, hipStream_t s);
// Sparse alternative for few sorted keys: offsets of the stars with insertions only, and hk[i] =
// star at the start of each run (STAR_TASK_NONE elsewhere; compact it for the active list).
hipError_t launch_sparse_runs(const uint64_t *keys, int64_t m, int64_t *off, uint64_t *hk, hipStream_t s);
hipError_t launch_u64_to_i32(const uint64_t *in, int64_t m, const int64_t *dev_m, int *out, hipStream_t s);
hipError_t launch_active_scatter(const uint32_t *flags, const uint32_t *pos, int n, int *list, hipStream_t s);
// Region sizes of the stars of a pass (list = nullptr: stars 0..nitems-1), per item: link
// vertices v = vn + m, triangle words 2v - 4 (twice that for stars that may need the wide
// format), history hn + m.
// pcount[s] = 0 for the pass's stars (their kernel writes the dense pair count).  dev_n (may be
// null): the list's length on the device when nitems is only its bound; entries beyond get zero.
hipError_t launch_capacities(const int *list, int nitems, const int64_t *dev_n, const int *vnum, const int *hnum,
                             const int64_t *ins_off, int64_t *vcap, int64_t *tcap, int64_t *hcap, int *pcount,
                             hipStream_t s);
// The pass's insertion pairs, star after star in list order (list = nullptr: stars 0..nitems-1):
// cnt64[i] = pcount[list[i]] for the offsets scan, then the copy to out[off[i] ...].
hipError_t launch_pair_counts(const int *list, int nitems, const int *pcount, int64_t *cnt64, hipStream_t s);
hipError_t launch_pair_gather(const int *list, int nitems, const int *pcount, const int64_t *off,
                              const int64_t *ntoff, int64_t tbase, const uint64_t *pairbuf, uint64_t *out,
                              hipStream_t s);
// nvoff[s] = vbase + vscan[i] (same for t, h) for s = list[i].
hipError_t launch_assign_regions(const int *list, int nitems, const int64_t *vscan, const int64_t *tscan,
                                 const int64_t *hscan, int64_t vbase, int64_t tbase, int64_t hbase, int64_t *nvoff,
                                 int64_t *ntoff, int64_t *nhoff, hipStream_t s);
// Live region sizes of every star (vertices, triangle words, history), for repacking.
hipError_t launch_live_sizes(int n, const int *vnum, const int *tnum, const int *hnum, int64_t *vsz, int64_t *tsz,
                             int64_t *hsz, hipStream_t s);
// Copies every star's live regions to the offsets vdst/tdst/hdst in the new arenas (the
// caller then adopts those offsets as the records' own).
hipError_t launch_repack(int n, const int *vnum, const int64_t *voff, const int *tnum, const int64_t *toff,
                         const int *hnum, const int64_t *hoff, const int *vpool, const uint64_t *tpool,
                         const int *hpool, const int64_t *vdst, const int64_t *tdst, const int64_t *hdst, int *vpool2,
                         uint64_t *tpool2, int *hpool2, hipStream_t s);

// Output: per-triangle keep flag for owner triangles on the lower hull (exact orient3d < 0)
struct OutParams {
    const double *pts;
    int n;
    const int *vnum;
    const int64_t *voff;
    const int *vpool;
    const int *tnum;
    const int64_t *toff;
    uint64_t *tpool;       // keep bits of inconclusive owner triangles are set in place
    int64_t *star_tet_count; // per star
    const int64_t *star_tet_off;
    int *tets;               // 4 per tetra
    unsigned long long *cnt;
};
// Decision pass over the owner triangles: filter-decided keeps get the keep bit in their word
// and are counted in star_tet_count; inconclusive ones -> list (star<<32 | triangle), counter
// G4D_CNT_APPEND5.  The exact pass keeps and counts those with exact orient3d < 0.  After the
// caller scans star_tet_count into star_tet_off, the write pass emits the kept triangles.
hipError_t launch_out_flag(const OutParams &p, uint64_t *list, int64_t cap, hipStream_t s);
hipError_t launch_out_exact(const OutParams &p, const uint64_t *list, int64_t m, hipStream_t s);
hipError_t launch_out_write(const OutParams &p, hipStream_t s);
// Star consistency of the stars in p (pts, star_tet_* and tets unused): every link triangle
// (a, b, c) of star s must appear as {s, b, c} in star a, {s, c, a} in star b and {s, a, b} in
// star c.  Adds the failing link triangles to cnt[G4D_CNT_APPEND5], the checked ones to
// cnt[G4D_CNT_APPEND4].
hipError_t launch_star_consistency(const OutParams &p, hipStream_t s);
// Face adjacency via an open-addressing hash of face triples.
hipError_t launch_adjacency(const int *tets, int64_t nt, int *opp, uint32_t *table, int64_t table_size,
                            unsigned long long *cnt, hipStream_t s);
// Same result for tetrahedra grouped spatially (the pipeline's output order): faces are first
// paired within tiles of 512 consecutive tetrahedra in LDS, the rest through `table`.
hipError_t launch_adjacency_tiled(const int *tets, int64_t nt, int *opp, uint32_t *table, int64_t table_size,
                                  unsigned long long *cnt, hipStream_t s);
hipError_t launch_map_perm(int *tets, int64_t m4, const int *perm, hipStream_t s);
// out[i] = a[idx[i]] for i < m.
hipError_t launch_gather_i64(const int64_t *a, const int64_t *idx, int64_t m, int64_t *out, hipStream_t s);
// Batched output: vertex ids (sorted positions) -> input index within the set; opp (global
// tetrahedron ids) -> index within the set's range of tetrahedra.
hipError_t launch_batch_localize(int *tets, int *opp, int64_t nt, const int *pt_set, const int64_t *pt_off,
                                 const int64_t *tet_off, const int *perm, hipStream_t s);

} // namespace g4d
