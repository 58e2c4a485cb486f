// gStar4D for MI355X — host orchestration behind the C ABI (include/gstar4d.h).
//
// Mirrors the reference's gdelInit / gdelCompute / getTetraFromGpu / gdelDeInit
// (GD/GDelaunay/GDelaunay.cpp:47-104) with the stages of GD/GDelaunay/GDelHost.cu:
// Morton sort -> PBA -> pairs and worksets -> star creation -> splaying rounds -> tetra output.
// All stage data stays in HBM; the host reads back only counters that size the next stage
// (one small read per stage, where the reference synchronises after every launch).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <cstring>
#include <string>
#include <vector>

#include "gstar4d.h"
#include "g4d_check.h"
#include "g4d_diag.h"
#include "g4d_grid.h"
#include "g4d_noalloc.h"
#include "g4d_prims.h"
#include "g4d_proposal.h"
#include "g4d_star.h"

using g4d::DevBuf;

namespace g4d {
ScratchArena *&scratch_arena_slot()
{
    static thread_local ScratchArena *arena = nullptr;
    return arena;
}
} // namespace g4d

namespace {

// Star storage: per-star records over three append-only arenas (link vertices as int,
// triangle words as uint64, history vertices as int).  top[k] counts the elements in use.
struct Store {
    static constexpr size_t esz[3] = {sizeof(int), sizeof(uint64_t), sizeof(int)};
    DevBuf vnum, voff, tnum, toff, hnum, hoff, nvoff, ntoff, nhoff;
    DevBuf arena[3], spare[3], live[3];
    int64_t top[3] = {0, 0, 0};
    int32_t repacks = 0;
    int64_t capacity(int k) const { return (int64_t)(arena[k].bytes / esz[k]); }
    void release()
    {
        DevBuf *b[] = {&vnum, &voff, &tnum, &toff, &hnum, &hoff, &nvoff, &ntoff, &nhoff};
        for (DevBuf *x : b) x->release();
        for (int k = 0; k < 3; k++) {
            arena[k].release();
            spare[k].release();
            live[k].release();
        }
        top[0] = top[1] = top[2] = 0;
    }
};
constexpr size_t Store::esz[3];

enum { EV_START, EV_MORTON, EV_PBA0, EV_PBA1, EV_PBA2, EV_PBA3, EV_PROP0, EV_PROP1, EV_PAIRS, EV_STAR, EV_SPLAY, EV_OUT,
       EV_ADJ, EV_NUM };

// Page-locked host buffer for results handed back through host pointers: device-to-host copies
// into it run at the link's rate instead of through the runtime's pageable staging.
struct PinnedBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need)
    {
        if (need <= bytes) return hipSuccess;
        release();
        const size_t cap = need + need / 8 + 256;
        hipError_t e = hipHostMalloc(&ptr, cap, hipHostMallocDefault);
        if (e != hipSuccess) {
            ptr = nullptr;
            return e;
        }
        bytes = cap;
        return hipSuccess;
    }
    void release()
    {
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    int32_t *i32() const { return static_cast<int32_t *>(ptr); }
};

} // namespace

struct g4d_ctx {
    g4d_options opt;
    hipStream_t stream = nullptr;
    std::string err;
    g4d::Prims P;
    DevBuf xyz_up, keys, pts, perm, vox, bounds, settab, ptset, grid, gtmp, idgrid, pbaspill, pbaenv, pbacnt, pcount, pairs,
        ws_off, cnt;
    Store store;
    DevBuf settet, drown, dflag, pairbuf, spcount, hkeys, dkeys, tasks, cand, keep, ins, ins_off, vcap, tcap, hcap, aflags, apos, active, mid, big, huge, xstars,
        xtasks, xtris, gwork, tetcnt, tetoff, tets, opp, table;
    // learned proposal: parameters, candidates, features, proposed pairs, capacity bins, labels
    DevBuf mlp_params, fkeys, ford, frank, pcand, pd2, pfeat, ppairs, pcap, plab, pvcap;
    bool prop_on = false, prop_export = false, have_samples = false;
    float prop_threshold = 0.f;
    int samples_n = 0;
    PinnedBuf h_tets, h_opp, h_perm; // results of g4d_triangulate / g4d_triangulate_batch
    // Page-locked staging of the small reads back to the host (counters and counts): a pageable
    // destination makes the runtime stage every such read itself, about 8 us more per read.
    // Slots 0..23 the counters, 24..39 scalars, 40..41 the primitives' counts.
    PinnedBuf hpin;
    int64_t *hstage = nullptr; // in hpin, or the calling context's for a temporary context
    std::vector<int32_t> h_set_grid;
    std::vector<int64_t> h_set_tet_off;
    const double *last_xyz = nullptr; // device points of the last triangulation (input order)
    int last_nsets = 1;
    int64_t num_tets = 0;
    int last_n = 0;
    bool have_result = false;
    g4d_stats st;
    hipEvent_t ev[EV_NUM];
    hipEvent_t ev_t0[2]; // around each launch of the 32-vertex star kernel
    bool ev_ok = false;
    hipError_t noalloc_status = hipErrorNotInitialized; // lookup of the scratch array's address
    // tuning (g4d_set_star_kernel, g4d_set_grid_table)
    int star_t0_variant = -1;
    int pba_bands = 0; // bands of the y and x PBA passes (0: by grid size)
    int stage_graphs = -1; // grid-stage chain as a hipGraph: 1 always, 0 never, -1 batched calls only
    // captured grid-stage chains (site scatter and the PBA passes), keyed by everything the
    // captured launches depend on: sizes, launch shapes and buffer addresses
    struct StageGraph {
        std::vector<int64_t> key;
        hipGraphExec_t exec = nullptr;
        uint64_t last_use = 0;
    };
    std::vector<StageGraph> graphs;
    uint64_t graph_clock = 0;
    std::vector<int64_t> grid_tab_n;   // ascending upper bounds of the set-size buckets
    std::vector<float> grid_tab_vpp;   // voxels per point of each bucket
};

// Stage trace on stderr when the environment sets G4D_TRACE (flushed per line).
static bool trace_on()
{
    static const bool on = std::getenv("G4D_TRACE") != nullptr;
    return on;
}

static void trace(const char *fmt, ...)
{
    if (!trace_on()) return;
    // wall time since the previous trace line (the traced stages end in a host synchronisation)
    static double last = 0.0;
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    const double now = ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
    va_list ap;
    va_start(ap, fmt);
    fprintf(stderr, "[g4d +%.1f ms] ", last > 0.0 ? now - last : 0.0);
    last = now;
    vfprintf(stderr, fmt, ap);
    fprintf(stderr, "\n");
    fflush(stderr);
    va_end(ap);
}

static int set_err(g4d_ctx *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    c->err = buf;
    return code;
}

static int hip_fail(g4d_ctx *c, hipError_t e, const char *what, int line)
{
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? G4D_ERR_OOM : G4D_ERR_DEVICE;
    return set_err(c, code, "%s failed at g4d_api.cpp:%d: %s", what, line, hipGetErrorString(e));
}

#define HCK(x)                                                                                                         \
    do {                                                                                                               \
        hipError_t e_ = (x);                                                                                           \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #x, __LINE__);                                                  \
    } while (0)

static int bits_for(int64_t n)
{
    int b = 1;
    while (((int64_t)1 << b) < n) b++;
    return b;
}

static uint64_t next_pow2(uint64_t v)
{
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

template <typename T>
static T *buf(DevBuf &b)
{
    return b.as<T>();
}

// Sort directed keys (a<<32 | b) with a, b < n < 2^nb.
static hipError_t sort_pairs(g4d::Prims &P, uint64_t *k, int64_t m, int nb, int64_t n, hipStream_t s)
{
    return g4d::sort_pair_keys(P, k, m, nb, n, s);
}

// The same, then without duplicates; *m becomes the distinct count.
static hipError_t sort_unique_pairs(g4d::Prims &P, uint64_t *k, int64_t *m, int nb, int64_t n, hipStream_t s)
{
    return g4d::sort_unique_pair_keys(P, k, *m, nb, n, m, s);
}

// Size limits of the device primitives: the radix sort and the compactions count with 32-bit
// histograms and offsets, and the face matching encodes (tetrahedron << 2 | face) + 1 in 32
// bits.  Lists beyond them are refused instead of wrapping.
static int guard_keys(g4d_ctx *ctx, int64_t keys, const char *what)
{
    if (g4d_check_sizes(keys, 0) == G4D_OK) return G4D_OK;
    return set_err(ctx, G4D_ERR_CAPACITY, "%s: %lld keys exceed the sort limit of %lld", what, (long long)keys,
                   (long long)G4D_MAX_SORT_KEYS);
}

static int guard_tets(g4d_ctx *ctx, int64_t nt)
{
    if (g4d_check_sizes(0, nt) == G4D_OK) return G4D_OK;
    return set_err(ctx, G4D_ERR_CAPACITY, "%lld tetrahedra exceed the limit of %lld", (long long)nt,
                   (long long)G4D_MAX_TETS);
}

static int read_counters(g4d_ctx *ctx, unsigned long long *h)
{
    unsigned long long *dst = ctx->hstage ? reinterpret_cast<unsigned long long *>(ctx->hstage) : h;
    HCK(hipMemcpyAsync(dst, ctx->cnt.ptr, sizeof(unsigned long long) * G4D_CNT_NUM, hipMemcpyDeviceToHost, ctx->stream));
    HCK(hipStreamSynchronize(ctx->stream));
    if (dst != h) memcpy(h, dst, sizeof(unsigned long long) * G4D_CNT_NUM);
    return G4D_OK;
}

// Scalar slot k (0..15) of the page-locked staging, or the caller's fallback variable.
static int64_t *hslot(g4d_ctx *ctx, int k, int64_t *fallback)
{
    return ctx->hstage ? ctx->hstage + G4D_CNT_NUM + k : fallback;
}

static int check_dev_error(g4d_ctx *ctx, unsigned long long flags, const char *stage)
{
    if (!flags) return G4D_OK;
    if (flags & G4D_DEV_ERR_DUPLICATE) return set_err(ctx, G4D_ERR_DUPLICATE_POINTS, "%s: duplicate input points", stage);
    if (flags & G4D_DEV_ERR_PREDICATE_RANGE)
        return set_err(ctx, G4D_ERR_PREDICATE_RANGE, "%s: coordinates outside the exact-arithmetic range", stage);
    if (flags & G4D_DEV_ERR_STAR_CAPACITY)
        return set_err(ctx, G4D_ERR_CAPACITY, "%s: a star exceeds its workspace", stage);
    return set_err(ctx, G4D_ERR_INTERNAL, "%s: device error flags 0x%llx", stage, flags);
}

#define CK(x)                                                                                                          \
    do {                                                                                                               \
        int r_ = (x);                                                                                                  \
        if (r_ != G4D_OK) return r_;                                                                                   \
    } while (0)

// Moves every star's live regions into fresh arenas with room for `need` more elements each
// (half as much again as live + need), dropping the regions earlier passes left behind.
static int repack_store(g4d_ctx *ctx, int n, const int64_t need[3])
{
    Store &st = ctx->store;
    hipStream_t s = ctx->stream;
    int64_t *sz[3];
    for (int k = 0; k < 3; k++) {
        HCK(st.live[k].ensure(sizeof(int64_t) * (size_t)(n + 1)));
        sz[k] = buf<int64_t>(st.live[k]);
    }
    HCK(g4d::launch_live_sizes(n, buf<int>(st.vnum), buf<int>(st.tnum), buf<int>(st.hnum), sz[0], sz[1], sz[2], s));
    int64_t live[3];
    for (int k = 0; k < 3; k++) {
        HCK(g4d::scan_excl_i64(ctx->P, sz[k], sz[k], n, sz[k] + n, s));
        HCK(hipMemcpyAsync(&live[k], sz[k] + n, 8, hipMemcpyDeviceToHost, s));
    }
    HCK(hipStreamSynchronize(s));
    for (int k = 0; k < 3; k++) {
        const int64_t cap = std::max<int64_t>((live[k] + need[k]) * 3 / 2, 1024);
        HCK(st.spare[k].ensure(Store::esz[k] * (size_t)cap));
    }
    HCK(g4d::launch_repack(n, buf<int>(st.vnum), buf<int64_t>(st.voff), buf<int>(st.tnum), buf<int64_t>(st.toff),
                           buf<int>(st.hnum), buf<int64_t>(st.hoff), buf<int>(st.arena[0]), buf<uint64_t>(st.arena[1]),
                           buf<int>(st.arena[2]), sz[0], sz[1], sz[2], buf<int>(st.spare[0]),
                           buf<uint64_t>(st.spare[1]), buf<int>(st.spare[2]), s));
    HCK(hipStreamSynchronize(s));
    std::swap(st.voff, st.live[0]);
    std::swap(st.toff, st.live[1]);
    std::swap(st.hoff, st.live[2]);
    for (int k = 0; k < 3; k++) {
        std::swap(st.arena[k], st.spare[k]);
        st.spare[k].release();
        st.top[k] = live[k];
    }
    st.repacks++;
    trace("star store repacked: live %lld / %lld / %lld, capacity %lld / %lld / %lld", (long long)live[0],
          (long long)live[1], (long long)live[2], (long long)st.capacity(0), (long long)st.capacity(1),
          (long long)st.capacity(2));
    return G4D_OK;
}

static void ev_record(g4d_ctx *ctx, int e)
{
    if (ctx->ev_ok) (void)hipEventRecord(ctx->ev[e], ctx->stream);
}

static double ev_ms(g4d_ctx *ctx, int a, int b)
{
    float ms = 0.f;
    if (!ctx->ev_ok) return 0.0;
    if (hipEventElapsedTime(&ms, ctx->ev[a], ctx->ev[b]) != hipSuccess) return 0.0;
    return ms;
}

// The whole path on device-resident FP64 points.
// Point sets of one call (host side): set k holds points [pt_off[k], pt_off[k+1]) with grid
// width g[k].  A single triangulation is one set.
struct SetSpec {
    int nsets = 1;
    std::vector<int64_t> pt_off;
    std::vector<int> g;
};

static SetSpec single_set(int n, int g)
{
    SetSpec sp;
    sp.nsets = 1;
    sp.pt_off = {0, (int64_t)n};
    sp.g = {g};
    return sp;
}

struct SetTotals {
    int64_t voxels = 0, pba_blocks = 0, spill = 0, pair_threads = 0;
};

// Uploads the set table (point ranges, grids, prefix sums of the grid stages' work) and, for
// several sets, the per-point set ids; GS.bounds points at room for 2 per set (or the
// single-set reduction's scratch).
static int setup_sets(g4d_ctx *ctx, const SetSpec &spec, int n, g4d::GridSets &GS, SetTotals &tot)
{
    hipStream_t s = ctx->stream;
    const int S = spec.nsets;
    std::vector<int64_t> tab(5 * (size_t)(S + 1));
    int64_t *h_pt = tab.data(), *h_vox = h_pt + (S + 1), *h_blk = h_vox + (S + 1), *h_spill = h_blk + (S + 1),
            *h_thr = h_spill + (S + 1);
    h_pt[0] = h_vox[0] = h_blk[0] = h_spill[0] = h_thr[0] = 0;
    for (int k = 0; k < S; k++) {
        const int64_t G = spec.g[k];
        h_pt[k + 1] = spec.pt_off[k + 1];
        h_vox[k + 1] = h_vox[k] + G * G * G;
        h_blk[k + 1] = h_blk[k] + g4d::pba_blocks((int)G);
        h_spill[k + 1] = h_spill[k] + g4d::pba_spill_ints((int)G);
        h_thr[k + 1] = h_thr[k] + g4d::grid_pair_threads((int)G, g4d::grid_pair_chunk((int)G));
    }
    tot.voxels = h_vox[S];
    tot.pba_blocks = h_blk[S];
    tot.spill = h_spill[S];
    tot.pair_threads = h_thr[S];
    const size_t tab_bytes = sizeof(int64_t) * tab.size();
    const size_t bounds_at = (tab_bytes + sizeof(int) * (size_t)S + 15) & ~(size_t)15;
    HCK(ctx->settab.ensure(bounds_at + sizeof(double) * (size_t)std::max(2 * S, 768)));
    char *dtab = buf<char>(ctx->settab);
    HCK(hipMemcpyAsync(dtab, tab.data(), tab_bytes, hipMemcpyHostToDevice, s));
    HCK(hipMemcpyAsync(dtab + tab_bytes, spec.g.data(), sizeof(int) * (size_t)S, hipMemcpyHostToDevice, s));
    GS.nsets = S;
    GS.pt_off = reinterpret_cast<const int64_t *>(dtab);
    GS.vox_off = GS.pt_off + (S + 1);
    GS.pba_blk = GS.vox_off + (S + 1);
    GS.spill = GS.pba_blk + (S + 1);
    GS.pair_thr = GS.spill + (S + 1);
    GS.g = reinterpret_cast<const int *>(dtab + tab_bytes);
    GS.bounds = reinterpret_cast<const double *>(dtab + bounds_at);
    GS.pt_set = nullptr;
    GS.key_shift = g4d::morton_key_shift(spec.g[0]);
    if (S > 1) {
        HCK(ctx->ptset.ensure(sizeof(int) * (size_t)n));
        HCK(g4d::launch_set_ids(GS.pt_off, S, buf<int>(ctx->ptset), s));
        GS.pt_set = buf<int>(ctx->ptset);
    }
    return G4D_OK;
}

// Learned proposal stage.  Candidates: the 8 nearest points within windows of two fine
// Morton orders; features of the digital Voronoi diagram; with the proposal on, the MLP's
// chosen candidates are added to the worksets (both directions, merged into the sorted keys
// pr[0, *m) without duplicates) and *pred_vertices receives the predicted total of final link
// vertices.  With export on, candidates and features stay for the training samples.
static int proposal_stage(g4d_ctx *ctx, int n, const g4d::GridSets &GS, const int *idg, uint64_t *pr, int64_t *m,
                          const int64_t *ws_off, int nb, int64_t *pred_vertices)
{
    hipStream_t s = ctx->stream;
    g4d::Prims &P = ctx->P;
    const int S = GS.nsets;
    const int sb = S > 1 ? bits_for(S) : 0;
    const int cb = std::min(21, (64 - nb - sb) / 3);
    if (cb < 6) return set_err(ctx, G4D_ERR_INVALID_INPUT, "proposal: too many points for fine Morton keys");
    const int hi_bit = sb + 3 * cb + nb;
    HCK(ctx->fkeys.ensure(sizeof(uint64_t) * (size_t)n));
    HCK(ctx->ford.ensure(2 * sizeof(int) * (size_t)n));
    HCK(ctx->frank.ensure(2 * sizeof(int) * (size_t)n));
    int *ord = buf<int>(ctx->ford), *rank = buf<int>(ctx->frank);
    uint64_t *fk = buf<uint64_t>(ctx->fkeys);
    for (int o = 0; o < 2; o++) {
        HCK(g4d::launch_fine_keys(buf<double>(ctx->pts), n, GS, cb, nb, o, fk, s));
        HCK(g4d::radix_sort_u64(P, fk, n, nb, hi_bit, s));
        HCK(g4d::launch_fine_ranks(fk, n, nb, ord + (size_t)o * n, rank + (size_t)o * n, s));
    }
    HCK(ctx->pcand.ensure(sizeof(int) * g4d::PROP_SLOTS * (size_t)n));
    HCK(ctx->pd2.ensure(sizeof(float) * g4d::PROP_SLOTS * (size_t)n));
    HCK(ctx->pfeat.ensure(sizeof(float) * g4d::PROP_FEATS * (size_t)n));
    HCK(g4d::launch_candidates(buf<double>(ctx->pts), n, GS, ord, rank, ord + n, rank + n, buf<int>(ctx->pcand),
                               buf<float>(ctx->pd2), s));
    HCK(g4d::launch_features(buf<double>(ctx->pts), buf<uint64_t>(ctx->vox), n, GS, idg, pr, ws_off,
                             buf<int>(ctx->pcand), buf<float>(ctx->pd2), buf<float>(ctx->pfeat), s));
    if (!ctx->prop_on) return G4D_OK;
    // proposals: per point up to 8 pairs (i << 32 | candidate), compacted, then both directions
    HCK(ctx->ppairs.ensure(sizeof(uint64_t) * 2 * g4d::PROP_SLOTS * (size_t)n));
    HCK(ctx->pcap.ensure((size_t)n));
    uint64_t *pp = buf<uint64_t>(ctx->ppairs);
    HCK(g4d::launch_mlp_propose(buf<float>(ctx->mlp_params), buf<float>(ctx->pfeat), buf<int>(ctx->pcand), n,
                                ctx->prop_threshold, pp, buf<uint8_t>(ctx->pcap), s));
    int64_t np = 0;
    HCK(g4d::compact_valid_u64(P, pp, pp, g4d::PROP_SLOTS * (int64_t)n, g4d::STAR_PAIR_NONE, &np, s));
    ctx->st.prop_pairs = np;
    // predicted final link sizes (at least the current workset) for the arena sizing
    HCK(ctx->pvcap.ensure(sizeof(int64_t) * (size_t)(n + 1)));
    int64_t *pv = buf<int64_t>(ctx->pvcap);
    HCK(g4d::launch_capacity_vertices(buf<uint8_t>(ctx->pcap), ws_off, n, pv, s));
    HCK(g4d::scan_excl_i64(P, pv, pv, n, pv + n, s));
    HCK(hipMemcpyAsync(pred_vertices, pv + n, 8, hipMemcpyDeviceToHost, s));
    if (np > 0) {
        CK(guard_keys(ctx, *m + 2 * np, "proposed worksets"));
        HCK(g4d::launch_reciprocate(pp, np, s));
        np *= 2;
        HCK(sort_unique_pairs(P, pp, &np, nb, n, s));
        HCK(P.sort_tmp_keys.ensure(sizeof(uint64_t) * (size_t)(*m + np)));
        uint64_t *merged = P.sort_tmp_keys.as<uint64_t>();
        HCK(g4d::merge_u64(pr, *m, pp, np, merged, s));
        int64_t mm = *m + np;
        // back into pr without the proposals already in a workset (out of place: unique_u64
        // would compact merged through the very buffer that holds it)
        HCK(g4d::unique_copy_u64(P, merged, pr, mm, &mm, s));
        *m = mm;
    }
    HCK(hipStreamSynchronize(s));
    return G4D_OK;
}

// Bands of the y and x PBA passes: the context's choice, else G4D_PBA_BANDS (measurement), else
// by grid width (bands pay once a column is long enough to give each band's wave real work).
static int pba_bands_for(const g4d_ctx *ctx, int gmax)
{
    int b = ctx->pba_bands;
    if (b <= 0) {
        static int env = -2;
        if (env == -2) {
            const char *e = getenv("G4D_PBA_BANDS");
            env = e ? atoi(e) : 0;
        }
        b = env > 0 ? env : (gmax >= 384 ? 4 : gmax >= 192 ? 2 : 1);
    }
    return std::max(1, std::min(b, g4d::PBA_MAX_BANDS));
}

// Device buffers of the PBA passes: the spill area and, for banded passes, the band
// envelopes and counts (allocated before any capture: a graph cannot allocate).
struct PbaBufs {
    int bands = 1;
    int *spill = nullptr, *benv = nullptr, *bcnt = nullptr, *fcnt = nullptr;
};

static int pba_buffers(g4d_ctx *ctx, const SetTotals &tot, int gmax, PbaBufs *pb)
{
    HCK(ctx->pbaspill.ensure(sizeof(int) * (size_t)tot.spill));
    pb->spill = buf<int>(ctx->pbaspill);
    pb->bands = pba_bands_for(ctx, gmax);
    if (pb->bands > 1) {
        HCK(ctx->pbaenv.ensure(sizeof(int) * (size_t)tot.spill));
        HCK(ctx->pbacnt.ensure(sizeof(int) * 64 * (size_t)tot.pba_blocks * (size_t)(pb->bands + 1)));
        pb->benv = buf<int>(ctx->pbaenv);
        pb->bcnt = buf<int>(ctx->pbacnt);
        pb->fcnt = pb->bcnt + 64 * tot.pba_blocks * pb->bands;
    }
    ctx->st.pba_bands = pb->bands;
    return G4D_OK;
}

// The three PBA axis passes: codes in grid -> (z) gtmp -> (y) grid -> (x) gtmp, the last with
// point ids (idgrid, or the ids the grids hold when vox is given).  The y and x passes run in
// bands when pba_bands_for says so.  Kernel launches only (it may run under stream capture).
static int pba_launch(g4d_ctx *ctx, const g4d::GridSets &GS, const SetTotals &tot, const PbaBufs &pb, int *grid,
                      int *gtmp, const int *idgrid_last, const uint64_t *pvox, bool events)
{
    hipStream_t s = ctx->stream;
    if (events) ev_record(ctx, EV_PBA0);
    HCK(g4d::pba_pass(2, GS, tot.pba_blocks, grid, gtmp, nullptr, pb.spill, pvox, s));
    if (events) ev_record(ctx, EV_PBA1);
    if (pb.bands > 1) HCK(g4d::pba_pass_banded(1, GS, tot.pba_blocks, pb.bands, gtmp, grid, nullptr, pb.spill, pb.benv,
                                                pb.bcnt, pb.fcnt, pvox, s));
    else HCK(g4d::pba_pass(1, GS, tot.pba_blocks, gtmp, grid, nullptr, pb.spill, pvox, s));
    if (events) ev_record(ctx, EV_PBA2);
    if (pb.bands > 1) HCK(g4d::pba_pass_banded(0, GS, tot.pba_blocks, pb.bands, grid, gtmp, idgrid_last, pb.spill,
                                                pb.benv, pb.bcnt, pb.fcnt, pvox, s));
    else HCK(g4d::pba_pass(0, GS, tot.pba_blocks, grid, gtmp, idgrid_last, pb.spill, pvox, s));
    if (events) ev_record(ctx, EV_PBA3);
    return G4D_OK;
}

static int run_pba(g4d_ctx *ctx, const g4d::GridSets &GS, const SetTotals &tot, int gmax, int *grid, int *gtmp,
                   const int *idgrid_last, const uint64_t *pvox, bool events)
{
    PbaBufs pb;
    CK(pba_buffers(ctx, tot, gmax, &pb));
    return pba_launch(ctx, GS, tot, pb, grid, gtmp, idgrid_last, pvox, events);
}

static void destroy_graphs(g4d_ctx *ctx)
{
    for (auto &g : ctx->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    ctx->graphs.clear();
}

// Grid stage as one hipGraph launch: the grid fill, the site scatter and the PBA passes
// (9-13 kernel launches) are captured once per key and replayed.  Every launch of the chain
// is a kernel (the fill included, no memset node), and every size and address the captured
// launches use is part of the key, so a replay runs exactly what an eager call would.
static int grid_stage_graph(g4d_ctx *ctx, const g4d::GridSets &GS, const SetTotals &tot, int gmax, int n,
                            const uint64_t *vox, int *grid, int *gtmp, int *idgrid, const uint64_t *pvox)
{
    hipStream_t s = ctx->stream;
    PbaBufs pb;
    CK(pba_buffers(ctx, tot, gmax, &pb));
    const std::vector<int64_t> key = {(int64_t)n, (int64_t)GS.nsets, tot.voxels, tot.pba_blocks, tot.spill,
                                      (int64_t)pb.bands, (int64_t)(intptr_t)vox, (int64_t)(intptr_t)grid,
                                      (int64_t)(intptr_t)gtmp, (int64_t)(intptr_t)idgrid, (int64_t)(intptr_t)pvox,
                                      (int64_t)(intptr_t)pb.spill, (int64_t)(intptr_t)pb.benv,
                                      (int64_t)(intptr_t)pb.bcnt, (int64_t)(intptr_t)GS.pt_off,
                                      (int64_t)(intptr_t)GS.pt_set, (int64_t)(intptr_t)GS.g};
    g4d_ctx::StageGraph *hit = nullptr;
    for (auto &g : ctx->graphs)
        if (g.key == key) hit = &g;
    if (!hit) {
        hipGraph_t graph = nullptr;
        HCK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        hipError_t e = g4d::launch_fill_i32(grid, tot.voxels, G4D_MARKER, s);
        if (e == hipSuccess) e = g4d::launch_scatter_sites(vox, n, GS, grid, idgrid, s);
        int r = G4D_OK;
        if (e == hipSuccess) r = pba_launch(ctx, GS, tot, pb, grid, gtmp, idgrid, pvox, false);
        const hipError_t ec = hipStreamEndCapture(s, &graph);
        if (e != hipSuccess) return hip_fail(ctx, e, "grid stage capture", __LINE__);
        if (r != G4D_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return r;
        }
        HCK(ec);
        hipGraphExec_t exec = nullptr;
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HCK(ei);
        if (ctx->graphs.size() >= 16) { // drop the least recently used
            size_t lru = 0;
            for (size_t k = 1; k < ctx->graphs.size(); k++)
                if (ctx->graphs[k].last_use < ctx->graphs[lru].last_use) lru = k;
            (void)hipGraphExecDestroy(ctx->graphs[lru].exec);
            ctx->graphs.erase(ctx->graphs.begin() + (ptrdiff_t)lru);
        }
        ctx->graphs.push_back(g4d_ctx::StageGraph{key, exec, 0});
        hit = &ctx->graphs.back();
        ctx->st.graph_captures++;
    }
    hit->last_use = ++ctx->graph_clock;
    ev_record(ctx, EV_PBA0);
    HCK(hipGraphLaunch(hit->exec, s));
    ev_record(ctx, EV_PBA1);
    ev_record(ctx, EV_PBA2);
    ev_record(ctx, EV_PBA3);
    ctx->st.graph_launches++;
    return G4D_OK;
}

static int run_device(g4d_ctx *ctx, const double *d_xyz, int n, const SetSpec &spec)
{
    ctx->last_xyz = d_xyz;
    ctx->last_nsets = spec.nsets;
    hipStream_t s = ctx->stream;
    g4d::Prims &P = ctx->P;
    const int S = spec.nsets;
    const bool batched = S > 1;
    int gmax = 0;
    for (int k = 0; k < S; k++) {
        const int G = spec.g[k];
        const int64_t nk = spec.pt_off[k + 1] - spec.pt_off[k];
        if (G < 4 || G > G4D_MAX_GRID)
            return set_err(ctx, G4D_ERR_GRID_LIMIT, "grid_size %d outside [4, %d]", G, G4D_MAX_GRID);
        if (batched && G > g4d::BATCH_MAX_GRID)
            return set_err(ctx, G4D_ERR_GRID_LIMIT, "set %d: batched grid %d above %d", k, G, g4d::BATCH_MAX_GRID);
        if (nk < 5) return set_err(ctx, G4D_ERR_INVALID_INPUT, "set %d: need at least 5 points, got %lld", k, (long long)nk);
        if (batched && nk > g4d::BATCH_MAX_SET_POINTS)
            return set_err(ctx, G4D_ERR_INVALID_INPUT, "set %d: %lld points, above the batched limit", k, (long long)nk);
        gmax = std::max(gmax, G);
    }
    if (n > (1 << 29)) return set_err(ctx, G4D_ERR_INVALID_INPUT, "too many points (%d)", n);
    if (ctx->opt.lds_star_limit != 0 && ctx->opt.lds_star_limit < 4)
        return set_err(ctx, G4D_ERR_INVALID_INPUT, "lds_star_limit %d below 4", ctx->opt.lds_star_limit);
    const int min_ws = ctx->opt.min_workset;
    if (min_ws < 4 || min_ws > 64) return set_err(ctx, G4D_ERR_INVALID_INPUT, "min_workset %d outside [4, 64]", min_ws);
    const int nb = bits_for(n);
    memset(&ctx->st, 0, sizeof(ctx->st));
    ctx->st.num_points = n;
    ctx->st.grid_size = gmax;
    ctx->have_result = false;
    ctx->have_samples = false;
    unsigned long long hc[G4D_CNT_NUM];

    HCK(ctx->cnt.ensure(sizeof(unsigned long long) * G4D_CNT_NUM));
    unsigned long long *cnt = buf<unsigned long long>(ctx->cnt);
    HCK(hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * G4D_CNT_NUM, s));
    ev_record(ctx, EV_START);

    // ---- set tables: points, grids and the work decompositions of the grid stages ----
    g4d::GridSets GS;
    SetTotals tot;
    CK(setup_sets(ctx, spec, n, GS, tot));
    double *dbounds = const_cast<double *>(GS.bounds);
    const int64_t G3 = tot.voxels;

    // ---- B: Morton order (GD/GDelaunay/GDelData.cu:91-131) ----
    if (batched) HCK(g4d::launch_set_bounds(d_xyz, GS, dbounds, s));
    else HCK(g4d::launch_bounds(d_xyz, n, dbounds, s));
    {
        std::vector<double> hb(2 * (size_t)S);
        HCK(hipMemcpyAsync(hb.data(), dbounds, sizeof(double) * 2 * (size_t)S, hipMemcpyDeviceToHost, s));
        HCK(hipStreamSynchronize(s));
        for (int k = 0; k < S; k++)
            if (!std::isfinite(hb[2 * k]) || !std::isfinite(hb[2 * k + 1]) || !(hb[2 * k + 1] > hb[2 * k]))
                return set_err(ctx, G4D_ERR_INVALID_INPUT, "set %d: coordinates must be finite and not all equal", k);
    }
    HCK(ctx->keys.ensure(sizeof(uint64_t) * (size_t)n));
    HCK(ctx->pts.ensure(sizeof(double) * 3 * (size_t)n));
    HCK(ctx->perm.ensure(sizeof(int) * (size_t)n));
    HCK(ctx->vox.ensure(sizeof(uint64_t) * (size_t)n));
    uint64_t *keys = buf<uint64_t>(ctx->keys);
    HCK(g4d::launch_morton_keys(d_xyz, n, GS, keys, s));
    if (batched) HCK(g4d::radix_sort_u64(P, keys, n, 24, 64, s));
    else HCK(g4d::radix_sort_u64(P, keys, n, GS.key_shift, 64, s));
    double *pts = buf<double>(ctx->pts);
    int *perm = buf<int>(ctx->perm);
    uint64_t *vox = buf<uint64_t>(ctx->vox);
    HCK(g4d::launch_gather_sorted(d_xyz, keys, n, GS, pts, perm, vox, s));
    HCK(g4d::launch_dup_check(pts, vox, n, GS, cnt, s));
    ev_record(ctx, EV_MORTON);

    // ---- C: PBA (GD/PBA/Pba.cu:105-122), every set's grid in one launch per pass ----
    // Grids up to G4D_CODE_GRID carry packed site codes (and a point-id grid for the last
    // pass); wider grids carry point ids throughout and read site voxels from `vox`.
    const bool id_mode = gmax > G4D_CODE_GRID;
    HCK(ctx->grid.ensure(sizeof(int) * (size_t)G3));
    HCK(ctx->gtmp.ensure(sizeof(int) * (size_t)G3));
    if (!id_mode) HCK(ctx->idgrid.ensure(sizeof(int) * (size_t)G3));
    int *grid = buf<int>(ctx->grid), *gtmp = buf<int>(ctx->gtmp), *idgrid = id_mode ? nullptr : buf<int>(ctx->idgrid);
    const uint64_t *pvox = id_mode ? vox : nullptr;
    if (ctx->stage_graphs == 1 || (ctx->stage_graphs < 0 && batched)) {
        // fill, scatter and PBA replayed from a captured graph (point ids in gtmp)
        CK(grid_stage_graph(ctx, GS, tot, gmax, n, vox, grid, gtmp, idgrid, pvox));
    } else {
        HCK(g4d::launch_fill_i32(grid, G3, G4D_MARKER, s));
        HCK(g4d::launch_scatter_sites(vox, n, GS, grid, idgrid, s));
        CK(run_pba(ctx, GS, tot, gmax, grid, gtmp, idgrid, pvox, true)); // point ids in gtmp
    }
    const int *idg = gtmp;

    // ---- D: missing points, grid pairs, worksets, pumping (GD/GDelaunay/GDelHost.cu:259-529) ----
    // Pairs are extracted once per unordered pair (min<<32 | max), deduplicated, then
    // reciprocated into directed workset keys (star<<32 | vertex) and grouped by star.
    // Raw pairs are counted per thread; each 256-thread tile then writes its pairs, deduplicated
    // on chip, into its raw region at pr[ucap...], and the tiles are packed into pr[0...).
    const int64_t nthr = tot.pair_threads, ntiles = nthr / 256;
    HCK(ctx->pcount.ensure(sizeof(int64_t) * (size_t)(nthr + 1 + 2 * ntiles + 1)));
    int64_t *pcnt = buf<int64_t>(ctx->pcount);
    int64_t *ttot = pcnt + nthr + 1;
    uint2 *tcnt = reinterpret_cast<uint2 *>(ttot + ntiles + 1);
    HCK(g4d::launch_grid_pairs(idg, GS, nthr, pcnt, nullptr, false, s));
    HCK(g4d::scan_excl_i64(P, pcnt, pcnt, nthr, pcnt + nthr, s));
    int64_t ngp_f = 0, *ngp_h = hslot(ctx, 0, &ngp_f);
    HCK(hipMemcpyAsync(ngp_h, pcnt + nthr, 8, hipMemcpyDeviceToHost, s));
    HCK(hipStreamSynchronize(s));
    const int64_t ngp = *ngp_h;
    CK(guard_keys(ctx, 2 * (ngp + (int64_t)n), "grid pairs"));
    const int64_t ucap = ngp + n;
    const int64_t prop_room = ctx->prop_on ? 2 * (int64_t)g4d::PROP_SLOTS * n : 0;
    HCK(ctx->pairs.ensure(sizeof(uint64_t) * (size_t)(2 * ucap + prop_room + 2 * (int64_t)n * min_ws + 64)));
    uint64_t *pr = buf<uint64_t>(ctx->pairs);
    HCK(g4d::launch_grid_pairs_dedup(idg, GS, nthr, pcnt, pr + ucap, tcnt, s));
    HCK(g4d::launch_tile_totals(tcnt, ntiles, ttot, s));
    HCK(g4d::scan_excl_i64(P, ttot, ttot, ntiles, ttot + ntiles, s));
    HCK(g4d::launch_gather_tiles(pr + ucap, pcnt, tcnt, ntiles, ttot, pr, s));
    HCK(hipMemsetAsync(cnt + G4D_CNT_APPEND0, 0, 8, s));
    HCK(g4d::launch_missing(vox, n, GS, idg, pr + ucap + ngp, cnt + G4D_CNT_APPEND0, n, s));
    int64_t ngd_f = 0, *ngd_h = hslot(ctx, 1, &ngd_f);
    HCK(hipMemcpyAsync(ngd_h, ttot + ntiles, 8, hipMemcpyDeviceToHost, s));
    CK(read_counters(ctx, hc));
    const int64_t ngd = *ngd_h;
    CK(check_dev_error(ctx, hc[G4D_CNT_ERROR], "morton"));
    const int64_t nmiss = (int64_t)hc[G4D_CNT_APPEND0];
    if (nmiss > 0)
        HCK(hipMemcpyAsync(pr + ngd, pr + ucap + ngp, sizeof(uint64_t) * (size_t)nmiss, hipMemcpyDeviceToDevice, s));
    trace("n %d in %d set(s), grid %d: grid pairs %lld raw, %lld after tile dedup, missing %lld", n, S, gmax,
          (long long)ngp, (long long)ngd, (long long)nmiss);
    // Both directions of every pair, then one sort and one unique: each distinct pair leaves
    // exactly its two directed keys (a missing point appears in no grid pair).
    const int64_t u = ngd + nmiss;
    HCK(g4d::launch_reciprocate(pr, u, s));
    int64_t m = 2 * u;
    HCK(sort_unique_pairs(P, pr, &m, nb, n, s));
    ctx->st.num_missing = nmiss;
    ctx->st.num_grid_pairs = m / 2 - nmiss;
    HCK(ctx->ws_off.ensure(sizeof(int64_t) * (size_t)(n + 1)));
    int64_t *ws_off = buf<int64_t>(ctx->ws_off);
    HCK(g4d::launch_csr_offsets(pr, m, n, ws_off, s));
    // learned proposal (plug-in point of _pumpWorksets, GD/GDelaunay/GDelHost.cu:352-529):
    // candidates and features from the digital Voronoi diagram, then extra workset pairs
    ev_record(ctx, EV_PROP0);
    int64_t pred_vertices = 0;
    if (ctx->prop_on || ctx->prop_export) {
        CK(proposal_stage(ctx, n, GS, idg, pr, &m, ws_off, nb, &pred_vertices));
        HCK(g4d::launch_csr_offsets(pr, m, n, ws_off, s));
    }
    ev_record(ctx, EV_PROP1);
    HCK(hipMemsetAsync(cnt + G4D_CNT_APPEND0, 0, 8, s));
    HCK(g4d::launch_pump(pr, ws_off, n, GS, min_ws, pr + m, cnt + G4D_CNT_APPEND0, 2 * (int64_t)n * min_ws + 64, s));
    CK(read_counters(ctx, hc));
    const int64_t npump = (int64_t)hc[G4D_CNT_APPEND0];
    if (npump > 0) {
        // A pumped candidate is not in the star's list, and the lists are symmetric, so no
        // pumped key is already present: sort and deduplicate the pumped keys alone (two stars
        // may pump each other) and merge them into the sorted worksets.
        int64_t np = 2 * npump;
        CK(guard_keys(ctx, m + np, "pumped worksets"));
        uint64_t *pk = pr + m;
        HCK(sort_unique_pairs(P, pk, &np, nb, n, s));
        HCK(P.sort_tmp_keys.ensure(sizeof(uint64_t) * (size_t)(m + np)));
        uint64_t *merged = P.sort_tmp_keys.as<uint64_t>();
        HCK(g4d::merge_u64(pr, m, pk, np, merged, s));
        HCK(hipMemcpyAsync(pr, merged, sizeof(uint64_t) * (size_t)(m + np), hipMemcpyDeviceToDevice, s));
        ctx->st.num_pumped = np;
        m += np;
        HCK(g4d::launch_csr_offsets(pr, m, n, ws_off, s));
    }
    ctx->st.num_workset = m;
    trace("worksets %lld keys (%lld pumped)", (long long)m, (long long)ctx->st.num_pumped);
    ev_record(ctx, EV_PAIRS);

    // ---- E: star creation (GD/GDelaunay/GDelHost.cu:543-743) ----
    HCK(ctx->vcap.ensure(sizeof(int64_t) * (size_t)(n + 1)));
    HCK(ctx->tcap.ensure(sizeof(int64_t) * (size_t)(n + 1)));
    HCK(ctx->hcap.ensure(sizeof(int64_t) * (size_t)(n + 1)));
    HCK(ctx->mid.ensure(sizeof(int) * (size_t)n));
    HCK(ctx->big.ensure(sizeof(int) * (size_t)n));
    HCK(ctx->huge.ensure(sizeof(int) * (size_t)n));
    HCK(ctx->xstars.ensure(sizeof(int) * (size_t)n));
    int64_t *vcap = buf<int64_t>(ctx->vcap), *tcap = buf<int64_t>(ctx->tcap), *hcap = buf<int64_t>(ctx->hcap);
    const int nblk_star = 5120; // 256 CUs x 20 resident waves of the 32-vertex tier (5 per SIMD)
    const int lds_limit = std::min(g4d::STAR_LDS_LIMIT, ctx->opt.lds_star_limit > 0 ? ctx->opt.lds_star_limit
                                                                                  : g4d::STAR_LDS_LIMIT);

    Store &store = ctx->store;
    HCK(store.vnum.ensure(sizeof(int) * (size_t)n));
    HCK(store.tnum.ensure(sizeof(int) * (size_t)n));
    HCK(store.hnum.ensure(sizeof(int) * (size_t)n));
    HCK(store.voff.ensure(sizeof(int64_t) * (size_t)n));
    HCK(store.toff.ensure(sizeof(int64_t) * (size_t)n));
    HCK(store.hoff.ensure(sizeof(int64_t) * (size_t)n));
    HCK(store.nvoff.ensure(sizeof(int64_t) * (size_t)n));
    HCK(store.ntoff.ensure(sizeof(int64_t) * (size_t)n));
    HCK(store.nhoff.ensure(sizeof(int64_t) * (size_t)n));
    // empty records until a star is built (a star that fails is reported, never read)
    HCK(hipMemsetAsync(store.vnum.ptr, 0, sizeof(int) * (size_t)n, s));
    HCK(hipMemsetAsync(store.tnum.ptr, 0, sizeof(int) * (size_t)n, s));
    HCK(hipMemsetAsync(store.hnum.ptr, 0, sizeof(int) * (size_t)n, s));
    store.top[0] = store.top[1] = store.top[2] = 0;
    store.repacks = 0;

    // One pass over the stars with insertions (mode 0: every star, from its workset).  Each
    // such star gets a fresh region at the arenas' tops; the others are not touched.
    auto star_pass = [&](int mode, const uint64_t *ins_keys, const int64_t *ins_off, int64_t nins, int64_t *drown_n,
                         int64_t *tp_n) -> int {
        const int *list = nullptr;
        int nitems = n;
        // sparse passes: the active list's length is known on the device first; the capacities
        // and their scans run over the bound nins, and the length comes back with the needs
        const int64_t *dev_na = nullptr;
        int64_t na_f = 0, *na_h = hslot(ctx, 9, &na_f);
        if (mode == 1 && nins < n / 2) {
            // few insertions: offsets and active list from the runs of the sorted keys, O(keys)
            // instead of O(stars)
            HCK(ctx->active.ensure(sizeof(int) * (size_t)n));
            HCK(ctx->hkeys.ensure(sizeof(uint64_t) * (size_t)std::max<int64_t>(nins, 1)));
            HCK(P.sort_tmp_keys.ensure(sizeof(uint64_t) * (size_t)std::max<int64_t>(nins, 1)));
            uint64_t *hk = buf<uint64_t>(ctx->hkeys);
            HCK(g4d::launch_sparse_runs(ins_keys, nins, const_cast<int64_t *>(ins_off), hk, s));
            uint64_t *runs = P.sort_tmp_keys.as<uint64_t>();
            HCK(g4d::compact_valid_u64_deferred(P, hk, runs, nins, g4d::STAR_TASK_NONE, na_h, &dev_na, s));
            HCK(g4d::launch_u64_to_i32(runs, nins, dev_na, buf<int>(ctx->active), s));
            nitems = (int)nins; // bound until the needs come back
            list = buf<int>(ctx->active);
        } else if (mode == 1) {
            HCK(g4d::launch_csr_offsets(ins_keys, nins, n, const_cast<int64_t *>(ins_off), s));
            HCK(ctx->aflags.ensure(sizeof(uint32_t) * (size_t)n));
            HCK(ctx->apos.ensure(sizeof(uint32_t) * (size_t)n));
            HCK(ctx->active.ensure(sizeof(int) * (size_t)n));
            uint32_t *fl = buf<uint32_t>(ctx->aflags), *pos = buf<uint32_t>(ctx->apos);
            HCK(g4d::launch_active_flags(ins_off, n, fl, s));
            HCK(g4d::scan_excl_u32(P, fl, pos, n, s));
            HCK(g4d::launch_active_scatter(fl, pos, n, buf<int>(ctx->active), s));
            int64_t last_f = 0, *last_h = hslot(ctx, 2, &last_f);
            uint32_t *last = reinterpret_cast<uint32_t *>(last_h);
            HCK(hipMemcpyAsync(&last[0], pos + n - 1, 4, hipMemcpyDeviceToHost, s));
            HCK(hipMemcpyAsync(&last[1], fl + n - 1, 4, hipMemcpyDeviceToHost, s));
            HCK(hipStreamSynchronize(s));
            nitems = (int)(last[0] + last[1]);
            list = buf<int>(ctx->active);
        }
        HCK(ctx->spcount.ensure(sizeof(int) * (size_t)n));
        int *spc = buf<int>(ctx->spcount);
        HCK(g4d::launch_capacities(list, nitems, dev_na, mode ? buf<int>(store.vnum) : nullptr,
                                   mode ? buf<int>(store.hnum) : nullptr, ins_off, vcap, tcap, hcap, spc, s));
        int64_t need[3] = {0, 0, 0};
        int64_t *caps[3] = {vcap, tcap, hcap};
        int64_t *need_h = hslot(ctx, 3, need); // three slots
        for (int k = 0; k < 3; k++) {
            HCK(g4d::scan_excl_i64(P, caps[k], caps[k], nitems, caps[k] + nitems, s));
            HCK(hipMemcpyAsync(&need_h[k], caps[k] + nitems, 8, hipMemcpyDeviceToHost, s));
        }
        HCK(hipStreamSynchronize(s));
        for (int k = 0; k < 3; k++) need[k] = need_h[k];
        if (dev_na) nitems = (int)*na_h;
        if (mode == 0) {
            // room for the splaying rounds: their regions come to about twice the creation's;
            // with the proposal's predicted link sizes, room for that many rebuilt stars too
            int64_t want[3] = {3 * need[0], 3 * need[1], 3 * need[2]};
            if (pred_vertices > 0) {
                want[0] = std::max(want[0], 3 * pred_vertices);
                want[1] = std::max(want[1], 3 * (2 * pred_vertices - 4 * (int64_t)n));
                want[2] = std::max(want[2], 3 * pred_vertices);
            }
            for (int k = 0; k < 3; k++)
                HCK(store.arena[k].ensure(Store::esz[k] * (size_t)std::max<int64_t>(want[k], 1024)));
        } else if (store.top[0] + need[0] > store.capacity(0) || store.top[1] + need[1] > store.capacity(1) ||
                   store.top[2] + need[2] > store.capacity(2)) {
            CK(repack_store(ctx, n, need));
        }
        HCK(g4d::launch_assign_regions(list, nitems, vcap, tcap, hcap, store.top[0], store.top[1], store.top[2],
                                       buf<int64_t>(store.nvoff), buf<int64_t>(store.ntoff), buf<int64_t>(store.nhoff),
                                       s));
        const int64_t tbase = store.top[1];
        for (int k = 0; k < 3; k++) store.top[k] += need[k];
        const int64_t dcap = std::max<int64_t>(nins, 1);
        const int64_t npb = 2 * need[1];
        CK(guard_keys(ctx, std::max(npb, nins), "star pass outputs"));
        const int64_t tpcap = npb + 4 * dcap + 64;
        HCK(ctx->drown.ensure(sizeof(uint64_t) * (size_t)dcap));
        HCK(ctx->dflag.ensure((size_t)dcap));
        HCK(ctx->pairbuf.ensure(sizeof(uint64_t) * (size_t)std::max<int64_t>(npb, 1)));
        HCK(ctx->cand.ensure(sizeof(uint64_t) * (size_t)tpcap));
        g4d::StarParams sp;
        memset(&sp, 0, sizeof(sp));
        sp.pts = pts;
        sp.n = n;
        sp.mode = mode;
        sp.ins_keys = ins_keys;
        sp.ins_off = ins_off;
        sp.vnum = buf<int>(store.vnum);
        sp.voff = buf<int64_t>(store.voff);
        sp.tnum = buf<int>(store.tnum);
        sp.toff = buf<int64_t>(store.toff);
        sp.hnum = buf<int>(store.hnum);
        sp.hoff = buf<int64_t>(store.hoff);
        sp.vpool = buf<int>(store.arena[0]);
        sp.tpool = buf<uint64_t>(store.arena[1]);
        sp.hpool = buf<int>(store.arena[2]);
        sp.nvoff = buf<int64_t>(store.nvoff);
        sp.ntoff = buf<int64_t>(store.ntoff);
        sp.nhoff = buf<int64_t>(store.nhoff);
        sp.dflag = buf<uint8_t>(ctx->dflag);
        sp.pairbuf = buf<uint64_t>(ctx->pairbuf);
        sp.tbase = tbase;
        sp.pcount = spc;
        sp.cnt = cnt;
        sp.lds_limit = lds_limit;
        sp.defer_list = buf<int>(ctx->mid);
        sp.defer_cap = n;
        sp.defer_counter = G4D_CNT_APPEND6;
        sp.exact_list = buf<int>(ctx->xstars);
        sp.exact_cap = n;
        sp.exact_counter = G4D_CNT_APPEND4;
        sp.t0_variant = ctx->star_t0_variant;
        // list counters, filter counts and the six work-queue heads of this pass, in one launch
        {
            uint32_t zmask = (1u << G4D_CNT_APPEND2) | (1u << G4D_CNT_APPEND3) | (1u << G4D_CNT_APPEND4) |
                             (1u << G4D_CNT_APPEND6) | (1u << G4D_CNT_XMID) | (1u << G4D_CNT_XBIG) |
                             (1u << G4D_CNT_ORIENT4) | (1u << G4D_CNT_COFACTOR);
            for (int q = 0; q < 6; q++) zmask |= 1u << (G4D_CNT_QUEUE0 + q);
            HCK(g4d::launch_zero_counters(cnt, zmask, s));
        }
        int queue = G4D_CNT_QUEUE0;
        sp.work_counter = queue++;
        sp.star_list = list;
        sp.nlist = nitems;
        sp.nlist_dev = nullptr;
        // Filter-only LDS kernels by size tier (32 -> 64 -> 256 link vertices); each passes the
        // stars it cannot hold to the next tier's list and the ones its filters cannot decide
        // to the exact list.  Then the exact-capable kernels of the same tiers over the exact
        // list.  Every later launch reads its list length on the device, so the chain runs
        // without host round trips; lists and counters are reused in stream order.
        if (ctx->ev_ok) (void)hipEventRecord(ctx->ev_t0[0], s);
        HCK(g4d::launch_star_kernel(sp, g4d::STAR_VMAX_T0, false, nblk_star, s));
        if (ctx->ev_ok) (void)hipEventRecord(ctx->ev_t0[1], s);
        auto chain = [&](int *lst, int counter, int *defer, int defer_counter, int vmax, bool exact) -> int {
            sp.star_list = lst;
            sp.nlist = n;
            sp.nlist_dev = cnt + counter;
            sp.defer_list = defer;
            sp.defer_counter = defer_counter;
            sp.work_counter = queue++;
            // a tier's list is a subset of the pass's stars: late rounds with a few thousand stars
            // dispatch a few hundred work-groups instead of thousands that would leave at once
            const int full = vmax == g4d::STAR_VMAX_T1 && !exact ? 4096 : 1024;
            const int nblk = std::max(64, std::min(full, (nitems + 7) / 8));
            HCK(g4d::launch_star_kernel(sp, vmax, exact, nblk, s));
            return G4D_OK;
        };
        // filter tiers: mid list (A6) -> 64, big list (A2) -> 256, beyond -> huge (A3)
        CK(chain(buf<int>(ctx->mid), G4D_CNT_APPEND6, buf<int>(ctx->big), G4D_CNT_APPEND2, g4d::STAR_VMAX_T1, false));
        CK(chain(buf<int>(ctx->big), G4D_CNT_APPEND2, buf<int>(ctx->huge), G4D_CNT_APPEND3, g4d::STAR_VMAX_T2, false));
        // the exact chain reuses the mid/big lists (the filter chain is done with them in stream
        // order) under counters of its own, so the filter tiers' counts stay for the statistics
        CK(chain(buf<int>(ctx->xstars), G4D_CNT_APPEND4, buf<int>(ctx->mid), G4D_CNT_XMID, g4d::STAR_VMAX_T0, true));
        CK(chain(buf<int>(ctx->mid), G4D_CNT_XMID, buf<int>(ctx->big), G4D_CNT_XBIG, g4d::STAR_VMAX_T1, true));
        CK(chain(buf<int>(ctx->big), G4D_CNT_XBIG, buf<int>(ctx->huge), G4D_CNT_APPEND3, g4d::STAR_VMAX_T2, true));
        sp.nlist_dev = nullptr;
        CK(read_counters(ctx, hc));
        {
            float ms = 0.f;
            if (ctx->ev_ok && hipEventElapsedTime(&ms, ctx->ev_t0[0], ctx->ev_t0[1]) == hipSuccess)
                ctx->st.t_star_t0_ms += ms;
            ctx->st.star_t0_filter_evals += (int64_t)hc[G4D_CNT_ORIENT4];
            ctx->st.star_t0_cofactors += (int64_t)hc[G4D_CNT_COFACTOR];
            ctx->st.star_t0_launches++;
        }
        const int64_t nmid = (int64_t)hc[G4D_CNT_APPEND6], nbig = (int64_t)hc[G4D_CNT_APPEND2];
        const int64_t nexact = (int64_t)hc[G4D_CNT_APPEND4];
        ctx->st.big_stars += (int32_t)nbig;
        ctx->st.exact_stars += nexact;
        const int64_t nhuge = (int64_t)hc[G4D_CNT_APPEND3];
        if (nhuge > 0) {
            ctx->st.huge_stars += (int32_t)nhuge;
            HCK(hipMemsetAsync(cnt + G4D_CNT_MISC, 0, 8, s));
            HCK(g4d::launch_need_max(sp, buf<int>(ctx->huge), (int)nhuge, cnt + G4D_CNT_MISC, s));
            CK(read_counters(ctx, hc));
            const int vcap_g = (int)hc[G4D_CNT_MISC];
            const int64_t stride = (g4d::star_global_bytes(vcap_g) + 4095) & ~(int64_t)4095;
            // bounded workspace: at most 256 blocks and 1 GiB
            const int nblk = (int)std::max<int64_t>(1, std::min<int64_t>({nhuge, 256, ((int64_t)1 << 30) / stride}));
            HCK(ctx->gwork.ensure((size_t)(stride * nblk)));
            HCK(g4d::launch_star_global(sp, buf<int>(ctx->huge), (int)nhuge, buf<char>(ctx->gwork), stride, vcap_g,
                                        nblk, s));
            CK(read_counters(ctx, hc));
            trace("global-memory star kernel: %lld stars, largest link plus insertions %d", (long long)nhuge, vcap_g);
        }
        CK(check_dev_error(ctx, hc[G4D_CNT_ERROR], mode ? "splaying" : "star creation"));
        // drowned insertions and triangle pairs, in star order (both counts read back together)
        int64_t cnt_f[2] = {0, 0}, *cnt_h = hslot(ctx, 6, cnt_f); // two slots
        HCK(g4d::compact_u64_deferred(P, ins_keys, buf<uint8_t>(ctx->dflag), buf<uint64_t>(ctx->drown), nins,
                                      &cnt_h[0], s));
        // each star wrote its pairs densely at the front of its region and their count: offsets
        // in list order (the vertex capacities are done with), then one copy per star
        HCK(g4d::launch_pair_counts(list, nitems, spc, vcap, s));
        HCK(g4d::scan_excl_i64(P, vcap, vcap, nitems, vcap + nitems, s));
        HCK(hipMemcpyAsync(&cnt_h[1], vcap + nitems, 8, hipMemcpyDeviceToHost, s));
        HCK(hipStreamSynchronize(s));
        *drown_n = cnt_h[0];
        const int64_t npairs = cnt_h[1];
        if (npairs > npb) return set_err(ctx, G4D_ERR_INTERNAL, "star pairs exceed their regions");
        HCK(g4d::launch_pair_gather(list, nitems, spc, vcap, buf<int64_t>(store.ntoff), tbase,
                                    buf<uint64_t>(ctx->pairbuf), buf<uint64_t>(ctx->cand), s));
        *tp_n = npairs;
        trace("star pass mode %d: %lld insertions into %d stars, over 32: %lld, over 64: %lld, exact %lld, "
              "global-memory %lld, drowned %lld, pairs %lld",
              mode, (long long)nins, nitems, (long long)nmid, (long long)nbig, (long long)nexact, (long long)nhuge,
              (long long)*drown_n, (long long)*tp_n);
        return G4D_OK;
    };

    int64_t nd = 0, ntp = 0;
    CK(star_pass(0, pr, ws_off, m, &nd, &ntp));
    ev_record(ctx, EV_STAR);

    // ---- F: splaying rounds (GD/GDelaunay/GDelHost.cu:1389-1470) ----
    int rounds = 0;
    bool converged = false;
    for (int it = 0; it < ctx->opt.max_rounds; it++) {
        // F1-F3: drowned insertions, mutual removal, containment proofs
        int64_t ntask = 0;
        if (nd > 0) {
            CK(guard_keys(ctx, nd, "drowned insertions"));
            HCK(ctx->dkeys.ensure(sizeof(uint64_t) * (size_t)nd));
            HCK(ctx->tasks.ensure(sizeof(uint64_t) * (size_t)nd));
            uint64_t *dk = buf<uint64_t>(ctx->dkeys);
            HCK(g4d::launch_drown_keys(buf<uint64_t>(ctx->drown), nd, dk, s));
            HCK(g4d::sort_keys_ab(P, dk, nd, 33, nb, 1 + nb, n, false, s));
            HCK(g4d::launch_nonmutual(dk, nd, buf<uint64_t>(ctx->tasks), s));
            HCK(g4d::compact_valid_u64(P, buf<uint64_t>(ctx->tasks), buf<uint64_t>(ctx->tasks), nd,
                                       g4d::STAR_TASK_NONE, &ntask, s));
            if (ntask > 0) {
                g4d::ProofParams pp;
                pp.pts = pts;
                pp.tasks = buf<uint64_t>(ctx->tasks);
                pp.ntasks = ntask;
                pp.vnum = buf<int>(store.vnum);
                pp.voff = buf<int64_t>(store.voff);
                pp.vpool = buf<int>(store.arena[0]);
                pp.tnum = buf<int>(store.tnum);
                pp.toff = buf<int64_t>(store.toff);
                pp.tpool = buf<uint64_t>(store.arena[1]);
                pp.out = buf<uint64_t>(ctx->cand) + ntp;
                pp.cnt = cnt;
                HCK(ctx->xtasks.ensure(sizeof(int64_t) * (size_t)ntask));
                pp.xlist = buf<int64_t>(ctx->xtasks);
                pp.xcap = ntask;
                pp.nx = 0;
                HCK(hipMemsetAsync(cnt + G4D_CNT_APPEND5, 0, 8, s));
                HCK(g4d::launch_proofs(pp, false, s));
                pp.nx = -1; // the exact pass reads the filter pass's count on the device
                HCK(g4d::launch_proofs(pp, true, s));
            }
        }
        // F4-F5: candidates = triangle pairs + proofs, history filter, sort-unique
        const int64_t ncand = ntp + 4 * ntask;
        int64_t nins = 0;
        if (ncand > 0) {
            CK(guard_keys(ctx, 2 * ncand, "splaying candidates"));
            HCK(ctx->keep.ensure((size_t)ncand));
            HCK(ctx->ins.ensure(sizeof(uint64_t) * (size_t)(2 * ncand + 2)));
            uint64_t *ins = buf<uint64_t>(ctx->ins);
            HCK(g4d::launch_hist_filter(buf<uint64_t>(ctx->cand), ncand, buf<int>(store.arena[2]),
                                        buf<int64_t>(store.hoff), buf<int>(store.hnum), buf<uint8_t>(ctx->keep), s));
            HCK(g4d::compact_u64(P, buf<uint64_t>(ctx->cand), buf<uint8_t>(ctx->keep), ins, ncand, &nins, s));
            if (nins > 0) {
                HCK(sort_unique_pairs(P, ins, &nins, nb, n, s));
            }
        }
        trace("round %d: %lld drowned, %lld proofs, %lld candidates, %lld new pairs", rounds + 1, (long long)nd,
              (long long)ntask, (long long)ncand, (long long)nins);
        if (nins == 0) {
            // the proofs' errors; otherwise the star pass below reads them with its own counters
            CK(read_counters(ctx, hc));
            CK(check_dev_error(ctx, hc[G4D_CNT_ERROR], "proofs"));
            converged = true;
            break;
        }
        rounds++;
        // F6-F9: reciprocate, group by star, insert
        uint64_t *ins = buf<uint64_t>(ctx->ins);
        HCK(g4d::launch_reciprocate(ins, nins, s));
        const int64_t nrec = 2 * nins;
        HCK(sort_pairs(P, ins, nrec, nb, n, s));
        ctx->st.splay_insertions += nrec;
        HCK(ctx->ins_off.ensure(sizeof(int64_t) * (size_t)(n + 1)));
        int64_t *ioff = buf<int64_t>(ctx->ins_off);
        CK(star_pass(1, ins, ioff, nrec, &nd, &ntp)); // builds the offsets of the stars with insertions
    }
    if (!converged) return set_err(ctx, G4D_ERR_NOT_CONVERGED, "splaying did not converge in %d rounds", rounds);
    ctx->st.splay_rounds = rounds;
    ev_record(ctx, EV_SPLAY);

    // ---- G: tetrahedra (GD/GDelaunay/GDelHost.cu:1520-1615) ----
    HCK(ctx->tetcnt.ensure(sizeof(int64_t) * (size_t)(n + 1)));
    HCK(ctx->tetoff.ensure(sizeof(int64_t) * (size_t)(n + 1)));
    g4d::OutParams op;
    op.pts = pts;
    op.n = n;
    op.vnum = buf<int>(store.vnum);
    op.voff = buf<int64_t>(store.voff);
    op.vpool = buf<int>(store.arena[0]);
    op.tnum = buf<int>(store.tnum);
    op.toff = buf<int64_t>(store.toff);
    op.tpool = buf<uint64_t>(store.arena[1]);
    op.star_tet_count = buf<int64_t>(ctx->tetcnt);
    op.star_tet_off = buf<int64_t>(ctx->tetoff);
    op.cnt = cnt;
    op.tets = nullptr;
    // owner triangles the orient3d filter cannot decide: exact pass marks the ones to keep
    int64_t xcap = std::max<int64_t>(n / 16, 4096);
    for (int attempt = 0; attempt < 2; attempt++) {
        HCK(ctx->xtris.ensure(sizeof(uint64_t) * (size_t)xcap));
        HCK(hipMemsetAsync(cnt + G4D_CNT_APPEND5, 0, 8, s));
        HCK(g4d::launch_out_flag(op, buf<uint64_t>(ctx->xtris), xcap, s));
        CK(read_counters(ctx, hc));
        if ((int64_t)hc[G4D_CNT_APPEND5] <= xcap) break;
        xcap = (int64_t)hc[G4D_CNT_APPEND5];
    }
    const int64_t nxt = (int64_t)hc[G4D_CNT_APPEND5];
    if (nxt > xcap) return set_err(ctx, G4D_ERR_INTERNAL, "output exact list sizing failed");
    HCK(g4d::launch_out_exact(op, buf<uint64_t>(ctx->xtris), nxt, s));
    int64_t *toffs = buf<int64_t>(ctx->tetoff);
    HCK(g4d::scan_excl_i64(P, buf<int64_t>(ctx->tetcnt), toffs, n, toffs + n, s));
    int64_t nt_f = 0, *nt_h = hslot(ctx, 8, &nt_f);
    HCK(hipMemcpyAsync(nt_h, toffs + n, 8, hipMemcpyDeviceToHost, s));
    HCK(hipStreamSynchronize(s));
    const int64_t nt = *nt_h;
    CK(guard_tets(ctx, nt));
    HCK(ctx->tets.ensure(sizeof(int) * 4 * (size_t)std::max<int64_t>(nt, 1)));
    op.tets = buf<int>(ctx->tets);
    HCK(g4d::launch_out_write(op, s));
    ev_record(ctx, EV_OUT);
    if (ctx->opt.compute_adjacency) {
        const uint64_t tsz = next_pow2((uint64_t)std::max<int64_t>(8 * nt, 1024));
        HCK(ctx->table.ensure(sizeof(uint32_t) * (size_t)tsz));
        HCK(ctx->opp.ensure(sizeof(int) * 4 * (size_t)std::max<int64_t>(nt, 1)));
        HCK(g4d::launch_adjacency_tiled(buf<int>(ctx->tets), nt, buf<int>(ctx->opp), buf<uint32_t>(ctx->table),
                                        (int64_t)tsz, cnt, s));
    }
    if (batched) {
        // per-set tetrahedron ranges (tetrahedra are grouped by owner star, stars by set) and
        // set-local vertex / tetrahedron ids in each set's input order
        HCK(ctx->settet.ensure(sizeof(int64_t) * (size_t)(S + 1)));
        int64_t *tet_off = buf<int64_t>(ctx->settet);
        HCK(g4d::launch_gather_i64(toffs, GS.pt_off, S + 1, tet_off, s));
        HCK(g4d::launch_batch_localize(buf<int>(ctx->tets), ctx->opt.compute_adjacency ? buf<int>(ctx->opp) : nullptr,
                                       nt, GS.pt_set, GS.pt_off, tet_off, perm, s));
        ctx->h_set_tet_off.resize((size_t)S + 1);
        HCK(hipMemcpyAsync(ctx->h_set_tet_off.data(), tet_off, sizeof(int64_t) * (size_t)(S + 1), hipMemcpyDeviceToHost,
                           s));
        ctx->h_set_grid = spec.g;
    } else if (ctx->opt.index_space == G4D_INDEX_INPUT) {
        HCK(g4d::launch_map_perm(buf<int>(ctx->tets), 4 * nt, perm, s));
    }
    ev_record(ctx, EV_ADJ);
    if (ctx->prop_export) {
        // training labels: which candidates ended in each star's final link, and its size bin
        HCK(ctx->plab.ensure(sizeof(uint32_t) * (size_t)n));
        HCK(g4d::launch_labels(buf<int>(ctx->pcand), n, buf<int>(store.vnum), buf<int64_t>(store.voff),
                               buf<int>(store.arena[0]), buf<uint32_t>(ctx->plab), s));
        ctx->have_samples = true;
        ctx->samples_n = n;
    }
    CK(read_counters(ctx, hc));
    CK(check_dev_error(ctx, hc[G4D_CNT_ERROR], "output"));
    ctx->num_tets = nt;
    ctx->last_n = n;
    ctx->have_result = true;
    ctx->st.num_tets = nt;
    ctx->st.exact_calls = (int64_t)hc[G4D_CNT_EXACT];
    ctx->st.sos_calls = (int64_t)hc[G4D_CNT_SOS];
    ctx->st.pba_bytes = 24 * G3;
    ctx->st.star_store_bytes = store.top[0] * 4 + store.top[1] * 8 + store.top[2] * 4;
    ctx->st.store_repacks = store.repacks;
    ctx->st.star_passes = rounds + 1;
    ctx->st.t_morton_ms = ev_ms(ctx, EV_START, EV_MORTON);
    ctx->st.t_pba_ms = ev_ms(ctx, EV_MORTON, EV_PBA3);
    ctx->st.t_pba_pass_ms[0] = ev_ms(ctx, EV_PBA0, EV_PBA1);
    ctx->st.t_pba_pass_ms[1] = ev_ms(ctx, EV_PBA1, EV_PBA2);
    ctx->st.t_pba_pass_ms[2] = ev_ms(ctx, EV_PBA2, EV_PBA3);
    ctx->st.t_pairs_ms = ev_ms(ctx, EV_PBA3, EV_PAIRS);
    ctx->st.t_prop_ms = ev_ms(ctx, EV_PROP0, EV_PROP1);
    ctx->st.prop_enabled = ctx->prop_on ? 1 : 0;
    ctx->st.t_star_ms = ev_ms(ctx, EV_PAIRS, EV_STAR);
    ctx->st.t_splay_ms = ev_ms(ctx, EV_STAR, EV_SPLAY);
    ctx->st.t_out_ms = ev_ms(ctx, EV_SPLAY, EV_OUT);
    ctx->st.t_adj_ms = ev_ms(ctx, EV_OUT, EV_ADJ);
    ctx->st.t_total_ms = ev_ms(ctx, EV_START, EV_ADJ);
    return G4D_OK;
}

// ------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------

extern "C" {

int g4d_default_options(g4d_options *opt)
{
    if (!opt) return G4D_ERR_INVALID_INPUT;
    opt->grid_size = 256;
    opt->min_workset = 8;
    opt->max_rounds = 1000;
    opt->index_space = G4D_INDEX_INPUT;
    opt->device = 0;
    opt->compute_adjacency = 1;
    opt->lds_star_limit = 0;
    return G4D_OK;
}

int g4d_device_count(void)
{
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

int g4d_create(g4d_ctx **out, const g4d_options *opt)
{
    if (!out) return G4D_ERR_INVALID_INPUT;
    *out = nullptr;
    g4d_ctx *ctx = new (std::nothrow) g4d_ctx();
    if (!ctx) return G4D_ERR_OOM;
    if (opt) ctx->opt = *opt;
    else g4d_default_options(&ctx->opt);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        delete ctx;
        return G4D_ERR_DEVICE;
    }
    if (ctx->opt.device < 0 || ctx->opt.device >= ndev) {
        delete ctx;
        return G4D_ERR_INVALID_INPUT;
    }
    if (hipSetDevice(ctx->opt.device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
        delete ctx;
        return G4D_ERR_DEVICE;
    }
    // device address of the allocation-free entry point's scratch array (once per device)
    ctx->noalloc_status = g4d::noalloc_setup(ctx->opt.device);
    ctx->ev_ok = true;
    for (int i = 0; i < EV_NUM; i++)
        if (hipEventCreate(&ctx->ev[i]) != hipSuccess) ctx->ev_ok = false;
    for (int i = 0; i < 2; i++)
        if (hipEventCreate(&ctx->ev_t0[i]) != hipSuccess) ctx->ev_ok = false;
    // staging of the small host reads; without it they go to pageable variables
    if (ctx->hpin.ensure(sizeof(int64_t) * 64) == hipSuccess) {
        ctx->hstage = static_cast<int64_t *>(ctx->hpin.ptr);
        ctx->P.hstage = ctx->hstage + G4D_CNT_NUM + 16;
    }
    *out = ctx;
    return G4D_OK;
}

void g4d_destroy(g4d_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->opt.device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    DevBuf *bufs[] = {&ctx->xyz_up, &ctx->keys, &ctx->pts, &ctx->perm, &ctx->vox, &ctx->bounds, &ctx->settab, &ctx->ptset, &ctx->grid,
                      &ctx->gtmp, &ctx->idgrid, &ctx->pbaspill, &ctx->pbaenv, &ctx->pbacnt, &ctx->pcount, &ctx->pairs, &ctx->ws_off, &ctx->cnt, &ctx->settet, &ctx->drown, &ctx->dflag, &ctx->pairbuf, &ctx->spcount, &ctx->hkeys, &ctx->dkeys,
                      &ctx->tasks, &ctx->cand, &ctx->keep, &ctx->ins, &ctx->ins_off, &ctx->vcap, &ctx->tcap,
                      &ctx->hcap, &ctx->aflags, &ctx->apos, &ctx->active, &ctx->mid, &ctx->big, &ctx->huge, &ctx->xstars, &ctx->xtasks, &ctx->xtris, &ctx->gwork, &ctx->tetcnt, &ctx->tetoff, &ctx->tets, &ctx->opp, &ctx->table,
                      &ctx->mlp_params, &ctx->fkeys, &ctx->ford, &ctx->frank, &ctx->pcand, &ctx->pd2, &ctx->pfeat,
                      &ctx->ppairs, &ctx->pcap, &ctx->plab, &ctx->pvcap};
    destroy_graphs(ctx);
    for (DevBuf *b : bufs) b->release();
    ctx->h_tets.release();
    ctx->h_opp.release();
    ctx->h_perm.release();
    ctx->hpin.release();
    ctx->store.release();
    ctx->P.release();
    if (ctx->ev_ok)
        for (int i = 0; i < EV_NUM; i++) (void)hipEventDestroy(ctx->ev[i]);
    if (ctx->ev_ok)
        for (int i = 0; i < 2; i++) (void)hipEventDestroy(ctx->ev_t0[i]);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *g4d_last_error(const g4d_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : "null context";
}

int g4d_set_options(g4d_ctx *ctx, const g4d_options *opt)
{
    if (!ctx || !opt) return G4D_ERR_INVALID_INPUT;
    if (opt->device != ctx->opt.device) return set_err(ctx, G4D_ERR_INVALID_INPUT, "device cannot change");
    ctx->opt = *opt;
    return G4D_OK;
}

int g4d_upload_points(g4d_ctx *ctx, const double *xyz, int32_t n, const double **d_xyz)
{
    if (!ctx || !xyz || n <= 0 || !d_xyz) return G4D_ERR_INVALID_INPUT;
    HCK(hipSetDevice(ctx->opt.device));
    HCK(ctx->xyz_up.ensure(sizeof(double) * 3 * (size_t)n));
    HCK(hipMemcpyAsync(ctx->xyz_up.ptr, xyz, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, ctx->stream));
    HCK(hipStreamSynchronize(ctx->stream));
    *d_xyz = buf<double>(ctx->xyz_up);
    return G4D_OK;
}

int g4d_triangulate_device(g4d_ctx *ctx, const double *d_xyz, int32_t n, int64_t *num_tets)
{
    if (!ctx || !d_xyz) return G4D_ERR_INVALID_INPUT;
    HCK(hipSetDevice(ctx->opt.device));
    const int r = run_device(ctx, d_xyz, n, single_set(n, ctx->opt.grid_size));
    if (r != G4D_OK) return r;
    if (num_tets) *num_tets = ctx->num_tets;
    return G4D_OK;
}

int g4d_check_sizes(int64_t sort_keys, int64_t num_tets)
{
    if (sort_keys < 0 || num_tets < 0) return G4D_ERR_INVALID_INPUT;
    if (sort_keys > G4D_MAX_SORT_KEYS || num_tets > G4D_MAX_TETS) return G4D_ERR_CAPACITY;
    return G4D_OK;
}

// Smallest grid of the batched sizes (multiples of 8 within [32, 256]) with at least vpp
// voxels per point.
static int grid_for_density(int64_t n, double vpp)
{
    int g = 32;
    while (g < g4d::BATCH_MAX_GRID && (double)g * g * g < vpp * (double)n) g += 8;
    return g;
}

int g4d_batch_grid_size(int64_t n)
{
    // about 24 voxels per point (C1 has 21, C2 32), in steps of 16, within [32, 256]
    int g = 32;
    while (g < g4d::BATCH_MAX_GRID && (int64_t)g * g * g < 24 * n) g += 16;
    return g;
}

int g4d_set_star_kernel(g4d_ctx *ctx, int32_t variant)
{
    if (!ctx || variant < -1 || variant >= g4d::STAR_T0_NUM) return G4D_ERR_INVALID_INPUT;
    ctx->star_t0_variant = variant;
    return G4D_OK;
}

int g4d_set_grid_table(g4d_ctx *ctx, int32_t nbuckets, const int64_t *max_points, const float *voxels_per_point)
{
    if (!ctx || nbuckets < 0 || nbuckets > 64 || (nbuckets > 0 && (!max_points || !voxels_per_point)))
        return G4D_ERR_INVALID_INPUT;
    for (int k = 0; k < nbuckets; k++) {
        if (!(voxels_per_point[k] > 0.f) || max_points[k] <= 0 || (k > 0 && max_points[k] <= max_points[k - 1]))
            return set_err(ctx, G4D_ERR_INVALID_INPUT, "grid table entry %d invalid", k);
    }
    ctx->grid_tab_n.assign(max_points, max_points + nbuckets);
    ctx->grid_tab_vpp.assign(voxels_per_point, voxels_per_point + nbuckets);
    return G4D_OK;
}

int g4d_set_stage_graphs(g4d_ctx *ctx, int32_t mode)
{
    if (!ctx || mode < -1 || mode > 1) return G4D_ERR_INVALID_INPUT;
    ctx->stage_graphs = mode;
    if (mode == 0) {
        HCK(hipSetDevice(ctx->opt.device));
        HCK(hipStreamSynchronize(ctx->stream));
        destroy_graphs(ctx);
    }
    return G4D_OK;
}

int g4d_set_pba_bands(g4d_ctx *ctx, int32_t bands)
{
    if (!ctx || bands < 0 || bands > g4d::PBA_MAX_BANDS) return G4D_ERR_INVALID_INPUT;
    ctx->pba_bands = bands;
    return G4D_OK;
}

int g4d_grid_for_set(const g4d_ctx *ctx, int64_t n)
{
    if (!ctx || ctx->grid_tab_n.empty()) return g4d_batch_grid_size(n);
    size_t k = 0;
    while (k + 1 < ctx->grid_tab_n.size() && n > ctx->grid_tab_n[k]) k++;
    return grid_for_density(n, (double)ctx->grid_tab_vpp[k]);
}

static int batch_spec(g4d_ctx *ctx, const int64_t *set_offsets, int32_t nsets, const int32_t *grids, int64_t *n_out,
                      SetSpec &spec)
{
    if (!set_offsets || nsets < 1 || nsets > g4d::BATCH_MAX_SETS)
        return set_err(ctx, G4D_ERR_INVALID_INPUT, "batch needs 1..%d sets", g4d::BATCH_MAX_SETS);
    if (set_offsets[0] != 0) return set_err(ctx, G4D_ERR_INVALID_INPUT, "set_offsets[0] must be 0");
    spec.nsets = nsets;
    spec.pt_off.assign(set_offsets, set_offsets + nsets + 1);
    spec.g.resize((size_t)nsets);
    for (int k = 0; k < nsets; k++) {
        if (set_offsets[k + 1] <= set_offsets[k]) return set_err(ctx, G4D_ERR_INVALID_INPUT, "set %d is empty", k);
        spec.g[(size_t)k] = grids ? grids[k] : g4d_grid_for_set(ctx, set_offsets[k + 1] - set_offsets[k]);
    }
    if (set_offsets[nsets] > (1 << 29)) return set_err(ctx, G4D_ERR_INVALID_INPUT, "too many points in the batch");
    *n_out = set_offsets[nsets];
    return G4D_OK;
}

int g4d_triangulate_batch_device(g4d_ctx *ctx, const double *d_xyz, const int64_t *set_offsets, int32_t nsets,
                                 const int32_t *grids, int64_t *num_tets)
{
    if (!ctx || !d_xyz) return G4D_ERR_INVALID_INPUT;
    HCK(hipSetDevice(ctx->opt.device));
    SetSpec spec;
    int64_t n = 0;
    CK(batch_spec(ctx, set_offsets, nsets, grids, &n, spec));
    if (nsets == 1) {
        // one set: the plain path in the input index space
        const int saved = ctx->opt.index_space;
        ctx->opt.index_space = G4D_INDEX_INPUT;
        const int r = run_device(ctx, d_xyz, (int)n, spec);
        ctx->opt.index_space = saved;
        CK(r);
        ctx->h_set_tet_off = {0, ctx->num_tets};
        ctx->h_set_grid = spec.g;
    } else {
        CK(run_device(ctx, d_xyz, (int)n, spec));
    }
    if (num_tets) *num_tets = ctx->num_tets;
    return G4D_OK;
}

int g4d_triangulate_batch(g4d_ctx *ctx, const double *xyz, const int64_t *set_offsets, int32_t nsets,
                          const int32_t *grids, g4d_batch_result *out)
{
    if (!ctx || !xyz || !out || !set_offsets || nsets < 1) return G4D_ERR_INVALID_INPUT;
    const int64_t n = set_offsets[nsets];
    if (n <= 0 || n > (1 << 29)) return set_err(ctx, G4D_ERR_INVALID_INPUT, "bad batch size");
    const double *d_xyz = nullptr;
    CK(g4d_upload_points(ctx, xyz, (int32_t)n, &d_xyz));
    CK(g4d_triangulate_batch_device(ctx, d_xyz, set_offsets, nsets, grids, nullptr));
    const int64_t nt = ctx->num_tets;
    HCK(ctx->h_tets.ensure(sizeof(int) * (size_t)std::max<int64_t>(4 * nt, 1)));
    if (ctx->opt.compute_adjacency) HCK(ctx->h_opp.ensure(sizeof(int) * (size_t)std::max<int64_t>(4 * nt, 1)));
    if (nt > 0)
        HCK(hipMemcpyAsync(ctx->h_tets.ptr, ctx->tets.ptr, sizeof(int) * 4 * (size_t)nt, hipMemcpyDeviceToHost,
                           ctx->stream));
    if (nt > 0 && ctx->opt.compute_adjacency)
        HCK(hipMemcpyAsync(ctx->h_opp.ptr, ctx->opp.ptr, sizeof(int) * 4 * (size_t)nt, hipMemcpyDeviceToHost,
                           ctx->stream));
    HCK(hipStreamSynchronize(ctx->stream));
    out->num_sets = nsets;
    out->pad = 0;
    out->num_points = n;
    out->num_tets = nt;
    out->tet_offsets = ctx->h_set_tet_off.data();
    out->tets = ctx->h_tets.i32();
    out->opp = ctx->opt.compute_adjacency ? ctx->h_opp.i32() : nullptr;
    out->grid_sizes = ctx->h_set_grid.data();
    return G4D_OK;
}

int g4d_copy_result(g4d_ctx *ctx, int32_t *tets, int32_t *opp, int32_t *perm)
{
    if (!ctx || !ctx->have_result) return G4D_ERR_INVALID_INPUT;
    const int64_t nt = ctx->num_tets;
    if (tets && nt > 0) HCK(hipMemcpyAsync(tets, ctx->tets.ptr, sizeof(int) * 4 * (size_t)nt, hipMemcpyDeviceToHost, ctx->stream));
    if (opp && nt > 0) {
        if (!ctx->opt.compute_adjacency) return set_err(ctx, G4D_ERR_INVALID_INPUT, "adjacency was not computed");
        HCK(hipMemcpyAsync(opp, ctx->opp.ptr, sizeof(int) * 4 * (size_t)nt, hipMemcpyDeviceToHost, ctx->stream));
    }
    if (perm) HCK(hipMemcpyAsync(perm, ctx->perm.ptr, sizeof(int) * (size_t)ctx->last_n, hipMemcpyDeviceToHost, ctx->stream));
    HCK(hipStreamSynchronize(ctx->stream));
    return G4D_OK;
}

int g4d_triangulate_device_noalloc(g4d_ctx *ctx, const double *d_xyz, int32_t n, int32_t *d_tets, int32_t *d_opp,
                                   int64_t max_tets, int32_t *d_perm, int64_t *num_tets, int64_t *scratch_used)
{
    if (!ctx || !d_xyz || !d_tets || max_tets < 0 || (ctx->opt.compute_adjacency && !d_opp))
        return G4D_ERR_INVALID_INPUT;
    if (num_tets) *num_tets = 0;
    if (scratch_used) *scratch_used = 0;
    if (n > G4D_NOALLOC_MAX_POINTS)
        return set_err(ctx, G4D_ERR_CAPACITY, "allocation-free call: %d points, at most %d", n,
                       G4D_NOALLOC_MAX_POINTS);
    if (ctx->opt.grid_size > G4D_NOALLOC_MAX_GRID)
        return set_err(ctx, G4D_ERR_CAPACITY, "allocation-free call: grid %d, at most %d", ctx->opt.grid_size,
                       G4D_NOALLOC_MAX_GRID);
    HCK(hipSetDevice(ctx->opt.device));
    void *base = g4d::noalloc_scratch(ctx->opt.device);
    if (ctx->noalloc_status != hipSuccess || !base)
        return set_err(ctx, G4D_ERR_DEVICE, "allocation-free call: scratch array unavailable (%s)",
                       hipGetErrorString(ctx->noalloc_status));
    // A context of its own for the run: same stream, options and events, every buffer carved
    // from the scratch array.  Nothing in it is freed (the arena owns the memory); the
    // caller's context keeps its buffers and its last result.
    g4d_ctx tmp;
    tmp.opt = ctx->opt;
    tmp.stream = ctx->stream;
    tmp.star_t0_variant = ctx->star_t0_variant;
    tmp.pba_bands = ctx->pba_bands;
    tmp.stage_graphs = 0; // instantiating a graph allocates
    tmp.ev_ok = ctx->ev_ok;
    tmp.hstage = ctx->hstage; // the caller's page-locked staging (allocated at g4d_create)
    tmp.P.hstage = ctx->P.hstage;
    for (int i = 0; i < EV_NUM; i++) tmp.ev[i] = ctx->ev[i];
    for (int i = 0; i < 2; i++) tmp.ev_t0[i] = ctx->ev_t0[i];
    g4d::ScratchArena arena;
    arena.base = static_cast<char *>(base);
    arena.cap = (size_t)G4D_NOALLOC_SCRATCH_BYTES;
    int r;
    {
        g4d::ScratchScope scope(&arena);
        r = run_device(&tmp, d_xyz, n, single_set(n, ctx->opt.grid_size));
        if (r == G4D_OK) {
            const int64_t nt = tmp.num_tets;
            if (num_tets) *num_tets = nt;
            if (nt > max_tets) {
                r = set_err(&tmp, G4D_ERR_CAPACITY, "allocation-free call: %lld tetrahedra, output holds %lld",
                            (long long)nt, (long long)max_tets);
            } else {
                hipStream_t s = ctx->stream;
                hipError_t e = hipSuccess;
                if (nt > 0)
                    e = hipMemcpyAsync(d_tets, tmp.tets.ptr, sizeof(int) * 4 * (size_t)nt, hipMemcpyDeviceToDevice, s);
                if (e == hipSuccess && nt > 0 && ctx->opt.compute_adjacency)
                    e = hipMemcpyAsync(d_opp, tmp.opp.ptr, sizeof(int) * 4 * (size_t)nt, hipMemcpyDeviceToDevice, s);
                if (e == hipSuccess && d_perm)
                    e = hipMemcpyAsync(d_perm, tmp.perm.ptr, sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice, s);
                if (e == hipSuccess) e = hipStreamSynchronize(s);
                if (e != hipSuccess) r = hip_fail(&tmp, e, "allocation-free output copy", __LINE__);
            }
        } else if (arena.exhausted) {
            r = set_err(&tmp, G4D_ERR_OOM, "allocation-free call: input needs more than the %lld-byte scratch array",
                        (long long)G4D_NOALLOC_SCRATCH_BYTES);
        }
    }
    if (scratch_used) *scratch_used = (int64_t)arena.top;
    ctx->st = tmp.st;
    if (r != G4D_OK) ctx->err = tmp.err;
    return r;
}

int g4d_triangulate(g4d_ctx *ctx, const double *xyz, int32_t n, g4d_result *out)
{
    if (!ctx || !xyz || !out) return G4D_ERR_INVALID_INPUT;
    const double *d_xyz = nullptr;
    int r = g4d_upload_points(ctx, xyz, n, &d_xyz);
    if (r != G4D_OK) return r;
    r = run_device(ctx, d_xyz, n, single_set(n, ctx->opt.grid_size));
    if (r != G4D_OK) return r;
    const int64_t nt = ctx->num_tets;
    HCK(ctx->h_tets.ensure(sizeof(int) * (size_t)std::max<int64_t>(4 * nt, 1)));
    if (ctx->opt.compute_adjacency) HCK(ctx->h_opp.ensure(sizeof(int) * (size_t)std::max<int64_t>(4 * nt, 1)));
    HCK(ctx->h_perm.ensure(sizeof(int) * (size_t)n));
    r = g4d_copy_result(ctx, ctx->h_tets.i32(), ctx->opt.compute_adjacency ? ctx->h_opp.i32() : nullptr,
                        ctx->h_perm.i32());
    if (r != G4D_OK) return r;
    out->num_tets = nt;
    out->tets = ctx->h_tets.i32();
    out->opp = ctx->opt.compute_adjacency ? ctx->h_opp.i32() : nullptr;
    out->perm = ctx->h_perm.i32();
    out->num_points = n;
    return G4D_OK;
}

int g4d_set_proposal(g4d_ctx *ctx, const float *params, int32_t nparams, float threshold, int32_t enable)
{
    if (!ctx) return G4D_ERR_INVALID_INPUT;
    if (!enable) {
        ctx->prop_on = false;
        return G4D_OK;
    }
    if (!params || nparams != g4d::MLP_NPARAM)
        return set_err(ctx, G4D_ERR_INVALID_INPUT, "proposal needs %d parameters", g4d::MLP_NPARAM);
    if (!std::isfinite(threshold)) return set_err(ctx, G4D_ERR_INVALID_INPUT, "threshold must be finite");
    HCK(hipSetDevice(ctx->opt.device));
    HCK(ctx->mlp_params.ensure(sizeof(float) * g4d::MLP_NPARAM));
    HCK(hipMemcpyAsync(ctx->mlp_params.ptr, params, sizeof(float) * g4d::MLP_NPARAM, hipMemcpyHostToDevice,
                       ctx->stream));
    HCK(hipStreamSynchronize(ctx->stream));
    ctx->prop_threshold = threshold;
    ctx->prop_on = true;
    return G4D_OK;
}

int g4d_set_sample_export(g4d_ctx *ctx, int32_t enable)
{
    if (!ctx) return G4D_ERR_INVALID_INPUT;
    ctx->prop_export = enable != 0;
    return G4D_OK;
}

int g4d_get_samples(g4d_ctx *ctx, int64_t offset, int64_t stride, float *features, uint32_t *labels,
                    int32_t *candidates, int64_t cap, int64_t *count)
{
    if (!ctx || !features || !labels || !count || offset < 0 || stride < 1 || cap < 0) return G4D_ERR_INVALID_INPUT;
    if (!ctx->have_samples) return set_err(ctx, G4D_ERR_INVALID_INPUT, "no samples: enable export before the call");
    HCK(hipSetDevice(ctx->opt.device));
    const int n = ctx->samples_n;
    std::vector<float> f((size_t)n * g4d::PROP_FEATS);
    std::vector<uint32_t> y((size_t)n);
    std::vector<int32_t> c(candidates ? (size_t)n * g4d::PROP_SLOTS : 0);
    hipStream_t s = ctx->stream;
    HCK(hipMemcpyAsync(f.data(), ctx->pfeat.ptr, sizeof(float) * f.size(), hipMemcpyDeviceToHost, s));
    HCK(hipMemcpyAsync(y.data(), ctx->plab.ptr, sizeof(uint32_t) * y.size(), hipMemcpyDeviceToHost, s));
    if (candidates) HCK(hipMemcpyAsync(c.data(), ctx->pcand.ptr, sizeof(int32_t) * c.size(), hipMemcpyDeviceToHost, s));
    HCK(hipStreamSynchronize(s));
    int64_t k = 0;
    for (int64_t i = offset; i < n && k < cap; i += stride, k++) {
        memcpy(features + k * g4d::PROP_FEATS, &f[(size_t)i * g4d::PROP_FEATS], sizeof(float) * g4d::PROP_FEATS);
        labels[k] = y[(size_t)i];
        if (candidates) memcpy(candidates + k * g4d::PROP_SLOTS, &c[(size_t)i * g4d::PROP_SLOTS], sizeof(int32_t) * g4d::PROP_SLOTS);
    }
    *count = k;
    return G4D_OK;
}

int g4d_mlp_forward(g4d_ctx *ctx, const float *params, const float *features, int64_t n, float *logits)
{
    if (!ctx || !params || !features || !logits || n < 0) return G4D_ERR_INVALID_INPUT;
    if (n == 0) return G4D_OK;
    HCK(hipSetDevice(ctx->opt.device));
    hipStream_t s = ctx->stream;
    DevBuf dp, df, dl;
    HCK(dp.ensure(sizeof(float) * g4d::MLP_NPARAM));
    HCK(df.ensure(sizeof(float) * g4d::PROP_FEATS * (size_t)n));
    HCK(dl.ensure(sizeof(float) * g4d::MLP_OUT * (size_t)n));
    HCK(hipMemcpyAsync(dp.ptr, params, sizeof(float) * g4d::MLP_NPARAM, hipMemcpyHostToDevice, s));
    HCK(hipMemcpyAsync(df.ptr, features, sizeof(float) * g4d::PROP_FEATS * (size_t)n, hipMemcpyHostToDevice, s));
    HCK(g4d::launch_mlp_forward(buf<float>(dp), buf<float>(df), n, buf<float>(dl), s));
    HCK(hipMemcpyAsync(logits, dl.ptr, sizeof(float) * g4d::MLP_OUT * (size_t)n, hipMemcpyDeviceToHost, s));
    HCK(hipStreamSynchronize(s));
    dp.release();
    df.release();
    dl.release();
    return G4D_OK;
}

int g4d_mlp_grad(g4d_ctx *ctx, const float *params, const float *features, const uint32_t *labels, int64_t n,
                 float *grad, double *loss)
{
    if (!ctx || !params || !features || !labels || !grad || !loss || n < 0) return G4D_ERR_INVALID_INPUT;
    HCK(hipSetDevice(ctx->opt.device));
    hipStream_t s = ctx->stream;
    DevBuf dp, df, dy, dg, dloss;
    HCK(dp.ensure(sizeof(float) * g4d::MLP_NPARAM));
    HCK(dg.ensure(sizeof(float) * g4d::MLP_NPARAM));
    HCK(dloss.ensure(sizeof(double)));
    HCK(df.ensure(sizeof(float) * g4d::PROP_FEATS * (size_t)std::max<int64_t>(n, 1)));
    HCK(dy.ensure(sizeof(uint32_t) * (size_t)std::max<int64_t>(n, 1)));
    HCK(hipMemcpyAsync(dp.ptr, params, sizeof(float) * g4d::MLP_NPARAM, hipMemcpyHostToDevice, s));
    if (n > 0) {
        HCK(hipMemcpyAsync(df.ptr, features, sizeof(float) * g4d::PROP_FEATS * (size_t)n, hipMemcpyHostToDevice, s));
        HCK(hipMemcpyAsync(dy.ptr, labels, sizeof(uint32_t) * (size_t)n, hipMemcpyHostToDevice, s));
    }
    HCK(g4d::launch_fill_f32(buf<float>(dg), g4d::MLP_NPARAM, 0.f, s));
    HCK(hipMemsetAsync(dloss.ptr, 0, sizeof(double), s));
    HCK(g4d::launch_mlp_grad(buf<float>(dp), buf<float>(df), buf<uint32_t>(dy), nullptr, n, buf<float>(dg),
                             buf<double>(dloss), s));
    HCK(hipMemcpyAsync(grad, dg.ptr, sizeof(float) * g4d::MLP_NPARAM, hipMemcpyDeviceToHost, s));
    HCK(hipMemcpyAsync(loss, dloss.ptr, sizeof(double), hipMemcpyDeviceToHost, s));
    HCK(hipStreamSynchronize(s));
    dp.release();
    df.release();
    dy.release();
    dg.release();
    dloss.release();
    return G4D_OK;
}

int g4d_mlp_train(g4d_ctx *ctx, float *params, const float *features, const uint32_t *labels, int64_t n,
                  int32_t epochs, int32_t batch, float lr, uint64_t seed, double *epoch_loss)
{
    if (!ctx || !params || !features || !labels || n <= 0 || epochs < 1 || batch < 32 || !(lr > 0.f))
        return G4D_ERR_INVALID_INPUT;
    if (n > INT32_MAX) return set_err(ctx, G4D_ERR_INVALID_INPUT, "at most 2^31 - 1 samples");
    HCK(hipSetDevice(ctx->opt.device));
    hipStream_t s = ctx->stream;
    DevBuf dp, dm, dv, dg, dloss, df, dy, didx;
    const size_t np = g4d::MLP_NPARAM;
    HCK(dp.ensure(sizeof(float) * np));
    HCK(dm.ensure(sizeof(float) * np));
    HCK(dv.ensure(sizeof(float) * np));
    HCK(dg.ensure(sizeof(float) * np));
    HCK(dloss.ensure(sizeof(double)));
    HCK(df.ensure(sizeof(float) * g4d::PROP_FEATS * (size_t)n));
    HCK(dy.ensure(sizeof(uint32_t) * (size_t)n));
    HCK(didx.ensure(sizeof(int) * (size_t)n));
    HCK(hipMemcpyAsync(dp.ptr, params, sizeof(float) * np, hipMemcpyHostToDevice, s));
    HCK(hipMemcpyAsync(df.ptr, features, sizeof(float) * g4d::PROP_FEATS * (size_t)n, hipMemcpyHostToDevice, s));
    HCK(hipMemcpyAsync(dy.ptr, labels, sizeof(uint32_t) * (size_t)n, hipMemcpyHostToDevice, s));
    HCK(g4d::launch_fill_f32(buf<float>(dm), (int64_t)np, 0.f, s));
    HCK(g4d::launch_fill_f32(buf<float>(dv), (int64_t)np, 0.f, s));
    HCK(g4d::launch_fill_f32(buf<float>(dg), (int64_t)np, 0.f, s));
    std::vector<int> order((size_t)n);
    for (int64_t i = 0; i < n; i++) order[(size_t)i] = (int)i;
    uint64_t rs = seed * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull;
    int step = 0;
    for (int e = 0; e < epochs; e++) {
        for (int64_t i = n - 1; i > 0; i--) { // Fisher-Yates with xorshift64*
            rs ^= rs >> 12;
            rs ^= rs << 25;
            rs ^= rs >> 27;
            const uint64_t r = (rs * 0x2545F4914F6CDD1Dull) % (uint64_t)(i + 1);
            std::swap(order[(size_t)i], order[(size_t)r]);
        }
        HCK(hipMemcpyAsync(didx.ptr, order.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, s));
        HCK(hipMemsetAsync(dloss.ptr, 0, sizeof(double), s));
        for (int64_t b0 = 0; b0 < n; b0 += batch) {
            const int64_t bs = std::min<int64_t>(batch, n - b0);
            HCK(g4d::launch_mlp_grad(buf<float>(dp), buf<float>(df), buf<uint32_t>(dy), buf<int>(didx) + b0, bs,
                                     buf<float>(dg), buf<double>(dloss), s));
            HCK(g4d::launch_adam(buf<float>(dp), buf<float>(dm), buf<float>(dv), buf<float>(dg), 1.f / (float)bs, lr,
                                 ++step, s));
        }
        double tot = 0.0;
        HCK(hipMemcpyAsync(&tot, dloss.ptr, sizeof(double), hipMemcpyDeviceToHost, s));
        HCK(hipStreamSynchronize(s));
        if (epoch_loss) epoch_loss[e] = tot / (double)n;
    }
    HCK(hipMemcpyAsync(params, dp.ptr, sizeof(float) * np, hipMemcpyDeviceToHost, s));
    HCK(hipStreamSynchronize(s));
    DevBuf *all[] = {&dp, &dm, &dv, &dg, &dloss, &df, &dy, &didx};
    for (DevBuf *b : all) b->release();
    return G4D_OK;
}

int g4d_get_stats(const g4d_ctx *ctx, g4d_stats *st)
{
    if (!ctx || !st) return G4D_ERR_INVALID_INPUT;
    *st = ctx->st;
    return G4D_OK;
}

// ---- diagnostics ----

int g4d_diag_orient4(g4d_ctx *ctx, const double *pts5, const int32_t *idx5, int64_t T, int32_t mode, int32_t *sign_out,
                     int32_t *depth_out)
{
    if (!ctx || !pts5 || !idx5 || !sign_out || T < 0) return G4D_ERR_INVALID_INPUT;
    if (T == 0) return G4D_OK;
    HCK(hipSetDevice(ctx->opt.device));
    DevBuf dp, di, ds, dd;
    HCK(dp.ensure(sizeof(double) * 15 * (size_t)T));
    HCK(di.ensure(sizeof(int) * 5 * (size_t)T));
    HCK(ds.ensure(sizeof(int) * (size_t)T));
    HCK(dd.ensure(sizeof(int) * (size_t)T));
    HCK(ctx->cnt.ensure(sizeof(unsigned long long) * G4D_CNT_NUM));
    HCK(hipMemsetAsync(ctx->cnt.ptr, 0, sizeof(unsigned long long) * G4D_CNT_NUM, ctx->stream));
    HCK(hipMemcpyAsync(dp.ptr, pts5, sizeof(double) * 15 * (size_t)T, hipMemcpyHostToDevice, ctx->stream));
    HCK(hipMemcpyAsync(di.ptr, idx5, sizeof(int) * 5 * (size_t)T, hipMemcpyHostToDevice, ctx->stream));
    HCK(g4d::launch_diag_orient4(buf<double>(dp), buf<int>(di), T, mode, buf<int>(ds), buf<int>(dd),
                                 buf<unsigned long long>(ctx->cnt), ctx->stream));
    HCK(hipMemcpyAsync(sign_out, ds.ptr, sizeof(int) * (size_t)T, hipMemcpyDeviceToHost, ctx->stream));
    if (depth_out) HCK(hipMemcpyAsync(depth_out, dd.ptr, sizeof(int) * (size_t)T, hipMemcpyDeviceToHost, ctx->stream));
    unsigned long long hc[G4D_CNT_NUM];
    CK(read_counters(ctx, hc));
    dp.release(); di.release(); ds.release(); dd.release();
    return check_dev_error(ctx, hc[G4D_CNT_ERROR], "diag_orient4");
}

int g4d_diag_orient3d(g4d_ctx *ctx, const double *pts4, int64_t T, int32_t *sign_out)
{
    if (!ctx || !pts4 || !sign_out || T < 0) return G4D_ERR_INVALID_INPUT;
    if (T == 0) return G4D_OK;
    HCK(hipSetDevice(ctx->opt.device));
    DevBuf dp, ds;
    HCK(dp.ensure(sizeof(double) * 12 * (size_t)T));
    HCK(ds.ensure(sizeof(int) * (size_t)T));
    HCK(ctx->cnt.ensure(sizeof(unsigned long long) * G4D_CNT_NUM));
    HCK(hipMemsetAsync(ctx->cnt.ptr, 0, sizeof(unsigned long long) * G4D_CNT_NUM, ctx->stream));
    HCK(hipMemcpyAsync(dp.ptr, pts4, sizeof(double) * 12 * (size_t)T, hipMemcpyHostToDevice, ctx->stream));
    HCK(g4d::launch_diag_orient3d(buf<double>(dp), T, buf<int>(ds), buf<unsigned long long>(ctx->cnt), ctx->stream));
    HCK(hipMemcpyAsync(sign_out, ds.ptr, sizeof(int) * (size_t)T, hipMemcpyDeviceToHost, ctx->stream));
    unsigned long long hc[G4D_CNT_NUM];
    CK(read_counters(ctx, hc));
    dp.release(); ds.release();
    return check_dev_error(ctx, hc[G4D_CNT_ERROR], "diag_orient3d");
}

int g4d_diag_morton(g4d_ctx *ctx, const double *xyz, int32_t n, int32_t grid, int32_t *perm_out, int32_t *vox_out)
{
    if (!ctx || !xyz || n <= 0 || !perm_out || !vox_out || grid < 4 || grid > G4D_MAX_GRID) return G4D_ERR_INVALID_INPUT;
    HCK(hipSetDevice(ctx->opt.device));
    hipStream_t s = ctx->stream;
    const double *d_xyz = nullptr;
    CK(g4d_upload_points(ctx, xyz, n, &d_xyz));
    HCK(ctx->keys.ensure(sizeof(uint64_t) * (size_t)n));
    HCK(ctx->pts.ensure(sizeof(double) * 3 * (size_t)n));
    HCK(ctx->perm.ensure(sizeof(int) * (size_t)n));
    HCK(ctx->vox.ensure(sizeof(uint64_t) * (size_t)n));
    g4d::GridSets GS;
    SetTotals tot;
    CK(setup_sets(ctx, single_set(n, grid), n, GS, tot));
    HCK(g4d::launch_bounds(d_xyz, n, const_cast<double *>(GS.bounds), s));
    HCK(g4d::launch_morton_keys(d_xyz, n, GS, buf<uint64_t>(ctx->keys), s));
    HCK(g4d::radix_sort_u64(ctx->P, buf<uint64_t>(ctx->keys), n, GS.key_shift, 64, s));
    HCK(g4d::launch_gather_sorted(d_xyz, buf<uint64_t>(ctx->keys), n, GS, buf<double>(ctx->pts), buf<int>(ctx->perm),
                                  buf<uint64_t>(ctx->vox), s));
    std::vector<uint64_t> vs((size_t)n);
    HCK(hipMemcpyAsync(perm_out, ctx->perm.ptr, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, s));
    HCK(hipMemcpyAsync(vs.data(), ctx->vox.ptr, sizeof(uint64_t) * (size_t)n, hipMemcpyDeviceToHost, s));
    HCK(hipStreamSynchronize(s));
    for (int k = 0; k < n; k++) {
        const int i = perm_out[k];
        const uint64_t c = vs[(size_t)k];
        vox_out[3 * (int64_t)i] = G4D_VOX_X(c);
        vox_out[3 * (int64_t)i + 1] = G4D_VOX_Y(c);
        vox_out[3 * (int64_t)i + 2] = G4D_VOX_Z(c);
    }
    return G4D_OK;
}

int g4d_diag_pba(g4d_ctx *ctx, const int32_t *vox_sorted, int32_t n, int32_t grid, int32_t codes, int32_t *out)
{
    if (!ctx || !vox_sorted || n <= 0 || !out || grid < 4 || grid > G4D_MAX_GRID) return G4D_ERR_INVALID_INPUT;
    if (codes < 0 || codes > 2) return set_err(ctx, G4D_ERR_INVALID_INPUT, "codes %d outside [0, 2]", codes);
    const bool id_mode = grid > G4D_CODE_GRID || codes == 2;
    if (grid > G4D_CODE_GRID && codes == 1)
        return set_err(ctx, G4D_ERR_INVALID_INPUT, "site codes exist only for grids up to %d", G4D_CODE_GRID);
    HCK(hipSetDevice(ctx->opt.device));
    hipStream_t s = ctx->stream;
    std::vector<uint64_t> enc((size_t)n);
    for (int k = 0; k < n; k++) {
        const int x = vox_sorted[3 * (int64_t)k], y = vox_sorted[3 * (int64_t)k + 1], z = vox_sorted[3 * (int64_t)k + 2];
        if (x < 0 || y < 0 || z < 0 || x >= grid || y >= grid || z >= grid)
            return set_err(ctx, G4D_ERR_INVALID_INPUT, "voxel outside the grid");
        enc[(size_t)k] = G4D_VOX(x, y, z);
    }
    const int64_t G3 = (int64_t)grid * grid * grid;
    HCK(ctx->vox.ensure(sizeof(uint64_t) * (size_t)n));
    HCK(ctx->grid.ensure(sizeof(int) * (size_t)G3));
    HCK(ctx->gtmp.ensure(sizeof(int) * (size_t)G3));
    if (!id_mode) HCK(ctx->idgrid.ensure(sizeof(int) * (size_t)G3));
    HCK(hipMemcpyAsync(ctx->vox.ptr, enc.data(), sizeof(uint64_t) * (size_t)n, hipMemcpyHostToDevice, s));
    g4d::GridSets GS;
    SetTotals tot;
    CK(setup_sets(ctx, single_set(n, grid), n, GS, tot));
    const uint64_t *vox = buf<uint64_t>(ctx->vox), *pvox = id_mode ? vox : nullptr;
    int *idgrid = id_mode ? nullptr : buf<int>(ctx->idgrid);
    HCK(g4d::launch_fill_i32(buf<int>(ctx->grid), G3, G4D_MARKER, s));
    HCK(g4d::launch_scatter_sites(vox, n, GS, buf<int>(ctx->grid), idgrid, s));
    CK(run_pba(ctx, GS, tot, grid, buf<int>(ctx->grid), buf<int>(ctx->gtmp), codes == 1 ? nullptr : idgrid, pvox,
               false));
    HCK(hipMemcpyAsync(out, ctx->gtmp.ptr, sizeof(int) * (size_t)G3, hipMemcpyDeviceToHost, s));
    HCK(hipStreamSynchronize(s));
    return G4D_OK;
}

} // extern "C"

// Validation of a device-resident mesh; see g4d_check_report.
static int check_device(g4d_ctx *ctx, const double *pts, int n, const int *tets, const int *opp, int64_t nt,
                        bool full, g4d_check_report *rep)
{
    hipStream_t s = ctx->stream;
    g4d::Prims &P = ctx->P;
    memset(rep, 0, sizeof(*rep));
    rep->num_tets = nt;
    rep->empty_sphere_failures = -1;
    if (nt <= 0) return G4D_OK;
    CK(guard_tets(ctx, nt));
    CK(guard_keys(ctx, 6 * nt, "validator edges"));
    HCK(ctx->cnt.ensure(sizeof(unsigned long long) * G4D_CNT_NUM));
    unsigned long long *cnt = buf<unsigned long long>(ctx->cnt);
    HCK(hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * G4D_CNT_NUM, s));
    DevBuf acc_b, used_b, ref_b, table_b, keys_b;
    HCK(acc_b.ensure(sizeof(unsigned long long) * g4d::CHECK_NUM));
    unsigned long long *acc = buf<unsigned long long>(acc_b);
    HCK(hipMemsetAsync(acc, 0, sizeof(unsigned long long) * g4d::CHECK_NUM, s));
    // independent face matching
    const uint64_t tsz = next_pow2((uint64_t)std::max<int64_t>(8 * nt, 1024));
    HCK(table_b.ensure(sizeof(uint32_t) * (size_t)tsz));
    HCK(ref_b.ensure(sizeof(int) * 4 * (size_t)nt));
    HCK(g4d::launch_adjacency(tets, nt, buf<int>(ref_b), buf<uint32_t>(table_b), (int64_t)tsz, cnt, s));
    HCK(used_b.ensure((size_t)n));
    HCK(hipMemsetAsync(used_b.ptr, 0, (size_t)n, s));
    HCK(g4d::launch_check_tets(pts, tets, opp, buf<int>(ref_b), nt, n, buf<unsigned char>(used_b), acc, cnt, s));
    HCK(g4d::launch_check_count_used(buf<unsigned char>(used_b), n, acc, s));
    // distinct edges
    HCK(keys_b.ensure(sizeof(uint64_t) * 6 * (size_t)nt));
    uint64_t *keys = buf<uint64_t>(keys_b);
    HCK(g4d::launch_check_edges(tets, nt, keys, s));
    const int nb = bits_for(n);
    int64_t ne = 6 * nt;
    HCK(sort_unique_pairs(P, keys, &ne, nb, n, s));
    unsigned long long h[g4d::CHECK_NUM];
    if (full) {
        // cell grid over the bounding box
        HCK(ctx->bounds.ensure(sizeof(double) * 768));
        HCK(g4d::launch_bounds(pts, n, buf<double>(ctx->bounds), s));
        double hb[2];
        HCK(hipMemcpyAsync(hb, ctx->bounds.ptr, sizeof(hb), hipMemcpyDeviceToHost, s));
        HCK(hipStreamSynchronize(s));
        const double box[6] = {hb[0], hb[0], hb[0], hb[1], hb[1], hb[1]};
        int R = (int)std::cbrt((double)n / 2.0);
        R = std::max(1, std::min(R, 256));
        const int64_t ncell = (int64_t)R * R * R;
        DevBuf box_b, cid_b, ccnt_b, start_b, fill_b, items_b;
        HCK(box_b.ensure(sizeof(box)));
        HCK(hipMemcpyAsync(box_b.ptr, box, sizeof(box), hipMemcpyHostToDevice, s));
        HCK(cid_b.ensure(sizeof(int) * (size_t)n));
        HCK(ccnt_b.ensure(sizeof(int64_t) * (size_t)ncell));
        HCK(start_b.ensure(sizeof(int64_t) * (size_t)(ncell + 1)));
        HCK(fill_b.ensure(sizeof(int) * (size_t)ncell));
        HCK(items_b.ensure(sizeof(int) * (size_t)n));
        HCK(hipMemsetAsync(ccnt_b.ptr, 0, sizeof(int64_t) * (size_t)ncell, s));
        HCK(hipMemsetAsync(fill_b.ptr, 0, sizeof(int) * (size_t)ncell, s));
        HCK(g4d::launch_check_cells(pts, n, buf<double>(box_b), R, buf<int>(cid_b), buf<unsigned long long>(ccnt_b),
                                    s));
        int64_t *start = buf<int64_t>(start_b);
        HCK(g4d::scan_excl_i64(P, buf<int64_t>(ccnt_b), start, ncell, start + ncell, s));
        HCK(g4d::launch_check_cell_fill(buf<int>(cid_b), n, start, buf<int>(fill_b), buf<int>(items_b), s));
        HCK(g4d::launch_check_empty(pts, tets, nt, buf<double>(box_b), R, start, buf<int>(items_b), acc, cnt, s));
        HCK(hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, s));
        HCK(hipStreamSynchronize(s));
        box_b.release();
        cid_b.release();
        ccnt_b.release();
        start_b.release();
        fill_b.release();
        items_b.release();
    } else {
        HCK(hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, s));
        HCK(hipStreamSynchronize(s));
    }
    unsigned long long hc[G4D_CNT_NUM];
    CK(read_counters(ctx, hc));
    acc_b.release();
    used_b.release();
    ref_b.release();
    table_b.release();
    keys_b.release();
    const unsigned long long other = hc[G4D_CNT_ERROR] & ~(unsigned long long)G4D_DEV_ERR_FACE;
    if (other) CK(check_dev_error(ctx, other, "check"));
    rep->num_vertices = (int64_t)h[g4d::CHECK_VERTS];
    rep->num_edges = ne;
    rep->hull_faces = (int64_t)h[g4d::CHECK_HULL];
    rep->num_faces = rep->hull_faces + (int64_t)h[g4d::CHECK_SHARED] / 2;
    rep->euler = rep->num_vertices - rep->num_edges + rep->num_faces - nt;
    rep->orientation_failures = (int64_t)h[g4d::CHECK_ORIENT];
    rep->flat_tets = (int64_t)h[g4d::CHECK_FLAT];
    rep->insphere_failures = (int64_t)h[g4d::CHECK_INSPHERE];
    rep->adjacency_failures = (int64_t)h[g4d::CHECK_ADJ] + ((hc[G4D_CNT_ERROR] & G4D_DEV_ERR_FACE) ? 1 : 0);
    if (full) {
        rep->empty_sphere_failures = (int64_t)h[g4d::CHECK_EMPTY];
        rep->empty_sphere_tests = (int64_t)h[g4d::CHECK_EMPTY_TESTS];
    }
    return G4D_OK;
}

extern "C" {

int g4d_check_result(g4d_ctx *ctx, int32_t full_empty_sphere, g4d_check_report *rep)
{
    if (!ctx || !rep || !ctx->have_result) return G4D_ERR_INVALID_INPUT;
    if (ctx->last_nsets > 1)
        return set_err(ctx, G4D_ERR_INVALID_INPUT, "the last result is a batch; check each set with g4d_check_mesh");
    HCK(hipSetDevice(ctx->opt.device));
    const double *pts = ctx->opt.index_space == G4D_INDEX_MORTON ? buf<double>(ctx->pts) : ctx->last_xyz;
    const int *opp = ctx->opt.compute_adjacency ? buf<int>(ctx->opp) : nullptr;
    return check_device(ctx, pts, ctx->last_n, buf<int>(ctx->tets), opp, ctx->num_tets, full_empty_sphere != 0, rep);
}

int g4d_check_mesh(g4d_ctx *ctx, const double *xyz, int32_t n, const int32_t *tets, const int32_t *opp,
                   int64_t num_tets, int32_t full_empty_sphere, g4d_check_report *rep)
{
    if (!ctx || !xyz || n <= 0 || !rep || num_tets < 0 || (num_tets > 0 && !tets)) return G4D_ERR_INVALID_INPUT;
    HCK(hipSetDevice(ctx->opt.device));
    hipStream_t s = ctx->stream;
    DevBuf dp, dt, dop;
    HCK(dp.ensure(sizeof(double) * 3 * (size_t)n));
    HCK(hipMemcpyAsync(dp.ptr, xyz, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, s));
    if (num_tets > 0) {
        HCK(dt.ensure(sizeof(int) * 4 * (size_t)num_tets));
        HCK(hipMemcpyAsync(dt.ptr, tets, sizeof(int) * 4 * (size_t)num_tets, hipMemcpyHostToDevice, s));
        if (opp) {
            HCK(dop.ensure(sizeof(int) * 4 * (size_t)num_tets));
            HCK(hipMemcpyAsync(dop.ptr, opp, sizeof(int) * 4 * (size_t)num_tets, hipMemcpyHostToDevice, s));
        }
    }
    HCK(hipStreamSynchronize(s));
    const int r = check_device(ctx, buf<double>(dp), n, buf<int>(dt), opp ? buf<int>(dop) : nullptr, num_tets,
                               full_empty_sphere != 0, rep);
    dp.release();
    dt.release();
    dop.release();
    return r;
}

static g4d::OutParams star_params(g4d_ctx *ctx)
{
    Store &store = ctx->store;
    g4d::OutParams op;
    memset(&op, 0, sizeof(op));
    op.n = ctx->last_n;
    op.vnum = buf<int>(store.vnum);
    op.voff = buf<int64_t>(store.voff);
    op.vpool = buf<int>(store.arena[0]);
    op.tnum = buf<int>(store.tnum);
    op.toff = buf<int64_t>(store.toff);
    op.tpool = buf<uint64_t>(store.arena[1]);
    op.cnt = buf<unsigned long long>(ctx->cnt);
    return op;
}

int g4d_check_stars(g4d_ctx *ctx, int64_t *inconsistent, int64_t *triangles)
{
    if (!ctx || !inconsistent || !ctx->have_result) return G4D_ERR_INVALID_INPUT;
    HCK(hipSetDevice(ctx->opt.device));
    hipStream_t s = ctx->stream;
    const g4d::OutParams op = star_params(ctx);
    HCK(g4d::launch_zero_counters(op.cnt, (1u << G4D_CNT_APPEND4) | (1u << G4D_CNT_APPEND5), s));
    HCK(g4d::launch_star_consistency(op, s));
    unsigned long long hc[G4D_CNT_NUM];
    CK(read_counters(ctx, hc));
    *inconsistent = (int64_t)hc[G4D_CNT_APPEND5];
    if (triangles) *triangles = (int64_t)hc[G4D_CNT_APPEND4];
    return G4D_OK;
}

int g4d_diag_set_link_vertex(g4d_ctx *ctx, int32_t star, int32_t local, int32_t v)
{
    if (!ctx || !ctx->have_result || star < 0 || star >= ctx->last_n) return G4D_ERR_INVALID_INPUT;
    HCK(hipSetDevice(ctx->opt.device));
    Store &store = ctx->store;
    int vn = 0;
    int64_t vo = 0;
    HCK(hipMemcpy(&vn, buf<int>(store.vnum) + star, sizeof(int), hipMemcpyDeviceToHost));
    HCK(hipMemcpy(&vo, buf<int64_t>(store.voff) + star, sizeof(int64_t), hipMemcpyDeviceToHost));
    if (local < 0 || local >= vn) return set_err(ctx, G4D_ERR_INVALID_INPUT, "star %d has %d link vertices", star, vn);
    HCK(hipMemcpy(buf<int>(store.arena[0]) + vo + local, &v, sizeof(int), hipMemcpyHostToDevice));
    return G4D_OK;
}

int g4d_diag_copy(g4d_ctx *ctx, int64_t bytes, double *ms)
{
    if (!ctx || bytes < 16) return G4D_ERR_INVALID_INPUT;
    HCK(hipSetDevice(ctx->opt.device));
    const int64_t n16 = bytes / 16;
    DevBuf a, b;
    HCK(a.ensure((size_t)n16 * 16));
    HCK(b.ensure((size_t)n16 * 16));
    HCK(hipMemsetAsync(a.ptr, 1, (size_t)n16 * 16, ctx->stream));
    if (ctx->ev_ok) HCK(hipEventRecord(ctx->ev[EV_START], ctx->stream));
    HCK(g4d::launch_diag_copy(buf<int4>(a), buf<int4>(b), n16, ctx->stream));
    if (ctx->ev_ok) HCK(hipEventRecord(ctx->ev[EV_MORTON], ctx->stream));
    HCK(hipStreamSynchronize(ctx->stream));
    if (ms) *ms = ev_ms(ctx, EV_START, EV_MORTON);
    a.release();
    b.release();
    return G4D_OK;
}

int g4d_diag_sort_pairs(g4d_ctx *ctx, const uint64_t *keys, int64_t m, int32_t nb, int32_t method, uint64_t *out,
                        double *ms, int32_t *shift)
{
    if (!ctx || m < 0 || (m > 0 && (!keys || !out)) || nb < 1 || nb > 31 || method < 0 || method > 4)
        return G4D_ERR_INVALID_INPUT;
    if (m > G4D_MAX_SORT_KEYS) return set_err(ctx, G4D_ERR_CAPACITY, "%lld keys exceed the sort limit", (long long)m);
    HCK(hipSetDevice(ctx->opt.device));
    hipStream_t s = ctx->stream;
    if (shift)
        *shift = method == 1 ? -1 : g4d::key_sort_shift(m, nb, (method == 2 || method == 4) ? nb + 1 : nb, 0, method >= 3);
    DevBuf d;
    HCK(d.ensure(sizeof(uint64_t) * (size_t)std::max<int64_t>(m, 1)));
    HCK(hipMemcpyAsync(d.ptr, keys, sizeof(uint64_t) * (size_t)m, hipMemcpyHostToDevice, s));
    if (ctx->ev_ok) HCK(hipEventRecord(ctx->ev[EV_START], s));
    if (method == 1) {
        if (m > 1) HCK(g4d::radix_sort_u64_2(ctx->P, buf<uint64_t>(d), m, 0, nb, 32, 32 + nb, s));
    } else if (method == 2 || method == 4) {
        HCK(g4d::sort_keys_ab(ctx->P, buf<uint64_t>(d), m, 33, nb, nb + 1, 0, method == 4, s));
    } else {
        HCK(g4d::sort_keys_ab(ctx->P, buf<uint64_t>(d), m, 32, nb, nb, 0, method == 3, s));
    }
    if (ctx->ev_ok) HCK(hipEventRecord(ctx->ev[EV_MORTON], s));
    HCK(hipMemcpyAsync(out, d.ptr, sizeof(uint64_t) * (size_t)m, hipMemcpyDeviceToHost, s));
    HCK(hipStreamSynchronize(s));
    if (ms) *ms = ctx->ev_ok ? ev_ms(ctx, EV_START, EV_MORTON) : 0.0;
    d.release();
    return G4D_OK;
}

int g4d_diag_grid_pairs(g4d_ctx *ctx, const int32_t *id_grid, int32_t grid, uint64_t *out, int64_t cap, int64_t *count)
{
    if (!ctx || !id_grid || !count || grid < 4 || grid > G4D_MAX_GRID) return G4D_ERR_INVALID_INPUT;
    HCK(hipSetDevice(ctx->opt.device));
    hipStream_t s = ctx->stream;
    const int64_t G3 = (int64_t)grid * grid * grid;
    HCK(ctx->gtmp.ensure(sizeof(int) * (size_t)G3));
    HCK(hipMemcpyAsync(ctx->gtmp.ptr, id_grid, sizeof(int) * (size_t)G3, hipMemcpyHostToDevice, s));
    HCK(ctx->cnt.ensure(sizeof(unsigned long long) * G4D_CNT_NUM));
    unsigned long long *cnt = buf<unsigned long long>(ctx->cnt);
    g4d::GridSets GS;
    SetTotals tot;
    CK(setup_sets(ctx, single_set(1, grid), 1, GS, tot));
    const int64_t nthr = tot.pair_threads;
    HCK(ctx->pcount.ensure(sizeof(int64_t) * (size_t)(nthr + 1)));
    int64_t *pcnt = buf<int64_t>(ctx->pcount);
    HCK(g4d::launch_grid_pairs(buf<int>(ctx->gtmp), GS, nthr, pcnt, nullptr, false, s));
    HCK(g4d::scan_excl_i64(ctx->P, pcnt, pcnt, nthr, pcnt + nthr, s));
    int64_t m = 0;
    HCK(hipMemcpyAsync(&m, pcnt + nthr, 8, hipMemcpyDeviceToHost, s));
    HCK(hipStreamSynchronize(s));
    HCK(ctx->pairs.ensure(sizeof(uint64_t) * (size_t)std::max<int64_t>(m, 1)));
    uint64_t *pr = buf<uint64_t>(ctx->pairs);
    HCK(g4d::launch_grid_pairs(buf<int>(ctx->gtmp), GS, nthr, pcnt, pr, true, s));
    int maxid = 0;
    for (int64_t i = 0; i < G3; i++) maxid = std::max(maxid, id_grid[i]);
    const int nb = bits_for((int64_t)maxid + 1);
    HCK(sort_unique_pairs(ctx->P, pr, &m, nb, (int64_t)maxid + 1, s));
    std::vector<uint64_t> h((size_t)std::max<int64_t>(m, 1));
    HCK(hipMemcpyAsync(h.data(), pr, sizeof(uint64_t) * (size_t)m, hipMemcpyDeviceToHost, s));
    HCK(hipStreamSynchronize(s));
    int64_t k = 0;
    for (int64_t i = 0; i < m; i++) {
        if (out && k < cap) out[k] = h[(size_t)i];
        k++;
    }
    *count = k;
    return k > cap && out ? set_err(ctx, G4D_ERR_CAPACITY, "output capacity %lld < %lld", (long long)cap, (long long)k)
                          : G4D_OK;
}

} // extern "C"
