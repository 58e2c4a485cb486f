// gStar4D for MI355X — 4D stars on the lifted points and parallel star splaying.
//
// Reference behaviour restated (GD/ = GDelaunay/):
//   initial cones            GD/GDelaunay/GDelPredKernels.cu:45-231
//   beneath marking          GD/GDelaunay/GDelPredKernels.cu:233-420
//   hole stitching           GD/GDelaunay/GDelKernels.cu:518-783
//   drowned insertions       GD/GDelaunay/GDelKernels.cu:1459-1527, GD/GDelaunay/GDelHost.cu:817-875
//   containment proofs       GD/GDelaunay/GDelPredKernels.cu:427-626
//   triangle insertions      GD/GDelaunay/GDelKernels.cu:818-939
//   history                  GD/GDelaunay/GDelKernels.cu:984-1068,1625-1677
//   tetra output             GD/GDelaunay/GDelKernels.cu:1097-1191, GD/GDelaunay/GDelPredKernels.cu:628-692
//
// MI355X design.  One wave64 owns one star at a time (dynamic work queue).  The star's link
// (a triangulated 2-sphere) lives in LDS: local vertex table with coordinates relative to the
// star point, triangles with 9-bit local vertices and 16-bit adjacency.  Each lane owns the
// triangle slots lane, lane+64, ...; for each slot LDS caches the four cofactors of the
// lifted 4x4 determinant, so the beneath test of a new vertex v against triangle
// (a,b,c) is one 4-term dot product:
//     orient4(a,b,c,s,v) = v'x*Mx - v'y*My + v'z*Mz - v'l*Ml      (primes: relative to s)
// with a forward-error bound from the cofactor permanents; uncertain lanes fall back to the
// exact integer predicate with SoS.  The hole boundary after freeing beneath triangles is
// found in parallel (edges of freed triangles whose neighbour survives), slots are assigned
// with a wave prefix sum over a free-slot stack, and the fan's adjacency is closed through
// two per-vertex maps (boundary edge starting / ending at each vertex).  The whole insertion
// list of a star runs inside one kernel, so there is no host loop per insertion index.
#include "g4d_predicates.h"
#include <algorithm>
#include <climits>
#include <cstdlib>
#include <type_traits>

#include "g4d_star.h"

namespace g4d {

static constexpr uint32_t TF_ALIVE = 1u << 27;
static constexpr uint32_t TF_CHECKED = 1u << 28;
static constexpr uint32_t TF_VIS = 1u << 29;
static constexpr uint32_t TF_FRESH = 1u << 30;
static constexpr uint32_t TF_UNSURE = 1u << 26; // global-memory kernel only (its flags word holds no vertices)

__device__ __forceinline__ int tvert(uint32_t t, int i) { return (int)((t >> (9 * i)) & 511u); }
__device__ __forceinline__ uint32_t tpack(int a, int b, int c, uint32_t fl)
{
    return (uint32_t)a | ((uint32_t)b << 9) | ((uint32_t)c << 18) | fl;
}

// HBM triangle: bits 0..26 vertices, 27..56 adjacency (10 bits each), 57 checked.
__device__ __forceinline__ uint64_t hpack(int a, int b, int c, int o0, int o1, int o2, bool checked)
{
    return (uint64_t)a | ((uint64_t)b << 9) | ((uint64_t)c << 18) | ((uint64_t)o0 << 27) | ((uint64_t)o1 << 37) |
           ((uint64_t)o2 << 47) | ((uint64_t)(checked ? 1 : 0) << 57);
}
__device__ __forceinline__ int hvert(uint64_t t, int i) { return (int)((t >> (9 * i)) & 511u); }
__device__ __forceinline__ int hopp(uint64_t t, int i) { return (int)((t >> (27 + 10 * i)) & 1023u); }
__device__ __forceinline__ bool hchecked(uint64_t t) { return (t >> 57) & 1u; }

// Wide format (stars with >= STAR_WIDE_MIN link vertices): two words per triangle.
static constexpr uint64_t W21 = (1ull << 21) - 1;
__device__ __forceinline__ uint64_t wpack3(int a, int b, int c)
{
    return (uint64_t)a | ((uint64_t)b << 21) | ((uint64_t)c << 42);
}
__device__ __forceinline__ int wfield(uint64_t w, int i) { return (int)((w >> (21 * i)) & W21); }

// Local vertex indices of triangle t of a star stored at tpool + to.
__device__ __forceinline__ void star_tri_verts(const uint64_t *tpool, int64_t to, int t, bool wide, int *v)
{
    if (wide) {
        const uint64_t w = tpool[to + 2 * (int64_t)t];
        v[0] = wfield(w, 0);
        v[1] = wfield(w, 1);
        v[2] = wfield(w, 2);
    } else {
        const uint64_t h = tpool[to + t];
        v[0] = hvert(h, 0);
        v[1] = hvert(h, 1);
        v[2] = hvert(h, 2);
    }
}

template <int VMAX>
struct StarLds {
    static constexpr int TMAX = 2 * VMAX;
    // per triangle slot: cofactors C[0..3] and permanents P[0..3], stored component-major so
    // that the lanes of a wave, each reading its own slot, touch consecutive 8-byte words
    double cof[8][TMAX];
    int vid[VMAX];
    double vc[VMAX][4];
    uint32_t tv[TMAX];
    union {
        uint16_t to[TMAX][4]; // neighbours across the edges opposite vertices 0-2 (fourth unused)
        uint64_t to64[TMAX];  // the same rows read as one word each
    };
    int16_t vstart[VMAX];
    int16_t vend[VMAX];
    uint16_t freel[TMAX];
    uint16_t be_a[TMAX], be_b[TMAX], be_u[TMAX], be_slot[TMAX];
    uint8_t be_j[TMAX];
    int ins[VMAX];
    int16_t ins_local[VMAX];
    uint8_t present[VMAX];
    uint16_t remap[TMAX];
    uint16_t vremap[VMAX];
    int sc[4];
};

// Cofactors and permanents of the triangle (r0, r1, r2) into the component-major slot table.
template <int TMAX>
__device__ __forceinline__ void lds_cofactors(const double *r0, const double *r1, const double *r2,
                                              double (*cof)[TMAX], int slot)
{
    double C[4], P[4];
    tri_cofactors(r0, r1, r2, C, P);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        cof[j][slot] = C[j];
        cof[4 + j][slot] = P[j];
    }
}

// Exclusive scan across the wave of per-lane counts below 2^BITS; returns the prefix, the total
// in *tot.  One ballot and a lane-masked bit count per bit: no cross-lane data movement.
template <int BITS>
__device__ __forceinline__ int wave_excl(int v, int *tot)
{
    int pre = 0, all = 0;
#pragma unroll
    for (int b = 0; b < BITS; b++) {
        const uint64_t m = __ballot((v >> b) & 1);
        pre += (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u)) << b;
        all += __popcll(m) << b;
    }
    *tot = all;
    return pre;
}

// Bits needed for counts up to x.
__host__ __device__ constexpr int count_bits(int x) { return x < 2 ? 1 : 1 + count_bits(x >> 1); }

// Insertion pairs of a link triangle with global vertices (a, b, c) (GD/GDelaunay/GDelKernels.cu:
// 818-939): when it is unchecked, each cyclic edge (u, w) with u < w gives (u<<32 | w); there
// are one or two such edges, the unused entry is STAR_PAIR_NONE.
struct TriPairs {
    uint64_t x, y;
};
__device__ __forceinline__ TriPairs star_tri_pairs(int a, int b, int c, bool unchecked)
{
    uint64_t e0 = STAR_PAIR_NONE, e1 = STAR_PAIR_NONE;
    if (unchecked) {
        const int u[3] = {a, b, c};
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int x = u[i], y = u[(i + 1) % 3];
            if (x < y) {
                const uint64_t k = ((uint64_t)(uint32_t)x << 32) | (uint32_t)y;
                if (e0 == STAR_PAIR_NONE) e0 = k;
                else e1 = k;
            }
        }
    }
    return TriPairs{e0, e1};
}

// The LDS star kernels run one wave per work-group.  A wave's LDS operations complete in
// order, so ordering its LDS traffic needs only a compiler-level barrier at wavefront scope;
// unlike __syncthreads() this does not wait for the wave's outstanding global stores.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Appends star s to a list (lane 0), flagging a buffer error when the list is full.
__device__ __forceinline__ void star_list_push(int *list, int64_t cap, unsigned long long *counter, int s,
                                               unsigned long long *cnt)
{
    if (__lane_id() == 0) {
        const unsigned long long k = atomicAdd(counter, 1ull);
        if ((int64_t)k < cap) list[k] = s;
        else atomicOr(&cnt[G4D_CNT_ERROR], (unsigned long long)G4D_DEV_ERR_BUFFER);
    }
}

// New history of star s: its old history (mode 1) merged with this pass's insertions ins[0, m)
// (sorted, disjoint from the old list) into the star's new region.  Returns the old length.
__device__ __forceinline__ int star_history_merge(const StarParams &p, int s, const int *ins, int m, int tid, int nt)
{
    const int64_t ho = p.nhoff[s];
    if (p.mode == 0) {
        for (int i = tid; i < m; i += nt) p.hpool[ho + i] = ins[i];
        return 0;
    }
    const int64_t hb = p.hoff[s];
    const int hn = p.hnum[s];
    const int *old = p.hpool + hb;
    for (int i = tid; i < hn; i += nt) {
        const int x = old[i];
        int lo = 0, hi = m;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (ins[mid] < x) lo = mid + 1;
            else hi = mid;
        }
        p.hpool[ho + i + lo] = x;
    }
    for (int i = tid; i < m; i += nt) {
        const int x = ins[i];
        int lo = 0, hi = hn;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (old[mid] < x) lo = mid + 1;
            else hi = mid;
        }
        p.hpool[ho + i + lo] = x;
    }
    return hn;
}

// Points star s's record at its rebuilt regions (one thread, after every read of the old
// record and regions).
__device__ __forceinline__ void star_commit(const StarParams &p, int s, int vcount, int64_t vo, int tcount, int64_t to,
                                            int hcount)
{
    p.vnum[s] = vcount;
    p.voff[s] = vo;
    p.tnum[s] = tcount;
    p.toff[s] = to;
    p.hnum[s] = hcount;
    p.hoff[s] = p.nhoff[s];
}

// EXACT = false: filters only.  A star with an inconclusive orientation is handed to
// p.exact_list untouched (its output is written only at the end), so the kernel carries no
// multiprecision code and no private scratch.
template <int VMAX, bool EXACT>
__device__ __forceinline__ void star_body(const StarParams &p, int nitems)
{
    constexpr int TMAX = StarLds<VMAX>::TMAX;
    constexpr int KPL = TMAX / 64;
    __shared__ StarLds<VMAX> L;
    const int lane = threadIdx.x;
    unsigned long long *cnt = p.cnt;

    if (p.nlist_dev) {
        const unsigned long long c = *p.nlist_dev;
        nitems = c < (unsigned long long)nitems ? (int)c : nitems;
    }
    // A list no longer than the grid is spread one star per work-group, without the queue;
    // work-groups beyond its length leave at once (thousands of idle work-groups contending for
    // the queue counter cost ~0.3 ms).  Longer lists go through the work queue in chunks: one
    // atomic per STAR_CHUNK stars for the filter-only kernels (many small stars), per star for
    // the exact ones (uneven cost).
    constexpr int CHUNK = EXACT ? 1 : STAR_CHUNK;
    const bool one_each = nitems <= (int)gridDim.x;
    if ((int)blockIdx.x >= nitems) return;
    bool first = true;
    int nfilt = 0, ncof = 0; // this lane's cofactor-filter evaluations and cofactor computations
    for (;;) {
        int base = 0;
        if (one_each) {
            base = first ? (int)blockIdx.x : nitems;
            first = false;
        } else {
            if (lane == 0) base = (int)atomicAdd(&cnt[p.work_counter], (unsigned long long)CHUNK);
            base = __shfl(base, 0, 64);
        }
        if (base >= nitems) break;
        const int bend = one_each ? base + 1 : (base + CHUNK < nitems ? base + CHUNK : nitems);
#pragma unroll 1
        for (int item = base; item < bend; item++) {
            const int s = p.star_list ? p.star_list[item] : item;
            const int64_t ib = p.ins_off[s];
            const int m = (int)(p.ins_off[s + 1] - ib);
            const int vn = p.mode ? p.vnum[s] : 0;
            const int need = vn + m;
            const int lim = VMAX < p.lds_limit ? VMAX : p.lds_limit;
            if (need > lim) {
                star_list_push(p.defer_list, p.defer_cap, &cnt[p.defer_counter], s, cnt);
                continue;
            }
            if (p.mode == 0 && m < 4) {
                if (lane == 0) {
                    atomicOr(&cnt[G4D_CNT_ERROR], (unsigned long long)G4D_DEV_ERR_CONE);
                    p.vnum[s] = 0;
                    p.tnum[s] = 0;
                }
                continue;
            }
            if (p.mode == 1 && m == 0) continue; // untouched star: stays where it is
            const double sx = p.pts[3 * (int64_t)s], sy = p.pts[3 * (int64_t)s + 1], sz = p.pts[3 * (int64_t)s + 2];
            // Insertion q takes link-vertex slot vn + q; every insertion's coordinates are fetched
            // here at once instead of one dependent load per insertion step.  (In mode 0 the
            // first four form the initial simplex, set up below.)
            const int q0 = p.mode ? 0 : 4;
            for (int q = lane; q < m; q += 64) {
                const int g = (int)(uint32_t)p.ins_keys[ib + q];
                L.ins[q] = g;
                L.ins_local[q] = -1;
                if (q >= q0) {
                    L.vid[vn + q] = g;
                    rel_coords(p.pts, g, sx, sy, sz, L.vc[vn + q]);
                }
            }
            int T, nfree = 0;
            if (p.mode == 1) {
                const int tn = p.tnum[s];
                const int64_t vi = p.voff[s], ti = p.toff[s];
                for (int v = lane; v < vn; v += 64) {
                    const int g = p.vpool[vi + v];
                    L.vid[v] = g;
                    rel_coords(p.pts, g, sx, sy, sz, L.vc[v]);
                }
                for (int t = lane; t < tn; t += 64) {
                    const uint64_t h = p.tpool[ti + t];
                    L.tv[t] = tpack(hvert(h, 0), hvert(h, 1), hvert(h, 2), TF_ALIVE | (hchecked(h) ? TF_CHECKED : 0u));
                    L.to[t][0] = (uint16_t)hopp(h, 0);
                    L.to[t][1] = (uint16_t)hopp(h, 1);
                    L.to[t][2] = (uint16_t)hopp(h, 2);
                }
                T = tn;
            } else {
                // initial 4-simplex (GD/GDelaunay/GDelPredKernels.cu:45-158)
                int l[4];
                for (int i = 0; i < 4; i++) l[i] = (int)(uint32_t)p.ins_keys[ib + i];
                const double *pp[5] = {p.pts + 3 * (int64_t)l[0], p.pts + 3 * (int64_t)l[1], p.pts + 3 * (int64_t)l[2],
                                       p.pts + 3 * (int64_t)l[3], p.pts + 3 * (int64_t)s};
                const int idx[5] = {l[0], l[1], l[2], l[3], s};
                const int o = EXACT ? orient4_full(pp, idx, cnt) : filter_insphere(pp);
                if (!EXACT && o == 0) {
                    star_list_push(p.exact_list, p.exact_cap, &cnt[p.exact_counter], s, cnt);
                    continue;
                }
                const bool sw = o < 0;
                if (sw) {
                    const int t = l[0];
                    l[0] = l[1];
                    l[1] = t;
                }
                if (lane < 4) {
                    L.vid[lane] = l[lane];
                    rel_coords(p.pts, l[lane], sx, sy, sz, L.vc[lane]);
                    // cone triangles (link triangles seen from the star point) and adjacency
                    const int lt[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
                    L.tv[lane] = tpack(lt[lane][0], lt[lane][1], lt[lane][2], TF_ALIVE);
                    L.to[lane][0] = (uint16_t)lt[lane][0];
                    L.to[lane][1] = (uint16_t)lt[lane][1];
                    L.to[lane][2] = (uint16_t)lt[lane][2];
                    const int local = (lane < 2 && sw) ? 1 - lane : lane;
                    L.ins_local[lane] = (int16_t)local;
                }
                T = 4;
            }
            wave_sync();
#pragma unroll 1
            for (int k = 0; k < KPL; k++) {
                const int slot = lane + 64 * k;
                if (slot < T) {
                    const uint32_t t = L.tv[slot];
                    lds_cofactors<TMAX>(L.vc[tvert(t, 0)], L.vc[tvert(t, 1)], L.vc[tvert(t, 2)], L.cof, slot);
                    ncof++;
                }
            }
            bool failed = false, unsure = false;
            for (int q = q0; q < m; q++) {
                const int gv = L.ins[q];
                const int Vn = vn + q;
                const double r[4] = {L.vc[Vn][0], L.vc[Vn][1], L.vc[Vn][2], L.vc[Vn][3]};
                int nv = 0;
                unsigned vm = 0; // bit k: slot lane + 64 k sees the vertex
                bool unc = false;
#pragma unroll 1
                for (int k = 0; k < KPL; k++) {
                    const int slot = lane + 64 * k;
                    if (slot < T) {
                        const uint32_t t = L.tv[slot];
                        if ((t & (TF_ALIVE | TF_VIS)) == TF_ALIVE) {
                            nfilt++;
                            double c[8];
#pragma unroll
                            for (int j = 0; j < 8; j++) c[j] = L.cof[j][slot];
                            int sg = cofactor_filter_sign(r, c, c + 4);
                            if (sg == 0) {
                                if (EXACT) {
                                    const int a = L.vid[tvert(t, 0)], b = L.vid[tvert(t, 1)], c = L.vid[tvert(t, 2)];
                                    const double *pp[5] = {p.pts + 3 * (int64_t)a, p.pts + 3 * (int64_t)b,
                                                           p.pts + 3 * (int64_t)c, p.pts + 3 * (int64_t)s,
                                                           p.pts + 3 * (int64_t)gv};
                                    const int idx[5] = {a, b, c, s, gv};
                                    sg = orient4_full(pp, idx, cnt);
                                } else {
                                    unc = true;
                                }
                            }
                            if (sg < 0) {
                                L.tv[slot] = t | TF_VIS;
                                vm |= 1u << k;
                                nv++;
                            }
                        }
                    }
                }
                if (!EXACT && __ballot(unc)) {
                    unsure = true;
                    break;
                }
                const int vinc = g4d_wave_incl_sum(nv);
                const int vbase = vinc - nv, nvis = __builtin_amdgcn_readlane(vinc, 63);
                if (nvis == 0) continue; // drowned at insertion
                if (lane == 0) L.ins_local[q] = (int16_t)Vn;
                // the visible slots join the free list (positions nfree .. nfree + nvis - 1)
                {
                    int fpos = nfree + vbase;
#pragma unroll 1
                    for (int k = 0; k < KPL; k++)
                        if ((vm >> k) & 1) L.freel[fpos++] = (uint16_t)(lane + 64 * k);
                }
                wave_sync();
                // Hole boundary: the edges of visible triangles whose neighbour is not visible, in
                // the order of the visible list; entry x is taken by lane x mod 64.  The edge's
                // position in the neighbour u is that of u's vertex off the edge.
                int nbtot = 0;
                for (int x0 = 0; x0 < nvis; x0 += 64) {
                    const int x = x0 + lane;
                    const bool own = x < nvis;
                    uint32_t t = 0u;
                    uint64_t o = 0ull;
                    if (own) {
                        const int slot = L.freel[nfree + x];
                        t = L.tv[slot];
                        o = L.to64[slot];
                    }
                    uint32_t tu[3] = {TF_VIS, TF_VIS, TF_VIS};
                    if (own)
#pragma unroll
                        for (int i = 0; i < 3; i++) tu[i] = L.tv[(int)(o >> (16 * i)) & (TMAX - 1)];
                    int nb = 0;
#pragma unroll
                    for (int i = 0; i < 3; i++) nb += (tu[i] & TF_VIS) ? 0 : 1;
                    const int binc = g4d_wave_incl_sum(nb);
                    int e = nbtot + binc - nb;
#pragma unroll
                    for (int i = 0; i < 3; i++) {
                        if (tu[i] & TF_VIS) continue;
                        const int a = tvert(t, (i + 1) % 3), b = tvert(t, (i + 2) % 3);
                        const int u0 = tvert(tu[i], 0), u1 = tvert(tu[i], 1);
                        if (e < TMAX) {
                            L.be_a[e] = (uint16_t)a;
                            L.be_b[e] = (uint16_t)b;
                            L.be_u[e] = (uint16_t)(o >> (16 * i));
                            L.be_j[e] = (uint8_t)((u0 != a && u0 != b) ? 0 : ((u1 != a && u1 != b) ? 1 : 2));
                        }
                        e++;
                    }
                    nbtot += __builtin_amdgcn_readlane(binc, 63);
                }
                // The visible triangles keep their words, TF_VIS marking them dead, until a new
                // triangle takes the slot: boundary edge k takes free slot nfree_tot - 1 - k (the
                // visible ones first) or, past the free list, slot T + k - nfree_tot.
                const int nfree_tot = nfree + nvis;
                const int newT = T + (nbtot > nfree_tot ? nbtot - nfree_tot : 0);
                if (nbtot > TMAX || nbtot < 3 || newT > TMAX || Vn + 1 > VMAX) {
                    failed = true;
                    break;
                }
                wave_sync();
                for (int k = lane; k < nbtot; k += 64) {
                    const int slot = k < nfree_tot ? L.freel[nfree_tot - 1 - k] : T + (k - nfree_tot);
                    const int a = L.be_a[k], b = L.be_b[k], u = L.be_u[k];
                    L.be_slot[k] = (uint16_t)slot;
                    L.tv[slot] = tpack(a, b, Vn, TF_ALIVE | TF_FRESH);
                    L.to[slot][2] = (uint16_t)u;
                    L.to[u][L.be_j[k]] = (uint16_t)slot;
                    L.vstart[a] = (int16_t)slot;
                    L.vend[b] = (int16_t)slot;
                }
                wave_sync();
                for (int k = lane; k < nbtot; k += 64) {
                    const int slot = L.be_slot[k];
                    L.to[slot][0] = (uint16_t)L.vstart[L.be_b[k]];
                    L.to[slot][1] = (uint16_t)L.vend[L.be_a[k]];
                }
                nfree = nfree_tot - (nbtot < nfree_tot ? nbtot : nfree_tot);
                T = newT;
                // the link writes above touch only the new slots' neighbour rows: the cofactor
                // pass below (triangle words, coordinates, this lane's cofactor columns) needs no
                // barrier after them
#pragma unroll 1
                for (int k = 0; k < KPL; k++) {
                    const int slot = lane + 64 * k;
                    if (slot < T) {
                        const uint32_t t = L.tv[slot];
                        if (t & TF_FRESH) {
                            lds_cofactors<TMAX>(L.vc[tvert(t, 0)], L.vc[tvert(t, 1)], L.vc[tvert(t, 2)], L.cof,
                                                slot);
                            ncof++;
                            L.tv[slot] = t & ~TF_FRESH;
                        }
                    }
                }
                // No barrier before the next step: its filter reads and writes only this lane's
                // own slots and cofactor columns and the listed coordinates, and its first barrier
                // precedes any read of another lane's words.
            }
            wave_sync();
            if (unsure) {
                star_list_push(p.exact_list, p.exact_cap, &cnt[p.exact_counter], s, cnt);
                wave_sync();
                continue;
            }
            if (failed) {
                if (lane == 0) {
                    atomicOr(&cnt[G4D_CNT_ERROR], (unsigned long long)G4D_DEV_ERR_HOLE);
                    p.vnum[s] = 0;
                    p.tnum[s] = 0;
                }
                wave_sync();
                continue;
            }
            // ---- finalize: link set, drowned insertions, compaction, pairs, history ----
            const int V = need; // drowned insertions' slots are never referenced and drop out
            for (int v = lane; v < V; v += 64) L.present[v] = 0;
            wave_sync();
#pragma unroll 1
            for (int k = 0; k < KPL; k++) {
                const int slot = lane + 64 * k;
                if (slot < T) {
                    const uint32_t t = L.tv[slot];
                    if ((t & (TF_ALIVE | TF_VIS)) == TF_ALIVE) {
                        L.present[tvert(t, 0)] = 1;
                        L.present[tvert(t, 1)] = 1;
                        L.present[tvert(t, 2)] = 1;
                    }
                }
            }
            wave_sync();
            for (int q = lane; q < m; q += 64) {
                const int lv = L.ins_local[q];
                p.dflag[ib + q] = (lv < 0 || !L.present[lv]) ? 1 : 0;
            }
            // vertex compaction
            const int64_t vo = p.nvoff[s], to = p.ntoff[s];
            int vcount = 0;
            for (int base = 0; base < V; base += 64) {
                const int v = base + lane;
                const bool pr = v < V && L.present[v];
                int tot;
                const int rk = g4d_wave_excl_count(pr, &tot);
                if (pr) {
                    L.vremap[v] = (uint16_t)(vcount + rk);
                    p.vpool[vo + vcount + rk] = L.vid[v];
                }
                vcount += tot;
            }
            int tcount = 0;
            for (int base = 0; base < T; base += 64) {
                const int t = base + lane;
                const bool al = t < T && (L.tv[t] & (TF_ALIVE | TF_VIS)) == TF_ALIVE;
                int tot;
                const int rk = g4d_wave_excl_count(al, &tot);
                if (al) L.remap[t] = (uint16_t)(tcount + rk);
                tcount += tot;
            }
            if (tcount > 2 * need - 4 || tcount != 2 * vcount - 4) {
                // not a triangulated 2-sphere: report instead of writing past the star's region
                if (lane == 0) {
                    atomicOr(&cnt[G4D_CNT_ERROR], (unsigned long long)G4D_DEV_ERR_HOLE);
                    p.vnum[s] = 0;
                    p.tnum[s] = 0;
                }
                wave_sync();
                continue;
            }
            wave_sync();
            const int64_t pb = 2 * (to - p.tbase);
            int npair = 0;
            for (int base = 0; base < T; base += 64) {
                const int t = base + lane;
                uint32_t tt = 0;
                if (t < T) tt = L.tv[t];
                TriPairs pr{STAR_PAIR_NONE, STAR_PAIR_NONE};
                if (t < T && (tt & (TF_ALIVE | TF_VIS)) == TF_ALIVE) {
                    const int a = tvert(tt, 0), b = tvert(tt, 1), c = tvert(tt, 2);
                    const int d = L.remap[t];
                    p.tpool[to + d] = hpack(L.vremap[a], L.vremap[b], L.vremap[c], L.remap[L.to[t][0]],
                                            L.remap[L.to[t][1]], L.remap[L.to[t][2]], true);
                    pr = star_tri_pairs(L.vid[a], L.vid[b], L.vid[c], !(tt & TF_CHECKED));
                }
                // the star's pairs, dense, in triangle order
                const int np = (pr.x != STAR_PAIR_NONE ? 1 : 0) + (pr.y != STAR_PAIR_NONE ? 1 : 0);
                int tot;
                const int at = npair + wave_excl<2>(np, &tot);
                if (np > 0) p.pairbuf[pb + at] = pr.x;
                if (np > 1) p.pairbuf[pb + at + 1] = pr.y;
                npair += tot;
            }
            if (lane == 0) p.pcount[s] = npair;
            const int hn = star_history_merge(p, s, L.ins, m, lane, 64);
            wave_sync();
            if (lane == 0) star_commit(p, s, vcount, vo, tcount, to, hn + m);
            wave_sync();
        }
    }
    const int wf = g4d_wave_sum(nfilt), wc = g4d_wave_sum(ncof);
    if (lane == 0) {
        atomicAdd(&cnt[G4D_CNT_ORIENT4], (unsigned long long)wf);
        atomicAdd(&cnt[G4D_CNT_COFACTOR], (unsigned long long)wc);
    }
}

// The 32-vertex tier (7 KB of LDS) is compiled for 5 waves per SIMD (96 VGPRs, a few spilled
// to scratch): measured at 4,194,304 uniform points, 5 waves beat 3, 4, 6 and 7 (49.8 ms per
// triangulation against 63.2, 54.2, 51.0 and 62.9).  The larger tiers are limited by their LDS
// footprint instead.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5, 8))) k_star32(StarParams p, int nitems)
{
    star_body<STAR_VMAX_T0, false>(p, nitems);
}

template <int VMAX, bool EXACT>
__global__ void __launch_bounds__(64) k_star(StarParams p, int nitems)
{
    star_body<VMAX, EXACT>(p, nitems);
}

// ------------------------------------------------------------------------------------
// Grouped 32-vertex star kernel: one star per group of GL lanes, 64 / GL stars per wave.
//
// The one-star-per-wave kernel leaves most lanes idle on a star of this tier (at most 60
// triangle slots, about 27 live) and pays its chain of ballots and wave-level barriers per
// insertion of one star; here the chain is shared by 64 / GL stars.  Each lane owns KPL =
// 64 / GL triangle slots (gl, gl + GL, ...) and keeps their cofactors and permanents in
// registers (only the owning lane ever reads or writes them), so a star's LDS holds only its
// link vertices, triangles and hole-boundary lists (about 3 KB).  Every barrier and ballot is
// reached by the whole wave in uniform control flow: the insertion loop runs to the largest
// insertion count of the wave's stars, and a group with nothing to do in a step has its work
// masked off.  Filters only (no multiprecision code): a star with an inconclusive orientation
// goes to p.exact_list untouched, a star beyond the tier to p.defer_list, as in k_star32.
// ------------------------------------------------------------------------------------

struct StarGrpLds {
    static constexpr int VMAX = STAR_VMAX_T0;
    static constexpr int TMAX = 2 * VMAX;
    double vc[VMAX][4];
    int vid[VMAX];
    int ins[VMAX];
    uint32_t tv[TMAX];
    union {
        uint16_t to[TMAX][4]; // neighbours across the edges opposite vertices 0-2 (fourth unused)
        uint64_t to64[TMAX];  // the same rows read as one word each
    };
    int16_t vstart[VMAX];
    int16_t vend[VMAX];
    uint16_t freel[TMAX];
    uint16_t be_a[TMAX], be_b[TMAX], be_u[TMAX], be_slot[TMAX];
    uint8_t be_j[TMAX];
    int16_t ins_local[VMAX];
    uint8_t ord[VMAX]; // insertion order: list indices by distance to the star point
    uint8_t present[VMAX];
    uint16_t remap[TMAX];
    uint16_t vremap[VMAX];
    // The fields above take 3200 bytes.  Padded to 3264 (192 mod 256) the four groups' blocks
    // start 48 banks apart (mod 64): group-parallel accesses (16 lanes of each group at the same
    // offsets) do not all land on the same banks, as they would four ways at a multiple of 256.
    uint8_t pad[64];
};
static_assert(sizeof(StarGrpLds) % 256 == 192, "group blocks start 48 banks apart");
static_assert(offsetof(StarGrpLds, to) % 8 == 0, "neighbour rows read as 64-bit words");

// Ballot of this lane's group of GL lanes, shifted down to bit 0.
template <int GL>
__device__ __forceinline__ uint64_t grp_ballot(bool pred, int g)
{
    const uint64_t m = __ballot(pred);
    return GL == 64 ? m : (m >> (g * GL)) & ((1ull << GL) - 1ull);
}

// Inclusive prefix sums over each DPP row of 16 lanes (row_shr 1, 2, 4, 8) and every lane's row
// total (row_ror 8, 4, 2, 1); lanes whose shifted source lies outside the row add 0.  Every lane
// of the wave must be active.
__device__ __forceinline__ int row16_scan(int v, int *tot)
{
    int x = v, a = v;
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);
    a += __builtin_amdgcn_update_dpp(0, a, 0x128, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);
    a += __builtin_amdgcn_update_dpp(0, a, 0x124, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);
    a += __builtin_amdgcn_update_dpp(0, a, 0x122, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);
    a += __builtin_amdgcn_update_dpp(0, a, 0x121, 0xf, 0xf, false);
    *tot = a;
    return x;
}

// Exclusive scan of per-lane counts below 2^BITS within the group; total in *tot.  Groups of 16
// lanes are DPP rows and scan any counts whose group sum stays below 2^31.
template <int GL, int BITS>
__device__ __forceinline__ int grp_excl(int v, int g, int gl, int *tot)
{
    if constexpr (GL == 16) return row16_scan(v, tot) - v;
    const uint64_t below = gl == 0 ? 0ull : (~0ull >> (64 - gl));
    int pre = 0, all = 0;
#pragma unroll
    for (int b = 0; b < BITS; b++) {
        const uint64_t m = grp_ballot<GL>((v >> b) & 1, g);
        pre += __popcll(m & below) << b;
        all += __popcll(m) << b;
    }
    *tot = all;
    return pre;
}

// Cofactors of a triangle slot, with either the four permanents in FP64 or only their largest,
// rounded up to FP32.  The filter's error bound 32 u sum_j |v'_j| P_j is then replaced by
// 32 u (sum_j |v'_j|) max_j P_j, which is no smaller, so a decided sign stays certain; an
// overflow to +inf or a bound that vanishes only makes the filter inconclusive.
__device__ __forceinline__ void grp_cofactors(const double *r0, const double *r1, const double *r2, double *C,
                                              double (&P)[4])
{
    tri_cofactors(r0, r1, r2, C, P);
}
__device__ __forceinline__ void grp_cofactors(const double *r0, const double *r1, const double *r2, double *C,
                                              float &P)
{
    double Pd[4];
    tri_cofactors(r0, r1, r2, C, Pd);
    P = __double2float_ru(fmax(fmax(Pd[0], Pd[1]), fmax(Pd[2], Pd[3])));
}
// rs = |v'_x| + |v'_y| + |v'_z| + v'_l of the vertex being inserted (used with the FP32 form).
__device__ __forceinline__ int grp_filter_sign(const double *r, double, const double *C, const double (&P)[4])
{
    return cofactor_filter_sign(r, C, P);
}
__device__ __forceinline__ int grp_filter_sign(const double *r, double rs, const double *C, const float &P)
{
    const double val = fma(r[3], C[3], fma(r[2], C[2], fma(r[1], C[1], r[0] * C[0])));
    const double per = rs * (double)P;
    const double bnd = 32.0 * 1.1102230246251565e-16 * per;
    if (!(per > 1e-250)) return 0;
    return val > bnd ? 1 : (-val > bnd ? -1 : 0);
}

// Largest value over the wave for non-negative v: maxima within each 16-lane row by DPP row
// rotations (8, 4, 2, 1), then the four rows' maxima read into scalars.  Every lane must be
// active.
__device__ __forceinline__ int wave_max_i(int v)
{
    int x = v;
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x128, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x124, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x122, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x121, 0xf, 0xf, false));
    const int r0 = __builtin_amdgcn_readlane(x, 0), r1 = __builtin_amdgcn_readlane(x, 16);
    const int r2 = __builtin_amdgcn_readlane(x, 32), r3 = __builtin_amdgcn_readlane(x, 48);
    return max(max(r0, r1), max(r2, r3));
}

template <int GL, bool PF>
__device__ __forceinline__ void star_body_grp(const StarParams &p, int nitems)
{
    constexpr int NG = 64 / GL;
    constexpr int VMAX = StarGrpLds::VMAX, TMAX = StarGrpLds::TMAX;
    constexpr int KPL = TMAX / GL;
    __shared__ StarGrpLds Ls[NG];
    __shared__ int order[NG * STAR_CHUNK]; // the chunk's items by insertion count
    const int lane = threadIdx.x, g = lane / GL, gl = lane % GL;
    StarGrpLds &L = Ls[g];
    unsigned long long *cnt = p.cnt;

    if (p.nlist_dev) {
        const unsigned long long c = *p.nlist_dev;
        nitems = c < (unsigned long long)nitems ? (int)c : nitems;
    }
    // Work unit: a batch of NG consecutive list entries, one per group.  As many batches as
    // work-groups or fewer: one batch each, no queue; otherwise STAR_CHUNK batches per atomic.
    const int nbatch = (nitems + NG - 1) / NG;
    const bool one_each = nbatch <= (int)gridDim.x;
    if ((int)blockIdx.x >= nbatch) return;
    bool first = true;
    int nfilt = 0, ncof = 0;
    double C[KPL][4];
    typename std::conditional<PF, float, double[4]>::type P[KPL];
    for (;;) {
        int base = 0;
        if (one_each) {
            base = first ? (int)blockIdx.x : nbatch;
            first = false;
        } else {
            if (lane == 0) base = (int)atomicAdd(&cnt[p.work_counter], (unsigned long long)STAR_CHUNK);
            base = __shfl(base, 0, 64);
        }
        if (base >= nbatch) break;
        const int bend = one_each ? base + 1 : (base + STAR_CHUNK < nbatch ? base + STAR_CHUNK : nbatch);
        // The chunk's stars are dealt to batches in order of their insertion counts, so that the
        // four stars of a batch (which step through their insertions in lock-step) need about as
        // many steps each; within a chunk of 32 consecutive stars the order keeps their locality.
        {
            const int nit = (bend - base) * NG;
            const int my = base * NG + lane;
            int key = INT_MAX;
            if (lane < nit && my < nitems) {
                const int s = p.star_list ? p.star_list[my] : my;
                key = (int)(p.ins_off[s + 1] - p.ins_off[s]);
            }
            int rank = 0;
#pragma unroll
            for (int j = 0; j < NG * STAR_CHUNK; j++) {
                const int kj = __shfl(key, j, 64);
                rank += (kj < key || (kj == key && j < lane)) ? 1 : 0;
            }
            if (lane < nit) order[rank] = my;
            wave_sync();
        }
#pragma unroll 1
        for (int batch = base; batch < bend; batch++) {
            const int item = order[(batch - base) * NG + g];
            // ---- this group's star: checks, then its link and insertions into LDS ----
            bool act = item < nitems;
            const int s = act ? (p.star_list ? p.star_list[item] : item) : 0;
            const int64_t ib = act ? p.ins_off[s] : 0;
            const int m = act ? (int)(p.ins_off[s + 1] - ib) : 0;
            const int vn = act && p.mode ? p.vnum[s] : 0;
            const int need = vn + m;
            const int lim = VMAX < p.lds_limit ? VMAX : p.lds_limit;
            if (act && need > lim) {
                if (gl == 0) {
                    const unsigned long long k = atomicAdd(&cnt[p.defer_counter], 1ull);
                    if ((int64_t)k < p.defer_cap) p.defer_list[k] = s;
                    else atomicOr(&cnt[G4D_CNT_ERROR], (unsigned long long)G4D_DEV_ERR_BUFFER);
                }
                act = false;
            }
            if (act && p.mode == 0 && m < 4) {
                if (gl == 0) {
                    atomicOr(&cnt[G4D_CNT_ERROR], (unsigned long long)G4D_DEV_ERR_CONE);
                    p.vnum[s] = 0;
                    p.tnum[s] = 0;
                }
                act = false;
            }
            if (act && p.mode == 1 && m == 0) act = false;
            double sx = 0.0, sy = 0.0, sz = 0.0;
            if (act) {
                sx = p.pts[3 * (int64_t)s];
                sy = p.pts[3 * (int64_t)s + 1];
                sz = p.pts[3 * (int64_t)s + 2];
            }
            const int q0 = p.mode ? 0 : 4;
            // Every listed vertex q takes link-vertex slot vn + q (the cone's four included).
            if (act) {
                for (int q = gl; q < m; q += GL) {
                    const int gv = (int)(uint32_t)p.ins_keys[ib + q];
                    L.ins[q] = gv;
                    L.ins_local[q] = -1;
                    L.vid[vn + q] = gv;
                    rel_coords(p.pts, gv, sx, sy, sz, L.vc[vn + q]);
                }
            }
            wave_sync();
            // Insertion order: nearest to the star point first (ties by list index).  The star
            // that results does not depend on the order (the link is the unique convex hull of
            // the lifted points as seen from the star point, the drowned flags mark listed
            // vertices absent from it, the pairs come from its unchecked triangles), but near
            // vertices are almost all in it: inserted first, they leave the far ones to be
            // rejected by the filter alone, and fewer triangles are made and undone.
            if (act)
                for (int q = gl; q < m; q += GL) {
                    const double dq = L.vc[vn + q][3];
                    int rank = 0;
                    for (int k0 = 0; k0 < m; k0 += 4) { // four reads in flight (indices kept below m)
                        double dk[4];
#pragma unroll
                        for (int i = 0; i < 4; i++) dk[i] = L.vc[vn + (k0 + i < m ? k0 + i : m - 1)][3];
#pragma unroll
                        for (int i = 0; i < 4; i++) {
                            const int k = k0 + i;
                            rank += (k < m && (dk[i] < dq || (dk[i] == dq && k < q))) ? 1 : 0;
                        }
                    }
                    L.ord[rank] = (uint8_t)q;
                }
            wave_sync();
            int T = 0, nfree = 0;
            bool unsure = false, failed = false;
            if (act && p.mode == 1) {
                const int tn = p.tnum[s];
                const int64_t vi = p.voff[s], ti = p.toff[s];
                for (int v = gl; v < vn; v += GL) {
                    const int gv = p.vpool[vi + v];
                    L.vid[v] = gv;
                    rel_coords(p.pts, gv, sx, sy, sz, L.vc[v]);
                }
                for (int t = gl; t < tn; t += GL) {
                    const uint64_t h = p.tpool[ti + t];
                    L.tv[t] = tpack(hvert(h, 0), hvert(h, 1), hvert(h, 2), TF_ALIVE | (hchecked(h) ? TF_CHECKED : 0u));
                    L.to[t][0] = (uint16_t)hopp(h, 0);
                    L.to[t][1] = (uint16_t)hopp(h, 1);
                    L.to[t][2] = (uint16_t)hopp(h, 2);
                }
                T = tn;
            } else if (act) {
                // initial 4-simplex (GD/GDelaunay/GDelPredKernels.cu:45-158) on the first four
                // vertices of the insertion order; slot c[i] holds cone vertex i
                int c[4];
                for (int i = 0; i < 4; i++) c[i] = L.ord[i];
                // the cone's coordinates relative to the star point, as listed in LDS (vn = 0)
                const double *pr[4] = {L.vc[c[0]], L.vc[c[1]], L.vc[c[2]], L.vc[c[3]]};
                const int o = filter_insphere_rel(pr);
                if (o == 0) {
                    unsure = true; // pushed to the exact list below, untouched
                } else {
                    if (o < 0) {
                        const int t = c[0];
                        c[0] = c[1];
                        c[1] = t;
                    }
                    if (gl < 4) {
                        const int lt[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
                        L.tv[gl] = tpack(c[lt[gl][0]], c[lt[gl][1]], c[lt[gl][2]], TF_ALIVE);
                        L.to[gl][0] = (uint16_t)lt[gl][0];
                        L.to[gl][1] = (uint16_t)lt[gl][1];
                        L.to[gl][2] = (uint16_t)lt[gl][2];
                        L.ins_local[c[gl]] = (int16_t)c[gl];
                    }
                    T = 4;
                }
            }
            wave_sync();
            bool live = act && !unsure; // the group is building its star
#pragma unroll
            for (int k = 0; k < KPL; k++) {
                const int slot = gl + GL * k;
                if (live && slot < T) {
                    const uint32_t t = L.tv[slot];
                    grp_cofactors(L.vc[tvert(t, 0)], L.vc[tvert(t, 1)], L.vc[tvert(t, 2)], C[k], P[k]);
                    ncof++;
                }
            }
            // ---- insertions: every group steps through its list in lock-step ----
            const int qend = wave_max_i(live ? m : 0);
            int jnext = L.ord[q0 < m ? q0 : 0];
            for (int q = q0; q < qend; q++) {
                bool step = live && q < m;
                const int j = step ? jnext : 0; // the list index inserted in this step
                jnext = L.ord[q + 1 < m ? q + 1 : 0]; // the next step's, read a step ahead
                const int Vn = vn + j;
                // this lane's triangle words, read together (every slot index is below TMAX)
                uint32_t tk[KPL];
#pragma unroll
                for (int k = 0; k < KPL; k++) tk[k] = L.tv[gl + GL * k];
                double r[4] = {0.0, 0.0, 0.0, 0.0};
                if (step) {
                    r[0] = L.vc[Vn][0];
                    r[1] = L.vc[Vn][1];
                    r[2] = L.vc[Vn][2];
                    r[3] = L.vc[Vn][3];
                }
                const double rs = fabs(r[0]) + fabs(r[1]) + fabs(r[2]) + r[3];
                int nv = 0;
                unsigned vm = 0; // bit k: slot gl + GL k sees the vertex
                bool unc = false;
#pragma unroll
                for (int k = 0; k < KPL; k++) {
                    const int slot = gl + GL * k;
                    if (step && slot < T && (tk[k] & (TF_ALIVE | TF_VIS)) == TF_ALIVE) {
                        nfilt++;
                        const int sg = grp_filter_sign(r, rs, C[k], P[k]);
                        if (sg == 0) unc = true;
                        if (sg < 0) {
                            L.tv[slot] = tk[k] | TF_VIS;
                            vm |= 1u << k;
                            nv++;
                        }
                    }
                }
                // every group drowned its vertex at once, or idles: nothing changes in this step
                if (__ballot(step && (nv > 0 || unc)) == 0ull) continue;
                int nvis, vbase;
                bool gunc;
                if constexpr (GL == 16) {
                    // one row scan of two counts: visible triangles (bits 0-7, at most TMAX) and
                    // inconclusive lanes (bits 8-12)
                    static_assert(TMAX < 256, "packed counts fit their fields");
                    const int pk = nv | (unc ? 1 << 8 : 0);
                    int tot;
                    vbase = (row16_scan(pk, &tot) - pk) & 255;
                    nvis = tot & 255;
                    gunc = (tot >> 8) != 0;
                } else {
                    gunc = grp_ballot<GL>(unc, g) != 0ull;
                    vbase = grp_excl<GL, count_bits(KPL)>(nv, g, gl, &nvis);
                }
                if (gunc) {
                    unsure = true;
                    live = false;
                    step = false;
                }
                if (nvis == 0) step = false; // drowned at insertion
                if (step && gl == 0) L.ins_local[j] = (int16_t)Vn;
                // the visible slots join the free list (positions nfree .. nfree + nvis - 1)
                if (step) {
                    int fpos = nfree + vbase;
#pragma unroll
                    for (int k = 0; k < KPL; k++)
                        if ((vm >> k) & 1) L.freel[fpos++] = (uint16_t)(gl + GL * k);
                }
                wave_sync();
                // Hole boundary: the edges of visible triangles whose neighbour is not visible, in
                // the order of the visible list; its entry x is taken by lane x mod GL.  The edge's
                // position in the neighbour u is that of u's vertex off the edge.
                int nbtot = 0;
                for (int x0 = 0; __ballot(step && x0 < nvis) != 0ull; x0 += GL) {
                    const int x = x0 + gl;
                    const bool own = step && x < nvis;
                    uint32_t t = 0u;
                    uint64_t o = 0ull;
                    if (own) {
                        const int slot = L.freel[nfree + x];
                        t = L.tv[slot];
                        o = L.to64[slot];
                    }
                    uint32_t tu[3] = {TF_VIS, TF_VIS, TF_VIS};
                    if (own)
#pragma unroll
                        for (int i = 0; i < 3; i++) tu[i] = L.tv[(int)(o >> (16 * i)) & (TMAX - 1)];
                    int nb = 0;
#pragma unroll
                    for (int i = 0; i < 3; i++) nb += (tu[i] & TF_VIS) ? 0 : 1;
                    int tot;
                    int e = nbtot + grp_excl<GL, 2>(nb, g, gl, &tot);
#pragma unroll
                    for (int i = 0; i < 3; i++) {
                        if (tu[i] & TF_VIS) continue;
                        const int a = tvert(t, (i + 1) % 3), b = tvert(t, (i + 2) % 3);
                        const int u0 = tvert(tu[i], 0), u1 = tvert(tu[i], 1);
                        if (e < TMAX) {
                            L.be_a[e] = (uint16_t)a;
                            L.be_b[e] = (uint16_t)b;
                            L.be_u[e] = (uint16_t)(o >> (16 * i));
                            L.be_j[e] = (uint8_t)((u0 != a && u0 != b) ? 0 : ((u1 != a && u1 != b) ? 1 : 2));
                        }
                        e++;
                    }
                    nbtot += tot;
                }
                // The visible triangles keep their words, TF_VIS marking them dead, until a new
                // triangle takes the slot: boundary edge k takes free slot nfree_tot - 1 - k (the
                // visible ones first) or, past the free list, slot T + k - nfree_tot.
                const int nfree_tot = nfree + nvis;
                const int newT = T + (nbtot > nfree_tot ? nbtot - nfree_tot : 0);
                if (step && (nbtot > TMAX || nbtot < 3 || newT > TMAX || Vn + 1 > VMAX)) {
                    failed = true;
                    live = false;
                    step = false;
                }
                wave_sync();
                if (step)
                    for (int k = gl; k < nbtot; k += GL) {
                        const int slot = k < nfree_tot ? L.freel[nfree_tot - 1 - k] : T + (k - nfree_tot);
                        const int a = L.be_a[k], b = L.be_b[k], u = L.be_u[k];
                        L.be_slot[k] = (uint16_t)slot;
                        L.tv[slot] = tpack(a, b, Vn, TF_ALIVE | TF_FRESH);
                        L.to[slot][2] = (uint16_t)u;
                        L.to[u][L.be_j[k]] = (uint16_t)slot;
                        L.vstart[a] = (int16_t)slot;
                        L.vend[b] = (int16_t)slot;
                    }
                wave_sync();
                if (step) {
                    for (int k = gl; k < nbtot; k += GL) {
                        const int slot = L.be_slot[k];
                        L.to[slot][0] = (uint16_t)L.vstart[L.be_b[k]];
                        L.to[slot][1] = (uint16_t)L.vend[L.be_a[k]];
                    }
                    nfree = nfree_tot - (nbtot < nfree_tot ? nbtot : nfree_tot);
                    T = newT;
                }
                // the link writes above touch only the new slots' neighbour rows: the cofactor
                // pass below (triangle words and coordinates) needs no barrier after them
#pragma unroll
                for (int k = 0; k < KPL; k++) tk[k] = L.tv[gl + GL * k];
#pragma unroll
                for (int k = 0; k < KPL; k++) {
                    const int slot = gl + GL * k;
                    if (step && slot < T && (tk[k] & TF_FRESH)) {
                        const uint32_t t = tk[k];
                        grp_cofactors(L.vc[tvert(t, 0)], L.vc[tvert(t, 1)], L.vc[tvert(t, 2)], C[k], P[k]);
                        ncof++;
                        L.tv[slot] = t & ~TF_FRESH;
                    }
                }
                // No barrier before the next step: its filter reads and writes only this lane's
                // own slots and the listed coordinates, and its first barrier precedes any read
                // of another lane's words.
            }
            wave_sync();
            // ---- outcome per group ----
            if (act && unsure && gl == 0) {
                const unsigned long long k = atomicAdd(&cnt[p.exact_counter], 1ull);
                if ((int64_t)k < p.exact_cap) p.exact_list[k] = s;
                else atomicOr(&cnt[G4D_CNT_ERROR], (unsigned long long)G4D_DEV_ERR_BUFFER);
            }
            if (act && failed && gl == 0) {
                atomicOr(&cnt[G4D_CNT_ERROR], (unsigned long long)G4D_DEV_ERR_HOLE);
                p.vnum[s] = 0;
                p.tnum[s] = 0;
            }
            bool fin = act && !unsure && !failed;
            // ---- finalize: link set, drowned insertions, compaction, pairs, history ----
            const int V = need; // drowned insertions' slots are never referenced and drop out
            if (fin)
                for (int v = gl; v < V; v += GL) L.present[v] = 0;
            // this lane's live triangle words (0 for dead and unused slots), read together
            uint32_t tf[KPL];
#pragma unroll
            for (int k = 0; k < KPL; k++) {
                const int slot = gl + GL * k;
                const uint32_t t = L.tv[slot];
                tf[k] = (fin && slot < T && (t & (TF_ALIVE | TF_VIS)) == TF_ALIVE) ? t : 0u;
            }
            wave_sync();
#pragma unroll
            for (int k = 0; k < KPL; k++)
                if (tf[k]) {
                    L.present[tvert(tf[k], 0)] = 1;
                    L.present[tvert(tf[k], 1)] = 1;
                    L.present[tvert(tf[k], 2)] = 1;
                }
            wave_sync();
            if (fin)
                for (int q = gl; q < m; q += GL) {
                    const int lv = L.ins_local[q];
                    p.dflag[ib + q] = (lv < 0 || !L.present[lv]) ? 1 : 0;
                }
            const int64_t vo = fin ? p.nvoff[s] : 0, to = fin ? p.ntoff[s] : 0;
            int vcount = 0;
#pragma unroll
            for (int b0 = 0; b0 < VMAX; b0 += GL) {
                const int v = b0 + gl;
                const bool pr = fin && v < V && L.present[v];
                int tot;
                const int rk = grp_excl<GL, 1>(pr ? 1 : 0, g, gl, &tot);
                if (pr) {
                    L.vremap[v] = (uint16_t)(vcount + rk);
                    p.vpool[vo + vcount + rk] = L.vid[v];
                }
                vcount += tot;
            }
            int tcount = 0;
#pragma unroll
            for (int b0 = 0; b0 < TMAX; b0 += GL) {
                const int t = b0 + gl;
                const bool al = tf[b0 / GL] != 0u;
                int tot;
                const int rk = grp_excl<GL, 1>(al ? 1 : 0, g, gl, &tot);
                if (al) L.remap[t] = (uint16_t)(tcount + rk);
                tcount += tot;
            }
            if (fin && (tcount > 2 * need - 4 || tcount != 2 * vcount - 4)) {
                // not a triangulated 2-sphere: report instead of writing past the star's region
                if (gl == 0) {
                    atomicOr(&cnt[G4D_CNT_ERROR], (unsigned long long)G4D_DEV_ERR_HOLE);
                    p.vnum[s] = 0;
                    p.tnum[s] = 0;
                }
                fin = false;
            }
            wave_sync();
            const int64_t pb = 2 * (to - p.tbase);
            // every slot's output word and pairs first (their reads overlap), then the pair scans
            TriPairs prk[KPL];
#pragma unroll
            for (int k = 0; k < KPL; k++) {
                const int t = gl + GL * k;
                const uint32_t tt = fin ? tf[k] : 0u;
                prk[k] = TriPairs{STAR_PAIR_NONE, STAR_PAIR_NONE};
                if (tt) {
                    const int a = tvert(tt, 0), b = tvert(tt, 1), c = tvert(tt, 2);
                    const int d = L.remap[t];
                    const uint64_t o = L.to64[t];
                    p.tpool[to + d] = hpack(L.vremap[a], L.vremap[b], L.vremap[c], L.remap[(int)(o & 0xffff)],
                                            L.remap[(int)((o >> 16) & 0xffff)], L.remap[(int)((o >> 32) & 0xffff)], true);
                    prk[k] = star_tri_pairs(L.vid[a], L.vid[b], L.vid[c], !(tt & TF_CHECKED));
                }
            }
            int npair = 0;
#pragma unroll
            for (int k = 0; k < KPL; k++) {
                const TriPairs pr = prk[k];
                // the star's pairs, dense, in triangle order
                const int np = (pr.x != STAR_PAIR_NONE ? 1 : 0) + (pr.y != STAR_PAIR_NONE ? 1 : 0);
                int tot;
                const int at = npair + grp_excl<GL, 2>(np, g, gl, &tot);
                if (np > 0) p.pairbuf[pb + at] = pr.x;
                if (np > 1) p.pairbuf[pb + at + 1] = pr.y;
                npair += tot;
            }
            int hn = 0;
            if (fin) {
                if (gl == 0) p.pcount[s] = npair;
                hn = star_history_merge(p, s, L.ins, m, gl, GL);
            }
            wave_sync();
            if (fin && gl == 0) star_commit(p, s, vcount, vo, tcount, to, hn + m);
            wave_sync();
        }
    }
    const int wf = g4d_wave_sum(nfilt), wc = g4d_wave_sum(ncof);
    if (lane == 0) {
        atomicAdd(&cnt[G4D_CNT_ORIENT4], (unsigned long long)wf);
        atomicAdd(&cnt[G4D_CNT_COFACTOR], (unsigned long long)wc);
    }
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 8))) k_star32g16(StarParams p, int nitems)
{
    star_body_grp<16, false>(p, nitems);
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 8))) k_star32g16f(StarParams p, int nitems)
{
    star_body_grp<16, true>(p, nitems);
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) k_star32g16f4(StarParams p, int nitems)
{
    star_body_grp<16, true>(p, nitems);
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) k_star32g32f(StarParams p, int nitems)
{
    star_body_grp<32, true>(p, nitems);
}

// The 32-vertex filter-only kernel variant of a launch: the caller's choice, else the
// G4D_STAR_T0 environment variable (measurement), else the measured best (four stars per wave
// with the FP32 largest permanent: 73.4 ms per triangulation at 4,194,304 uniform points, G =
// 512, against 89.1 for one star per wave).
static int star_t0_variant(int requested)
{
    if (requested >= 0 && requested < STAR_T0_NUM) return requested;
    static int env = -2;
    if (env == -2) {
        const char *e = getenv("G4D_STAR_T0");
        env = e ? atoi(e) : -1;
        if (env < 0 || env >= STAR_T0_NUM) env = -1;
    }
    return env >= 0 ? env : STAR_T0_G16F;
}

static_assert(G4D_CNT_QUEUE0 + 6 <= G4D_CNT_NUM, "six queue heads must fit the counter block");
static_assert(G4D_CNT_NUM <= 32, "counter masks are 32 bits wide");

__global__ void k_zero_counters(unsigned long long *cnt, uint32_t mask)
{
    const int i = threadIdx.x;
    if (i < G4D_CNT_NUM && ((mask >> i) & 1u)) cnt[i] = 0ull;
}

hipError_t launch_zero_counters(unsigned long long *cnt, uint32_t mask, hipStream_t s)
{
    if (G4D_CNT_NUM < 32 && (mask >> G4D_CNT_NUM) != 0u) return hipErrorInvalidValue; // beyond the block
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s, cnt, mask);
    return hipGetLastError();
}

hipError_t launch_star_kernel(const StarParams &p, int vmax, bool exact, int nblocks, hipStream_t s)
{
    const int nitems = p.star_list ? p.nlist : p.n;
    if (nitems <= 0) return hipSuccess;
    if (p.nlist_dev && !p.star_list) return hipErrorInvalidValue;
    if (p.lds_limit > STAR_LDS_LIMIT) return hipErrorInvalidValue; // local ids must fit 9 bits
    if (nblocks > nitems) nblocks = nitems;
    if (exact) {
        if (vmax == STAR_VMAX_T0)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_star<STAR_VMAX_T0, true>), dim3(nblocks), dim3(64), 0, s, p, nitems);
        else if (vmax == STAR_VMAX_T1)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_star<STAR_VMAX_T1, true>), dim3(nblocks), dim3(64), 0, s, p, nitems);
        else if (vmax == STAR_VMAX_T2)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_star<STAR_VMAX_T2, true>), dim3(nblocks), dim3(64), 0, s, p, nitems);
        else
            return hipErrorInvalidValue;
    } else if (vmax == STAR_VMAX_T0) {
        const int v = star_t0_variant(p.t0_variant);
        if (v == STAR_T0_WAVE) {
            hipLaunchKernelGGL(k_star32, dim3(nblocks), dim3(64), 0, s, p, nitems);
        } else {
            // one work-group per resident wave: 256 CUs x 4 SIMDs x the variant's occupancy
            // one work-group per resident wave (256 CUs x 4 SIMDs x the variant's occupancy), or
            // one per batch of stars when the list is shorter
            const int per_simd = v <= STAR_T0_G16F ? 3 : 4;
            const int ng = v == STAR_T0_G32F ? 2 : 4;
            const int nb = std::max(1, std::min(1024 * per_simd, (nitems + ng - 1) / ng));
            if (v == STAR_T0_G16) hipLaunchKernelGGL(k_star32g16, dim3(nb), dim3(64), 0, s, p, nitems);
            else if (v == STAR_T0_G16F) hipLaunchKernelGGL(k_star32g16f, dim3(nb), dim3(64), 0, s, p, nitems);
            else if (v == STAR_T0_G16F4) hipLaunchKernelGGL(k_star32g16f4, dim3(nb), dim3(64), 0, s, p, nitems);
            else hipLaunchKernelGGL(k_star32g32f, dim3(nb), dim3(64), 0, s, p, nitems);
        }
    } else if (vmax == STAR_VMAX_T1) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_star<STAR_VMAX_T1, false>), dim3(nblocks), dim3(64), 0, s, p, nitems);
    } else if (vmax == STAR_VMAX_T2) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_star<STAR_VMAX_T2, false>), dim3(nblocks), dim3(64), 0, s, p, nitems);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------
// Global-memory star kernel: the same insertion / finalisation as k_star for stars whose
// link exceeds the LDS kernels' capacity.  One wave per star; the star's working arrays
// (32-bit local indices, cached cofactors per triangle slot) live in a per-block HBM
// workspace sized for the largest such star of the pass.
// ------------------------------------------------------------------------------------

static constexpr int GSTAR_NT = 512; // threads per star in the global-memory kernel

// Exclusive prefix sum of v over the work-group's threads in thread order; total in *tot.
// red: GSTAR_NT / 64 ints of LDS.  Every thread of the work-group must call it.
__device__ __forceinline__ int block_excl(int v, int *tot, int *red)
{
    const int lane = __lane_id(), w = threadIdx.x >> 6;
    const int inc = g4d_wave_incl_sum(v);
    if (lane == 63) red[w] = inc;
    __syncthreads();
    int base = 0, all = 0;
    for (int i = 0; i < GSTAR_NT / 64; i++) {
        const int x = red[i];
        if (i < w) base += x;
        all += x;
    }
    __syncthreads();
    *tot = all;
    return base + inc - v;
}

struct GWork {
    int *vid, *vstart, *vend, *ins, *ins_local, *vremap;
    double *vc;   // 4 per vertex
    int *tva;     // 3 local vertices per triangle slot
    uint32_t *tfl;
    int *to;      // 3 adjacency slots per triangle slot
    double *cof;  // 8 per triangle slot: cofactors, permanents
    int *freel, *be_a, *be_b, *be_u, *be_slot, *be_j, *remap;
    uint8_t *present;
};

__host__ __device__ inline int64_t gwork_layout(char *base, int vcap, GWork *w)
{
    const int64_t V = vcap, T = 2 * (int64_t)vcap;
    int64_t off = 0;
    auto take = [&](int64_t bytes) -> char * {
        char *r = base ? base + off : nullptr;
        off += (bytes + 255) & ~(int64_t)255;
        return r;
    };
    GWork t;
    t.vc = (double *)take(32 * V);
    t.cof = (double *)take(64 * T);
    t.vid = (int *)take(4 * V);
    t.vstart = (int *)take(4 * V);
    t.vend = (int *)take(4 * V);
    t.ins = (int *)take(4 * V);
    t.ins_local = (int *)take(4 * V);
    t.vremap = (int *)take(4 * V);
    t.tva = (int *)take(12 * T);
    t.tfl = (uint32_t *)take(4 * T);
    t.to = (int *)take(12 * T);
    t.freel = (int *)take(4 * T);
    t.be_a = (int *)take(4 * T);
    t.be_b = (int *)take(4 * T);
    t.be_u = (int *)take(4 * T);
    t.be_slot = (int *)take(4 * T);
    t.be_j = (int *)take(4 * T);
    t.remap = (int *)take(4 * T);
    t.present = (uint8_t *)take(V);
    if (w) *w = t;
    return off;
}

int64_t star_global_bytes(int vcap) { return gwork_layout(nullptr, vcap, nullptr); }

__global__ void __launch_bounds__(GSTAR_NT) k_star_global(StarParams p, const int *list, int nitems, char *ws,
                                                          int64_t ws_stride, int vcap)
{
    __shared__ int red[GSTAR_NT / 64];
    const int tid = threadIdx.x;
    unsigned long long *cnt = p.cnt;
    GWork L;
    gwork_layout(ws + (int64_t)blockIdx.x * ws_stride, vcap, &L);
    const int TMAX = 2 * vcap;

    for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
        const int s = list[item];
        const int64_t ib = p.ins_off[s];
        const int m = (int)(p.ins_off[s + 1] - ib);
        const int vn = p.mode ? p.vnum[s] : 0;
        const bool in_wide = vn >= STAR_WIDE_MIN;
        if (vn + m > vcap || (p.mode == 0 && m < 4)) {
            if (tid == 0) {
                atomicOr(&cnt[G4D_CNT_ERROR], (unsigned long long)(vn + m > vcap ? G4D_DEV_ERR_STAR_CAPACITY
                                                                                  : G4D_DEV_ERR_CONE));
                p.vnum[s] = 0;
                p.tnum[s] = 0;
            }
            continue;
        }
        if (p.mode == 1 && m == 0) continue; // untouched star: stays where it is
        const double sx = p.pts[3 * (int64_t)s], sy = p.pts[3 * (int64_t)s + 1], sz = p.pts[3 * (int64_t)s + 2];
        // insertion q takes link-vertex slot vn + q (mode 0: the first four form the simplex)
        const int q0 = p.mode ? 0 : 4;
        for (int q = tid; q < m; q += GSTAR_NT) {
            const int g = (int)(uint32_t)p.ins_keys[ib + q];
            L.ins[q] = g;
            L.ins_local[q] = -1;
            if (q >= q0) {
                L.vid[vn + q] = g;
                rel_coords(p.pts, g, sx, sy, sz, L.vc + 4 * (vn + q));
            }
        }
        int T, nfree = 0;
        if (p.mode == 1) {
            const int tn = p.tnum[s];
            const int64_t vi = p.voff[s], ti = p.toff[s];
            for (int v = tid; v < vn; v += GSTAR_NT) {
                const int g = p.vpool[vi + v];
                L.vid[v] = g;
                rel_coords(p.pts, g, sx, sy, sz, L.vc + 4 * v);
            }
            for (int t = tid; t < tn; t += GSTAR_NT) {
                int a[3], o[3];
                bool chk;
                if (in_wide) {
                    const uint64_t w0 = p.tpool[ti + 2 * (int64_t)t], w1 = p.tpool[ti + 2 * (int64_t)t + 1];
                    for (int i = 0; i < 3; i++) {
                        a[i] = wfield(w0, i);
                        o[i] = wfield(w1, i);
                    }
                    chk = (w0 >> 63) & 1u;
                } else {
                    const uint64_t h = p.tpool[ti + t];
                    for (int i = 0; i < 3; i++) {
                        a[i] = hvert(h, i);
                        o[i] = hopp(h, i);
                    }
                    chk = hchecked(h);
                }
                for (int i = 0; i < 3; i++) {
                    L.tva[3 * t + i] = a[i];
                    L.to[3 * t + i] = o[i];
                }
                L.tfl[t] = TF_ALIVE | (chk ? TF_CHECKED : 0u);
            }
            T = tn;
        } else {
            // initial 4-simplex (GD/GDelaunay/GDelPredKernels.cu:45-158)
            int l[4];
            for (int i = 0; i < 4; i++) l[i] = (int)(uint32_t)p.ins_keys[ib + i];
            const double *pp[5] = {p.pts + 3 * (int64_t)l[0], p.pts + 3 * (int64_t)l[1], p.pts + 3 * (int64_t)l[2],
                                   p.pts + 3 * (int64_t)l[3], p.pts + 3 * (int64_t)s};
            const int idx[5] = {l[0], l[1], l[2], l[3], s};
            const bool sw = orient4_full(pp, idx, cnt) < 0;
            if (sw) {
                const int t = l[0];
                l[0] = l[1];
                l[1] = t;
            }
            if (tid < 4) {
                L.vid[tid] = l[tid];
                rel_coords(p.pts, l[tid], sx, sy, sz, L.vc + 4 * tid);
                const int lt[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
                for (int i = 0; i < 3; i++) {
                    L.tva[3 * tid + i] = lt[tid][i];
                    L.to[3 * tid + i] = lt[tid][i];
                }
                L.tfl[tid] = TF_ALIVE;
                L.ins_local[tid] = (tid < 2 && sw) ? 1 - tid : tid;
            }
            T = 4;
        }
        __syncthreads();
        for (int slot = tid; slot < T; slot += GSTAR_NT)
            tri_cofactors(L.vc + 4 * L.tva[3 * slot], L.vc + 4 * L.tva[3 * slot + 1], L.vc + 4 * L.tva[3 * slot + 2],
                          L.cof + 8 * slot, L.cof + 8 * slot + 4);
        bool failed = false, unsure = false;
        for (int q = q0; q < m; q++) {
            const int gv = L.ins[q];
            const int Vn = vn + q;
            const double r[4] = {L.vc[4 * Vn], L.vc[4 * Vn + 1], L.vc[4 * Vn + 2], L.vc[4 * Vn + 3]};
            int nv = 0, nunc = 0;
            for (int slot = tid; slot < T; slot += GSTAR_NT) {
                const uint32_t f = L.tfl[slot];
                if (!(f & TF_ALIVE)) continue;
                const int sg = cofactor_filter_sign(r, L.cof + 8 * slot, L.cof + 8 * slot + 4);
                if (sg == 0) {
                    L.tfl[slot] = f | TF_UNSURE;
                    nunc++;
                } else if (sg < 0) {
                    L.tfl[slot] = f | TF_VIS;
                    nv++;
                }
            }
            // Triangles the filter cannot decide are gathered into one list (be_slot is free until
            // the hole boundary) and decided exactly one per thread, so that the block's threads
            // share the exact work instead of each wave stalling on its own few slots.  Below 2^15
            // slots one scan serves both counts (each total is at most T), packed in 16-bit halves.
            int vbase, nvis, nunc_tot, ub;
            if (T < 32768) {
                int both;
                const int pre = block_excl(nv | (nunc << 16), &both, red);
                vbase = pre & 0xffff;
                nvis = both & 0xffff;
                ub = pre >> 16;
                nunc_tot = both >> 16;
            } else {
                ub = block_excl(nunc, &nunc_tot, red);
                vbase = block_excl(nv, &nvis, red);
            }
            if (nunc_tot > 0) {
                for (int slot = tid; slot < T && nunc > 0; slot += GSTAR_NT)
                    if (L.tfl[slot] & TF_UNSURE) {
                        L.be_slot[ub++] = slot;
                        nunc--;
                    }
                __syncthreads();
                for (int k = tid; k < nunc_tot; k += GSTAR_NT) {
                    const int slot = L.be_slot[k];
                    const int a = L.vid[L.tva[3 * slot]], b = L.vid[L.tva[3 * slot + 1]], c = L.vid[L.tva[3 * slot + 2]];
                    const double *pp[5] = {p.pts + 3 * (int64_t)a, p.pts + 3 * (int64_t)b, p.pts + 3 * (int64_t)c,
                                           p.pts + 3 * (int64_t)s, p.pts + 3 * (int64_t)gv};
                    const int idx[5] = {a, b, c, s, gv};
                    const int sg = orient4_full(pp, idx, cnt);
                    const uint32_t f = L.tfl[slot] & ~TF_UNSURE;
                    L.tfl[slot] = sg < 0 ? (f | TF_VIS) : f;
                }
                __syncthreads();
                // each thread counts the beneath triangles among its own slots again
                nv = 0;
                for (int slot = tid; slot < T; slot += GSTAR_NT) nv += (L.tfl[slot] & TF_VIS) ? 1 : 0;
                vbase = block_excl(nv, &nvis, red);
            }
            if (nvis == 0) continue; // drowned at insertion
            if (tid == 0) L.ins_local[q] = Vn;
            __syncthreads();
            // hole boundary: edges of freed triangles whose neighbour survives
            int nb = 0;
            for (int slot = tid; slot < T; slot += GSTAR_NT)
                if (L.tfl[slot] & TF_VIS)
                    for (int i = 0; i < 3; i++) nb += (L.tfl[L.to[3 * slot + i]] & TF_VIS) ? 0 : 1;
            int nbtot;
            int e = block_excl(nb, &nbtot, red);
            if (nbtot > TMAX || nbtot < 3) {
                failed = true;
                break;
            }
            int fpos = nfree + vbase;
            for (int slot = tid; slot < T; slot += GSTAR_NT) {
                if (!(L.tfl[slot] & TF_VIS)) continue;
                for (int i = 0; i < 3; i++) {
                    const int u = L.to[3 * slot + i];
                    if (L.tfl[u] & TF_VIS) continue;
                    int j = 0;
                    while (j < 2 && L.to[3 * u + j] != slot) j++;
                    L.be_a[e] = L.tva[3 * slot + (i + 1) % 3];
                    L.be_b[e] = L.tva[3 * slot + (i + 2) % 3];
                    L.be_u[e] = u;
                    L.be_j[e] = j;
                    e++;
                }
                L.freel[fpos++] = slot;
            }
            __syncthreads();
            const int nfree_tot = nfree + nvis;
            for (int slot = tid; slot < T; slot += GSTAR_NT)
                if (L.tfl[slot] & TF_VIS) L.tfl[slot] = 0u;
            for (int k = tid; k < nbtot; k += GSTAR_NT)
                L.be_slot[k] = k < nfree_tot ? L.freel[nfree_tot - 1 - k] : T + (k - nfree_tot);
            const int newT = T + (nbtot > nfree_tot ? nbtot - nfree_tot : 0);
            if (newT > TMAX || Vn + 1 > vcap) {
                failed = true;
                break;
            }
            __syncthreads();
            for (int k = tid; k < nbtot; k += GSTAR_NT) {
                const int slot = L.be_slot[k], a = L.be_a[k], b = L.be_b[k], u = L.be_u[k];
                L.tva[3 * slot] = a;
                L.tva[3 * slot + 1] = b;
                L.tva[3 * slot + 2] = Vn;
                L.tfl[slot] = TF_ALIVE | TF_FRESH;
                L.to[3 * slot + 2] = u;
                L.to[3 * u + L.be_j[k]] = slot;
                L.vstart[a] = slot;
                L.vend[b] = slot;
            }
            __syncthreads();
            for (int k = tid; k < nbtot; k += GSTAR_NT) {
                const int slot = L.be_slot[k];
                L.to[3 * slot] = L.vstart[L.be_b[k]];
                L.to[3 * slot + 1] = L.vend[L.be_a[k]];
            }
            nfree = nfree_tot - (nbtot < nfree_tot ? nbtot : nfree_tot);
            T = newT;
            __syncthreads();
            for (int slot = tid; slot < T; slot += GSTAR_NT) {
                const uint32_t f = L.tfl[slot];
                if (f & TF_FRESH) {
                    tri_cofactors(L.vc + 4 * L.tva[3 * slot], L.vc + 4 * L.tva[3 * slot + 1],
                                  L.vc + 4 * L.tva[3 * slot + 2], L.cof + 8 * slot, L.cof + 8 * slot + 4);
                    L.tfl[slot] = f & ~TF_FRESH;
                }
            }
            __syncthreads();
        }
        if (unsure) {
            star_list_push(p.exact_list, p.exact_cap, &cnt[p.exact_counter], s, cnt);
            __syncthreads();
            continue;
        }
        if (failed) {
            if (tid == 0) {
                atomicOr(&cnt[G4D_CNT_ERROR], (unsigned long long)G4D_DEV_ERR_HOLE);
                p.vnum[s] = 0;
                p.tnum[s] = 0;
            }
            __syncthreads();
            continue;
        }
        // ---- finalize: link set, drowned insertions, compaction, pairs, history ----
        const int V = vn + m; // drowned insertions' slots are never referenced and drop out
        for (int v = tid; v < V; v += GSTAR_NT) L.present[v] = 0;
        __syncthreads();
        for (int slot = tid; slot < T; slot += GSTAR_NT)
            if (L.tfl[slot] & TF_ALIVE)
                for (int i = 0; i < 3; i++) L.present[L.tva[3 * slot + i]] = 1;
        __syncthreads();
        for (int q = tid; q < m; q += GSTAR_NT) {
            const int lv = L.ins_local[q];
            p.dflag[ib + q] = (lv < 0 || !L.present[lv]) ? 1 : 0;
        }
        const int64_t vo = p.nvoff[s], to = p.ntoff[s];
        int vcount = 0;
        for (int base = 0; base < V; base += GSTAR_NT) {
            const int v = base + tid;
            const bool pr = v < V && L.present[v];
            int tot;
            const int rk = block_excl(pr ? 1 : 0, &tot, red);
            if (pr) {
                L.vremap[v] = vcount + rk;
                p.vpool[vo + vcount + rk] = L.vid[v];
            }
            vcount += tot;
        }
        int tcount = 0;
        for (int base = 0; base < T; base += GSTAR_NT) {
            const int t = base + tid;
            const bool al = t < T && (L.tfl[t] & TF_ALIVE);
            int tot;
            const int rk = block_excl(al ? 1 : 0, &tot, red);
            if (al) L.remap[t] = tcount + rk;
            tcount += tot;
        }
        const bool wide = vcount >= STAR_WIDE_MIN;
        const int64_t words = wide ? 2 * (int64_t)tcount : tcount;
        if (words > star_tcap_words(vn + m) || tcount != 2 * vcount - 4) {
            if (tid == 0) {
                atomicOr(&cnt[G4D_CNT_ERROR], (unsigned long long)G4D_DEV_ERR_HOLE);
                p.vnum[s] = 0;
                p.tnum[s] = 0;
            }
            __syncthreads();
            continue;
        }
        __syncthreads();
        const int64_t pb = 2 * (to - p.tbase);
        int64_t npair = 0;
        for (int base = 0; base < T; base += GSTAR_NT) {
            const int t = base + tid;
            const uint32_t f = t < T ? L.tfl[t] : 0u;
            TriPairs pr{STAR_PAIR_NONE, STAR_PAIR_NONE};
            if (f & TF_ALIVE) {
                const int a = L.tva[3 * t], b = L.tva[3 * t + 1], c = L.tva[3 * t + 2];
                const int ra = L.vremap[a], rb = L.vremap[b], rc = L.vremap[c];
                const int o0 = L.remap[L.to[3 * t]], o1 = L.remap[L.to[3 * t + 1]], o2 = L.remap[L.to[3 * t + 2]];
                const int64_t d = L.remap[t];
                if (wide) {
                    p.tpool[to + 2 * d] = wpack3(ra, rb, rc) | (1ull << 63);
                    p.tpool[to + 2 * d + 1] = wpack3(o0, o1, o2);
                } else {
                    p.tpool[to + d] = hpack(ra, rb, rc, o0, o1, o2, true);
                }
                pr = star_tri_pairs(L.vid[a], L.vid[b], L.vid[c], !(f & TF_CHECKED));
            }
            // the star's pairs, dense, in triangle order
            const int np = (pr.x != STAR_PAIR_NONE ? 1 : 0) + (pr.y != STAR_PAIR_NONE ? 1 : 0);
            int tot;
            const int64_t at = npair + block_excl(np, &tot, red);
            if (np > 0) p.pairbuf[pb + at] = pr.x;
            if (np > 1) p.pairbuf[pb + at + 1] = pr.y;
            npair += tot;
        }
        if (tid == 0) p.pcount[s] = (int)npair;
        const int hn = star_history_merge(p, s, L.ins, m, tid, GSTAR_NT);
        __syncthreads();
        if (tid == 0) star_commit(p, s, vcount, vo, tcount, to, hn + m);
        __syncthreads();
    }
}

hipError_t launch_star_global(const StarParams &p, const int *list, int nitems, char *ws, int64_t ws_stride, int vcap,
                              int nblocks, hipStream_t s)
{
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_star_global, dim3(nblocks), dim3(GSTAR_NT), 0, s, p, list, nitems, ws, ws_stride, vcap);
    return hipGetLastError();
}

__global__ void k_need_max(StarParams p, const int *list, int nitems, unsigned long long *out)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nitems; i += gridDim.x * blockDim.x) {
        const int s = list[i];
        const int64_t need = (p.ins_off[s + 1] - p.ins_off[s]) + (p.mode ? p.vnum[s] : 0);
        atomicMax(out, (unsigned long long)need);
    }
}

hipError_t launch_need_max(const StarParams &p, const int *list, int nitems, unsigned long long *out, hipStream_t s)
{
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_need_max, dim3(g4d_blocks(nitems, 256, 1024)), dim3(256), 0, s, p, list, nitems, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------
// Containment proofs for drowned insertions (GD/GDelaunay/GDelPredKernels.cu:427-520).
// exVert is the lowest link vertex; the proof triangle is the unique live triangle
// without exVert whose three edges give one orient4(s, v, exVert, edge) sign.
// ------------------------------------------------------------------------------------

// EXACT = false: filters only, over all tasks; a task with an inconclusive orientation goes to
// p.xlist.  EXACT = true: the tasks listed in p.xlist, with exact predicates and SoS.
// Minimum over the GL lanes of a group (GL a power of two dividing 64).
template <int GL>
__device__ __forceinline__ int grp_min(int v)
{
    // minima within each 16-lane DPP row (rotations 8, 4, 2, 1); a whole-wave group then reads
    // the four rows' minima into scalars.  Every lane of the group must be active.
    static_assert(GL == 16 || GL == 64, "groups are DPP rows or the whole wave");
    int x = v;
    x = min(x, __builtin_amdgcn_update_dpp(x, x, 0x128, 0xf, 0xf, false));
    x = min(x, __builtin_amdgcn_update_dpp(x, x, 0x124, 0xf, 0xf, false));
    x = min(x, __builtin_amdgcn_update_dpp(x, x, 0x122, 0xf, 0xf, false));
    x = min(x, __builtin_amdgcn_update_dpp(x, x, 0x121, 0xf, 0xf, false));
    if constexpr (GL == 64)
        x = min(min(__builtin_amdgcn_readlane(x, 0), __builtin_amdgcn_readlane(x, 16)),
                min(__builtin_amdgcn_readlane(x, 32), __builtin_amdgcn_readlane(x, 48)));
    return x;
}

// One task per group of GL lanes: the filter pass takes 16 lanes per task (four tasks per
// wave; a star has about 27 link triangles, so a whole wave per task left most lanes idle),
// the exact pass a whole wave.  The lanes of a group stay together: every loop that ends in a
// group reduction has a trip count uniform over the group.
template <bool EXACT, int GL>
__global__ void __launch_bounds__(64) k_proofs(ProofParams p)
{
    constexpr int NG = 64 / GL;
    const int lane = threadIdx.x, g = lane / GL, gl = lane % GL;
    // the exact pass reads its list length from the counter the filter pass appended with
    // (nx < 0), so that no host round trip separates the two launches
    int64_t nt = p.ntasks;
    if (EXACT) {
        nt = p.nx;
        if (nt < 0) {
            const int64_t c = (int64_t)p.cnt[G4D_CNT_APPEND5];
            nt = c < p.xcap ? c : p.xcap;
        }
    }
    for (int64_t it = (int64_t)blockIdx.x * NG + g; it < nt; it += (int64_t)gridDim.x * NG) {
        const int64_t task = EXACT ? p.xlist[it] : it;
        const uint64_t tk = p.tasks[task];
        const int s = (int)(tk >> 32), v = (int)(uint32_t)tk;
        const int vn = p.vnum[s], tn = p.tnum[s];
        const bool wide = vn >= STAR_WIDE_MIN;
        const int64_t vo = p.voff[s], to = p.toff[s];
        int mn = 0x7fffffff;
        for (int i = gl; i < vn; i += GL) mn = min(mn, p.vpool[vo + i]);
        const int ex = grp_min<GL>(mn);
        int k0 = 0x7fffffff, k1 = 0x7fffffff, k2 = 0x7fffffff;
        bool unc = false;
        for (int t = gl; t < tn; t += GL) {
            int lv[3];
            star_tri_verts(p.tpool, to, t, wide, lv);
            const int a = p.vpool[vo + lv[0]], b = p.vpool[vo + lv[1]], c = p.vpool[vo + lv[2]];
            if (a == ex || b == ex || c == ex) continue;
            const int tri[3] = {a, b, c};
            int o[3];
            bool ok = true;
            for (int i = 0; i < 3 && ok; i++) {
                const int pv = tri[i], tv2 = tri[(i + 1) % 3];
                const double *pp[5] = {p.pts + 3 * (int64_t)s, p.pts + 3 * (int64_t)v, p.pts + 3 * (int64_t)ex,
                                       p.pts + 3 * (int64_t)pv, p.pts + 3 * (int64_t)tv2};
                if (EXACT) {
                    const int idx[5] = {s, v, ex, pv, tv2};
                    o[i] = orient4_full(pp, idx, p.cnt);
                } else {
                    o[i] = filter_insphere(pp);
                    if (o[i] == 0) {
                        unc = true;
                        ok = false;
                    }
                }
                if (i > 0 && o[i] != o[0]) ok = false;
            }
            if (!ok) continue;
            int x0 = a, x1 = b, x2 = c, tmp;
            if (x0 > x1) { tmp = x0; x0 = x1; x1 = tmp; }
            if (x1 > x2) { tmp = x1; x1 = x2; x2 = tmp; }
            if (x0 > x1) { tmp = x0; x0 = x1; x1 = tmp; }
            if (x0 < k0 || (x0 == k0 && (x1 < k1 || (x1 == k1 && x2 < k2)))) {
                k0 = x0;
                k1 = x1;
                k2 = x2;
            }
        }
        // the group's verdict: any inconclusive filter sends the task to the exact pass
        if (!EXACT && grp_min<GL>(unc ? 0 : 1) == 0) {
            if (gl == 0) {
                const unsigned long long k = atomicAdd(&p.cnt[G4D_CNT_APPEND5], 1ull);
                if ((int64_t)k < p.xcap) p.xlist[k] = task;
                else atomicOr(&p.cnt[G4D_CNT_ERROR], (unsigned long long)G4D_DEV_ERR_BUFFER);
            }
            continue;
        }
        const int b0 = grp_min<GL>(k0);
        const int b1 = grp_min<GL>(k0 == b0 ? k1 : 0x7fffffff);
        const int b2 = grp_min<GL>((k0 == b0 && k1 == b1) ? k2 : 0x7fffffff);
        if (gl == 0) {
            uint64_t *o = p.out + 4 * task;
            if (b0 == 0x7fffffff) {
                atomicOr(&p.cnt[G4D_CNT_ERROR], (unsigned long long)G4D_DEV_ERR_PROOF);
                o[0] = o[1] = o[2] = o[3] = ~0ull;
            } else {
                const int xs[4] = {ex, b0, b1, b2};
                for (int j = 0; j < 4; j++) {
                    const int a = v < xs[j] ? v : xs[j], b = v < xs[j] ? xs[j] : v;
                    o[j] = ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
                }
            }
        }
    }
}

hipError_t launch_proofs(const ProofParams &p, bool exact, hipStream_t s)
{
    const int64_t nt = exact ? (p.nx < 0 ? p.xcap : p.nx) : p.ntasks;
    if (nt <= 0) return hipSuccess;
    if (exact) {
        const int nb = (int)(nt < 1024 ? nt : 1024);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_proofs<true, 64>), dim3(nb), dim3(64), 0, s, p);
    } else {
        const int64_t nw = (nt + 3) / 4; // four tasks per wave
        const int nb = (int)(nw < 65536 ? nw : 65536);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_proofs<false, 16>), dim3(nb), dim3(64), 0, s, p);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------
// Drowned pairs, mutual removal (GD/GDelaunay/GDelHost.cu:832-875), history filter,
// reciprocation, capacities
// ------------------------------------------------------------------------------------

__global__ void k_drown_keys(const uint64_t *drown, int64_t m, uint64_t *keys)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t s = (uint32_t)(drown[i] >> 32), v = (uint32_t)drown[i];
        const uint32_t mn = s < v ? s : v, mx = s < v ? v : s;
        keys[i] = ((uint64_t)mn << 33) | ((uint64_t)mx << 1) | (uint64_t)(s == mn ? 0 : 1);
    }
}

hipError_t launch_drown_keys(const uint64_t *drown, int64_t m, uint64_t *keys, hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_drown_keys, dim3(g4d_blocks(m, 256, 65536)), dim3(256), 0, s, drown, m, keys);
    return hipGetLastError();
}

// tasks[i] = (star << 32 | vertex) of the drowned insertion i unless it is mutual, else
// STAR_TASK_NONE; the caller compacts (order-preserving, no atomics).
__global__ void k_nonmutual(const uint64_t *keys, int64_t m, uint64_t *tasks)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i] >> 1;
        const bool mutual = (i > 0 && (keys[i - 1] >> 1) == k) || (i + 1 < m && (keys[i + 1] >> 1) == k);
        const uint32_t mn = (uint32_t)(keys[i] >> 33), mx = (uint32_t)((keys[i] >> 1) & 0xffffffffu);
        const bool dir = keys[i] & 1;
        const uint32_t st = dir ? mx : mn, v = dir ? mn : mx;
        tasks[i] = mutual ? STAR_TASK_NONE : (((uint64_t)st << 32) | v);
    }
}

hipError_t launch_nonmutual(const uint64_t *keys, int64_t m, uint64_t *tasks, hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_nonmutual, dim3(g4d_blocks(m, 256, 65536)), dim3(256), 0, s, keys, m, tasks);
    return hipGetLastError();
}

__global__ void k_hist_filter(const uint64_t *pairs, int64_t m, const int *hpool, const int64_t *hoff, const int *hnum,
                              uint8_t *keep)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = pairs[i];
        const uint32_t a = (uint32_t)(k >> 32), b = (uint32_t)k;
        if (!(a < b)) {
            keep[i] = 0;
            continue;
        }
        const int *h = hpool + hoff[a];
        const int hn = hnum[a];
        int lo = 0, hi = hn;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if ((uint32_t)h[mid] < b) lo = mid + 1;
            else hi = mid;
        }
        keep[i] = (lo < hn && (uint32_t)h[lo] == b) ? 0 : 1;
    }
}

hipError_t launch_hist_filter(const uint64_t *pairs, int64_t m, const int *hpool, const int64_t *hoff, const int *hnum,
                              uint8_t *keep, hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_hist_filter, dim3(g4d_blocks(m, 256, 65536)), dim3(256), 0, s, pairs, m, hpool, hoff, hnum,
                       keep);
    return hipGetLastError();
}

__global__ void k_reciprocate(uint64_t *keys, int64_t m)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i];
        keys[m + i] = (k << 32) | (k >> 32);
    }
}

hipError_t launch_reciprocate(uint64_t *keys, int64_t m, hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reciprocate, dim3(g4d_blocks(m, 256, 65536)), dim3(256), 0, s, keys, m);
    return hipGetLastError();
}

// Sparse form of a pass's insertion offsets and active list, for sorted keys much fewer than the
// stars: the key that starts a star's run writes the star's first offset, the key that ends it
// the next one (equal to the following star's first when that star has insertions too), and
// hk[i] holds the star at a run's start (STAR_TASK_NONE elsewhere), so the order-preserving
// compaction of hk lists the stars with insertions in ascending order.  Offsets of stars
// without insertions are not written (nothing reads them).
__global__ void k_sparse_runs(const uint64_t *keys, int64_t m, int64_t *off, uint64_t *hk)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t st = (uint32_t)(keys[i] >> 32);
        const bool head = i == 0 || (uint32_t)(keys[i - 1] >> 32) != st;
        const bool tail = i == m - 1 || (uint32_t)(keys[i + 1] >> 32) != st;
        if (head) off[st] = i;
        if (tail) off[st + 1] = i + 1;
        hk[i] = head ? (uint64_t)st : STAR_TASK_NONE;
    }
}

hipError_t launch_sparse_runs(const uint64_t *keys, int64_t m, int64_t *off, uint64_t *hk, hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sparse_runs, dim3(g4d_blocks(m, 256, 65536)), dim3(256), 0, s, keys, m, off, hk);
    return hipGetLastError();
}

// dev_m (may be null): the count on the device, at most m.
__global__ void k_u64_to_i32(const uint64_t *in, int64_t m, const int64_t *dev_m, int *out)
{
    if (dev_m) m = min(m, *dev_m);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (int)in[i];
}

hipError_t launch_u64_to_i32(const uint64_t *in, int64_t m, const int64_t *dev_m, int *out, hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_u64_to_i32, dim3(g4d_blocks(m, 256, 65536)), dim3(256), 0, s, in, m, dev_m, out);
    return hipGetLastError();
}

__global__ void k_active_flags(const int64_t *ins_off, int n, uint32_t *flags)
{
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x)
        flags[s] = ins_off[s + 1] > ins_off[s] ? 1u : 0u;
}

hipError_t launch_active_flags(const int64_t *ins_off, int n, uint32_t *flags, hipStream_t s)
{
    hipLaunchKernelGGL(k_active_flags, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, ins_off, n, flags);
    return hipGetLastError();
}

__global__ void k_active_scatter(const uint32_t *flags, const uint32_t *pos, int n, int *list)
{
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x)
        if (flags[s]) list[pos[s]] = s;
}

hipError_t launch_active_scatter(const uint32_t *flags, const uint32_t *pos, int n, int *list, hipStream_t s)
{
    hipLaunchKernelGGL(k_active_scatter, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, flags, pos, n, list);
    return hipGetLastError();
}

// dev_n (may be null): the list's length on the device, at most nitems; the entries from there
// to nitems get zero capacities (so that scans over nitems total the list's needs).
__global__ void k_capacities(const int *list, int nitems, const int64_t *dev_n, const int *vnum, const int *hnum,
                             const int64_t *ins_off, int64_t *vcap, int64_t *tcap, int64_t *hcap, int *pcount)
{
    const int nl = dev_n ? (int)min((int64_t)nitems, *dev_n) : nitems;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nitems; i += gridDim.x * blockDim.x) {
        if (i >= nl) {
            vcap[i] = 0;
            tcap[i] = 0;
            hcap[i] = 0;
            continue;
        }
        const int s = list ? list[i] : i;
        pcount[s] = 0;
        const int64_t m = ins_off[s + 1] - ins_off[s];
        const int64_t v = (vnum ? vnum[s] : 0) + m;
        vcap[i] = v;
        tcap[i] = star_tcap_words(v);
        hcap[i] = (hnum ? hnum[s] : 0) + m;
    }
}

hipError_t launch_capacities(const int *list, int nitems, const int64_t *dev_n, const int *vnum, const int *hnum,
                             const int64_t *ins_off, int64_t *vcap, int64_t *tcap, int64_t *hcap, int *pcount,
                             hipStream_t s)
{
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_capacities, dim3(g4d_blocks(nitems, 256, 65536)), dim3(256), 0, s, list, nitems, dev_n, vnum,
                       hnum, ins_off, vcap, tcap, hcap, pcount);
    return hipGetLastError();
}

__global__ void k_pair_counts(const int *list, int nitems, const int *pcount, int64_t *cnt64)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nitems; i += gridDim.x * blockDim.x)
        cnt64[i] = pcount[list ? list[i] : i];
}

hipError_t launch_pair_counts(const int *list, int nitems, const int *pcount, int64_t *cnt64, hipStream_t s)
{
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pair_counts, dim3(g4d_blocks(nitems, 256, 65536)), dim3(256), 0, s, list, nitems, pcount,
                       cnt64);
    return hipGetLastError();
}

// One quarter-wave (16 lanes) per star copies its dense pairs.
__global__ void __launch_bounds__(256) k_pair_gather(const int *list, int nitems, const int *pcount,
                                                     const int64_t *off, const int64_t *ntoff, int64_t tbase,
                                                     const uint64_t *pairbuf, uint64_t *out)
{
    const int sub = threadIdx.x & 15;
    for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4; i < nitems;
         i += ((int64_t)gridDim.x * blockDim.x) >> 4) {
        const int s = list ? list[i] : (int)i;
        const int c = pcount[s];
        const uint64_t *src = pairbuf + 2 * (ntoff[s] - tbase);
        uint64_t *dst = out + off[i];
        for (int k = sub; k < c; k += 16) dst[k] = src[k];
    }
}

hipError_t launch_pair_gather(const int *list, int nitems, const int *pcount, const int64_t *off,
                              const int64_t *ntoff, int64_t tbase, const uint64_t *pairbuf, uint64_t *out,
                              hipStream_t s)
{
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pair_gather, dim3(g4d_blocks((int64_t)nitems * 16, 256, 65536)), dim3(256), 0, s, list,
                       nitems, pcount, off, ntoff, tbase, pairbuf, out);
    return hipGetLastError();
}

__global__ void k_assign_regions(const int *list, int nitems, const int64_t *vscan, const int64_t *tscan,
                                 const int64_t *hscan, int64_t vbase, int64_t tbase, int64_t hbase, int64_t *nvoff,
                                 int64_t *ntoff, int64_t *nhoff)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nitems; i += gridDim.x * blockDim.x) {
        const int s = list ? list[i] : i;
        nvoff[s] = vbase + vscan[i];
        ntoff[s] = tbase + tscan[i];
        nhoff[s] = hbase + hscan[i];
    }
}

hipError_t launch_assign_regions(const int *list, int nitems, const int64_t *vscan, const int64_t *tscan,
                                 const int64_t *hscan, int64_t vbase, int64_t tbase, int64_t hbase, int64_t *nvoff,
                                 int64_t *ntoff, int64_t *nhoff, hipStream_t s)
{
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_assign_regions, dim3(g4d_blocks(nitems, 256, 65536)), dim3(256), 0, s, list, nitems, vscan,
                       tscan, hscan, vbase, tbase, hbase, nvoff, ntoff, nhoff);
    return hipGetLastError();
}

__global__ void k_live_sizes(int n, const int *vnum, const int *tnum, const int *hnum, int64_t *vsz, int64_t *tsz,
                             int64_t *hsz)
{
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        const int v = vnum[s];
        vsz[s] = v;
        tsz[s] = v >= STAR_WIDE_MIN ? 2 * (int64_t)tnum[s] : tnum[s];
        hsz[s] = hnum[s];
    }
}

hipError_t launch_live_sizes(int n, const int *vnum, const int *tnum, const int *hnum, int64_t *vsz, int64_t *tsz,
                             int64_t *hsz, hipStream_t s)
{
    hipLaunchKernelGGL(k_live_sizes, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, n, vnum, tnum, hnum, vsz, tsz,
                       hsz);
    return hipGetLastError();
}

// One wave per star.
__global__ void __launch_bounds__(256) k_repack(int n, const int *vnum, const int64_t *voff, const int *tnum,
                                                const int64_t *toff, const int *hnum, const int64_t *hoff,
                                                const int *vpool, const uint64_t *tpool, const int *hpool,
                                                const int64_t *vdst, const int64_t *tdst, const int64_t *hdst,
                                                int *vpool2, uint64_t *tpool2, int *hpool2)
{
    const int lane = threadIdx.x & 63;
    const int w0 = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nw = (gridDim.x * blockDim.x) >> 6;
    for (int s = w0; s < n; s += nw) {
        const int v = vnum[s], h = hnum[s];
        const int64_t tw = v >= STAR_WIDE_MIN ? 2 * (int64_t)tnum[s] : tnum[s];
        const int64_t vi = voff[s], ti = toff[s], hi = hoff[s], vo = vdst[s], to = tdst[s], ho = hdst[s];
        for (int i = lane; i < v; i += 64) vpool2[vo + i] = vpool[vi + i];
        for (int64_t i = lane; i < tw; i += 64) tpool2[to + i] = tpool[ti + i];
        for (int i = lane; i < h; i += 64) hpool2[ho + i] = hpool[hi + i];
    }
}

hipError_t launch_repack(int n, const int *vnum, const int64_t *voff, const int *tnum, const int64_t *toff,
                         const int *hnum, const int64_t *hoff, const int *vpool, const uint64_t *tpool,
                         const int *hpool, const int64_t *vdst, const int64_t *tdst, const int64_t *hdst, int *vpool2,
                         uint64_t *tpool2, int *hpool2, hipStream_t s)
{
    hipLaunchKernelGGL(k_repack, dim3(g4d_blocks(n, 4, 65536)), dim3(256), 0, s, n, vnum, voff, tnum, toff, hnum, hoff,
                       vpool, tpool, hpool, vdst, tdst, hdst, vpool2, tpool2, hpool2);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------
// Output tetrahedra: owner triangles (star id below its three link vertices) whose exact
// orient3d(v0, v1, v2, star) is negative; upper-hull and flat ones are dropped.
// ------------------------------------------------------------------------------------

// Keep flag of an owner triangle decided by the exact orient3d (bit 58 of a packed word, bit
// 63 of the second word of a wide one); clear as written by the star kernels.
__device__ __forceinline__ uint64_t *out_keep_word(const OutParams &p, int64_t to, int t, bool wide)
{
    return p.tpool + to + (wide ? 2 * (int64_t)t + 1 : (int64_t)t);
}
__device__ __forceinline__ uint64_t out_keep_bit(bool wide) { return wide ? (1ull << 63) : (1ull << 58); }

// The output passes give each star a group of OUT_GL lanes, one lane per link triangle in turn
// (a star has ~27 link triangles): four stars per wave step, so that each wave keeps four
// stars' chains of dependent reads (star record, triangle words, link ids, points) in flight
// instead of two.
static constexpr int OUT_GL = 16;
static constexpr int OUT_NG = 64 / OUT_GL;

// Decision pass (coalesced reads of each star's triangle words): every owner triangle whose
// orient3d filter says "lower hull" gets its keep bit and is counted for its star;
// inconclusive ones go to the list for the exact pass.
__global__ void __launch_bounds__(256) k_out_flag(OutParams p, uint64_t *list, int64_t cap)
{
    const int lane = threadIdx.x & 63, grp = lane / OUT_GL, gl = lane % OUT_GL;
    const uint64_t gmask = ((OUT_GL == 64 ? 0ull : (1ull << OUT_GL)) - 1ull) << (grp * OUT_GL);
    const int w0 = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6), nw = (int)((gridDim.x * blockDim.x) >> 6);
    for (int s0 = OUT_NG * w0; s0 < p.n; s0 += OUT_NG * nw) {
        const int s = s0 + grp;
        const bool has = s < p.n;
        const int tn = has ? p.tnum[s] : 0;
        const bool wide = has && p.vnum[s] >= STAR_WIDE_MIN;
        const int64_t vo = has ? p.voff[s] : 0, to = has ? p.toff[s] : 0;
        const int tmax = wave_max_i(tn);
        int64_t c = 0;
        for (int t0 = 0; t0 < tmax; t0 += OUT_GL) {
            const int t = t0 + gl;
            int o = 1;
            if (t < tn) {
                int lv[3];
                star_tri_verts(p.tpool, to, t, wide, lv);
                const int a = p.vpool[vo + lv[0]], b = p.vpool[vo + lv[1]], cc = p.vpool[vo + lv[2]];
                if (s < a && s < b && s < cc) {
                    const double *pp[4] = {p.pts + 3 * (int64_t)a, p.pts + 3 * (int64_t)b, p.pts + 3 * (int64_t)cc,
                                           p.pts + 3 * (int64_t)s};
                    o = filter_orient3d(pp);
                    if (o < 0) *out_keep_word(p, to, t, wide) |= out_keep_bit(wide);
                }
            }
            c += __popcll(__ballot(o < 0) & gmask);
            const long long k = g4d_wave_append(o == 0, &p.cnt[G4D_CNT_APPEND5]);
            if (o == 0 && k < cap) list[k] = ((uint64_t)(uint32_t)s << 32) | (uint32_t)t;
        }
        if (has && gl == 0) p.star_tet_count[s] = c;
    }
}

// Exact pass over the inconclusive owner triangles: keep (and count) those with orient3d < 0.
__global__ void k_out_exact(OutParams p, const uint64_t *list, int64_t m)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const int s = (int)(list[i] >> 32), t = (int)(uint32_t)list[i];
        const bool wide = p.vnum[s] >= STAR_WIDE_MIN;
        const int64_t vo = p.voff[s], to = p.toff[s];
        int lv[3];
        star_tri_verts(p.tpool, to, t, wide, lv);
        const int a = p.vpool[vo + lv[0]], b = p.vpool[vo + lv[1]], cc = p.vpool[vo + lv[2]];
        const double *pp[4] = {p.pts + 3 * (int64_t)a, p.pts + 3 * (int64_t)b, p.pts + 3 * (int64_t)cc,
                               p.pts + 3 * (int64_t)s};
        if (exact_orient3d(pp, p.cnt) < 0) {
            *out_keep_word(p, to, t, wide) |= out_keep_bit(wide);
            atomicAdd(reinterpret_cast<unsigned long long *>(&p.star_tet_count[s]), 1ull);
        }
    }
}

// Write pass, a group of OUT_GL lanes per star: the kept triangles of each star at its scanned offset,
// in triangle order.
__global__ void __launch_bounds__(256) k_out_write(OutParams p)
{
    const int lane = threadIdx.x & 63, grp = lane / OUT_GL, gl = lane % OUT_GL;
    const uint64_t gmask = ((OUT_GL == 64 ? 0ull : (1ull << OUT_GL)) - 1ull) << (grp * OUT_GL);
    const int w0 = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6), nw = (int)((gridDim.x * blockDim.x) >> 6);
    const uint64_t below = ((1ull << lane) - 1ull) & gmask;
    for (int s0 = OUT_NG * w0; s0 < p.n; s0 += OUT_NG * nw) {
        const int s = s0 + grp;
        const bool has = s < p.n;
        const int tn = has ? p.tnum[s] : 0;
        const bool wide = has && p.vnum[s] >= STAR_WIDE_MIN;
        const int64_t vo = has ? p.voff[s] : 0, to = has ? p.toff[s] : 0;
        int64_t k = has ? p.star_tet_off[s] : 0;
        const int tmax = wave_max_i(tn);
        for (int t0 = 0; t0 < tmax; t0 += OUT_GL) {
            const int t = t0 + gl;
            const bool keep = t < tn && (*out_keep_word(p, to, t, wide) & out_keep_bit(wide));
            const uint64_t m = __ballot(keep);
            if (keep) {
                int lv[3];
                star_tri_verts(p.tpool, to, t, wide, lv);
                int4 q;
                q.x = p.vpool[vo + lv[0]];
                q.y = p.vpool[vo + lv[1]];
                q.z = p.vpool[vo + lv[2]];
                q.w = s;
                reinterpret_cast<int4 *>(p.tets)[k + __popcll(m & below)] = q;
            }
            k += __popcll(m & gmask);
        }
    }
}

hipError_t launch_out_flag(const OutParams &p, uint64_t *list, int64_t cap, hipStream_t s)
{
    hipLaunchKernelGGL(k_out_flag, dim3(g4d_blocks((int64_t)p.n * 64, 256, 16384)), dim3(256), 0, s, p, list, cap);
    return hipGetLastError();
}

hipError_t launch_out_exact(const OutParams &p, const uint64_t *list, int64_t m, hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_out_exact, dim3(g4d_blocks(m, 64, 1024)), dim3(64), 0, s, p, list, m);
    return hipGetLastError();
}

hipError_t launch_out_write(const OutParams &p, hipStream_t s)
{
    hipLaunchKernelGGL(k_out_write, dim3(g4d_blocks((int64_t)p.n * 64, 256, 16384)), dim3(256), 0, s, p);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------
// Star consistency (kerCheckStarConsistency / checkIfStarHasTriangle, GD/GDelaunay/
// GDelKernels.cu:1193-1271): link triangle (a, b, c) of star s stands for tetrahedron
// {s, a, b, c}, so star a must hold a link triangle with the vertex set {s, b, c}, and likewise
// stars b and c.  Vertex sets are compared without orientation, as the reference does.
// ------------------------------------------------------------------------------------

// Does star u hold a link triangle with vertex set {x, y, z} (global ids)?
__device__ bool star_has_triangle(const OutParams &p, int u, int x, int y, int z)
{
    if (u < 0 || u >= p.n) return false;
    const int vn = p.vnum[u], tn = p.tnum[u];
    const int64_t vo = p.voff[u], to = p.toff[u];
    int lx = -1, ly = -1, lz = -1;
    for (int v = 0; v < vn; v++) {
        const int g = p.vpool[vo + v];
        lx = g == x ? v : lx;
        ly = g == y ? v : ly;
        lz = g == z ? v : lz;
    }
    if (lx < 0 || ly < 0 || lz < 0) return false;
    const bool wide = vn >= STAR_WIDE_MIN;
    for (int t = 0; t < tn; t++) {
        int lv[3];
        star_tri_verts(p.tpool, to, t, wide, lv);
        const bool hx = lv[0] == lx || lv[1] == lx || lv[2] == lx;
        const bool hy = lv[0] == ly || lv[1] == ly || lv[2] == ly;
        const bool hz = lv[0] == lz || lv[1] == lz || lv[2] == lz;
        if (hx && hy && hz) return true;
    }
    return false;
}

// One half-wave per star, one lane per link triangle; cnt[G4D_CNT_APPEND5] counts the link
// triangles that fail, cnt[G4D_CNT_APPEND4] those checked.
__global__ void __launch_bounds__(256) k_star_consistency(OutParams p)
{
    const int lane = threadIdx.x & 63, half = lane >> 5, hl = lane & 31;
    const int w0 = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6), nw = (int)((gridDim.x * blockDim.x) >> 6);
    int bad = 0, seen = 0;
    for (int s0 = 2 * w0; s0 < p.n; s0 += 2 * nw) {
        const int s = s0 + half;
        const bool has = s < p.n;
        const int tn = has ? p.tnum[s] : 0;
        const bool wide = has && p.vnum[s] >= STAR_WIDE_MIN;
        const int64_t vo = has ? p.voff[s] : 0, to = has ? p.toff[s] : 0;
        const int tmax = max(tn, __shfl_xor(tn, 32, 64));
        for (int t0 = 0; t0 < tmax; t0 += 32) {
            const int t = t0 + hl;
            if (t >= tn) continue;
            int lv[3];
            star_tri_verts(p.tpool, to, t, wide, lv);
            const int g[3] = {p.vpool[vo + lv[0]], p.vpool[vo + lv[1]], p.vpool[vo + lv[2]]};
            bool ok = true;
            for (int i = 0; i < 3 && ok; i++) ok = star_has_triangle(p, g[i], s, g[(i + 1) % 3], g[(i + 2) % 3]);
            seen++;
            bad += ok ? 0 : 1;
        }
    }
    const int wb = g4d_wave_sum(bad), ws = g4d_wave_sum(seen);
    if (lane == 0) {
        if (wb) atomicAdd(&p.cnt[G4D_CNT_APPEND5], (unsigned long long)wb);
        atomicAdd(&p.cnt[G4D_CNT_APPEND4], (unsigned long long)ws);
    }
}

hipError_t launch_star_consistency(const OutParams &p, hipStream_t s)
{
    if (p.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_star_consistency, dim3(g4d_blocks((int64_t)p.n * 32, 256, 16384)), dim3(256), 0, s, p);
    return hipGetLastError();
}

__device__ __forceinline__ void face_of(const int *tt, int i, int *f)
{
    int q = 0;
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (j != i) f[q++] = tt[j];
    int tmp;
    if (f[0] > f[1]) { tmp = f[0]; f[0] = f[1]; f[1] = tmp; }
    if (f[1] > f[2]) { tmp = f[1]; f[1] = f[2]; f[2] = tmp; }
    if (f[0] > f[1]) { tmp = f[0]; f[0] = f[1]; f[1] = tmp; }
}

__global__ void k_adjacency(const int *tets, int64_t nt, int *opp, uint32_t *table, uint64_t mask,
                            unsigned long long *cnt)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += (int64_t)gridDim.x * blockDim.x) {
        const int *tt = tets + 4 * t;
        for (int i = 0; i < 4; i++) {
            int f[3];
            face_of(tt, i, f);
            uint64_t h = (uint64_t)(uint32_t)f[0] * 0x9E3779B97F4A7C15ull ^ (uint64_t)(uint32_t)f[1] * 0xC2B2AE3D27D4EB4Full ^
                         (uint64_t)(uint32_t)f[2] * 0x165667B19E3779F9ull;
            h ^= h >> 29;
            h *= 0xBF58476D1CE4E5B9ull;
            h ^= h >> 32;
            uint64_t slot = h & mask;
            const uint32_t mine = (uint32_t)((t << 2) | i) + 1u;
            for (uint64_t probe = 0; probe <= mask; probe++) {
                uint32_t cur = atomicCAS(&table[slot], 0u, mine);
                if (cur == 0u) break; // first occurrence stored
                const int64_t t2 = (int64_t)((cur - 1u) >> 2);
                const int i2 = (int)((cur - 1u) & 3u);
                int g[3];
                face_of(tets + 4 * t2, i2, g);
                if (g[0] == f[0] && g[1] == f[1] && g[2] == f[2]) {
                    if (atomicCAS(&opp[4 * t2 + i2], -1, (int)t) != -1)
                        atomicOr(&cnt[G4D_CNT_ERROR], (unsigned long long)G4D_DEV_ERR_FACE);
                    opp[4 * t + i] = (int)t2;
                    break;
                }
                slot = (slot + 1) & mask;
            }
        }
    }
}

__device__ __forceinline__ uint64_t face_hash(const int *f)
{
    uint64_t h = (uint64_t)(uint32_t)f[0] * 0x9E3779B97F4A7C15ull ^ (uint64_t)(uint32_t)f[1] * 0xC2B2AE3D27D4EB4Full ^
                 (uint64_t)(uint32_t)f[2] * 0x165667B19E3779F9ull;
    h ^= h >> 29;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 32;
    return h;
}

// Tiled variant for the pipeline's output, whose tetrahedra come grouped by owner star in
// Morton order: a work-group first pairs the faces of its ADJ_TETS consecutive tetrahedra in an
// LDS hash table (most interior faces have both tetrahedra in the same tile), and only the
// faces left unpaired go through the global table as in k_adjacency.  512-tetrahedron tiles
// (24.6 KB of LDS) measured faster than 1,024 and 2,048 at 4M points: residency wins over the
// extra global inserts of the smaller tiles.
static constexpr int ADJ_TETS = 512;
static constexpr int ADJ_HASH = 4096;

__global__ void __launch_bounds__(256) k_adjacency_tiled(const int *tets, int64_t nt, int *opp, uint32_t *table,
                                                         uint64_t mask, unsigned long long *cnt)
{
    __shared__ int4 lt[ADJ_TETS];
    __shared__ uint32_t tab[ADJ_HASH];
    __shared__ uint32_t done[ADJ_TETS * 4 / 32];
    for (int64_t tile = blockIdx.x; tile * ADJ_TETS < nt; tile += gridDim.x) {
        const int64_t t0 = tile * ADJ_TETS;
        const int nloc = (int)(nt - t0 < ADJ_TETS ? nt - t0 : ADJ_TETS);
        for (int k = threadIdx.x; k < nloc; k += blockDim.x) lt[k] = reinterpret_cast<const int4 *>(tets)[t0 + k];
        for (int k = threadIdx.x; k < ADJ_HASH; k += blockDim.x) tab[k] = 0u;
        for (int k = threadIdx.x; k < ADJ_TETS * 4 / 32; k += blockDim.x) done[k] = 0u;
        __syncthreads();
        for (int lf = threadIdx.x; lf < 4 * nloc; lf += blockDim.x) {
            const int k = lf >> 2, i = lf & 3;
            int f[3];
            face_of(reinterpret_cast<const int *>(&lt[k]), i, f);
            uint32_t h = (uint32_t)face_hash(f) & (ADJ_HASH - 1);
            for (int probe = 0; probe < ADJ_HASH; probe++) { // fewer faces than slots: ends
                const uint32_t cur = atomicCAS(&tab[h], 0u, (uint32_t)lf + 1u);
                if (cur == 0u) break;
                const int other = (int)cur - 1;
                int g[3];
                face_of(reinterpret_cast<const int *>(&lt[other >> 2]), other & 3, g);
                if (g[0] == f[0] && g[1] == f[1] && g[2] == f[2]) {
                    const uint32_t was = atomicOr(&done[other >> 5], 1u << (other & 31));
                    if (was & (1u << (other & 31))) atomicOr(&cnt[G4D_CNT_ERROR], (unsigned long long)G4D_DEV_ERR_FACE);
                    atomicOr(&done[lf >> 5], 1u << (lf & 31));
                    opp[4 * (t0 + (other >> 2)) + (other & 3)] = (int)(t0 + k);
                    opp[4 * (t0 + k) + i] = (int)(t0 + (other >> 2));
                    break;
                }
                h = (h + 1) & (ADJ_HASH - 1);
            }
        }
        __syncthreads();
        for (int lf = threadIdx.x; lf < 4 * nloc; lf += blockDim.x) {
            if (done[lf >> 5] & (1u << (lf & 31))) continue;
            const int k = lf >> 2, i = lf & 3;
            const int64_t t = t0 + k;
            int f[3];
            face_of(reinterpret_cast<const int *>(&lt[k]), i, f);
            uint64_t slot = face_hash(f) & mask;
            const uint32_t mine = (uint32_t)((t << 2) | i) + 1u;
            for (uint64_t probe = 0; probe <= mask; probe++) {
                const uint32_t cur = atomicCAS(&table[slot], 0u, mine);
                if (cur == 0u) break;
                const int64_t t2 = (int64_t)((cur - 1u) >> 2);
                const int i2 = (int)((cur - 1u) & 3u);
                int g[3];
                face_of(tets + 4 * t2, i2, g);
                if (g[0] == f[0] && g[1] == f[1] && g[2] == f[2]) {
                    if (atomicCAS(&opp[4 * t2 + i2], -1, (int)t) != -1)
                        atomicOr(&cnt[G4D_CNT_ERROR], (unsigned long long)G4D_DEV_ERR_FACE);
                    opp[4 * t + i] = (int)t2;
                    break;
                }
                slot = (slot + 1) & mask;
            }
        }
        __syncthreads();
    }
}

hipError_t launch_adjacency_tiled(const int *tets, int64_t nt, int *opp, uint32_t *table, int64_t table_size,
                                  unsigned long long *cnt, hipStream_t s)
{
    if (nt <= 0) return hipSuccess;
    HIP_CHECK_RET(hipMemsetAsync(table, 0, sizeof(uint32_t) * (size_t)table_size, s));
    HIP_CHECK_RET(hipMemsetAsync(opp, 0xff, sizeof(int) * 4 * (size_t)nt, s));
    const int64_t tiles = (nt + ADJ_TETS - 1) / ADJ_TETS;
    hipLaunchKernelGGL(k_adjacency_tiled, dim3((unsigned)(tiles < 65536 ? tiles : 65536)), dim3(256), 0, s, tets, nt,
                       opp, table, (uint64_t)(table_size - 1), cnt);
    return hipGetLastError();
}

hipError_t launch_adjacency(const int *tets, int64_t nt, int *opp, uint32_t *table, int64_t table_size,
                            unsigned long long *cnt, hipStream_t s)
{
    if (nt <= 0) return hipSuccess;
    HIP_CHECK_RET(hipMemsetAsync(table, 0, sizeof(uint32_t) * (size_t)table_size, s));
    HIP_CHECK_RET(hipMemsetAsync(opp, 0xff, sizeof(int) * 4 * (size_t)nt, s));
    hipLaunchKernelGGL(k_adjacency, dim3(g4d_blocks(nt, 256, 65536)), dim3(256), 0, s, tets, nt, opp, table,
                       (uint64_t)(table_size - 1), cnt);
    return hipGetLastError();
}

__global__ void k_map_perm(int *a, int64_t m, const int *perm)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
        a[i] = perm[a[i]];
}

__global__ void k_gather_i64(const int64_t *a, const int64_t *idx, int64_t m, int64_t *out)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = a[idx[i]];
}

hipError_t launch_gather_i64(const int64_t *a, const int64_t *idx, int64_t m, int64_t *out, hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_i64, dim3(g4d_blocks(m, 256, 65536)), dim3(256), 0, s, a, idx, m, out);
    return hipGetLastError();
}

__global__ void k_batch_localize(int *tets, int *opp, int64_t nt, const int *pt_set, const int64_t *pt_off,
                                 const int64_t *tet_off, const int *perm)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += (int64_t)gridDim.x * blockDim.x) {
        const int st = pt_set[tets[4 * t]];
        const int base = (int)pt_off[st];
        for (int i = 0; i < 4; i++) tets[4 * t + i] = perm[tets[4 * t + i]] - base;
        if (opp)
            for (int i = 0; i < 4; i++) {
                const int o = opp[4 * t + i];
                opp[4 * t + i] = o < 0 ? -1 : (int)(o - tet_off[st]);
            }
    }
}

hipError_t launch_batch_localize(int *tets, int *opp, int64_t nt, const int *pt_set, const int64_t *pt_off,
                                 const int64_t *tet_off, const int *perm, hipStream_t s)
{
    if (nt <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_localize, dim3(g4d_blocks(nt, 256, 65536)), dim3(256), 0, s, tets, opp, nt, pt_set,
                       pt_off, tet_off, perm);
    return hipGetLastError();
}

hipError_t launch_map_perm(int *tets, int64_t m4, const int *perm, hipStream_t s)
{
    if (m4 <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_map_perm, dim3(g4d_blocks(m4, 256, 65536)), dim3(256), 0, s, tets, m4, perm);
    return hipGetLastError();
}

} // namespace g4d
