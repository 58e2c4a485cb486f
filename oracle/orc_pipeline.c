/*
 * gStar4D CPU oracle: the triangulation pipeline — TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 * Restates, stage by stage, the reference path
 *   gdelInit  (GD/GDelaunay/GDelaunay.cpp:47-60)  -> Morton sort   GD/GDelaunay/GDelData.cu:51-131
 *   doPba     (GD/PBA/Pba.cu:105-122)             -> PBA           GD/PBA/pba3DHost.cu:200-295
 *   makeStarsFromGrid (GD/GDelaunay/GDelHost.cu:746-810)
 *   processFacets     (GD/GDelaunay/GDelHost.cu:1389-1470)
 *   makeTetraFromStars(GD/GDelaunay/GDelHost.cu:1617-1623)
 * sequentially on the host.  Deterministic (min-index) choices replace the reference's
 * storage-order / race dependent ones; see orc.h.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <omp.h>

#include "orc.h"

#define MARKER (-1)
#define ENC(x, y, z) (((x) << 20) | ((y) << 10) | (z))
#define DEC_X(v) ((v) >> 20)
#define DEC_Y(v) (((v) >> 10) & 0x3ff)
#define DEC_Z(v) ((v)&0x3ff)

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void *xmalloc(size_t n)
{
    void *p = malloc(n ? n : 1);
    if (!p) {
        fprintf(stderr, "gstar4d oracle: out of memory (%zu bytes)\n", n);
        abort();
    }
    return p;
}
static void *xrealloc(void *p, size_t n)
{
    void *q = realloc(p, n ? n : 1);
    if (!q) {
        fprintf(stderr, "gstar4d oracle: out of memory (%zu bytes)\n", n);
        abort();
    }
    return q;
}

/* LSD radix sort of uint64 keys on the low `bits` bits (stable). */
static void radix_u64(uint64_t *a, int64_t n, int bits)
{
    if (n < 2) return;
    uint64_t *tmp = xmalloc(sizeof(uint64_t) * n);
    uint64_t *src = a, *dst = tmp;
    for (int shift = 0; shift < bits; shift += 11) {
        int64_t cnt[2048 + 1];
        memset(cnt, 0, sizeof(cnt));
        for (int64_t i = 0; i < n; i++) cnt[((src[i] >> shift) & 2047) + 1]++;
        for (int d = 0; d < 2048; d++) cnt[d + 1] += cnt[d];
        for (int64_t i = 0; i < n; i++) dst[cnt[(src[i] >> shift) & 2047]++] = src[i];
        uint64_t *t = src; src = dst; dst = t;
    }
    if (src != a) memcpy(a, src, sizeof(uint64_t) * n);
    free(tmp);
}

static int64_t unique_u64(uint64_t *a, int64_t n)
{
    if (n == 0) return 0;
    int64_t m = 1;
    for (int64_t i = 1; i < n; i++)
        if (a[i] != a[m - 1]) a[m++] = a[i];
    return m;
}

static int bits_for(int64_t n)
{
    int b = 1;
    while (((int64_t)1 << b) < n) b++;
    return b;
}

/* ------------------------------------------------------------------------------------ */
/* B. Scaling and Morton order                                                          */
/* ------------------------------------------------------------------------------------ */

/* Grid units as in Application::_scalePoint (GD/Main/Application.cpp:589-603): translate by
   the minimum coordinate, scale the full coordinate range onto [1, G-2]; all in FP64 here. */
void orc_scale_to_voxels(const double *xyz, int32_t n, int32_t grid, int32_t *vox)
{
    double mn = xyz[0], mx = xyz[0];
    for (int64_t i = 0; i < 3 * (int64_t)n; i++) {
        if (xyz[i] < mn) mn = xyz[i];
        if (xyz[i] > mx) mx = xyz[i];
    }
    const double range = mx - mn;
    const double span = (double)(grid - 3);
    for (int64_t i = 0; i < 3 * (int64_t)n; i++) {
        const double t = span * (xyz[i] - mn);
        const double v = t / range + 1.0;
        int iv = (int)v;
        if (iv < 1) iv = 1;
        if (iv > grid - 2) iv = grid - 2;
        vox[i] = iv;
    }
}

/* Morton code: bit 3b = x_b, 3b+1 = y_b, 3b+2 = z_b (GD/GDelaunay/GDelData.cu:51-88, which
   interleaves 10 bits per axis); 11 bits per axis here so that grids up to 2048 order the same
   way (the eleventh bits are zero below 1024). */
static uint64_t morton33(int x, int y, int z)
{
    uint64_t m = 0;
    for (int b = 0; b < 11; b++) {
        m |= (uint64_t)((x >> b) & 1) << (3 * b);
        m |= (uint64_t)((y >> b) & 1) << (3 * b + 1);
        m |= (uint64_t)((z >> b) & 1) << (3 * b + 2);
    }
    return m;
}

int orc_morton_order(const double *xyz, int32_t n, int32_t grid, int32_t *perm, int32_t *vox)
{
    if (n < 1 || grid < 4 || grid > ORC_MAX_GRID) return -1;
    orc_scale_to_voxels(xyz, n, grid, vox);
    uint64_t *key = xmalloc(sizeof(uint64_t) * n);
    for (int32_t i = 0; i < n; i++)
        key[i] = (morton33(vox[3 * i], vox[3 * i + 1], vox[3 * i + 2]) << 31) | (uint32_t)i;
    radix_u64(key, n, 64); /* (code, input index): a stable order by code */
    for (int32_t k = 0; k < n; k++) perm[k] = (int32_t)(key[k] & 0x7fffffffu);
    free(key);
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* C. PBA digital Voronoi: three axis passes of an exact 1D lower envelope.             */
/* Candidate k sits at axis position q_k with offset d_k (squared distance in the other  */
/* two axes); f_k(t) = (t - q_k)^2 + d_k.  Stack pruning (Maurer, GD/PBA/pba3DKernel.h:  */
/* 215-272) uses exact integer bisector comparisons; the colour walk (:419-487) advances  */
/* while the next candidate is at least as close, so ties go to the larger position.    */
/* ------------------------------------------------------------------------------------ */

static int dominated(int qu, int64_t du, int qv, int64_t dv, int qw, int64_t dw)
{
    /* v is useless if bisector(u,v) >= bisector(v,w) */
    const int64_t n_uv = ((int64_t)qv * qv + dv) - ((int64_t)qu * qu + du);
    const int64_t d_uv = qv - qu;
    const int64_t n_vw = ((int64_t)qw * qw + dw) - ((int64_t)qv * qv + dv);
    const int64_t d_vw = qw - qv;
    return n_uv * d_vw >= n_vw * d_uv;
}

/* cand_site[t] (MARKER if none), cand_d[t]; writes out[t] = winning site or MARKER. */
static void envelope_1d(int g, const int32_t *cand_site, const int64_t *cand_d, int32_t *out,
                        int32_t *sq, int64_t *sd, int32_t *ss)
{
    int m = 0;
    for (int q = 0; q < g; q++) {
        if (cand_site[q] == MARKER) continue;
        const int64_t d = cand_d[q];
        while (m >= 2 && dominated(sq[m - 2], sd[m - 2], sq[m - 1], sd[m - 1], q, d)) m--;
        sq[m] = q;
        sd[m] = d;
        ss[m] = cand_site[q];
        m++;
    }
    int idx = 0;
    for (int t = 0; t < g; t++) {
        if (m == 0) {
            out[t] = MARKER;
            continue;
        }
        while (idx + 1 < m) {
            const int64_t a = (int64_t)(t - sq[idx]) * (t - sq[idx]) + sd[idx];
            const int64_t b = (int64_t)(t - sq[idx + 1]) * (t - sq[idx + 1]) + sd[idx + 1];
            if (b <= a)
                idx++;
            else
                break;
        }
        out[t] = ss[idx];
    }
}

/* Site coordinate c (0 x, 1 y, 2 z) of a grid label: a site code, or (by_id) a point index
   into vox_sorted. */
static inline int64_t site_coord(int32_t s, int c, const int32_t *vox_sorted, int by_id)
{
    if (by_id) return vox_sorted[3 * (int64_t)s + c];
    return c == 0 ? DEC_X(s) : c == 1 ? DEC_Y(s) : DEC_Z(s);
}

/* The three sweeps over a grid of site labels (MARKER where empty).  Only site positions enter
   the envelope decisions, so the result does not depend on how sites are labelled. */
static void pba_sweeps(int32_t g, int32_t *codes, const int32_t *vox_sorted, int by_id)
{
    const int64_t G = g;
    /* the columns of one pass are independent: each thread sweeps whole columns with
       scratch of its own */
#pragma omp parallel
    {
        int32_t *cs = xmalloc(sizeof(int32_t) * g), *out = xmalloc(sizeof(int32_t) * g);
        int64_t *cd = xmalloc(sizeof(int64_t) * g);
        int32_t *sq = xmalloc(sizeof(int32_t) * g), *ss = xmalloc(sizeof(int32_t) * g);
        int64_t *sd = xmalloc(sizeof(int64_t) * g);
        /* Z pass: column (x,y), candidates are the sites in the column (offset 0) */
#pragma omp for schedule(static)
        for (int64_t y = 0; y < G; y++)
            for (int64_t x = 0; x < G; x++) {
                for (int64_t z = 0; z < G; z++) {
                    cs[z] = codes[(z * G + y) * G + x];
                    cd[z] = 0;
                }
                envelope_1d(g, cs, cd, out, sq, sd, ss);
                for (int64_t z = 0; z < G; z++) codes[(z * G + y) * G + x] = out[z];
            }
        /* Y pass: column (x,z); candidate at row y is the Z-pass site of (x,y,z) */
#pragma omp for schedule(static)
        for (int64_t z = 0; z < G; z++)
            for (int64_t x = 0; x < G; x++) {
                for (int64_t y = 0; y < G; y++) {
                    const int32_t s = codes[(z * G + y) * G + x];
                    cs[y] = s;
                    if (s != MARKER) {
                        const int64_t dz = z - site_coord(s, 2, vox_sorted, by_id);
                        cd[y] = dz * dz;
                    }
                }
                envelope_1d(g, cs, cd, out, sq, sd, ss);
                for (int64_t y = 0; y < G; y++) codes[(z * G + y) * G + x] = out[y];
            }
        /* X pass: row (y,z); candidate at x is the Y-pass site of (x,y,z) */
#pragma omp for schedule(static)
        for (int64_t z = 0; z < G; z++)
            for (int64_t y = 0; y < G; y++) {
                for (int64_t x = 0; x < G; x++) {
                    const int32_t s = codes[(z * G + y) * G + x];
                    cs[x] = s;
                    if (s != MARKER) {
                        const int64_t dy = y - site_coord(s, 1, vox_sorted, by_id),
                                      dz = z - site_coord(s, 2, vox_sorted, by_id);
                        cd[x] = dy * dy + dz * dz;
                    }
                }
                envelope_1d(g, cs, cd, out, sq, sd, ss);
                for (int64_t x = 0; x < G; x++) codes[(z * G + y) * G + x] = out[x];
            }
        free(cs); free(out); free(cd); free(sq); free(ss); free(sd);
    }
}

int orc_pba_codes(const int32_t *vox_sorted, int32_t n, int32_t g, int32_t *codes)
{
    if (g < 4 || g > 1024) return -1;
    const int64_t G = g, G3 = G * G * G;
    for (int64_t i = 0; i < G3; i++) codes[i] = MARKER;
    for (int32_t i = 0; i < n; i++) {
        const int x = vox_sorted[3 * i], y = vox_sorted[3 * i + 1], z = vox_sorted[3 * i + 2];
        codes[(z * G + y) * G + x] = ENC(x, y, z);
    }
    pba_sweeps(g, codes, vox_sorted, 0);
    return 0;
}

int orc_pba(const int32_t *vox_sorted, int32_t n, int32_t g, int32_t *out_grid)
{
    if (g < 4 || g > ORC_MAX_GRID) return -1;
    const int64_t G = g, G3 = G * G * G;
    /* leader of a voxel = lowest Morton-sorted index of the points in it (the reference's
       racy last-writer scatter GD/PBA/pba3DKernel.h:518-542 made deterministic); the sweeps
       carry leaders as labels */
    for (int64_t i = 0; i < G3; i++) out_grid[i] = MARKER;
    for (int32_t i = n - 1; i >= 0; i--) {
        const int64_t x = vox_sorted[3 * i], y = vox_sorted[3 * i + 1], z = vox_sorted[3 * i + 2];
        out_grid[(z * G + y) * G + x] = i;
    }
    pba_sweeps(g, out_grid, vox_sorted, 1);
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* D2. Voronoi-neighbour pairs from 2x2x2 voxel cubes (GD/GDelaunay/GDelKernels.cu:137-299) */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    uint64_t *a;
    int64_t n, cap;
} u64vec;

static void u64_push(u64vec *v, uint64_t x)
{
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 1024;
        v->a = xrealloc(v->a, sizeof(uint64_t) * v->cap);
    }
    v->a[v->n++] = x;
}

static inline uint64_t pair_key(int a, int b)
{
    return a < b ? (((uint64_t)(uint32_t)a << 32) | (uint32_t)b)
                 : (((uint64_t)(uint32_t)b << 32) | (uint32_t)a);
}

static inline int32_t grid_at(const int32_t *grid, int g, int x, int y, int z)
{
    if (x < 0 || y < 0 || z < 0 || x >= g || y >= g || z >= g) return MARKER;
    return grid[((int64_t)z * g + y) * g + x];
}

/* grabPair (GD/GDelaunay/GDelKernels.cu:155-176): skipped when the first value is Marker */
static int grab(uint64_t *out, int m, int a, int b)
{
    if (a == MARKER) return m;
    out[m] = pair_key(a, b);
    return m + 1;
}

/* The pairs of one 2x2x2 voxel cube (GD/GDelaunay/GDelKernels.cu:196-270): corner c holds the
   label of voxel (x + (c & 1), y + (c >> 1 & 1), z + (c >> 2)); root = corner 0, opp = corner
   7, link cycle 6, 2, 3, 1, 5, 4.  At most 6 * 5 + 1 pairs. */
int orc_cube_pairs(const int32_t *v, uint64_t *out)
{
    static const int link[7] = {6, 2, 3, 1, 5, 4, 6};
    const int root = v[0], opp = v[7];
    int m = 0;
    if (root == opp) return 0;
    const int has_marker = (root == MARKER) || (opp == MARKER);
    int has_quad = 0;
    int a = v[link[0]];
    for (int i = 0; i < 6; i++) {
        const int b = v[link[i + 1]];
        if (a != b && a != root && a != opp && b != root && b != opp) {
            m = grab(out, m, root, a);
            m = grab(out, m, opp, a);
            m = grab(out, m, root, b);
            m = grab(out, m, opp, b);
            m = grab(out, m, a, b);
            has_quad = 1;
        }
        a = b;
    }
    if (has_quad && !has_marker) m = grab(out, m, root, opp);
    return m;
}

static void read_grid_pairs_rows(const int32_t *grid, int g, int y, u64vec *out)
{
    uint64_t pk[32];
    for (int x = -1; x < g; x++)
        for (int z = -1; z < g; z++) {
            int32_t v[8];
            for (int c = 0; c < 8; c++)
                v[c] = grid_at(grid, g, x + (c & 1), y + ((c >> 1) & 1), z + ((c >> 2) & 1));
            const int m = orc_cube_pairs(v, pk);
            for (int k = 0; k < m; k++) u64_push(out, pk[k]);
        }
}

/* Every 2x2x2 cube, including the outer Marker layer; the pairs of each y slab of cubes are
   collected per thread and concatenated (the caller sorts them). */
static void read_grid_pairs(const int32_t *grid, int g, u64vec *out)
{
    const int nt = omp_get_max_threads();
    u64vec *part = calloc((size_t)nt, sizeof(u64vec));
    if (!part) abort();
#pragma omp parallel for schedule(dynamic, 4)
    for (int y = -1; y < g; y++) read_grid_pairs_rows(grid, g, y, &part[omp_get_thread_num()]);
    for (int t = 0; t < nt; t++) {
        for (int64_t k = 0; k < part[t].n; k++) u64_push(out, part[t].a[k]);
        free(part[t].a);
    }
    free(part);
}

int64_t orc_grid_pairs(const int32_t *grid, int32_t g, uint64_t *out, int64_t cap)
{
    u64vec v = {0};
    read_grid_pairs(grid, g, &v);
    radix_u64(v.a, v.n, 64);
    const int64_t m = unique_u64(v.a, v.n);
    if (m > cap) {
        free(v.a);
        return -m;
    }
    if (out) memcpy(out, v.a, sizeof(uint64_t) * m);
    free(v.a);
    return m;
}

/* ------------------------------------------------------------------------------------ */
/* E/F. Stars                                                                           */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    int v[3];
    int opp[3]; /* slot of the triangle across the edge opposite v[i] */
} otri;

enum { TRI_ALIVE = 1, TRI_CHECKED = 2 };

typedef struct {
    otri *t;
    unsigned char *fl;
    int nt, cap;
    int *hist; /* sorted vertex ids ever inserted (GD/GDelaunay/GDelData.h:113-125) */
    int nhist, caphist;
} ostar;

typedef struct {
    const double *pts; /* Morton-sorted FP64 points */
    int n;
    ostar *st;
} octx;

/* Per-thread scratch of the star loops (stars are independent, so each thread owns the
   stars it works on and only its own scratch). */
typedef struct {
    unsigned char *vis;
    int vis_cap;
    int *lbuf, lcap; /* sorted link of a star */
    int *wbuf, wcap; /* a star's insertion list */
} oscratch;

static void scratch_free(oscratch *w)
{
    free(w->vis);
    free(w->lbuf);
    free(w->wbuf);
    memset(w, 0, sizeof(*w));
}

static int *scratch_w(oscratch *w, int m)
{
    if (w->wcap < m) {
        w->wcap = m;
        w->wbuf = xrealloc(w->wbuf, sizeof(int) * m);
    }
    return w->wbuf;
}

static int orient4(const octx *c, int a, int b, int d, int e, int f)
{
    return orc_orient4_sos(c->pts, a, b, d, e, f, NULL);
}

static void star_reserve(ostar *S, int need)
{
    if (need <= S->cap) return;
    int nc = S->cap ? S->cap : 16;
    while (nc < need) nc *= 2;
    S->t = xrealloc(S->t, sizeof(otri) * nc);
    S->fl = xrealloc(S->fl, nc);
    S->cap = nc;
}

static int star_has_vertex(const ostar *S, int v)
{
    for (int t = 0; t < S->nt; t++)
        if ((S->fl[t] & TRI_ALIVE) && (S->t[t].v[0] == v || S->t[t].v[1] == v || S->t[t].v[2] == v))
            return 1;
    return 0;
}

/* Initial cone: 4-simplex of the star point and 4 workset vertices
   (GD/GDelaunay/GDelPredKernels.cu:45-158). */
static void star_make_cone(octx *c, int s, const int *w4)
{
    static const int LinkTri[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
    int l[4] = {w4[0], w4[1], w4[2], w4[3]};
    if (orient4(c, l[0], l[1], l[2], l[3], s) < 0) {
        int t = l[0]; l[0] = l[1]; l[1] = t;
    }
    ostar *S = &c->st[s];
    star_reserve(S, 4);
    S->nt = 4;
    for (int ti = 0; ti < 4; ti++) {
        for (int k = 0; k < 3; k++) S->t[ti].v[k] = l[LinkTri[ti][k]];
        S->fl[ti] = TRI_ALIVE;
    }
    /* adjacency by edge matching: edge opposite v[i] is (v[i+1], v[i+2]); its twin is the
       reversed directed edge in another triangle */
    for (int ti = 0; ti < 4; ti++)
        for (int i = 0; i < 3; i++) {
            const int a = S->t[ti].v[(i + 1) % 3], b = S->t[ti].v[(i + 2) % 3];
            S->t[ti].opp[i] = -1;
            for (int tj = 0; tj < 4; tj++) {
                if (tj == ti) continue;
                for (int k = 0; k < 3; k++)
                    if (S->t[tj].v[(k + 1) % 3] == b && S->t[tj].v[(k + 2) % 3] == a) S->t[ti].opp[i] = tj;
            }
        }
}

/* Beneath/beyond insertion of v into star s (GD/GDelaunay/GDelPredKernels.cu:233-420 and
   GD/GDelaunay/GDelKernels.cu:518-783): free every triangle (a,b,c) with
   orient4(a,b,c,s,v) < 0 and fan the hole boundary to v.  Returns 0 if v is drowned. */
static int star_insert(octx *c, oscratch *w, int s, int v)
{
    ostar *S = &c->st[s];
    if (w->vis_cap < S->nt) {
        w->vis_cap = 2 * S->nt + 64;
        w->vis = xrealloc(w->vis, w->vis_cap);
    }
    unsigned char *vis = w->vis;
    int nvis = 0;
    for (int t = 0; t < S->nt; t++) {
        vis[t] = 0;
        if (!(S->fl[t] & TRI_ALIVE)) continue;
        if (orient4(c, S->t[t].v[0], S->t[t].v[1], S->t[t].v[2], s, v) < 0) {
            vis[t] = 1;
            nvis++;
        }
    }
    if (nvis == 0) return 0;
    /* boundary edges */
    int nb = 0, bcap = 3 * nvis;
    int *ba = xmalloc(sizeof(int) * bcap), *bb = xmalloc(sizeof(int) * bcap);
    int *bu = xmalloc(sizeof(int) * bcap), *bj = xmalloc(sizeof(int) * bcap);
    int *slots = xmalloc(sizeof(int) * (bcap + nvis));
    int nfree = 0;
    int *freel = xmalloc(sizeof(int) * nvis);
    for (int t = 0; t < S->nt; t++) {
        if (!vis[t]) continue;
        freel[nfree++] = t;
        for (int i = 0; i < 3; i++) {
            const int u = S->t[t].opp[i];
            if (vis[u]) continue;
            int j = 0;
            while (j < 3 && S->t[u].opp[j] != t) j++;
            if (j == 3) {
                fprintf(stderr, "gstar4d oracle: broken star adjacency (star %d)\n", s);
                abort();
            }
            ba[nb] = S->t[t].v[(i + 1) % 3];
            bb[nb] = S->t[t].v[(i + 2) % 3];
            bu[nb] = u;
            bj[nb] = j;
            nb++;
        }
    }
    for (int t = 0; t < S->nt; t++)
        if (vis[t]) S->fl[t] = 0;
    star_reserve(S, S->nt + nb);
    int nt = S->nt;
    for (int k = 0; k < nb; k++) slots[k] = (k < nfree) ? freel[k] : nt++;
    S->nt = nt;
    for (int k = 0; k < nb; k++) {
        otri *T = &S->t[slots[k]];
        T->v[0] = ba[k];
        T->v[1] = bb[k];
        T->v[2] = v;
        T->opp[2] = bu[k];
        S->t[bu[k]].opp[bj[k]] = slots[k];
        S->fl[slots[k]] = TRI_ALIVE; /* new: unchecked */
    }
    for (int k = 0; k < nb; k++) {
        otri *T = &S->t[slots[k]];
        T->opp[0] = T->opp[1] = -1;
        for (int m = 0; m < nb; m++) {
            if (ba[m] == bb[k]) T->opp[0] = slots[m]; /* across (b, v) */
            if (bb[m] == ba[k]) T->opp[1] = slots[m]; /* across (v, a) */
        }
        if (T->opp[0] < 0 || T->opp[1] < 0) {
            fprintf(stderr, "gstar4d oracle: hole boundary is not a cycle (star %d)\n", s);
            abort();
        }
    }
    free(ba); free(bb); free(bu); free(bj); free(slots); free(freel);
    return 1;
}

/* Insertion order of a star's list (insert_order 1): nearest to the star point first, ties by
   id.  The star that results is the same in any order: its link is the unique hull of the
   lifted points as seen from the star point (SoS decides the degenerate cases consistently),
   a drowned insertion is a listed vertex absent from it, and its pairs come from its triangles
   that the pass created; the order changes only how many triangles are made and undone. */
typedef struct {
    double d;
    int v;
} odist;

static int cmp_dist(const void *a, const void *b)
{
    const odist *x = a, *y = b;
    if (x->d != y->d) return x->d < y->d ? -1 : 1;
    return (x->v > y->v) - (x->v < y->v);
}

static void order_nearest(const double *pts, int s, const int *in, int m, int *out)
{
    odist *t = xmalloc(sizeof(odist) * (size_t)(m > 0 ? m : 1));
    const double *ps = pts + 3 * (int64_t)s;
    for (int i = 0; i < m; i++) {
        const double *pv = pts + 3 * (int64_t)in[i];
        const double x = pv[0] - ps[0], y = pv[1] - ps[1], z = pv[2] - ps[2];
        t[i].d = x * x + y * y + z * z;
        t[i].v = in[i];
    }
    qsort(t, (size_t)m, sizeof(odist), cmp_dist);
    for (int i = 0; i < m; i++) out[i] = t[i].v;
    free(t);
}

static int cmp_int(const void *a, const void *b)
{
    const int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* sorted distinct link vertices of star s */
static int star_link(const ostar *S, int **buf, int *cap)
{
    int m = 0;
    if (*cap < 3 * S->nt + 1) {
        *cap = 3 * S->nt + 64;
        *buf = xrealloc(*buf, sizeof(int) * (*cap));
    }
    for (int t = 0; t < S->nt; t++)
        if (S->fl[t] & TRI_ALIVE)
            for (int k = 0; k < 3; k++) (*buf)[m++] = S->t[t].v[k];
    qsort(*buf, m, sizeof(int), cmp_int);
    int u = 0;
    for (int i = 0; i < m; i++)
        if (u == 0 || (*buf)[u - 1] != (*buf)[i]) (*buf)[u++] = (*buf)[i];
    return u;
}

static int in_sorted(const int *a, int n, int x)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < n && a[lo] == x;
}

/* Containment proof for a drowned insertion v of star s
   (GD/GDelaunay/GDelPredKernels.cu:427-520).  exVert is the lowest link vertex (the
   reference takes vertex 0 of the first live triangle in slot order); the proof triangle is
   the unique live triangle without exVert whose three edges all give the same
   orient4(s, v, exVert, edge) sign.  Writes 4 ordered pairs. */
static int star_proof(octx *c, int s, int v, int exv, uint64_t *out4)
{
    ostar *S = &c->st[s];
    int best = -1;
    int bkey[3] = {0, 0, 0};
    for (int t = 0; t < S->nt; t++) {
        if (!(S->fl[t] & TRI_ALIVE)) continue;
        const int *tv = S->t[t].v;
        if (tv[0] == exv || tv[1] == exv || tv[2] == exv) continue;
        int o0 = orient4(c, s, v, exv, tv[0], tv[1]);
        int o1 = orient4(c, s, v, exv, tv[1], tv[2]);
        if (o1 != o0) continue;
        int o2 = orient4(c, s, v, exv, tv[2], tv[0]);
        if (o2 != o0) continue;
        int k[3] = {tv[0], tv[1], tv[2]};
        qsort(k, 3, sizeof(int), cmp_int);
        if (best < 0 || k[0] < bkey[0] || (k[0] == bkey[0] && (k[1] < bkey[1] || (k[1] == bkey[1] && k[2] < bkey[2])))) {
            best = t;
            memcpy(bkey, k, sizeof(k));
        }
    }
    if (best < 0) return -1;
    out4[0] = pair_key(v, exv);
    for (int k = 0; k < 3; k++) out4[1 + k] = pair_key(v, S->t[best].v[k]);
    return 0;
}

static void hist_add(ostar *S, const int *vs, int m)
{
    /* merge sorted vs into sorted hist */
    if (S->nhist + m > S->caphist) {
        S->caphist = 2 * (S->nhist + m) + 8;
        S->hist = xrealloc(S->hist, sizeof(int) * S->caphist);
    }
    int i = S->nhist - 1, j = m - 1, k = S->nhist + m - 1;
    while (j >= 0) {
        if (i >= 0 && S->hist[i] > vs[j])
            S->hist[k--] = S->hist[i--];
        else
            S->hist[k--] = vs[j--];
    }
    S->nhist += m;
}

/* CSR of ordered pairs grouped by first element (pairs sorted). */
static void pairs_to_csr(const uint64_t *p, int64_t m, int n, int64_t *off)
{
    for (int i = 0; i <= n; i++) off[i] = 0;
    for (int64_t k = 0; k < m; k++) off[(p[k] >> 32) + 1]++;
    for (int i = 0; i < n; i++) off[i + 1] += off[i];
}

int orc_set_threads(int nthreads)
{
    if (nthreads < 1) nthreads = omp_get_num_procs();
    if (nthreads > 256) nthreads = 256;
    omp_set_num_threads(nthreads);
    return nthreads;
}

int64_t orc_triangulate(const double *xyz, int32_t n, const orc_options *opt, int32_t *perm_out,
                        int32_t *tets_out, int64_t cap, orc_stats *stats)
{
    return orc_triangulate_ex(xyz, n, opt, NULL, 0, perm_out, tets_out, cap, stats);
}

int64_t orc_triangulate_ex(const double *xyz, int32_t n, const orc_options *opt, const uint64_t *extra,
                           int64_t n_extra, int32_t *perm_out, int32_t *tets_out, int64_t cap, orc_stats *stats)
{
    orc_stats st;
    memset(&st, 0, sizeof(st));
    st.num_points = n;
    const int g = opt->grid_size;
    const int min_ws = opt->min_workset > 0 ? opt->min_workset : 8;
    if (n < 5 || g < 4 || g > ORC_MAX_GRID) return -1;
    orc_pred_counter_reset();
    double t0 = now_s();

    /* ---- B: Morton order, sorted points ---- */
    int32_t *perm = xmalloc(sizeof(int32_t) * n);
    int32_t *vox = xmalloc(sizeof(int32_t) * 3 * n);
    orc_morton_order(xyz, n, g, perm, vox);
    double *pts = xmalloc(sizeof(double) * 3 * n);
    int32_t *vs = xmalloc(sizeof(int32_t) * 3 * n);
    for (int32_t k = 0; k < n; k++)
        for (int c = 0; c < 3; c++) {
            pts[3 * k + c] = xyz[3 * (int64_t)perm[k] + c];
            vs[3 * k + c] = vox[3 * (int64_t)perm[k] + c];
        }
    free(vox);
    if (perm_out) memcpy(perm_out, perm, sizeof(int32_t) * n);
    /* duplicates: equal points share a voxel, so they are adjacent runs of equal voxels */
    for (int32_t k = 1; k < n; k++) {
        int32_t j = k - 1;
        while (j >= 0 && vs[3 * j] == vs[3 * k] && vs[3 * j + 1] == vs[3 * k + 1] && vs[3 * j + 2] == vs[3 * k + 2]) {
            if (pts[3 * j] == pts[3 * k] && pts[3 * j + 1] == pts[3 * k + 1] && pts[3 * j + 2] == pts[3 * k + 2]) {
                free(perm); free(pts); free(vs);
                return -2; /* duplicate points */
            }
            j--;
        }
    }
    double t1 = now_s();
    st.t_morton = t1 - t0;

    /* ---- C: PBA ---- */
    const int64_t G3 = (int64_t)g * g * g;
    int32_t *grid = xmalloc(sizeof(int32_t) * G3);
    orc_pba(vs, n, g, grid);
    double t2 = now_s();
    st.t_pba = t2 - t1;

    /* ---- D1: missing points (GD/GDelaunay/GDelKernels.cu:109-135) ---- */
    u64vec ws = {0};
    for (int32_t i = 0; i < n; i++) {
        const int32_t leader = grid[((int64_t)vs[3 * i + 2] * g + vs[3 * i + 1]) * g + vs[3 * i]];
        if (leader != i) {
            st.num_missing++;
            u64_push(&ws, ((uint64_t)i << 32) | (uint32_t)leader);
            u64_push(&ws, ((uint64_t)(uint32_t)leader << 32) | (uint32_t)i);
        }
    }
    /* ---- D2/D3: grid pairs, sort-unique ---- */
    {
        u64vec gp = {0};
        read_grid_pairs(grid, g, &gp);
        radix_u64(gp.a, gp.n, 64);
        gp.n = unique_u64(gp.a, gp.n);
        st.num_grid_pairs = gp.n;
        for (int64_t k = 0; k < gp.n; k++) {
            const uint32_t a = (uint32_t)(gp.a[k] >> 32), b = (uint32_t)gp.a[k];
            u64_push(&ws, ((uint64_t)a << 32) | b);
            u64_push(&ws, ((uint64_t)b << 32) | a);
        }
        free(gp.a);
    }
    free(grid);
    /* extra candidate pairs (a learned proposal, Morton index space), both directions; they
       change which vertices the stars start from, never the result */
    for (int64_t k = 0; k < n_extra; k++) {
        const uint32_t a = (uint32_t)(extra[k] >> 32), b = (uint32_t)extra[k];
        if (a == b || a >= (uint32_t)n || b >= (uint32_t)n) continue;
        u64_push(&ws, ((uint64_t)a << 32) | b);
        u64_push(&ws, ((uint64_t)b << 32) | a);
    }
    const int nb = bits_for(n);
    radix_u64(ws.a, ws.n, 32 + nb);
    ws.n = unique_u64(ws.a, ws.n);
    /* ---- D4: pumping (GD/GDelaunay/GDelHost.cu:352-529, GD/GDelaunay/GDelKernels.cu:347-477) ---- */
    int64_t *off = xmalloc(sizeof(int64_t) * (n + 1));
    pairs_to_csr(ws.a, ws.n, n, off);
    {
        u64vec pump = {0};
        int *cand = NULL;
        int ccap = 0;
        for (int s = 0; s < n; s++) {
            const int have = (int)(off[s + 1] - off[s]);
            if (have >= min_ws) continue;
            const int need = min_ws - have;
            if (ccap < need) {
                ccap = need;
                cand = xrealloc(cand, sizeof(int) * ccap);
            }
            int got = 0;
            /* neighbours' worksets, in ascending order */
            for (int64_t k = off[s]; k < off[s + 1] && got < need; k++) {
                const int nei = (int)(uint32_t)ws.a[k];
                for (int64_t q = off[nei]; q < off[nei + 1] && got < need; q++) {
                    const int cv = (int)(uint32_t)ws.a[q];
                    if (cv == s) continue;
                    int dup = 0;
                    for (int64_t r = off[s]; r < off[s + 1] && !dup; r++) dup = ((int)(uint32_t)ws.a[r] == cv);
                    for (int r = 0; r < got && !dup; r++) dup = (cand[r] == cv);
                    if (!dup) cand[got++] = cv;
                }
            }
            /* then vertex ids 0, 1, 2, ... (the reference's Q2 spin on star==id is skipped) */
            for (int cv = 0; got < need && cv < n; cv++) {
                if (cv == s) continue;
                int dup = 0;
                for (int64_t r = off[s]; r < off[s + 1] && !dup; r++) dup = ((int)(uint32_t)ws.a[r] == cv);
                for (int r = 0; r < got && !dup; r++) dup = (cand[r] == cv);
                if (!dup) cand[got++] = cv;
            }
            for (int r = 0; r < got; r++) {
                u64_push(&pump, ((uint64_t)(uint32_t)s << 32) | (uint32_t)cand[r]);
                u64_push(&pump, ((uint64_t)(uint32_t)cand[r] << 32) | (uint32_t)s);
            }
        }
        free(cand);
        radix_u64(pump.a, pump.n, 32 + nb);
        pump.n = unique_u64(pump.a, pump.n);
        st.num_pumped = (int)pump.n;
        for (int64_t k = 0; k < pump.n; k++) u64_push(&ws, pump.a[k]);
        free(pump.a);
        radix_u64(ws.a, ws.n, 32 + nb);
        ws.n = unique_u64(ws.a, ws.n);
        pairs_to_csr(ws.a, ws.n, n, off);
    }
    st.num_workset = ws.n;
    for (int s = 0; s < n; s++)
        if (off[s + 1] - off[s] > st.max_workset) st.max_workset = (int)(off[s + 1] - off[s]);

    /* ---- E: stars (each star is built from its own workset alone: one per thread) ---- */
    octx c;
    memset(&c, 0, sizeof(c));
    c.pts = pts;
    c.n = n;
    c.st = calloc(n, sizeof(ostar));
    if (!c.st) abort();
#pragma omp parallel
    {
        oscratch w = {0};
#pragma omp for schedule(dynamic, 256)
        for (int s = 0; s < n; s++) {
            const int m = (int)(off[s + 1] - off[s]);
            int *wb = scratch_w(&w, m);
            for (int k = 0; k < m; k++) wb[k] = (int)(uint32_t)ws.a[off[s] + k];
            const int *ord = wb;
            int *ob = NULL;
            if (opt->insert_order == 1) {
                ob = xmalloc(sizeof(int) * (size_t)m);
                order_nearest(pts, s, wb, m, ob);
                ord = ob;
            }
            star_make_cone(&c, s, ord);
            for (int k = 4; k < m; k++) star_insert(&c, &w, s, ord[k]);
            free(ob);
            ostar *S = &c.st[s];
            S->hist = xmalloc(sizeof(int) * (m + 8));
            S->caphist = m + 8;
            memcpy(S->hist, wb, sizeof(int) * m);
            S->nhist = m;
        }
        scratch_free(&w);
    }
    double t3 = now_s();
    st.t_star = t3 - t2;

    /* ---- F: splaying (GD/GDelaunay/GDelHost.cu:1389-1470) ---- */
    /* "last batch": CSR of the insertions of the previous step (initially the worksets) */
    uint64_t *batch = ws.a; /* sorted (star, vert) pairs */
    int64_t nbatch = ws.n;
    ws.a = NULL;
    int rounds = 0;
    for (;;) {
        if (rounds >= opt->max_rounds) {
            fprintf(stderr, "gstar4d oracle: splaying did not converge in %d rounds\n", rounds);
            break;
        }
        /* F1: drowned insertions (GD/GDelaunay/GDelKernels.cu:1492-1527): a flag per entry of
           the last batch, star groups in parallel, then an in-order compaction */
        u64vec drown = {0};
        {
            unsigned char *dfl = xmalloc((size_t)nbatch + 1);
            int64_t *gs = xmalloc(sizeof(int64_t) * ((size_t)nbatch + 1));
            int64_t ng = 0;
            for (int64_t k = 0; k < nbatch; k++)
                if (k == 0 || (batch[k] >> 32) != (batch[k - 1] >> 32)) gs[ng++] = k;
            gs[ng] = nbatch;
#pragma omp parallel
            {
                oscratch w = {0};
#pragma omp for schedule(dynamic, 256)
                for (int64_t q = 0; q < ng; q++) {
                    const int s = (int)(batch[gs[q]] >> 32);
                    const int nl = star_link(&c.st[s], &w.lbuf, &w.lcap);
                    for (int64_t e = gs[q]; e < gs[q + 1]; e++)
                        dfl[e] = !in_sorted(w.lbuf, nl, (int)(uint32_t)batch[e]);
                }
                scratch_free(&w);
            }
            for (int64_t k = 0; k < nbatch; k++)
                if (dfl[k]) u64_push(&drown, batch[k]);
            free(dfl);
            free(gs);
        }
        /* F2: remove mutually drowned pairs (GD/GDelaunay/GDelHost.cu:832-875) */
        u64vec ins = {0};
        {
            uint64_t *key = xmalloc(sizeof(uint64_t) * (drown.n + 1));
            for (int64_t k = 0; k < drown.n; k++) {
                const uint32_t s = (uint32_t)(drown.a[k] >> 32), v = (uint32_t)drown.a[k];
                key[k] = pair_key((int)s, (int)v);
            }
            radix_u64(key, drown.n, 64);
            /* F3: a containment proof for every drowned insertion whose pair is drowned one
               way only; proofs land in per-entry slots, then in order into `ins` */
            uint64_t *proof = xmalloc(sizeof(uint64_t) * 4 * (drown.n + 1));
            unsigned char *has = xmalloc((size_t)drown.n + 1);
            int failed = 0;
#pragma omp parallel
            {
                oscratch w = {0};
#pragma omp for schedule(dynamic, 64)
                for (int64_t k = 0; k < drown.n; k++) {
                    const uint32_t s = (uint32_t)(drown.a[k] >> 32), v = (uint32_t)drown.a[k];
                    const uint64_t pk = pair_key((int)s, (int)v);
                    /* binary search count of pk in key[] */
                    int64_t lo = 0, hi = drown.n;
                    while (lo < hi) {
                        int64_t mid = (lo + hi) / 2;
                        if (key[mid] < pk) lo = mid + 1; else hi = mid;
                    }
                    const int mutual = (lo + 1 < drown.n && key[lo + 1] == pk);
                    has[k] = 0;
                    if (mutual) continue;
                    const int nl = star_link(&c.st[s], &w.lbuf, &w.lcap);
                    if (nl == 0 || star_proof(&c, (int)s, (int)v, w.lbuf[0], proof + 4 * k) != 0) {
#pragma omp critical
                        {
                            fprintf(stderr, "gstar4d oracle: no containment proof (star %u vert %u)\n", s, v);
                            failed = 1;
                        }
                        continue;
                    }
                    has[k] = 1;
                }
                scratch_free(&w);
            }
            if (failed) abort();
            for (int64_t k = 0; k < drown.n; k++)
                if (has[k])
                    for (int q = 0; q < 4; q++) u64_push(&ins, proof[4 * k + q]);
            free(has);
            free(proof);
            free(key);
        }
        free(drown.a);
        /* F4: pairs from unchecked triangles (GD/GDelaunay/GDelKernels.cu:818-939) */
        for (int s = 0; s < n; s++) {
            ostar *S = &c.st[s];
            for (int t = 0; t < S->nt; t++) {
                if ((S->fl[t] & TRI_ALIVE) && !(S->fl[t] & TRI_CHECKED)) {
                    for (int i = 0; i < 3; i++) {
                        const int a = S->t[t].v[i], b = S->t[t].v[(i + 1) % 3];
                        if (a < b) u64_push(&ins, ((uint64_t)(uint32_t)a << 32) | (uint32_t)b);
                    }
                    S->fl[t] |= TRI_CHECKED;
                }
            }
        }
        /* F5: sort-unique, history filter, reciprocate */
        radix_u64(ins.a, ins.n, 32 + nb);
        ins.n = unique_u64(ins.a, ins.n);
        int64_t m = 0;
        for (int64_t k = 0; k < ins.n; k++) {
            const int a = (int)(ins.a[k] >> 32), b = (int)(uint32_t)ins.a[k];
            if (!in_sorted(c.st[a].hist, c.st[a].nhist, b)) ins.a[m++] = ins.a[k];
        }
        ins.n = m;
        free(batch);
        if (m == 0) {
            free(ins.a);
            batch = NULL;
            break;
        }
        rounds++;
        for (int64_t k = 0; k < m; k++) {
            const uint32_t a = (uint32_t)(ins.a[k] >> 32), b = (uint32_t)ins.a[k];
            u64_push(&ins, ((uint64_t)b << 32) | a);
        }
        radix_u64(ins.a, ins.n, 32 + nb);
        st.splay_insertions += ins.n;
        /* F6-F9: insert per star in ascending vertex order; extend history (star groups in
           parallel: a group touches only its own star) */
        {
            int64_t *gs = xmalloc(sizeof(int64_t) * ((size_t)ins.n + 1));
            int64_t ng = 0;
            for (int64_t k = 0; k < ins.n; k++)
                if (k == 0 || (ins.a[k] >> 32) != (ins.a[k - 1] >> 32)) gs[ng++] = k;
            gs[ng] = ins.n;
#pragma omp parallel
            {
                oscratch w = {0};
#pragma omp for schedule(dynamic, 64)
                for (int64_t q = 0; q < ng; q++) {
                    const int s = (int)(ins.a[gs[q]] >> 32);
                    const int cnt = (int)(gs[q + 1] - gs[q]);
                    int *wb = scratch_w(&w, cnt);
                    for (int r = 0; r < cnt; r++) wb[r] = (int)(uint32_t)ins.a[gs[q] + r];
                    if (opt->insert_order == 1) {
                        int *ob = xmalloc(sizeof(int) * (size_t)cnt);
                        order_nearest(pts, s, wb, cnt, ob);
                        for (int r = 0; r < cnt; r++) star_insert(&c, &w, s, ob[r]);
                        free(ob);
                    } else {
                        for (int r = 0; r < cnt; r++) star_insert(&c, &w, s, wb[r]);
                    }
                    hist_add(&c.st[s], wb, cnt);
                }
                scratch_free(&w);
            }
            free(gs);
        }
        batch = ins.a;
        nbatch = ins.n;
    }
    st.splay_rounds = rounds;
    double t4 = now_s();
    st.t_splay = t4 - t3;

    /* ---- G: tetra from owner triangles, lower hull by exact orient3d
       (GD/GDelaunay/GDelKernels.cu:1097-1191, GD/GDelaunay/GDelPredKernels.cu:628-665) ---- */
    int64_t nt = 0;
    for (int s = 0; s < n; s++) {
        ostar *S = &c.st[s];
        for (int t = 0; t < S->nt; t++) {
            if (!(S->fl[t] & TRI_ALIVE)) continue;
            const int *tv = S->t[t].v;
            if (!(s < tv[0] && s < tv[1] && s < tv[2])) continue;
            st.num_lowup_tetra++;
            const int o = orc_orient3d(pts + 3 * (int64_t)tv[0], pts + 3 * (int64_t)tv[1],
                                       pts + 3 * (int64_t)tv[2], pts + 3 * (int64_t)s);
            if (o >= 0) continue; /* upper hull and flat tetra dropped */
            if (tets_out && nt < cap) {
                tets_out[4 * nt + 0] = tv[0];
                tets_out[4 * nt + 1] = tv[1];
                tets_out[4 * nt + 2] = tv[2];
                tets_out[4 * nt + 3] = s;
            }
            nt++;
        }
    }
    st.num_tetra = (int32_t)nt;
    double t5 = now_s();
    st.t_out = t5 - t4;
    st.orient4_calls = orc_pred_counter(0);
    st.exact_calls = orc_pred_counter(1);
    st.sos_calls = orc_pred_counter(2);

    for (int s = 0; s < n; s++) {
        free(c.st[s].t);
        free(c.st[s].fl);
        free(c.st[s].hist);
    }
    free(c.st);
    free(off);
    free(pts);
    free(vs);
    free(perm);
    if (stats) *stats = st;
    return nt;
}

/* ------------------------------------------------------------------------------------ */
/* Canonical form, adjacency, validator                                                 */
/* ------------------------------------------------------------------------------------ */

static int cmp_quad(const void *a, const void *b)
{
    const int32_t *x = a, *y = b;
    for (int i = 0; i < 4; i++)
        if (x[i] != y[i]) return x[i] < y[i] ? -1 : 1;
    return 0;
}

void orc_canonicalize(int32_t *tets, int64_t nt, const int32_t *perm)
{
    for (int64_t t = 0; t < nt; t++) {
        int32_t *q = tets + 4 * t;
        if (perm)
            for (int i = 0; i < 4; i++) q[i] = perm[q[i]];
        qsort(q, 4, sizeof(int32_t), cmp_int);
    }
    qsort(tets, nt, 4 * sizeof(int32_t), cmp_quad);
}

typedef struct {
    int32_t a, b, c;
    int32_t tet;
    int32_t face;
} oface;

static int cmp_face(const void *x, const void *y)
{
    const oface *p = x, *q = y;
    if (p->a != q->a) return p->a < q->a ? -1 : 1;
    if (p->b != q->b) return p->b < q->b ? -1 : 1;
    if (p->c != q->c) return p->c < q->c ? -1 : 1;
    return 0;
}

int orc_adjacency(const int32_t *tets, int64_t nt, int32_t n, int32_t *opp)
{
    (void)n;
    oface *f = xmalloc(sizeof(oface) * 4 * nt + 1);
    for (int64_t t = 0; t < nt; t++)
        for (int i = 0; i < 4; i++) {
            int k[3], q = 0;
            for (int j = 0; j < 4; j++)
                if (j != i) k[q++] = tets[4 * t + j];
            qsort(k, 3, sizeof(int), cmp_int);
            oface *e = &f[4 * t + i];
            e->a = k[0]; e->b = k[1]; e->c = k[2];
            e->tet = (int32_t)t;
            e->face = i;
            opp[4 * t + i] = -1;
        }
    qsort(f, 4 * nt, sizeof(oface), cmp_face);
    int bad = 0;
    for (int64_t k = 0; k < 4 * nt;) {
        int64_t e = k + 1;
        while (e < 4 * nt && cmp_face(&f[k], &f[e]) == 0) e++;
        if (e - k == 2) {
            opp[4 * f[k].tet + f[k].face] = f[k + 1].tet;
            opp[4 * f[k + 1].tet + f[k + 1].face] = f[k].tet;
        } else if (e - k > 2) {
            bad++;
        }
        k = e;
    }
    free(f);
    return bad;
}

/* Uniform grid over points for the empty-sphere test. */
int orc_validate(const double *xyz, int32_t n, const int32_t *tets, int64_t nt, int full_empty,
                 int64_t *counts)
{
    int flags = 0;
    int64_t c_euler = 0, c_orient = 0, c_insph = 0, c_empty = 0;
    /* Euler characteristic via unique edges and faces (GD/Common/Geometry.cu:211-317) */
    {
        uint64_t *e = xmalloc(sizeof(uint64_t) * 6 * nt + 1);
        int64_t ne = 0;
        for (int64_t t = 0; t < nt; t++)
            for (int i = 0; i < 4; i++)
                for (int j = i + 1; j < 4; j++) e[ne++] = pair_key(tets[4 * t + i], tets[4 * t + j]);
        radix_u64(e, ne, 64);
        ne = unique_u64(e, ne);
        free(e);
        int32_t *opp = xmalloc(sizeof(int32_t) * 4 * nt + 1);
        if (orc_adjacency(tets, nt, n, opp) != 0) flags |= 16;
        int64_t nf_shared = 0, nf_single = 0;
        for (int64_t k = 0; k < 4 * nt; k++) {
            if (opp[k] >= 0)
                nf_shared++;
            else
                nf_single++;
        }
        const int64_t nf = nf_single + nf_shared / 2;
        /* vertices used */
        unsigned char *used = calloc(n, 1);
        int64_t nv = 0;
        for (int64_t k = 0; k < 4 * nt; k++)
            if (!used[tets[k]]) {
                used[tets[k]] = 1;
                nv++;
            }
        free(used);
        c_euler = nv - ne + nf - nt;
        if (c_euler != 1) flags |= 1;
        /* orientation and adjacent-pair insphere (GD/Common/Geometry.cu:319-393) */
        for (int64_t t = 0; t < nt; t++) {
            const int32_t *q = tets + 4 * t;
            const double *p0 = xyz + 3 * (int64_t)q[0], *p1 = xyz + 3 * (int64_t)q[1],
                         *p2 = xyz + 3 * (int64_t)q[2], *p3 = xyz + 3 * (int64_t)q[3];
            const int o = orc_orient3d(p0, p1, p2, p3);
            if (o == 0) c_orient++;
            for (int i = 0; i < 4; i++) {
                const int64_t u = opp[4 * t + i];
                if (u < 0) continue;
                /* vertex of u not in t */
                int w = -1;
                for (int j = 0; j < 4; j++) {
                    const int cand = tets[4 * u + j];
                    if (cand != q[0] && cand != q[1] && cand != q[2] && cand != q[3]) w = cand;
                }
                const int s = orc_insphere(p0, p1, p2, p3, xyz + 3 * (int64_t)w);
                if (s * o > 0) c_insph++; /* strictly inside */
            }
        }
        free(opp);
        if (c_orient) flags |= 2;
        if (c_insph) flags |= 4;
    }
    if (full_empty && nt > 0) {
        /* bucket grid over the bounding box */
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int32_t i = 0; i < n; i++)
            for (int c = 0; c < 3; c++) {
                if (xyz[3 * i + c] < lo[c]) lo[c] = xyz[3 * i + c];
                if (xyz[3 * i + c] > hi[c]) hi[c] = xyz[3 * i + c];
            }
        int R = (int)cbrt((double)n / 2.0);
        if (R < 1) R = 1;
        if (R > 256) R = 256;
        double cell[3];
        for (int c = 0; c < 3; c++) cell[c] = (hi[c] - lo[c]) / R + 1e-300;
        int64_t ncell = (int64_t)R * R * R;
        int64_t *start = calloc(ncell + 1, sizeof(int64_t));
        int32_t *items = xmalloc(sizeof(int32_t) * n);
        int32_t *cid = xmalloc(sizeof(int32_t) * n);
        for (int32_t i = 0; i < n; i++) {
            int ci[3];
            for (int c = 0; c < 3; c++) {
                ci[c] = (int)((xyz[3 * i + c] - lo[c]) / cell[c]);
                if (ci[c] >= R) ci[c] = R - 1;
                if (ci[c] < 0) ci[c] = 0;
            }
            cid[i] = (int32_t)(((int64_t)ci[2] * R + ci[1]) * R + ci[0]);
            start[cid[i] + 1]++;
        }
        for (int64_t k = 0; k < ncell; k++) start[k + 1] += start[k];
        int64_t *fill = xmalloc(sizeof(int64_t) * ncell);
        memcpy(fill, start, sizeof(int64_t) * ncell);
        for (int32_t i = 0; i < n; i++) items[fill[cid[i]]++] = i;
        free(fill);
        free(cid);
        for (int64_t t = 0; t < nt; t++) {
            const int32_t *q = tets + 4 * t;
            const double *p[4];
            for (int i = 0; i < 4; i++) p[i] = xyz + 3 * (int64_t)q[i];
            /* circumcentre (approximate; the query box is padded) */
            double a[3][3], b[3];
            for (int i = 0; i < 3; i++) {
                double s = 0;
                for (int c = 0; c < 3; c++) {
                    a[i][c] = p[i + 1][c] - p[0][c];
                    s += a[i][c] * a[i][c];
                }
                b[i] = 0.5 * s;
            }
            const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                               a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                               a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
            double cc[3], r2 = 0;
            int whole = (fabs(det) < 1e-300);
            if (!whole) {
                double inv[3][3];
                inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
                inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
                inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
                inv[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
                inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
                inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
                inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
                inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
                inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
                for (int c = 0; c < 3; c++) {
                    cc[c] = p[0][c] + inv[c][0] * b[0] + inv[c][1] * b[1] + inv[c][2] * b[2];
                    r2 += (cc[c] - p[0][c]) * (cc[c] - p[0][c]);
                }
                if (!isfinite(r2)) whole = 1;
            }
            int lo_c[3], hi_c[3];
            for (int c = 0; c < 3; c++) {
                if (whole) {
                    lo_c[c] = 0;
                    hi_c[c] = R - 1;
                    continue;
                }
                const double r = sqrt(r2) * (1 + 1e-6) + 1e-12;
                double l = (cc[c] - r - lo[c]) / cell[c], h = (cc[c] + r - lo[c]) / cell[c];
                lo_c[c] = l < 0 ? 0 : (l >= R ? R : (int)l);
                hi_c[c] = h < 0 ? -1 : (h >= R ? R - 1 : (int)h);
            }
            const int o = orc_orient3d(p[0], p[1], p[2], p[3]);
            for (int z = lo_c[2]; z <= hi_c[2]; z++)
                for (int y = lo_c[1]; y <= hi_c[1]; y++)
                    for (int x = lo_c[0]; x <= hi_c[0]; x++) {
                        const int64_t k = ((int64_t)z * R + y) * R + x;
                        for (int64_t it = start[k]; it < start[k + 1]; it++) {
                            const int32_t w = items[it];
                            if (w == q[0] || w == q[1] || w == q[2] || w == q[3]) continue;
                            const double *pw = xyz + 3 * (int64_t)w;
                            if (!whole) {
                                double d2 = 0;
                                for (int c = 0; c < 3; c++) d2 += (pw[c] - cc[c]) * (pw[c] - cc[c]);
                                if (d2 > r2 * (1 + 1e-6) + 1e-24) continue;
                            }
                            const int s = orc_insphere(p[0], p[1], p[2], p[3], pw);
                            if (s * o > 0) c_empty++;
                        }
                    }
        }
        free(start);
        free(items);
        if (c_empty) flags |= 8;
    }
    if (counts) {
        counts[0] = c_euler;
        counts[1] = c_orient;
        counts[2] = c_insph;
        counts[3] = c_empty;
    }
    return flags;
}
