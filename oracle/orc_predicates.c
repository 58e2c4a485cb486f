/*
 * gStar4D CPU oracle: exact geometric predicates — TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 * Exactness follows Shewchuk's published method (J. R. Shewchuk, "Adaptive Precision
 * Floating-Point Arithmetic and Fast Robust Geometric Predicates", DCG 18, 1997):
 * nonoverlapping floating-point expansions built with Two-Sum / Two-Product and grown one
 * component at a time (Grow-Expansion, Theorem 10).  Two-Product uses fma() for the error
 * term.  The code is this project's own; it mirrors the semantics of
 *   - orient3dfast/orient3dexact   GD/GDelaunay/GDelShewchukDevice.h:664-794
 *   - insphere/insphereexact       GD/GDelaunay/GDelShewchukDevice.h:823-1144
 *   - orient{1,2,3}dexact_lifted   GD/GDelaunay/GDelShewchukDevice.h:1146-1419
 *   - _orientation4SoS             GD/GDelaunay/GDelPredDevice.h:142-669
 *   - orientation4SoS              GD/GDelaunay/GDelPredDevice.h:671-698
 *
 * Sign conventions (all verified by tests against the reference's predicates.c and against
 * rational arithmetic):
 *   orient3d(a,b,c,d)    = det[[a 1][b 1][c 1][d 1]]        (x,y,z,1 columns)
 *   insphere(a,b,c,d,e)  = det[[p |p|^2 1]] over rows a..e  (x,y,z,lift,1 columns)
 *   orientation4(p0..p4) = sign insphere(p0..p4), SoS when zero.
 *
 * Compile with -ffp-contract=off: the filters' error bounds assume separately rounded
 * products, and Two-Sum needs unfused additions.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <omp.h>

#include "orc.h"

#define ORC_EPS 1.1102230246251565e-16 /* 2^-53, unit roundoff of FP64 */

/* Call counters, one padded row per OpenMP thread (summed when read), so parallel star
   loops count without contention. */
#define ORC_MAX_THREADS 256
static struct {
    int64_t orient4, exact, sos, pad[5];
} g_cnt[ORC_MAX_THREADS];
#define CNT (g_cnt[omp_get_thread_num() & (ORC_MAX_THREADS - 1)])

int64_t orc_pred_counter(int which)
{
    int64_t s = 0;
    for (int t = 0; t < ORC_MAX_THREADS; t++)
        s += which == 0 ? g_cnt[t].orient4 : which == 1 ? g_cnt[t].exact : g_cnt[t].sos;
    return s;
}
void orc_pred_counter_reset(void) { memset(g_cnt, 0, sizeof(g_cnt)); }

/* ------------------------------------------------------------------------------------ */
/* Expansion arithmetic                                                                  */
/* ------------------------------------------------------------------------------------ */

#define EXP_CAP 1024

typedef struct {
    int n;
    double v[EXP_CAP]; /* nonoverlapping, increasing magnitude, no zeros (except n==1) */
} expn;

static void orc_die(const char *msg)
{
    fprintf(stderr, "gstar4d oracle: fatal: %s\n", msg);
    abort();
}

/* e <- e + b exactly (Grow-Expansion with zero elimination; in place is safe because the
   output index never passes the input index). */
static void ex_grow(expn *e, double b)
{
    double q = b;
    int hn = 0;
    for (int i = 0; i < e->n; i++) {
        const double ei = e->v[i];
        const double s = q + ei;
        const double bv = s - q;
        const double av = s - bv;
        const double t = (q - av) + (ei - bv);
        if (t != 0.0) e->v[hn++] = t;
        q = s;
    }
    if (q != 0.0 || hn == 0) {
        if (hn >= EXP_CAP) orc_die("expansion capacity exceeded");
        e->v[hn++] = q;
    }
    e->n = hn;
}

static void ex_set(expn *e, double b)
{
    e->n = 1;
    e->v[0] = b;
}

/* acc += sgn * (a * b), exactly. */
static void ex_addprod(expn *acc, const expn *a, const expn *b, int sgn)
{
    for (int i = 0; i < a->n; i++) {
        for (int j = 0; j < b->n; j++) {
            const double p = a->v[i] * b->v[j];
            const double err = fma(a->v[i], b->v[j], -p);
            if (sgn > 0) {
                if (err != 0.0) ex_grow(acc, err);
                ex_grow(acc, p);
            } else {
                if (err != 0.0) ex_grow(acc, -err);
                ex_grow(acc, -p);
            }
        }
    }
}

static int ex_sign(const expn *e)
{
    const double top = e->v[e->n - 1];
    return top > 0.0 ? 1 : (top < 0.0 ? -1 : 0);
}

/* |p|^2 as an exact expansion. */
static void ex_lift(expn *out, const double *p)
{
    ex_set(out, 0.0);
    for (int c = 0; c < 3; c++) {
        const double sq = p[c] * p[c];
        const double err = fma(p[c], p[c], -sq);
        if (err != 0.0) ex_grow(out, err);
        ex_grow(out, sq);
    }
}

/* ------------------------------------------------------------------------------------ */
/* Exact sign of det[[c_0 ... c_{k-2} 1]] for k <= 5 rows (memoised Laplace expansion    */
/* along the columns in order; the implicit last column is all ones).                    */
/* ------------------------------------------------------------------------------------ */

static __thread expn t_memo[32];
static __thread expn t_entry[5][4];

int orc_sign_minor(int k, const double *const *rows, const int *cols)
{
    if (k < 1 || k > 5) orc_die("orc_sign_minor: bad size");
    /* entries */
    for (int r = 0; r < k; r++) {
        for (int j = 0; j < k - 1; j++) {
            const int c = cols[j];
            if (c == 3)
                ex_lift(&t_entry[r][j], rows[r]);
            else
                ex_set(&t_entry[r][j], rows[r][c]);
        }
    }
    /* memo over row subsets ordered by popcount: subset of size s uses column k - s */
    const int full = (1 << k) - 1;
    for (int size = 1; size <= k; size++) {
        for (int mask = 1; mask <= full; mask++) {
            if (__builtin_popcount((unsigned)mask) != size) continue;
            expn *d = &t_memo[mask];
            if (size == 1) {
                ex_set(d, 1.0);
                continue;
            }
            const int col = k - size;
            ex_set(d, 0.0);
            int pos = 0;
            for (int r = 0; r < k; r++) {
                if (!(mask & (1 << r))) continue;
                const expn *sub = &t_memo[mask & ~(1 << r)];
                ex_addprod(d, &t_entry[r][col], sub, (pos & 1) ? -1 : 1);
                pos++;
            }
        }
    }
    return ex_sign(&t_memo[full]);
}

int orc_orient3d_exact(const double *pa, const double *pb, const double *pc, const double *pd)
{
    const double *rows[4] = {pa, pb, pc, pd};
    const int cols[3] = {0, 1, 2};
    return orc_sign_minor(4, rows, cols);
}

int orc_insphere_exact(const double *pa, const double *pb, const double *pc, const double *pd,
                       const double *pe)
{
    const double *rows[5] = {pa, pb, pc, pd, pe};
    const int cols[4] = {0, 1, 2, 3};
    return orc_sign_minor(5, rows, cols);
}

/* ------------------------------------------------------------------------------------ */
/* Filters.  Forward error analysis of the evaluation order below (each product and each */
/* addition rounded once, inputs rounded once by the translation) bounds the error of    */
/* every monomial by gamma_7 (orient3d) and gamma_17 (insphere) times the permanent; the */
/* constants used (12 eps, 32 eps) are deliberately looser.  Tiny permanents (possible   */
/* underflow) always go exact.                                                          */
/* ------------------------------------------------------------------------------------ */

int orc_orient3d(const double *pa, const double *pb, const double *pc, const double *pd)
{
    const double adx = pa[0] - pd[0], ady = pa[1] - pd[1], adz = pa[2] - pd[2];
    const double bdx = pb[0] - pd[0], bdy = pb[1] - pd[1], bdz = pb[2] - pd[2];
    const double cdx = pc[0] - pd[0], cdy = pc[1] - pd[1], cdz = pc[2] - pd[2];
    const double m_bc = bdx * cdy - cdx * bdy, p_bc = fabs(bdx * cdy) + fabs(cdx * bdy);
    const double m_ca = cdx * ady - adx * cdy, p_ca = fabs(cdx * ady) + fabs(adx * cdy);
    const double m_ab = adx * bdy - bdx * ady, p_ab = fabs(adx * bdy) + fabs(bdx * ady);
    const double det = adz * m_bc + bdz * m_ca + cdz * m_ab;
    const double perm = fabs(adz) * p_bc + fabs(bdz) * p_ca + fabs(cdz) * p_ab;
    const double bound = 12.0 * ORC_EPS * perm;
    if (perm > 1e-250) {
        if (det > bound) return 1;
        if (-det > bound) return -1;
    }
    CNT.exact++;
    return orc_orient3d_exact(pa, pb, pc, pd);
}

int orc_insphere(const double *pa, const double *pb, const double *pc, const double *pd,
                 const double *pe)
{
    /* rows a..d translated by e; det4 of [x y z lift] equals the 5x5 lifted determinant */
    const double *p[4] = {pa, pb, pc, pd};
    double x[4], y[4], z[4], l[4];
    for (int i = 0; i < 4; i++) {
        x[i] = p[i][0] - pe[0];
        y[i] = p[i][1] - pe[1];
        z[i] = p[i][2] - pe[2];
        l[i] = x[i] * x[i] + y[i] * y[i] + z[i] * z[i];
    }
    /* 2x2 xy minors m[i][j] = x_i y_j - x_j y_i */
    double m2[4][4], p2[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = i + 1; j < 4; j++) {
            const double u = x[i] * y[j], v = x[j] * y[i];
            m2[i][j] = u - v;
            p2[i][j] = fabs(u) + fabs(v);
        }
    /* 3x3 xyz minors over row triples (i<j<k), expanded along z */
    double m3[4], p3[4]; /* index = row left out */
    for (int out = 0; out < 4; out++) {
        int r[3], q = 0;
        for (int i = 0; i < 4; i++)
            if (i != out) r[q++] = i;
        const int i = r[0], j = r[1], k = r[2];
        m3[out] = z[i] * m2[j][k] - z[j] * m2[i][k] + z[k] * m2[i][j];
        p3[out] = fabs(z[i]) * p2[j][k] + fabs(z[j]) * p2[i][k] + fabs(z[k]) * p2[i][j];
    }
    /* det4 along the lift column (column index 3): sum_i (-1)^(i+3) l_i M_i3 */
    const double det = -l[0] * m3[0] + l[1] * m3[1] - l[2] * m3[2] + l[3] * m3[3];
    const double perm = l[0] * p3[0] + l[1] * p3[1] + l[2] * p3[2] + l[3] * p3[3];
    const double bound = 32.0 * ORC_EPS * perm;
    if (perm > 1e-250) {
        if (det > bound) return 1;
        if (-det > bound) return -1;
    }
    CNT.exact++;
    return orc_insphere_exact(pa, pb, pc, pd, pe);
}

/* ------------------------------------------------------------------------------------ */
/* Simulation of Simplicity for the lifted orientation (GD/GDelaunay/GDelPredDevice.h).   */
/* Each depth is the minor det[[cols... 1]] over a subset of the index-sorted rows.       */
/* Columns: 0=x 1=y 2=z 3=lift.  The table restates the reference's switch statements    */
/* (:194-499 rows/columns, :503-612 lifted flags and minor sizes, :618-660 sign flips).  */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    const char *rows; /* digits of sorted-row positions */
    const char *cols; /* 'x','y','z','l' */
    int flip;         /* result negated at this depth */
} sos_depth;

static const sos_depth SOS_TABLE[49] = {
    {"", "", 0},          /* 0 unused */
    {"1234", "xyz", 1},   /* 1 */
    {"1234", "xyl", 0},   /* 2 */
    {"1234", "xzl", 1},   /* 3 */
    {"1234", "yzl", 0},   /* 4 */
    {"0234", "xyz", 0},   /* 5 */
    {"234", "xy", 0},     /* 6 */
    {"234", "xz", 1},     /* 7 */
    {"234", "yz", 0},     /* 8 */
    {"0234", "xyl", 1},   /* 9 */
    {"234", "xl", 0},     /* 10 */
    {"234", "yl", 1},     /* 11 */
    {"0234", "xzl", 0},   /* 12 */
    {"234", "zl", 0},     /* 13 */
    {"0234", "yzl", 1},   /* 14 */
    {"0134", "xyz", 1},   /* 15 */
    {"134", "xy", 1},     /* 16 */
    {"134", "xz", 0},     /* 17 */
    {"134", "yz", 1},     /* 18 */
    {"034", "xy", 0},     /* 19 */
    {"34", "x", 1},       /* 20 */
    {"34", "y", 0},       /* 21 */
    {"034", "xz", 1},     /* 22 */
    {"34", "z", 1},       /* 23 */
    {"034", "yz", 0},     /* 24 */
    {"0134", "xyl", 0},   /* 25 */
    {"134", "xl", 1},     /* 26 */
    {"134", "yl", 0},     /* 27 */
    {"034", "xl", 0},     /* 28 */
    {"34", "l", 0},       /* 29 */
    {"034", "yl", 1},     /* 30 */
    {"0134", "xzl", 1},   /* 31 */
    {"134", "zl", 1},     /* 32 */
    {"034", "zl", 0},     /* 33 */
    {"0134", "yzl", 0},   /* 34 */
    {"0124", "xyz", 0},   /* 35 */
    {"124", "xy", 0},     /* 36 */
    {"124", "xz", 1},     /* 37 */
    {"124", "yz", 0},     /* 38 */
    {"024", "xy", 1},     /* 39 */
    {"24", "x", 0},       /* 40 */
    {"24", "y", 1},       /* 41 */
    {"024", "xz", 0},     /* 42 */
    {"24", "z", 0},       /* 43 */
    {"024", "yz", 1},     /* 44 */
    {"014", "xy", 0},     /* 45 */
    {"14", "x", 1},       /* 46 */
    {"14", "y", 0},       /* 47 */
    {"04", "x", 0},       /* 48 */
};

static int col_code(char c) { return c == 'x' ? 0 : c == 'y' ? 1 : c == 'z' ? 2 : 3; }

int orc_orient4_sos_xyz(const double *p0, const double *p1, const double *p2, const double *p3,
                        const double *p4, int i0, int i1, int i2, int i3, int i4, int *depth_out)
{
    CNT.orient4++;
    const int s = orc_insphere(p0, p1, p2, p3, p4);
    if (s != 0) {
        if (depth_out) *depth_out = 0;
        return s;
    }
    CNT.sos++;
    /* sort the indices ascending, tracking the permutation parity */
    int idx[5] = {i0, i1, i2, i3, i4};
    const double *pt[5] = {p0, p1, p2, p3, p4};
    int parity = 0;
    for (int i = 1; i < 5; i++) {
        for (int j = i; j > 0 && idx[j] < idx[j - 1]; j--) {
            int ti = idx[j]; idx[j] = idx[j - 1]; idx[j - 1] = ti;
            const double *tp = pt[j]; pt[j] = pt[j - 1]; pt[j - 1] = tp;
            parity ^= 1;
        }
    }
    int depth = 0, sign = 0;
    for (int d = 1; d <= 48 && sign == 0; d++) {
        const sos_depth *e = &SOS_TABLE[d];
        const int k = (int)strlen(e->rows);
        const double *rows[5];
        int cols[4];
        for (int r = 0; r < k; r++) rows[r] = pt[e->rows[r] - '0'];
        for (int c = 0; c < k - 1; c++) cols[c] = col_code(e->cols[c]);
        sign = orc_sign_minor(k, rows, cols);
        depth = d;
        if (sign != 0 && e->flip) sign = -sign;
    }
    if (sign == 0) {
        depth = 49;
        sign = 1;
    }
    if (parity) sign = -sign;
    if (depth_out) *depth_out = depth;
    return sign;
}

int orc_orient4_sos(const double *pts, int i0, int i1, int i2, int i3, int i4, int *depth_out)
{
    return orc_orient4_sos_xyz(pts + 3 * (int64_t)i0, pts + 3 * (int64_t)i1, pts + 3 * (int64_t)i2,
                               pts + 3 * (int64_t)i3, pts + 3 * (int64_t)i4, i0, i1, i2, i3, i4,
                               depth_out);
}
