// gStar4D for MI355X — exact geometric predicates on the GPU.
//
// Semantics (must equal the reference's, GD/GDelaunay/GDelPredDevice.h and
// GD/GDelaunay/GDelShewchukDevice.h, with REAL = double):
//   orient3d(a,b,c,d)    sign det[[a 1][b 1][c 1][d 1]]                  (:664-794)
//   insphere(a,b,c,d,e)  sign det[[p |p|^2 1]] over rows a..e             (:823-1144)
//   orient4(p0..p4)      insphere sign, and when it is zero Simulation of Simplicity over
//                        the index-sorted rows with the reference's 48-depth minor table
//                        and sign flips (GD/GDelaunay/GDelPredDevice.h:142-698).
//
// Method (this project's own): a semi-static FP64 filter with a forward-error bound, and on
// an uncertain filter an exact evaluation in fixed-point integer arithmetic.  All FP64
// inputs of one predicate are rescaled to integers sharing the lowest set bit of any of
// them, so every determinant is an exact signed big integer (32-bit limbs, two's
// complement, v_mad_u64_u32 products).  Coordinates whose integer form needs more than
// 125 bits raise G4D_DEV_ERR_PREDICATE_RANGE instead of returning a wrong sign.
#pragma once

#include "g4d_common.h"

namespace g4d {

// ------------------------------------------------------------------------------------
// Fixed-width signed big integers
// ------------------------------------------------------------------------------------

template <int N>
struct BI {
    uint32_t w[N];
};

template <int N>
__device__ __forceinline__ BI<N> bi_from_i64(int64_t v)
{
    BI<N> r;
    r.w[0] = (uint32_t)(uint64_t)v;
    if (N > 1) r.w[1] = (uint32_t)((uint64_t)v >> 32);
    const uint32_t ext = v < 0 ? 0xffffffffu : 0u;
#pragma unroll
    for (int i = 2; i < N; i++) r.w[i] = ext;
    return r;
}

template <int NR, int N>
__device__ __forceinline__ BI<NR> bi_ext(const BI<N> &a)
{
    BI<NR> r;
    const uint32_t ext = (a.w[N - 1] >> 31) ? 0xffffffffu : 0u;
#pragma unroll
    for (int i = 0; i < NR; i++) r.w[i] = i < N ? a.w[i] : ext;
    return r;
}

template <int N>
__device__ __forceinline__ BI<N> bi_add(const BI<N> &a, const BI<N> &b)
{
    BI<N> r;
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < N; i++) {
        const uint64_t t = (uint64_t)a.w[i] + b.w[i] + c;
        r.w[i] = (uint32_t)t;
        c = t >> 32;
    }
    return r;
}

template <int N>
__device__ __forceinline__ BI<N> bi_sub(const BI<N> &a, const BI<N> &b)
{
    BI<N> r;
    uint64_t c = 1;
#pragma unroll
    for (int i = 0; i < N; i++) {
        const uint64_t t = (uint64_t)a.w[i] + (uint32_t)~b.w[i] + c;
        r.w[i] = (uint32_t)t;
        c = t >> 32;
    }
    return r;
}

template <int N>
__device__ __forceinline__ BI<N> bi_neg(const BI<N> &a)
{
    BI<N> r;
    uint64_t c = 1;
#pragma unroll
    for (int i = 0; i < N; i++) {
        const uint64_t t = (uint64_t)(uint32_t)~a.w[i] + c;
        r.w[i] = (uint32_t)t;
        c = t >> 32;
    }
    return r;
}

template <int N>
__device__ __forceinline__ bool bi_isneg(const BI<N> &a)
{
    return (a.w[N - 1] >> 31) != 0;
}

template <int N>
__device__ __forceinline__ int bi_sign(const BI<N> &a)
{
    if (bi_isneg(a)) return -1;
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < N; i++) o |= a.w[i];
    return o ? 1 : 0;
}

// Exact signed product; |a| < 2^(32NA-1), |b| < 2^(32NB-1) so it fits in NA+NB limbs.
template <int NA, int NB>
__device__ __forceinline__ BI<NA + NB> bi_mul(const BI<NA> &a_in, const BI<NB> &b_in)
{
    const bool na = bi_isneg(a_in), nb = bi_isneg(b_in);
    const BI<NA> a = na ? bi_neg(a_in) : a_in;
    const BI<NB> b = nb ? bi_neg(b_in) : b_in;
    BI<NA + NB> r;
#pragma unroll
    for (int i = 0; i < NA + NB; i++) r.w[i] = 0;
#pragma unroll
    for (int i = 0; i < NA; i++) {
        uint64_t carry = 0;
#pragma unroll
        for (int j = 0; j < NB; j++) {
            const uint64_t t = (uint64_t)a.w[i] * b.w[j] + r.w[i + j] + carry;
            r.w[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        r.w[i + NB] = (uint32_t)carry;
    }
    return (na != nb) ? bi_neg(r) : r;
}

// ------------------------------------------------------------------------------------
// Fixed-point conversion of the FP64 inputs of one predicate
// ------------------------------------------------------------------------------------

// Decompose finite nonzero x = m * 2^e with m an odd-free-trailing 53-bit integer (trailing
// zero bits stripped into e).  Returns false for zero.
__device__ __forceinline__ bool fp_decompose(double x, int64_t *m, int *e)
{
    const uint64_t bits = (uint64_t)__double_as_longlong(x);
    const int bexp = (int)((bits >> 52) & 0x7ff);
    uint64_t mant = bits & ((1ull << 52) - 1);
    int ex;
    if (bexp == 0) {
        if (mant == 0) return false;
        ex = -1074;
    } else {
        mant |= (1ull << 52);
        ex = bexp - 1075;
    }
    const int tz = __builtin_ctzll(mant);
    mant >>= tz;
    ex += tz;
    *m = (bits >> 63) ? -(int64_t)mant : (int64_t)mant;
    *e = ex;
    return true;
}

// Integer coordinates of NP points (3 each) on a common scale.  K = limbs per coordinate.
template <int NP>
struct FixedPts {
    int64_t m[NP][3];
    int e[NP][3];
    bool nz[NP][3];
    int emin;
    int bits; // bits of the largest magnitude after scaling
};

template <int NP>
__device__ __forceinline__ void fixed_analyse(const double *const *p, FixedPts<NP> &f)
{
    f.emin = 1 << 20;
    int top = -(1 << 20);
#pragma unroll
    for (int i = 0; i < NP; i++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            f.nz[i][c] = fp_decompose(p[i][c], &f.m[i][c], &f.e[i][c]);
            if (f.nz[i][c]) {
                f.emin = min(f.emin, f.e[i][c]);
                const uint64_t am = (uint64_t)(f.m[i][c] < 0 ? -f.m[i][c] : f.m[i][c]);
                top = max(top, f.e[i][c] + 64 - __builtin_clzll(am));
            }
        }
    f.bits = (top == -(1 << 20)) ? 0 : top - f.emin;
}

// Coordinate as a K-limb signed integer scaled by 2^-emin.
template <int K, int NP>
__device__ __forceinline__ BI<K> fixed_coord(const FixedPts<NP> &f, int i, int c)
{
    BI<K> r;
#pragma unroll
    for (int q = 0; q < K; q++) r.w[q] = 0;
    if (!f.nz[i][c]) return r;
    const int64_t mm = f.m[i][c];
    const uint64_t am = (uint64_t)(mm < 0 ? -mm : mm);
    const int sh = f.e[i][c] - f.emin; // >= 0, and am << sh < 2^(32K-2) by the caller's check
    const int wq = sh >> 5, bq = sh & 31;
#pragma unroll
    for (int q = 0; q < K; q++) {
        // limb q receives bits [32q, 32q+32) of am << sh
        uint32_t v = 0;
        const int lo = q - wq; // which 32-bit chunk of (am << bq)
        const uint64_t lo64 = am << bq;
        const uint64_t hi64 = bq ? (am >> (64 - bq)) : 0; // bits 64.. of am << bq
        if (lo == 0) v = (uint32_t)lo64;
        else if (lo == 1) v = (uint32_t)(lo64 >> 32);
        else if (lo == 2) v = (uint32_t)hi64;
        r.w[q] = v;
    }
    return mm < 0 ? bi_neg(r) : r;
}

// ------------------------------------------------------------------------------------
// Exact determinants over K-limb coordinates
// ------------------------------------------------------------------------------------

// Lifted determinant det[[p |p|^2 1]] of 5 points (rows), via translation by the last row:
// det4[[d |d|^2]] with d = p_i - p_4.  Sign equals Shewchuk's insphere(p0,p1,p2,p3,p4).
template <int K>
__device__ __noinline__ int exact_insphere_k(const FixedPts<5> &f)
{
    // |coordinates| < 2^(32K-3) so differences fit K limbs; tight limb counts follow from
    // |d| < 2^(32K-2): 2x2 minors < 2^(64K-3), 3x3 minors < 2^(96K-3), lifts < 2^(64K-2),
    // determinant < 2^(160K-1).
    constexpr int ND = K;
    BI<ND> d[4][3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const BI<ND> e = bi_ext<ND>(fixed_coord<K>(f, 4, c));
#pragma unroll
        for (int i = 0; i < 4; i++) d[i][c] = bi_sub(bi_ext<ND>(fixed_coord<K>(f, i, c)), e);
    }
    // 2x2 xy minors
    constexpr int N2 = 2 * ND;
    BI<N2> m2[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = i + 1; j < 4; j++)
            m2[i][j] = bi_sub(bi_ext<N2>(bi_mul(d[i][0], d[j][1])), bi_ext<N2>(bi_mul(d[j][0], d[i][1])));
    // 3x3 xyz minors, index = row left out; expanded along z
    constexpr int N3 = ND + N2;
    BI<N3> m3[4];
#pragma unroll
    for (int out = 0; out < 4; out++) {
        int r[3], q = 0;
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (i != out) r[q++] = i;
        BI<N3> t0 = bi_ext<N3>(bi_mul(d[r[0]][2], m2[r[1]][r[2]]));
        BI<N3> t1 = bi_ext<N3>(bi_mul(d[r[1]][2], m2[r[0]][r[2]]));
        BI<N3> t2 = bi_ext<N3>(bi_mul(d[r[2]][2], m2[r[0]][r[1]]));
        m3[out] = bi_add(bi_sub(t0, t1), t2);
    }
    // lifts
    constexpr int NL = 2 * ND;
    BI<NL> l[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        l[i] = bi_add(bi_add(bi_ext<NL>(bi_mul(d[i][0], d[i][0])), bi_ext<NL>(bi_mul(d[i][1], d[i][1]))),
                      bi_ext<NL>(bi_mul(d[i][2], d[i][2])));
    }
    constexpr int NF = NL + N3;
    BI<NF> acc = bi_ext<NF>(bi_mul(l[1], m3[1]));
    acc = bi_sub(acc, bi_ext<NF>(bi_mul(l[0], m3[0])));
    acc = bi_sub(acc, bi_ext<NF>(bi_mul(l[2], m3[2])));
    acc = bi_add(acc, bi_ext<NF>(bi_mul(l[3], m3[3])));
    return bi_sign(acc);
}

// orient3d exact: det[[p 1]] over 4 points, translated by the last.
template <int K>
__device__ __noinline__ int exact_orient3d_k(const FixedPts<4> &f)
{
    constexpr int ND = K;
    BI<ND> d[3][3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const BI<ND> e = bi_ext<ND>(fixed_coord<K>(f, 3, c));
#pragma unroll
        for (int i = 0; i < 3; i++) d[i][c] = bi_sub(bi_ext<ND>(fixed_coord<K>(f, i, c)), e);
    }
    constexpr int N2 = 2 * ND;
    const BI<N2> m12 = bi_sub(bi_mul(d[1][0], d[2][1]), bi_mul(d[2][0], d[1][1]));
    const BI<N2> m02 = bi_sub(bi_mul(d[0][0], d[2][1]), bi_mul(d[2][0], d[0][1]));
    const BI<N2> m01 = bi_sub(bi_mul(d[0][0], d[1][1]), bi_mul(d[1][0], d[0][1]));
    constexpr int N3 = ND + N2;
    BI<N3> acc = bi_mul(d[0][2], m12);
    acc = bi_sub(acc, bi_mul(d[1][2], m02));
    acc = bi_add(acc, bi_mul(d[2][2], m01));
    return bi_sign(acc);
}

// SoS minor: sign det[[cols... 1]] over k rows (k = 2..4) of the 5 fixed points;
// rows/cols are packed nibbles (row i = (rows >> 4i) & 15, col c = (cols >> 4c) & 15,
// col 3 = lift |p|^2 of the absolute coordinates).
template <int K>
__device__ __noinline__ int exact_minor_k(const FixedPts<5> &f, int k, uint32_t rows, uint32_t cols)
{
    // entries as absolute integers: |coord| < 2^(32K-3), |lift| < 2^(64K-4); all
    // differences then fit 2K limbs, 2x2 minors 4K limbs and 3x3 minors 6K limbs.
    constexpr int NE = 2 * K;
    BI<NE> a[4][3];
    for (int i = 0; i < k; i++) {
        const int r = (rows >> (4 * i)) & 15;
        for (int j = 0; j < k - 1; j++) {
            const int c = (cols >> (4 * j)) & 15;
            if (c < 3) {
                a[i][j] = bi_ext<NE>(fixed_coord<K>(f, r, c));
            } else {
                const BI<K> x = fixed_coord<K>(f, r, 0), y = fixed_coord<K>(f, r, 1), z = fixed_coord<K>(f, r, 2);
                a[i][j] = bi_add(bi_add(bi_ext<NE>(bi_mul(x, x)), bi_ext<NE>(bi_mul(y, y))), bi_ext<NE>(bi_mul(z, z)));
            }
        }
    }
    // translate by the last row (the ones column drops out)
    BI<NE> d[3][3];
    for (int i = 0; i < k - 1; i++)
        for (int j = 0; j < k - 1; j++) d[i][j] = bi_sub(a[i][j], a[k - 1][j]);
    if (k == 2) return bi_sign(d[0][0]);
    if (k == 3) return bi_sign(bi_sub(bi_mul(d[0][0], d[1][1]), bi_mul(d[1][0], d[0][1])));
    constexpr int N2 = 2 * NE;
    const BI<N2> m12 = bi_sub(bi_mul(d[1][0], d[2][1]), bi_mul(d[2][0], d[1][1]));
    const BI<N2> m02 = bi_sub(bi_mul(d[0][0], d[2][1]), bi_mul(d[2][0], d[0][1]));
    const BI<N2> m01 = bi_sub(bi_mul(d[0][0], d[1][1]), bi_mul(d[1][0], d[0][1]));
    constexpr int N3 = NE + N2;
    BI<N3> acc = bi_mul(d[0][2], m12);
    acc = bi_sub(acc, bi_mul(d[1][2], m02));
    acc = bi_add(acc, bi_mul(d[2][2], m01));
    return bi_sign(acc);
}

// ------------------------------------------------------------------------------------
// SoS depth table (GD/GDelaunay/GDelPredDevice.h:194-660): for depth 1..48, the rows of the
// index-sorted points, the columns (0=x 1=y 2=z 3=lift) and whether the sign flips.
// Packed: bits 0..15 rows (nibbles), 16..27 cols (nibbles), 28..30 k (rows), 31 flip.
// ------------------------------------------------------------------------------------

#define G4D_SOS(r, c, k, fl) ((uint32_t)(r) | ((uint32_t)(c) << 16) | ((uint32_t)(k) << 28) | ((uint32_t)(fl) << 31))
__device__ __constant__ static const uint32_t kSosTable[49] = {
    0,
    G4D_SOS(0x4321, 0x210, 4, 1), G4D_SOS(0x4321, 0x310, 4, 0), G4D_SOS(0x4321, 0x320, 4, 1),
    G4D_SOS(0x4321, 0x321, 4, 0), G4D_SOS(0x4320, 0x210, 4, 0), G4D_SOS(0x432, 0x10, 3, 0),
    G4D_SOS(0x432, 0x20, 3, 1),   G4D_SOS(0x432, 0x21, 3, 0),   G4D_SOS(0x4320, 0x310, 4, 1),
    G4D_SOS(0x432, 0x30, 3, 0),   G4D_SOS(0x432, 0x31, 3, 1),   G4D_SOS(0x4320, 0x320, 4, 0),
    G4D_SOS(0x432, 0x32, 3, 0),   G4D_SOS(0x4320, 0x321, 4, 1), G4D_SOS(0x4310, 0x210, 4, 1),
    G4D_SOS(0x431, 0x10, 3, 1),   G4D_SOS(0x431, 0x20, 3, 0),   G4D_SOS(0x431, 0x21, 3, 1),
    G4D_SOS(0x430, 0x10, 3, 0),   G4D_SOS(0x43, 0x0, 2, 1),     G4D_SOS(0x43, 0x1, 2, 0),
    G4D_SOS(0x430, 0x20, 3, 1),   G4D_SOS(0x43, 0x2, 2, 1),     G4D_SOS(0x430, 0x21, 3, 0),
    G4D_SOS(0x4310, 0x310, 4, 0), G4D_SOS(0x431, 0x30, 3, 1),   G4D_SOS(0x431, 0x31, 3, 0),
    G4D_SOS(0x430, 0x30, 3, 0),   G4D_SOS(0x43, 0x3, 2, 0),     G4D_SOS(0x430, 0x31, 3, 1),
    G4D_SOS(0x4310, 0x320, 4, 1), G4D_SOS(0x431, 0x32, 3, 1),   G4D_SOS(0x430, 0x32, 3, 0),
    G4D_SOS(0x4310, 0x321, 4, 0), G4D_SOS(0x4210, 0x210, 4, 0), G4D_SOS(0x421, 0x10, 3, 0),
    G4D_SOS(0x421, 0x20, 3, 1),   G4D_SOS(0x421, 0x21, 3, 0),   G4D_SOS(0x420, 0x10, 3, 1),
    G4D_SOS(0x42, 0x0, 2, 0),     G4D_SOS(0x42, 0x1, 2, 1),     G4D_SOS(0x420, 0x20, 3, 0),
    G4D_SOS(0x42, 0x2, 2, 0),     G4D_SOS(0x420, 0x21, 3, 1),   G4D_SOS(0x410, 0x10, 3, 0),
    G4D_SOS(0x41, 0x0, 2, 1),     G4D_SOS(0x41, 0x1, 2, 0),     G4D_SOS(0x40, 0x0, 2, 0),
};
#undef G4D_SOS

// ------------------------------------------------------------------------------------
// Predicate entry points
// ------------------------------------------------------------------------------------

struct PredCounters {
    unsigned long long *cnt; // counter block or nullptr
};

__device__ __forceinline__ void pred_flag_error(unsigned long long *cnt, uint32_t flag)
{
    if (cnt) atomicOr(&cnt[G4D_CNT_ERROR], (unsigned long long)flag);
}

// Exact insphere sign of 5 points; flags range errors.
__device__ __noinline__ int exact_insphere(const double *const *p, unsigned long long *cnt)
{
    FixedPts<5> f;
    fixed_analyse<5>(p, f);
    if (f.bits <= 61) return exact_insphere_k<2>(f);
    if (f.bits <= 125) return exact_insphere_k<4>(f);
    pred_flag_error(cnt, G4D_DEV_ERR_PREDICATE_RANGE);
    return 0;
}

__device__ __noinline__ int exact_orient3d(const double *const *p, unsigned long long *cnt)
{
    FixedPts<4> f;
    fixed_analyse<4>(p, f);
    if (f.bits <= 61) return exact_orient3d_k<2>(f);
    if (f.bits <= 125) return exact_orient3d_k<4>(f);
    pred_flag_error(cnt, G4D_DEV_ERR_PREDICATE_RANGE);
    return 0;
}

// SoS resolution of a zero lifted determinant (GD/GDelaunay/GDelPredDevice.h:142-669).
__device__ __noinline__ int sos_orient4(const double *const *pin, const int *idx_in, int *depth_out,
                                        unsigned long long *cnt)
{
    int idx[5];
    const double *pt[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
        idx[i] = idx_in[i];
        pt[i] = pin[i];
    }
    int parity = 0;
    for (int i = 1; i < 5; i++)
        for (int j = i; j > 0 && idx[j] < idx[j - 1]; j--) {
            const int ti = idx[j]; idx[j] = idx[j - 1]; idx[j - 1] = ti;
            const double *tp = pt[j]; pt[j] = pt[j - 1]; pt[j - 1] = tp;
            parity ^= 1;
        }
    FixedPts<5> f;
    fixed_analyse<5>(pt, f);
    const int kk = f.bits <= 61 ? 2 : (f.bits <= 125 ? 4 : 0);
    if (kk == 0) pred_flag_error(cnt, G4D_DEV_ERR_PREDICATE_RANGE);
    int sign = 0, depth = 49;
    for (int d = 1; d <= 48; d++) {
        const uint32_t e = kSosTable[d];
        const int k = (e >> 28) & 7;
        const uint32_t rows = e & 0xffff, cols = (e >> 16) & 0xfff;
        int s = 0;
        if (kk == 2) s = exact_minor_k<2>(f, k, rows, cols);
        else if (kk == 4) s = exact_minor_k<4>(f, k, rows, cols);
        if (s != 0) {
            sign = (e >> 31) ? -s : s;
            depth = d;
            break;
        }
    }
    if (sign == 0) sign = 1;
    if (parity) sign = -sign;
    if (depth_out) *depth_out = depth;
    return sign;
}

// Filtered lifted determinant sign of (p0..p4) with the 4x4 form translated by p4.
// Returns 0 when the forward-error bound cannot certify the sign.  Every monomial carries
// at most 17 roundings (inputs 1+1+1+5 via the lift, minors 5, final sum 4), so the error
// is below 17.1 u times the computed permanent; 32 u is used.
__device__ __forceinline__ int filter_insphere(const double *const *p)
{
    double x[4], y[4], z[4], l[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        x[i] = p[i][0] - p[4][0];
        y[i] = p[i][1] - p[4][1];
        z[i] = p[i][2] - p[4][2];
        l[i] = x[i] * x[i] + y[i] * y[i] + z[i] * z[i];
    }
    double m2[4][4], q2[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = i + 1; j < 4; j++) {
            const double u = x[i] * y[j], v = x[j] * y[i];
            m2[i][j] = u - v;
            q2[i][j] = fabs(u) + fabs(v);
        }
    double m3[4], q3[4];
#pragma unroll
    for (int out = 0; out < 4; out++) {
        int r[3], q = 0;
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (i != out) r[q++] = i;
        m3[out] = z[r[0]] * m2[r[1]][r[2]] - z[r[1]] * m2[r[0]][r[2]] + z[r[2]] * m2[r[0]][r[1]];
        q3[out] = fabs(z[r[0]]) * q2[r[1]][r[2]] + fabs(z[r[1]]) * q2[r[0]][r[2]] + fabs(z[r[2]]) * q2[r[0]][r[1]];
    }
    const double det = -l[0] * m3[0] + l[1] * m3[1] - l[2] * m3[2] + l[3] * m3[3];
    const double perm = l[0] * q3[0] + l[1] * q3[1] + l[2] * q3[2] + l[3] * q3[3];
    const double bound = 32.0 * 1.1102230246251565e-16 * perm;
    if (!(perm > 1e-250)) return 0;
    if (det > bound) return 1;
    if (-det > bound) return -1;
    return 0;
}

__device__ __forceinline__ int filter_orient3d(const double *const *p)
{
    const double adx = p[0][0] - p[3][0], ady = p[0][1] - p[3][1], adz = p[0][2] - p[3][2];
    const double bdx = p[1][0] - p[3][0], bdy = p[1][1] - p[3][1], bdz = p[1][2] - p[3][2];
    const double cdx = p[2][0] - p[3][0], cdy = p[2][1] - p[3][1], cdz = p[2][2] - p[3][2];
    const double u0 = bdx * cdy, v0 = cdx * bdy, u1 = cdx * ady, v1 = adx * cdy, u2 = adx * bdy, v2 = bdx * ady;
    const double det = adz * (u0 - v0) + bdz * (u1 - v1) + cdz * (u2 - v2);
    const double perm = fabs(adz) * (fabs(u0) + fabs(v0)) + fabs(bdz) * (fabs(u1) + fabs(v1)) +
                        fabs(cdz) * (fabs(u2) + fabs(v2));
    const double bound = 12.0 * 1.1102230246251565e-16 * perm;
    if (!(perm > 1e-250)) return 0;
    if (det > bound) return 1;
    if (-det > bound) return -1;
    return 0;
}

// Full orient4 with exact fallback and SoS; never returns 0.
__device__ __forceinline__ int orient4_full(const double *const *p, const int *idx, unsigned long long *cnt)
{
    int s = filter_insphere(p);
    if (s != 0) return s;
    if (cnt) atomicAdd(&cnt[G4D_CNT_EXACT], 1ull);
    s = exact_insphere(p, cnt);
    if (s != 0) return s;
    if (cnt) atomicAdd(&cnt[G4D_CNT_SOS], 1ull);
    return sos_orient4(p, idx, nullptr, cnt);
}

__device__ __forceinline__ int orient3d_full(const double *const *p, unsigned long long *cnt)
{
    const int s = filter_orient3d(p);
    if (s != 0) return s;
    return exact_orient3d(p, cnt);
}

// Cofactors of the lifted determinant for triangle rows r0, r1, r2 (coordinates relative
// to the star point, 4th entry the lift).  C is signed so that orient4 = v' . C; P are the
// matching permanents (nonnegative).
__device__ __forceinline__ void tri_cofactors(const double *r0, const double *r1, const double *r2, double *C,
                                              double *P)
{
    // 2x2 minors of rows (r1, r2) over column pairs (p, q): r1p r2q - r1q r2p
    double m[4][4], pm[4][4];
#pragma unroll
    for (int p = 0; p < 4; p++)
#pragma unroll
        for (int q = p + 1; q < 4; q++) {
            const double u = r1[p] * r2[q], w = r1[q] * r2[p];
            m[p][q] = u - w;
            pm[p][q] = fabs(u) + fabs(w);
        }
    // M_x: cols (y,z,l); M_y: (x,z,l); M_z: (x,y,l); M_l: (x,y,z)
    const double Mx = r0[1] * m[2][3] - r0[2] * m[1][3] + r0[3] * m[1][2];
    const double My = r0[0] * m[2][3] - r0[2] * m[0][3] + r0[3] * m[0][2];
    const double Mz = r0[0] * m[1][3] - r0[1] * m[0][3] + r0[3] * m[0][1];
    const double Ml = r0[0] * m[1][2] - r0[1] * m[0][2] + r0[2] * m[0][1];
    P[0] = fabs(r0[1]) * pm[2][3] + fabs(r0[2]) * pm[1][3] + fabs(r0[3]) * pm[1][2];
    P[1] = fabs(r0[0]) * pm[2][3] + fabs(r0[2]) * pm[0][3] + fabs(r0[3]) * pm[0][2];
    P[2] = fabs(r0[0]) * pm[1][3] + fabs(r0[1]) * pm[0][3] + fabs(r0[3]) * pm[0][1];
    P[3] = fabs(r0[0]) * pm[1][2] + fabs(r0[1]) * pm[0][2] + fabs(r0[2]) * pm[0][1];
    C[0] = Mx;
    C[1] = -My;
    C[2] = Mz;
    C[3] = -Ml;
}

__device__ __forceinline__ void rel_coords(const double *pts, int v, double sx, double sy, double sz, double *r)
{
    const double x = pts[3 * (int64_t)v] - sx, y = pts[3 * (int64_t)v + 1] - sy, z = pts[3 * (int64_t)v + 2] - sz;
    r[0] = x;
    r[1] = y;
    r[2] = z;
    r[3] = x * x + y * y + z * z;
}

// Sign of orient4(a,b,c,s,v) = v'.C from cached cofactors, or 0 when the forward-error bound
// cannot certify it.  Each monomial of the 4x4 lifted determinant carries at most 17
// roundings (coordinates relative to s: 1 each; lifts: 5; 2x2 minors: 2; 3x3 minors: 3 more;
// the dot product: 4), so |error| < 17.1 u sum_j |v'_j| P_j with P_j the computed cofactor
// permanents; 32 u is used.  Fused multiply-adds only remove roundings from this count.
__device__ __forceinline__ int cofactor_filter_sign(const double *r, const double *C, const double *P)
{
    const double val = r[0] * C[0] + r[1] * C[1] + r[2] * C[2] + r[3] * C[3];
    const double per = fabs(r[0]) * P[0] + fabs(r[1]) * P[1] + fabs(r[2]) * P[2] + r[3] * P[3];
    const double bnd = 32.0 * 1.1102230246251565e-16 * per;
    if (!(per > 1e-250)) return 0;
    return val > bnd ? 1 : (-val > bnd ? -1 : 0);
}

} // namespace g4d
