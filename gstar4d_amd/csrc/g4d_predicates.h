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
// Runtime-width signed big integers (32-bit limbs, two's complement).  The exact path is
// rare, so it is written for a small register footprint: non-unrolled loops over private
// arrays, every primitive out of line.  Kernels that can reach it keep their own register
// budget instead of inheriting a wide unrolled multiprecision body.
// ------------------------------------------------------------------------------------

namespace rt {

constexpr int MAXL = 24; // 768 bits: the widest value (a 3x3 lifted minor at 125-bit inputs)

__device__ __noinline__ void sext(uint32_t *r, int nr, const uint32_t *a, int na)
{
    const uint32_t ext = (a[na - 1] >> 31) ? 0xffffffffu : 0u;
#pragma unroll 1
    for (int i = nr - 1; i >= 0; i--) r[i] = i < na ? a[i] : ext;
}

__device__ __noinline__ void add(uint32_t *r, const uint32_t *a, const uint32_t *b, int n)
{
    uint64_t c = 0;
#pragma unroll 1
    for (int i = 0; i < n; i++) {
        const uint64_t t = (uint64_t)a[i] + b[i] + c;
        r[i] = (uint32_t)t;
        c = t >> 32;
    }
}

__device__ __noinline__ void sub(uint32_t *r, const uint32_t *a, const uint32_t *b, int n)
{
    uint64_t c = 1;
#pragma unroll 1
    for (int i = 0; i < n; i++) {
        const uint64_t t = (uint64_t)a[i] + (uint32_t)~b[i] + c;
        r[i] = (uint32_t)t;
        c = t >> 32;
    }
}

__device__ __noinline__ void neg(uint32_t *r, int n)
{
    uint64_t c = 1;
#pragma unroll 1
    for (int i = 0; i < n; i++) {
        const uint64_t t = (uint64_t)(uint32_t)~r[i] + c;
        r[i] = (uint32_t)t;
        c = t >> 32;
    }
}

__device__ __noinline__ int sign(const uint32_t *a, int n)
{
    if (a[n - 1] >> 31) return -1;
    uint32_t o = 0;
#pragma unroll 1
    for (int i = 0; i < n; i++) o |= a[i];
    return o ? 1 : 0;
}

// r[na+nb] = a[na] * b[nb], exact and signed (|a| < 2^(32na-1), |b| < 2^(32nb-1)).
__device__ __noinline__ void mul(uint32_t *r, const uint32_t *a, int na, const uint32_t *b, int nb)
{
    uint32_t ma[MAXL], mb[MAXL];
    const bool sa = a[na - 1] >> 31, sb = b[nb - 1] >> 31;
#pragma unroll 1
    for (int i = 0; i < na; i++) ma[i] = a[i];
#pragma unroll 1
    for (int i = 0; i < nb; i++) mb[i] = b[i];
    if (sa) neg(ma, na);
    if (sb) neg(mb, nb);
#pragma unroll 1
    for (int i = 0; i < na + nb; i++) r[i] = 0;
#pragma unroll 1
    for (int i = 0; i < na; i++) {
        uint64_t carry = 0;
#pragma unroll 1
        for (int j = 0; j < nb; j++) {
            const uint64_t t = (uint64_t)ma[i] * mb[j] + r[i + j] + carry;
            r[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        r[i + nb] = (uint32_t)carry;
    }
    if (sa != sb) neg(r, na + nb);
}

// r[n] = a*b - c*d with a, b, c, d of na limbs each (n = 2na).
__device__ __noinline__ void det2(uint32_t *r, const uint32_t *a, const uint32_t *b, const uint32_t *c,
                                  const uint32_t *d, int na)
{
    uint32_t t[rt::MAXL];
    mul(r, a, na, b, na);
    mul(t, c, na, d, na);
    sub(r, r, t, 2 * na);
}

} // namespace rt

// ------------------------------------------------------------------------------------
// Fixed-width signed integers of N 64-bit limbs (two's complement), fully unrolled in
// registers.  Used when every coordinate of a predicate fits 61 bits on the common scale
// (the usual case: FP64 inputs spanning fewer than 8 binades), where the widths below are
// known at compile time; wider inputs take the runtime-width path above.
// ------------------------------------------------------------------------------------

namespace fx {

typedef unsigned __int128 u128;

template <int N>
struct I {
    uint64_t w[N];
};

template <int N>
__device__ __forceinline__ I<N> from64(int64_t v)
{
    I<N> r;
    r.w[0] = (uint64_t)v;
#pragma unroll
    for (int i = 1; i < N; i++) r.w[i] = v < 0 ? ~0ull : 0ull;
    return r;
}

template <int N, int M>
__device__ __forceinline__ I<N> ext(const I<M> &a)
{
    I<N> r;
    const uint64_t e = (a.w[M - 1] >> 63) ? ~0ull : 0ull;
#pragma unroll
    for (int i = 0; i < N; i++) r.w[i] = i < M ? a.w[i] : e;
    return r;
}

template <int N>
__device__ __forceinline__ I<N> add(const I<N> &a, const I<N> &b)
{
    I<N> r;
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < N; i++) {
        const u128 t = (u128)a.w[i] + b.w[i] + c;
        r.w[i] = (uint64_t)t;
        c = (uint64_t)(t >> 64);
    }
    return r;
}

template <int N>
__device__ __forceinline__ I<N> sub(const I<N> &a, const I<N> &b)
{
    I<N> r;
    uint64_t c = 1;
#pragma unroll
    for (int i = 0; i < N; i++) {
        const u128 t = (u128)a.w[i] + (~b.w[i]) + c;
        r.w[i] = (uint64_t)t;
        c = (uint64_t)(t >> 64);
    }
    return r;
}

template <int N>
__device__ __forceinline__ I<N> neg(const I<N> &a)
{
    I<N> r;
    uint64_t c = 1;
#pragma unroll
    for (int i = 0; i < N; i++) {
        const u128 t = (u128)(~a.w[i]) + c;
        r.w[i] = (uint64_t)t;
        c = (uint64_t)(t >> 64);
    }
    return r;
}

template <int N>
__device__ __forceinline__ int sign(const I<N> &a)
{
    if (a.w[N - 1] >> 63) return -1;
    uint64_t o = 0;
#pragma unroll
    for (int i = 0; i < N; i++) o |= a.w[i];
    return o ? 1 : 0;
}

// Exact signed product (|a| < 2^(64NA-1), |b| < 2^(64NB-1)): magnitudes, schoolbook, sign.
template <int NA, int NB>
__device__ __forceinline__ I<NA + NB> mul(const I<NA> &a, const I<NB> &b)
{
    const bool sa = a.w[NA - 1] >> 63, sb = b.w[NB - 1] >> 63;
    const I<NA> ma = sa ? neg(a) : a;
    const I<NB> mb = sb ? neg(b) : b;
    I<NA + NB> r;
#pragma unroll
    for (int i = 0; i < NA + NB; i++) r.w[i] = 0;
#pragma unroll
    for (int i = 0; i < NA; i++) {
        uint64_t carry = 0;
#pragma unroll
        for (int j = 0; j < NB; j++) {
            const u128 t = (u128)ma.w[i] * mb.w[j] + r.w[i + j] + carry;
            r.w[i + j] = (uint64_t)t;
            carry = (uint64_t)(t >> 64);
        }
        r.w[i + NB] = carry;
    }
    return sa != sb ? neg(r) : r;
}

template <int N>
__device__ __forceinline__ I<2 * N> det2(const I<N> &a, const I<N> &b, const I<N> &c, const I<N> &d)
{
    return sub(mul(a, b), mul(c, d));
}

} // namespace fx

// ------------------------------------------------------------------------------------
// Fixed-point conversion of the FP64 inputs of one predicate
// ------------------------------------------------------------------------------------

// Decompose finite nonzero x = m * 2^e with m an integer of at most 53 bits whose trailing
// zero bits are moved into e.  Returns false for zero.
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

// Integer coordinates of up to 5 points (3 each) on a common scale 2^emin.
struct FixedPts {
    int64_t m[5][3];
    int e[5][3];
    int emin;
    int bits; // bits of the largest magnitude after scaling
    int limbs; // 2 (bits <= 61) or 4 (bits <= 125); 0 when out of range
};

__device__ __noinline__ void fixed_analyse(const double *const *p, int np, FixedPts &f)
{
    f.emin = 1 << 20;
    int top = -(1 << 20);
#pragma unroll 1
    for (int i = 0; i < np; i++)
#pragma unroll 1
        for (int c = 0; c < 3; c++) {
            int64_t m;
            int e;
            if (!fp_decompose(p[i][c], &m, &e)) {
                f.m[i][c] = 0;
                f.e[i][c] = 0;
                continue;
            }
            f.m[i][c] = m;
            f.e[i][c] = e;
            f.emin = min(f.emin, e);
            const uint64_t am = (uint64_t)(m < 0 ? -m : m);
            top = max(top, e + 64 - __builtin_clzll(am));
        }
    f.bits = (top == -(1 << 20)) ? 0 : top - f.emin;
    f.limbs = f.bits <= 61 ? 2 : (f.bits <= 125 ? 4 : 0);
}

// Coordinate (i, c) as a K-limb signed integer scaled by 2^-emin.
__device__ __noinline__ void fixed_coord(const FixedPts &f, int i, int c, int K, uint32_t *r)
{
#pragma unroll 1
    for (int q = 0; q < K; q++) r[q] = 0;
    const int64_t mm = f.m[i][c];
    if (mm == 0) return;
    const uint64_t am = (uint64_t)(mm < 0 ? -mm : mm);
    const int sh = f.e[i][c] - f.emin; // >= 0; am << sh < 2^(32K-3) by the range check
    const int wq = sh >> 5, bq = sh & 31;
    const uint64_t lo64 = am << bq;
    const uint64_t hi64 = bq ? (am >> (64 - bq)) : 0;
    if (wq < K) r[wq] = (uint32_t)lo64;
    if (wq + 1 < K) r[wq + 1] = (uint32_t)(lo64 >> 32);
    if (wq + 2 < K) r[wq + 2] = (uint32_t)hi64;
    if (mm < 0) rt::neg(r, K);
}

// ------------------------------------------------------------------------------------
// Exact determinants on K-limb coordinates (K = 2 or 4).  With |coordinate| < 2^(32K-3):
// differences fit K limbs, 2x2 minors 2K, 3x3 minors 3K, lifts 2K, the lifted 4x4
// determinant 5K; for SoS minors every entry (coordinate or lift, translated) fits 2K limbs,
// so 2x2 minors fit 4K and 3x3 minors 6K <= 24 limbs.
// ------------------------------------------------------------------------------------

// det[[p |p|^2 1]] over rows 0..4, translated by row 4: det4[[d |d|^2]], d = p_i - p_4.
// Its sign equals Shewchuk's insphere(p0, p1, p2, p3, p4).
__device__ __noinline__ int exact_insphere_rt(const FixedPts &f)
{
    const int K = f.limbs;
    uint32_t d[4][3][4], e[4];
#pragma unroll 1
    for (int c = 0; c < 3; c++) {
        fixed_coord(f, 4, c, K, e);
#pragma unroll 1
        for (int i = 0; i < 4; i++) {
            fixed_coord(f, i, c, K, d[i][c]);
            rt::sub(d[i][c], d[i][c], e, K);
        }
    }
    // 2x2 xy minors for the pairs (i<j), index i*4+j
    uint32_t m2[16][8];
#pragma unroll 1
    for (int i = 0; i < 4; i++)
#pragma unroll 1
        for (int j = i + 1; j < 4; j++) rt::det2(m2[4 * i + j], d[i][0], d[j][1], d[j][0], d[i][1], K);
    // lifted column expansion: det4 = sum_i (-1)^(i+3) l_i M3_i, M3_i the xyz minor without row i
    uint32_t acc[20], m3[12], t[rt::MAXL], l[8], sq[8];
#pragma unroll 1
    for (int q = 0; q < 5 * K; q++) acc[q] = 0;
#pragma unroll 1
    for (int out = 0; out < 4; out++) {
        int r[3], n = 0;
#pragma unroll 1
        for (int i = 0; i < 4; i++)
            if (i != out) r[n++] = i;
        rt::mul(m3, d[r[0]][2], K, m2[4 * r[1] + r[2]], 2 * K);
        rt::mul(t, d[r[1]][2], K, m2[4 * r[0] + r[2]], 2 * K);
        rt::sub(m3, m3, t, 3 * K);
        rt::mul(t, d[r[2]][2], K, m2[4 * r[0] + r[1]], 2 * K);
        rt::add(m3, m3, t, 3 * K);
        rt::mul(l, d[out][0], K, d[out][0], K);
        rt::mul(sq, d[out][1], K, d[out][1], K);
        rt::add(l, l, sq, 2 * K);
        rt::mul(sq, d[out][2], K, d[out][2], K);
        rt::add(l, l, sq, 2 * K);
        rt::mul(t, l, 2 * K, m3, 3 * K);
        if (out & 1) rt::add(acc, acc, t, 5 * K);
        else rt::sub(acc, acc, t, 5 * K);
    }
    return rt::sign(acc, 5 * K);
}

// orient3d exact: det[[p 1]] over rows 0..3, translated by row 3.
__device__ __noinline__ int exact_orient3d_rt(const FixedPts &f)
{
    const int K = f.limbs;
    uint32_t d[3][3][4], e[4];
#pragma unroll 1
    for (int c = 0; c < 3; c++) {
        fixed_coord(f, 3, c, K, e);
#pragma unroll 1
        for (int i = 0; i < 3; i++) {
            fixed_coord(f, i, c, K, d[i][c]);
            rt::sub(d[i][c], d[i][c], e, K);
        }
    }
    uint32_t m[8], t[rt::MAXL], acc[12];
    rt::det2(m, d[1][0], d[2][1], d[2][0], d[1][1], K);
    rt::mul(acc, d[0][2], K, m, 2 * K);
    rt::det2(m, d[0][0], d[2][1], d[2][0], d[0][1], K);
    rt::mul(t, d[1][2], K, m, 2 * K);
    rt::sub(acc, acc, t, 3 * K);
    rt::det2(m, d[0][0], d[1][1], d[1][0], d[0][1], K);
    rt::mul(t, d[2][2], K, m, 2 * K);
    rt::add(acc, acc, t, 3 * K);
    return rt::sign(acc, 3 * K);
}

// SoS minor: sign det[[cols... 1]] over k rows (k = 2..4) of the fixed points; rows/cols are
// packed nibbles (row i = (rows >> 4i) & 15, col c = (cols >> 4c) & 15, col 3 = |p|^2).
__device__ __noinline__ int exact_minor_rt(const FixedPts &f, int k, uint32_t rows, uint32_t cols)
{
    const int K = f.limbs, NE = 2 * K;
    uint32_t a[4][3][8], x[4], y[4], z[4], t[rt::MAXL];
#pragma unroll 1
    for (int i = 0; i < k; i++) {
        const int r = (rows >> (4 * i)) & 15;
#pragma unroll 1
        for (int j = 0; j < k - 1; j++) {
            const int c = (cols >> (4 * j)) & 15;
            if (c < 3) {
                fixed_coord(f, r, c, K, x);
                rt::sext(a[i][j], NE, x, K);
            } else {
                fixed_coord(f, r, 0, K, x);
                fixed_coord(f, r, 1, K, y);
                fixed_coord(f, r, 2, K, z);
                rt::mul(a[i][j], x, K, x, K);
                rt::mul(t, y, K, y, K);
                rt::add(a[i][j], a[i][j], t, NE);
                rt::mul(t, z, K, z, K);
                rt::add(a[i][j], a[i][j], t, NE);
            }
        }
    }
    // translate by the last row (the ones column drops out)
#pragma unroll 1
    for (int i = 0; i < k - 1; i++)
#pragma unroll 1
        for (int j = 0; j < k - 1; j++) rt::sub(a[i][j], a[i][j], a[k - 1][j], NE);
    if (k == 2) return rt::sign(a[0][0], NE);
    uint32_t m[16], acc[rt::MAXL];
    if (k == 3) {
        rt::det2(m, a[0][0], a[1][1], a[1][0], a[0][1], NE);
        return rt::sign(m, 2 * NE);
    }
    rt::det2(m, a[1][0], a[2][1], a[2][0], a[1][1], NE);
    rt::mul(acc, a[0][2], NE, m, 2 * NE);
    rt::det2(m, a[0][0], a[2][1], a[2][0], a[0][1], NE);
    rt::mul(t, a[1][2], NE, m, 2 * NE);
    rt::sub(acc, acc, t, 3 * NE);
    rt::det2(m, a[0][0], a[1][1], a[1][0], a[0][1], NE);
    rt::mul(t, a[2][2], NE, m, 2 * NE);
    rt::add(acc, acc, t, 3 * NE);
    return rt::sign(acc, 3 * NE);
}

// ------------------------------------------------------------------------------------
// The same determinants at compile-time widths for coordinates of at most 61 bits on the
// common scale (FixedPts::limbs == 2): a coordinate is one 64-bit limb, a translated
// coordinate fits 62 bits, a 2x2 minor 125, a 3x3 minor 189, a lift 126 and the lifted 4x4
// determinant 317 bits.
// ------------------------------------------------------------------------------------

__device__ __forceinline__ int64_t fixed_coord64(const FixedPts &f, int i, int c)
{
    const int64_t mm = f.m[i][c];
    return mm == 0 ? 0 : mm * ((int64_t)1 << (f.e[i][c] - f.emin));
}

// Translated coordinates d[i] = p_i - p_last (i < np - 1) as 64-bit integers when every one fits
// 62 bits.  Coordinates of up to 125 bits are formed exactly in 128-bit integers, so clustered
// points whose absolute coordinates span many binades still take the fixed-width path.
__device__ __forceinline__ bool fixed_diffs64(const FixedPts &f, int np, int64_t (*d)[3])
{
    const __int128 lim = (__int128)1 << 62;
#pragma unroll 1
    for (int c = 0; c < 3; c++) {
        const int64_t ml = f.m[np - 1][c];
        const __int128 xl = ml == 0 ? 0 : (__int128)ml * ((__int128)1 << (f.e[np - 1][c] - f.emin));
#pragma unroll 1
        for (int i = 0; i < np - 1; i++) {
            const int64_t mi = f.m[i][c];
            const __int128 xi = mi == 0 ? 0 : (__int128)mi * ((__int128)1 << (f.e[i][c] - f.emin));
            const __int128 x = xi - xl;
            if (x >= lim || x <= -lim) return false;
            d[i][c] = (int64_t)x;
        }
    }
    return true;
}

// Lifted 4x4 determinant of translated coordinates of up to 62 bits (sign of insphere).
__device__ __noinline__ int exact_insphere_fx(const int64_t (*dd)[3])
{
    fx::I<1> d[4][3];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int c = 0; c < 3; c++) d[i][c] = fx::from64<1>(dd[i][c]);
    fx::I<5> acc = fx::from64<5>(0);
#pragma unroll
    for (int out = 0; out < 4; out++) {
        const int r0 = out == 0 ? 1 : 0, r1 = out <= 1 ? 2 : 1, r2 = out <= 2 ? 3 : 2;
        const fx::I<3> m3 = fx::add(fx::sub(fx::mul(d[r0][2], fx::det2(d[r1][0], d[r2][1], d[r2][0], d[r1][1])),
                                            fx::mul(d[r1][2], fx::det2(d[r0][0], d[r2][1], d[r2][0], d[r0][1]))),
                                    fx::mul(d[r2][2], fx::det2(d[r0][0], d[r1][1], d[r1][0], d[r0][1])));
        const fx::I<2> l = fx::add(fx::add(fx::mul(d[out][0], d[out][0]), fx::mul(d[out][1], d[out][1])),
                                   fx::mul(d[out][2], d[out][2]));
        const fx::I<5> t = fx::mul(l, m3);
        acc = (out & 1) ? fx::add(acc, t) : fx::sub(acc, t);
    }
    return fx::sign(acc);
}

// 3x3 determinant of translated coordinates of up to 62 bits (sign of orient3d).
__device__ __noinline__ int exact_orient3d_fx(const int64_t (*dd)[3])
{
    fx::I<1> d[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int c = 0; c < 3; c++) d[i][c] = fx::from64<1>(dd[i][c]);
    const fx::I<3> acc = fx::add(fx::sub(fx::mul(d[0][2], fx::det2(d[1][0], d[2][1], d[2][0], d[1][1])),
                                         fx::mul(d[1][2], fx::det2(d[0][0], d[2][1], d[2][0], d[0][1]))),
                                 fx::mul(d[2][2], fx::det2(d[0][0], d[1][1], d[1][0], d[0][1])));
    return fx::sign(acc);
}

// SoS minor (as exact_minor_rt) with every entry, coordinate or lift, in two limbs.
__device__ __noinline__ int exact_minor_fx(const FixedPts &f, int k, uint32_t rows, uint32_t cols)
{
    fx::I<2> a[4][3];
#pragma unroll 1
    for (int i = 0; i < k; i++) {
        const int r = (rows >> (4 * i)) & 15;
#pragma unroll 1
        for (int j = 0; j < k - 1; j++) {
            const int c = (cols >> (4 * j)) & 15;
            if (c < 3) {
                a[i][j] = fx::from64<2>(fixed_coord64(f, r, c));
            } else {
                const fx::I<1> x = fx::from64<1>(fixed_coord64(f, r, 0)), y = fx::from64<1>(fixed_coord64(f, r, 1)),
                               z = fx::from64<1>(fixed_coord64(f, r, 2));
                a[i][j] = fx::add(fx::add(fx::mul(x, x), fx::mul(y, y)), fx::mul(z, z));
            }
        }
    }
#pragma unroll 1
    for (int i = 0; i < k - 1; i++)
#pragma unroll 1
        for (int j = 0; j < k - 1; j++) a[i][j] = fx::sub(a[i][j], a[k - 1][j]);
    if (k == 2) return fx::sign(a[0][0]);
    if (k == 3) return fx::sign(fx::det2(a[0][0], a[1][1], a[1][0], a[0][1]));
    const fx::I<6> acc =
        fx::add(fx::sub(fx::mul(a[0][2], fx::det2(a[1][0], a[2][1], a[2][0], a[1][1])),
                        fx::mul(a[1][2], fx::det2(a[0][0], a[2][1], a[2][0], a[0][1]))),
                fx::mul(a[2][2], fx::det2(a[0][0], a[1][1], a[1][0], a[0][1])));
    return fx::sign(acc);
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
    FixedPts f;
    fixed_analyse(p, 5, f);
    int64_t d[4][3];
    if (f.limbs && fixed_diffs64(f, 5, d)) return exact_insphere_fx(d);
    if (f.limbs) return exact_insphere_rt(f);
    pred_flag_error(cnt, G4D_DEV_ERR_PREDICATE_RANGE);
    return 0;
}

__device__ __noinline__ int exact_orient3d(const double *const *p, unsigned long long *cnt)
{
    FixedPts f;
    fixed_analyse(p, 4, f);
    int64_t d[3][3];
    if (f.limbs && fixed_diffs64(f, 4, d)) return exact_orient3d_fx(d);
    if (f.limbs) return exact_orient3d_rt(f);
    pred_flag_error(cnt, G4D_DEV_ERR_PREDICATE_RANGE);
    return 0;
}

// SoS resolution of a zero lifted determinant (GD/GDelaunay/GDelPredDevice.h:142-669).
__device__ __noinline__ int sos_orient4(const double *const *pin, const int *idx_in, int *depth_out,
                                        unsigned long long *cnt)
{
    int idx[5];
    const double *pt[5];
#pragma unroll 1
    for (int i = 0; i < 5; i++) {
        idx[i] = idx_in[i];
        pt[i] = pin[i];
    }
    int parity = 0;
#pragma unroll 1
    for (int i = 1; i < 5; i++)
#pragma unroll 1
        for (int j = i; j > 0 && idx[j] < idx[j - 1]; j--) {
            const int ti = idx[j]; idx[j] = idx[j - 1]; idx[j - 1] = ti;
            const double *tp = pt[j]; pt[j] = pt[j - 1]; pt[j - 1] = tp;
            parity ^= 1;
        }
    FixedPts f;
    fixed_analyse(pt, 5, f);
    if (!f.limbs) pred_flag_error(cnt, G4D_DEV_ERR_PREDICATE_RANGE);
    int sign = 0, depth = 49;
#pragma unroll 1
    for (int d = 1; d <= 48 && f.limbs; d++) {
        const uint32_t e = kSosTable[d];
        const int k = (e >> 28) & 7;
        const int s = f.limbs == 2 ? exact_minor_fx(f, k, e & 0xffff, (e >> 16) & 0xfff)
                                   : exact_minor_rt(f, k, e & 0xffff, (e >> 16) & 0xfff);
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
__device__ __forceinline__ int filter_insphere_xyzl(const double *x, const double *y, const double *z, const double *l);
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
    return filter_insphere_xyzl(x, y, z, l);
}

// filter_insphere from rows r[i] = (x, y, z, x^2 + y^2 + z^2) of p0..p3 already translated by
// p4 with the same operations as above (rel_coords computes them so): the same sign, bit for bit.
__device__ __forceinline__ int filter_insphere_rel(const double *const *r)
{
    double x[4], y[4], z[4], l[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        x[i] = r[i][0];
        y[i] = r[i][1];
        z[i] = r[i][2];
        l[i] = r[i][3];
    }
    return filter_insphere_xyzl(x, y, z, l);
}

__device__ __forceinline__ int filter_insphere_xyzl(const double *x, const double *y, const double *z, const double *l)
{
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
    // 2x2 minors of rows (r1, r2) over column pairs (p, q): r1p r2q - r1q r2p, and their
    // permanents |r1p r2q| + |r1q r2p|, each with one rounded product and one fused step
    double m[4][4], pm[4][4];
#pragma unroll
    for (int p = 0; p < 4; p++)
#pragma unroll
        for (int q = p + 1; q < 4; q++) {
            const double w = r1[q] * r2[p];
            m[p][q] = fma(r1[p], r2[q], -w);
            pm[p][q] = fma(fabs(r1[p]), fabs(r2[q]), fabs(w));
        }
    // M_x: cols (y,z,l); M_y: (x,z,l); M_z: (x,y,l); M_l: (x,y,z)
    const double Mx = fma(r0[3], m[1][2], fma(-r0[2], m[1][3], r0[1] * m[2][3]));
    const double My = fma(r0[3], m[0][2], fma(-r0[2], m[0][3], r0[0] * m[2][3]));
    const double Mz = fma(r0[3], m[0][1], fma(-r0[1], m[0][3], r0[0] * m[1][3]));
    const double Ml = fma(r0[2], m[0][1], fma(-r0[1], m[0][2], r0[0] * m[1][2]));
    P[0] = fma(fabs(r0[3]), pm[1][2], fma(fabs(r0[2]), pm[1][3], fabs(r0[1]) * pm[2][3]));
    P[1] = fma(fabs(r0[3]), pm[0][2], fma(fabs(r0[2]), pm[0][3], fabs(r0[0]) * pm[2][3]));
    P[2] = fma(fabs(r0[3]), pm[0][1], fma(fabs(r0[1]), pm[0][3], fabs(r0[0]) * pm[1][3]));
    P[3] = fma(fabs(r0[2]), pm[0][1], fma(fabs(r0[1]), pm[0][2], fabs(r0[0]) * pm[1][2]));
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
    const double val = fma(r[3], C[3], fma(r[2], C[2], fma(r[1], C[1], r[0] * C[0])));
    const double per = fma(r[3], P[3], fma(fabs(r[2]), P[2], fma(fabs(r[1]), P[1], fabs(r[0]) * P[0])));
    const double bnd = 32.0 * 1.1102230246251565e-16 * per;
    if (!(per > 1e-250)) return 0;
    return val > bnd ? 1 : (-val > bnd ? -1 : 0);
}

} // namespace g4d
