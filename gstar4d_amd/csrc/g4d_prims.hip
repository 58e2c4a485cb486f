// gStar4D for MI355X — device primitives.
//
// The reference leans on thrust for sort_by_key / exclusive_scan / remove_if
// (GD/GDelaunay/GDelCommon.cu:85-277).  Here: a three-phase tiled scan, a stable LSD radix
// sort whose per-wave ranking uses ballot-matched digit peers (8 ballots per 8-bit digit),
// order-preserving compaction by tile counts and ballot ranks, a merge of a long and a short
// sorted list, and single-work-group versions of the sort and the compaction for short lists.
#include "g4d_prims.h"

#include <algorithm>

namespace g4d {

static constexpr int SCAN_THREADS = 256;
static constexpr int SCAN_ITEMS = 8;
static constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

template <typename T>
__device__ __forceinline__ T wave_incl_scan_t(T v)
{
    const int lane = __lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const T o = __shfl_up(v, off, 64);
        if (lane >= off) v += o;
    }
    return v;
}

// Block-wide exclusive scan over 256 threads; returns exclusive prefix, total in *tot.
template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T *lds4, T *tot)
{
    const int lane = __lane_id(), wid = threadIdx.x >> 6;
    const T inc = wave_incl_scan_t<T>(v);
    if (lane == 63) lds4[wid] = inc;
    __syncthreads();
    T wbase = 0, all = 0;
#pragma unroll
    for (int w = 0; w < SCAN_THREADS / 64; w++) {
        const T x = lds4[w];
        if (w < wid) wbase += x;
        all += x;
    }
    __syncthreads();
    *tot = all;
    return wbase + inc - v;
}

template <typename T>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_reduce(const T *in, int64_t n, T *bsum)
{
    __shared__ T lds[4];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    T s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        const int64_t k = base + (int64_t)i * SCAN_THREADS + threadIdx.x;
        if (k < n) s += in[k];
    }
    T tot;
    (void)block_excl_scan<T>(s, lds, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

template <typename T>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_apply(const T *in, T *out, int64_t n,
                                                             const T *boff, T *total)
{
    __shared__ T lds[4];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    T v[SCAN_ITEMS];
    T s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        const int64_t k = base + i;
        v[i] = (k < n) ? in[k] : (T)0;
        s += v[i];
    }
    T tot;
    T run = block_excl_scan<T>(s, lds, &tot) + (boff ? boff[blockIdx.x] : (T)0);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        const int64_t k = base + i;
        const T x = v[i];
        if (k < n) out[k] = run;
        run += x;
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_THREADS - 1) *total = run;
}

template <typename T>
static hipError_t scan_excl_impl(Prims &P, const T *in, T *out, int64_t n, T *total, hipStream_t s, int depth)
{
    if (n <= 0) {
        if (total) return hipMemsetAsync(total, 0, sizeof(T), s);
        return hipSuccess;
    }
    const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb == 1) {
        hipLaunchKernelGGL(k_scan_apply<T>, dim3(1), dim3(SCAN_THREADS), 0, s, in, out, n, (const T *)nullptr, total);
        return hipGetLastError();
    }
    if (depth >= 4) return hipErrorInvalidValue;
    DevBuf &tmp = P.scan_tmp[depth];
    HIP_CHECK_RET(tmp.ensure(sizeof(T) * (size_t)nb));
    T *bs = tmp.as<T>();
    hipLaunchKernelGGL(k_scan_reduce<T>, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, s, in, n, bs);
    HIP_CHECK_RET(hipGetLastError());
    HIP_CHECK_RET(scan_excl_impl<T>(P, bs, bs, nb, (T *)nullptr, s, depth + 1));
    hipLaunchKernelGGL(k_scan_apply<T>, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, s, in, out, n, (const T *)bs, total);
    return hipGetLastError();
}

hipError_t scan_excl_i64(Prims &P, const int64_t *in, int64_t *out, int64_t n, int64_t *total, hipStream_t s)
{
    return scan_excl_impl<int64_t>(P, in, out, n, total, s, 0);
}

hipError_t scan_excl_u32(Prims &P, const uint32_t *in, uint32_t *out, int64_t n, hipStream_t s)
{
    return scan_excl_impl<uint32_t>(P, in, out, n, (uint32_t *)nullptr, s, 0);
}

// ------------------------------------------------------------------------------------
// Radix sort
// ------------------------------------------------------------------------------------

static constexpr int RS_THREADS = 256;
static constexpr int RS_ITEMS = 8;
static constexpr int RS_TILE = RS_THREADS * RS_ITEMS; // 2048 keys per block
static constexpr int RS_WAVES = RS_THREADS / 64;

__global__ void __launch_bounds__(RS_THREADS) k_rs_hist(const uint64_t *keys, int64_t n, int shift, uint32_t *hist,
                                                        int nblocks)
{
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll
    for (int i = 0; i < RS_ITEMS; i++) {
        const int64_t k = base + (int64_t)i * RS_THREADS + threadIdx.x;
        if (k < n) atomicAdd(&cnt[(keys[k] >> shift) & 255], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// Each wave owns RS_ITEMS consecutive rounds of 64 keys (wave-contiguous sub-tile), so
// (wave, round, lane) order equals input order: ranking in that order is stable.  SELF: the
// histogram is raw (few tiles) and each work-group derives its own digit offsets from it,
// which saves the scan's launches.
static constexpr int RS_SELF_MAX_TILES = 128;

template <bool SELF>
__global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(const uint64_t *keys, uint64_t *out, int64_t n, int shift,
                                                           const uint32_t *hist_scan, int nblocks)
{
    __shared__ uint32_t cnt[RS_WAVES][256];
    __shared__ uint32_t goff[256];
    __shared__ uint32_t part[RS_WAVES];
    const int lane = __lane_id(), wid = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < RS_WAVES * 256; i += RS_THREADS) (&cnt[0][0])[i] = 0;
    if (SELF) {
        // digit d: keys of digit d in earlier tiles, plus all keys of smaller digits
        const uint32_t *h = hist_scan + (int64_t)threadIdx.x * nblocks;
        uint32_t before = 0, total = 0;
        for (int b = 0; b < nblocks; b++) {
            const uint32_t c = h[b];
            if (b < (int)blockIdx.x) before += c;
            total += c;
        }
        uint32_t inc = total;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(inc, o, 64);
            if (lane >= o) inc += t;
        }
        if (lane == 63) part[wid] = inc;
        __syncthreads();
        uint32_t wb = 0;
        for (int w = 0; w < wid; w++) wb += part[w];
        goff[threadIdx.x] = wb + inc - total + before;
    } else {
        goff[threadIdx.x] = hist_scan[(int64_t)threadIdx.x * nblocks + blockIdx.x];
    }
    __syncthreads();
    const int64_t wbase = (int64_t)blockIdx.x * RS_TILE + (int64_t)wid * (RS_ITEMS * 64);
    uint64_t kv[RS_ITEMS];
    uint32_t rank[RS_ITEMS];
    int dg[RS_ITEMS];
    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        const int64_t k = wbase + (int64_t)r * 64 + lane;
        const bool valid = k < n;
        kv[r] = valid ? keys[k] : 0ull;
        const int d = valid ? (int)((kv[r] >> shift) & 255) : 256;
        dg[r] = d;
        uint64_t peers = __ballot(1);
#pragma unroll
        for (int b = 0; b < 9; b++) {
            const bool bit = (d >> b) & 1;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        uint32_t before = 0;
        if (valid) before = cnt[wid][d];
        const uint32_t inwave = (uint32_t)__popcll(peers & lt);
        rank[r] = before + inwave;
        // lowest lane of each peer group bumps the counter after everyone read it
        __builtin_amdgcn_wave_barrier();
        if (valid && (peers & lt) == 0) cnt[wid][d] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // exclusive prefix over waves, per digit
    {
        const int d = threadIdx.x;
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < RS_WAVES; w++) {
            const uint32_t t = cnt[w][d];
            cnt[w][d] = run;
            run += t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        const int d = dg[r];
        if (d < 256) out[goff[d] + cnt[wid][d] + rank[r]] = kv[r];
    }
}

// Lists up to SMALL_SORT keys are sorted by one work-group in a single launch: per 8-bit
// digit a histogram sweep and a scatter sweep over tiles of 1024 keys in order (ballot-matched
// ranks within a wave, wave offsets per digit, running digit offsets), ping-ponging between
// keys and tmp; the result ends in keys.
static constexpr int64_t SMALL_SORT = 3072;

__global__ void __launch_bounds__(1024) k_sort_small(uint64_t *keys, uint64_t *tmp, int n, int lo_bit, int hi_bit,
                                                     int lo2, int hi2)
{
    __shared__ uint32_t goff[256];
    __shared__ uint32_t wcnt[16][256];
    __shared__ uint32_t part[4];
    const int lane = __lane_id(), wid = threadIdx.x >> 6;
    const uint64_t lt = (1ull << lane) - 1ull;
    uint64_t *src = keys, *dst = tmp;
    // digits of [lo_bit, hi_bit), then of [lo2, hi2)
    const int np1 = hi_bit > lo_bit ? (hi_bit - lo_bit + 7) / 8 : 0, np2 = hi2 > lo2 ? (hi2 - lo2 + 7) / 8 : 0;
    for (int pass = 0; pass < np1 + np2; pass++) {
        const int shift = pass < np1 ? lo_bit + 8 * pass : lo2 + 8 * (pass - np1);
        if (threadIdx.x < 256) goff[threadIdx.x] = 0;
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += 1024) atomicAdd(&goff[(src[i] >> shift) & 255], 1u);
        __syncthreads();
        // exclusive scan of the 256 digit counts (4 waves)
        uint32_t c = 0, inc = 0;
        if (threadIdx.x < 256) {
            c = goff[threadIdx.x];
            inc = c;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t t = __shfl_up(inc, o, 64);
                if (lane >= o) inc += t;
            }
            if (lane == 63) part[wid] = inc;
        }
        __syncthreads();
        if (threadIdx.x < 256) {
            uint32_t b = 0;
            for (int w = 0; w < wid; w++) b += part[w];
            goff[threadIdx.x] = b + inc - c;
        }
        for (int t0 = 0; t0 < n; t0 += 1024) {
            for (int i = threadIdx.x; i < 16 * 256; i += 1024) (&wcnt[0][0])[i] = 0;
            __syncthreads();
            const int k = t0 + threadIdx.x;
            const bool valid = k < n;
            const uint64_t key = valid ? src[k] : 0ull;
            const int d = valid ? (int)((key >> shift) & 255) : 256;
            uint64_t peers = __ballot(1);
#pragma unroll
            for (int b = 0; b < 9; b++) {
                const bool bit = (d >> b) & 1;
                const uint64_t bal = __ballot(bit);
                peers &= bit ? bal : ~bal;
            }
            const uint32_t rank = (uint32_t)__popcll(peers & lt);
            if (valid && (peers & lt) == 0) wcnt[wid][d] = (uint32_t)__popcll(peers);
            __syncthreads();
            if (threadIdx.x < 256) { // exclusive prefix over waves per digit; advance the digit offset
                uint32_t r = 0;
#pragma unroll
                for (int w = 0; w < 16; w++) {
                    const uint32_t t = wcnt[w][threadIdx.x];
                    wcnt[w][threadIdx.x] = r;
                    r += t;
                }
                c = r; // this tile's count of the digit
            }
            __syncthreads();
            if (valid) dst[goff[d] + wcnt[wid][d] + rank] = key;
            __syncthreads();
            if (threadIdx.x < 256) goff[threadIdx.x] += c;
        }
        __syncthreads();
        uint64_t *t = src;
        src = dst;
        dst = t;
    }
    if (src != keys) {
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += 1024) keys[i] = src[i];
    }
}

hipError_t radix_sort_u64(Prims &P, uint64_t *keys, int64_t n, int lo_bit, int hi_bit, hipStream_t s)
{
    return radix_sort_u64_2(P, keys, n, lo_bit, hi_bit, 0, 0, s);
}

hipError_t radix_sort_u64_2(Prims &P, uint64_t *keys, int64_t n, int lo_bit, int hi_bit, int lo2, int hi2,
                            hipStream_t s)
{
    if (n <= 1 || (hi_bit <= lo_bit && hi2 <= lo2)) return hipSuccess;
    if (n > (int64_t)0xFFFFFFFF) return hipErrorInvalidValue; // 32-bit histograms and digit offsets
    if (hi_bit <= lo_bit) {
        lo_bit = lo2;
        hi_bit = hi2;
        lo2 = hi2 = 0;
    }
    if (n <= SMALL_SORT) {
        HIP_CHECK_RET(P.sort_tmp_keys.ensure(sizeof(uint64_t) * (size_t)n));
        hipLaunchKernelGGL(k_sort_small, dim3(1), dim3(1024), 0, s, keys, P.sort_tmp_keys.as<uint64_t>(), (int)n, lo_bit,
                           hi_bit, lo2, hi2);
        return hipGetLastError();
    }
    std::vector<int> shifts;
    for (int sh = lo_bit; sh < hi_bit; sh += 8) shifts.push_back(sh);
    for (int sh = lo2; sh < hi2; sh += 8) shifts.push_back(sh);
    const int64_t nb64 = (n + RS_TILE - 1) / RS_TILE;
    if (nb64 > (1 << 30)) return hipErrorInvalidValue;
    const int nb = (int)nb64;
    HIP_CHECK_RET(P.sort_tmp_keys.ensure(sizeof(uint64_t) * (size_t)n));
    HIP_CHECK_RET(P.sort_hist.ensure(sizeof(uint32_t) * (size_t)256 * nb));
    uint64_t *src = keys, *dst = P.sort_tmp_keys.as<uint64_t>();
    uint32_t *hist = P.sort_hist.as<uint32_t>();
    int passes = 0;
    for (const int shift : shifts) {
        hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(RS_THREADS), 0, s, src, n, shift, hist, nb);
        HIP_CHECK_RET(hipGetLastError());
        if (nb <= RS_SELF_MAX_TILES) {
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_rs_scatter<true>), dim3(nb), dim3(RS_THREADS), 0, s, src, dst, n, shift,
                               hist, nb);
        } else {
            HIP_CHECK_RET(scan_excl_u32(P, hist, hist, (int64_t)256 * nb, s));
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_rs_scatter<false>), dim3(nb), dim3(RS_THREADS), 0, s, src, dst, n, shift,
                               hist, nb);
        }
        HIP_CHECK_RET(hipGetLastError());
        uint64_t *t = src;
        src = dst;
        dst = t;
        passes++;
    }
    if (src != keys) HIP_CHECK_RET(hipMemcpyAsync(keys, src, sizeof(uint64_t) * (size_t)n, hipMemcpyDeviceToDevice, s));
    return hipSuccess;
}

// ------------------------------------------------------------------------------------
// Sort of keys a << ash | b (a < 2^na, b < 2^nbb, the bits between zero) by buckets of
// consecutive a: directed pairs a << 32 | b, and the drowned-insertion keys.  The LSD sort
// above moves every key once per 8-bit digit (six times for 22-bit ids, with a histogram read
// before each).  Here a key moves twice: a scatter into the bucket of its 2^sb consecutive a
// values (bucket offsets from a count pass over the same buckets; a work-group's place inside
// each bucket it touches is reserved with one atomic), then one work-group per bucket sorts it
// in LDS by its remaining sb + nbb <= 32 bits and writes it back in order.  The
// bucket width is chosen so that buckets average a few thousand keys; a bucket beyond the LDS
// capacity (stars with very long lists) is sorted by its work-group through global memory
// instead.  The order inside a bucket after the scatter depends on the atomics; the final
// order does not (keys end sorted by their whole value).
// ------------------------------------------------------------------------------------
static constexpr int BK_THREADS = 1024;
static constexpr int BK_MAX_BUCKETS = 16384;
static constexpr int BK_TILE = 16384;  // keys per work-group of the count and scatter passes
static constexpr int BK_TARGET = 4096;        // average bucket aimed for
static constexpr int64_t BK_MIN_KEYS = SMALL_SORT + 1; // shorter lists: one work-group

__device__ __forceinline__ uint32_t bk_bucket(uint64_t key, int sh, int nbkt)
{
    return min((uint32_t)(key >> sh), (uint32_t)(nbkt - 1)); // keys outside the contract stay in bounds
}

// The count and scatter passes take tiles of `tile` keys (a multiple of BK_THREADS up to
// BK_TILE), shorter for short lists so that the launch still fills the chip.
__global__ void __launch_bounds__(BK_THREADS) k_bk_count(const uint64_t *keys, int64_t m, int sh, int nbkt, int tile,
                                                         uint32_t *count)
{
    __shared__ uint32_t h[BK_MAX_BUCKETS];
    for (int i = threadIdx.x; i < nbkt; i += BK_THREADS) h[i] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * tile;
    const int R = tile / BK_THREADS;
#pragma unroll 4
    for (int r = 0; r < R; r++) {
        const int64_t k = base + (int64_t)r * BK_THREADS + threadIdx.x;
        if (k < m) atomicAdd(&h[bk_bucket(keys[k], sh, nbkt)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nbkt; i += BK_THREADS)
        if (h[i]) atomicAdd(&count[i], h[i]);
}

__global__ void __launch_bounds__(BK_THREADS) k_bk_scatter(const uint64_t *keys, int64_t m, int sh, int nbkt,
                                                           int tile, uint32_t *cursor, uint64_t *out)
{
    constexpr int R = BK_TILE / BK_THREADS;
    __shared__ uint32_t h[BK_MAX_BUCKETS];
    for (int i = threadIdx.x; i < nbkt; i += BK_THREADS) h[i] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * tile;
    const int64_t end = min(m, base + tile);
    uint64_t kv[R];
    uint32_t rk[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int64_t k = base + (int64_t)r * BK_THREADS + threadIdx.x;
        kv[r] = k < end ? keys[k] : 0ull;
        rk[r] = k < end ? atomicAdd(&h[bk_bucket(kv[r], sh, nbkt)], 1u) : 0u;
    }
    __syncthreads();
    // this work-group's place in each bucket it touches
    for (int i = threadIdx.x; i < nbkt; i += BK_THREADS)
        if (h[i]) h[i] = atomicAdd(&cursor[i], h[i]);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int64_t k = base + (int64_t)r * BK_THREADS + threadIdx.x;
        if (k < end) out[h[bk_bucket(kv[r], sh, nbkt)] + rk[r]] = kv[r];
    }
}

// Bits of a key below its bucket: (a mod 2^sb) << nbb | b.
__device__ __forceinline__ uint32_t bk_local(uint64_t key, int ash, uint32_t smask, int nbb, uint32_t bmask)
{
    return (((uint32_t)(key >> ash) & smask) << nbb) | ((uint32_t)key & bmask);
}

// Per-bucket sort: a work-group of T threads (T / 64 waves) holds up to T * BKS_ITEMS keys in
// registers, BKS_ITEMS per lane.  256-thread groups (8,192 keys) for buckets of a few thousand
// keys, 512-thread groups (16,384) for larger ones: several buckets share a CU, so one bucket's
// loads and stores overlap another's digit passes.
static constexpr int BKS_ITEMS = 32;

template <int T>
struct BkLds {
    static constexpr int W = T / 64;
    uint16_t wcnt[W][256]; // per wave and digit: count, then the prefix over earlier waves
    uint32_t dbase[256];   // keys of smaller digits
    uint32_t dcnt[256];    // keys of the digit
    uint32_t part[4];
};

// Stable ranking of one 8-bit digit of up to T * BKS_ITEMS keys held (wave, item, lane) in key
// order, nvw of them valid in this wave (a prefix): rank[r] = the key's place among the keys of
// its digit (ballot-matched peers within the wave, then per-wave counts); L.dbase and L.dcnt
// describe the digits.  Every thread of the work-group must call it.
template <int T>
__device__ __forceinline__ void bk_rank_digit(const uint32_t (&kv)[BKS_ITEMS], int shift, int nvw,
                                              uint32_t (&rank)[BKS_ITEMS], BkLds<T> &L)
{
    constexpr int W = T / 64;
    const int lane = __lane_id(), wid = threadIdx.x >> 6;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (int i = threadIdx.x; i < W * 256; i += T) (&L.wcnt[0][0])[i] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < BKS_ITEMS; r++) {
        if (r * 64 >= nvw) break; // wave-uniform
        const bool valid = r * 64 + lane < nvw;
        const int d = (int)((kv[r] >> shift) & 255u);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t before = valid ? L.wcnt[wid][d] : 0u;
        rank[r] = before + (uint32_t)__popcll(peers & lt);
        __builtin_amdgcn_wave_barrier();
        if (valid && (peers & lt) == 0) L.wcnt[wid][d] = (uint16_t)(before + (uint32_t)__popcll(peers));
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // threads 0..255: one digit each (T >= 256)
    uint32_t run = 0, inc = 0;
    if (threadIdx.x < 256) {
        const int d = threadIdx.x;
#pragma unroll
        for (int w = 0; w < W; w++) {
            const uint32_t t = L.wcnt[w][d];
            L.wcnt[w][d] = (uint16_t)run;
            run += t;
        }
        L.dcnt[d] = run;
        inc = run;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(inc, o, 64);
            if (lane >= o) inc += t;
        }
        if (lane == 63) L.part[wid] = inc;
    }
    __syncthreads();
    if (threadIdx.x < 256) {
        uint32_t wb = 0;
        for (int w = 0; w < wid; w++) wb += L.part[w];
        L.dbase[threadIdx.x] = wb + inc - run;
    }
#pragma unroll
    for (int r = 0; r < BKS_ITEMS; r++) {
        if (r * 64 >= nvw) break;
        if (r * 64 + lane < nvw) rank[r] += L.wcnt[wid][(kv[r] >> shift) & 255u];
    }
    __syncthreads();
}

// A tile of n <= T * BKS_ITEMS keys is laid out (wave, item, lane) with ipw = ceil(n / T) items
// per lane, so that every wave has work; the valid keys of a wave are a prefix of its ipw * 64
// places.
__device__ __forceinline__ int bk_wave_valid(uint32_t n, int ipw, int wid)
{
    const uint32_t lo = (uint32_t)(wid * ipw * 64);
    return n > lo ? (int)min((uint32_t)(ipw * 64), n - lo) : 0;
}

// Bucket blockIdx.x holds src[off[b] .. off[b+1]); a bucket of at most T * BKS_ITEMS keys is
// sorted into dst at the same places, or back into src when to_src.  Longer buckets are left to
// k_bk_sort_big (a separate kernel, so that its registers do not cost this one occupancy).
template <int T>
__global__ void __launch_bounds__(T) k_bk_sort(uint64_t *src, uint64_t *dst, const uint32_t *off, int ash, int sb,
                                               int nbb, int to_src)
{
    constexpr int CAP = T * BKS_ITEMS;
    __shared__ uint32_t buf[CAP];
    __shared__ BkLds<T> L;
    const uint32_t o = off[blockIdx.x], sz = off[blockIdx.x + 1] - o;
    if (sz == 0 || sz > (uint32_t)CAP) return; // work-group-uniform
    const int lane = __lane_id(), wid = threadIdx.x >> 6;
    const uint32_t smask = (1u << sb) - 1u, bmask = (1u << nbb) - 1u;
    const int bits = sb + nbb;
    const uint64_t hi = (uint64_t)blockIdx.x << sb; // the bucket's a values above the local bits
    uint64_t *res = (to_src ? src : dst) + o;
    if (sz <= (uint32_t)T) {
        // A bucket of at most one key per thread (short lists spread over many buckets): each key's
        // place is the count of smaller keys plus that of equal keys before it, read from LDS.
        const uint32_t i = threadIdx.x;
        uint32_t key = 0u;
        if (i < sz) {
            key = bk_local(src[o + i], ash, smask, nbb, bmask);
            buf[i] = key;
        }
        __syncthreads();
        if (i < sz) {
            uint32_t rk = 0;
            for (uint32_t j = 0; j < sz; j++) {
                const uint32_t kj = buf[j];
                rk += (kj < key || (kj == key && j < i)) ? 1u : 0u;
            }
            res[rk] = ((hi | (key >> nbb)) << ash) | (key & bmask);
        }
        return;
    }
    uint32_t kv[BKS_ITEMS], rank[BKS_ITEMS];
    const int ipw = (int)((sz + T - 1) / T);
    const int nvw = bk_wave_valid(sz, ipw, wid);
    const int tbase = wid * ipw * 64 + lane;
#pragma unroll
    for (int r = 0; r < BKS_ITEMS; r++)
        kv[r] = r * 64 + lane < nvw ? bk_local(src[o + tbase + 64 * r], ash, smask, nbb, bmask) : 0u;
    for (int shift = 0; shift < bits; shift += 8) {
        bk_rank_digit<T>(kv, shift, nvw, rank, L);
#pragma unroll
        for (int r = 0; r < BKS_ITEMS; r++)
            if (r * 64 + lane < nvw) buf[L.dbase[(kv[r] >> shift) & 255u] + rank[r]] = kv[r];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < BKS_ITEMS; r++)
            if (r * 64 + lane < nvw) kv[r] = buf[tbase + 64 * r];
        __syncthreads();
    }
    // every key of the bucket is in registers by now
#pragma unroll
    for (int r = 0; r < BKS_ITEMS; r++)
        if (r * 64 + lane < nvw) res[tbase + 64 * r] = ((hi | (kv[r] >> nbb)) << ash) | (kv[r] & bmask);
}

// Buckets longer than cap keys: digit passes through global memory, alternating between the
// bucket's places in src and dst, each a bucket-wide digit histogram and then tiles of T *
// BKS_ITEMS keys in order, each ranked stably and written after the tiles before it.  The
// result ends where k_bk_sort puts it.  Work-groups of other buckets leave at once.
template <int T>
__global__ void __launch_bounds__(T) k_bk_sort_big(uint64_t *src, uint64_t *dst, const uint32_t *off, int ash, int sb,
                                                   int nbb, int to_src, uint32_t cap)
{
    constexpr int CAP = T * BKS_ITEMS;
    __shared__ uint32_t goff[256]; // running digit offsets
    __shared__ BkLds<T> L;
    const uint32_t o = off[blockIdx.x], sz = off[blockIdx.x + 1] - o;
    if (sz <= cap) return; // work-group-uniform
    const int lane = __lane_id(), wid = threadIdx.x >> 6;
    const uint32_t smask = (1u << sb) - 1u, bmask = (1u << nbb) - 1u;
    const int bits = sb + nbb;
    uint64_t *res = (to_src ? src : dst) + o;
    uint64_t *a = src + o, *q = dst + o;
    uint32_t kv[BKS_ITEMS], rank[BKS_ITEMS];
    const int tbase = wid * BKS_ITEMS * 64 + lane;
    for (int shift = 0; shift < bits; shift += 8) {
        if (threadIdx.x < 256) goff[threadIdx.x] = 0;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < sz; i += T)
            atomicAdd(&goff[(bk_local(a[i], ash, smask, nbb, bmask) >> shift) & 255u], 1u);
        __syncthreads();
        uint32_t c = 0, inc = 0;
        if (threadIdx.x < 256) {
            c = goff[threadIdx.x];
            inc = c;
#pragma unroll
            for (int t = 1; t < 64; t <<= 1) {
                const uint32_t u = __shfl_up(inc, t, 64);
                if (lane >= t) inc += u;
            }
            if (lane == 63) L.part[wid] = inc;
        }
        __syncthreads();
        if (threadIdx.x < 256) {
            uint32_t wb = 0;
            for (int w = 0; w < wid; w++) wb += L.part[w];
            goff[threadIdx.x] = wb + inc - c;
        }
        __syncthreads();
        for (uint32_t t0 = 0; t0 < sz; t0 += CAP) {
            const int nvw = bk_wave_valid(min((uint32_t)CAP, sz - t0), BKS_ITEMS, wid);
#pragma unroll
            for (int r = 0; r < BKS_ITEMS; r++)
                kv[r] = r * 64 + lane < nvw ? bk_local(a[t0 + tbase + 64 * r], ash, smask, nbb, bmask) : 0u;
            bk_rank_digit<T>(kv, shift, nvw, rank, L);
            // the key is rebuilt from its local bits and the bucket (as k_bk_sort writes it)
            const uint64_t hi = (uint64_t)blockIdx.x << sb;
#pragma unroll
            for (int r = 0; r < BKS_ITEMS; r++)
                if (r * 64 + lane < nvw)
                    q[goff[(kv[r] >> shift) & 255u] + rank[r]] = ((hi | (kv[r] >> nbb)) << ash) | (kv[r] & bmask);
            __syncthreads();
            if (threadIdx.x < 256) goff[threadIdx.x] += L.dcnt[threadIdx.x];
            __syncthreads();
        }
        uint64_t *t = a;
        a = q;
        q = t;
    }
    if (a != res)
        for (uint32_t i = threadIdx.x; i < sz; i += T) res[i] = a[i];
}

// Bucket shift for m keys over nids values of a (0: unknown, taken as 2^(na-1)), or -1 when the
// bucket sort does not apply: lists below BK_MIN_KEYS (unless forced), or 2^14 buckets that
// would not leave 32-bit local keys.  The widest buckets that average at most BK_TARGET keys,
// else the narrowest allowed.
int key_sort_shift(int64_t m, int na, int nbb, int64_t nids, bool force)
{
    if (m <= SMALL_SORT || m > (int64_t)0xFFFFFFFF || na < 1 || na > 31 || nbb < 1 || nbb > 32) return -1;
    if (!force && m < BK_MIN_KEYS) return -1;
    const int lo = std::max(0, na - 14), hi = std::min(na, 32 - nbb);
    if (lo > hi) return -1;
    const int64_t ids = nids > 0 ? nids : (int64_t)1 << (na - 1);
    int sb = lo;
    for (int t = hi; t > lo; t--) {
        const int64_t nb = std::max<int64_t>(1, (ids + ((int64_t)1 << t) - 1) >> t);
        if (m <= (int64_t)BK_TARGET * nb) {
            sb = t;
            break;
        }
    }
    return sb;
}

int pair_sort_shift(int64_t m, int nb, int64_t nids) { return key_sort_shift(m, nb, nb, nids, false); }

// Bucket sort proper; the result ends in P.sort_tmp_keys when to_tmp, else in keys.
static hipError_t bucket_sort(Prims &P, uint64_t *keys, int64_t m, int ash, int na, int nbb, int sb, int64_t nids,
                              bool to_tmp, hipStream_t s)
{
    const int nbkt = 1 << (na - sb);
    HIP_CHECK_RET(P.sort_tmp_keys.ensure(sizeof(uint64_t) * (size_t)m));
    HIP_CHECK_RET(P.sort_hist.ensure(sizeof(uint32_t) * (size_t)(2 * nbkt + 2)));
    uint64_t *tmp = P.sort_tmp_keys.as<uint64_t>();
    uint32_t *count = P.sort_hist.as<uint32_t>(), *off = count + nbkt + 1;
    HIP_CHECK_RET(hipMemsetAsync(count, 0, sizeof(uint32_t) * (size_t)(nbkt + 1), s));
    // tiles of up to BK_TILE keys, at least ~512 of them
    const int64_t want = (m + 511) / 512;
    const int tile = (int)std::min<int64_t>(
        BK_TILE, std::max<int64_t>(BK_THREADS, (want + BK_THREADS - 1) / BK_THREADS * BK_THREADS));
    const int64_t nblk = (m + tile - 1) / tile;
    const int sh = ash + sb;
    hipLaunchKernelGGL(k_bk_count, dim3((unsigned)nblk), dim3(BK_THREADS), 0, s, keys, m, sh, nbkt, tile, count);
    HIP_CHECK_RET(hipGetLastError());
    HIP_CHECK_RET(scan_excl_u32(P, count, off, nbkt + 1, s)); // off[nbkt] = m
    HIP_CHECK_RET(hipMemcpyAsync(count, off, sizeof(uint32_t) * (size_t)nbkt, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_bk_scatter, dim3((unsigned)nblk), dim3(BK_THREADS), 0, s, keys, m, sh, nbkt, tile, count,
                       tmp);
    HIP_CHECK_RET(hipGetLastError());
    // the smaller work-group when the average bucket is at most half its capacity
    const int64_t ids = nids > 0 ? nids : (int64_t)1 << (na - 1);
    const int64_t used = std::max<int64_t>(1, (ids + ((int64_t)1 << sb) - 1) >> sb);
    uint32_t cap;
    if (m <= (int64_t)128 * BKS_ITEMS * used) {
        hipLaunchKernelGGL(k_bk_sort<256>, dim3((unsigned)nbkt), dim3(256), 0, s, tmp, keys, off, ash, sb, nbb,
                           to_tmp ? 1 : 0);
        cap = 256 * BKS_ITEMS;
    } else {
        hipLaunchKernelGGL(k_bk_sort<512>, dim3((unsigned)nbkt), dim3(512), 0, s, tmp, keys, off, ash, sb, nbb,
                           to_tmp ? 1 : 0);
        cap = 512 * BKS_ITEMS;
    }
    HIP_CHECK_RET(hipGetLastError());
    // buckets beyond LDS, if any (the others' work-groups leave at once)
    hipLaunchKernelGGL(k_bk_sort_big<512>, dim3((unsigned)nbkt), dim3(512), 0, s, tmp, keys, off, ash, sb, nbb,
                       to_tmp ? 1 : 0, cap);
    return hipGetLastError();
}

hipError_t sort_keys_ab(Prims &P, uint64_t *keys, int64_t m, int ash, int na, int nbb, int64_t nids, bool force,
                        hipStream_t s)
{
    if (m <= 1) return hipSuccess;
    const int sb = key_sort_shift(m, na, nbb, nids, force);
    if (sb < 0 || ash < nbb || ash + na > 64) return radix_sort_u64_2(P, keys, m, 0, nbb, ash, ash + na, s);
    return bucket_sort(P, keys, m, ash, na, nbb, sb, nids, false, s);
}

hipError_t sort_pair_keys(Prims &P, uint64_t *keys, int64_t m, int nb, int64_t nids, hipStream_t s)
{
    return sort_keys_ab(P, keys, m, 32, nb, nb, nids, false, s);
}

hipError_t sort_unique_pair_keys(Prims &P, uint64_t *keys, int64_t m, int nb, int64_t nids, int64_t *m_out,
                                 hipStream_t s)
{
    const int sb = key_sort_shift(m, nb, nb, nids, false);
    if (m <= 1 || sb < 0) {
        HIP_CHECK_RET(radix_sort_u64_2(P, keys, m, 0, nb, 32, 32 + nb, s));
        return unique_u64(P, keys, m, m_out, s);
    }
    HIP_CHECK_RET(bucket_sort(P, keys, m, 32, nb, nb, sb, nids, true, s));
    return unique_copy_u64(P, P.sort_tmp_keys.as<uint64_t>(), keys, m, m_out, s);
}

// ------------------------------------------------------------------------------------
// Merge of two sorted lists, for a long list `a` and a short one `b`.  Work-group w writes
// outputs [w*MG_TILE, (w+1)*MG_TILE): its first thread finds where the two tile boundaries cross
// the inputs (merge path, ties take `a` first), the tile's part of `b` is staged in LDS, and
// every element goes straight to its output: an `a` element after the staged `b` elements
// smaller than it, a `b` element after the tile's `a` elements not greater than it.
// ------------------------------------------------------------------------------------

static constexpr int MG_TILE = 4096;

__device__ __forceinline__ int64_t merge_split(const uint64_t *a, int64_t na, const uint64_t *b, int64_t nb, int64_t d)
{
    int64_t lo = d > nb ? d - nb : 0, hi = d < na ? d : na;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (a[mid] <= b[d - mid - 1]) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(256) k_merge_u64(const uint64_t *a, int64_t na, const uint64_t *b, int64_t nb,
                                                   uint64_t *out)
{
    __shared__ uint64_t bw[MG_TILE];
    __shared__ int64_t sp[4];
    const int64_t total = na + nb;
    const int64_t o0 = (int64_t)blockIdx.x * MG_TILE;
    const int64_t o1 = o0 + MG_TILE < total ? o0 + MG_TILE : total;
    if (threadIdx.x == 0) {
        const int64_t i0 = merge_split(a, na, b, nb, o0), i1 = merge_split(a, na, b, nb, o1);
        sp[0] = i0;
        sp[1] = o0 - i0;
        sp[2] = i1;
        sp[3] = o1 - i1;
    }
    __syncthreads();
    const int64_t i0 = sp[0], j0 = sp[1], i1 = sp[2], j1 = sp[3];
    const int nbw = (int)(j1 - j0);
    for (int k = threadIdx.x; k < nbw; k += blockDim.x) bw[k] = b[j0 + k];
    __syncthreads();
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        const uint64_t v = a[i];
        int lo = 0, hi = nbw; // staged b elements < v
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (bw[mid] < v) lo = mid + 1;
            else hi = mid;
        }
        out[o0 + (i - i0) + lo] = v;
    }
    for (int k = threadIdx.x; k < nbw; k += blockDim.x) {
        const uint64_t v = bw[k];
        int64_t lo = i0, hi = i1; // tile's a elements <= v
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (a[mid] <= v) lo = mid + 1;
            else hi = mid;
        }
        out[o0 + k + (lo - i0)] = v;
    }
}

hipError_t merge_u64(const uint64_t *a, int64_t na, const uint64_t *b, int64_t nb, uint64_t *out, hipStream_t s)
{
    const int64_t nt = (na + nb + MG_TILE - 1) / MG_TILE;
    if (nt <= 0) return hipSuccess;
    if (nt > INT32_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_merge_u64, dim3((unsigned)nt), dim3(256), 0, s, a, na, b, nb, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------
// Compaction
// ------------------------------------------------------------------------------------

// Lists up to SMALL_COMPACT entries are compacted by one work-group in a single launch
// (tiles of 1024 in order, ballot ranks, a running offset), which also works in place: an
// entry kept from index j lands at a position <= j, and later tiles read beyond every
// position written so far.  MODE 0: adjacent-unique, 1: keep[i] != 0, 2: in[i] != none.
static constexpr int64_t SMALL_COMPACT = 32768;

template <int MODE>
__global__ void __launch_bounds__(1024) k_compact_small(const uint64_t *in, const uint8_t *keep, uint64_t none,
                                                        uint64_t *out, int64_t n, uint32_t *count)
{
    __shared__ uint32_t wsum[16];
    __shared__ uint64_t last;
    const int lane = __lane_id(), wid = threadIdx.x >> 6;
    uint32_t run = 0;
    for (int64_t t0 = 0; t0 < n; t0 += 1024) {
        const int64_t i = t0 + threadIdx.x;
        uint64_t v = 0;
        bool f = false;
        if (i < n) {
            v = in[i];
            if (MODE == 0) f = i == 0 || v != (threadIdx.x == 0 ? last : in[i - 1]);
            else if (MODE == 1) f = keep[i] != 0;
            else f = v != none;
        }
        const uint64_t m = __ballot(f);
        if (lane == 0) wsum[wid] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < 16; w++) {
            const uint32_t c = wsum[w];
            if (w < wid) before += c;
            tot += c;
        }
        if (f) out[run + before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = v;
        if (MODE == 0 && threadIdx.x == 1023) last = v;
        run += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        count[0] = run;
        count[1] = 0; // the two words read as an int64 count too
    }
}

template <int MODE>
static hipError_t compact_small(Prims &P, const uint64_t *in, const uint8_t *keep, uint64_t none, uint64_t *out,
                                int64_t n, int64_t *n_out, hipStream_t s)
{
    HIP_CHECK_RET(P.pos.ensure(sizeof(uint32_t) * 4));
    uint32_t *dcount = P.pos.as<uint32_t>();
    hipLaunchKernelGGL(k_compact_small<MODE>, dim3(1), dim3(1024), 0, s, in, keep, none, out, n, dcount);
    HIP_CHECK_RET(hipGetLastError());
    uint32_t c = 0;
    uint32_t *hc = P.hstage ? reinterpret_cast<uint32_t *>(P.hstage) : &c;
    HIP_CHECK_RET(hipMemcpyAsync(hc, dcount, 4, hipMemcpyDeviceToHost, s));
    HIP_CHECK_RET(hipStreamSynchronize(s));
    *n_out = *hc;
    return hipSuccess;
}

// Longer lists: a counting pass per tile of CMP_TILE entries, a scan of the tile counts, and
// a writing pass that ranks each round of 256 entries with ballots, so no flag or position
// array goes through HBM.  Out of place only (tiles write concurrently).
static constexpr int CMP_ROUNDS = 16;
static constexpr int CMP_TILE = 256 * CMP_ROUNDS;

template <int MODE>
__device__ __forceinline__ bool cmp_keep(const uint64_t *in, const uint8_t *keep, uint64_t none, int64_t i,
                                         uint64_t v)
{
    if (MODE == 0) return i == 0 || v != in[i - 1];
    if (MODE == 1) return keep[i] != 0;
    return v != none;
}

template <int MODE>
__global__ void __launch_bounds__(256) k_cmp_count(const uint64_t *in, const uint8_t *keep, uint64_t none, int64_t n,
                                                   int64_t *tile_cnt)
{
    __shared__ uint32_t wsum[4];
    const int64_t t0 = (int64_t)blockIdx.x * CMP_TILE;
    uint32_t c = 0;
#pragma unroll 4
    for (int r = 0; r < CMP_ROUNDS; r++) {
        const int64_t i = t0 + r * 256 + threadIdx.x;
        if (i < n) c += cmp_keep<MODE>(in, keep, none, i, in[i]) ? 1u : 0u;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) tile_cnt[blockIdx.x] = (int64_t)wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

template <int MODE>
__global__ void __launch_bounds__(256) k_cmp_write(const uint64_t *in, const uint8_t *keep, uint64_t none, int64_t n,
                                                   const int64_t *tile_off, uint64_t *out)
{
    __shared__ uint32_t wsum[4];
    const int lane = __lane_id(), wid = threadIdx.x >> 6;
    const int64_t t0 = (int64_t)blockIdx.x * CMP_TILE;
    int64_t run = tile_off[blockIdx.x];
    for (int r = 0; r < CMP_ROUNDS; r++) {
        const int64_t i = t0 + r * 256 + threadIdx.x;
        uint64_t v = 0;
        bool f = false;
        if (i < n) {
            v = in[i];
            f = cmp_keep<MODE>(in, keep, none, i, v);
        }
        const uint64_t m = __ballot(f);
        if (lane == 0) wsum[wid] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < 4; w++) {
            const uint32_t x = wsum[w];
            if (w < wid) before += x;
            tot += x;
        }
        if (f) out[run + before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = v;
        run += tot;
        __syncthreads();
    }
}

template <int MODE>
static hipError_t compact_big(Prims &P, const uint64_t *in, const uint8_t *keep, uint64_t none, uint64_t *out,
                              int64_t n, int64_t *n_out, hipStream_t s)
{
    const int64_t nt = (n + CMP_TILE - 1) / CMP_TILE;
    if (nt > INT32_MAX) return hipErrorInvalidValue;
    HIP_CHECK_RET(P.flags.ensure(sizeof(int64_t) * (size_t)(nt + 1)));
    int64_t *tc = P.flags.as<int64_t>();
    hipLaunchKernelGGL(k_cmp_count<MODE>, dim3((unsigned)nt), dim3(256), 0, s, in, keep, none, n, tc);
    HIP_CHECK_RET(hipGetLastError());
    HIP_CHECK_RET(scan_excl_i64(P, tc, tc, nt, tc + nt, s));
    uint64_t *dst = out;
    if (in == out) {
        if (in == P.sort_tmp_keys.ptr) return hipErrorInvalidValue; // would read and write one buffer
        HIP_CHECK_RET(P.sort_tmp_keys.ensure(sizeof(uint64_t) * (size_t)n));
        dst = P.sort_tmp_keys.as<uint64_t>();
    }
    hipLaunchKernelGGL(k_cmp_write<MODE>, dim3((unsigned)nt), dim3(256), 0, s, in, keep, none, n, tc, dst);
    HIP_CHECK_RET(hipGetLastError());
    int64_t m = 0;
    int64_t *hm = P.hstage ? P.hstage : &m;
    HIP_CHECK_RET(hipMemcpyAsync(hm, tc + nt, 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK_RET(hipStreamSynchronize(s));
    m = *hm;
    if (dst != out && m > 0) HIP_CHECK_RET(hipMemcpyAsync(out, dst, sizeof(uint64_t) * (size_t)m, hipMemcpyDeviceToDevice, s));
    *n_out = m;
    return hipSuccess;
}

// The same kernels as compact_small / compact_big with in != out, the count left for the caller:
// on the device at *dev_count (int64; valid until the next compaction) and, once the stream
// reaches it, at *host_count (page-locked memory, or the copy blocks).
template <int MODE>
static hipError_t compact_deferred(Prims &P, const uint64_t *in, const uint8_t *keep, uint64_t none, uint64_t *out,
                                   int64_t n, int64_t *host_count, const int64_t **dev_count, hipStream_t s)
{
    *host_count = 0;
    if (n <= 0 || in == out) return hipErrorInvalidValue;
    if (n <= SMALL_COMPACT) {
        HIP_CHECK_RET(P.pos.ensure(sizeof(uint32_t) * 4));
        uint32_t *dcount = P.pos.as<uint32_t>();
        hipLaunchKernelGGL(k_compact_small<MODE>, dim3(1), dim3(1024), 0, s, in, keep, none, out, n, dcount);
        HIP_CHECK_RET(hipGetLastError());
        if (dev_count) *dev_count = reinterpret_cast<const int64_t *>(dcount);
        return hipMemcpyAsync(host_count, dcount, 8, hipMemcpyDeviceToHost, s);
    }
    const int64_t nt = (n + CMP_TILE - 1) / CMP_TILE;
    if (nt > INT32_MAX) return hipErrorInvalidValue;
    HIP_CHECK_RET(P.flags.ensure(sizeof(int64_t) * (size_t)(nt + 1)));
    int64_t *tc = P.flags.as<int64_t>();
    hipLaunchKernelGGL(k_cmp_count<MODE>, dim3((unsigned)nt), dim3(256), 0, s, in, keep, none, n, tc);
    HIP_CHECK_RET(hipGetLastError());
    HIP_CHECK_RET(scan_excl_i64(P, tc, tc, nt, tc + nt, s));
    hipLaunchKernelGGL(k_cmp_write<MODE>, dim3((unsigned)nt), dim3(256), 0, s, in, keep, none, n, tc, out);
    HIP_CHECK_RET(hipGetLastError());
    if (dev_count) *dev_count = tc + nt;
    return hipMemcpyAsync(host_count, tc + nt, 8, hipMemcpyDeviceToHost, s);
}

hipError_t compact_u64_deferred(Prims &P, const uint64_t *in, const uint8_t *keep, uint64_t *out, int64_t n,
                                int64_t *host_count, hipStream_t s)
{
    if (n <= 0) {
        *host_count = 0;
        return hipSuccess;
    }
    return compact_deferred<1>(P, in, keep, 0, out, n, host_count, nullptr, s);
}

hipError_t compact_valid_u64_deferred(Prims &P, const uint64_t *in, uint64_t *out, int64_t n, uint64_t none,
                                      int64_t *host_count, const int64_t **dev_count, hipStream_t s)
{
    return compact_deferred<2>(P, in, nullptr, none, out, n, host_count, dev_count, s);
}

hipError_t unique_u64(Prims &P, uint64_t *keys, int64_t n, int64_t *n_out, hipStream_t s)
{
    if (n <= 1) {
        *n_out = n;
        return hipSuccess;
    }
    if (n <= SMALL_COMPACT) return compact_small<0>(P, keys, nullptr, 0, keys, n, n_out, s);
    return compact_big<0>(P, keys, nullptr, 0, keys, n, n_out, s);
}

hipError_t unique_copy_u64(Prims &P, const uint64_t *in, uint64_t *out, int64_t n, int64_t *n_out, hipStream_t s)
{
    if (n <= 0) {
        *n_out = 0;
        return hipSuccess;
    }
    if (in == out) return hipErrorInvalidValue;
    if (n <= SMALL_COMPACT) return compact_small<0>(P, in, nullptr, 0, out, n, n_out, s);
    return compact_big<0>(P, in, nullptr, 0, out, n, n_out, s);
}

hipError_t compact_u64(Prims &P, const uint64_t *in, const uint8_t *keep, uint64_t *out, int64_t n, int64_t *n_out,
                       hipStream_t s)
{
    if (n <= 0) {
        *n_out = 0;
        return hipSuccess;
    }
    if (n <= SMALL_COMPACT) return compact_small<1>(P, in, keep, 0, out, n, n_out, s);
    return compact_big<1>(P, in, keep, 0, out, n, n_out, s);
}

hipError_t compact_valid_u64(Prims &P, const uint64_t *in, uint64_t *out, int64_t n, uint64_t none, int64_t *n_out,
                             hipStream_t s)
{
    if (n <= 0) {
        *n_out = 0;
        return hipSuccess;
    }
    if (n <= SMALL_COMPACT) return compact_small<2>(P, in, nullptr, none, out, n, n_out, s);
    return compact_big<2>(P, in, nullptr, none, out, n, n_out, s);
}

} // namespace g4d
