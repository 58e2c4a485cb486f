// gStar4D for MI355X — device primitives: scans, stable LSD radix sort, compaction.
#pragma once

#include <vector>

#include "g4d_common.h"

namespace g4d {

// Bump allocator over a caller-supplied device region.  While one is installed on the calling
// thread (ScratchScope), DevBuf::ensure carves from it instead of calling hipMalloc and
// DevBuf::release forgets the pointer instead of calling hipFree, so a whole pipeline run can
// take every buffer from one preallocated region.
struct ScratchArena {
    char *base = nullptr;
    size_t cap = 0;
    size_t top = 0;
    bool exhausted = false;
};
ScratchArena *&scratch_arena_slot();
struct ScratchScope {
    ScratchArena *prev;
    explicit ScratchScope(ScratchArena *a) : prev(scratch_arena_slot()) { scratch_arena_slot() = a; }
    ~ScratchScope() { scratch_arena_slot() = prev; }
};

// Growable device buffer owned by a context; never shrinks, reused across calls.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need)
    {
        if (need <= bytes) return hipSuccess;
        if (ScratchArena *a = scratch_arena_slot()) {
            // the most recent region grows in place
            if (ptr && static_cast<char *>(ptr) + bytes == a->base + a->top &&
                need <= a->cap - (size_t)(static_cast<char *>(ptr) - a->base)) {
                a->top = (size_t)(static_cast<char *>(ptr) - a->base) + need;
                bytes = need;
                return hipSuccess;
            }
            const size_t off = (a->top + 255) & ~(size_t)255;
            if (off > a->cap || need > a->cap - off) {
                a->exhausted = true;
                return hipErrorOutOfMemory;
            }
            ptr = a->base + off;
            bytes = need;
            a->top = off + need;
            return hipSuccess;
        }
        if (ptr) {
            hipError_t e = hipFree(ptr);
            if (e != hipSuccess) return e;
            ptr = nullptr;
            bytes = 0;
        }
        size_t cap = need + need / 8 + 256;
        hipError_t e = hipMalloc(&ptr, cap);
        if (e != hipSuccess) {
            ptr = nullptr;
            return e;
        }
        bytes = cap;
        return hipSuccess;
    }
    void release()
    {
        if (ptr && !scratch_arena_slot()) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <typename T>
    T *as() const
    {
        return reinterpret_cast<T *>(ptr);
    }
};

struct Prims {
    DevBuf scan_tmp[4];
    DevBuf sort_tmp_keys, sort_hist, flags, pos;
    // Page-locked host slots (two) for the counts the compactions read back, owned by the
    // caller; null: stack variables (a pageable destination costs the runtime a staged copy).
    int64_t *hstage = nullptr;
    void release()
    {
        for (auto &b : scan_tmp) b.release();
        sort_tmp_keys.release();
        sort_hist.release();
        flags.release();
        pos.release();
    }
};

// Exclusive prefix sum of n int64 values; out may alias in.  If total != nullptr the device
// total is written there.
hipError_t scan_excl_i64(Prims &P, const int64_t *in, int64_t *out, int64_t n, int64_t *total,
                         hipStream_t s);
hipError_t scan_excl_u32(Prims &P, const uint32_t *in, uint32_t *out, int64_t n, hipStream_t s);

// Stable LSD radix sort of uint64 keys on bits [lo_bit, hi_bit).  Result in keys.
hipError_t radix_sort_u64(Prims &P, uint64_t *keys, int64_t n, int lo_bit, int hi_bit, hipStream_t s);
// The same over the digits of [lo_bit, hi_bit) and then of [lo2, hi2) in one ping-pong
// sequence (an even total number of passes ends in keys without a copy).
hipError_t radix_sort_u64_2(Prims &P, uint64_t *keys, int64_t n, int lo_bit, int hi_bit, int lo2, int hi2,
                            hipStream_t s);

// Sort of directed pair keys a << 32 | b (a, b < 2^nb) by a, then b: a scatter into buckets of
// consecutive a values and one LDS sort per bucket; nids (0: unknown) is the number of values a
// takes and sizes the buckets.  Lists below 4 M keys, and ids wider than 23 bits, go through
// radix_sort_u64_2 instead.  Result in keys.
hipError_t sort_pair_keys(Prims &P, uint64_t *keys, int64_t m, int nb, int64_t nids, hipStream_t s);
// sort_pair_keys, then removal of adjacent duplicates (count via host sync); the sorted list is
// compacted from the sort's scratch buffer straight into keys.  keys must not be P.sort_tmp_keys.
hipError_t sort_unique_pair_keys(Prims &P, uint64_t *keys, int64_t m, int nb, int64_t nids, int64_t *m_out,
                                 hipStream_t s);
// Bucket shift sort_pair_keys uses for m keys (buckets of 2^shift ids), -1 for the LSD sort.
int pair_sort_shift(int64_t m, int nb, int64_t nids);
// The same sort for keys a << ash | b with a < 2^na, b < 2^nbb and zeros between (nbb <= ash):
// by a, then b.  force: the bucket sort also below its minimum length.
hipError_t sort_keys_ab(Prims &P, uint64_t *keys, int64_t m, int ash, int na, int nbb, int64_t nids, bool force,
                        hipStream_t s);
int key_sort_shift(int64_t m, int na, int nbb, int64_t nids, bool force);

// Merge of two sorted lists into out (no aliasing); equal keys keep a's before b's.
hipError_t merge_u64(const uint64_t *a, int64_t na, const uint64_t *b, int64_t nb, uint64_t *out, hipStream_t s);

// Order-preserving removal of adjacent duplicates; returns the new count via host sync.
// keys must not be P.sort_tmp_keys (longer lists go through that buffer when compacted in place).
hipError_t unique_u64(Prims &P, uint64_t *keys, int64_t n, int64_t *n_out, hipStream_t s);
// The same from in to out (distinct buffers; in may be P.sort_tmp_keys).
hipError_t unique_copy_u64(Prims &P, const uint64_t *in, uint64_t *out, int64_t n, int64_t *n_out, hipStream_t s);

// Order-preserving compaction keep[i] != 0.  Count returned via host sync.
// compact_u64 without the host synchronisation (in and out distinct): the count is copied to
// *host_count (page-locked memory, or the copy blocks) once the stream reaches it.
hipError_t compact_u64_deferred(Prims &P, const uint64_t *in, const uint8_t *keep, uint64_t *out, int64_t n,
                                int64_t *host_count, hipStream_t s);
// compact_valid_u64 from in to a distinct out without the host synchronisation (n > 0): the
// count at *dev_count on the device (int64, valid until the next compaction) and, once the
// stream reaches it, at *host_count (page-locked).
hipError_t compact_valid_u64_deferred(Prims &P, const uint64_t *in, uint64_t *out, int64_t n, uint64_t none,
                                      int64_t *host_count, const int64_t **dev_count, hipStream_t s);
hipError_t compact_u64(Prims &P, const uint64_t *in, const uint8_t *keep, uint64_t *out, int64_t n,
                       int64_t *n_out, hipStream_t s);

// Order-preserving compaction of the entries different from `none`.  Count via host sync.
hipError_t compact_valid_u64(Prims &P, const uint64_t *in, uint64_t *out, int64_t n, uint64_t none,
                             int64_t *n_out, hipStream_t s);

// Kernel launch helpers
inline int g4d_blocks(int64_t n, int per_block, int cap = 1 << 20)
{
    int64_t b = (n + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

} // namespace g4d
