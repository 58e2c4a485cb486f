// gStar4D for MI355X — shared device/host definitions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define G4D_WAVE 64

#define G4D_MARKER (-1)
// Voxel site code: x<<20 | y<<10 | z, 10 bits per axis (grid width <= 1024), as the
// reference's PBA encoding (GD/PBA/pba3D.h:61-72).
#define G4D_ENC(x, y, z) (((x) << 20) | ((y) << 10) | (z))
#define G4D_DEC_X(v) ((v) >> 20)
#define G4D_DEC_Y(v) (((v) >> 10) & 0x3ff)
#define G4D_DEC_Z(v) ((v)&0x3ff)
// Voxel of a point: 21 bits per axis, x | y << 21 | z << 42.  Grids up to G4D_CODE_GRID wide
// hold site codes in their cells; wider grids (up to G4D_MAX_GRID) hold point ids, and the site
// coordinates come from the points' voxels.
#define G4D_VOX(x, y, z) ((uint64_t)(x) | ((uint64_t)(y) << 21) | ((uint64_t)(z) << 42))
#define G4D_VOX_X(v) ((int)((v)&0x1fffff))
#define G4D_VOX_Y(v) ((int)(((v) >> 21) & 0x1fffff))
#define G4D_VOX_Z(v) ((int)(((v) >> 42) & 0x1fffff))
// A build may lower G4D_CODE_GRID (never raise it: codes have 10 bits per axis) to run the
// point-id path on small grids.
#ifndef G4D_CODE_GRID
#define G4D_CODE_GRID 1024
#endif
#define G4D_MAX_GRID 2048

// Device-side error flags (bitwise OR into a counter block word).
enum g4d_dev_error : uint32_t {
    G4D_DEV_ERR_NONE = 0,
    G4D_DEV_ERR_STAR_CAPACITY = 1u << 0,   // star exceeds the LDS star capacity
    G4D_DEV_ERR_PREDICATE_RANGE = 1u << 1, // coordinates outside the exact-arithmetic range
    G4D_DEV_ERR_HOLE = 1u << 2,            // hole boundary not a single cycle
    G4D_DEV_ERR_PROOF = 1u << 3,           // no containment proof found
    G4D_DEV_ERR_BUFFER = 1u << 4,          // an output buffer overflowed
    G4D_DEV_ERR_DUPLICATE = 1u << 5,       // duplicate input points
    G4D_DEV_ERR_CONE = 1u << 6,            // workset smaller than 4
    G4D_DEV_ERR_FACE = 1u << 7,            // a face shared by more than two tetrahedra
};

// Counter block layout (int64 words), device resident, read back once per stage.
enum g4d_counter : int {
    G4D_CNT_ERROR = 0,
    G4D_CNT_APPEND0 = 1,  // generic append counters
    G4D_CNT_APPEND1 = 2,
    G4D_CNT_APPEND2 = 3,
    G4D_CNT_ORIENT4 = 4,  // cofactor-filter orient4 evaluations in the LDS star kernels
    G4D_CNT_EXACT = 5,    // exact insphere evaluations
    G4D_CNT_SOS = 6,      // SoS resolutions
    G4D_CNT_ORIENT3 = 7,
    G4D_CNT_MISC = 8,
    G4D_CNT_WORK = 9,     // spare
    G4D_CNT_APPEND3 = 10,
    G4D_CNT_APPEND4 = 11,
    G4D_CNT_APPEND5 = 12,
    G4D_CNT_APPEND6 = 13,
    G4D_CNT_COFACTOR = 14, // cached-cofactor computations in the LDS star kernels
    G4D_CNT_XMID = 15,     // list lengths of the exact chain's 64- and 256-vertex tiers
    G4D_CNT_XBIG = 16,
    G4D_CNT_QUEUE0 = 17,   // work-queue heads of the six star-kernel launches of one pass (17..22)
    G4D_CNT_NUM = 24
};

#define HIP_CHECK_RET(expr)                                                                    \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess) return _e;                                                       \
    } while (0)

// Wave-level helpers ------------------------------------------------------------------

__device__ __forceinline__ int g4d_lane() { return __lane_id(); }

// Exclusive prefix count of `pred` over lanes below this one; total in *tot.
__device__ __forceinline__ int g4d_wave_excl_count(bool pred, int *tot)
{
    const uint64_t m = __ballot(pred);
    *tot = __popcll(m);
    return __popcll(m & ((1ull << __lane_id()) - 1ull));
}

// Wave-inclusive scan of small non-negative ints via DPP-free shuffles.
// Inclusive prefix sum over the 64 lanes: DPP row shifts (1, 2, 4, 8) within each row of 16
// lanes, then the totals of rows 0-2 (lanes 15, 31, 47) read into scalars and added to the rows
// above them.  No LDS traffic; every lane of the wave must be active.
__device__ __forceinline__ int g4d_wave_incl_sum(int v)
{
    int x = v;
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);
    const int r0 = __builtin_amdgcn_readlane(x, 15), r1 = __builtin_amdgcn_readlane(x, 31);
    const int r2 = __builtin_amdgcn_readlane(x, 47);
    const int row = __lane_id() >> 4;
    return x + (row >= 1 ? r0 : 0) + (row >= 2 ? r1 : 0) + (row >= 3 ? r2 : 0);
}

__device__ __forceinline__ int g4d_wave_sum(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ int g4d_wave_min(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const int o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    return v;
}

__device__ __forceinline__ uint64_t g4d_wave_min_u64(uint64_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint64_t o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    return v;
}

__device__ __forceinline__ int g4d_wave_max(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const int o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

// Wave-aggregated append: one atomic per wave; returns this lane's slot (or -1 if !pred).
__device__ __forceinline__ long long g4d_wave_append(bool pred, unsigned long long *counter)
{
    int tot;
    const int rank = g4d_wave_excl_count(pred, &tot);
    unsigned long long base = 0;
    const int leader = __ffsll((long long)__ballot(1)) - 1;
    if (__lane_id() == leader && tot > 0) base = atomicAdd(counter, (unsigned long long)tot);
    base = __shfl(base, leader, 64);
    return pred ? (long long)(base + rank) : -1;
}

// Wave-aggregated append of cnt >= 0 items per lane: one atomic per wave; returns the first
// slot of this lane's items.  Every lane of the wave must call it.
__device__ __forceinline__ unsigned long long g4d_wave_append_n(unsigned cnt, unsigned long long *counter)
{
    const int lane = __lane_id();
    unsigned inc = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    const unsigned tot = __shfl(inc, 63, 64);
    unsigned long long base = 0;
    if (lane == 0 && tot > 0) base = atomicAdd(counter, (unsigned long long)tot);
    base = __shfl(base, 0, 64);
    return base + (inc - cnt);
}
