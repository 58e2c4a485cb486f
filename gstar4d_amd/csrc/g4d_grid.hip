// gStar4D for MI355X — Morton order, PBA digital Voronoi, Voronoi-neighbour pairs, worksets.
//
// Reference stages restated here:
//   Morton sort          GD/GDelaunay/GDelData.cu:51-131
//   PBA                  GD/PBA/pba3DHost.cu:200-295, GD/PBA/pba3DKernel.h:40-579
//   missing points       GD/GDelaunay/GDelKernels.cu:109-135
//   grid pairs           GD/GDelaunay/GDelKernels.cu:137-299
//   worksets + pumping   GD/GDelaunay/GDelHost.cu:352-529, GD/GDelaunay/GDelKernels.cu:347-477
//
// PBA design for MI355X: each of the three axis passes is ONE read and ONE write of the
// int32 grid.  A 64-lane workgroup stages TX whole columns (G voxels each) in LDS with
// coalesced row loads, runs an exact-integer lower envelope per column in LDS (the Maurer
// stack lives in place in the consumed part of the column), and writes the coloured column
// back with coalesced stores.  The X pass reads rows (contiguous) and so needs no transpose;
// its final store converts site codes to point ids.
#include "g4d_grid.h"

namespace g4d {

// ------------------------------------------------------------------------------------
// Bounds and Morton keys
// ------------------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_bounds_partial(const double *xyz, int64_t m, double *part)
{
    __shared__ double smin[256], smax[256];
    __shared__ int sbad;
    if (threadIdx.x == 0) sbad = 0;
    __syncthreads();
    double lo = 1e308, hi = -1e308;
    bool bad = false;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < m; i += (int64_t)gridDim.x * 256) {
        const double v = xyz[i];
        bad |= !isfinite(v);
        lo = fmin(lo, v);
        hi = fmax(hi, v);
    }
    if (bad) atomicOr(&sbad, 1);
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + off]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = smin[0];
        part[3 * blockIdx.x + 1] = smax[0];
        part[3 * blockIdx.x + 2] = sbad ? 1.0 : 0.0;
    }
}

// Reduces the per-block partials into part[0] = min, part[1] = max; a non-finite input
// coordinate makes part[0] NaN so the host rejects the input.
__global__ void k_bounds_final(double *part, int nb)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double lo = part[0], hi = part[1];
    bool bad = part[2] != 0.0;
    for (int b = 1; b < nb; b++) {
        lo = fmin(lo, part[3 * b]);
        hi = fmax(hi, part[3 * b + 1]);
        bad |= part[3 * b + 2] != 0.0;
    }
    part[0] = bad ? __longlong_as_double(0x7ff8000000000000ll) : lo;
    part[1] = hi;
}

__device__ __forceinline__ int set_of(const GridSets &S, int64_t k) { return S.pt_set ? S.pt_set[k] : 0; }

// First set index whose prefix entry exceeds x (sets of a work decomposition).
__device__ __forceinline__ int set_of_work(const int64_t *off, int nsets, int64_t x)
{
    int lo = 0, hi = nsets - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// One work-group per set: min and max over the set's coordinates (NaN minimum on a
// non-finite coordinate).
__global__ void __launch_bounds__(256) k_set_bounds(const double *xyz, GridSets S, double *bounds)
{
    __shared__ double smin[256], smax[256];
    __shared__ int sbad;
    const int k = blockIdx.x;
    if (threadIdx.x == 0) sbad = 0;
    __syncthreads();
    double lo = 1e308, hi = -1e308;
    bool bad = false;
    for (int64_t i = 3 * S.pt_off[k] + threadIdx.x; i < 3 * S.pt_off[k + 1]; i += 256) {
        const double v = xyz[i];
        bad |= !isfinite(v);
        lo = fmin(lo, v);
        hi = fmax(hi, v);
    }
    if (bad) atomicOr(&sbad, 1);
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + off]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        bounds[2 * k] = sbad ? __longlong_as_double(0x7ff8000000000000ll) : smin[0];
        bounds[2 * k + 1] = smax[0];
    }
}

__global__ void k_set_ids(const int64_t *pt_off, int *pt_set)
{
    const int k = blockIdx.x;
    for (int64_t i = pt_off[k] + threadIdx.x; i < pt_off[k + 1]; i += blockDim.x) pt_set[i] = k;
}

// Spreads the low 21 bits of v to every third bit.
__device__ __forceinline__ uint64_t spread21(uint64_t v)
{
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

// Voxel of a coordinate, as Application::_scalePoint (GD/Main/Application.cpp:589-603) in
// FP64: (G-3) * (x - min) / (max - min) + 1, truncated; compiled without FP contraction so it
// is bit-identical to the host computation.
__device__ __forceinline__ int voxel_of(double x, double mn, double range, double span, int g)
{
    const double t = span * (x - mn);
    const double v = t / range + 1.0;
    int iv = (int)v;
    iv = iv < 1 ? 1 : iv;
    iv = iv > g - 2 ? g - 2 : iv;
    return iv;
}

__global__ void k_morton_keys(const double *xyz, int n, GridSets S, uint64_t *keys)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int st = set_of(S, i);
        const int g = S.g[st];
        const double mn = S.bounds[2 * st], range = S.bounds[2 * st + 1] - mn, span = (double)(g - 3);
        const int x = voxel_of(xyz[3 * (int64_t)i], mn, range, span, g);
        const int y = voxel_of(xyz[3 * (int64_t)i + 1], mn, range, span, g);
        const int z = voxel_of(xyz[3 * (int64_t)i + 2], mn, range, span, g);
        const uint64_t code = spread21(x) | (spread21(y) << 1) | (spread21(z) << 2);
        if (S.nsets == 1)
            keys[i] = (code << S.key_shift) | (uint64_t)i;
        else
            keys[i] = ((uint64_t)st << 48) | ((uint64_t)code << 24) | (uint64_t)(i - S.pt_off[st]);
    }
}

__global__ void k_gather_sorted(const double *xyz, const uint64_t *keys, int n, GridSets S, double *pts, int *perm,
                                uint64_t *vox)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint64_t key = keys[k];
        int st = 0, i;
        if (S.nsets == 1) {
            i = (int)(key & ((1ull << S.key_shift) - 1ull));
        } else {
            st = (int)(key >> 48);
            i = (int)(S.pt_off[st] + (int64_t)(key & 0xffffffull));
        }
        const int g = S.g[st];
        const double mn = S.bounds[2 * st], range = S.bounds[2 * st + 1] - mn, span = (double)(g - 3);
        const double x = xyz[3 * (int64_t)i], y = xyz[3 * (int64_t)i + 1], z = xyz[3 * (int64_t)i + 2];
        pts[3 * (int64_t)k] = x;
        pts[3 * (int64_t)k + 1] = y;
        pts[3 * (int64_t)k + 2] = z;
        perm[k] = i;
        vox[k] = G4D_VOX(voxel_of(x, mn, range, span, g), voxel_of(y, mn, range, span, g),
                         voxel_of(z, mn, range, span, g));
    }
}

// Equal points share a voxel, and equal voxels are adjacent in Morton order.
__global__ void k_dup_check(const double *pts, const uint64_t *vox, int n, GridSets S, unsigned long long *cnt)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int64_t first = S.pt_off[set_of(S, k)];
        for (int j = k - 1; j >= first && vox[j] == vox[k]; j--) {
            if (pts[3 * (int64_t)j] == pts[3 * (int64_t)k] && pts[3 * (int64_t)j + 1] == pts[3 * (int64_t)k + 1] &&
                pts[3 * (int64_t)j + 2] == pts[3 * (int64_t)k + 2]) {
                atomicOr(&cnt[G4D_CNT_ERROR], (unsigned long long)G4D_DEV_ERR_DUPLICATE);
                break;
            }
        }
    }
}

// ------------------------------------------------------------------------------------
// PBA
// ------------------------------------------------------------------------------------

__global__ void k_fill_i32(int *a, int64_t n, int v)
{
    int4 *a4 = reinterpret_cast<int4 *>(a);
    const int64_t n4 = n >> 2;
    const int4 vv = make_int4(v, v, v, v);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (int64_t)gridDim.x * blockDim.x)
        a4[i] = vv;
    for (int64_t i = (n4 << 2) + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

// Site scatter: the first point of each equal-voxel run (lowest Morton index) is the voxel's
// leader; the reference picks a racy last writer (GD/PBA/pba3DKernel.h:518-542).  Without an
// id grid the cells take the leader's point id instead of its site code.
__global__ void k_scatter_sites(const uint64_t *vox, int n, GridSets S, int *grid, int *idgrid)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint64_t v = vox[k];
        const int st = set_of(S, k);
        if (k > S.pt_off[st] && vox[k - 1] == v) continue;
        const int64_t g = S.g[st];
        const int x = G4D_VOX_X(v), y = G4D_VOX_Y(v), z = G4D_VOX_Z(v);
        const int64_t idx = S.vox_off[st] + ((int64_t)z * g + y) * g + x;
        if (!idgrid) { // point-id grid
            grid[idx] = k;
            continue;
        }
        const int c = G4D_ENC(x, y, z);
        grid[idx] = c;
        idgrid[idx] = k;
    }
}

// Maurer's test: site v (between u and w along the sweep) leaves the lower envelope.  Each site
// enters as its position q along the sweep and f = q^2 + (squared offset from the column), both
// integers: coordinates are below 2048, so f < 3 * 2^22, the differences below 2^24 and the
// products below 2^36.  In FP64 all of it is exact, and an FP64 multiply costs a quarter of a
// 32 x 32 -> 64-bit integer multiply.
__device__ __forceinline__ bool env_dominated(double qu, double fu, double qv, double fv, double qw, double fw)
{
    return (fv - fu) * (qw - qv) >= (fw - fv) * (qv - qu);
}

// Offset of a site from the column line (squared distance in the two non-sweep axes) and
// its position along the sweep axis.
template <int AXIS, bool IDS>
__device__ __forceinline__ void site_pos_off(int s, const uint64_t *vox, int ca, int cb, int *q, int *d)
{
    int sx, sy, sz;
    if (IDS) {
        const uint64_t v = vox[s];
        sx = G4D_VOX_X(v);
        sy = G4D_VOX_Y(v);
        sz = G4D_VOX_Z(v);
    } else {
        sx = G4D_DEC_X(s);
        sy = G4D_DEC_Y(s);
        sz = G4D_DEC_Z(s);
    }
    if (AXIS == 2) { // sweep z; column (x=ca, y=cb)
        *q = sz;
        const int dx = ca - sx, dy = cb - sy;
        *d = __mul24(dx, dx) + __mul24(dy, dy);
    } else if (AXIS == 1) { // sweep y; column (x=ca, z=cb)
        *q = sy;
        const int dx = ca - sx, dz = cb - sz;
        *d = __mul24(dx, dx) + __mul24(dz, dz);
    } else { // sweep x; column (y=ca, z=cb)
        *q = sx;
        const int dy = ca - sy, dz = cb - sz;
        *d = __mul24(dy, dy) + __mul24(dz, dz);
    }
}

// (q, f) of a site for env_dominated.
template <int AXIS, bool IDS>
__device__ __forceinline__ void site_qf(int s, const uint64_t *vox, int ca, int cb, double *q, double *f)
{
    int qi, di;
    site_pos_off<AXIS, IDS>(s, vox, ca, cb, &qi, &di);
    *q = (double)qi;
    *f = (double)(__mul24(qi, qi) + di);
}

// One wave per tile of 64 columns along the sweep axis, one lane per column.  The column is
// streamed through an LDS tile of PBA_CHUNK positions x 64 columns (coalesced loads and stores
// for every axis).  The lower envelope of the candidate sites (Maurer stack, exact integer
// dominance test) keeps its two top entries in registers, so a candidate costs one stack read
// per pop; the stack itself lives in LDS up to pba_stack(AXIS) entries, deeper entries in a
// per-block region of `spill` laid out [depth][lane] so that a wave's accesses share cache
// lines.  The LDS size sets how many waves share a CU, so it differs per pass: the z pass sees
// only the sites themselves (a shallow envelope); at 32 voxels per site the y pass's envelope
// of a 512-long column is about 115 entries deep (at most ~140), where a deep LDS stack pays;
// in the x pass more resident waves pay more than the spill they avoid (measured at 512^3).
// The colour walk writes the nearest site of every voxel back through the same tile.  Ties: a
// later site along the axis wins (the walk advances on <=).
static constexpr int PBA_COLS = 64;
static constexpr int PBA_CHUNK = 16;
__host__ __device__ constexpr int pba_stack(int axis) { return axis == 2 ? 32 : axis == 1 ? 120 : 56; }

template <int STACK>
struct PbaLds {
    int tile[PBA_CHUNK][PBA_COLS + 1];
    int scode[STACK][PBA_COLS];
};

template <int AXIS>
__device__ __forceinline__ int64_t pba_gidx64(int g, int fixed, int c, int q)
{
    const int64_t G = g;
    if (AXIS == 2) return ((int64_t)q * G + fixed) * G + c;   // z = q, y = fixed, x = c
    if (AXIS == 1) return ((int64_t)fixed * G + q) * G + c;   // z = fixed, y = q, x = c
    return ((int64_t)fixed * G + c) * G + q;                  // z = fixed, y = c, x = q
}

// Tile element k of a lane: column cc and position qq (position fastest for the X pass, whose
// positions are contiguous in memory; column fastest otherwise).
static constexpr int PBA_PER_LANE = PBA_CHUNK * PBA_COLS / 64;
template <int AXIS>
__device__ __forceinline__ int pba_tile_c(int k)
{
    const int i = (int)threadIdx.x + 64 * k;
    return AXIS == 0 ? i / PBA_CHUNK : i % PBA_COLS;
}
template <int AXIS>
__device__ __forceinline__ int pba_tile_q(int k)
{
    const int i = (int)threadIdx.x + 64 * k;
    return AXIS == 0 ? i % PBA_CHUNK : i / PBA_COLS;
}

// Issues every load of a chunk before any is consumed.
template <int AXIS>
__device__ __forceinline__ void pba_load_chunk(int (&pre)[PBA_PER_LANE], const int *in, int g, int fixed, int c0,
                                               int ncols, int r0)
{
    const int nq = min(PBA_CHUNK, g - r0);
#pragma unroll
    for (int k = 0; k < PBA_PER_LANE; k++) {
        const int cc = pba_tile_c<AXIS>(k), qq = pba_tile_q<AXIS>(k);
        pre[k] = (qq < nq && cc < ncols) ? in[pba_gidx64<AXIS>(g, fixed, c0 + cc, r0 + qq)] : G4D_MARKER;
    }
}

// Stack entry below the LDS part.  Kept out of line: inlined next to the LDS read, the
// compiler (ROCm 7.2, gfx950) emits an illegal operand for the unrolled tile loops.
__device__ __attribute__((noinline)) int pba_spill_get(const int *sp, int k, int stack)
{
    return sp[(int64_t)(k - stack) * PBA_COLS];
}

template <int AXIS, bool IDS>
__global__ void __launch_bounds__(64) k_pba_pass(const int *in_all, int *out_all, GridSets S, const int *idgrid_all,
                                                 int *spill_all, const uint64_t *vox)
{
    constexpr int PBA_STACK = pba_stack(AXIS);
    __shared__ PbaLds<PBA_STACK> L;
    const int lane = threadIdx.x;
    const int st = S.nsets == 1 ? 0 : set_of_work(S.pba_blk, S.nsets, blockIdx.x);
    const int64_t lb = blockIdx.x - S.pba_blk[st];
    const int g = S.g[st];
    const int *in = in_all + S.vox_off[st];
    int *out = out_all + S.vox_off[st];
    const int *idgrid = idgrid_all ? idgrid_all + S.vox_off[st] : nullptr;
    const int ntiles = (g + PBA_COLS - 1) / PBA_COLS;
    const int fixed = (int)(lb / ntiles);
    const int c0 = (int)(lb % ntiles) * PBA_COLS;
    const int ncols = min(PBA_COLS, g - c0);
    const bool active = lane < ncols;
    const int c = c0 + lane;
    const int ca = c, cb = fixed; // x (Z, Y passes) or y (X pass); y (Z), z (Y), z (X)
    int *sp = spill_all + S.spill[st] + lb * PBA_COLS * g + lane; // entry k >= PBA_STACK at (k - PBA_STACK) * 64
#define get(k) ((k) < PBA_STACK ? L.scode[(k)][lane] : pba_spill_get(sp, (k), PBA_STACK))

    // Tile element k of this lane: (cc, qq) with every load of a chunk issued before the first
    // is consumed, and the next chunk's loads in flight while the current one is processed.
    int pre[PBA_PER_LANE];

    // ---- lower envelope; (q0, f0) = entry m-2, (q1, f1) = entry m-1 ----
    int m = 0;
    double q0 = 0, f0 = 0, q1 = 0, f1 = 0;
    pba_load_chunk<AXIS>(pre, in, g, fixed, c0, ncols, 0);
    for (int r0 = 0; r0 < g; r0 += PBA_CHUNK) {
        const int nq = min(PBA_CHUNK, g - r0);
#pragma unroll
        for (int k = 0; k < PBA_PER_LANE; k++) {
            const int cc = pba_tile_c<AXIS>(k), qq = pba_tile_q<AXIS>(k);
            L.tile[qq][cc] = pre[k];
        }
        __syncthreads();
        if (r0 + PBA_CHUNK < g) pba_load_chunk<AXIS>(pre, in, g, fixed, c0, ncols, r0 + PBA_CHUNK);
        if (active) {
            for (int qq = 0; qq < nq; qq++) {
                const int s = L.tile[qq][lane];
                if (s == G4D_MARKER) continue;
                double qs, fs;
                site_qf<AXIS, IDS>(s, vox, ca, cb, &qs, &fs);
                while (m >= 2 && env_dominated(q0, f0, q1, f1, qs, fs)) {
                    m--;
                    q1 = q0;
                    f1 = f0;
                    if (m >= 2) site_qf<AXIS, IDS>(get(m - 2), vox, ca, cb, &q0, &f0);
                }
                if (m < PBA_STACK) L.scode[m][lane] = s;
                else sp[(int64_t)(m - PBA_STACK) * PBA_COLS] = s;
                q0 = q1;
                f0 = f1;
                q1 = qs;
                f1 = fs;
                m++;
            }
        }
        __syncthreads();
    }
    // ---- colour walk ----
    // With an id grid (the last pass), each envelope entry's point id is looked up once, when
    // it becomes the next entry of the walk (the load is in flight while the walk covers the
    // current entry's run of voxels), instead of once per voxel.
    const bool ids_out = !IDS && idgrid;
    auto id_of = [&](int code) {
        return idgrid[((int64_t)G4D_DEC_Z(code) * g + G4D_DEC_Y(code)) * g + G4D_DEC_X(code)];
    };
    int idx = 0, qa = 0, qb = 0, sa = G4D_MARKER, sb = G4D_MARKER;
    int da = 0, db = 0;
    if (active && m > 0) {
        sa = L.scode[0][lane];
        site_pos_off<AXIS, IDS>(sa, vox, ca, cb, &qa, &da);
        if (ids_out) sa = id_of(sa);
    }
    if (active && m > 1) {
        sb = get(1);
        site_pos_off<AXIS, IDS>(sb, vox, ca, cb, &qb, &db);
        if (ids_out) sb = id_of(sb);
    }
    for (int t0 = 0; t0 < g; t0 += PBA_CHUNK) {
        const int nt = min(PBA_CHUNK, g - t0);
        if (active) {
            for (int tt = 0; tt < nt; tt++) {
                const int t = t0 + tt;
                if (m == 0) {
                    L.tile[tt][lane] = ids_out ? -1 : G4D_MARKER;
                    continue;
                }
                while (idx + 1 < m) {
                    const int fa = __mul24(t - qa, t - qa) + da;
                    const int fb = __mul24(t - qb, t - qb) + db;
                    if (fb <= fa) {
                        idx++;
                        qa = qb;
                        da = db;
                        sa = sb;
                        if (idx + 1 < m) {
                            sb = get(idx + 1);
                            site_pos_off<AXIS, IDS>(sb, vox, ca, cb, &qb, &db);
                            if (ids_out) sb = id_of(sb);
                        }
                    } else {
                        break;
                    }
                }
                L.tile[tt][lane] = sa;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PBA_PER_LANE; k++) {
            const int cc = pba_tile_c<AXIS>(k), qq = pba_tile_q<AXIS>(k);
            pre[k] = L.tile[qq][cc];
        }
#pragma unroll
        for (int k = 0; k < PBA_PER_LANE; k++) {
            const int cc = pba_tile_c<AXIS>(k), qq = pba_tile_q<AXIS>(k);
            if (qq < nt && cc < ncols) out[pba_gidx64<AXIS>(g, fixed, c0 + cc, t0 + qq)] = pre[k];
        }
        __syncthreads();
    }
#undef get
}

// ------------------------------------------------------------------------------------
// Banded axis pass (the reference's bands: kernelMaurerAxis over m2 bands, kernelMergeBands,
// kernelColorAxis over m3 bands, GD/PBA/pba3DKernel.h:215-487).  A column of g positions is cut
// into `bands` bands of pba_band_len(g, bands) positions, and the pass runs in three launches:
//   1. k_pba_band_env: one wave per (column tile, band) builds the lower envelope of the band's
//      candidate sites (a band's envelope is short, so the LDS stack is small and many waves
//      share a CU) and writes it to `benv`;
//   2. k_pba_band_merge: one wave per column tile pushes the bands' envelopes, in order, through
//      one Maurer stack: a site hidden inside its band is hidden on the whole line, so the
//      column's envelope is the envelope of the bands' envelopes, and this walk sees only their
//      entries instead of every position; the result goes back to `benv`;
//   3. k_pba_band_color: one wave per (column tile, band) colours the band's positions, starting
//      at the envelope entry that owns the band's first position (binary search: entry i owns
//      position t up to the first i whose successor is strictly farther at t).
// `benv` and `spill` have the layout of the spill area: per column tile 64 g ints, band b's
// region at 64 b L, entry k of lane l at 64 k + l.  bcnt holds 64 counts per (tile, band),
// fcnt 64 per tile.
// ------------------------------------------------------------------------------------

__host__ __device__ inline int pba_band_len(int g, int bands)
{
    const int L = (g + bands - 1) / bands;
    return (L + PBA_CHUNK - 1) / PBA_CHUNK * PBA_CHUNK;
}
static constexpr int PBA_BAND_STACK = 48;
static constexpr int PBA_MERGE_PF = 8; // band entries in flight per lane in the merge

template <int AXIS, bool IDS>
__global__ void __launch_bounds__(64) k_pba_band_env(const int *in_all, GridSets S, int *spill_all, int *benv_all,
                                                     int *bcnt, const uint64_t *vox, int bands)
{
    __shared__ PbaLds<PBA_BAND_STACK> L;
    const int lane = threadIdx.x;
    const int64_t bb = blockIdx.x / bands;
    const int band = (int)(blockIdx.x % bands);
    const int st = S.nsets == 1 ? 0 : set_of_work(S.pba_blk, S.nsets, bb);
    const int64_t lb = bb - S.pba_blk[st];
    const int g = S.g[st];
    const int *in = in_all + S.vox_off[st];
    const int ntiles = (g + PBA_COLS - 1) / PBA_COLS;
    const int fixed = (int)(lb / ntiles);
    const int c0 = (int)(lb % ntiles) * PBA_COLS;
    const int ncols = min(PBA_COLS, g - c0);
    const bool active = lane < ncols;
    const int ca = c0 + lane, cb = fixed;
    const int Lb = pba_band_len(g, bands);
    const int rb = band * Lb, re = min(g, rb + Lb);
    const int64_t region = S.spill[st] + lb * PBA_COLS * g + (int64_t)band * PBA_COLS * Lb + lane;
    int *sp = spill_all + region;
    int *ev = benv_all + region;
#define get(k) ((k) < PBA_BAND_STACK ? L.scode[(k)][lane] : pba_spill_get(sp, (k), PBA_BAND_STACK))
    int pre[PBA_PER_LANE];
    int m = 0;
    double q0 = 0, f0 = 0, q1 = 0, f1 = 0;
    if (rb < re) pba_load_chunk<AXIS>(pre, in, g, fixed, c0, ncols, rb);
    for (int r0 = rb; r0 < re; r0 += PBA_CHUNK) {
        const int nq = min(PBA_CHUNK, re - r0);
#pragma unroll
        for (int k = 0; k < PBA_PER_LANE; k++) {
            const int cc = pba_tile_c<AXIS>(k), qq = pba_tile_q<AXIS>(k);
            L.tile[qq][cc] = pre[k];
        }
        __syncthreads();
        if (r0 + PBA_CHUNK < re) pba_load_chunk<AXIS>(pre, in, g, fixed, c0, ncols, r0 + PBA_CHUNK);
        if (active) {
            for (int qq = 0; qq < nq; qq++) {
                const int sv = L.tile[qq][lane];
                if (sv == G4D_MARKER) continue;
                double qs, fs;
                site_qf<AXIS, IDS>(sv, vox, ca, cb, &qs, &fs);
                while (m >= 2 && env_dominated(q0, f0, q1, f1, qs, fs)) {
                    m--;
                    q1 = q0;
                    f1 = f0;
                    if (m >= 2) site_qf<AXIS, IDS>(get(m - 2), vox, ca, cb, &q0, &f0);
                }
                if (m < PBA_BAND_STACK) L.scode[m][lane] = sv;
                else sp[(int64_t)(m - PBA_BAND_STACK) * PBA_COLS] = sv;
                q0 = q1;
                f0 = f1;
                q1 = qs;
                f1 = fs;
                m++;
            }
        }
        __syncthreads();
    }
    if (active) {
        for (int k = 0; k < m; k++) ev[(int64_t)k * PBA_COLS] = get(k);
        bcnt[(bb * bands + band) * PBA_COLS + lane] = m;
    }
#undef get
}

template <int AXIS, bool IDS>
__global__ void __launch_bounds__(64) k_pba_band_merge(GridSets S, int *spill_all, int *benv_all, const int *bcnt,
                                                       int *fcnt, const uint64_t *vox, int bands)
{
    constexpr int STACK = pba_stack(AXIS);
    __shared__ int scode[STACK][PBA_COLS];
    const int lane = threadIdx.x;
    const int64_t bb = blockIdx.x;
    const int st = S.nsets == 1 ? 0 : set_of_work(S.pba_blk, S.nsets, bb);
    const int64_t lb = bb - S.pba_blk[st];
    const int g = S.g[st];
    const int ntiles = (g + PBA_COLS - 1) / PBA_COLS;
    const int fixed = (int)(lb / ntiles);
    const int c0 = (int)(lb % ntiles) * PBA_COLS;
    if (lane >= min(PBA_COLS, g - c0)) return; // no barriers below: lanes are independent
    const int ca = c0 + lane, cb = fixed;
    const int Lb = pba_band_len(g, bands);
    const int64_t base = S.spill[st] + lb * PBA_COLS * g + lane;
    int *sp = spill_all + base; // entry k >= STACK of the merged stack at (k - STACK) * 64
    int *ev = benv_all + base;
#define get(k) ((k) < STACK ? scode[(k)][lane] : pba_spill_get(sp, (k), STACK))
    int m = 0;
    double q0 = 0, f0 = 0, q1 = 0, f1 = 0;
    // A band's entries are read PBA_MERGE_PF at a time, the next group's loads in flight while
    // the current group goes through the stack (each entry's load is otherwise one dependent
    // round trip to memory).
    for (int band = 0; band < bands; band++) {
        const int nb = band * Lb < g ? bcnt[(bb * bands + band) * PBA_COLS + lane] : 0;
        const int *bev = ev + (int64_t)band * PBA_COLS * Lb;
        int nxt[PBA_MERGE_PF];
#pragma unroll
        for (int i = 0; i < PBA_MERGE_PF; i++) nxt[i] = i < nb ? bev[(int64_t)i * PBA_COLS] : 0;
        for (int k0 = 0; k0 < nb; k0 += PBA_MERGE_PF) {
            int cur[PBA_MERGE_PF];
#pragma unroll
            for (int i = 0; i < PBA_MERGE_PF; i++) {
                cur[i] = nxt[i];
                const int kn = k0 + PBA_MERGE_PF + i;
                nxt[i] = kn < nb ? bev[(int64_t)kn * PBA_COLS] : 0;
            }
#pragma unroll
            for (int i = 0; i < PBA_MERGE_PF; i++) {
            if (k0 + i >= nb) break;
            const int sv = cur[i];
            double qs, fs;
            site_qf<AXIS, IDS>(sv, vox, ca, cb, &qs, &fs);
            while (m >= 2 && env_dominated(q0, f0, q1, f1, qs, fs)) {
                m--;
                q1 = q0;
                f1 = f0;
                if (m >= 2) site_qf<AXIS, IDS>(get(m - 2), vox, ca, cb, &q0, &f0);
            }
            if (m < STACK) scode[m][lane] = sv;
            else sp[(int64_t)(m - STACK) * PBA_COLS] = sv;
            q0 = q1;
            f0 = f1;
            q1 = qs;
            f1 = fs;
            m++;
            }
        }
    }
    // every band entry has been read: the merged envelope goes to the front of the region
    for (int k = 0; k < m; k++) ev[(int64_t)k * PBA_COLS] = get(k);
    fcnt[bb * PBA_COLS + lane] = m;
#undef get
}

template <int AXIS, bool IDS>
__global__ void __launch_bounds__(64) k_pba_band_color(GridSets S, const int *benv_all, const int *fcnt, int *out_all,
                                                       const int *idgrid_all, const uint64_t *vox, int bands)
{
    __shared__ int tile[PBA_CHUNK][PBA_COLS + 1];
    const int lane = threadIdx.x;
    const int64_t bb = blockIdx.x / bands;
    const int band = (int)(blockIdx.x % bands);
    const int st = S.nsets == 1 ? 0 : set_of_work(S.pba_blk, S.nsets, bb);
    const int64_t lb = bb - S.pba_blk[st];
    const int g = S.g[st];
    int *out = out_all + S.vox_off[st];
    const int *idgrid = idgrid_all ? idgrid_all + S.vox_off[st] : nullptr;
    const int ntiles = (g + PBA_COLS - 1) / PBA_COLS;
    const int fixed = (int)(lb / ntiles);
    const int c0 = (int)(lb % ntiles) * PBA_COLS;
    const int ncols = min(PBA_COLS, g - c0);
    const bool active = lane < ncols;
    const int ca = c0 + lane, cb = fixed;
    const int Lb = pba_band_len(g, bands);
    const int rb = band * Lb, re = min(g, rb + Lb);
    if (rb >= re) return; // an empty band: the whole work-group leaves before any barrier
    const int *ev = benv_all + S.spill[st] + lb * PBA_COLS * g + lane;
    const int m = active ? fcnt[bb * PBA_COLS + lane] : 0;
    const bool ids_out = !IDS && idgrid;
    auto id_of = [&](int code) {
        return idgrid[((int64_t)G4D_DEC_Z(code) * g + G4D_DEC_Y(code)) * g + G4D_DEC_X(code)];
    };
    // owner of position rb: the first entry whose successor is strictly farther there
    int idx = 0;
    if (active && m > 1) {
        int lo = 0, hi = m - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            int qa, da, qb, db;
            site_pos_off<AXIS, IDS>(ev[(int64_t)mid * PBA_COLS], vox, ca, cb, &qa, &da);
            site_pos_off<AXIS, IDS>(ev[(int64_t)(mid + 1) * PBA_COLS], vox, ca, cb, &qb, &db);
            const int fa = __mul24(rb - qa, rb - qa) + da, fb = __mul24(rb - qb, rb - qb) + db;
            if (fb > fa) hi = mid;
            else lo = mid + 1;
        }
        idx = lo;
    }
    // The walk holds the current entry (sa), the next one decoded (sb, qb, db) and the raw
    // entry after it (sc), loaded one advance ahead so that an advance waits on no load.
    int qa = 0, qb = 0, sa = G4D_MARKER, sb = G4D_MARKER, sc = G4D_MARKER, da = 0, db = 0;
    if (active && m > 0) {
        sa = ev[(int64_t)idx * PBA_COLS];
        site_pos_off<AXIS, IDS>(sa, vox, ca, cb, &qa, &da);
        if (ids_out) sa = id_of(sa);
    }
    if (active && idx + 1 < m) {
        sb = ev[(int64_t)(idx + 1) * PBA_COLS];
        site_pos_off<AXIS, IDS>(sb, vox, ca, cb, &qb, &db);
        if (ids_out) sb = id_of(sb);
    }
    if (active && idx + 2 < m) sc = ev[(int64_t)(idx + 2) * PBA_COLS];
    int pre[PBA_PER_LANE];
    for (int t0 = rb; t0 < re; t0 += PBA_CHUNK) {
        const int nt = min(PBA_CHUNK, re - t0);
        if (active) {
            for (int tt = 0; tt < nt; tt++) {
                const int t = t0 + tt;
                if (m == 0) {
                    tile[tt][lane] = ids_out ? -1 : G4D_MARKER;
                    continue;
                }
                while (idx + 1 < m) {
                    const int fa = __mul24(t - qa, t - qa) + da;
                    const int fb = __mul24(t - qb, t - qb) + db;
                    if (fb <= fa) {
                        idx++;
                        qa = qb;
                        da = db;
                        sa = sb;
                        if (idx + 1 < m) {
                            site_pos_off<AXIS, IDS>(sc, vox, ca, cb, &qb, &db);
                            sb = ids_out ? id_of(sc) : sc;
                            sc = idx + 2 < m ? ev[(int64_t)(idx + 2) * PBA_COLS] : G4D_MARKER;
                        }
                    } else {
                        break;
                    }
                }
                tile[tt][lane] = sa;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PBA_PER_LANE; k++) {
            const int cc = pba_tile_c<AXIS>(k), qq = pba_tile_q<AXIS>(k);
            pre[k] = tile[qq][cc];
        }
#pragma unroll
        for (int k = 0; k < PBA_PER_LANE; k++) {
            const int cc = pba_tile_c<AXIS>(k), qq = pba_tile_q<AXIS>(k);
            if (qq < nt && cc < ncols) out[pba_gidx64<AXIS>(g, fixed, c0 + cc, t0 + qq)] = pre[k];
        }
        __syncthreads();
    }
}

template <int AXIS, bool IDS>
static void launch_band_pass(const GridSets &S, int64_t nblocks, int bands, const int *in, int *out,
                             const int *idgrid, int *spill, int *benv, int *bcnt, int *fcnt, const uint64_t *vox,
                             hipStream_t s)
{
    const dim3 bl(64);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pba_band_env<AXIS, IDS>), dim3((unsigned)(nblocks * bands)), bl, 0, s, in, S,
                       spill, benv, bcnt, vox, bands);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pba_band_merge<AXIS, IDS>), dim3((unsigned)nblocks), bl, 0, s, S, spill, benv,
                       bcnt, fcnt, vox, bands);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pba_band_color<AXIS, IDS>), dim3((unsigned)(nblocks * bands)), bl, 0, s, S,
                       benv, fcnt, out, idgrid, vox, bands);
}

hipError_t pba_pass_banded(int axis, const GridSets &S, int64_t nblocks, int bands, const int *in, int *out,
                           const int *idgrid, int *spill, int *benv, int *bcnt, int *fcnt, const uint64_t *vox,
                           hipStream_t s)
{
    if (nblocks <= 0 || bands < 2 || bands > PBA_MAX_BANDS || nblocks * bands > INT32_MAX) return hipErrorInvalidValue;
    if (axis != 0 && axis != 1) return hipErrorInvalidValue;
    if (vox && S.nsets != 1) return hipErrorInvalidValue;
    if (vox) {
        if (axis == 1) launch_band_pass<1, true>(S, nblocks, bands, in, out, idgrid, spill, benv, bcnt, fcnt, vox, s);
        else launch_band_pass<0, true>(S, nblocks, bands, in, out, idgrid, spill, benv, bcnt, fcnt, vox, s);
    } else {
        if (axis == 1) launch_band_pass<1, false>(S, nblocks, bands, in, out, idgrid, spill, benv, bcnt, fcnt, vox, s);
        else launch_band_pass<0, false>(S, nblocks, bands, in, out, idgrid, spill, benv, bcnt, fcnt, vox, s);
    }
    return hipGetLastError();
}

int64_t pba_blocks(int g) { return (int64_t)g * ((g + PBA_COLS - 1) / PBA_COLS); }

int64_t pba_spill_ints(int g) { return pba_blocks(g) * PBA_COLS * g; }

hipError_t pba_pass(int axis, const GridSets &S, int64_t nblocks, const int *in, int *out, const int *idgrid,
                    int *spill, const uint64_t *vox, hipStream_t s)
{
    if (nblocks <= 0 || nblocks > INT32_MAX) return hipErrorInvalidValue;
    const dim3 gr((unsigned)nblocks), bl(64);
    if (vox) {
        if (S.nsets != 1) return hipErrorInvalidValue;
        if (axis == 2) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pba_pass<2, true>), gr, bl, 0, s, in, out, S, idgrid, spill, vox);
        else if (axis == 1) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pba_pass<1, true>), gr, bl, 0, s, in, out, S, idgrid, spill, vox);
        else hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pba_pass<0, true>), gr, bl, 0, s, in, out, S, idgrid, spill, vox);
    } else {
        if (axis == 2) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pba_pass<2, false>), gr, bl, 0, s, in, out, S, idgrid, spill, vox);
        else if (axis == 1) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pba_pass<1, false>), gr, bl, 0, s, in, out, S, idgrid, spill, vox);
        else hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pba_pass<0, false>), gr, bl, 0, s, in, out, S, idgrid, spill, vox);
    }
    return hipGetLastError();
}

int morton_key_shift(int g)
{
    int b = 10;
    while ((1 << b) < g) b++;
    return 64 - 3 * b;
}

// ------------------------------------------------------------------------------------
// Missing points and Voronoi-neighbour pairs.  Keys are directed (star<<32 | vertex); both
// directions are emitted (the reference reciprocates at GD/GDelaunay/GDelHost.cu:380-386).
// ------------------------------------------------------------------------------------

__device__ __forceinline__ uint64_t upair(int a, int b)
{
    return a < b ? (((uint64_t)(uint32_t)a << 32) | (uint32_t)b) : (((uint64_t)(uint32_t)b << 32) | (uint32_t)a);
}

// Missing points (GD/GDelaunay/GDelKernels.cu:109-135): a point whose voxel is led by another
// point pairs with that leader.  Emits ordered pairs (min<<32 | max).
__global__ void k_missing(const uint64_t *vox, int n, GridSets S, const int *grid, uint64_t *out,
                          unsigned long long *counter, int64_t cap)
{
    for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const int k = base + threadIdx.x;
        bool miss = false;
        int leader = 0;
        if (k < n) {
            const uint64_t v = vox[k];
            const int st = set_of(S, k);
            const int64_t g = S.g[st];
            leader = grid[S.vox_off[st] + ((int64_t)G4D_VOX_Z(v) * g + G4D_VOX_Y(v)) * g + G4D_VOX_X(v)];
            miss = leader != k;
        }
        const long long slot = g4d_wave_append(miss, counter);
        if (miss && slot < cap) out[slot] = upair(k, leader);
    }
}

__device__ __forceinline__ int grid_at(const int *grid, int g, int x, int y, int z)
{
    if (x < 0 || y < 0 || z < 0 || x >= g || y >= g || z >= g) return G4D_MARKER;
    return grid[((int64_t)z * g + y) * g + x];
}

// Voronoi-neighbour pairs of one 2x2x2 cube (GD/GDelaunay/GDelKernels.cu:181-290): corner
// c = ix + 2 iy + 4 iz, root = corner 0, opp = corner 7, link cycle 6,2,3,1,5,4.  Slot
// 5i+{0..4} holds the pairs of link edge i (root-a, opp-a, root-b, opp-b, a-b); slot 30 is
// root-opp.  Returns the bit mask of emitted slots.
__device__ __forceinline__ uint32_t cube_pairs(const int *v, uint64_t *pr)
{
    const int root = v[0], opp = v[7];
    if (root == opp) return 0u;
    const bool has_marker = (root == G4D_MARKER) || (opp == G4D_MARKER);
    const int link[7] = {v[6], v[2], v[3], v[1], v[5], v[4], v[6]};
    uint32_t mask = 0;
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const int a = link[i], b = link[i + 1];
        if (a != b && a != root && a != opp && b != root && b != opp) {
            // grabPair skips a pair whose first value is the Marker (:155-176)
            if (root != G4D_MARKER) { mask |= 1u << (5 * i + 0); pr[5 * i + 0] = upair(root, a); }
            if (opp != G4D_MARKER) { mask |= 1u << (5 * i + 1); pr[5 * i + 1] = upair(opp, a); }
            if (root != G4D_MARKER) { mask |= 1u << (5 * i + 2); pr[5 * i + 2] = upair(root, b); }
            if (opp != G4D_MARKER) { mask |= 1u << (5 * i + 3); pr[5 * i + 3] = upair(opp, b); }
            mask |= 1u << (5 * i + 4);
            pr[5 * i + 4] = upair(a, b);
        }
    }
    if (mask && !has_marker) {
        mask |= 1u << 30;
        pr[30] = upair(root, opp);
    }
    return mask;
}

__host__ __device__ inline int grid_pair_chunk_dev(int g) { return g > 512 ? 32 : 16; }

// Threads are laid out in tiles of GP_TILE x GP_TILE (x, y) columns of cubes (x, y from -1,
// the outer Marker layer) and a chunk of zc cubes along z; a set's thread count is a whole
// number of tiles, so a 256-thread work-group never spans two sets.
static constexpr int GP_TILE = 16;
static constexpr int GP_THREADS = GP_TILE * GP_TILE;

__device__ __forceinline__ bool gp_coords(int64_t lt, int g, int zc, int *x, int *y, int *z0, int *z1)
{
    const int w = g + 1;
    const int ntx = (w + GP_TILE - 1) / GP_TILE;
    const int64_t tile = lt / GP_THREADS;
    const int t = (int)(lt % GP_THREADS);
    *x = (int)(tile % ntx) * GP_TILE + t % GP_TILE - 1;
    *y = (int)((tile / ntx) % ntx) * GP_TILE + t / GP_TILE - 1;
    const int zb = (int)(tile / ((int64_t)ntx * ntx));
    *z0 = zb * zc - 1;
    *z1 = min(*z0 + zc, g);
    return *x < g && *y < g;
}

// Pass 0 counts each thread's raw pairs, pass 1 writes them at the scanned offset
// (deterministic order, no atomics).
template <bool WRITE>
__global__ void __launch_bounds__(256) k_grid_pairs(const int *grid_all, GridSets S, int64_t nthr, int64_t *count,
                                                    uint64_t *out)
{
    for (int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; tid < nthr; tid += (int64_t)gridDim.x * blockDim.x) {
        const int st = S.nsets == 1 ? 0 : set_of_work(S.pair_thr, S.nsets, tid);
        const int64_t lt = tid - S.pair_thr[st];
        const int g = S.g[st];
        const int *grid = grid_all + S.vox_off[st];
        int x, y, z0, z1;
        if (!gp_coords(lt, g, grid_pair_chunk_dev(g), &x, &y, &z0, &z1)) {
            if (!WRITE) count[tid] = 0;
            continue;
        }
        int v[8];
        v[0] = grid_at(grid, g, x, y, z0);
        v[1] = grid_at(grid, g, x + 1, y, z0);
        v[2] = grid_at(grid, g, x, y + 1, z0);
        v[3] = grid_at(grid, g, x + 1, y + 1, z0);
        int64_t k = WRITE ? count[tid] : 0;
        uint64_t pr[31];
        for (int z = z0; z < z1; z++) {
            v[4] = grid_at(grid, g, x, y, z + 1);
            v[5] = grid_at(grid, g, x + 1, y, z + 1);
            v[6] = grid_at(grid, g, x, y + 1, z + 1);
            v[7] = grid_at(grid, g, x + 1, y + 1, z + 1);
            const uint32_t mask = cube_pairs(v, pr);
            if (WRITE) {
#pragma unroll
                for (int j = 0; j < 31; j++)
                    if (mask & (1u << j)) out[k++] = pr[j];
            } else {
                k += __popc(mask);
            }
            v[0] = v[4];
            v[1] = v[5];
            v[2] = v[6];
            v[3] = v[7];
        }
        if (!WRITE) count[tid] = k;
    }
}

// Pass 1 with the work-group's pairs deduplicated in an LDS hash set: one work-group per tile,
// whose raw pairs have the region [off[first thread], off[first thread of the next tile]) of
// `out`.  Distinct pairs go to the bottom of the region in table order; a pair that finds no
// free slot within GP_PROBES probes is written as is from the top of the region down.  Either
// way the region holds at most the tile's raw count.  cnt[tile] = (table count, overflow count).
// Adjacent columns see the same Voronoi faces, so a 16 x 16 x zc tile keeps roughly one pair in
// four; the global sort and unique that follow remove what remains duplicated across tiles.
static constexpr int GP_HASH = 4096;
static constexpr int GP_PROBES = 32;
static constexpr int GP_STAGE = 256;
static constexpr uint64_t GP_EMPTY = ~0ull;

__device__ __forceinline__ void gp_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inserts a pair into the tile's hash set (a plain read settles a repeat without an atomic);
// a pair that finds no free slot within GP_PROBES probes goes straight to the top of the region.
__device__ __forceinline__ void gp_insert(unsigned long long *tab, uint64_t key, uint64_t *out, int64_t rend,
                                          unsigned int *nover)
{
    uint32_t h = (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 52);
    bool done = false;
    for (int p = 0; p < GP_PROBES && !done; p++) {
        unsigned long long cur = ((volatile unsigned long long *)tab)[h];
        if (cur == GP_EMPTY) cur = atomicCAS(&tab[h], GP_EMPTY, (unsigned long long)key);
        done = cur == GP_EMPTY || cur == key;
        h = (h + 1) & (GP_HASH - 1);
    }
    if (!done) out[rend - 1 - atomicAdd(nover, 1u)] = key;
}

__global__ void __launch_bounds__(GP_THREADS) k_grid_pairs_dedup(const int *grid_all, GridSets S, const int64_t *off,
                                                                 uint64_t *out, uint2 *cnt)
{
    __shared__ unsigned long long tab[GP_HASH];
    __shared__ unsigned long long stage[GP_THREADS / 64][GP_STAGE];
    __shared__ unsigned int nover;
    __shared__ unsigned int wsum[GP_THREADS / 64];
    for (int i = threadIdx.x; i < GP_HASH; i += GP_THREADS) tab[i] = GP_EMPTY;
    if (threadIdx.x == 0) nover = 0;
    __syncthreads();
    const int64_t tid = (int64_t)blockIdx.x * GP_THREADS + threadIdx.x;
    const int st = S.nsets == 1 ? 0 : set_of_work(S.pair_thr, S.nsets, tid);
    const int64_t lt = tid - S.pair_thr[st];
    const int g = S.g[st];
    const int *grid = grid_all + S.vox_off[st];
    const int64_t rbeg = off[(int64_t)blockIdx.x * GP_THREADS];
    const int64_t rend = off[(int64_t)(blockIdx.x + 1) * GP_THREADS];
    // Every lane walks its column (a lane outside the grid sees only Marker cubes, which emit
    // nothing), so the wave steps through its cubes in lockstep: a step's pairs are first
    // packed into the wave's LDS stage (exclusive scan of the lanes' counts) and then inserted
    // by all 64 lanes together, instead of one pass per pair slot that any lane uses.  A step
    // with more pairs than the stage holds inserts slot by slot.
    int x, y, z0, z1;
    (void)gp_coords(lt, g, grid_pair_chunk_dev(g), &x, &y, &z0, &z1);
    const int lane = __lane_id(), wid = threadIdx.x >> 6;
    unsigned long long *stg = stage[wid];
    int v[8];
    v[0] = grid_at(grid, g, x, y, z0);
    v[1] = grid_at(grid, g, x + 1, y, z0);
    v[2] = grid_at(grid, g, x, y + 1, z0);
    v[3] = grid_at(grid, g, x + 1, y + 1, z0);
    uint64_t pr[31];
    for (int z = z0; z < z1; z++) {
        v[4] = grid_at(grid, g, x, y, z + 1);
        v[5] = grid_at(grid, g, x + 1, y, z + 1);
        v[6] = grid_at(grid, g, x, y + 1, z + 1);
        v[7] = grid_at(grid, g, x + 1, y + 1, z + 1);
        const uint32_t mask = cube_pairs(v, pr);
        const uint32_t c = (uint32_t)__popc(mask);
        const uint32_t inc = (uint32_t)g4d_wave_incl_sum((int)c);
        const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
        if (tot <= GP_STAGE) {
            uint32_t pos = inc - c;
#pragma unroll
            for (int j = 0; j < 31; j++)
                if (mask & (1u << j)) stg[pos++] = pr[j];
            gp_wave_sync();
            for (uint32_t k = lane; k < tot; k += 64) gp_insert(tab, stg[k], out, rend, &nover);
            gp_wave_sync();
        } else {
#pragma unroll
            for (int j = 0; j < 31; j++)
                if (mask & (1u << j)) gp_insert(tab, pr[j], out, rend, &nover);
        }
        v[0] = v[4];
        v[1] = v[5];
        v[2] = v[6];
        v[3] = v[7];
    }
    __syncthreads();
    // compact the table: GP_HASH / GP_THREADS slots per thread, block-wide exclusive scan
    constexpr int PER = GP_HASH / GP_THREADS;
    unsigned long long e[PER];
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < PER; i++) {
        e[i] = tab[i * GP_THREADS + threadIdx.x];
        c += e[i] != GP_EMPTY;
    }
    const uint32_t inc = (uint32_t)g4d_wave_incl_sum((int)c);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    uint32_t base = 0, total = 0;
#pragma unroll
    for (int w = 0; w < GP_THREADS / 64; w++) {
        if (w < wid) base += wsum[w];
        total += wsum[w];
    }
    uint64_t *dst = out + rbeg + base + inc - c;
#pragma unroll
    for (int i = 0; i < PER; i++)
        if (e[i] != GP_EMPTY) *dst++ = e[i];
    if (threadIdx.x == 0) cnt[blockIdx.x] = make_uint2(total, nover);
}

// Packs every tile's table pairs and overflow pairs densely: tile b's pairs go to
// dst + doff[b], doff being the exclusive scan of the tiles' total counts.
__global__ void __launch_bounds__(256) k_gather_tiles(const uint64_t *src, const int64_t *off, const uint2 *cnt,
                                                      const int64_t *doff, uint64_t *dst)
{
    const int64_t b = blockIdx.x;
    const int64_t rbeg = off[b * GP_THREADS], rend = off[(b + 1) * GP_THREADS];
    const uint2 c = cnt[b];
    uint64_t *d = dst + doff[b];
    for (uint32_t i = threadIdx.x; i < c.x; i += blockDim.x) d[i] = src[rbeg + i];
    for (uint32_t i = threadIdx.x; i < c.y; i += blockDim.x) d[c.x + i] = src[rend - c.y + i];
}

__global__ void k_tile_totals(const uint2 *cnt, int64_t nb, int64_t *tot)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += (int64_t)gridDim.x * blockDim.x)
        tot[i] = (int64_t)cnt[i].x + cnt[i].y;
}

// CSR offsets of sorted directed keys: off[s] = first index with star >= s (off has n+1).
__global__ void k_csr_offsets(const uint64_t *keys, int64_t m, int n, int64_t *off)
{
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s <= n; s += gridDim.x * blockDim.x) {
        const uint64_t want = (uint64_t)(uint32_t)s << 32;
        int64_t lo = 0, hi = m;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (keys[mid] < want) lo = mid + 1;
            else hi = mid;
        }
        off[s] = lo;
    }
}

__device__ __forceinline__ bool list_has(const uint64_t *keys, int64_t b, int64_t e, int v)
{
    for (int64_t i = b; i < e; i++)
        if ((int)(uint32_t)keys[i] == v) return true;
    return false;
}

// Pumping of stars with fewer than min_ws candidates (GD/GDelaunay/GDelKernels.cu:347-477):
// neighbours' candidates in ascending order, then ids 0,1,2,...  Emits both directions.
__global__ void k_pump(const uint64_t *keys, const int64_t *off, int n, GridSets S, int min_ws, uint64_t *out,
                       unsigned long long *counter, int64_t cap)
{
    // the loop bound is uniform per work-group, so whole waves reach the aggregated append
    for (int s0 = blockIdx.x * blockDim.x; s0 < n; s0 += gridDim.x * blockDim.x) {
        const int s = s0 + threadIdx.x;
        int got[64];
        int ng = 0;
        const int64_t b = s < n ? off[s] : 0, e = s < n ? off[s + 1] : 0;
        const int have = (int)(e - b);
        if (s < n && have < min_ws) {
            const int need = min_ws - have;
            for (int64_t k = b; k < e && ng < need; k++) {
                const int nei = (int)(uint32_t)keys[k];
                for (int64_t q = off[nei]; q < off[nei + 1] && ng < need; q++) {
                    const int cv = (int)(uint32_t)keys[q];
                    if (cv == s || list_has(keys, b, e, cv)) continue;
                    bool dup = false;
                    for (int r = 0; r < ng; r++) dup |= (got[r] == cv);
                    if (!dup) got[ng++] = cv;
                }
            }
            const int st = set_of(S, s);
            for (int cv = (int)S.pt_off[st]; ng < need && cv < (int)S.pt_off[st + 1]; cv++) {
                if (cv == s || list_has(keys, b, e, cv)) continue;
                bool dup = false;
                for (int r = 0; r < ng; r++) dup |= (got[r] == cv);
                if (!dup) got[ng++] = cv;
            }
        }
        const unsigned long long slot = g4d_wave_append_n((unsigned)ng, counter);
        for (int r = 0; r < ng; r++) {
            const int64_t k = 2 * (int64_t)(slot + r);
            if (k + 1 < cap) {
                out[k] = ((uint64_t)(uint32_t)s << 32) | (uint32_t)got[r];
                out[k + 1] = ((uint64_t)(uint32_t)got[r] << 32) | (uint32_t)s;
            }
        }
    }
}

// ------------------------------------------------------------------------------------
// Host-side launchers
// ------------------------------------------------------------------------------------

hipError_t launch_bounds(const double *xyz, int n, double *part, hipStream_t s)
{
    const int nb = 256;
    hipLaunchKernelGGL(k_bounds_partial, dim3(nb), dim3(256), 0, s, xyz, 3 * (int64_t)n, part);
    HIP_CHECK_RET(hipGetLastError());
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(64), 0, s, part, nb);
    return hipGetLastError();
}

hipError_t launch_set_bounds(const double *xyz, const GridSets &S, double *bounds, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_bounds, dim3(S.nsets), dim3(256), 0, s, xyz, S, bounds);
    return hipGetLastError();
}

hipError_t launch_set_ids(const int64_t *pt_off, int nsets, int *pt_set, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_ids, dim3(nsets), dim3(256), 0, s, pt_off, pt_set);
    return hipGetLastError();
}

hipError_t launch_morton_keys(const double *xyz, int n, const GridSets &S, uint64_t *keys, hipStream_t s)
{
    hipLaunchKernelGGL(k_morton_keys, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, xyz, n, S, keys);
    return hipGetLastError();
}

hipError_t launch_gather_sorted(const double *xyz, const uint64_t *keys, int n, const GridSets &S, double *pts,
                                int *perm, uint64_t *vox, hipStream_t s)
{
    hipLaunchKernelGGL(k_gather_sorted, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, xyz, keys, n, S, pts, perm,
                       vox);
    return hipGetLastError();
}

hipError_t launch_dup_check(const double *pts, const uint64_t *vox, int n, const GridSets &S, unsigned long long *cnt,
                            hipStream_t s)
{
    hipLaunchKernelGGL(k_dup_check, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, pts, vox, n, S, cnt);
    return hipGetLastError();
}

hipError_t launch_fill_i32(int *a, int64_t n, int v, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_i32, dim3(g4d_blocks(n / 4 + 1, 256, 65536)), dim3(256), 0, s, a, n, v);
    return hipGetLastError();
}

hipError_t launch_scatter_sites(const uint64_t *vox, int n, const GridSets &S, int *grid, int *idgrid, hipStream_t s)
{
    hipLaunchKernelGGL(k_scatter_sites, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, vox, n, S, grid, idgrid);
    return hipGetLastError();
}

hipError_t launch_missing(const uint64_t *vox, int n, const GridSets &S, const int *grid, uint64_t *out,
                          unsigned long long *counter, int64_t cap, hipStream_t s)
{
    hipLaunchKernelGGL(k_missing, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, vox, n, S, grid, out, counter, cap);
    return hipGetLastError();
}

int grid_pair_chunk(int g) { return grid_pair_chunk_dev(g); }

int64_t grid_pair_threads(int g, int zc)
{
    const int64_t w = g + 1, ntx = (w + GP_TILE - 1) / GP_TILE;
    return ntx * ntx * ((w + zc - 1) / zc) * GP_THREADS;
}

hipError_t launch_grid_pairs_dedup(const int *grid, const GridSets &S, int64_t nthr, const int64_t *off, uint64_t *out,
                                   uint2 *cnt, hipStream_t s)
{
    if (nthr <= 0) return hipSuccess;
    if (nthr % GP_THREADS || nthr / GP_THREADS > INT32_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_grid_pairs_dedup, dim3((unsigned)(nthr / GP_THREADS)), dim3(GP_THREADS), 0, s, grid, S, off, out,
                       cnt);
    return hipGetLastError();
}

hipError_t launch_tile_totals(const uint2 *cnt, int64_t ntiles, int64_t *tot, hipStream_t s)
{
    hipLaunchKernelGGL(k_tile_totals, dim3(g4d_blocks(ntiles, 256, 65536)), dim3(256), 0, s, cnt, ntiles, tot);
    return hipGetLastError();
}

hipError_t launch_gather_tiles(const uint64_t *src, const int64_t *off, const uint2 *cnt, int64_t ntiles,
                               const int64_t *doff, uint64_t *dst, hipStream_t s)
{
    if (ntiles <= 0) return hipSuccess;
    if (ntiles > INT32_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gather_tiles, dim3((unsigned)ntiles), dim3(256), 0, s, src, off, cnt, doff, dst);
    return hipGetLastError();
}

hipError_t launch_grid_pairs(const int *grid, const GridSets &S, int64_t nthr, int64_t *count, uint64_t *out,
                             bool write, hipStream_t s)
{
    if (nthr <= 0) return hipSuccess;
    if (write)
        hipLaunchKernelGGL(k_grid_pairs<true>, dim3(g4d_blocks(nthr, 256, 65536)), dim3(256), 0, s, grid, S, nthr, count,
                           out);
    else
        hipLaunchKernelGGL(k_grid_pairs<false>, dim3(g4d_blocks(nthr, 256, 65536)), dim3(256), 0, s, grid, S, nthr, count,
                           out);
    return hipGetLastError();
}

hipError_t launch_csr_offsets(const uint64_t *keys, int64_t m, int n, int64_t *off, hipStream_t s)
{
    hipLaunchKernelGGL(k_csr_offsets, dim3(g4d_blocks((int64_t)n + 1, 256, 65536)), dim3(256), 0, s, keys, m, n, off);
    return hipGetLastError();
}

hipError_t launch_pump(const uint64_t *keys, const int64_t *off, int n, const GridSets &S, int min_ws, uint64_t *out,
                       unsigned long long *counter, int64_t cap, hipStream_t s)
{
    hipLaunchKernelGGL(k_pump, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, keys, off, n, S, min_ws, out, counter,
                       cap);
    return hipGetLastError();
}

} // namespace g4d
