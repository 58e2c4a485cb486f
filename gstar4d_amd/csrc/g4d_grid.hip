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

// Coordinates are below 2048, so squared distances stay below 2^23 and the numerators below
// 2^24: 32-bit terms, 64-bit products.
__device__ __forceinline__ bool env_dominated(int qu, int du, int qv, int dv, int qw, int dw)
{
    const int n_uv = (qv * qv + dv) - (qu * qu + du);
    const int d_uv = qv - qu;
    const int n_vw = (qw * qw + dw) - (qv * qv + dv);
    const int d_vw = qw - qv;
    return (int64_t)n_uv * d_vw >= (int64_t)n_vw * d_uv;
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
        *d = dx * dx + dy * dy;
    } else if (AXIS == 1) { // sweep y; column (x=ca, z=cb)
        *q = sy;
        const int dx = ca - sx, dz = cb - sz;
        *d = dx * dx + dz * dz;
    } else { // sweep x; column (y=ca, z=cb)
        *q = sx;
        const int dy = ca - sy, dz = cb - sz;
        *d = dy * dy + dz * dz;
    }
}

// Global index of element q of column c in this workgroup's tile.
template <int AXIS>
__device__ __forceinline__ int64_t pba_gidx(int g, int fixed, int c0, int c, int q)
{
    const int64_t G = g;
    if (AXIS == 2) return ((int64_t)q * G + fixed) * G + c0 + c;   // z=q, y=fixed, x=c0+c
    if (AXIS == 1) return ((int64_t)fixed * G + q) * G + c0 + c;   // z=fixed, y=q, x=c0+c
    return ((int64_t)fixed * G + c0 + c) * G + q;                  // z=fixed, y=c0+c, x=q
}

// One wave per tile of 64 columns along the sweep axis, one lane per column.  The column is
// streamed through an LDS tile of 64 positions x 64 columns (coalesced loads and stores for
// every axis), the lower envelope of the candidate sites (Maurer stack, exact integer
// dominance test) is kept per lane in LDS -- entries beyond PBA_STACK spill to a per-column
// region of `spill` -- and the colour walk writes the nearest site of every voxel back through
// the same tile.  Ties: a later site along the axis wins (colour walk advances on <=).
static constexpr int PBA_COLS = 64;
static constexpr int PBA_CHUNK = 64;
static constexpr int PBA_STACK = 64;

struct PbaLds {
    int tile[PBA_CHUNK][PBA_COLS + 1];
    int scode[PBA_STACK][PBA_COLS];
};

template <int AXIS>
__device__ __forceinline__ int64_t pba_gidx64(int g, int fixed, int c, int q)
{
    const int64_t G = g;
    if (AXIS == 2) return ((int64_t)q * G + fixed) * G + c;   // z = q, y = fixed, x = c
    if (AXIS == 1) return ((int64_t)fixed * G + q) * G + c;   // z = fixed, y = q, x = c
    return ((int64_t)fixed * G + c) * G + q;                  // z = fixed, y = c, x = q
}

template <int AXIS, bool IDS>
__global__ void __launch_bounds__(64) k_pba_pass(const int *in_all, int *out_all, GridSets S, const int *idgrid_all,
                                                 int *spill_all, const uint64_t *vox)
{
    __shared__ PbaLds L;
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
    int *sp = spill_all + S.spill[st] + (lb * PBA_COLS + lane) * g;

    // ---- lower envelope ----
    int m = 0;
    for (int q0 = 0; q0 < g; q0 += PBA_CHUNK) {
        const int nq = min(PBA_CHUNK, g - q0);
        for (int i = lane; i < PBA_CHUNK * PBA_COLS; i += 64) {
            int cc, qq;
            if (AXIS == 0) { cc = i / PBA_CHUNK; qq = i % PBA_CHUNK; } else { qq = i / PBA_COLS; cc = i % PBA_COLS; }
            if (qq < nq && cc < ncols) L.tile[qq][cc] = in[pba_gidx64<AXIS>(g, fixed, c0 + cc, q0 + qq)];
        }
        __syncthreads();
        if (active) {
            for (int qq = 0; qq < nq; qq++) {
                const int s = L.tile[qq][lane];
                if (s == G4D_MARKER) continue;
                int qs, ds;
                site_pos_off<AXIS, IDS>(s, vox, ca, cb, &qs, &ds);
                while (m >= 2) {
                    int qu, qv, du, dv;
                    const int su = m - 2 < PBA_STACK ? L.scode[m - 2][lane] : sp[m - 2];
                    const int sv = m - 1 < PBA_STACK ? L.scode[m - 1][lane] : sp[m - 1];
                    site_pos_off<AXIS, IDS>(su, vox, ca, cb, &qu, &du);
                    site_pos_off<AXIS, IDS>(sv, vox, ca, cb, &qv, &dv);
                    if (env_dominated(qu, du, qv, dv, qs, ds)) m--;
                    else break;
                }
                if (m < PBA_STACK) L.scode[m][lane] = s;
                else sp[m] = s;
                m++;
            }
        }
        __syncthreads();
    }
    // ---- colour walk ----
    int idx = 0, qa = 0, qb = 0, sa = G4D_MARKER;
    int da = 0, db = 0;
    if (active && m > 0) {
        sa = L.scode[0][lane];
        site_pos_off<AXIS, IDS>(sa, vox, ca, cb, &qa, &da);
    }
    if (active && m > 1) {
        const int sb = 1 < PBA_STACK ? L.scode[1][lane] : sp[1];
        site_pos_off<AXIS, IDS>(sb, vox, ca, cb, &qb, &db);
    }
    for (int t0 = 0; t0 < g; t0 += PBA_CHUNK) {
        const int nt = min(PBA_CHUNK, g - t0);
        if (active) {
            for (int tt = 0; tt < nt; tt++) {
                const int t = t0 + tt;
                if (m == 0) {
                    L.tile[tt][lane] = G4D_MARKER;
                    continue;
                }
                while (idx + 1 < m) {
                    const int fa = (t - qa) * (t - qa) + da;
                    const int fb = (t - qb) * (t - qb) + db;
                    if (fb <= fa) {
                        idx++;
                        qa = qb;
                        da = db;
                        sa = idx < PBA_STACK ? L.scode[idx][lane] : sp[idx];
                        if (idx + 1 < m) {
                            const int sb = idx + 1 < PBA_STACK ? L.scode[idx + 1][lane] : sp[idx + 1];
                            site_pos_off<AXIS, IDS>(sb, vox, ca, cb, &qb, &db);
                        }
                    } else {
                        break;
                    }
                }
                L.tile[tt][lane] = sa;
            }
        }
        __syncthreads();
        for (int i = lane; i < PBA_CHUNK * PBA_COLS; i += 64) {
            int cc, qq;
            if (AXIS == 0) { cc = i / PBA_CHUNK; qq = i % PBA_CHUNK; } else { qq = i / PBA_COLS; cc = i % PBA_COLS; }
            if (qq < nt && cc < ncols) {
                int v = L.tile[qq][cc];
                if (!IDS && idgrid)
                    v = (v == G4D_MARKER) ? -1
                                          : idgrid[((int64_t)G4D_DEC_Z(v) * g + G4D_DEC_Y(v)) * g + G4D_DEC_X(v)];
                out[pba_gidx64<AXIS>(g, fixed, c0 + cc, t0 + qq)] = v;
            }
        }
        __syncthreads();
    }
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

// One thread per (x, y) column of cubes and a chunk of zc cubes along z; pass 0 counts the
// thread's pairs, pass 1 writes them at the scanned offset (deterministic order, no atomics).
template <bool WRITE>
__global__ void __launch_bounds__(256) k_grid_pairs(const int *grid_all, GridSets S, int64_t nthr, int64_t *count,
                                                    uint64_t *out)
{
    for (int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; tid < nthr; tid += (int64_t)gridDim.x * blockDim.x) {
        const int st = S.nsets == 1 ? 0 : set_of_work(S.pair_thr, S.nsets, tid);
        const int64_t lt = tid - S.pair_thr[st];
        const int g = S.g[st];
        const int *grid = grid_all + S.vox_off[st];
        const int zc = grid_pair_chunk_dev(g);
        const int w = g + 1;
        const int64_t ncol = (int64_t)w * w;
        const int64_t col = lt % ncol;
        const int zb = (int)(lt / ncol);
        const int x = (int)(col % w) - 1, y = (int)(col / w) - 1;
        const int z0 = zb * zc - 1, z1 = min(z0 + zc, g);
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
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        const int64_t b = off[s], e = off[s + 1];
        const int have = (int)(e - b);
        if (have >= min_ws) continue;
        const int need = min_ws - have;
        int got[64];
        int ng = 0;
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
        for (int r = 0; r < ng; r++) {
            const unsigned long long slot = atomicAdd(counter, 1ull);
            const int64_t k = 2 * (int64_t)slot;
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
    const int64_t w = g + 1;
    return w * w * ((w + zc - 1) / zc);
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
