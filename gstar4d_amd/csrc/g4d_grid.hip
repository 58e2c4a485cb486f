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

__device__ __forceinline__ uint32_t spread10(uint32_t v)
{
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
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

__global__ void k_morton_keys(const double *xyz, int n, int g, const double *bounds, uint64_t *keys)
{
    const double mn = bounds[0], range = bounds[1] - bounds[0], span = (double)(g - 3);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int x = voxel_of(xyz[3 * (int64_t)i], mn, range, span, g);
        const int y = voxel_of(xyz[3 * (int64_t)i + 1], mn, range, span, g);
        const int z = voxel_of(xyz[3 * (int64_t)i + 2], mn, range, span, g);
        const uint32_t code = spread10(x) | (spread10(y) << 1) | (spread10(z) << 2);
        keys[i] = ((uint64_t)code << 32) | (uint32_t)i;
    }
}

__global__ void k_gather_sorted(const double *xyz, const uint64_t *keys, int n, int g, const double *bounds,
                                double *pts, int *perm, int *vox)
{
    const double mn = bounds[0], range = bounds[1] - bounds[0], span = (double)(g - 3);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int i = (int)(uint32_t)keys[k];
        const double x = xyz[3 * (int64_t)i], y = xyz[3 * (int64_t)i + 1], z = xyz[3 * (int64_t)i + 2];
        pts[3 * (int64_t)k] = x;
        pts[3 * (int64_t)k + 1] = y;
        pts[3 * (int64_t)k + 2] = z;
        perm[k] = i;
        vox[k] = G4D_ENC(voxel_of(x, mn, range, span, g), voxel_of(y, mn, range, span, g),
                         voxel_of(z, mn, range, span, g));
    }
}

// Equal points share a voxel, and equal voxels are adjacent in Morton order.
__global__ void k_dup_check(const double *pts, const int *vox, int n, unsigned long long *cnt)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        for (int j = k - 1; j >= 0 && vox[j] == vox[k]; j--) {
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
// leader; the reference picks a racy last writer (GD/PBA/pba3DKernel.h:518-542).
__global__ void k_scatter_sites(const int *vox, int n, int g, int *grid, int *idgrid)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int c = vox[k];
        if (k > 0 && vox[k - 1] == c) continue;
        const int64_t idx = ((int64_t)G4D_DEC_Z(c) * g + G4D_DEC_Y(c)) * g + G4D_DEC_X(c);
        grid[idx] = c;
        idgrid[idx] = k;
    }
}

__device__ __forceinline__ bool env_dominated(int qu, int64_t du, int qv, int64_t dv, int qw, int64_t dw)
{
    const int64_t n_uv = ((int64_t)qv * qv + dv) - ((int64_t)qu * qu + du);
    const int64_t d_uv = qv - qu;
    const int64_t n_vw = ((int64_t)qw * qw + dw) - ((int64_t)qv * qv + dv);
    const int64_t d_vw = qw - qv;
    return n_uv * d_vw >= n_vw * d_uv;
}

// Offset of a site from the column line (squared distance in the two non-sweep axes) and
// its position along the sweep axis.
template <int AXIS>
__device__ __forceinline__ void site_pos_off(int s, int ca, int cb, int *q, int64_t *d)
{
    const int sx = G4D_DEC_X(s), sy = G4D_DEC_Y(s), sz = G4D_DEC_Z(s);
    if (AXIS == 2) { // sweep z; column (x=ca, y=cb)
        *q = sz;
        const int64_t dx = ca - sx, dy = cb - sy;
        *d = dx * dx + dy * dy;
    } else if (AXIS == 1) { // sweep y; column (x=ca, z=cb)
        *q = sy;
        const int64_t dx = ca - sx, dz = cb - sz;
        *d = dx * dx + dz * dz;
    } else { // sweep x; column (y=ca, z=cb)
        *q = sx;
        const int64_t dy = ca - sy, dz = cb - sz;
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

template <int AXIS>
__global__ void __launch_bounds__(64) k_pba_pass(const int *in, int *out, int g, int tx, const int *idgrid)
{
    extern __shared__ int lds[];
    int *col = lds;            // [g][tx] candidates, then the in-place stack
    int *res = lds + g * tx;   // [g][tx] coloured output
    const int ncb = g / tx;
    const int fixed = blockIdx.x / ncb;
    const int c0 = (blockIdx.x % ncb) * tx;
    const int total = g * tx;
    // coalesced load
    for (int i = threadIdx.x; i < total; i += 64) {
        int c, q;
        if (AXIS == 0) { c = i / g; q = i % g; } else { q = i / tx; c = i % tx; }
        col[q * tx + c] = in[pba_gidx<AXIS>(g, fixed, c0, c, q)];
    }
    __syncthreads();
    if (threadIdx.x < tx) {
        const int c = threadIdx.x;
        const int ca = (AXIS == 2 || AXIS == 1) ? c0 + c : c0 + c; // x for Z/Y passes, y for X pass
        const int cb = fixed;                                       // y (Z), z (Y), z (X)
        int m = 0;
        for (int q = 0; q < g; q++) {
            const int s = col[q * tx + c];
            if (s == G4D_MARKER) continue;
            int qs;
            int64_t ds;
            site_pos_off<AXIS>(s, ca, cb, &qs, &ds);
            while (m >= 2) {
                int qu, qv;
                int64_t du, dv;
                site_pos_off<AXIS>(col[(m - 2) * tx + c], ca, cb, &qu, &du);
                site_pos_off<AXIS>(col[(m - 1) * tx + c], ca, cb, &qv, &dv);
                if (env_dominated(qu, du, qv, dv, qs, ds)) m--;
                else break;
            }
            col[m * tx + c] = s;
            m++;
        }
        int idx = 0;
        int qa = 0, qb = 0;
        int64_t da = 0, db = 0;
        if (m > 0) site_pos_off<AXIS>(col[c], ca, cb, &qa, &da);
        if (m > 1) site_pos_off<AXIS>(col[tx + c], ca, cb, &qb, &db);
        for (int t = 0; t < g; t++) {
            if (m == 0) {
                res[t * tx + c] = G4D_MARKER;
                continue;
            }
            while (idx + 1 < m) {
                const int64_t fa = (int64_t)(t - qa) * (t - qa) + da;
                const int64_t fb = (int64_t)(t - qb) * (t - qb) + db;
                if (fb <= fa) {
                    idx++;
                    qa = qb;
                    da = db;
                    if (idx + 1 < m) site_pos_off<AXIS>(col[(idx + 1) * tx + c], ca, cb, &qb, &db);
                } else {
                    break;
                }
            }
            res[t * tx + c] = col[idx * tx + c];
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < total; i += 64) {
        int c, q;
        if (AXIS == 0) { c = i / g; q = i % g; } else { q = i / tx; c = i % tx; }
        int v = res[q * tx + c];
        if (idgrid) {
            v = (v == G4D_MARKER) ? -1
                                  : idgrid[((int64_t)G4D_DEC_Z(v) * g + G4D_DEC_Y(v)) * g + G4D_DEC_X(v)];
        }
        out[pba_gidx<AXIS>(g, fixed, c0, c, q)] = v;
    }
}

int pba_tile_columns(int g)
{
    int tx = 8192 / g; // two int buffers of g*tx in 64 KiB of LDS
    if (tx > 64) tx = 64;
    if (tx < 1) tx = 1;
    while (g % tx) tx--;
    return tx;
}

hipError_t pba_pass(int axis, int g, const int *in, int *out, const int *idgrid, hipStream_t s)
{
    const int tx = pba_tile_columns(g);
    const size_t lds = sizeof(int) * 2 * (size_t)g * tx;
    const int nblk = g * (g / tx);
    if (axis == 2)
        hipLaunchKernelGGL(k_pba_pass<2>, dim3(nblk), dim3(64), lds, s, in, out, g, tx, idgrid);
    else if (axis == 1)
        hipLaunchKernelGGL(k_pba_pass<1>, dim3(nblk), dim3(64), lds, s, in, out, g, tx, idgrid);
    else
        hipLaunchKernelGGL(k_pba_pass<0>, dim3(nblk), dim3(64), lds, s, in, out, g, tx, idgrid);
    return hipGetLastError();
}

hipError_t pba_run(int g, int *grid, int *tmp, const int *idgrid, hipStream_t s)
{
    HIP_CHECK_RET(pba_pass(2, g, grid, tmp, nullptr, s));
    HIP_CHECK_RET(pba_pass(1, g, tmp, grid, nullptr, s));
    return pba_pass(0, g, grid, tmp, idgrid, s);
}

// ------------------------------------------------------------------------------------
// Missing points and Voronoi-neighbour pairs.  Keys are directed (star<<32 | vertex); both
// directions are emitted (the reference reciprocates at GD/GDelaunay/GDelHost.cu:380-386).
// ------------------------------------------------------------------------------------

__device__ __forceinline__ void emit2(uint64_t *out, unsigned long long *counter, int64_t cap, bool pred, int a, int b)
{
    const long long slot = g4d_wave_append(pred, counter);
    if (pred) {
        const int64_t k = 2 * (int64_t)slot;
        if (k + 1 < cap) {
            out[k] = ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
            out[k + 1] = ((uint64_t)(uint32_t)b << 32) | (uint32_t)a;
        }
    }
}

__global__ void k_missing(const int *vox, int n, int g, const int *grid, uint64_t *out, unsigned long long *counter,
                          int64_t cap)
{
    for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const int k = base + threadIdx.x;
        bool miss = false;
        int leader = 0;
        if (k < n) {
            const int c = vox[k];
            leader = grid[((int64_t)G4D_DEC_Z(c) * g + G4D_DEC_Y(c)) * g + G4D_DEC_X(c)];
            miss = leader != k;
        }
        emit2(out, counter, cap, miss, k, leader);
    }
}

__device__ __forceinline__ int grid_at(const int *grid, int g, int x, int y, int z)
{
    if (x < 0 || y < 0 || z < 0 || x >= g || y >= g || z >= g) return G4D_MARKER;
    return grid[((int64_t)z * g + y) * g + x];
}

// One thread per (x, y) column of 2x2x2 cubes, sweeping z (GD/GDelaunay/GDelKernels.cu:181-290).
// Corner c = ix + 2 iy + 4 iz; root = corner 0, opp = corner 7; link cycle 6,2,3,1,5,4.
__global__ void __launch_bounds__(256) k_grid_pairs(const int *grid, int g, uint64_t *out, unsigned long long *counter,
                                                    int64_t cap)
{
    const int w = g + 1;
    const int64_t ncol = (int64_t)w * w;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < ncol; base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t col = base + threadIdx.x;
        const bool active = col < ncol;
        const int x = active ? (int)(col % w) - 1 : -1;
        const int y = active ? (int)(col / w) - 1 : -1;
        int v[8];
        v[0] = active ? grid_at(grid, g, x, y, -1) : G4D_MARKER;
        v[1] = active ? grid_at(grid, g, x + 1, y, -1) : G4D_MARKER;
        v[2] = active ? grid_at(grid, g, x, y + 1, -1) : G4D_MARKER;
        v[3] = active ? grid_at(grid, g, x + 1, y + 1, -1) : G4D_MARKER;
        for (int z = -1; z < g; z++) {
            v[4] = active ? grid_at(grid, g, x, y, z + 1) : G4D_MARKER;
            v[5] = active ? grid_at(grid, g, x + 1, y, z + 1) : G4D_MARKER;
            v[6] = active ? grid_at(grid, g, x, y + 1, z + 1) : G4D_MARKER;
            v[7] = active ? grid_at(grid, g, x + 1, y + 1, z + 1) : G4D_MARKER;
            const int root = v[0], opp = v[7];
            const bool differ = active && root != opp;
            const bool has_marker = (root == G4D_MARKER) || (opp == G4D_MARKER);
            bool has_quad = false;
            const int link[7] = {v[6], v[2], v[3], v[1], v[5], v[4], v[6]};
#pragma unroll
            for (int i = 0; i < 6; i++) {
                const int a = link[i], b = link[i + 1];
                const bool quad = differ && a != b && a != root && a != opp && b != root && b != opp;
                has_quad |= quad;
                // grabPair skips a pair whose first value is the Marker (:155-176)
                emit2(out, counter, cap, quad && root != G4D_MARKER, root, a);
                emit2(out, counter, cap, quad && opp != G4D_MARKER, opp, a);
                emit2(out, counter, cap, quad && root != G4D_MARKER, root, b);
                emit2(out, counter, cap, quad && opp != G4D_MARKER, opp, b);
                emit2(out, counter, cap, quad, a, b);
            }
            emit2(out, counter, cap, has_quad && !has_marker, root, opp);
            v[0] = v[4];
            v[1] = v[5];
            v[2] = v[6];
            v[3] = v[7];
        }
    }
}

// CSR offsets of sorted directed keys: off[s] = first index with star >= s (off has n+1).
__global__ void k_csr_offsets(const uint64_t *keys, int64_t m, int n, int64_t *off)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= m; i += (int64_t)gridDim.x * blockDim.x) {
        const int prev = (i == 0) ? -1 : (int)(keys[i - 1] >> 32);
        const int cur = (i == m) ? n : (int)(keys[i] >> 32);
        for (int s = prev + 1; s <= cur; s++) off[s] = i;
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
__global__ void k_pump(const uint64_t *keys, const int64_t *off, int n, int min_ws, uint64_t *out,
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
        for (int cv = 0; ng < need && cv < n; cv++) {
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

hipError_t launch_morton_keys(const double *xyz, int n, int g, const double *bounds, uint64_t *keys, hipStream_t s)
{
    hipLaunchKernelGGL(k_morton_keys, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, xyz, n, g, bounds, keys);
    return hipGetLastError();
}

hipError_t launch_gather_sorted(const double *xyz, const uint64_t *keys, int n, int g, const double *bounds,
                                double *pts, int *perm, int *vox, hipStream_t s)
{
    hipLaunchKernelGGL(k_gather_sorted, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, xyz, keys, n, g, bounds, pts,
                       perm, vox);
    return hipGetLastError();
}

hipError_t launch_dup_check(const double *pts, const int *vox, int n, unsigned long long *cnt, hipStream_t s)
{
    hipLaunchKernelGGL(k_dup_check, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, pts, vox, n, cnt);
    return hipGetLastError();
}

hipError_t launch_fill_i32(int *a, int64_t n, int v, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_i32, dim3(g4d_blocks(n / 4 + 1, 256, 65536)), dim3(256), 0, s, a, n, v);
    return hipGetLastError();
}

hipError_t launch_scatter_sites(const int *vox, int n, int g, int *grid, int *idgrid, hipStream_t s)
{
    hipLaunchKernelGGL(k_scatter_sites, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, vox, n, g, grid, idgrid);
    return hipGetLastError();
}

hipError_t launch_missing(const int *vox, int n, int g, const int *grid, uint64_t *out, unsigned long long *counter,
                          int64_t cap, hipStream_t s)
{
    hipLaunchKernelGGL(k_missing, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, vox, n, g, grid, out, counter, cap);
    return hipGetLastError();
}

hipError_t launch_grid_pairs(const int *grid, int g, uint64_t *out, unsigned long long *counter, int64_t cap,
                             hipStream_t s)
{
    const int64_t ncol = (int64_t)(g + 1) * (g + 1);
    hipLaunchKernelGGL(k_grid_pairs, dim3(g4d_blocks(ncol, 256, 65536)), dim3(256), 0, s, grid, g, out, counter, cap);
    return hipGetLastError();
}

hipError_t launch_csr_offsets(const uint64_t *keys, int64_t m, int n, int64_t *off, hipStream_t s)
{
    hipLaunchKernelGGL(k_csr_offsets, dim3(g4d_blocks(m + 1, 256, 65536)), dim3(256), 0, s, keys, m, n, off);
    return hipGetLastError();
}

hipError_t launch_pump(const uint64_t *keys, const int64_t *off, int n, int min_ws, uint64_t *out,
                       unsigned long long *counter, int64_t cap, hipStream_t s)
{
    hipLaunchKernelGGL(k_pump, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, keys, off, n, min_ws, out, counter,
                       cap);
    return hipGetLastError();
}

} // namespace g4d
