// gStar4D for MI355X — mesh validation on the GPU.
//
// Restates TetraMesh::check (GDelaunay/Common/Geometry.cu:200-393): Euler characteristic
// V - E + F - T over distinct vertices, edges and faces (:211-317), orientation (:319-345,
// a tetrahedron fails when orient3d > 0) and the adjacent-pair in-sphere test (:347-393, the
// vertex across each shared face must not lie strictly inside the tetrahedron's sphere).
// Added: flat tetrahedra, adjacency verification against an independent face hash, and a
// full empty-sphere test of every point near each tetrahedron's circumsphere.
//
// Every kernel here evaluates exact predicates inline, so each runs on a grid of at most
// CHECK_MAX_BLOCKS work-groups (grid-stride), which bounds its private scratch footprint.
#include "g4d_check.h"
#include "g4d_predicates.h"

namespace g4d {

static constexpr int CHECK_MAX_BLOCKS = 1024;

// Orientation, flat tetrahedra, adjacency agreement and the adjacent-pair in-sphere test.
// ref_opp: adjacency computed here by face hashing; opp: the adjacency under test.
__global__ void k_check_tets(const double *pts, const int *tets, const int *opp, const int *ref_opp, int64_t nt, int n,
                             unsigned char *used, unsigned long long *acc, unsigned long long *cnt)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += (int64_t)gridDim.x * blockDim.x) {
        const int *q = tets + 4 * t;
        bool vok = true;
        for (int i = 0; i < 4; i++) {
            if (q[i] < 0 || q[i] >= n) vok = false;
            else used[q[i]] = 1;
        }
        if (!vok) {
            atomicAdd(&acc[CHECK_ADJ], 1ull);
            continue;
        }
        const double *p[4] = {pts + 3 * (int64_t)q[0], pts + 3 * (int64_t)q[1], pts + 3 * (int64_t)q[2],
                              pts + 3 * (int64_t)q[3]};
        const int o = orient3d_full(p, cnt);
        if (o > 0) atomicAdd(&acc[CHECK_ORIENT], 1ull);
        if (o == 0) atomicAdd(&acc[CHECK_FLAT], 1ull);
        for (int i = 0; i < 4; i++) {
            const int r = ref_opp[4 * t + i];
            if (r < 0) atomicAdd(&acc[CHECK_HULL], 1ull);
            else atomicAdd(&acc[CHECK_SHARED], 1ull);
            if (opp && opp[4 * t + i] != r) atomicAdd(&acc[CHECK_ADJ], 1ull);
            const int u = opp ? opp[4 * t + i] : r;
            if (u < 0 || u >= nt || (int64_t)u <= t) continue; // each pair once
            // vertex of u not in t
            int w = -1;
            for (int j = 0; j < 4; j++) {
                const int c = tets[4 * (int64_t)u + j];
                if (c != q[0] && c != q[1] && c != q[2] && c != q[3]) w = c;
            }
            if (w < 0 || w >= n) continue; // counted by the adjacency comparison
            const double *pe[5] = {p[0], p[1], p[2], p[3], pts + 3 * (int64_t)w};
            int s = filter_insphere(pe);
            if (s == 0) s = exact_insphere(pe, cnt);
            if (s * o > 0) atomicAdd(&acc[CHECK_INSPHERE], 1ull);
        }
    }
}

__global__ void k_check_edges(const int *tets, int64_t nt, uint64_t *keys)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += (int64_t)gridDim.x * blockDim.x) {
        const int *q = tets + 4 * t;
        int k = 0;
        for (int i = 0; i < 4; i++)
            for (int j = i + 1; j < 4; j++) {
                const uint32_t a = (uint32_t)q[i], b = (uint32_t)q[j];
                keys[6 * t + k++] = a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a;
            }
    }
}

__global__ void k_check_count_used(const unsigned char *used, int n, unsigned long long *acc)
{
    for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const int i = base + threadIdx.x;
        const bool u = i < n && used[i];
        const long long slot = g4d_wave_append(u, &acc[CHECK_VERTS]);
        (void)slot;
    }
}

// Uniform cell grid over the points' bounding box for the empty-sphere test.
__global__ void k_check_cell_ids(const double *pts, int n, const double *box, int R, int *cid,
                                 unsigned long long *cellcnt)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int ci[3];
        for (int c = 0; c < 3; c++) {
            const double w = (box[3 + c] - box[c]) / R;
            int k = w > 0 ? (int)((pts[3 * (int64_t)i + c] - box[c]) / w) : 0;
            ci[c] = k < 0 ? 0 : (k >= R ? R - 1 : k);
        }
        const int id = (ci[2] * R + ci[1]) * R + ci[0];
        cid[i] = id;
        atomicAdd(&cellcnt[id], 1ull);
    }
}

__global__ void k_check_cell_fill(const int *cid, int n, const int64_t *start, int *fill, int *items)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int id = cid[i];
        const int k = atomicAdd(&fill[id], 1);
        items[start[id] + k] = i;
    }
}

// One wave per tetrahedron: every point in the cells overlapping the (padded) circumsphere's
// bounding box, other than the tetrahedron's vertices, must not be strictly inside.
__global__ void __launch_bounds__(64) k_check_empty(const double *pts, const int *tets, int64_t nt, const double *box,
                                                    int R, const int64_t *start, const int *items,
                                                    unsigned long long *acc, unsigned long long *cnt)
{
    const int lane = threadIdx.x;
    for (int64_t t = blockIdx.x; t < nt; t += gridDim.x) {
        const int *q = tets + 4 * t;
        const double *p[4] = {pts + 3 * (int64_t)q[0], pts + 3 * (int64_t)q[1], pts + 3 * (int64_t)q[2],
                              pts + 3 * (int64_t)q[3]};
        double a[3][3], b[3];
        for (int i = 0; i < 3; i++) {
            double s2 = 0;
            for (int c = 0; c < 3; c++) {
                a[i][c] = p[i + 1][c] - p[0][c];
                s2 += a[i][c] * a[i][c];
            }
            b[i] = 0.5 * s2;
        }
        const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                           a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                           a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
        int lo[3] = {0, 0, 0}, hi[3] = {R - 1, R - 1, R - 1};
        if (fabs(det) > 1e-300) {
            // circumcentre c = p0 + A^-1 b (approximate; the box is padded by a cell)
            double x[3];
            for (int k = 0; k < 3; k++) {
                double m[3][3];
                for (int i = 0; i < 3; i++)
                    for (int c = 0; c < 3; c++) m[i][c] = (c == k) ? b[i] : a[i][c];
                x[k] = (m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                        m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                        m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0])) /
                       det;
            }
            const double rad = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
            for (int c = 0; c < 3; c++) {
                const double w = (box[3 + c] - box[c]) / R;
                const double cc = p[0][c] + x[c];
                if (w > 0 && isfinite(cc) && isfinite(rad)) {
                    const double l = floor((cc - rad - box[c]) / w) - 1, h = floor((cc + rad - box[c]) / w) + 1;
                    lo[c] = l < 0 ? 0 : (l > R - 1 ? R - 1 : (int)l);
                    hi[c] = h < 0 ? 0 : (h > R - 1 ? R - 1 : (int)h);
                }
            }
        }
        int o = 0;
        if (lane == 0) o = orient3d_full(p, cnt);
        o = __shfl(o, 0, 64);
        if (o == 0) continue; // flat: counted elsewhere, no sphere
        unsigned long long tests = 0, bad = 0;
        for (int z = lo[2]; z <= hi[2]; z++)
            for (int y = lo[1]; y <= hi[1]; y++)
                for (int x = lo[0]; x <= hi[0]; x++) {
                    const int id = (z * R + y) * R + x;
                    for (int64_t k = start[id] + lane; k < start[id + 1]; k += 64) {
                        const int v = items[k];
                        if (v == q[0] || v == q[1] || v == q[2] || v == q[3]) continue;
                        const double *pe[5] = {p[0], p[1], p[2], p[3], pts + 3 * (int64_t)v};
                        int s = filter_insphere(pe);
                        if (s == 0) s = exact_insphere(pe, cnt);
                        tests++;
                        if (s * o > 0) bad++;
                    }
                }
        if (tests) atomicAdd(&acc[CHECK_EMPTY_TESTS], tests);
        if (bad) atomicAdd(&acc[CHECK_EMPTY], bad);
    }
}

static int grid_of(int64_t n, int per) { return g4d_blocks(n, per, CHECK_MAX_BLOCKS); }

hipError_t launch_check_tets(const double *pts, const int *tets, const int *opp, const int *ref_opp, int64_t nt, int n,
                             unsigned char *used, unsigned long long *acc, unsigned long long *cnt, hipStream_t s)
{
    if (nt <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_check_tets, dim3(grid_of(nt, 64)), dim3(64), 0, s, pts, tets, opp, ref_opp, nt, n, used, acc,
                       cnt);
    return hipGetLastError();
}

hipError_t launch_check_edges(const int *tets, int64_t nt, uint64_t *keys, hipStream_t s)
{
    if (nt <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_check_edges, dim3(g4d_blocks(nt, 256, 65536)), dim3(256), 0, s, tets, nt, keys);
    return hipGetLastError();
}

hipError_t launch_check_count_used(const unsigned char *used, int n, unsigned long long *acc, hipStream_t s)
{
    hipLaunchKernelGGL(k_check_count_used, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, used, n, acc);
    return hipGetLastError();
}

hipError_t launch_check_cells(const double *pts, int n, const double *box, int R, int *cid,
                              unsigned long long *cellcnt, hipStream_t s)
{
    hipLaunchKernelGGL(k_check_cell_ids, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, pts, n, box, R, cid,
                       cellcnt);
    return hipGetLastError();
}

hipError_t launch_check_cell_fill(const int *cid, int n, const int64_t *start, int *fill, int *items, hipStream_t s)
{
    hipLaunchKernelGGL(k_check_cell_fill, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, cid, n, start, fill, items);
    return hipGetLastError();
}

hipError_t launch_check_empty(const double *pts, const int *tets, int64_t nt, const double *box, int R,
                              const int64_t *start, const int *items, unsigned long long *acc, unsigned long long *cnt,
                              hipStream_t s)
{
    if (nt <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_check_empty, dim3((unsigned)(nt < CHECK_MAX_BLOCKS ? nt : CHECK_MAX_BLOCKS)), dim3(64), 0, s,
                       pts, tets, nt, box, R, start, items, acc, cnt);
    return hipGetLastError();
}

} // namespace g4d
