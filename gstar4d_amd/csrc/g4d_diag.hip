// gStar4D for MI355X — kernel-level diagnostic entry points (predicate parity checks).
#include "g4d_diag.h"
#include "g4d_predicates.h"

namespace g4d {

__global__ void k_diag_orient4(const double *pts5, const int *idx5, int64_t T, int mode, int *sign, int *depth,
                               unsigned long long *cnt)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < T; t += (int64_t)gridDim.x * blockDim.x) {
        const double *p[5];
        int idx[5];
        for (int i = 0; i < 5; i++) {
            p[i] = pts5 + 15 * t + 3 * i;
            idx[i] = idx5[5 * t + i];
        }
        int s = 0, d = 0;
        if (mode == 0) {
            s = filter_insphere(p);
            if (s == 0) {
                s = exact_insphere(p, cnt);
                if (s == 0) s = sos_orient4(p, idx, &d, cnt);
            }
        } else if (mode == 1) {
            s = exact_insphere(p, cnt);
            if (s == 0) s = sos_orient4(p, idx, &d, cnt);
        } else if (mode == 2) {
            // the star kernel's cached-cofactor filter for orient4(a, b, c, s, v)
            const double sx = p[3][0], sy = p[3][1], sz = p[3][2];
            double ra[4], rb[4], rc[4], rv[4], C[4], P[4];
            for (int i = 0; i < 3; i++) {
                ra[i] = p[0][i] - p[3][i];
                rb[i] = p[1][i] - p[3][i];
                rc[i] = p[2][i] - p[3][i];
                rv[i] = p[4][i] - p[3][i];
            }
            (void)sx; (void)sy; (void)sz;
            ra[3] = ra[0] * ra[0] + ra[1] * ra[1] + ra[2] * ra[2];
            rb[3] = rb[0] * rb[0] + rb[1] * rb[1] + rb[2] * rb[2];
            rc[3] = rc[0] * rc[0] + rc[1] * rc[1] + rc[2] * rc[2];
            rv[3] = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
            tri_cofactors(ra, rb, rc, C, P);
            s = cofactor_filter_sign(rv, C, P);
        } else {
            s = filter_insphere(p);
        }
        sign[t] = s;
        if (depth) depth[t] = d;
    }
}

__global__ void k_diag_orient3d(const double *pts4, int64_t T, int *sign, unsigned long long *cnt)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < T; t += (int64_t)gridDim.x * blockDim.x) {
        const double *p[4];
        for (int i = 0; i < 4; i++) p[i] = pts4 + 12 * t + 3 * i;
        sign[t] = orient3d_full(p, cnt);
    }
}

hipError_t launch_diag_orient4(const double *pts5, const int *idx5, int64_t T, int mode, int *sign, int *depth,
                               unsigned long long *cnt, hipStream_t s)
{
    if (T <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_diag_orient4, dim3(g4d_blocks(T, 64, 65536)), dim3(64), 0, s, pts5, idx5, T, mode, sign,
                       depth, cnt);
    return hipGetLastError();
}

hipError_t launch_diag_orient3d(const double *pts4, int64_t T, int *sign, unsigned long long *cnt, hipStream_t s)
{
    if (T <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_diag_orient3d, dim3(g4d_blocks(T, 64, 65536)), dim3(64), 0, s, pts4, T, sign, cnt);
    return hipGetLastError();
}

__global__ void k_diag_copy(const int4 *in, int4 *out, int64_t n16)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

hipError_t launch_diag_copy(const int4 *in, int4 *out, int64_t n16, hipStream_t s)
{
    if (n16 <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_diag_copy, dim3(g4d_blocks(n16, 256, 8192)), dim3(256), 0, s, in, out, n16);
    return hipGetLastError();
}

} // namespace g4d
