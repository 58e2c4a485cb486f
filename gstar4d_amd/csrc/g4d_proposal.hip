// gStar4D for MI355X — learned proposal kernels (candidates, features, MLP inference and
// training on the 16x16x32 BF16 MFMA).
//
// MFMA operand layout (v_mfma_f32_16x16x32_bf16, 64 lanes): lane l holds A[l % 16][8 (l / 16)
// + 0..7] and B[8 (l / 16) + 0..7][l % 16] (8 BF16 each) and receives D[4 (l / 16) + 0..3]
// [l % 16] (4 FP32).  Activations, weights and gradients are staged in LDS as BF16 row-major
// tiles; every operand is gathered from LDS in that layout.
#include "g4d_proposal.h"

namespace g4d {

// ------------------------------------------------------------------------------------
// Candidates
// ------------------------------------------------------------------------------------

__device__ __forceinline__ int prop_set(const GridSets &S, int64_t i) { return S.pt_set ? S.pt_set[i] : 0; }

__device__ __forceinline__ uint64_t prop_spread21(uint64_t v)
{
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__global__ void k_fine_keys(const double *pts, int n, GridSets S, int cb, int pb, int shifted, uint64_t *keys)
{
    const uint64_t cells = 1ull << cb;
    const uint64_t sh = shifted ? cells / 3 : 0;
    const double span = (double)(cells - 1 - sh);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int st = prop_set(S, i);
        const double lo = S.bounds[2 * st], range = S.bounds[2 * st + 1] - lo;
        uint64_t code = 0;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double t = (pts[3 * (int64_t)i + c] - lo) / range * span;
            t = t < 0.0 ? 0.0 : (t > span ? span : t);
            code |= prop_spread21((uint64_t)t + sh) << c;
        }
        keys[i] = ((uint64_t)st << (3 * cb + pb)) | (code << pb) | (uint64_t)i;
    }
}

__global__ void k_fine_ranks(const uint64_t *keys, int n, uint64_t mask, int *ord, int *rank)
{
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        const int i = (int)(keys[r] & mask);
        ord[r] = i;
        rank[i] = r;
    }
}

__global__ void __launch_bounds__(256) k_candidates(const double *pts, int n, GridSets S, const int *ord0,
                                                    const int *rank0, const int *ord1, const int *rank1, int *cand,
                                                    float *d2out)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int st = prop_set(S, i);
        const int64_t lo = S.pt_off[st], hi = S.pt_off[st + 1];
        const double px = pts[3 * (int64_t)i], py = pts[3 * (int64_t)i + 1], pz = pts[3 * (int64_t)i + 2];
        float bd[PROP_SLOTS];
        int bc[PROP_SLOTS];
#pragma unroll
        for (int k = 0; k < PROP_SLOTS; k++) {
            bd[k] = INFINITY;
            bc[k] = -1;
        }
        for (int o = 0; o < 2; o++) {
            const int *ord = o ? ord1 : ord0;
            const int r = o ? rank1[i] : rank0[i];
            for (int d = -PROP_WIN; d <= PROP_WIN; d++) {
                const int rr = r + d;
                if (d == 0 || rr < 0 || rr >= n) continue;
                const int j = ord[rr];
                if (j < lo || j >= hi) continue; // another set
                const double dx = pts[3 * (int64_t)j] - px, dy = pts[3 * (int64_t)j + 1] - py,
                             dz = pts[3 * (int64_t)j + 2] - pz;
                const float dd = (float)(dx * dx + dy * dy + dz * dz);
                if (!(dd < bd[PROP_SLOTS - 1])) continue;
                bool dup = false;
#pragma unroll
                for (int k = 0; k < PROP_SLOTS; k++) dup |= bc[k] == j;
                if (dup) continue;
                // insert in ascending order (register resident: fixed trip counts)
                float cd = dd;
                int cc = j;
#pragma unroll
                for (int k = 0; k < PROP_SLOTS; k++) {
                    if (cd < bd[k]) {
                        const float td = bd[k];
                        const int tc = bc[k];
                        bd[k] = cd;
                        bc[k] = cc;
                        cd = td;
                        cc = tc;
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < PROP_SLOTS; k++) {
            cand[PROP_SLOTS * (int64_t)i + k] = bc[k];
            d2out[PROP_SLOTS * (int64_t)i + k] = bd[k];
        }
    }
}

// ------------------------------------------------------------------------------------
// Features
// ------------------------------------------------------------------------------------

__device__ __forceinline__ uint64_t vox_code(uint64_t v)
{
    return prop_spread21(G4D_VOX_X(v)) | (prop_spread21(G4D_VOX_Y(v)) << 1) | (prop_spread21(G4D_VOX_Z(v)) << 2);
}

__global__ void __launch_bounds__(256) k_features(const double *pts, const uint64_t *vox, int n, GridSets S,
                                                  const int *idgrid, const uint64_t *ws_keys, const int64_t *ws_off,
                                                  const int *cand, const float *d2, float *feat)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int st = prop_set(S, i);
        const int64_t lo = S.pt_off[st], hi = S.pt_off[st + 1];
        const int g = S.g[st];
        const uint64_t v = vox[i];
        const int vx = G4D_VOX_X(v), vy = G4D_VOX_Y(v), vz = G4D_VOX_Z(v);
        const int leader = idgrid[S.vox_off[st] + ((int64_t)vz * g + vy) * g + vx];
        // points sharing the voxel: the run of equal Morton codes around i (sorted within the set)
        const uint64_t code = vox_code(v);
        int64_t a = lo, b = i; // first position with code >= code
        while (a < b) {
            const int64_t m = (a + b) >> 1;
            if (vox_code(vox[m]) < code) a = m + 1;
            else b = m;
        }
        const int64_t first = a;
        a = i + 1;
        b = hi;
        while (a < b) {
            const int64_t m = (a + b) >> 1;
            if (vox_code(vox[m]) <= code) a = m + 1;
            else b = m;
        }
        const int64_t occ = a - first;
        const int64_t w0 = ws_off[i], w1 = ws_off[i + 1];
        float msum = 0.f;
        int mcnt = 0;
        for (int64_t k = w0; k < w1 && mcnt < 32; k++, mcnt++) {
            const uint64_t u = vox[(int)(uint32_t)ws_keys[k]];
            const float dx = (float)(G4D_VOX_X(u) - vx), dy = (float)(G4D_VOX_Y(u) - vy), dz = (float)(G4D_VOX_Z(u) - vz);
            msum += sqrtf(dx * dx + dy * dy + dz * dz);
        }
        const double range = S.bounds[2 * st + 1] - S.bounds[2 * st];
        const float vsz = (float)(range / (double)(g - 3)); // voxel edge in input units
        float f[PROP_FEATS];
        f[0] = leader != i ? 1.f : 0.f;
        f[1] = log2f(1.f + (float)occ);
        f[2] = log2f(1.f + (float)(w1 - w0));
        f[3] = mcnt ? msum / (float)mcnt : 0.f;
        const float dn = d2[PROP_SLOTS * (int64_t)i];
        const float d1 = isfinite(dn) ? sqrtf(dn) : 0.f;
        f[4] = d1 > 0.f ? log2f(fmaxf(d1 / vsz, 1e-6f)) : -20.f;
#pragma unroll
        for (int j = 1; j < PROP_SLOTS; j++) {
            const float dj = d2[PROP_SLOTS * (int64_t)i + j];
            f[4 + j] = (isfinite(dj) && d1 > 0.f) ? 0.5f * log2f(dj / dn) : 8.f;
        }
#pragma unroll
        for (int k = 0; k < PROP_FEATS; k++) feat[PROP_FEATS * (int64_t)i + k] = f[k];
    }
}

// ------------------------------------------------------------------------------------
// MLP on the MFMA
// ------------------------------------------------------------------------------------

typedef short short8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u); // round to nearest even (inputs are finite)
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ f32x4 mfma16x16x32(short8 a, short8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}

// 8 consecutive BF16 of a row-major LDS tile (row r, columns c .. c+7); columns at or beyond
// `cols` read as zero (K padding).
__device__ __forceinline__ short8 lds_row8(const uint16_t *t, int ld, int cols, int r, int c)
{
    short8 v;
#pragma unroll
    for (int q = 0; q < 8; q++) v[q] = c + q < cols ? (short)t[r * ld + c + q] : (short)0;
    return v;
}
// 8 BF16 down a column of a row-major LDS tile (rows r .. r+7, column c); rows at or beyond
// `rows` read as zero.
__device__ __forceinline__ short8 lds_col8(const uint16_t *t, int ld, int rows, int r, int c)
{
    short8 v;
#pragma unroll
    for (int q = 0; q < 8; q++) v[q] = r + q < rows ? (short)t[(r + q) * ld + c] : (short)0;
    return v;
}

// Weights in LDS (BF16): transposed copies for the forward B operands, row-major copies for
// the backward ones; biases FP32.
struct MlpW {
    uint16_t w1t[MLP_H][MLP_IN];   // [out][in]
    uint16_t w2t[MLP_H][MLP_H];
    uint16_t w2[MLP_H][MLP_H];     // [in][out]
    uint16_t w3t[MLP_OUT][MLP_H];
    uint16_t w3[MLP_H][MLP_OUT];
    float b1[MLP_H], b2[MLP_H], b3[MLP_OUT];
};

__device__ void stage_weights(const float *p, MlpW &W, bool backward)
{
    for (int e = threadIdx.x; e < MLP_IN * MLP_H; e += blockDim.x) {
        const int i = e / MLP_H, o = e % MLP_H;
        W.w1t[o][i] = f2bf(p[MLP_OFF_W1 + e]);
    }
    for (int e = threadIdx.x; e < MLP_H * MLP_H; e += blockDim.x) {
        const int i = e / MLP_H, o = e % MLP_H;
        const uint16_t w = f2bf(p[MLP_OFF_W2 + e]);
        W.w2t[o][i] = w;
        if (backward) W.w2[i][o] = w;
    }
    for (int e = threadIdx.x; e < MLP_H * MLP_OUT; e += blockDim.x) {
        const int i = e / MLP_OUT, o = e % MLP_OUT;
        const uint16_t w = f2bf(p[MLP_OFF_W3 + e]);
        W.w3t[o][i] = w;
        if (backward) W.w3[i][o] = w;
    }
    for (int e = threadIdx.x; e < MLP_H; e += blockDim.x) {
        W.b1[e] = p[MLP_OFF_B1 + e];
        W.b2[e] = p[MLP_OFF_B2 + e];
    }
    for (int e = threadIdx.x; e < MLP_OUT; e += blockDim.x) W.b3[e] = p[MLP_OFF_B3 + e];
}

// Per-wave tiles of 32 samples.
struct MlpTile {
    uint16_t x[32][MLP_IN];
    uint16_t h1[32][MLP_H];
    uint16_t h2[32][MLP_H];
    uint16_t dz3[32][MLP_OUT];
    uint16_t dz2[32][MLP_H];
    uint16_t dz1[32][MLP_H];
    float z3[32][MLP_OUT];
    uint32_t lab[32];
};

// One dense layer for 16 samples (rows m0 .. m0+15 of `in`): out = act(in . W + b), W given
// transposed as wt[n][k]; FP32 results rounded to BF16 into `out` (ReLU) or kept in FP32
// (`outf`, no activation).
template <int K, int N>
__device__ __forceinline__ void dense16(const uint16_t *in, int ldin, int kcols, const uint16_t *wt, const float *bias,
                                        int m0, uint16_t *out, float *outf, bool relu)
{
    const int l = __lane_id();
    const int r = l & 15, kq = 8 * (l >> 4);
#pragma unroll
    for (int nt = 0; nt < N / 16; nt++) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < K / 32; kt++) {
            const short8 a = lds_row8(in, ldin, kcols, m0 + r, kt * 32 + kq);
            const short8 b = lds_row8(wt, K, K, nt * 16 + r, kt * 32 + kq);
            acc = mfma16x16x32(a, b, acc);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int row = m0 + 4 * (l >> 4) + q, col = nt * 16 + r;
            float v = acc[q] + bias[col];
            if (relu) {
                v = v > 0.f ? v : 0.f;
                out[row * N + col] = f2bf(v);
            } else {
                outf[row * N + col] = v;
            }
        }
    }
}

// Orders one wave's LDS writes before its later LDS reads (a wave's LDS operations complete in
// order; this keeps the compiler from moving them across).
__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Forward pass of the wave's 32 samples: x staged, h1, h2 (BF16) and z3 (FP32 logits) out.
__device__ __forceinline__ void mlp_forward_tile(const MlpW &W, MlpTile &T)
{
    for (int m0 = 0; m0 < 32; m0 += 16)
        dense16<MLP_IN, MLP_H>(&T.x[0][0], MLP_IN, MLP_IN, &W.w1t[0][0], W.b1, m0, &T.h1[0][0], nullptr, true);
    wave_sync_lds();
    for (int m0 = 0; m0 < 32; m0 += 16)
        dense16<MLP_H, MLP_H>(&T.h1[0][0], MLP_H, MLP_H, &W.w2t[0][0], W.b2, m0, &T.h2[0][0], nullptr, true);
    wave_sync_lds();
    for (int m0 = 0; m0 < 32; m0 += 16)
        dense16<MLP_H, MLP_OUT>(&T.h2[0][0], MLP_H, MLP_H, &W.w3t[0][0], W.b3, m0, nullptr, &T.z3[0][0], false);
    wave_sync_lds();
}

// Stages the features of samples s0 .. s0+31 (rows beyond `cnt` zero) as BF16.
__device__ __forceinline__ void stage_x(MlpTile &T, const float *feat, const int *idx, int64_t s0, int cnt)
{
    const int l = __lane_id();
    for (int e = l; e < 32 * MLP_IN; e += 64) {
        const int r = e / MLP_IN, c = e % MLP_IN;
        float v = 0.f;
        if (r < cnt && c < PROP_FEATS) {
            const int64_t row = idx ? (int64_t)idx[s0 + r] : s0 + r;
            v = feat[PROP_FEATS * row + c];
        }
        T.x[r][c] = f2bf(v);
    }
}

static constexpr int MLP_THREADS = 256;
static constexpr int MLP_WAVES = MLP_THREADS / 64;

__global__ void __launch_bounds__(MLP_THREADS) k_mlp_forward(const float *params, const float *feat, int64_t n,
                                                             float *logits, const int *cand, float threshold,
                                                             uint64_t *pairs, uint8_t *capv)
{
    __shared__ MlpW W;
    __shared__ MlpTile tiles[MLP_WAVES];
    stage_weights(params, W, false);
    __syncthreads();
    MlpTile &T = tiles[threadIdx.x / 64];
    const int l = __lane_id();
    const int64_t ntiles = (n + 31) / 32;
    for (int64_t t = (int64_t)blockIdx.x * MLP_WAVES + threadIdx.x / 64; t < ntiles; t += (int64_t)gridDim.x * MLP_WAVES) {
        const int64_t s0 = 32 * t;
        const int cnt = (int)min<int64_t>(32, n - s0);
        stage_x(T, feat, nullptr, s0, cnt);
        wave_sync_lds();
        mlp_forward_tile(W, T);
        wave_sync_lds();
        if (logits) {
            for (int e = l; e < cnt * MLP_OUT; e += 64) logits[MLP_OUT * s0 + e] = T.z3[e / MLP_OUT][e % MLP_OUT];
        }
        if (pairs) {
            // 32 samples x 8 slots = 4 per lane
            for (int e = l; e < 32 * PROP_SLOTS; e += 64) {
                const int r = e / PROP_SLOTS, j = e % PROP_SLOTS;
                if (r >= cnt) continue;
                const int64_t i = s0 + r;
                const int c = cand[PROP_SLOTS * i + j];
                pairs[PROP_SLOTS * i + j] = (c >= 0 && T.z3[r][j] > threshold)
                                                ? (((uint64_t)(uint32_t)i << 32) | (uint32_t)c)
                                                : ~0ull;
            }
            if (l < cnt) {
                int best = 0;
                float bv = T.z3[l][PROP_SLOTS];
                for (int b = 1; b < PROP_BINS; b++)
                    if (T.z3[l][PROP_SLOTS + b] > bv) {
                        bv = T.z3[l][PROP_SLOTS + b];
                        best = b;
                    }
                capv[s0 + l] = (uint8_t)best;
            }
        }
        wave_sync_lds();
    }
}

// Loss and output gradient of one sample's logits z[16] with label word `lab` (FP32).
__device__ __forceinline__ float sample_loss_grad(const float *z, uint32_t lab, float *dz)
{
    float loss = 0.f;
#pragma unroll
    for (int j = 0; j < PROP_SLOTS; j++) {
        const bool valid = (lab >> (8 + j)) & 1u;
        const float y = (float)((lab >> j) & 1u);
        const float x = z[j];
        if (valid) {
            loss += fmaxf(x, 0.f) - x * y + log1pf(expf(-fabsf(x)));
            dz[j] = 1.f / (1.f + expf(-x)) - y;
        } else {
            dz[j] = 0.f;
        }
    }
    const int bin = (int)((lab >> 16) & 7u);
    float mx = z[PROP_SLOTS];
#pragma unroll
    for (int b = 1; b < PROP_BINS; b++) mx = fmaxf(mx, z[PROP_SLOTS + b]);
    float se = 0.f;
#pragma unroll
    for (int b = 0; b < PROP_BINS; b++) se += expf(z[PROP_SLOTS + b] - mx);
    loss += mx + logf(se) - z[PROP_SLOTS + bin];
#pragma unroll
    for (int b = 0; b < PROP_BINS; b++) dz[PROP_SLOTS + b] = expf(z[PROP_SLOTS + b] - mx) / se - (b == bin ? 1.f : 0.f);
    return loss;
}

// Weight-gradient tile: acc[M/16][N/16] += A^T . B over the tile's 32 samples, A = act
// [32][M] and B = dz [32][N] (row-major BF16 in LDS); K runs over the samples.
template <int M, int N>
__device__ __forceinline__ void wgrad(const uint16_t *act, int lda, int acols, const uint16_t *dz, f32x4 (&acc)[M / 16][N / 16])
{
    const int l = __lane_id();
    const int r = l & 15, kq = 8 * (l >> 4);
#pragma unroll
    for (int mt = 0; mt < M / 16; mt++) {
        short8 a;
#pragma unroll
        for (int q = 0; q < 8; q++) a[q] = mt * 16 + r < acols ? (short)act[(kq + q) * lda + mt * 16 + r] : (short)0;
#pragma unroll
        for (int nt = 0; nt < N / 16; nt++) {
            const short8 b = lds_col8(dz, N, 32, kq, nt * 16 + r);
            acc[mt][nt] = mfma16x16x32(a, b, acc[mt][nt]);
        }
    }
}

// Backward dh[16 rows][N] = dz[16][K] . W^T with W row-major [N][K] (the layer's [in][out]),
// masked by the ReLU of h (h > 0), rounded to BF16 into dzprev.
template <int K, int N>
__device__ __forceinline__ void dgrad16(const uint16_t *dz, int kcols, const uint16_t *w, const uint16_t *h, int m0,
                                        uint16_t *dzprev)
{
    const int l = __lane_id();
    const int r = l & 15, kq = 8 * (l >> 4);
#pragma unroll
    for (int nt = 0; nt < N / 16; nt++) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < (K + 31) / 32; kt++) {
            const short8 a = lds_row8(dz, K, kcols, m0 + r, kt * 32 + kq);
            const short8 b = lds_row8(w, K, kcols, nt * 16 + r, kt * 32 + kq);
            acc = mfma16x16x32(a, b, acc);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int row = m0 + 4 * (l >> 4) + q, col = nt * 16 + r;
            const float v = bf2f(h[row * N + col]) > 0.f ? acc[q] : 0.f;
            dzprev[row * N + col] = f2bf(v);
        }
    }
}

__global__ void __launch_bounds__(MLP_THREADS) k_mlp_grad(const float *params, const float *feat,
                                                          const uint32_t *labels, const int *idx, int64_t n,
                                                          float *grad, double *loss)
{
    __shared__ MlpW W;
    __shared__ MlpTile tiles[MLP_WAVES];
    stage_weights(params, W, true);
    __syncthreads();
    MlpTile &T = tiles[threadIdx.x / 64];
    const int l = __lane_id();
    f32x4 g1[MLP_IN / 16][MLP_H / 16], g2[MLP_H / 16][MLP_H / 16], g3[MLP_H / 16][MLP_OUT / 16];
#pragma unroll
    for (int a = 0; a < MLP_IN / 16; a++)
#pragma unroll
        for (int b = 0; b < MLP_H / 16; b++) g1[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < MLP_H / 16; a++)
#pragma unroll
        for (int b = 0; b < MLP_H / 16; b++) g2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < MLP_H / 16; a++) g3[a][0] = f32x4{0.f, 0.f, 0.f, 0.f};
    float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, lsum = 0.f;
    const int64_t ntiles = (n + 31) / 32;
    for (int64_t t = (int64_t)blockIdx.x * MLP_WAVES + threadIdx.x / 64; t < ntiles; t += (int64_t)gridDim.x * MLP_WAVES) {
        const int64_t s0 = 32 * t;
        const int cnt = (int)min<int64_t>(32, n - s0);
        stage_x(T, feat, idx, s0, cnt);
        if (l < 32) T.lab[l] = l < cnt ? labels[idx ? (int64_t)idx[s0 + l] : s0 + l] : 0u;
        wave_sync_lds();
        mlp_forward_tile(W, T);
        wave_sync_lds();
        // output gradient per sample (rows beyond cnt contribute nothing)
        if (l < 32) {
            float z[MLP_OUT], dz[MLP_OUT];
#pragma unroll
            for (int k = 0; k < MLP_OUT; k++) z[k] = T.z3[l][k];
            if (l < cnt) {
                lsum += sample_loss_grad(z, T.lab[l], dz);
            } else {
#pragma unroll
                for (int k = 0; k < MLP_OUT; k++) dz[k] = 0.f;
            }
#pragma unroll
            for (int k = 0; k < MLP_OUT; k++) T.dz3[l][k] = f2bf(dz[k]);
        }
        wave_sync_lds();
        for (int m0 = 0; m0 < 32; m0 += 16)
            dgrad16<MLP_OUT, MLP_H>(&T.dz3[0][0], MLP_OUT, &W.w3[0][0], &T.h2[0][0], m0, &T.dz2[0][0]);
        wave_sync_lds();
        for (int m0 = 0; m0 < 32; m0 += 16)
            dgrad16<MLP_H, MLP_H>(&T.dz2[0][0], MLP_H, &W.w2[0][0], &T.h1[0][0], m0, &T.dz1[0][0]);
        wave_sync_lds();
        wgrad<MLP_IN, MLP_H>(&T.x[0][0], MLP_IN, MLP_IN, &T.dz1[0][0], g1);
        wgrad<MLP_H, MLP_H>(&T.h1[0][0], MLP_H, MLP_H, &T.dz2[0][0], g2);
        wgrad<MLP_H, MLP_OUT>(&T.h2[0][0], MLP_H, MLP_H, &T.dz3[0][0], g3);
        // bias gradients: column sums of the BF16 output gradients
        for (int r = 0; r < 32; r++) {
            gb1 += bf2f(T.dz1[r][l]);
            gb2 += bf2f(T.dz2[r][l]);
            if (l < MLP_OUT) gb3 += bf2f(T.dz3[r][l]);
        }
        wave_sync_lds();
    }
    // accumulate into the global sums: D[4 (l/16) + q][l % 16] of each tile
    const int r = l & 15, q0 = 4 * (l >> 4);
#pragma unroll
    for (int mt = 0; mt < MLP_IN / 16; mt++)
#pragma unroll
        for (int nt = 0; nt < MLP_H / 16; nt++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = mt * 16 + q0 + q, o = nt * 16 + r;
                if (i < PROP_FEATS) atomicAdd(&grad[MLP_OFF_W1 + i * MLP_H + o], g1[mt][nt][q]);
            }
#pragma unroll
    for (int mt = 0; mt < MLP_H / 16; mt++)
#pragma unroll
        for (int nt = 0; nt < MLP_H / 16; nt++)
#pragma unroll
            for (int q = 0; q < 4; q++)
                atomicAdd(&grad[MLP_OFF_W2 + (mt * 16 + q0 + q) * MLP_H + nt * 16 + r], g2[mt][nt][q]);
#pragma unroll
    for (int mt = 0; mt < MLP_H / 16; mt++)
#pragma unroll
        for (int q = 0; q < 4; q++) atomicAdd(&grad[MLP_OFF_W3 + (mt * 16 + q0 + q) * MLP_OUT + r], g3[mt][0][q]);
    atomicAdd(&grad[MLP_OFF_B1 + l], gb1);
    atomicAdd(&grad[MLP_OFF_B2 + l], gb2);
    if (l < MLP_OUT) atomicAdd(&grad[MLP_OFF_B3 + l], gb3);
    // loss: wave sum, one atomic per wave
    float v = lsum;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (l == 0) atomicAdd(loss, (double)v);
}

__global__ void k_adam(float *p, float *m, float *v, float *g, float inv_count, float lr, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= MLP_NPARAM) return;
    const bool pad = i < MLP_OFF_B1 && (i / MLP_H) >= PROP_FEATS; // padded input rows stay zero
    const float gr = pad ? 0.f : g[i] * inv_count;
    const float mi = 0.9f * m[i] + 0.1f * gr;
    const float vi = 0.999f * v[i] + 0.001f * gr * gr;
    m[i] = mi;
    v[i] = vi;
    if (!pad) p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
    g[i] = 0.f;
}

__global__ void k_labels(const int *cand, int n, const int *vnum, const int64_t *voff, const int *vpool,
                         uint32_t *labels)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int vn = vnum[i];
        const int64_t vo = voff[i];
        uint32_t w = (uint32_t)prop_bin(vn) << 16;
        for (int j = 0; j < PROP_SLOTS; j++) {
            const int c = cand[PROP_SLOTS * (int64_t)i + j];
            if (c < 0) continue;
            w |= 1u << (8 + j);
            for (int k = 0; k < vn; k++)
                if (vpool[vo + k] == c) {
                    w |= 1u << j;
                    break;
                }
        }
        labels[i] = w;
    }
}

__global__ void k_capacity_vertices(const uint8_t *capv, const int64_t *ws_off, int n, int64_t *vcap)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int64_t e = prop_bin_edge(capv[i]);
        const int64_t w = ws_off[i + 1] - ws_off[i];
        vcap[i] = e > w ? e : w;
    }
}

__global__ void k_fill_f32(float *a, int64_t n, float v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

// ------------------------------------------------------------------------------------
// Launchers
// ------------------------------------------------------------------------------------

hipError_t launch_fine_keys(const double *pts, int n, const GridSets &S, int code_bits, int pos_bits, int shifted,
                            uint64_t *keys, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (code_bits < 1 || code_bits > 21) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_fine_keys, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, pts, n, S, code_bits, pos_bits,
                       shifted, keys);
    return hipGetLastError();
}

hipError_t launch_fine_ranks(const uint64_t *keys, int n, int pos_bits, int *ord, int *rank, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fine_ranks, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, keys, n,
                       (1ull << pos_bits) - 1ull, ord, rank);
    return hipGetLastError();
}

hipError_t launch_candidates(const double *pts, int n, const GridSets &S, const int *ord0, const int *rank0,
                             const int *ord1, const int *rank1, int *cand, float *d2, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_candidates, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, pts, n, S, ord0, rank0, ord1,
                       rank1, cand, d2);
    return hipGetLastError();
}

hipError_t launch_features(const double *pts, const uint64_t *vox, int n, const GridSets &S, const int *idgrid,
                           const uint64_t *ws_keys, const int64_t *ws_off, const int *cand, const float *d2,
                           float *feat, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_features, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, pts, vox, n, S, idgrid, ws_keys,
                       ws_off, cand, d2, feat);
    return hipGetLastError();
}

static int mlp_blocks(int64_t n)
{
    const int64_t tiles = (n + 31) / 32;
    int64_t b = (tiles + MLP_WAVES - 1) / MLP_WAVES;
    return (int)(b < 1 ? 1 : (b > 2048 ? 2048 : b));
}

hipError_t launch_mlp_propose(const float *params, const float *feat, const int *cand, int n, float threshold,
                              uint64_t *pairs, uint8_t *capv, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_forward, dim3(mlp_blocks(n)), dim3(MLP_THREADS), 0, s, params, feat, (int64_t)n, nullptr,
                       cand, threshold, pairs, capv);
    return hipGetLastError();
}

hipError_t launch_mlp_forward(const float *params, const float *feat, int64_t n, float *logits, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_forward, dim3(mlp_blocks(n)), dim3(MLP_THREADS), 0, s, params, feat, n, logits,
                       (const int *)nullptr, 0.f, (uint64_t *)nullptr, (uint8_t *)nullptr);
    return hipGetLastError();
}

hipError_t launch_mlp_grad(const float *params, const float *feat, const uint32_t *labels, const int *idx, int64_t n,
                           float *grad, double *loss, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_grad, dim3(mlp_blocks(n)), dim3(MLP_THREADS), 0, s, params, feat, labels, idx, n, grad,
                       loss);
    return hipGetLastError();
}

hipError_t launch_adam(float *params, float *m, float *v, float *grad, float inv_count, float lr, int step,
                       hipStream_t s)
{
    const float c1 = 1.f - powf(0.9f, (float)step), c2 = 1.f - powf(0.999f, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s, params, m, v, grad, inv_count, lr, c1,
                       c2);
    return hipGetLastError();
}

hipError_t launch_labels(const int *cand, int n, const int *vnum, const int64_t *voff, const int *vpool,
                         uint32_t *labels, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_labels, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, cand, n, vnum, voff, vpool, labels);
    return hipGetLastError();
}

hipError_t launch_capacity_vertices(const uint8_t *capv, const int64_t *ws_off, int n, int64_t *vcap, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_capacity_vertices, dim3(g4d_blocks(n, 256, 65536)), dim3(256), 0, s, capv, ws_off, n, vcap);
    return hipGetLastError();
}

hipError_t launch_fill_f32(float *a, int64_t n, float v, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_f32, dim3(g4d_blocks(n, 256, 4096)), dim3(256), 0, s, a, n, v);
    return hipGetLastError();
}

} // namespace g4d
