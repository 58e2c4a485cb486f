// gStar4D for MI355X — learned proposal: candidate neighbours, per-point features of the
// digital Voronoi diagram and a small MLP that picks extra workset candidates and predicts
// each star's final link size.
//
// It plugs in where the reference builds worksets (_pumpWorksets, GD/GDelaunay/GDelHost.cu:
// 352-529) and sizes star storage (__expandStarsForInsertion, GD/GDelaunay/GDelHost.cu:
// 1139-1186).  Proposed pairs only add vertices to the stars' starting worksets; splaying
// still runs to full consistency, so the tetrahedra never depend on the proposal.
#pragma once

#include "g4d_grid.h"

namespace g4d {

// Candidates: the PROP_SLOTS nearest points (squared distance) among the points within
// +-PROP_WIN positions of a point in two fine Morton orders (one shifted by a third of the
// range), restricted to the point's set.
static constexpr int PROP_SLOTS = 8;
static constexpr int PROP_WIN = 12;
static constexpr int PROP_FEATS = 12;
static constexpr int PROP_BINS = 8;
// Capacity bins of the final link size: <= 10, 13, 16, 19, 22, 26, 32 vertices, and above.
__host__ __device__ inline int prop_bin_edge(int b)
{
    return b == 0 ? 10 : b == 1 ? 13 : b == 2 ? 16 : b == 3 ? 19 : b == 4 ? 22 : b == 5 ? 26 : b == 6 ? 32 : 48;
}

// MLP 12 -> 64 -> 64 -> 16 (ReLU), weights BF16 on the 16x16x32 MFMA, FP32 accumulation.
// Parameter block (FP32): W1[MLP_IN][MLP_H] (rows >= PROP_FEATS zero), b1[MLP_H],
// W2[MLP_H][MLP_H], b2[MLP_H], W3[MLP_H][MLP_OUT], b3[MLP_OUT]; outputs 0..7 are the slot
// logits, 8..15 the capacity-bin logits.
static constexpr int MLP_IN = 32;
static constexpr int MLP_H = 64;
static constexpr int MLP_OUT = 16;
static constexpr int MLP_OFF_W1 = 0;
static constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_IN * MLP_H;
static constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
static constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
static constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
static constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H * MLP_OUT;
static constexpr int MLP_NPARAM = MLP_OFF_B3 + MLP_OUT;

// Label word of a training sample: bit j (j < 8) slot j's candidate is in the final link,
// bit 8 + j slot j holds a candidate, bits 16..18 the capacity bin of the final link size.
__host__ __device__ inline int prop_bin(int link_vertices)
{
    int b = 0;
    while (b < PROP_BINS - 1 && link_vertices > prop_bin_edge(b)) b++;
    return b;
}

// Sort keys of one fine Morton order over every set: set | code | position (bits chosen by
// the caller so they fit 64).  shift != 0 offsets the lattice by a third of its range.
hipError_t launch_fine_keys(const double *pts, int n, const GridSets &S, int code_bits, int pos_bits, int shifted,
                            uint64_t *keys, hipStream_t s);
// ord[r] = position of rank r (low pos_bits of the sorted keys); rank[ord[r]] = r.
hipError_t launch_fine_ranks(const uint64_t *keys, int n, int pos_bits, int *ord, int *rank, hipStream_t s);
// cand[8i + j], d2[8i + j]: the PROP_SLOTS nearest of the window candidates (-1 / +inf when
// fewer), nearest first.
hipError_t launch_candidates(const double *pts, int n, const GridSets &S, const int *ord0, const int *rank0,
                             const int *ord1, const int *rank1, int *cand, float *d2, hipStream_t s);
// Features feat[12 i + f] (input-independent scale: voxel units and logarithms):
//  0 the point lost its voxel (not the voxel's leader), 1 log2(1 + points in its voxel),
//  2 log2(1 + workset size before pumping), 3 mean voxel distance to its first 32 workset
//  vertices, 4 log2 of the nearest candidate's distance in voxels, 5..11 log2 of the distance
//  ratio of candidate j + 1 to the nearest (8 when absent).
hipError_t launch_features(const double *pts, const uint64_t *vox, int n, const GridSets &S, const int *idgrid,
                           const uint64_t *ws_keys, const int64_t *ws_off, const int *cand, const float *d2,
                           float *feat, hipStream_t s);
// Inference: per point, slot j proposes (i << 32 | cand) into pairs[8i + j] when its logit
// exceeds `threshold` (STAR_PAIR_NONE otherwise); capv[i] = predicted capacity bin.
hipError_t launch_mlp_propose(const float *params, const float *feat, const int *cand, int n, float threshold,
                              uint64_t *pairs, uint8_t *capv, hipStream_t s);
// Logits (16 per sample) of n samples.
hipError_t launch_mlp_forward(const float *params, const float *feat, int64_t n, float *logits, hipStream_t s);
// Gradient sums over samples [0, n) (or the minibatch rows idx[0..n) when idx is set) added
// into grad[MLP_NPARAM]; the loss sum is added into *loss.  Loss per sample: binary
// cross-entropy with logits over the valid slots plus the capacity-bin cross-entropy.
hipError_t launch_mlp_grad(const float *params, const float *feat, const uint32_t *labels, const int *idx, int64_t n,
                           float *grad, double *loss, hipStream_t s);
// Adam step on params from grad / count; clears grad.
hipError_t launch_adam(float *params, float *m, float *v, float *grad, float inv_count, float lr, int step,
                       hipStream_t s);
// Label word of every point from the final star links (link vertices of star i are
// vpool[voff[i] .. voff[i] + vnum[i])).
hipError_t launch_labels(const int *cand, int n, const int *vnum, const int64_t *voff, const int *vpool,
                         uint32_t *labels, hipStream_t s);
// Predicted final link vertices of every star from its capacity bin (the bin's upper edge; 48
// for the open bin), at least the star's current workset size.
hipError_t launch_capacity_vertices(const uint8_t *capv, const int64_t *ws_off, int n, int64_t *vcap, hipStream_t s);
hipError_t launch_fill_f32(float *a, int64_t n, float v, hipStream_t s);

} // namespace g4d
