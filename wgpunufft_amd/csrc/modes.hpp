// Mode-space and type-3 elementwise stages, FP64 direct sums.
#pragma once

#include "nufft_internal.hpp"

namespace mnufft {

struct ModeParams {
    int64_t n_modes[kMaxDims];
    int64_t fine[kMaxDims];
    int64_t phihat_off[kMaxDims];  // offset of each axis in the 1/phi_hat table
    int64_t mode_total;
    int64_t fine_total;
    int centered;
};

constexpr int kMaxQuadrature = 20;

struct Type3Params {
    double src_center[kMaxDims];
    double inv_gamma[kMaxDims];
    double tgt_center[kMaxDims];
    double tgt_scale[kMaxDims];  // h * gamma
    double sign;
    int nquad;
    double quad_x[kMaxQuadrature];  // positive Gauss-Legendre nodes times w/2
    double quad_w[kMaxQuadrature];  // matching weights times 2 (w/2) phi(node)
};

// Target boxes of a partitioned type-3 plan: axis a is cut into k[a] boxes
// of width[a] from lo[a]; box index = b0 + k0 (b1 + k1 b2).
struct Type3Boxes {
    int k[kMaxDims];
    double lo[kMaxDims];
    double width[kMaxDims];
    int64_t count;
};

struct DirectParams {
    int64_t n_modes[kMaxDims];
    int64_t mode_total;
    int64_t M;
    int64_t nout;
    int centered;
    double sign;
};

// Mode stage of one problem of a multi-plan execution: its modes take
// [mode_base, mode_base + mp.mode_total) and its fine cells [fine_base,
// fine_base + mp.fine_total) of the combined lists; `zero` (type 1 without
// points) writes zeros as the plan's own execute does.
struct ModeProblem {
    ModeParams mp;
    const double *inv_phihat;
    int64_t mode_base, fine_base;
    int zero;
};

template <int D, typename T>
bool deconvolve(const ModeParams &mp, const double *inv_phihat, const T *grid, T *modes,
                int64_t vectors, hipStream_t stream);
// Deconvolution (type 1: probs[q].grid -> probs[q].out) and
// predeconvolution (type 2: probs[q].in -> probs[q].grid) of every problem
// in one launch each, the same bits as the per-plan kernels.
template <int D, typename T>
bool deconvolve_many(const BatchProblem *probs, const ModeProblem *mprobs, int nprob,
                     int64_t total_modes, hipStream_t stream);
template <int D, typename T>
bool predeconvolve_many(const BatchProblem *probs, const ModeProblem *mprobs, int nprob,
                        int64_t total_cells, hipStream_t stream);
template <int D, typename T>
bool predeconvolve(const ModeParams &mp, const double *inv_phihat, const T *modes, T *grid,
                   int64_t vectors, hipStream_t stream);
template <typename T>
bool fill_zero(T *p, int64_t n, hipStream_t stream);
template <int D, typename T>
bool type3_source(const Type3Params &tp, int64_t M, int64_t vectors, const T *x, const T *c, T *xr,
                  T *cp, hipStream_t stream);
template <int D, typename T>
bool type3_target(const Type3Params &tp, int64_t N, const T *s, T *sr, double *factor,
                  hipStream_t stream);
// out[v * n_out + (perm ? perm[k] : k)] for k < N (n_out = N without perm).
template <typename T>
bool type3_final(int64_t N, int64_t vectors, const T *interp, const double *factor, T *out,
                 hipStream_t stream, const uint32_t *perm = nullptr, int64_t n_out = 0);
// Target boxes: box[k] of every target and counts[box] (counts cleared
// here first), then, with offsets[box] = the exclusive scan of the counts,
// perm[] = the targets grouped by box (any order inside a box) and
// s_sorted[] = their coordinates in that order; cursor (count values) is
// scratch.
template <int D, typename T>
bool type3_box_count(const Type3Boxes &bx, int64_t N, const T *s, uint32_t *box,
                     uint32_t *counts, hipStream_t stream);
template <int D, typename T>
bool type3_box_scatter(const Type3Boxes &bx, int64_t N, const T *s, const uint32_t *box,
                       const uint32_t *offsets, uint32_t *cursor, uint32_t *perm, T *s_sorted,
                       hipStream_t stream);
// Per vector v < vectors: out[v][c][r] = in[v][r][c] for a rows x cols
// matrix of complex values (the rank 4-8 FFT moves the axes it has
// transformed behind the others with it).
template <typename T>
bool transpose_complex(const T *in, T *out, int64_t rows, int64_t cols, int64_t vectors,
                       hipStream_t stream);
// kind 1: out[o] = sum_j in[j] e^{i sign k_o.x_j} over the linear mode
// indices sel (or all); kind 2: out[o] = sum_k in[k] e^{i sign k.x_o} for
// the points sel; kind 3: out[o] = sum_j in[j] e^{i sign s_o.x_j}.
bool direct_sum(int kind, int d, const DirectParams &dp, const double *x, const double *in,
                const int64_t *sel, const double *s, double *out, hipStream_t stream);

}  // namespace mnufft
