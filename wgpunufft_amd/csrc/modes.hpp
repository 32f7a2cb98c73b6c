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

struct DirectParams {
    int64_t n_modes[kMaxDims];
    int64_t mode_total;
    int64_t M;
    int64_t nout;
    int centered;
    double sign;
};

template <int D, typename T>
bool deconvolve(const ModeParams &mp, const double *inv_phihat, const T *grid, T *modes,
                int64_t vectors, hipStream_t stream);
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
template <typename T>
bool type3_final(int64_t N, int64_t vectors, const T *interp, const double *factor, T *out,
                 hipStream_t stream);
// kind 1: out[o] = sum_j in[j] e^{i sign k_o.x_j} over the linear mode
// indices sel (or all); kind 2: out[o] = sum_k in[k] e^{i sign k.x_o} for
// the points sel; kind 3: out[o] = sum_j in[j] e^{i sign s_o.x_j}.
bool direct_sum(int kind, int d, const DirectParams &dp, const double *x, const double *in,
                const int64_t *sel, const double *s, double *out, hipStream_t stream);

}  // namespace mnufft
