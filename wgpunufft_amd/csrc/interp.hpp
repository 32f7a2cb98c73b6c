// Type-2 interpolation from the fine grid at the points.
#pragma once

#include "nufft_internal.hpp"

namespace mnufft {

// out[v * M + j] = sum over the support of point j of grid[v] * weights, for
// the `vectors` grids; points are visited in the order of `sorted`.
template <int D, typename T>
bool interpolate(int64_t M, int64_t vectors, const uint32_t *sorted, const T *coords,
                 const GridParams &g, const T *grid, T *out, hipStream_t stream);

// One launch for many problems (1D/2D): problem p's points, at slots
// [m_off, m_off + M) of `sorted` (indices of the concatenated point set),
// are interpolated from its grid into its output, in work items of
// 256 / next_pow2(w) points from item_base on (items over all problems).
// The same per-point sums as interpolate().
template <int D, typename T>
bool interpolate_many(const uint32_t *sorted, const BatchProblem *probs, int nprob,
                      int64_t item_begin, int64_t items, int width, hipStream_t stream);

// 2D, width <= 8, bins of 8 x 8 cells: tiles staged in LDS (like the 3D
// tiled interpolation), one workgroup per bin.
template <typename T>
bool interpolate_tiled2(int64_t M, int64_t vectors, const uint32_t *sorted,
                        const uint32_t *offsets, const T *coords, const GridParams &g,
                        const T *grid, T *out, hipStream_t stream);

// The same for the problems of a batch (F32): offsets over the combined bin
// list, one launch per kernel width c over bins [bin_lo[c], bin_hi[c]) (the
// problems of one width take consecutive bins).
bool interpolate_tiled2_many(const uint32_t *sorted, const uint32_t *offsets,
                             const BatchProblem *probs, int nprob, const int *widths,
                             const int64_t *bin_lo, const int64_t *bin_hi, int nclass,
                             hipStream_t stream);

// 3D, width <= 8: points binned by 8^3 tiles of support starts (the bin
// geometry must have sides 8), grid tiles staged in LDS.  item_off needs
// nbins + 1 + scan_tiles(nbins + 1) words of scratch.
template <typename T>
bool interpolate_tiled3(int64_t M, int64_t vectors, const uint32_t *sorted, const uint32_t *offsets,
                        uint32_t *item_off, const T *coords, const GridParams &g, const T *grid,
                        T *out, hipStream_t stream);

// The same for the problems of a batch (F32, one vector each): offsets over
// the combined bin list (every problem's bins of 8^3 cells), one launch per
// kernel width c over bins [bin_lo[c], bin_hi[c]) (the problems of one width
// take consecutive bins; points[c] bounds their points).  item_off needs
// nbins + 1 + scan_tiles(nbins + 1) words of scratch.
bool interpolate_tiled3_many(const uint32_t *sorted, const uint32_t *offsets, uint32_t *item_off,
                             int64_t nbins, const BatchProblem *probs, int nprob,
                             const int *widths, const int64_t *bin_lo, const int64_t *bin_hi,
                             const int64_t *points, int nclass, hipStream_t stream);

}  // namespace mnufft
