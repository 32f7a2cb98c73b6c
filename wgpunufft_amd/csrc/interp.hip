// Type-2 interpolation on gfx950.
//
// Reference behaviour (src/gpu_3d.rs:1238-1359 and src/gpu_2d.rs:1022-1113):
// points are visited in bin order so that neighbouring lanes read
// overlapping fine-grid rows from cache; `G = next_pow2(w)` lanes evaluate
// one point, lane j owning support offset j along axis 0, summing its column
// over the other axes; the lane sums are then combined in a fixed order, so
// each output is one fixed-order sum that does not depend on the point order.
//
// MI355X design: the G lanes of a point are a sub-group of a 64-lane
// wavefront; the per-axis kernel weights are computed once per lane (weight
// j on lane j) and broadcast with cross-lane shuffles instead of being
// recomputed, and the lane sums combine in a fixed butterfly.
#include <hip/hip_runtime.h>

#include "fold.hpp"
#include "interp.hpp"

namespace mnufft {

namespace {

template <typename T>
__device__ inline T es_w(T d, int w, T beta)
{
    T s = (T)2 * d / (T)w;
    T q = (T)1 - s * s;
    if (!(q > (T)0)) return (T)0;
    if constexpr (sizeof(T) == 4)
        return expf(beta * (sqrtf(q) - 1.0f));
    else
        return exp(beta * (sqrt(q) - 1.0));
}

__device__ inline int64_t wrap_idx(int64_t i, int64_t n)
{
    if (i < 0) i += n;
    if (i >= n) i -= n;
    return i;
}

template <int D, typename T, int G>
__global__ __launch_bounds__(256) void k_interp(int64_t M, int64_t vectors,
                                                const uint32_t *__restrict__ sorted,
                                                const T *__restrict__ coords, GridParams g,
                                                const T *__restrict__ grid, T *__restrict__ out)
{
    constexpr int P = 256 / G;
    const int j = threadIdx.x % G;
    const int64_t item = (int64_t)blockIdx.x * P + threadIdx.x / G;
    const bool has = item < M * vectors;
    const int w = g.width;
    const T beta = (T)g.beta;
    int64_t v = 0, s = 0;
    if (has) {
        v = item / M;
        s = item - v * M;
    }
    uint32_t idx = has ? sorted[s] : 0u;
    bool reach = has && point_in_reach<D, T>(coords, idx);
    int64_t st[3] = {0, 0, 0};
    T wt[3] = {(T)0, (T)0, (T)0};
#pragma unroll
    for (int a = 0; a < D; ++a) {
        int64_t start = 0;
        double off = kOutsideOffset;
        if (has) support_start<T>(coords[(int64_t)idx * D + a], reach, g, a, &start, &off);
        st[a] = wrapped_start(start, g.ax[a].n);
        wt[a] = j < w ? es_w<T>((T)j + (T)off, w, beta) : (T)0;
    }
    T are = (T)0, aim = (T)0;
    const T *gv = grid + 2 * v * g.fine_total;
    if (has && j < w) {
        int64_t x = wrap_idx(st[0] + j, g.ax[0].n);
        if (D == 1) {
            are = gv[2 * x];
            aim = gv[2 * x + 1];
        } else if (D == 2) {
            int64_t n0 = g.ax[0].n;
            for (int dy = 0; dy < w; ++dy) {
                T wy = __shfl(wt[1], dy, G);
                int64_t y = wrap_idx(st[1] + dy, g.ax[1].n);
                const T *p = gv + 2 * (x + n0 * y);
                are += wy * p[0];
                aim += wy * p[1];
            }
        } else {
            int64_t n0 = g.ax[0].n, n01 = n0 * g.ax[1].n;
            int64_t ys[kMaxWidth];
            T wys[kMaxWidth];
            for (int dy = 0; dy < w; ++dy) {
                ys[dy] = n0 * wrap_idx(st[1] + dy, g.ax[1].n);
                wys[dy] = __shfl(wt[1], dy, G);
            }
            for (int dz = 0; dz < w; ++dz) {
                T wz = __shfl(wt[2], dz, G);
                int64_t z = n01 * wrap_idx(st[2] + dz, g.ax[2].n);
                T pre = (T)0, pim = (T)0;
                for (int dy = 0; dy < w; ++dy) {
                    const T *p = gv + 2 * (x + ys[dy] + z);
                    pre += wys[dy] * p[0];
                    pim += wys[dy] * p[1];
                }
                are += wz * pre;
                aim += wz * pim;
            }
        }
        are *= wt[0];
        aim *= wt[0];
    }
    // (the shuffles above read lanes 0..w-1 of the point's group, which are
    // exactly the lanes inside this branch)
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        are += __shfl_xor(are, o, G);
        aim += __shfl_xor(aim, o, G);
    }
    if (has && j == 0) {
        T *dst = out + 2 * (v * M + (int64_t)idx);
        dst[0] = are;
        dst[1] = aim;
    }
}

template <int D, typename T, int G>
bool launch(int64_t M, int64_t vectors, const uint32_t *sorted, const T *coords,
            const GridParams &g, const T *grid, T *out, hipStream_t stream)
{
    constexpr int P = 256 / G;
    int64_t items = M * vectors;
    unsigned blocks = (unsigned)((items + P - 1) / P);
    k_interp<D, T, G><<<blocks, 256, 0, stream>>>(M, vectors, sorted, coords, g, grid, out);
    return hip_ok(hipGetLastError(), "interpolation launch");
}

}  // namespace

template <int D, typename T>
bool interpolate(int64_t M, int64_t vectors, const uint32_t *sorted, const T *coords,
                 const GridParams &g, const T *grid, T *out, hipStream_t stream)
{
    if (M == 0) return true;
    if (g.width <= 4) return launch<D, T, 4>(M, vectors, sorted, coords, g, grid, out, stream);
    if (g.width <= 8) return launch<D, T, 8>(M, vectors, sorted, coords, g, grid, out, stream);
    return launch<D, T, 16>(M, vectors, sorted, coords, g, grid, out, stream);
}

#define MNUFFT_INTERP_INSTANTIATE(D, T)                                                           \
    template bool interpolate<D, T>(int64_t, int64_t, const uint32_t *, const T *,                 \
                                    const GridParams &, const T *, T *, hipStream_t);
MNUFFT_INTERP_INSTANTIATE(1, float)
MNUFFT_INTERP_INSTANTIATE(2, float)
MNUFFT_INTERP_INSTANTIATE(3, float)
MNUFFT_INTERP_INSTANTIATE(1, double)
MNUFFT_INTERP_INSTANTIATE(2, double)
MNUFFT_INTERP_INSTANTIATE(3, double)

}  // namespace mnufft
