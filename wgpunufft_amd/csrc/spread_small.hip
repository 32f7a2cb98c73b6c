// Deterministic spreading of many points onto a small F32 grid: the outer
// grids of 1D and 2D type-3 plans, which hold up to millions of sources on a
// few thousand cells (reference src/gpu_dense_spread.rs:1-48).
//
// A binned output-stationary spread parallelizes over cells and starves on
// such grids.  Here the points, in their original order, are cut into at
// most 1024 equal groups; a workgroup takes one group and one tile of 64
// segments of 8 cells along axis 0 (each lane owns one segment of one row),
// stages the group's points 256 at a time in LDS (support starts, weights,
// strengths) and adds every point of the chunk, in point order, to the cells
// of its segment, then writes the group's partial grid.  A reduction adds
// the partial grids of every cell in group order.  The groups depend only on
// the point count, so the result is bitwise repeatable with neither atomics
// nor a point sort.
#include <hip/hip_runtime.h>

#include "fold.hpp"
#include "spread.hpp"
#include "spread_device.hpp"

namespace mnufft {

namespace {

constexpr int kSmallLanes = 64;
constexpr int kSmallSeg = 8;       // cells of a lane's segment along axis 0
constexpr int kSmallChunk = 256;   // points staged in LDS at a time
constexpr int kSmallMaxW = 16;

struct SmallPoint {
    int sx, sy;     // wrapped support starts
    float cre, cim;
};

// Periodic offset of cell c from support start s, in [0, n).
__device__ inline int wrap_off(int c, int s, int n)
{
    int d = c - s;
    d += d < 0 ? n : 0;
    d -= d >= n ? n : 0;
    return d;
}

// One workgroup: group `grp` of the points, tile `tile` of the segments,
// vector blockIdx.y.  partial[(v * groups + grp) * cells + cell].
template <int D>
__global__ __launch_bounds__(kSmallLanes) void k_small_spread(
    GridParams g, int64_t M, int64_t groups, int64_t group_points, int64_t tiles,
    const float *__restrict__ pts, const float *__restrict__ c, float *__restrict__ partial)
{
    __shared__ SmallPoint sp[kSmallChunk];
    __shared__ float wts[kSmallChunk][D][kSmallMaxW];
    const int W = g.width;
    const int n0 = (int)g.ax[0].n, n1 = D > 1 ? (int)g.ax[1].n : 1;
    const int64_t cells = (int64_t)n0 * n1;
    const int segs_row = (n0 + kSmallSeg - 1) / kSmallSeg;
    const int64_t tile = blockIdx.x % tiles;
    const int64_t grp = blockIdx.x / tiles;
    const int64_t v = blockIdx.y;
    const int64_t seg = tile * kSmallLanes + threadIdx.x;
    const bool live = seg < (int64_t)segs_row * n1;
    const int cx = (int)(seg % segs_row) * kSmallSeg;
    const int row = live ? (int)(seg / segs_row) : 0;
    const float beta = (float)g.beta;
    const int64_t first = grp * group_points;
    const int64_t last = first + group_points < M ? first + group_points : M;
    float sre[kSmallSeg], sim[kSmallSeg];
#pragma unroll
    for (int i = 0; i < kSmallSeg; ++i) sre[i] = sim[i] = 0.0f;
    for (int64_t c0 = first; c0 < last; c0 += kSmallChunk) {
        const int n = (int)(last - c0 < kSmallChunk ? last - c0 : kSmallChunk);
        for (int k = threadIdx.x; k < n; k += kSmallLanes) {
            const int64_t j = c0 + k;
            float x[D];
#pragma unroll
            for (int a = 0; a < D; ++a) x[a] = pts[j * D + a];
            bool reach = true;
#pragma unroll
            for (int a = 0; a < D; ++a) reach = reach && FoldTraits<float>::in_reach(x[a]);
            int st[2] = {0, 0};
#pragma unroll
            for (int a = 0; a < D; ++a) {
                int64_t s;
                double off;
                support_start<float>(x[a], reach, g, a, &s, &off);
                st[a] = (int)wrapped_start(s, g.ax[a].n);
                for (int t = 0; t < W; ++t)
                    wts[k][a][t] = es_weight<float>((float)t + (float)off, W, beta);
            }
            const float2 cv = *(const float2 *)(c + 2 * (v * M + j));
            sp[k] = SmallPoint{st[0], st[1], cv.x, cv.y};
        }
        __syncthreads();
        if (live) {
            for (int k = 0; k < n; ++k) {
                const SmallPoint p = sp[k];
                float wy = 1.0f;
                if constexpr (D == 2) {
                    const int dy = wrap_off(row, p.sy, n1);
                    if (dy >= W) continue;
                    wy = wts[k][1][dy];
                }
                const float vre = p.cre * wy, vim = p.cim * wy;
#pragma unroll
                for (int i = 0; i < kSmallSeg; ++i) {
                    const int dx = wrap_off(cx + i, p.sx, n0);
                    if (dx < W) {
                        const float wx = wts[k][0][dx];
                        sre[i] += vre * wx;
                        sim[i] += vim * wx;
                    }
                }
            }
        }
        __syncthreads();
    }
    if (live) {
        float *dst = partial + 2 * ((v * groups + grp) * cells + (int64_t)row * n0 + cx);
#pragma unroll
        for (int i = 0; i < kSmallSeg; ++i) {
            if (cx + i < n0) {
                dst[2 * i] = sre[i];
                dst[2 * i + 1] = sim[i];
            }
        }
    }
}

// grid[v][cell] = sum over groups, in group order, of the partial grids.
__global__ __launch_bounds__(256) void k_small_reduce(int64_t cells, int64_t groups,
                                                      int64_t vectors,
                                                      const float *__restrict__ partial,
                                                      float *__restrict__ grid)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells * vectors) return;
    const int64_t v = i / cells, cell = i - v * cells;
    const float *src = partial + 2 * (v * groups * cells + cell);
    float re = 0.0f, im = 0.0f;
    for (int64_t q = 0; q < groups; ++q) {
        const float2 p = *(const float2 *)(src + 2 * q * cells);
        re += p.x;
        im += p.y;
    }
    *(float2 *)(grid + 2 * i) = make_float2(re, im);
}

}  // namespace

bool small_spread_supported(int dim, const int64_t *fine, int width)
{
    if (dim < 1 || dim > 2 || width > kSmallMaxW) return false;
    int64_t cells = 1;
    for (int a = 0; a < dim; ++a) {
        if (fine[a] < 2 * width) return false;
        cells *= fine[a];
    }
    return cells <= kSmallSpreadMaxCells;
}

int64_t small_spread_groups(int64_t M)
{
    int64_t g = (M + kSmallGroupPoints - 1) / kSmallGroupPoints;
    return g < 1 ? 1 : (g > kSmallMaxGroups ? kSmallMaxGroups : g);
}

bool SmallSpreader::spread(const GridParams &g, int64_t M, const float *pts, const float *c,
                           int64_t vectors, float *grid, hipStream_t stream)
{
    const int dim = g.dims;
    if (dim < 1 || dim > 2 || g.width > kSmallMaxW) {
        set_error("small-grid spread needs a 1D or 2D grid and a width <= 16");
        return false;
    }
    if (vectors > 65535) {
        set_error("at most 65535 vectors per execution");
        return false;
    }
    const int64_t cells = g.fine_total;
    const int64_t groups = small_spread_groups(M);
    const int64_t group_points = (M + groups - 1) / groups;
    const int64_t segs = ((g.ax[0].n + kSmallSeg - 1) / kSmallSeg) * (dim > 1 ? g.ax[1].n : 1);
    const int64_t tiles = (segs + kSmallLanes - 1) / kSmallLanes;
    // vectors in chunks whose partial grids stay within the budget
    const int64_t per_vector = groups * cells * 2 * (int64_t)sizeof(float);
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(vectors, budget_bytes / per_vector));
    const int64_t need = per_vector * chunk;
    if (need > partial_bytes) {
        if (partial) (void)hipFree(partial);
        partial = nullptr;
        partial_bytes = 0;
        if (!hip_ok(hipMalloc(&partial, need), "small-grid partial grids")) return false;
        partial_bytes = need;
    }
    if (groups * tiles > 0x7fffffff) {
        set_error("small-grid spread: too many workgroups");
        return false;
    }
    for (int64_t v0 = 0; v0 < vectors; v0 += chunk) {
        const int64_t vc = std::min<int64_t>(chunk, vectors - v0);
        const dim3 grid3((unsigned)(groups * tiles), (unsigned)vc);
        const float *cv = c + 2 * v0 * M;
        if (dim == 1)
            k_small_spread<1><<<grid3, kSmallLanes, 0, stream>>>(g, M, groups, group_points, tiles,
                                                                 pts, cv, (float *)partial);
        else
            k_small_spread<2><<<grid3, kSmallLanes, 0, stream>>>(g, M, groups, group_points, tiles,
                                                                 pts, cv, (float *)partial);
        const int64_t n = cells * vc;
        k_small_reduce<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(
            cells, groups, vc, (const float *)partial, grid + 2 * v0 * cells);
    }
    return hip_ok(hipGetLastError(), "small-grid spread launch");
}

SmallSpreader::~SmallSpreader()
{
    if (partial) (void)hipFree(partial);
}

}  // namespace mnufft
