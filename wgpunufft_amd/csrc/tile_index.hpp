// Index arithmetic of the output-stationary spreaders, shared by the device
// kernels and host-side replays of their addressing: the bin intervals that
// reach a block of cells, the reach rows of a tile, the record slot of the
// k-th record of a tile, and the work item -> (tile, part, records) map of
// the tile plan.  Everything here is plain integer arithmetic on values the
// caller passes in.
#pragma once

#include <hip/hip_runtime.h>

#include "nufft_internal.hpp"

namespace mnufft {

// Bin intervals of one axis that hold the support starts reaching a block
// (starts b0-w+1 .. b0+B-1, periodic).  A reach within one bin side of the
// whole axis takes every bin once.
struct AxisReach {
    int64_t lo0, len0, lo1, len1;
    __host__ __device__ int64_t count() const { return len0 + len1; }
    __host__ __device__ int64_t bin(int64_t k) const { return k < len0 ? lo0 + k : lo1 + (k - len0); }
};

__host__ __device__ inline AxisReach axis_reach(int64_t b0, int64_t B, int w, const AxisParams &a)
{
    AxisReach r{0, 0, 0, 0};
    int64_t n = a.n;
    int64_t L = B + w - 1 < n ? B + w - 1 : n;
    int64_t side = (int64_t)1 << a.shift;
    if (L + side > n) {
        r.lo0 = 0;
        r.len0 = a.bins;
        return r;
    }
    int64_t r0 = b0 - (w - 1);
    if (r0 < 0) r0 += n;
    if (r0 + L <= n) {
        r.lo0 = r0 >> a.shift;
        r.len0 = ((r0 + L - 1) >> a.shift) - r.lo0 + 1;
    } else {
        r.lo0 = r0 >> a.shift;
        r.len0 = a.bins - r.lo0;
        r.lo1 = 0;
        r.len1 = ((r0 + L - n - 1) >> a.shift) + 1;
    }
    return r;
}

// Bin segments in reach of a block of the output-stationary spreaders: one
// x interval (two when it wraps) per bin row of the other axes' reach.
template <int D>
struct BlockReach {
    AxisReach ax[D];
    int64_t b0[D];
    int64_t nseg;
    // segment t -> record slots [first, last)
    __host__ __device__ void segment(int64_t t, const GridParams &g, const uint32_t *offsets,
                            int64_t *first, int64_t *last) const
    {
        int64_t nxi = ax[0].len1 > 0 ? 2 : 1;
        int64_t ix = t % nxi;
        int64_t rest = t / nxi;
        int64_t row = 0, stride = g.ax[0].bins;
#pragma unroll
        for (int a = 1; a < D; ++a) {
            int64_t cnt = ax[a].count();
            int64_t k = rest % cnt;
            rest /= cnt;
            row += ax[a].bin(k) * stride;
            stride *= g.ax[a].bins;
        }
        int64_t lo = ix == 0 ? ax[0].lo0 : ax[0].lo1;
        int64_t len = ix == 0 ? ax[0].len0 : ax[0].len1;
        *first = offsets[row + lo];
        *last = offsets[row + lo + len];
    }
};

template <int D>
__host__ __device__ inline BlockReach<D> block_reach(int64_t block, const GridParams &g,
                                            const SpreadParams &sp)
{
    BlockReach<D> r;
    int64_t rest = block;
    r.nseg = 1;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        int64_t bi = rest % sp.nblocks[a];
        rest /= sp.nblocks[a];
        r.b0[a] = bi * sp.block[a];
        r.ax[a] = axis_reach(r.b0[a], sp.block[a], g.width, g.ax[a]);
        if (a == 0)
            r.nseg *= r.ax[0].len1 > 0 ? 2 : 1;
        else
            r.nseg *= r.ax[a].count();
    }
    return r;
}

// Every reach row of a row tile (y start, then z start; bins one cell along
// y and z): its one or two slot ranges along x (the second when the reach
// wraps) and the exclusive scan of the record counts.
template <int N>
struct TileRows {
    uint32_t f0[N], n0[N], f1[N], excl[N + 1];
};

// Offset index of the first x bin of reach row r of the tile at (y0, z0):
// rows run over NRY y starts, then z starts, from w - 1 cells before the
// tile's first cell (periodic).
template <int D, int NRY, int W>
__host__ __device__ inline int64_t tile_row_base(int r, int64_t y0, int64_t z0,
                                                 const GridParams &g)
{
    const int64_t n1 = g.ax[1].n, n2 = g.ax[2].n;
    int64_t sy = y0 - (W - 1) + r % NRY;
    if (sy < 0) sy += n1;
    if (sy >= n1) sy -= n1;
    int64_t line = sy;
    if constexpr (D == 3) {
        int64_t sz = z0 - (W - 1) + r / NRY;
        if (sz < 0) sz += n2;
        if (sz >= n2) sz -= n2;
        line += n1 * sz;
    }
    return g.ax[0].bins * line;
}

// Slot of record k of the tile; *q = its reach row.
template <int NROWS>
__host__ __device__ inline uint32_t tile_slot(const TileRows<NROWS> &rows, uint32_t k, int *q)
{
    int lo = 0, hi = NROWS - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (rows.excl[mid] <= k)
            lo = mid;
        else
            hi = mid - 1;
    }
    *q = lo;
    const uint32_t j = k - rows.excl[lo];
    return j < rows.n0[lo] ? rows.f0[lo] + j : rows.f1[lo] + (j - rows.n0[lo]);
}

// Work item -> (tile, part, record range) of the tile plan.
struct TileItem {
    int64_t b, p, rlo, rhi;
    uint32_t np;
};

__host__ __device__ inline TileItem tile_item(int64_t item, int64_t block_count,
                                              const uint32_t *__restrict__ item_off,
                                              const uint32_t *__restrict__ parts,
                                              const uint32_t *__restrict__ totals,
                                              uint32_t part_len_cap)
{
    int64_t lo = 0, hi = block_count - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if ((int64_t)item_off[mid] <= item)
            lo = mid;
        else
            hi = mid - 1;
    }
    TileItem t;
    t.b = lo;
    t.p = item - item_off[lo];
    t.np = parts[lo];
    const int64_t total = totals[lo];
    t.rlo = 0;
    t.rhi = total;
    if (t.np > 1) {
        t.rlo = t.p * (int64_t)part_len_cap;
        t.rhi = t.rlo + part_len_cap < total ? t.rlo + part_len_cap : total;
    }
    return t;
}

// Problem of a batch's combined element t (work item, partial slot, heavy
// tile): the last problem q whose base[4q + k] <= t.
__host__ __device__ inline int batch_problem_of(const int64_t *__restrict__ base, int nprob, int k,
                                                int64_t t)
{
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base[4 * mid + k] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Float index of a heavy part's partial sum of cell (x, y, z) of a tile of
// TX x TY x TZ cells (real part; the imaginary part follows): slot `pslot`,
// vector v of `vectors`.
template <int TX, int TY, int TZ>
__host__ __device__ inline int64_t partial_index(int64_t pslot, int64_t vectors, int64_t v,
                                                 int x, int y, int z)
{
    return 2 * (((pslot * vectors + v) * TZ + z) * ((int64_t)TX * TY) + (int64_t)y * TX + x);
}

}  // namespace mnufft
