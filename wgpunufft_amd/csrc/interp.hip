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

#include "bins.hpp"
#include "fold.hpp"
#include "interp.hpp"
#include "spread_device.hpp"

namespace mnufft {

namespace {

// FP32 weights from the hardware sqrt/exp2 (es_weight_fast), FP64 exact.
template <typename T>
__device__ inline T es_w(T d, int w, T beta)
{
    if constexpr (sizeof(T) == 4) {
        return es_weight_fast(d, 2.0f / (float)w, beta * 1.44269504088896341f);
    } else {
        T s = (T)2 * d / (T)w;
        T q = (T)1 - s * s;
        if (!(q > (T)0)) return (T)0;
        return exp(beta * (sqrt(q) - 1.0));
    }
}

__device__ inline int64_t wrap_idx(int64_t i, int64_t n)
{
    if (i < 0) i += n;
    if (i >= n) i -= n;
    return i;
}

// Ranks 4-8: the weighted sum over the support of axes 1..A of one lane's
// axis-0 cell (grid offset `base` so far), axis A outermost; each level sums
// its w inner sums in order, weighted by the axis-A weight of lane d of the
// point's group.
template <int A, typename T, int G>
__device__ inline void nest_sum(const T *__restrict__ gv, int64_t base, const int64_t *st,
                                const T *wt, const int64_t *stride, const GridParams &g, int w,
                                T &re, T &im)
{
    re = (T)0;
    im = (T)0;
    for (int d = 0; d < w; ++d) {
        const T wa = __shfl(wt[A], d, G);
        const int64_t off = base + stride[A] * wrap_idx(st[A] + d, g.ax[A].n);
        T pr, pi;
        if constexpr (A == 1) {
            pr = gv[2 * off];
            pi = gv[2 * off + 1];
        } else {
            nest_sum<A - 1, T, G>(gv, off, st, wt, stride, g, w, pr, pi);
        }
        re += wa * pr;
        im += wa * pi;
    }
}

// The G lanes of a point group: sorted slot s0 (valid below end) holds
// point sorted[s0] - base of `coords`, written to out[2 (vo + point)].
template <int D, typename T, int G>
__device__ inline void interp_group(int64_t s0, int64_t end, uint32_t base,
                                    const uint32_t *__restrict__ sorted,
                                    const T *__restrict__ coords, const GridParams &g,
                                    const T *__restrict__ gv, T *__restrict__ out, int64_t vo)
{
    const int j = threadIdx.x % G;
    const bool has = s0 < end;
    const int64_t s = has ? s0 : 0;
    const int w = g.width;
    const T beta = (T)g.beta;
    uint32_t idx = has ? sorted[s] - base : 0u;
    T x[D];
#pragma unroll
    for (int a = 0; a < D; ++a) x[a] = has ? coords[(int64_t)idx * D + a] : (T)0;
    bool reach = has;
#pragma unroll
    for (int a = 0; a < D; ++a) reach = reach && FoldTraits<T>::in_reach(x[a]);
    int64_t st[D];
    T wt[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        int64_t start = 0;
        double off = kOutsideOffset;
        if (has) support_start<T>(x[a], reach, g, a, &start, &off);
        st[a] = wrapped_start(start, g.ax[a].n);
        wt[a] = j < w ? es_w<T>((T)j + (T)off, w, beta) : (T)0;
    }
    T are = (T)0, aim = (T)0;
    if (has && j < w) {
        int64_t xc = wrap_idx(st[0] + j, g.ax[0].n);
        if (D == 1) {
            are = gv[2 * xc];
            aim = gv[2 * xc + 1];
        } else if (D == 2) {
            int64_t n0 = g.ax[0].n;
            for (int dy = 0; dy < w; ++dy) {
                T wy = __shfl(wt[1], dy, G);
                int64_t y = wrap_idx(st[1] + dy, g.ax[1].n);
                const T *p = gv + 2 * (xc + n0 * y);
                are += wy * p[0];
                aim += wy * p[1];
            }
        } else if constexpr (D >= 4) {
            int64_t stride[D];
            stride[0] = 1;
#pragma unroll
            for (int a = 1; a < D; ++a) stride[a] = stride[a - 1] * g.ax[a - 1].n;
            nest_sum<D - 1, T, G>(gv, xc, st, wt, stride, g, w, are, aim);
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
                    const T *p = gv + 2 * (xc + ys[dy] + z);
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
        T *dst = out + 2 * (vo + (int64_t)idx);
        dst[0] = are;
        dst[1] = aim;
    }
}

template <int D, typename T, int G>
__global__ __launch_bounds__(256) void k_interp(int64_t M, int64_t vectors,
                                                const uint32_t *__restrict__ sorted,
                                                const T *__restrict__ coords, GridParams g,
                                                const T *__restrict__ grid, T *__restrict__ out)
{
    constexpr int P = 256 / G;
    // blockIdx.y = vector, so each launch dimension stays within 2^32 lanes
    const int64_t v = blockIdx.y;
    interp_group<D, T, G>((int64_t)blockIdx.x * P + threadIdx.x / G, M, 0u, sorted, coords, g,
                          grid + 2 * v * g.fine_total, out, v * M);
}

// Many problems in one launch: the work items of problem p (P points each,
// in the order of the batched sort) start at item_base; every workgroup
// takes items with a stride.
template <int D, typename T, int G>
__global__ __launch_bounds__(256) void k_interp_many(const uint32_t *__restrict__ sorted,
                                                     const BatchProblem *__restrict__ probs,
                                                     int nprob, int64_t item_begin,
                                                     int64_t items)
{
    constexpr int P = 256 / G;
    for (int64_t item = item_begin + blockIdx.x; item < items; item += gridDim.x) {
        int lo = 0, hi = nprob - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (probs[mid].item_base <= item)
                lo = mid;
            else
                hi = mid - 1;
        }
        const BatchProblem &pb = probs[lo];
        const int64_t s0 = pb.m_off + (item - pb.item_base) * P + threadIdx.x / G;
        interp_group<D, T, G>(s0, pb.m_off + pb.M, (uint32_t)pb.m_off, sorted,
                              (const T *)pb.coords, pb.g, (const T *)pb.grid, (T *)pb.out, 0);
    }
}

// ---------------------------------------------------------------------
// LDS-tiled 3D interpolation.  Type-2 bins here are tiles of TX x TY x TZ
// (8^3) cells of support starts; a work item is one tile and up to P of its
// points, which the bin sort leaves ordered by support-start cell, so the
// lanes of a wavefront read neighbouring windows.  The workgroup stages the
// tile plus its (w-1)-cell halo (periodic) in LDS once, then every lane
// evaluates whole points from LDS: w^2 rows of w cells (F32: read as aligned
// pairs of cells, 16 bytes per LDS read) and 2 w^3 FMAs per point, with the
// fine grid read from HBM about (TX+w-1)(TY+w-1)(TZ+w-1)/(TX TY TZ) times
// per item instead of w^3 times per point.  Work items are dealt to the
// workgroups in order, so the items of a heavy tile spread over every XCD.
// Each output is one lane's fixed-order sum.
// ---------------------------------------------------------------------
template <typename T>
struct Complex2;
template <>
struct Complex2<float> {
    using type = float2;
};
template <>
struct Complex2<double> {
    using type = double2;
};

// Work items per bin, ceil(count / P), into item_off[b] (item_off[nbins] =
// 0); the exclusive scan (scan_u32, in place) then makes item_off[b] the
// first item of bin b and item_off[nbins] the number of items.
__global__ __launch_bounds__(256) void k_item_counts(int64_t nbins, uint32_t points_per_item,
                                                     const uint32_t *__restrict__ offsets,
                                                     uint32_t *__restrict__ item_off)
{
    const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (b > nbins) return;
    uint32_t t = 0;
    if (b < nbins) {
        const uint32_t n = offsets[b + 1] - offsets[b];
        t = (n + points_per_item - 1) / points_per_item;
    }
    item_off[b] = t;
}

// Tile and part of work item `item`.
struct TileItem3 {
    int64_t b, k, x0, y0, z0;
};

__device__ inline TileItem3 tile_item3(int64_t item, const uint32_t *__restrict__ item_off,
                                       const GridParams &g, int tx, int ty, int tz)
{
    int64_t lo = 0, hi = g.nbins - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if ((int64_t)item_off[mid] <= item)
            lo = mid;
        else
            hi = mid - 1;
    }
    TileItem3 t;
    t.b = lo;
    t.k = item - item_off[lo];
    t.x0 = (lo % g.ax[0].bins) * tx;
    t.y0 = ((lo / g.ax[0].bins) % g.ax[1].bins) * ty;
    t.z0 = lo / (g.ax[0].bins * g.ax[1].bins) * tz;
    return t;
}

// LDS tile of the 3D tiled interpolation: TX x TY x TZ cells and their
// halo.  F32: every row of a point's window is read as NP aligned pairs of
// cells (16-byte LDS reads), starting at the even cell at or before the
// window, with a zero weight on the cell outside it.
template <typename T, int W, int TX, int TY, int TZ>
struct Tile3 {
    static constexpr bool kPairs = sizeof(T) == 4;
    static constexpr int NP = (W + 2) / 2;
    static constexpr int EX = TX + W - 1, EY = TY + W - 1, EZ = TZ + W - 1;
    static constexpr int EXR = kPairs && TX + 2 * NP - 2 > EX ? TX + 2 * NP - 2 : EX;
    static constexpr int RX = kPairs ? (EXR + 1) / 2 * 2 : EX;  // LDS row stride (cells)
    static constexpr int CELLS = RX * EY * EZ;
};

// One work item: the tile at (x0, y0, z0) of grid gv staged in LDS, then the
// points at sorted slots [first, last) (point index sorted[s] - base of
// coords) interpolated from it into out_v.
template <typename T, int W, int TX, int TY, int TZ, int THREADS>
__device__ __forceinline__ void interp_tile3(typename Complex2<T>::type *tile, const GridParams &g,
                                             const T *__restrict__ gv, int64_t x0, int64_t y0,
                                             int64_t z0, int64_t first, int64_t last,
                                             const uint32_t *__restrict__ sorted, uint32_t base,
                                             const T *__restrict__ coords, T *__restrict__ out_v)
{
    using Shape = Tile3<T, W, TX, TY, TZ>;
    constexpr bool kPairs = Shape::kPairs;
    constexpr int NP = Shape::NP;
    constexpr int EY = Shape::EY, EZ = Shape::EZ;
    constexpr int RX = Shape::RX;
    const int64_t n0 = g.ax[0].n, n1 = g.ax[1].n, n2 = g.ax[2].n;
    const T beta = (T)g.beta;
    // tile loads: every wavefront fills RPW rows of RX cells at a time, each
    // lane keeping its x cell
    constexpr int ROWS = EY * EZ;
    static_assert(RX <= 64, "a tile row must fit one wavefront");
    constexpr int RPW = 64 / RX;
    const int lane = threadIdx.x & 63;
    const int rsub = lane / RX, lx = lane % RX;
    const int r_first = (int)(threadIdx.x >> 6) * RPW + rsub;
    constexpr int R_STEP = (THREADS / 64) * RPW;
    using T2 = typename Complex2<T>::type;
    if (rsub < RPW) {
        int64_t gx = x0 + lx;
        gx = gx >= n0 ? gx - n0 : gx;
        for (int r = r_first; r < ROWS; r += R_STEP) {
            const int ly = r % EY, lz = r / EY;
            int64_t gy = y0 + ly, gz = z0 + lz;
            gy = gy >= n1 ? gy - n1 : gy;
            gz = gz >= n2 ? gz - n2 : gz;
            tile[r * RX + lx] = *(const T2 *)(gv + 2 * (gx + n0 * (gy + n1 * gz)));
        }
    }
    __syncthreads();
    for (int64_t s = first + threadIdx.x; s < last; s += THREADS) {
        const uint32_t idx = sorted[s] - base;
        T x[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) x[a] = coords[(int64_t)idx * 3 + a];
        const bool reach = FoldTraits<T>::in_reach(x[0]) && FoldTraits<T>::in_reach(x[1]) &&
                           FoldTraits<T>::in_reach(x[2]);
        int l[3];
        T wt[3][W];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            int64_t start;
            double off;
            support_start<T>(x[a], reach, g, a, &start, &off);
            int64_t ws = wrapped_start(start, g.ax[a].n);
            int64_t o = a == 0 ? x0 : (a == 1 ? y0 : z0);
            l[a] = (int)(ws - o);
            if constexpr (sizeof(T) == 4) {
                // the kernel at distances j + off: scaled distance off * 2/w +
                // j * 2/w, exponent as one FMA (hardware sqrt / exp2)
                constexpr float two_over_w = 2.0f / (float)W;
                const float u0 = (float)off * two_over_w;
                const float bl = (float)beta * 1.44269504088896341f;
#pragma unroll
                for (int j = 0; j < W; ++j) {
                    const float sv = u0 + (float)j * two_over_w;
                    const float q = __builtin_fmaf(-sv, sv, 1.0f);
                    const float e = __builtin_amdgcn_exp2f(
                        __builtin_fmaf(__builtin_amdgcn_sqrtf(q), bl, -bl));
                    wt[a][j] = q > 0.0f ? e : 0.0f;
                }
            } else {
#pragma unroll
                for (int j = 0; j < W; ++j) wt[a][j] = es_w<T>((T)j + (T)off, W, beta);
            }
        }
        T sre = (T)0, sim = (T)0;
        if constexpr (kPairs) {
            const int sh = l[0] & 1;
            T wx[2 * NP];
#pragma unroll
            for (int k = 0; k < 2 * NP; ++k) {
                const T a0 = k < W ? wt[0][k] : (T)0;
                const T a1 = k >= 1 && k - 1 < W ? wt[0][k - 1] : (T)0;
                wx[k] = sh ? a1 : a0;
            }
#pragma unroll
            for (int dz = 0; dz < W; ++dz) {
                T pre = (T)0, pim = (T)0;
#pragma unroll
                for (int dy = 0; dy < W; ++dy) {
                    const int rowb = ((l[2] + dz) * EY + (l[1] + dy)) * RX + l[0] - sh;
                    const float4 *rp = (const float4 *)(tile + rowb);
                    T rre = (T)0, rim = (T)0;
#pragma unroll
                    for (int q = 0; q < NP; ++q) {
                        const float4 c2 = rp[q];
                        rre += wx[2 * q] * c2.x;
                        rim += wx[2 * q] * c2.y;
                        rre += wx[2 * q + 1] * c2.z;
                        rim += wx[2 * q + 1] * c2.w;
                    }
                    pre += wt[1][dy] * rre;
                    pim += wt[1][dy] * rim;
                }
                sre += wt[2][dz] * pre;
                sim += wt[2][dz] * pim;
            }
        } else {
#pragma unroll
            for (int dz = 0; dz < W; ++dz) {
                T pre = (T)0, pim = (T)0;
#pragma unroll
                for (int dy = 0; dy < W; ++dy) {
                    const int rowb = ((l[2] + dz) * EY + (l[1] + dy)) * RX + l[0];
                    T rre = (T)0, rim = (T)0;
#pragma unroll
                    for (int dx = 0; dx < W; ++dx) {
                        const T2 cv = tile[rowb + dx];
                        rre += wt[0][dx] * cv.x;
                        rim += wt[0][dx] * cv.y;
                    }
                    pre += wt[1][dy] * rre;
                    pim += wt[1][dy] * rim;
                }
                sre += wt[2][dz] * pre;
                sim += wt[2][dz] * pim;
            }
        }
        T *dst = out_v + 2 * (int64_t)idx;
        dst[0] = sre;
        dst[1] = sim;
    }
    __syncthreads();
}

template <typename T, int W, int TX, int TY, int TZ, int THREADS, int P>
__global__ __launch_bounds__(THREADS) void k_interp_tiled3(
    int64_t M, const uint32_t *__restrict__ sorted, const uint32_t *__restrict__ offsets,
    const uint32_t *__restrict__ item_off, const T *__restrict__ coords, GridParams g,
    const T *__restrict__ grid, T *__restrict__ out)
{
    using T2 = typename Complex2<T>::type;
    __shared__ __align__(16) T2 tile[Tile3<T, W, TX, TY, TZ>::CELLS];
    const int64_t n_items = item_off[g.nbins];
    const int64_t v = blockIdx.y;
    const T *gv = grid + 2 * v * g.fine_total;
    for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const TileItem3 ti = tile_item3(item, item_off, g, TX, TY, TZ);
        const int64_t first = (int64_t)offsets[ti.b] + ti.k * P;
        int64_t last = first + P;
        if (last > (int64_t)offsets[ti.b + 1]) last = offsets[ti.b + 1];
        interp_tile3<T, W, TX, TY, TZ, THREADS>(tile, g, gv, ti.x0, ti.y0, ti.z0, first, last,
                                                sorted, 0u, coords, out + 2 * v * M);
    }
}

// Many problems (one vector each): the work items of the combined bins
// [bin_lo, bin_hi) of a batched sort (per bin ceil(points / P) items in
// item_off); an item's bin -> its problem (bin bases) -> the problem's tile.
template <typename T, int W, int TX, int TY, int TZ, int THREADS, int P>
__global__ __launch_bounds__(THREADS) void k_interp_tiled3_many(
    const uint32_t *__restrict__ sorted, const uint32_t *__restrict__ offsets,
    const uint32_t *__restrict__ item_off, const BatchProblem *__restrict__ probs, int nprob,
    int64_t bin_lo, int64_t bin_hi)
{
    using T2 = typename Complex2<T>::type;
    __shared__ __align__(16) T2 tile[Tile3<T, W, TX, TY, TZ>::CELLS];
    const int64_t i0 = item_off[bin_lo], i1 = item_off[bin_hi];
    for (int64_t item = i0 + blockIdx.x; item < i1; item += gridDim.x) {
        int64_t lo = bin_lo, hi = bin_hi - 1;
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if ((int64_t)item_off[mid] <= item)
                lo = mid;
            else
                hi = mid - 1;
        }
        const int64_t b = lo;
        int qlo = 0, qhi = nprob - 1;
        while (qlo < qhi) {
            const int mid = (qlo + qhi + 1) >> 1;
            if ((int64_t)probs[mid].bin_base <= b)
                qlo = mid;
            else
                qhi = mid - 1;
        }
        const BatchProblem &pb = probs[qlo];
        const int64_t lb = b - (int64_t)pb.bin_base;
        const int64_t x0 = (lb % pb.g.ax[0].bins) * TX;
        const int64_t y0 = ((lb / pb.g.ax[0].bins) % pb.g.ax[1].bins) * TY;
        const int64_t z0 = lb / (pb.g.ax[0].bins * pb.g.ax[1].bins) * TZ;
        const int64_t first = (int64_t)offsets[b] + (item - item_off[b]) * P;
        int64_t last = first + P;
        if (last > (int64_t)offsets[b + 1]) last = offsets[b + 1];
        interp_tile3<T, W, TX, TY, TZ, THREADS>(tile, pb.g, (const T *)pb.grid, x0, y0, z0, first,
                                                last, sorted, (uint32_t)pb.m_off,
                                                (const T *)pb.coords, (T *)pb.out);
    }
}

// ---------------------------------------------------------------------
// LDS-tiled 2D interpolation: type-2 bins are tiles of TX x TY cells of
// support starts; workgroups take bins with a stride (a 2D tile holds a few
// hundred points at most densities, so a bin is one unit of work).  The
// workgroup stages the tile and its (w-1)-cell halo (periodic) in LDS, then
// every lane evaluates whole points from LDS (w rows of w cells, F32 rows
// read as aligned pairs of cells).  With MANY, bins run over the problems of
// a batch (bins [bin_lo, bin_hi) of the combined bin list, one kernel
// width), every tile taking its problem's grid, points and output.  Each
// output is one lane's fixed-order sum.
// ---------------------------------------------------------------------
template <typename T, int W, int TX, int TY, int THREADS, bool MANY>
__global__ __launch_bounds__(THREADS) void k_interp_tiled2(
    int64_t M, const uint32_t *__restrict__ sorted, const uint32_t *__restrict__ offsets,
    int64_t bin_lo, int64_t bin_hi, const T *__restrict__ coords0, GridParams g0,
    const T *__restrict__ grid0, T *__restrict__ out0, const BatchProblem *__restrict__ probs,
    int nprob)
{
    constexpr bool kPairs = sizeof(T) == 4;
    constexpr int NP = (W + 2) / 2;
    constexpr int EX = TX + W - 1, EY = TY + W - 1;
    constexpr int EXR = kPairs && TX + 2 * NP - 2 > EX ? TX + 2 * NP - 2 : EX;
    constexpr int RX = kPairs ? (EXR + 1) / 2 * 2 : EX;  // LDS row stride (cells)
    using T2 = typename Complex2<T>::type;
    __shared__ __align__(16) T2 tile[RX * EY];
    static_assert(RX <= 64, "a tile row must fit one wavefront");
    constexpr int RPW = 64 / RX;
    const int lane = threadIdx.x & 63;
    const int rsub = lane / RX, lx = lane % RX;
    const int r_first = (int)(threadIdx.x >> 6) * RPW + rsub;
    constexpr int R_STEP = (THREADS / 64) * RPW;
    const int64_t v = blockIdx.y;
    for (int64_t B = bin_lo + blockIdx.x; B < bin_hi; B += gridDim.x) {
        const int64_t first = offsets[B], last = offsets[B + 1];
        if (first == last) continue;
        GridParams g = g0;
        const T *coords = coords0;
        const T *gv = grid0;
        T *out = out0;
        int64_t bin = B, vo = 0;
        uint32_t base = 0;
        if constexpr (MANY) {
            int plo = 0, phi = nprob - 1;
            while (plo < phi) {
                const int mid = (plo + phi + 1) >> 1;
                if ((int64_t)probs[mid].bin_base <= B)
                    plo = mid;
                else
                    phi = mid - 1;
            }
            const BatchProblem &pb = probs[plo];
            g = pb.g;
            coords = (const T *)pb.coords;
            gv = (const T *)pb.grid;
            out = (T *)pb.out;
            bin = B - (int64_t)pb.bin_base;
            base = (uint32_t)pb.m_off;
        } else {
            gv = grid0 + 2 * v * g0.fine_total;
            vo = v * M;
        }
        const int64_t n0 = g.ax[0].n, n1 = g.ax[1].n;
        const int64_t x0 = (bin % g.ax[0].bins) * TX, y0 = (bin / g.ax[0].bins) * TY;
        if (rsub < RPW) {
            int64_t gx = x0 + lx;
            gx = gx >= n0 ? gx - n0 : gx;
            for (int r = r_first; r < EY; r += R_STEP) {
                int64_t gy = y0 + r;
                gy = gy >= n1 ? gy - n1 : gy;
                tile[r * RX + lx] = *(const T2 *)(gv + 2 * (gx + n0 * gy));
            }
        }
        __syncthreads();
        for (int64_t s = first + threadIdx.x; s < last; s += THREADS) {
            const uint32_t idx = sorted[s] - base;
            T x[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) x[a] = coords[(int64_t)idx * 2 + a];
            const bool reach = FoldTraits<T>::in_reach(x[0]) && FoldTraits<T>::in_reach(x[1]);
            int l[2];
            T wt[2][W];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                int64_t start;
                double off;
                support_start<T>(x[a], reach, g, a, &start, &off);
                const int64_t ws = wrapped_start(start, g.ax[a].n);
                l[a] = (int)(ws - (a == 0 ? x0 : y0));
                if constexpr (sizeof(T) == 4) {
                    constexpr float two_over_w = 2.0f / (float)W;
                    const float u0 = (float)off * two_over_w;
                    const float bl = (float)g.beta * 1.44269504088896341f;
#pragma unroll
                    for (int j = 0; j < W; ++j) {
                        const float sv = u0 + (float)j * two_over_w;
                        const float q = __builtin_fmaf(-sv, sv, 1.0f);
                        const float e = __builtin_amdgcn_exp2f(
                            __builtin_fmaf(__builtin_amdgcn_sqrtf(q), bl, -bl));
                        wt[a][j] = q > 0.0f ? e : 0.0f;
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < W; ++j) wt[a][j] = es_w<T>((T)j + (T)off, W, (T)g.beta);
                }
            }
            T sre = (T)0, sim = (T)0;
            if constexpr (kPairs) {
                const int sh = l[0] & 1;
                T wx[2 * NP];
#pragma unroll
                for (int q = 0; q < 2 * NP; ++q) {
                    const T a0 = q < W ? wt[0][q] : (T)0;
                    const T a1 = q >= 1 && q - 1 < W ? wt[0][q - 1] : (T)0;
                    wx[q] = sh ? a1 : a0;
                }
#pragma unroll
                for (int dy = 0; dy < W; ++dy) {
                    const float4 *rp = (const float4 *)(tile + (l[1] + dy) * RX + l[0] - sh);
                    T rre = (T)0, rim = (T)0;
#pragma unroll
                    for (int q = 0; q < NP; ++q) {
                        const float4 c2 = rp[q];
                        rre += wx[2 * q] * c2.x;
                        rim += wx[2 * q] * c2.y;
                        rre += wx[2 * q + 1] * c2.z;
                        rim += wx[2 * q + 1] * c2.w;
                    }
                    sre += wt[1][dy] * rre;
                    sim += wt[1][dy] * rim;
                }
            } else {
#pragma unroll
                for (int dy = 0; dy < W; ++dy) {
                    const int rowb = (l[1] + dy) * RX + l[0];
                    T rre = (T)0, rim = (T)0;
#pragma unroll
                    for (int dx = 0; dx < W; ++dx) {
                        const T2 cv = tile[rowb + dx];
                        rre += wt[0][dx] * cv.x;
                        rim += wt[0][dx] * cv.y;
                    }
                    sre += wt[1][dy] * rre;
                    sim += wt[1][dy] * rim;
                }
            }
            T *dst = out + 2 * (vo + (int64_t)idx);
            dst[0] = sre;
            dst[1] = sim;
        }
        __syncthreads();
    }
}

constexpr int kTile2Threads = 128;

template <typename T, int W, bool MANY>
void launch_tiled2_kernel(int64_t M, int64_t vectors, const uint32_t *sorted,
                          const uint32_t *offsets, int64_t bin_lo, int64_t bin_hi,
                          const T *coords, const GridParams &g, const T *grid, T *out,
                          const BatchProblem *probs, int nprob, hipStream_t stream)
{
    const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>(bin_hi - bin_lo, 32768));
    k_interp_tiled2<T, W, 8, 8, kTile2Threads, MANY>
        <<<dim3(gx, (unsigned)vectors), kTile2Threads, 0, stream>>>(
            M, sorted, offsets, bin_lo, bin_hi, coords, g, grid, out, probs, nprob);
}

template <typename T, bool MANY>
bool launch_tiled2_w(int w, int64_t M, int64_t vectors, const uint32_t *sorted,
                     const uint32_t *offsets, int64_t bin_lo, int64_t bin_hi, const T *coords,
                     const GridParams &g, const T *grid, T *out, const BatchProblem *probs,
                     int nprob, hipStream_t stream)
{
#define MNUFFT_TILED2(Wd)                                                                         \
    case Wd:                                                                                      \
        launch_tiled2_kernel<T, Wd, MANY>(M, vectors, sorted, offsets, bin_lo, bin_hi, coords,   \
                                          g, grid, out, probs, nprob, stream);                   \
        return true;
    switch (w) {
        MNUFFT_TILED2(2)
        MNUFFT_TILED2(3)
        MNUFFT_TILED2(4)
        MNUFFT_TILED2(5)
        MNUFFT_TILED2(6)
        MNUFFT_TILED2(7)
        MNUFFT_TILED2(8)
    default:
        set_error("tiled interpolation needs a kernel width <= 8");
        return false;
    }
#undef MNUFFT_TILED2
}

template <typename T, int W, int TX, int TY, int TZ, int THREADS, int P>
bool launch_tiled(int64_t M, int64_t vectors, const uint32_t *sorted, const uint32_t *offsets,
                  uint32_t *item_off, const T *coords, const GridParams &g, const T *grid, T *out,
                  hipStream_t stream)
{
    // item_off holds nbins + 1 offsets, then the scan's tile sums
    const int64_t n = g.nbins + 1;
    k_item_counts<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(g.nbins, (uint32_t)P, offsets,
                                                                  item_off);
    scan_u32(item_off, n, item_off + n, item_off, stream);
    int64_t items = std::min<int64_t>(g.nbins + M / P + 64, 16384);
    dim3 grid3((unsigned)items, (unsigned)vectors);
    k_interp_tiled3<T, W, TX, TY, TZ, THREADS, P><<<grid3, THREADS, 0, stream>>>(
        M, sorted, offsets, item_off, coords, g, grid, out);
    return hip_ok(hipGetLastError(), "tiled interpolation launch");
}

template <typename T, int W>
bool launch_tiled_w(int64_t M, int64_t vectors, const uint32_t *sorted, const uint32_t *offsets,
                    uint32_t *item_off, const T *coords, const GridParams &g, const T *grid, T *out,
                    hipStream_t stream)
{
    // the bin sides of the type-2 geometry are the tile
    if (g.ax[0].shift == 3 && g.ax[1].shift == 3 && g.ax[2].shift == 3)
        return launch_tiled<T, W, 8, 8, 8, 256, 1024>(M, vectors, sorted, offsets, item_off,
                                                      coords, g, grid, out, stream);
    set_error("tiled interpolation: unsupported tile geometry");
    return false;
}

template <int D, typename T, int G>
bool launch(int64_t M, int64_t vectors, const uint32_t *sorted, const T *coords,
            const GridParams &g, const T *grid, T *out, hipStream_t stream)
{
    constexpr int P = 256 / G;
    dim3 blocks((unsigned)((M + P - 1) / P), (unsigned)vectors);
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

template <int D, typename T>
bool interpolate_many(const uint32_t *sorted, const BatchProblem *probs, int nprob,
                      int64_t item_begin, int64_t items, int width, hipStream_t stream)
{
    if (items <= item_begin) return true;
    const unsigned grid = (unsigned)std::min<int64_t>(items - item_begin, 65535);
    if (width <= 4)
        k_interp_many<D, T, 4><<<grid, 256, 0, stream>>>(sorted, probs, nprob, item_begin, items);
    else if (width <= 8)
        k_interp_many<D, T, 8><<<grid, 256, 0, stream>>>(sorted, probs, nprob, item_begin, items);
    else
        k_interp_many<D, T, 16><<<grid, 256, 0, stream>>>(sorted, probs, nprob, item_begin,
                                                          items);
    return hip_ok(hipGetLastError(), "batched interpolation launch");
}
template bool interpolate_many<1, float>(const uint32_t *, const BatchProblem *, int, int64_t,
                                         int64_t, int, hipStream_t);
template bool interpolate_many<2, float>(const uint32_t *, const BatchProblem *, int, int64_t,
                                         int64_t, int, hipStream_t);
template bool interpolate_many<3, float>(const uint32_t *, const BatchProblem *, int, int64_t,
                                         int64_t, int, hipStream_t);

template <typename T>
bool interpolate_tiled2(int64_t M, int64_t vectors, const uint32_t *sorted,
                        const uint32_t *offsets, const T *coords, const GridParams &g,
                        const T *grid, T *out, hipStream_t stream)
{
    if (M == 0) return true;
    if (!launch_tiled2_w<T, false>(g.width, M, vectors, sorted, offsets, 0, g.nbins, coords, g,
                                   grid, out, nullptr, 0, stream))
        return false;
    return hip_ok(hipGetLastError(), "tiled 2D interpolation launch");
}
template bool interpolate_tiled2<float>(int64_t, int64_t, const uint32_t *, const uint32_t *,
                                        const float *, const GridParams &, const float *,
                                        float *, hipStream_t);
template bool interpolate_tiled2<double>(int64_t, int64_t, const uint32_t *, const uint32_t *,
                                         const double *, const GridParams &, const double *,
                                         double *, hipStream_t);

bool interpolate_tiled2_many(const uint32_t *sorted, const uint32_t *offsets,
                             const BatchProblem *probs, int nprob, const int *widths,
                             const int64_t *bin_lo, const int64_t *bin_hi, int nclass,
                             hipStream_t stream)
{
    for (int c = 0; c < nclass; ++c)
        if (!launch_tiled2_w<float, true>(widths[c], 0, 1, sorted, offsets, bin_lo[c], bin_hi[c],
                                          nullptr, GridParams{}, nullptr, nullptr, probs, nprob,
                                          stream))
            return false;
    return hip_ok(hipGetLastError(), "batched tiled interpolation launch");
}

template <typename T>
bool interpolate_tiled3(int64_t M, int64_t vectors, const uint32_t *sorted, const uint32_t *offsets,
                        uint32_t *item_off, const T *coords, const GridParams &g, const T *grid,
                        T *out, hipStream_t stream)
{
    if (M == 0) return true;
    switch (g.width) {
    case 2: return launch_tiled_w<T, 2>(M, vectors, sorted, offsets, item_off, coords, g, grid, out, stream);
    case 3: return launch_tiled_w<T, 3>(M, vectors, sorted, offsets, item_off, coords, g, grid, out, stream);
    case 4: return launch_tiled_w<T, 4>(M, vectors, sorted, offsets, item_off, coords, g, grid, out, stream);
    case 5: return launch_tiled_w<T, 5>(M, vectors, sorted, offsets, item_off, coords, g, grid, out, stream);
    case 6: return launch_tiled_w<T, 6>(M, vectors, sorted, offsets, item_off, coords, g, grid, out, stream);
    case 7: return launch_tiled_w<T, 7>(M, vectors, sorted, offsets, item_off, coords, g, grid, out, stream);
    case 8: return launch_tiled_w<T, 8>(M, vectors, sorted, offsets, item_off, coords, g, grid, out, stream);
    default: set_error("tiled interpolation needs a kernel width <= 8"); return false;
    }
}
template bool interpolate_tiled3<float>(int64_t, int64_t, const uint32_t *, const uint32_t *,
                                        uint32_t *, const float *, const GridParams &,
                                        const float *, float *, hipStream_t);
template bool interpolate_tiled3<double>(int64_t, int64_t, const uint32_t *, const uint32_t *,
                                         uint32_t *, const double *, const GridParams &,
                                         const double *, double *, hipStream_t);

bool interpolate_tiled3_many(const uint32_t *sorted, const uint32_t *offsets, uint32_t *item_off,
                             int64_t nbins, const BatchProblem *probs, int nprob,
                             const int *widths, const int64_t *bin_lo, const int64_t *bin_hi,
                             const int64_t *points, int nclass, hipStream_t stream)
{
    constexpr int P = 1024;
    const int64_t n = nbins + 1;
    k_item_counts<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(nbins, (uint32_t)P, offsets,
                                                                  item_off);
    scan_u32(item_off, n, item_off + n, item_off, stream);
    for (int c = 0; c < nclass; ++c) {
        const int64_t items =
            std::min<int64_t>(bin_hi[c] - bin_lo[c] + points[c] / P + 64, 16384);
        const dim3 grid3((unsigned)std::max<int64_t>(items, 1), 1);
#define MNUFFT_TILED3_MANY(WW)                                                                    \
    k_interp_tiled3_many<float, WW, 8, 8, 8, 256, P><<<grid3, 256, 0, stream>>>(                \
        sorted, offsets, item_off, probs, nprob, bin_lo[c], bin_hi[c])
        switch (widths[c]) {
        case 2: MNUFFT_TILED3_MANY(2); break;
        case 3: MNUFFT_TILED3_MANY(3); break;
        case 4: MNUFFT_TILED3_MANY(4); break;
        case 5: MNUFFT_TILED3_MANY(5); break;
        case 6: MNUFFT_TILED3_MANY(6); break;
        case 7: MNUFFT_TILED3_MANY(7); break;
        case 8: MNUFFT_TILED3_MANY(8); break;
        default: set_error("batched tiled interpolation needs a kernel width <= 8"); return false;
        }
#undef MNUFFT_TILED3_MANY
    }
    return hip_ok(hipGetLastError(), "batched tiled 3D interpolation launch");
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
MNUFFT_INTERP_INSTANTIATE(4, float)
MNUFFT_INTERP_INSTANTIATE(5, float)
MNUFFT_INTERP_INSTANTIATE(6, float)
MNUFFT_INTERP_INSTANTIATE(7, float)
MNUFFT_INTERP_INSTANTIATE(8, float)
MNUFFT_INTERP_INSTANTIATE(4, double)
MNUFFT_INTERP_INSTANTIATE(5, double)
MNUFFT_INTERP_INSTANTIATE(6, double)
MNUFFT_INTERP_INSTANTIATE(7, double)
MNUFFT_INTERP_INSTANTIATE(8, double)

}  // namespace mnufft
