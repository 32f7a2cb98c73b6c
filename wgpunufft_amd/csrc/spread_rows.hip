// Plain (FP32 / FP64 vector) type-1 spreading of 3D row tiles (kernel
// widths 2..8 in FP32, 2..16 in FP64) and of 2D row tiles (widths 2..8), for
// one plan or for the plans of a batch.
//
// Same contract as the reference's output-stationary spreader
// (src/gpu_nd_spread.rs:1-36, 1583-1667): every cell of an 8 x 8 x 8 tile
// (x, y, z) is owned by one lane of one wavefront, which adds the records
// whose support reaches the tile in a fixed order (reach row by reach row,
// slot order inside a row) and writes the cell once; tiles with too many
// records are split into parts whose partial sums are added in part order.
// No atomics, so results are bitwise repeatable.
//
// Mapping: bins are one cell along y and z, so the records of one reach row
// share their y and z support.  Lane l owns the row of cells (y = l % 8,
// z = l / 8) across the tile's TX x cells, held in registers.  Every staged
// record carries its weights laid out over the tile's own cells (zero off
// its support), so a record costs each lane two weight reads (its y and z
// cell), one product, the record's TX x weights (broadcast vector reads) and
// TX complex FMAs on registers with static indices: no branch on the
// record's position and no dynamically indexed registers.  Records are
// staged 64 at a time (their global loads issued one chunk ahead),
// compacted to those whose x support touches the tile, with their weights
// computed once per tile visit.
#include <hip/hip_runtime.h>

#include "spread.hpp"
#include "spread_device.hpp"
#include "tile_stage.hpp"

namespace mnufft {

namespace {

constexpr int kRowsTX = 8, kRowsTY = 8, kRowsTZ = 8;

// A staged record: its weights over the tile's cells of each axis (zero
// off the support) and its strength.
template <typename T, int TX>
struct RowRec {
    T kx[TX];
    T ky[kRowsTY];
    T kz[kRowsTZ];
    T cre, cim;
};

// Weights of a record over the n cells of a tile axis from `rel` (the
// support start relative to the tile origin) and its offset.
template <typename T, int W, int N>
__device__ inline void tile_weights(T *dst, int rel, T off, T beta)
{
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const int t = i - rel;
        dst[i] = (t >= 0 && t < W) ? es_weight<T>((T)t + off, W, beta) : (T)0;
    }
}

// Sums the records [klo, khi) of the reach of the tile at (x0, y0, z0) into
// the lane's row of TX cells (ar, ai).  `offsets` are the bin offsets of the
// tile's own grid; slots index the records and strengths `cv` directly.
template <typename T, int W, int TX, int NROWS>
__device__ __forceinline__ void rows3_accumulate(RowRec<T, TX> (&st)[64], TileRows<NROWS> &rows,
                                                 int lane, const GridParams &g, int64_t x0,
                                                 int64_t y0, int64_t z0, int64_t klo, int64_t khi,
                                                 const uint32_t *__restrict__ offsets,
                                                 const int4 *__restrict__ rec_start,
                                                 const T *__restrict__ rec_off,
                                                 const T *__restrict__ cv, T (&ar)[TX],
                                                 T (&ai)[TX])
{
    constexpr int TY = kRowsTY, TZ = kRowsTZ;
    constexpr int NRY = TY + W - 1;
    const int ly = lane & 7, lz = lane >> 3;  // the lane's row of the tile
    const T beta = (T)g.beta;
    const int64_t n0 = g.ax[0].n;
    const AxisReach rx = axis_reach(x0, TX, W, g.ax[0]);
    int64_t xs = x0 - (W - 1);
    if (xs < 0) xs += n0;

    stage_tile_rows<3, TY, TZ, W, NROWS>(rows, lane, y0, z0, rx, g, offsets);
#pragma unroll
    for (int i = 0; i < TX; ++i) ar[i] = ai[i] = (T)0;
    // records in chunks of 64, the next chunk's loads in flight while the
    // current one is added
    Fetched<T> nx = fetch_record<T, NROWS>(rows, klo + lane, khi, rec_start, rec_off, cv);
    for (int64_t k0 = klo; k0 < khi; k0 += 64) {
        const Fetched<T> cur = nx;
        nx = fetch_record<T, NROWS>(rows, k0 + 64 + lane, khi, rec_start, rec_off, cv);
        bool take = false;
        int ee = 0;
        if (cur.valid) {
            int64_t e64 = (int64_t)cur.sx0 - xs;
            if (e64 < 0) e64 += n0;
            take = e64 < TX + W - 1;
            ee = (int)e64;
        }
        const uint64_t mask = __ballot(take);
        const int pos = __popcll(mask & ((1ull << lane) - 1ull));
        const int nst = __popcll(mask);
        if (take) {
            RowRec<T, TX> &R = st[pos];
            tile_weights<T, W, TX>(R.kx, ee - (W - 1), (T)cur.off.x, beta);
            tile_weights<T, W, TY>(R.ky, cur.q % NRY - (W - 1), (T)cur.off.y, beta);
            tile_weights<T, W, TZ>(R.kz, cur.q / NRY - (W - 1), (T)cur.off.z, beta);
            R.cre = cur.c.x;
            R.cim = cur.c.y;
        }
        wave_sync();
#pragma unroll 2
        for (int s = 0; s < nst; ++s) {
            const RowRec<T, TX> &S = st[s];
            const T wyz = S.ky[ly] * S.kz[lz];
            const T vre = S.cre * wyz, vim = S.cim * wyz;
#pragma unroll
            for (int i = 0; i < TX; ++i) {
                const T kx = S.kx[i];
                ar[i] += vre * kx;
                ai[i] += vim * kx;
            }
        }
        wave_sync();
    }
}

// The lane's row of TX cells: into the grid `gv` (one vector) when the tile
// is one part, else into the part's partial-sum slot `slot` (TX x 8 x 8).
template <typename T, int TX>
__device__ __forceinline__ void rows3_store(int lane, const GridParams &g, int64_t x0, int64_t y0,
                                            int64_t z0, bool whole, T *__restrict__ gv,
                                            T *__restrict__ slot, const T (&ar)[TX],
                                            const T (&ai)[TX])
{
    const int ly = lane & 7, lz = lane >> 3;
    const int64_t n0 = g.ax[0].n, n1 = g.ax[1].n;
    if (whole) {
        const int64_t gy = y0 + ly, gz = z0 + lz;
        if (gy < n1 && gz < g.ax[2].n) {
            T *row = gv + 2 * (x0 + n0 * (gy + n1 * gz));
#pragma unroll
            for (int i = 0; i < TX; ++i) {
                if (x0 + i < n0) {
                    row[2 * i] = ar[i];
                    row[2 * i + 1] = ai[i];
                }
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < TX; ++i) {
            T *o = slot + partial_index<TX, kRowsTY, kRowsTZ>(0, 1, 0, i, ly, lz);
            o[0] = ar[i];
            o[1] = ai[i];
        }
    }
}

template <typename T, int W, int TX>
__global__ __launch_bounds__(64) void k_spread_rows(
    GridParams g, SpreadParams sp, const uint32_t *__restrict__ offsets,
    const int4 *__restrict__ rec_start, const T *__restrict__ rec_off,
    const T *__restrict__ c_sorted, int64_t M, const uint32_t *__restrict__ totals,
    const uint32_t *__restrict__ parts, const uint32_t *__restrict__ item_off,
    const uint32_t *__restrict__ heavy_base, const uint32_t *__restrict__ ctrl,
    T *__restrict__ grid, T *__restrict__ partial)
{
    static_assert(W <= kMaxRowsWidth, "widths beyond the row tiles");
    constexpr int TY = kRowsTY, TZ = kRowsTZ;
    constexpr int NROWS = (TY + W - 1) * (TZ + W - 1);
    __shared__ RowRec<T, TX> st[64];
    __shared__ TileRows<NROWS> rows;

    const int lane = threadIdx.x;
    const int64_t v = blockIdx.y;
    const int64_t n_items = ctrl[0];
    const uint32_t part_len_cap = ctrl[3];
    const T *cv = c_sorted + 2 * v * M;

    for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const TileItem ti = tile_item(item, sp.block_count, item_off, parts, totals, part_len_cap);
        const int64_t b = ti.b;
        const int64_t x0 = (b % sp.nblocks[0]) * TX;
        const int64_t y0 = ((b / sp.nblocks[0]) % sp.nblocks[1]) * TY;
        const int64_t z0 = (b / (sp.nblocks[0] * sp.nblocks[1])) * TZ;
        T ar[TX], ai[TX];
        rows3_accumulate<T, W, TX, NROWS>(st, rows, lane, g, x0, y0, z0, ti.rlo, ti.rhi, offsets,
                                          rec_start, rec_off, cv, ar, ai);
        const bool whole = ti.np == 1;
        T *slot = whole ? nullptr
                        : partial + partial_index<TX, TY, TZ>((int64_t)heavy_base[b] + ti.p,
                                                              gridDim.y, v, 0, 0, 0);
        rows3_store<T, TX>(lane, g, x0, y0, z0, whole, grid + 2 * v * g.fine_total, slot, ar,
                           ai);
    }
}

// The 3D row tiles of every problem of a batch class (one vector each, one
// kernel width), as k_spread_rows2_many does for 2D.
template <typename T, int W>
__global__ __launch_bounds__(64) void k_spread_rows3_many(
    const BatchProblem *__restrict__ probs, const SpreadProblem *__restrict__ sprobs, int nprob,
    const uint32_t *__restrict__ offsets, const int4 *__restrict__ rec_start,
    const T *__restrict__ rec_off, const T *__restrict__ c_sorted,
    const uint32_t *__restrict__ totals, const uint32_t *__restrict__ parts,
    const uint32_t *__restrict__ item_off, const uint32_t *__restrict__ heavy_base,
    const uint32_t *__restrict__ pctrl, const int64_t *__restrict__ pbase,
    T *__restrict__ partial)
{
    static_assert(W <= kMaxRowsWidth, "widths beyond the row tiles");
    constexpr int TX = kRowsTX, TY = kRowsTY, TZ = kRowsTZ;
    constexpr int NROWS = (TY + W - 1) * (TZ + W - 1);
    __shared__ RowRec<T, TX> st[64];
    __shared__ TileRows<NROWS> rows;

    const int lane = threadIdx.x;
    const int64_t n_items = pbase[4 * nprob];
    for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const int q = batch_problem_of(pbase, nprob, 0, item);
        const BatchProblem &pb = probs[q];
        const SpreadProblem &spb = sprobs[q];
        const int64_t bb = spb.block_base;
        const TileItem ti = tile_item(item - pbase[4 * q], spb.sp.block_count, item_off + bb,
                                      parts + bb, totals + bb, pctrl[8 * q + 3]);
        const int64_t b = ti.b;
        const int64_t nb0 = spb.sp.nblocks[0], nb1 = spb.sp.nblocks[1];
        const int64_t x0 = (b % nb0) * TX, y0 = ((b / nb0) % nb1) * TY, z0 = (b / (nb0 * nb1)) * TZ;
        T ar[TX], ai[TX];
        rows3_accumulate<T, W, TX, NROWS>(st, rows, lane, pb.g, x0, y0, z0, ti.rlo, ti.rhi,
                                          offsets + pb.bin_base, rec_start, rec_off, c_sorted,
                                          ar, ai);
        const bool whole = ti.np == 1;
        T *slot = whole ? nullptr
                        : partial + partial_index<TX, TY, TZ>(
                                        pbase[4 * q + 1] + (int64_t)heavy_base[bb + b] + ti.p, 1,
                                        0, 0, 0, 0);
        rows3_store<T, TX>(lane, pb.g, x0, y0, z0, whole, (T *)pb.grid, slot, ar, ai);
    }
}

// 2D row tiles of 16 x 16 cells: lane l owns the four x cells 4 (l / 16) ..
// + 3 of row y = l % 16, in registers.  Bins are one cell along y, so the
// records of a reach row share their y support; a staged record carries its
// weights over the tile's 16 x and 16 y cells (zero off its support), so a
// record costs each lane one weight read, one product, a four-weight read
// and eight FMAs.  Same order, parts and partial-sum layout as the 3D tiles.
constexpr int kRows2T = 16;

template <typename T>
struct RowRec2 {
    T kx[kRows2T];
    T ky[kRows2T];
    T cre, cim;
};

// Sums the records [klo, khi) of the reach of the tile at (x0, y0) into the
// lane's four cells (ar, ai).  `offsets` are the bin offsets of the tile's
// own grid; slots index the records and strengths `cv` directly.
template <typename T, int W, int NROWS>
__device__ __forceinline__ void rows2_accumulate(RowRec2<T> (&st)[64], TileRows<NROWS> &rows,
                                                 int lane, const GridParams &g, int64_t x0,
                                                 int64_t y0, int64_t klo, int64_t khi,
                                                 const uint32_t *__restrict__ offsets,
                                                 const int4 *__restrict__ rec_start,
                                                 const T *__restrict__ rec_off,
                                                 const T *__restrict__ cv, T (&ar)[4], T (&ai)[4])
{
    constexpr int TX = kRows2T, TY = kRows2T;
    const int ly = lane & 15, xc = 4 * (lane >> 4);
    const T beta = (T)g.beta;
    const int64_t n0 = g.ax[0].n;
    const AxisReach rx = axis_reach(x0, TX, W, g.ax[0]);
    int64_t xs = x0 - (W - 1);
    if (xs < 0) xs += n0;

    stage_tile_rows<2, TY, 1, W, NROWS>(rows, lane, y0, 0, rx, g, offsets);
#pragma unroll
    for (int i = 0; i < 4; ++i) ar[i] = ai[i] = (T)0;
    Fetched<T> nx = fetch_record<T, NROWS>(rows, klo + lane, khi, rec_start, rec_off, cv);
    for (int64_t k0 = klo; k0 < khi; k0 += 64) {
        const Fetched<T> cur = nx;
        nx = fetch_record<T, NROWS>(rows, k0 + 64 + lane, khi, rec_start, rec_off, cv);
        bool take = false;
        int ee = 0;
        if (cur.valid) {
            int64_t e64 = (int64_t)cur.sx0 - xs;
            if (e64 < 0) e64 += n0;
            take = e64 < TX + W - 1;
            ee = (int)e64;
        }
        const uint64_t mask = __ballot(take);
        const int pos = __popcll(mask & ((1ull << lane) - 1ull));
        const int nst = __popcll(mask);
        if (take) {
            RowRec2<T> &R = st[pos];
            tile_weights<T, W, TX>(R.kx, ee - (W - 1), (T)cur.off.x, beta);
            tile_weights<T, W, TY>(R.ky, cur.q - (W - 1), (T)cur.off.y, beta);
            R.cre = cur.c.x;
            R.cim = cur.c.y;
        }
        wave_sync();
#pragma unroll 2
        for (int s = 0; s < nst; ++s) {
            const RowRec2<T> &S = st[s];
            const T wy = S.ky[ly];
            const T vre = S.cre * wy, vim = S.cim * wy;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const T kx = S.kx[xc + i];
                ar[i] += vre * kx;
                ai[i] += vim * kx;
            }
        }
        wave_sync();
    }
}

// The lane's four cells: into the grid `gv` (one vector) when the tile is
// one part, else into the part's 16 x 16 partial-sum slot `slot`.
template <typename T>
__device__ __forceinline__ void rows2_store(int lane, const GridParams &g, int64_t x0, int64_t y0,
                                            bool whole, T *__restrict__ gv, T *__restrict__ slot,
                                            const T (&ar)[4], const T (&ai)[4])
{
    const int ly = lane & 15, xc = 4 * (lane >> 4);
    const int64_t n0 = g.ax[0].n, gy = y0 + ly;
    if (whole) {
        if (gy < g.ax[1].n) {
            T *row = gv + 2 * (x0 + xc + n0 * gy);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (x0 + xc + i < n0) {
                    row[2 * i] = ar[i];
                    row[2 * i + 1] = ai[i];
                }
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            T *o = slot + 2 * (ly * kRows2T + xc + i);
            o[0] = ar[i];
            o[1] = ai[i];
        }
    }
}

// The single-plan 2D kernel keeps its loop written out in full: built on
// rows2_accumulate / rows2_store it ran 7% slower (3.95 vs 3.68 ms on 5e7
// points over 2048^2 cells, width 7).
template <typename T, int W>
__global__ __launch_bounds__(64) void k_spread_rows2(
    GridParams g, SpreadParams sp, const uint32_t *__restrict__ offsets,
    const int4 *__restrict__ rec_start, const T *__restrict__ rec_off,
    const T *__restrict__ c_sorted, int64_t M, const uint32_t *__restrict__ totals,
    const uint32_t *__restrict__ parts, const uint32_t *__restrict__ item_off,
    const uint32_t *__restrict__ heavy_base, const uint32_t *__restrict__ ctrl,
    T *__restrict__ grid, T *__restrict__ partial)
{
    static_assert(W <= 8, "widths beyond the 2D row tiles");
    constexpr int TX = kRows2T, TY = kRows2T;
    constexpr int NROWS = TY + W - 1;
    __shared__ RowRec2<T> st[64];
    __shared__ TileRows<NROWS> rows;

    const int lane = threadIdx.x;
    const int ly = lane & 15, xc = 4 * (lane >> 4);  // the lane's row and first x cell
    const int64_t v = blockIdx.y;
    const T beta = (T)g.beta;
    const int64_t n_items = ctrl[0];
    const uint32_t part_len_cap = ctrl[3];
    const T *cv = c_sorted + 2 * v * M;
    const int64_t n0 = g.ax[0].n, n1 = g.ax[1].n;

    for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const TileItem ti = tile_item(item, sp.block_count, item_off, parts, totals, part_len_cap);
        const int64_t b = ti.b;
        const int64_t x0 = (b % sp.nblocks[0]) * TX, y0 = (b / sp.nblocks[0]) * TY;
        const AxisReach rx = axis_reach(x0, TX, W, g.ax[0]);
        int64_t xs = x0 - (W - 1);
        if (xs < 0) xs += n0;

        stage_tile_rows<2, TY, 1, W, NROWS>(rows, lane, y0, 0, rx, g, offsets);
        T ar[4], ai[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) ar[i] = ai[i] = (T)0;
        const int64_t klo = ti.rlo, khi = ti.rhi;
        Fetched<T> nx = fetch_record<T, NROWS>(rows, klo + lane, khi, rec_start, rec_off, cv);
        for (int64_t k0 = klo; k0 < khi; k0 += 64) {
            const Fetched<T> cur = nx;
            nx = fetch_record<T, NROWS>(rows, k0 + 64 + lane, khi, rec_start, rec_off, cv);
            bool take = false;
            int ee = 0;
            if (cur.valid) {
                int64_t e64 = (int64_t)cur.sx0 - xs;
                if (e64 < 0) e64 += n0;
                take = e64 < TX + W - 1;
                ee = (int)e64;
            }
            const uint64_t mask = __ballot(take);
            const int pos = __popcll(mask & ((1ull << lane) - 1ull));
            const int nst = __popcll(mask);
            if (take) {
                RowRec2<T> &R = st[pos];
                tile_weights<T, W, TX>(R.kx, ee - (W - 1), (T)cur.off.x, beta);
                tile_weights<T, W, TY>(R.ky, cur.q - (W - 1), (T)cur.off.y, beta);
                R.cre = cur.c.x;
                R.cim = cur.c.y;
            }
            wave_sync();
#pragma unroll 2
            for (int s = 0; s < nst; ++s) {
                const RowRec2<T> &S = st[s];
                const T wy = S.ky[ly];
                const T vre = S.cre * wy, vim = S.cim * wy;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const T kx = S.kx[xc + i];
                    ar[i] += vre * kx;
                    ai[i] += vim * kx;
                }
            }
            wave_sync();
        }

        const int64_t gy = y0 + ly;
        if (ti.np == 1) {
            if (gy < n1) {
                T *row = grid + 2 * (v * g.fine_total + x0 + xc + n0 * gy);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if (x0 + xc + i < n0) {
                        row[2 * i] = ar[i];
                        row[2 * i + 1] = ai[i];
                    }
                }
            }
        } else {
            const int64_t pslot = (int64_t)heavy_base[b] + ti.p;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                T *o = partial + partial_index<TX, TY, 1>(pslot, gridDim.y, v, xc + i, ly, 0);
                o[0] = ar[i];
                o[1] = ai[i];
            }
        }
    }
}

// The 2D row tiles of every problem of a batch class (one vector each, one
// kernel width): item -> problem (item bases) -> the problem's tile and
// part, summed as the problem's own spread sums it.
template <typename T, int W>
__global__ __launch_bounds__(64) void k_spread_rows2_many(
    const BatchProblem *__restrict__ probs, const SpreadProblem *__restrict__ sprobs, int nprob,
    const uint32_t *__restrict__ offsets, const int4 *__restrict__ rec_start,
    const T *__restrict__ rec_off, const T *__restrict__ c_sorted,
    const uint32_t *__restrict__ totals, const uint32_t *__restrict__ parts,
    const uint32_t *__restrict__ item_off, const uint32_t *__restrict__ heavy_base,
    const uint32_t *__restrict__ pctrl, const int64_t *__restrict__ pbase,
    T *__restrict__ partial)
{
    static_assert(W <= 8, "widths beyond the 2D row tiles");
    constexpr int NROWS = kRows2T + W - 1;
    __shared__ RowRec2<T> st[64];
    __shared__ TileRows<NROWS> rows;

    const int lane = threadIdx.x;
    const int64_t n_items = pbase[4 * nprob];
    for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const int q = batch_problem_of(pbase, nprob, 0, item);
        const BatchProblem &pb = probs[q];
        const SpreadProblem &spb = sprobs[q];
        const int64_t bb = spb.block_base;
        const TileItem ti = tile_item(item - pbase[4 * q], spb.sp.block_count, item_off + bb,
                                      parts + bb, totals + bb, pctrl[8 * q + 3]);
        const int64_t b = ti.b;
        const int64_t x0 = (b % spb.sp.nblocks[0]) * kRows2T;
        const int64_t y0 = (b / spb.sp.nblocks[0]) * kRows2T;
        T ar[4], ai[4];
        rows2_accumulate<T, W, NROWS>(st, rows, lane, pb.g, x0, y0, ti.rlo, ti.rhi,
                                      offsets + pb.bin_base, rec_start, rec_off, c_sorted, ar,
                                      ai);
        const bool whole = ti.np == 1;
        T *slot = whole ? nullptr
                        : partial + partial_index<kRows2T, kRows2T, 1>(
                                        pbase[4 * q + 1] + (int64_t)heavy_base[bb + b] + ti.p, 1,
                                        0, 0, 0, 0);
        rows2_store<T>(lane, pb.g, x0, y0, whole, (T *)pb.grid, slot, ar, ai);
    }
}

template <typename T, int W>
bool launch_rows_width(int tx, dim3 grid3, const GridParams &g, const SpreadParams &sp,
                       const uint32_t *offsets, const int4 *rec_start, const T *rec_off,
                       const T *c_sorted, int64_t M, const TileWork &w, T *grid, T *partial,
                       hipStream_t stream)
{
    if (tx != kRowsTX) return false;
    k_spread_rows<T, W, kRowsTX><<<grid3, 64, 0, stream>>>(
        g, sp, offsets, rec_start, rec_off, c_sorted, M, w.totals, w.parts, w.item_off,
        w.heavy_base, w.ctrl, grid, partial);
    return true;
}

}  // namespace

template <typename T>
bool launch_rows2_tiles(const GridParams &g, const SpreadParams &sp, const uint32_t *offsets,
                        const int4 *rec_start, const T *rec_off, const T *c_sorted, int64_t M,
                        int64_t vectors, const TileWork &w, T *grid, T *partial,
                        hipStream_t stream)
{
    const dim3 grid3(w.items_grid, (unsigned)vectors);
    if (sp.block[0] != kRows2T || sp.block[1] != kRows2T) {
        set_error("plain 2D row tiles are 16 x 16 cells");
        return false;
    }
#define MNUFFT_ROWS2(W)                                                                             \
    case W:                                                                                         \
        k_spread_rows2<T, W><<<grid3, 64, 0, stream>>>(g, sp, offsets, rec_start, rec_off,          \
                                                       c_sorted, M, w.totals, w.parts, w.item_off,  \
                                                       w.heavy_base, w.ctrl, grid, partial);        \
        break;
    switch (g.width) {
        MNUFFT_ROWS2(2)
        MNUFFT_ROWS2(3)
        MNUFFT_ROWS2(4)
        MNUFFT_ROWS2(5)
        MNUFFT_ROWS2(6)
        MNUFFT_ROWS2(7)
        MNUFFT_ROWS2(8)
    default:
        set_error("plain 2D row tiles need a width of 2..8");
        return false;
    }
#undef MNUFFT_ROWS2
    return hip_ok(hipGetLastError(), "2D row-tile spread launch");
}

template <typename T>
bool launch_rows2_many(int width, const BatchProblem *probs, const SpreadProblem *sprobs,
                       int nprob, unsigned items_grid, const uint32_t *offsets,
                       const int4 *rec_start, const T *rec_off, const T *c_sorted,
                       const uint32_t *totals, const uint32_t *parts, const uint32_t *item_off,
                       const uint32_t *heavy_base, const uint32_t *pctrl, const int64_t *pbase,
                       T *partial, hipStream_t stream)
{
#define MNUFFT_ROWS2_MANY(W)                                                                        \
    case W:                                                                                         \
        k_spread_rows2_many<T, W><<<items_grid, 64, 0, stream>>>(                                   \
            probs, sprobs, nprob, offsets, rec_start, rec_off, c_sorted, totals, parts, item_off,   \
            heavy_base, pctrl, pbase, partial);                                                     \
        break;
    switch (width) {
        MNUFFT_ROWS2_MANY(2)
        MNUFFT_ROWS2_MANY(3)
        MNUFFT_ROWS2_MANY(4)
        MNUFFT_ROWS2_MANY(5)
        MNUFFT_ROWS2_MANY(6)
        MNUFFT_ROWS2_MANY(7)
        MNUFFT_ROWS2_MANY(8)
    default:
        set_error("batched 2D row tiles need a width of 2..8");
        return false;
    }
#undef MNUFFT_ROWS2_MANY
    return hip_ok(hipGetLastError(), "batched 2D row-tile spread launch");
}

template bool launch_rows2_many<float>(int, const BatchProblem *, const SpreadProblem *, int,
                                       unsigned, const uint32_t *, const int4 *, const float *,
                                       const float *, const uint32_t *, const uint32_t *,
                                       const uint32_t *, const uint32_t *, const uint32_t *,
                                       const int64_t *, float *, hipStream_t);
template bool launch_rows2_many<double>(int, const BatchProblem *, const SpreadProblem *, int,
                                        unsigned, const uint32_t *, const int4 *, const double *,
                                        const double *, const uint32_t *, const uint32_t *,
                                        const uint32_t *, const uint32_t *, const uint32_t *,
                                        const int64_t *, double *, hipStream_t);

template <typename T>
bool launch_rows3_many(int width, const BatchProblem *probs, const SpreadProblem *sprobs,
                       int nprob, unsigned items_grid, const uint32_t *offsets,
                       const int4 *rec_start, const T *rec_off, const T *c_sorted,
                       const uint32_t *totals, const uint32_t *parts, const uint32_t *item_off,
                       const uint32_t *heavy_base, const uint32_t *pctrl, const int64_t *pbase,
                       T *partial, hipStream_t stream)
{
    bool ok = true;
#define MNUFFT_ROWS3_MANY(W)                                                                        \
    case W:                                                                                         \
        k_spread_rows3_many<T, W><<<items_grid, 64, 0, stream>>>(                                   \
            probs, sprobs, nprob, offsets, rec_start, rec_off, c_sorted, totals, parts, item_off,   \
            heavy_base, pctrl, pbase, partial);                                                     \
        break;
    switch (width) {
        MNUFFT_ROWS3_MANY(2)
        MNUFFT_ROWS3_MANY(3)
        MNUFFT_ROWS3_MANY(4)
        MNUFFT_ROWS3_MANY(5)
        MNUFFT_ROWS3_MANY(6)
        MNUFFT_ROWS3_MANY(7)
        MNUFFT_ROWS3_MANY(8)
    default:
        ok = false;
        if constexpr (sizeof(T) == 8) {
            ok = true;
            switch (width) {
                MNUFFT_ROWS3_MANY(9)
                MNUFFT_ROWS3_MANY(10)
                MNUFFT_ROWS3_MANY(11)
                MNUFFT_ROWS3_MANY(12)
                MNUFFT_ROWS3_MANY(13)
                MNUFFT_ROWS3_MANY(14)
                MNUFFT_ROWS3_MANY(15)
                MNUFFT_ROWS3_MANY(16)
            default: ok = false; break;
            }
        }
        break;
    }
#undef MNUFFT_ROWS3_MANY
    if (!ok) {
        set_error("batched 3D row tiles need a width of 2..8 (FP32) or 2..16 (FP64)");
        return false;
    }
    return hip_ok(hipGetLastError(), "batched 3D row-tile spread launch");
}

template bool launch_rows3_many<float>(int, const BatchProblem *, const SpreadProblem *, int,
                                       unsigned, const uint32_t *, const int4 *, const float *,
                                       const float *, const uint32_t *, const uint32_t *,
                                       const uint32_t *, const uint32_t *, const uint32_t *,
                                       const int64_t *, float *, hipStream_t);
template bool launch_rows3_many<double>(int, const BatchProblem *, const SpreadProblem *, int,
                                        unsigned, const uint32_t *, const int4 *, const double *,
                                        const double *, const uint32_t *, const uint32_t *,
                                        const uint32_t *, const uint32_t *, const uint32_t *,
                                        const int64_t *, double *, hipStream_t);

template <typename T>
bool launch_rows_tiles(const GridParams &g, const SpreadParams &sp, const uint32_t *offsets,
                       const int4 *rec_start, const T *rec_off, const T *c_sorted, int64_t M,
                       int64_t vectors, const TileWork &w, T *grid, T *partial, hipStream_t stream)
{
    if (g.dims == 2)
        return launch_rows2_tiles<T>(g, sp, offsets, rec_start, rec_off, c_sorted, M, vectors, w,
                                     grid, partial, stream);
    const dim3 grid3(w.items_grid, (unsigned)vectors);
    const int tx = (int)sp.block[0];
    if (sp.block[1] != kRowsTY || sp.block[2] != kRowsTZ) {
        set_error("plain row tiles are 8 x 8 x 8 cells");
        return false;
    }
    bool ok = false;
    switch (g.width) {
    case 2: ok = launch_rows_width<T, 2>(tx, grid3, g, sp, offsets, rec_start, rec_off, c_sorted, M, w, grid, partial, stream); break;
    case 3: ok = launch_rows_width<T, 3>(tx, grid3, g, sp, offsets, rec_start, rec_off, c_sorted, M, w, grid, partial, stream); break;
    case 4: ok = launch_rows_width<T, 4>(tx, grid3, g, sp, offsets, rec_start, rec_off, c_sorted, M, w, grid, partial, stream); break;
    case 5: ok = launch_rows_width<T, 5>(tx, grid3, g, sp, offsets, rec_start, rec_off, c_sorted, M, w, grid, partial, stream); break;
    case 6: ok = launch_rows_width<T, 6>(tx, grid3, g, sp, offsets, rec_start, rec_off, c_sorted, M, w, grid, partial, stream); break;
    case 7: ok = launch_rows_width<T, 7>(tx, grid3, g, sp, offsets, rec_start, rec_off, c_sorted, M, w, grid, partial, stream); break;
    case 8: ok = launch_rows_width<T, 8>(tx, grid3, g, sp, offsets, rec_start, rec_off, c_sorted, M, w, grid, partial, stream); break;
    default:
        if constexpr (sizeof(T) == 8) {
            switch (g.width) {
            case 9: ok = launch_rows_width<T, 9>(tx, grid3, g, sp, offsets, rec_start, rec_off, c_sorted, M, w, grid, partial, stream); break;
            case 10: ok = launch_rows_width<T, 10>(tx, grid3, g, sp, offsets, rec_start, rec_off, c_sorted, M, w, grid, partial, stream); break;
            case 11: ok = launch_rows_width<T, 11>(tx, grid3, g, sp, offsets, rec_start, rec_off, c_sorted, M, w, grid, partial, stream); break;
            case 12: ok = launch_rows_width<T, 12>(tx, grid3, g, sp, offsets, rec_start, rec_off, c_sorted, M, w, grid, partial, stream); break;
            case 13: ok = launch_rows_width<T, 13>(tx, grid3, g, sp, offsets, rec_start, rec_off, c_sorted, M, w, grid, partial, stream); break;
            case 14: ok = launch_rows_width<T, 14>(tx, grid3, g, sp, offsets, rec_start, rec_off, c_sorted, M, w, grid, partial, stream); break;
            case 15: ok = launch_rows_width<T, 15>(tx, grid3, g, sp, offsets, rec_start, rec_off, c_sorted, M, w, grid, partial, stream); break;
            case 16: ok = launch_rows_width<T, 16>(tx, grid3, g, sp, offsets, rec_start, rec_off, c_sorted, M, w, grid, partial, stream); break;
            default: break;
            }
        }
        break;
    }
    if (!ok) {
        set_error("plain row tiles need a 3D plan, a width of 2..8 (FP32) or 2..16 (FP64) and "
                  "8 x 8 x 8 tiles");
        return false;
    }
    return hip_ok(hipGetLastError(), "row-tile spread launch");
}

template bool launch_rows_tiles<float>(const GridParams &, const SpreadParams &, const uint32_t *,
                                       const int4 *, const float *, const float *, int64_t,
                                       int64_t, const TileWork &, float *, float *, hipStream_t);
template bool launch_rows_tiles<double>(const GridParams &, const SpreadParams &,
                                        const uint32_t *, const int4 *, const double *,
                                        const double *, int64_t, int64_t, const TileWork &,
                                        double *, double *, hipStream_t);

}  // namespace mnufft
