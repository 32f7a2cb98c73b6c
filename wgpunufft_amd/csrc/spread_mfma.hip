// Type-1 spreading of 3D tiles as dense contractions on the matrix cores.
//
// Same contract as the plain spreaders (src/gpu_nd_spread.rs:1-36): every
// fine-grid cell of a 16 x 16 x 8 tile (x, y, z) is owned by one workgroup, which sums the
// records whose support reaches the tile in a fixed order and writes the cell
// once; tiles with too many records are split into parts whose partial sums
// are added in part order.  Results are bitwise repeatable and do not depend
// on the input order of the points (the records are visited in canonical
// order).
//
// The tile's sum is recast as a matrix product over its records j:
//
//   G[(y,z), (x,re|im)] += sum_j  Ky_j[y] Kz_j[z]  *  Kx_j[x] c_j
//                          \_____ A (128 x K) ____/  \__ B (K x 32) __/
//
// with Kx_j, Ky_j, Kz_j the record's separable kernel weights laid out over
// the tile's cells of each axis (zero off the support).  The 128 (y,z)
// rows form four 32-row blocks (16 y x 2 z), each held in registers as the
// accumulator of a v_mfma_f32_32x32x16 product; B's 32 columns are the 16 x
// cells, real and imaginary parts interleaved, so the accumulator rows are
// written out as contiguous 128-byte rows of the interleaved complex grid.
// Records arrive in row order (z start, then y start), so the 16 records of
// a k-block touch a narrow band of z and the row blocks outside that band
// are skipped.
//
// Precision: A and B are formed in FP32 and rounded to FP16 (one product,
// relative error about 2^-11), or split into hi + lo halves in FP16 or
// bfloat16 with the three products hi*hi + hi*lo + lo*hi accumulated in FP32
// (relative error about 2^-21 for FP16, 2^-16 for bfloat16).  Strengths are
// scaled by a power of two so that the largest magnitude is below one, which
// keeps the FP16 operands in range; the scale is undone exactly on output.
#include <hip/hip_runtime.h>

#include "mfma_ops.hpp"
#include "spread.hpp"
#include "spread_device.hpp"
#include "tile_stage.hpp"

namespace mnufft {

namespace {

using namespace mfma;

constexpr int kTile = 16;
constexpr int kPitch = 20;  // LDS pitch of the weight tables, in floats

struct MfmaRec {
    int sx, sy, sz;  // support start relative to the tile origin
    float ox, oy, oz;
    float cre, cim;  // scaled strength
};

// Largest |re|, |im| of each vector's strengths (blockIdx.y), as float bits
// (NaN above Inf).
__global__ __launch_bounds__(256) void k_absmax(const float *__restrict__ c, int64_t n,
                                                uint32_t *__restrict__ out)
{
    const float *cv = c + (int64_t)blockIdx.y * n;
    uint32_t m = 0;
    const int64_t n4 = n / 4;  // n is even (complex pairs); the tail is < 4
    const float4 *c4 = (const float4 *)cv;
    const bool aligned = ((uintptr_t)cv & 15) == 0;
    if (aligned) {
#pragma unroll 4
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
             i += (int64_t)gridDim.x * blockDim.x) {
            const float4 q = c4[i];
            const uint32_t a = __float_as_uint(q.x) & 0x7fffffffu;
            const uint32_t b = __float_as_uint(q.y) & 0x7fffffffu;
            const uint32_t d = __float_as_uint(q.z) & 0x7fffffffu;
            const uint32_t e = __float_as_uint(q.w) & 0x7fffffffu;
            const uint32_t ab = a > b ? a : b, de = d > e ? d : e;
            const uint32_t u = ab > de ? ab : de;
            m = u > m ? u : m;
        }
    }
    for (int64_t i = (aligned ? 4 * n4 : 0) + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
         i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t u = __float_as_uint(cv[i]) & 0x7fffffffu;
        m = u > m ? u : m;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t y = __shfl_xor(m, o, 64);
        m = y > m ? y : m;
    }
    // one atomic per workgroup: atomics on one address serialize
    __shared__ uint32_t wmax[4];
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t b = wmax[0];
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) b = wmax[i] > b ? wmax[i] : b;
        if (b) atomicMax(out + blockIdx.y, b);
    }
}

template <int W, int MODE, int TZ>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void
k_spread_mfma3(
    GridParams g, SpreadParams sp, const uint32_t *__restrict__ offsets,
    const int4 *__restrict__ rec_start, const float *__restrict__ rec_off,
    const float *__restrict__ c_sorted, int64_t M, const uint32_t *__restrict__ totals,
    const uint32_t *__restrict__ parts, const uint32_t *__restrict__ item_off,
    const uint32_t *__restrict__ heavy_base, const uint32_t *__restrict__ ctrl,
    const uint32_t *__restrict__ cmax, float *__restrict__ grid, float *__restrict__ partial)
{
    using Op = Operand<MODE>;
    using V = typename Op::V;
    constexpr bool kSplit = MODE != (int)SpreadMethod::MfmaF16;
    constexpr int NRY = kTile + W - 1;
    constexpr int NROWS = NRY * (TZ + W - 1);
    constexpr int kBlocks = TZ / 2;  // 32-row blocks (16 y x 2 z) of the tile
    // one spare row per table takes the weights that fall outside the tile
    __shared__ __align__(16) float KX[(kTile + 1) * kPitch];
    __shared__ __align__(16) float KY[(kTile + 1) * kPitch];
    __shared__ __align__(16) float KZ[(kTile + 1) * kPitch];
    __shared__ __align__(16) float CC[2 * kPitch];
    __shared__ MfmaRec st[64];
    __shared__ TileRows<NROWS> rows;

    const int lane = threadIdx.x;
    const int r32 = lane & 31;
    const int half = lane >> 5;
    const int j0 = 8 * half;   // first of the lane's 8 records in a k-block
    const int ay = r32 & 15;   // A row: y cell, and z offset within the pair
    const int az = r32 >> 4;
    const int bxc = r32 >> 1;  // B column: x cell, real (0) or imaginary (1)
    const int bpart = r32 & 1;

    float scale, unscale;
    strength_scale(cmax + blockIdx.y, &scale, &unscale);

    const int64_t v = blockIdx.y;
    const float two_over_w = 2.0f / (float)W;
    const float beta_l2e = (float)(g.beta * 1.4426950408889634);
    const int64_t n_items = ctrl[0];
    const uint32_t part_len_cap = ctrl[3];
    const float *cv = c_sorted + 2 * v * M;
    const int64_t n0 = g.ax[0].n, n1 = g.ax[1].n, n2 = g.ax[2].n;

    for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const TileItem ti = tile_item(item, sp.block_count, item_off, parts, totals, part_len_cap);
        const int64_t b = ti.b, p = ti.p, rlo = ti.rlo, rhi = ti.rhi;
        const uint32_t np = ti.np;
        const int64_t tbx = b % sp.nblocks[0];
        const int64_t tby = (b / sp.nblocks[0]) % sp.nblocks[1];
        const int64_t tbz = b / (sp.nblocks[0] * sp.nblocks[1]);
        const int64_t x0 = tbx * kTile, y0 = tby * kTile, z0 = tbz * TZ;
        const AxisReach rx = axis_reach(x0, kTile, W, g.ax[0]);
        int64_t xs = x0 - (W - 1);
        if (xs < 0) xs += n0;

        stage_tile_rows<3, kTile, TZ, W, NROWS>(rows, lane, y0, z0, rx, g, offsets);
        f32x16 acc[kBlocks];
#pragma unroll
        for (int m = 0; m < kBlocks; ++m) acc[m] = f32x16{};
        // records in chunks of 64, the next chunk's loads in flight while
        // the current one is contracted
        const int64_t klo = rlo, khi = rhi;
        Fetched<float> nx = fetch_record<float, NROWS>(rows, klo + lane, khi, rec_start, rec_off, cv);
        for (int64_t k0 = klo; k0 < khi; k0 += 64) {
            const Fetched<float> cur = nx;
            nx = fetch_record<float, NROWS>(rows, k0 + 64 + lane, khi, rec_start, rec_off, cv);
            bool take = false;
            MfmaRec R{};
            if (cur.valid) {
                int64_t ee = (int64_t)cur.sx0 - xs;
                if (ee < 0) ee += n0;
                take = ee < kTile + W - 1;
                R.sx = (int)ee - (W - 1);
                R.sy = cur.q % NRY - (W - 1);
                R.sz = cur.q / NRY - (W - 1);
                R.ox = cur.off.x;
                R.oy = cur.off.y;
                R.oz = cur.off.z;
                R.cre = cur.c.x * scale;
                R.cim = cur.c.y * scale;
            }
            const uint64_t mask = __ballot(take);
            const int pos = __popcll(mask & ((1ull << lane) - 1ull));
            const int nst = __popcll(mask);
            if (take) st[pos] = R;
            wave_sync();
            for (int kb = 0; kb < nst; kb += 16) {
                const int kn = nst - kb < 16 ? nst - kb : 16;
                // clear the weight tables, then drop each record's
                // in-tile weights into its column
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    float *tb = t == 0 ? KX : (t == 1 ? KY : KZ);
                    *(float4 *)(tb + (lane >> 2) * kPitch + 4 * (lane & 3)) =
                        float4{0.0f, 0.0f, 0.0f, 0.0f};
                }
                wave_sync();
                if (lane < 48) {
                    const int j = lane & 15, axis = lane >> 4;
                    if (j < kn) {
                        const MfmaRec &S = st[kb + j];
                        const int s = axis == 0 ? S.sx : (axis == 1 ? S.sy : S.sz);
                        const float o = axis == 0 ? S.ox : (axis == 1 ? S.oy : S.oz);
                        float *tb = axis == 0 ? KX : (axis == 1 ? KY : KZ);
                        const unsigned lim = axis == 2 ? TZ : kTile;
#pragma unroll
                        for (int t = 0; t < W; ++t) {
                            const unsigned c = (unsigned)(s + t);
                            const unsigned row = c < lim ? c : (unsigned)kTile;
                            tb[row * kPitch + j] = es_weight_fast((float)t + o, two_over_w, beta_l2e);
                        }
                    }
                } else {
                    const int j = lane - 48;
                    CC[j] = j < kn ? st[kb + j].cre : 0.0f;
                    CC[kPitch + j] = j < kn ? st[kb + j].cim : 0.0f;
                }
                wave_sync();
                // band of z cells the k-block touches (records in row order)
                int zlo = __builtin_amdgcn_readfirstlane(st[kb].sz);
                int zhi = __builtin_amdgcn_readfirstlane(st[kb + kn - 1].sz) + W - 1;
                zlo = zlo < 0 ? 0 : zlo;
                zhi = zhi > TZ - 1 ? TZ - 1 : zhi;
                const int mlo = zlo >> 1, mhi = zhi >> 1;

                V bh, bl;
                {
                    const f32x8 kx = load8(KX + bxc * kPitch + j0);
                    const f32x8 cc = load8(CC + bpart * kPitch + j0);
                    Op::split(kx * cc, bh, bl);
                }
                const f32x8 ky = load8(KY + ay * kPitch + j0);
#pragma unroll
                for (int m = 0; m < kBlocks; ++m) {
                    if (m < mlo || m > mhi) continue;
                    const f32x8 kz = load8(KZ + (2 * m + az) * kPitch + j0);
                    V ah, al;
                    Op::split(ky * kz, ah, al);
                    acc[m] = Op::mma(ah, bh, acc[m]);
                    if constexpr (kSplit) {
                        acc[m] = Op::mma(ah, bl, acc[m]);
                        acc[m] = Op::mma(al, bh, acc[m]);
                    }
                }
                wave_sync();
            }
            wave_sync();
        }

        // accumulator element i of block m: row 8*(i/4) + 4*half + i%4 of
        // the block (y = row % 16, z = 2m + row / 16), column r32
        const int64_t pslot = np > 1 ? (int64_t)heavy_base[b] + p : 0;
#pragma unroll
        for (int m = 0; m < kBlocks; ++m) {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int row = 8 * (i >> 2) + 4 * half + (i & 3);
                const int yy = row & 15, zz = 2 * m + (row >> 4);
                const float val = acc[m][i] * unscale;
                if (np == 1) {
                    const int64_t gx = x0 + bxc, gy = y0 + yy, gz = z0 + zz;
                    if (gx < n0 && gy < n1 && gz < n2)
                        grid[2 * (v * g.fine_total + gx + n0 * (gy + n1 * gz)) + bpart] = val;
                } else {
                    partial[partial_index<kTile, kTile, TZ>(pslot, gridDim.y, v, bxc, yy, zz) +
                            bpart] = val;
                }
            }
        }
    }
}

// 2D tiles of 32 x 32 cells: G[y, (x, re|im)] += sum_j Ky_j[y] * Kx_j[x] c_j,
// one 32-row block (the tile's y cells) against two 32-column blocks (x cells
// 0..15 and 16..31, real and imaginary parts interleaved).  The 16 records of
// a k-block share one A operand (their y weights) across both column blocks.
constexpr int kTile2 = 32;

template <int W, int MODE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_spread_mfma2(
    GridParams g, SpreadParams sp, const uint32_t *__restrict__ offsets,
    const int4 *__restrict__ rec_start, const float *__restrict__ rec_off,
    const float *__restrict__ c_sorted, int64_t M, const uint32_t *__restrict__ totals,
    const uint32_t *__restrict__ parts, const uint32_t *__restrict__ item_off,
    const uint32_t *__restrict__ heavy_base, const uint32_t *__restrict__ ctrl,
    const uint32_t *__restrict__ cmax, float *__restrict__ grid, float *__restrict__ partial)
{
    using Op = Operand<MODE>;
    using V = typename Op::V;
    constexpr bool kSplit = MODE != (int)SpreadMethod::MfmaF16;
    constexpr int NROWS = kTile2 + W - 1;
    // one spare row per table takes the weights that fall outside the tile
    __shared__ __align__(16) float KX[(kTile2 + 1) * kPitch];
    __shared__ __align__(16) float KY[(kTile2 + 1) * kPitch];
    __shared__ __align__(16) float CC[2 * kPitch];
    __shared__ MfmaRec st[64];
    __shared__ TileRows<NROWS> rows;

    const int lane = threadIdx.x;
    const int r32 = lane & 31;
    const int half = lane >> 5;
    const int j0 = 8 * half;
    const int bxc = r32 >> 1;  // B column: x cell within the column block
    const int bpart = r32 & 1;

    float scale, unscale;
    strength_scale(cmax + blockIdx.y, &scale, &unscale);
    const int64_t v = blockIdx.y;
    const float two_over_w = 2.0f / (float)W;
    const float beta_l2e = (float)(g.beta * 1.4426950408889634);
    const int64_t n_items = ctrl[0];
    const uint32_t part_len_cap = ctrl[3];
    const float *cv = c_sorted + 2 * v * M;
    const int64_t n0 = g.ax[0].n, n1 = g.ax[1].n;

    for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const TileItem ti = tile_item(item, sp.block_count, item_off, parts, totals, part_len_cap);
        const int64_t b = ti.b;
        const int64_t x0 = (b % sp.nblocks[0]) * kTile2, y0 = (b / sp.nblocks[0]) * kTile2;
        const AxisReach rx = axis_reach(x0, kTile2, W, g.ax[0]);
        int64_t xs = x0 - (W - 1);
        if (xs < 0) xs += n0;

        f32x16 acc[2] = {f32x16{}, f32x16{}};
        stage_tile_rows<2, kTile2, 1, W, NROWS>(rows, lane, y0, 0, rx, g, offsets);
        // records in chunks of 64, the next chunk's loads in flight while
        // the current one is contracted
        const int64_t klo = ti.rlo, khi = ti.rhi;
        Fetched<float> nx = fetch_record<float, NROWS>(rows, klo + lane, khi, rec_start, rec_off, cv);
        for (int64_t k0 = klo; k0 < khi; k0 += 64) {
            const Fetched<float> cur = nx;
            nx = fetch_record<float, NROWS>(rows, k0 + 64 + lane, khi, rec_start, rec_off, cv);
            bool take = false;
            MfmaRec R{};
            if (cur.valid) {
                int64_t ee = (int64_t)cur.sx0 - xs;
                if (ee < 0) ee += n0;
                take = ee < kTile2 + W - 1;
                R.sx = (int)ee - (W - 1);
                R.sy = cur.q - (W - 1);
                R.ox = cur.off.x;
                R.oy = cur.off.y;
                R.cre = cur.c.x * scale;
                R.cim = cur.c.y * scale;
            }
            const uint64_t mask = __ballot(take);
            const int pos = __popcll(mask & ((1ull << lane) - 1ull));
            const int nst = __popcll(mask);
            if (take) st[pos] = R;
            wave_sync();
            for (int kb = 0; kb < nst; kb += 16) {
                const int kn = nst - kb < 16 ? nst - kb : 16;
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    float *tb = t == 0 ? KX : KY;
                    *(float4 *)(tb + (lane >> 1) * kPitch + 8 * (lane & 1)) =
                        float4{0.0f, 0.0f, 0.0f, 0.0f};
                    *(float4 *)(tb + (lane >> 1) * kPitch + 8 * (lane & 1) + 4) =
                        float4{0.0f, 0.0f, 0.0f, 0.0f};
                }
                wave_sync();
                if (lane < 32) {
                    const int j = lane & 15, axis = lane >> 4;
                    if (j < kn) {
                        const MfmaRec &S = st[kb + j];
                        const int s0 = axis == 0 ? S.sx : S.sy;
                        const float o = axis == 0 ? S.ox : S.oy;
                        float *tb = axis == 0 ? KX : KY;
#pragma unroll
                        for (int t = 0; t < W; ++t) {
                            const unsigned c = (unsigned)(s0 + t);
                            const unsigned row = c < (unsigned)kTile2 ? c : (unsigned)kTile2;
                            tb[row * kPitch + j] = es_weight_fast((float)t + o, two_over_w, beta_l2e);
                        }
                    }
                } else if (lane < 48) {
                    const int j = lane - 32;
                    CC[j] = j < kn ? st[kb + j].cre : 0.0f;
                    CC[kPitch + j] = j < kn ? st[kb + j].cim : 0.0f;
                }
                wave_sync();
                V ah, al;
                Op::split(load8(KY + r32 * kPitch + j0), ah, al);
                const f32x8 cc = load8(CC + bpart * kPitch + j0);
#pragma unroll
                for (int nb = 0; nb < 2; ++nb) {
                    V bh, bl;
                    Op::split(load8(KX + (16 * nb + bxc) * kPitch + j0) * cc, bh, bl);
                    acc[nb] = Op::mma(ah, bh, acc[nb]);
                    if constexpr (kSplit) {
                        acc[nb] = Op::mma(ah, bl, acc[nb]);
                        acc[nb] = Op::mma(al, bh, acc[nb]);
                    }
                }
                wave_sync();
            }
            wave_sync();
        }

        // accumulator element i: row (y) 8*(i/4) + 4*half + i%4, column r32
        const int64_t pslot = ti.np > 1 ? (int64_t)heavy_base[b] + ti.p : 0;
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int yy = 8 * (i >> 2) + 4 * half + (i & 3);
                const int xx = 16 * nb + bxc;
                const float val = acc[nb][i] * unscale;
                if (ti.np == 1) {
                    const int64_t gx = x0 + xx, gy = y0 + yy;
                    if (gx < n0 && gy < n1)
                        grid[2 * (v * g.fine_total + gx + n0 * gy) + bpart] = val;
                } else {
                    partial[partial_index<kTile2, kTile2, 1>(pslot, gridDim.y, v, xx, yy, 0) +
                            bpart] = val;
                }
            }
        }
    }
}

template <int D, int W, int MODE>
void launch_kernel(dim3 grid3, const GridParams &g, const SpreadParams &sp,
                   const uint32_t *offsets, const int4 *rec_start, const float *rec_off,
                   const float *c_sorted, int64_t M, const TileWork &w, float *grid,
                   float *partial, hipStream_t stream)
{
    if constexpr (D == 3)
        k_spread_mfma3<W, MODE, 8><<<grid3, 64, 0, stream>>>(
            g, sp, offsets, rec_start, rec_off, c_sorted, M, w.totals, w.parts, w.item_off,
            w.heavy_base, w.ctrl, w.cmax, grid, partial);
    else
        k_spread_mfma2<W, MODE><<<grid3, 64, 0, stream>>>(
            g, sp, offsets, rec_start, rec_off, c_sorted, M, w.totals, w.parts, w.item_off,
            w.heavy_base, w.ctrl, w.cmax, grid, partial);
}

template <int D, int W>
bool launch_width(SpreadMethod method, dim3 grid3, const GridParams &g, const SpreadParams &sp,
                  const uint32_t *offsets, const int4 *rec_start, const float *rec_off,
                  const float *c_sorted, int64_t M, const TileWork &w, float *grid,
                  float *partial, hipStream_t stream)
{
    switch (method) {
    case SpreadMethod::MfmaF16:
        launch_kernel<D, W, (int)SpreadMethod::MfmaF16>(grid3, g, sp, offsets, rec_start, rec_off,
                                                        c_sorted, M, w, grid, partial, stream);
        return true;
    case SpreadMethod::MfmaF16x3:
        launch_kernel<D, W, (int)SpreadMethod::MfmaF16x3>(grid3, g, sp, offsets, rec_start,
                                                          rec_off, c_sorted, M, w, grid, partial,
                                                          stream);
        return true;
    case SpreadMethod::MfmaBf16x3:
        launch_kernel<D, W, (int)SpreadMethod::MfmaBf16x3>(grid3, g, sp, offsets, rec_start,
                                                           rec_off, c_sorted, M, w, grid, partial,
                                                           stream);
        return true;
    default:
        return false;
    }
}

template <int D>
bool launch_dim(SpreadMethod method, dim3 grid3, const GridParams &g, const SpreadParams &sp,
                const uint32_t *offsets, const int4 *rec_start, const float *rec_off,
                const float *c_sorted, int64_t M, const TileWork &w, float *grid, float *partial,
                hipStream_t stream)
{
#define MNUFFT_MFMA_WIDTH(Wd)                                                                     \
    case Wd:                                                                                      \
        return launch_width<D, Wd>(method, grid3, g, sp, offsets, rec_start, rec_off, c_sorted,  \
                                   M, w, grid, partial, stream);
    switch (g.width) {
        MNUFFT_MFMA_WIDTH(2)
        MNUFFT_MFMA_WIDTH(3)
        MNUFFT_MFMA_WIDTH(4)
        MNUFFT_MFMA_WIDTH(5)
        MNUFFT_MFMA_WIDTH(6)
        MNUFFT_MFMA_WIDTH(7)
        MNUFFT_MFMA_WIDTH(8)
    default:
        return false;
    }
#undef MNUFFT_MFMA_WIDTH
}

}  // namespace

bool launch_absmax(const float *c, int64_t n, int64_t vectors, uint32_t *out,
                   hipStream_t stream)
{
    int64_t blocks = (n / 4 + 255) / 256;
    if (blocks > 512) blocks = 512;
    if (blocks < 1) blocks = 1;
    k_absmax<<<dim3((unsigned)blocks, (unsigned)vectors), 256, 0, stream>>>(c, n, out);
    return hip_ok(hipGetLastError(), "strength range launch");
}

bool launch_mfma_tiles(SpreadMethod method, int dim, const GridParams &g, const SpreadParams &sp,
                       const uint32_t *offsets, const int4 *rec_start, const float *rec_off,
                       const float *c_sorted, int64_t M, int64_t vectors, const TileWork &w,
                       float *grid, float *partial, hipStream_t stream)
{
    dim3 grid3(w.items_grid, (unsigned)vectors);
    bool ok = false;
    if (dim == 3)
        ok = launch_dim<3>(method, grid3, g, sp, offsets, rec_start, rec_off, c_sorted, M, w, grid,
                           partial, stream);
    else if (dim == 2)
        ok = launch_dim<2>(method, grid3, g, sp, offsets, rec_start, rec_off, c_sorted, M, w, grid,
                           partial, stream);
    if (!ok) {
        set_error("matrix-core spread needs a 2D or 3D plan, a width <= 8 and an MFMA method");
        return false;
    }
    return hip_ok(hipGetLastError(), "matrix-core spread launch");
}

}  // namespace mnufft
