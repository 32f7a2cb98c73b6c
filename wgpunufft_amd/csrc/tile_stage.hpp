// Record staging of the row-tile spreaders (the plain row tiles and the
// matrix-core tiles): the reach rows of a tile with their slot ranges and
// the scan of their record counts, and one lane's record fetched with its
// loads in flight until used.  One-wavefront workgroups.
#pragma once

#include <hip/hip_runtime.h>

#include "nufft_internal.hpp"
#include "tile_index.hpp"

namespace mnufft {
namespace {

// Wavefront-wide LDS ordering for one-wavefront workgroups: LDS
// instructions of a wavefront execute in order, so a compiler barrier is
// enough (no s_barrier, no wait for outstanding global loads, which stay in
// flight across it).
__device__ inline void wave_sync()
{
    __builtin_amdgcn_wave_barrier();
    __asm__ volatile("" ::: "memory");
}

// Stages every reach row of the tile at (y0, z0): rows run over the TY + W - 1
// y starts, then (3D) the TZ + W - 1 z starts; row r's x slot ranges are
// [f0, f0 + n0) and, when the x reach wraps, [f1, f1 + (c1 - f1)); excl[r]
// counts the records of the rows before r, excl[NROWS] all of them.
template <int D, int TY, int TZ, int W, int N>
__device__ inline uint32_t stage_tile_rows(TileRows<N> &rows, int lane, int64_t y0, int64_t z0,
                                           const AxisReach &rx, const GridParams &g,
                                           const uint32_t *__restrict__ offsets)
{
    constexpr int NRY = TY + W - 1;
    constexpr int NROWS = D == 3 ? NRY * (TZ + W - 1) : NRY;
    constexpr int NB = (NROWS + 63) / 64;
    static_assert(NROWS <= N, "reach rows beyond the staging arrays");
    // all loads first, then the scans
    uint32_t f0[NB], c0[NB], f1[NB], c1[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int r = 64 * i + lane;
        f0[i] = c0[i] = f1[i] = c1[i] = 0;
        if (r < NROWS) {
            const int64_t row = tile_row_base<D, NRY, W>(r, y0, z0, g);
            f0[i] = offsets[row + rx.lo0];
            c0[i] = offsets[row + rx.lo0 + rx.len0];
            if (rx.len1 > 0) {
                f1[i] = offsets[row + rx.lo1];
                c1[i] = offsets[row + rx.lo1 + rx.len1];
            }
        }
    }
    uint32_t carry = 0;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int r = 64 * i + lane;
        const uint32_t n0 = c0[i] - f0[i], n1c = c1[i] - f1[i];
        const uint32_t cnt = n0 + n1c;
        uint32_t x = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (r < NROWS) {
            rows.f0[r] = f0[i];
            rows.n0[r] = n0;
            rows.f1[r] = f1[i];
            rows.excl[r] = carry + x - cnt;
        }
        carry += __shfl(x, 63, 64);
    }
    if (lane == 0) rows.excl[NROWS] = carry;
    wave_sync();
    return carry;
}

template <typename T>
struct Vec4Of;
template <>
struct Vec4Of<float> {
    using V4 = float4;
    using V2 = float2;
};
template <>
struct Vec4Of<double> {
    using V4 = double4;
    using V2 = double2;
};

// One lane's record of a chunk (the tile's k-th record), with its global
// loads in flight until used.
template <typename T>
struct Fetched {
    int q, sx0;
    typename Vec4Of<T>::V4 off;
    typename Vec4Of<T>::V2 c;
    bool valid;
};

template <typename T, int NROWS>
__device__ inline Fetched<T> fetch_record(const TileRows<NROWS> &rows, int64_t k, int64_t khi,
                                          const int4 *__restrict__ rec_start,
                                          const T *__restrict__ rec_off,
                                          const T *__restrict__ cv)
{
    using V4 = typename Vec4Of<T>::V4;
    using V2 = typename Vec4Of<T>::V2;
    Fetched<T> f;
    f.valid = k < khi;
    f.q = 0;
    f.sx0 = 0;
    f.off = V4{(T)0, (T)0, (T)0, (T)0};
    f.c = V2{(T)0, (T)0};
    if (f.valid) {
        const uint32_t slot = tile_slot<NROWS>(rows, (uint32_t)k, &f.q);
        f.sx0 = rec_start[slot].x;
        f.off = *(const V4 *)(rec_off + 4 * (int64_t)slot);
        f.c = *(const V2 *)(cv + 2 * (int64_t)slot);
    }
    return f;
}

}  // namespace
}  // namespace mnufft
