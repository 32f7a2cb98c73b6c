// Device-side fold of a coordinate onto the fine grid and its support start
// (reference src/gpu_nd_wgsl.rs:138-188, 198-310 and the count shader
// src/gpu_nd_bins.rs:474-554).
//
// F32 coordinates fold in f64: x * (s_hi + s_lo) with s = n / (2 pi) split
// into an f32 head and tail as the reference does; both partial products are
// exact in f64, their sum and the origin n/2 are each rounded once.  The
// support start is ceil(pos - w/2); the prepared offset is start - pos.
// A coordinate whose magnitude exceeds 4 pi (NaN included) is out of the
// fold's reach: start 0 and an offset past every support, so it weighs zero.
#pragma once

#include "nufft_internal.hpp"

namespace mnufft {

constexpr uint32_t kFoldReachBitsF32 = 0x41490fdbu;  // bits of (float)(4 pi)
constexpr double kFoldReachF64 = 12.566370614359172;
constexpr double kOutsideOffset = 1024.0;

template <typename T>
struct FoldTraits;

template <>
struct FoldTraits<float> {
    __device__ static inline bool in_reach(float x)
    {
        return (__float_as_uint(x) & 0x7fffffffu) <= kFoldReachBitsF32;
    }
    __device__ static inline double position(float x, const AxisParams &a)
    {
#pragma clang fp contract(off)
        double t1 = (double)x * (double)a.scale_hi;
        double t2 = (double)x * (double)a.scale_lo;
        return (t1 + t2) + a.origin;
    }
};

template <>
struct FoldTraits<double> {
    __device__ static inline bool in_reach(double x) { return fabs(x) <= kFoldReachF64; }
    __device__ static inline double position(double x, const AxisParams &a)
    {
#pragma clang fp contract(off)
        return x * a.scale + a.origin;
    }
};

// Folded position in [0, n) of an in-reach coordinate.
template <typename T>
__device__ inline double fold_position(T x, const AxisParams &a)
{
#pragma clang fp contract(off)
    double n = (double)a.n;
    double pos = FoldTraits<T>::position(x, a);
    if (pos < 0.0) pos += n;
    if (pos < 0.0) pos += n;
    if (pos >= n) pos -= n;
    if (pos >= n) pos -= n;
    return pos;
}

// Support start (unwrapped, >= -w/2) and offset start - pos of one axis.
template <typename T>
__device__ inline void support_start(T x, bool reach, const GridParams &g, int axis,
                                     int64_t *start, double *offset)
{
#pragma clang fp contract(off)
    if (!reach) {
        *start = 0;
        *offset = kOutsideOffset;
        return;
    }
    double pos = fold_position<T>(x, g.ax[axis]);
    // pos is in [0, n), so the start fits 32 bits (a direct f64 -> i32
    // conversion)
    const int s = (int)ceil(pos - g.half_width);
    *start = s;
    *offset = (double)s - pos;
}

// Start wrapped onto the grid, clamped so that any input lands in a bin.
__device__ inline int64_t wrapped_start(int64_t start, int64_t n)
{
    int64_t w = start < 0 ? start + n : start;
    return w < 0 ? 0 : (w > n - 1 ? n - 1 : w);
}

template <int D, typename T>
__device__ inline bool point_in_reach(const T *coords, int64_t j)
{
    bool r = true;
#pragma unroll
    for (int a = 0; a < D; ++a) r = r && FoldTraits<T>::in_reach(coords[j * D + a]);
    return r;
}

// Bits of the in-bin start cell (the bin sides' bits).
template <int D>
__device__ __host__ inline int cell_bits(const GridParams &g)
{
    int b = 0;
    for (int a = 0; a < D; ++a) b += g.ax[a].shift;
    return b;
}

}  // namespace mnufft
