// Matrix-core operand helpers shared by the 3D tile and dense-bin spreads.
//
// An FP32 operand value a is fed to the matrix cores as FP16 (one product,
// relative error about 2^-11) or as a hi + lo pair in FP16 or bfloat16 whose
// three products hi*hi + hi*lo + lo*hi are accumulated in FP32 (relative
// error about 2^-21 for FP16, 2^-16 for bfloat16).
#pragma once

#include <hip/hip_runtime.h>

#include "nufft_internal.hpp"
#include "tile_stage.hpp"

namespace mnufft {
namespace mfma {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// mma: v_mfma_f32_32x32x16 (lane l holds A row l % 32, B column l % 32, K
// elements 8 (l / 32) .. +7); mma16: v_mfma_f32_16x16x32 (lane l holds A row
// l % 16, B column l % 16, K elements 8 (l / 16) .. +7; D column l % 16, rows
// 4 (l / 16) .. +3).
template <int MODE>
struct Operand;

template <>
struct Operand<(int)SpreadMethod::MfmaF16> {
    using V = f16x8;
    static constexpr bool kSplit = false;
    __device__ static void split(const f32x8 &a, V &hi, V &lo)
    {
        hi = __builtin_convertvector(a, V);
        (void)lo;
    }
    __device__ static f32x16 mma(const V &a, const V &b, const f32x16 &c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
    __device__ static f32x4 mma16(const V &a, const V &b, const f32x4 &c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

template <>
struct Operand<(int)SpreadMethod::MfmaF16x3> {
    using V = f16x8;
    static constexpr bool kSplit = true;
    __device__ static void split(const f32x8 &a, V &hi, V &lo)
    {
        hi = __builtin_convertvector(a, V);
        lo = __builtin_convertvector(a - __builtin_convertvector(hi, f32x8), V);
    }
    __device__ static f32x16 mma(const V &a, const V &b, const f32x16 &c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
    __device__ static f32x4 mma16(const V &a, const V &b, const f32x4 &c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

template <>
struct Operand<(int)SpreadMethod::MfmaBf16x3> {
    using V = bf16x8;
    static constexpr bool kSplit = true;
    __device__ static void split(const f32x8 &a, V &hi, V &lo)
    {
        hi = __builtin_convertvector(a, V);
        lo = __builtin_convertvector(a - __builtin_convertvector(hi, f32x8), V);
    }
    __device__ static f32x16 mma(const V &a, const V &b, const f32x16 &c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
    __device__ static f32x4 mma16(const V &a, const V &b, const f32x4 &c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

__device__ inline f32x8 load8(const float *p)
{
    const float4 a = *(const float4 *)p;
    const float4 b = *(const float4 *)(p + 4);
    return f32x8{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
}

// Power-of-two scale that brings the largest strength magnitude below one.
__device__ inline void strength_scale(const uint32_t *cmax, float *scale, float *unscale)
{
    *scale = 1.0f;
    *unscale = 1.0f;
    const float m = __uint_as_float(cmax[0]);
    if (m > 0.0f && m <= 3.4028234663852886e38f) {
        int e = 0;
        (void)frexpf(m, &e);
        *scale = ldexpf(1.0f, -e);
        *unscale = ldexpf(1.0f, e);
    }
}

}  // namespace mfma
}  // namespace mnufft
