// Device helpers shared by the spreaders: the exponential-of-semicircle
// weight and the bin intervals that reach a block of cells.
#pragma once

#include <hip/hip_runtime.h>

#include "nufft_internal.hpp"
#include "tile_index.hpp"

namespace mnufft {
namespace {

template <typename T>
__device__ inline T es_weight(T d, int w, T beta)
{
    T s = (T)2 * d / (T)w;
    T q = (T)1 - s * s;
    if (!(q > (T)0)) return (T)0;
    if constexpr (sizeof(T) == 4)
        return expf(beta * (sqrtf(q) - 1.0f));
    else
        return exp(beta * (sqrt(q) - 1.0));
}

// The FP32 weight from the hardware square root and base-2 exponential
// (v_sqrt_f32, v_exp_f32): a few ulp where the weight is near one and an
// absolute error below 1e-12 at the support edges (weight about e^-beta),
// against the correctly rounded sqrtf/expf sequence above.  Used where the
// weights feed FP32 sums of many terms (interpolation) or operands rounded
// to FP16 (matrix-core spread).
__device__ inline float es_weight_fast(float d, float two_over_w, float beta_log2e)
{
    const float s = d * two_over_w;
    const float q = 1.0f - s * s;
    if (!(q > 0.0f)) return 0.0f;
    return __builtin_amdgcn_exp2f(beta_log2e * (__builtin_amdgcn_sqrtf(q) - 1.0f));
}

}  // namespace
}  // namespace mnufft
