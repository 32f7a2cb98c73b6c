// Device helpers shared by the spreaders: the exponential-of-semicircle
// weight and the bin intervals that reach a block of cells.
#pragma once

#include <hip/hip_runtime.h>

#include "nufft_internal.hpp"

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

// Bin intervals of one axis that hold the support starts reaching a block
// (starts b0-w+1 .. b0+B-1, periodic).  A reach within one bin side of the
// whole axis takes every bin once.
struct AxisReach {
    int64_t lo0, len0, lo1, len1;
    __device__ int64_t count() const { return len0 + len1; }
    __device__ int64_t bin(int64_t k) const { return k < len0 ? lo0 + k : lo1 + (k - len0); }
};

__device__ inline AxisReach axis_reach(int64_t b0, int64_t B, int w, const AxisParams &a)
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

}  // namespace
}  // namespace mnufft
