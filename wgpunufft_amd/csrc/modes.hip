// Mode-space stages: type-1 deconvolution / truncation, type-2
// predeconvolution onto the fine grid, type-3 rescale and phase stages, and
// FP64 direct sums.
//
// Reference behaviour:
//  * deconvolution (src/gpu_type1_nd.rs:617-687): modes[v,k] =
//    fine[v, fold(k)] * amp[k], fold centered: k < N/2 ? n - (N/2 - k) :
//    k - N/2; FFT order: k < ceil(N/2) ? k : n - (N - k);
//  * predeconvolution (src/gpu_3d.rs:902-970, src/gpu_nd.rs:618-697): the
//    inverse mapping onto a cleared grid;
//  * amplitudes (src/gpu_nd.rs:910-1014): amp[k] = prod_a 1/phi_hat_a(|k_a|),
//    formed in f64;
//  * type 3 (src/gpu_type3.rs:741-1297): x' = (x - C)/gamma and
//    c' = c exp(i isign D.x); s' = h gamma (s - D) and
//    factor = exp(i isign C.(s - D)) / prod_a phi_hat(s'_a); out = interp * factor.
//
// MI355X design: every stage is one HBM-streaming pass, one lane per output
// element.  Predeconvolution writes every fine cell once (the mode or zero),
// which fuses the reference's clear pass into the scatter and keeps the
// sequence free of memset nodes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "modes.hpp"

namespace mnufft {

namespace {

inline unsigned grid_for(int64_t n, int threads)
{
    // the elementwise kernels loop with a grid stride: at most 2^20
    // workgroups keeps blocks x threads far below 2^32 work-items
    int64_t b = (n + threads - 1) / threads;
    if (b > ((int64_t)1 << 20)) b = (int64_t)1 << 20;
    return (unsigned)(b < 1 ? 1 : b);
}

// Mode k of one vector: out[2k..] = grid[fold(k)] * amplitude (the
// amplitude multiplied in axis order).
template <int D, typename T>
__device__ inline void deconv_mode(const ModeParams &mp, const double *__restrict__ inv_phihat,
                                   const T *__restrict__ grid_v, int64_t k, T *__restrict__ out)
{
    int64_t rest = k, fine_lin = 0, stride = 1;
    double amp = 1.0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        int64_t N = mp.n_modes[a], n = mp.fine[a];
        int64_t ka = rest % N;
        rest /= N;
        int64_t m = mp.centered ? ka - N / 2 : (ka < (N + 1) / 2 ? ka : ka - N);
        int64_t f = m < 0 ? m + n : m;
        fine_lin += f * stride;
        stride *= n;
        amp *= inv_phihat[mp.phihat_off[a] + (m < 0 ? -m : m)];
    }
    const T *src = grid_v + 2 * fine_lin;
    T a = (T)amp;
    out[0] = src[0] * a;
    out[1] = src[1] * a;
}

template <int D, typename T>
__global__ __launch_bounds__(256) void k_deconv(ModeParams mp, const double *__restrict__ inv_phihat,
                                                const T *__restrict__ grid, T *__restrict__ modes,
                                                int64_t vectors)
{
    const int64_t n_total = mp.mode_total * vectors;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_total;
         t += (int64_t)gridDim.x * blockDim.x) {
        int64_t v = t / mp.mode_total;
        int64_t k = t - v * mp.mode_total;
        deconv_mode<D, T>(mp, inv_phihat, grid + 2 * v * mp.fine_total, k, modes + 2 * t);
    }
}

// Problem of element t of a combined list (the last problem whose base <= t).
template <bool Modes>
__device__ inline int problem_of(const ModeProblem *__restrict__ mprobs, int nprob, int64_t t)
{
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((Modes ? mprobs[mid].mode_base : mprobs[mid].fine_base) <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

template <int D, typename T>
__global__ __launch_bounds__(256) void k_deconv_many(const BatchProblem *__restrict__ probs,
                                                     const ModeProblem *__restrict__ mprobs,
                                                     int nprob, int64_t total)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int q = problem_of<true>(mprobs, nprob, t);
        const ModeProblem &mq = mprobs[q];
        const int64_t k = t - mq.mode_base;
        T *out = (T *)probs[q].out + 2 * k;
        if (mq.zero) {
            out[0] = (T)0;
            out[1] = (T)0;
        } else {
            deconv_mode<D, T>(mq.mp, mq.inv_phihat, (const T *)probs[q].grid, k, out);
        }
    }
}

// Row form of the deconvolution: workgroup (b, c) writes modes
// [256 c, 256 c + 256) of the mode rows b, b + gridDim.x, ... (a row: the
// modes of one vector with every axis above 0 fixed).  A lane forms its
// axis-0 fine index and factor once; each row's part takes 32-bit divisions
// only, and each mode is one vector store.  The amplitude is multiplied in
// the axis order of k_deconv, so both kernels give the same bits.
template <int D, typename T>
__global__ __launch_bounds__(256) void k_deconv_rows(ModeParams mp,
                                                     const double *__restrict__ inv_phihat,
                                                     const T *__restrict__ grid,
                                                     T *__restrict__ modes,
                                                     uint32_t rows_per_vector, uint32_t rows)
{
    using V2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
    const uint32_t N0 = (uint32_t)mp.n_modes[0], n0 = (uint32_t)mp.fine[0];
    const uint32_t k = blockIdx.y * 256u + threadIdx.x;
    if (k >= N0) return;
    const int64_t m0 = mp.centered ? (int64_t)k - N0 / 2
                                   : (k < (N0 + 1) / 2 ? (int64_t)k : (int64_t)k - N0);
    const int64_t f0 = m0 < 0 ? m0 + n0 : m0;
    const double ph0 = inv_phihat[mp.phihat_off[0] + (m0 < 0 ? -m0 : m0)];
    for (uint32_t r = blockIdx.x; r < rows; r += gridDim.x) {
        const uint32_t v = r / rows_per_vector;
        const uint32_t row = r - v * rows_per_vector;
        uint32_t rest = row;
        int64_t fine_row = 0, stride = n0;
        double amp = 1.0;
        amp *= ph0;
#pragma unroll
        for (int a = 1; a < D; ++a) {
            const uint32_t N = (uint32_t)mp.n_modes[a], n = (uint32_t)mp.fine[a];
            const uint32_t ka = rest % N;
            rest /= N;
            const int64_t m = mp.centered ? (int64_t)ka - N / 2
                                          : (ka < (N + 1) / 2 ? (int64_t)ka : (int64_t)ka - N);
            fine_row += (m < 0 ? m + n : m) * stride;
            stride *= n;
            amp *= inv_phihat[mp.phihat_off[a] + (m < 0 ? -m : m)];
        }
        const T *src = grid + 2 * ((int64_t)v * mp.fine_total + fine_row + f0);
        const T ta = (T)amp;
        V2 val;
        val.x = src[0] * ta;
        val.y = src[1] * ta;
        *(V2 *)(modes + 2 * ((int64_t)v * mp.mode_total + (int64_t)row * N0 + k)) = val;
    }
}

// Fine cell f of one vector: the mode folded onto it times its amplitude
// (amplitude in axis order), zero off the modes.
template <int D, typename T>
__device__ inline void predeconv_cell(const ModeParams &mp, const double *__restrict__ inv_phihat,
                                      const T *__restrict__ modes_v, int64_t f, T *__restrict__ out)
{
    int64_t rest = f, mode_lin = 0, stride = 1;
    bool has = true;
    double amp = 1.0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        int64_t N = mp.n_modes[a], n = mp.fine[a];
        int64_t fa = rest % n;
        rest /= n;
        int64_t hi = N - 1 - N / 2;  // largest frequency
        int64_t m;
        if (fa <= hi)
            m = fa;
        else if (fa >= n - N / 2)
            m = fa - n;
        else {
            has = false;
            m = 0;
        }
        int64_t ka = mp.centered ? m + N / 2 : (m >= 0 ? m : m + N);
        mode_lin += ka * stride;
        stride *= N;
        amp *= inv_phihat[mp.phihat_off[a] + (m < 0 ? -m : m)];
    }
    T re = (T)0, im = (T)0;
    if (has) {
        const T *src = modes_v + 2 * mode_lin;
        T a = (T)amp;
        re = src[0] * a;
        im = src[1] * a;
    }
    out[0] = re;
    out[1] = im;
}

template <int D, typename T>
__global__ __launch_bounds__(256) void k_predeconv(ModeParams mp,
                                                   const double *__restrict__ inv_phihat,
                                                   const T *__restrict__ modes,
                                                   T *__restrict__ grid, int64_t vectors)
{
    const int64_t n_total = mp.fine_total * vectors;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_total;
         t += (int64_t)gridDim.x * blockDim.x) {
        int64_t v = t / mp.fine_total;
        int64_t f = t - v * mp.fine_total;
        predeconv_cell<D, T>(mp, inv_phihat, modes + 2 * v * mp.mode_total, f, grid + 2 * t);
    }
}

template <int D, typename T>
__global__ __launch_bounds__(256) void k_predeconv_many(const BatchProblem *__restrict__ probs,
                                                        const ModeProblem *__restrict__ mprobs,
                                                        int nprob, int64_t total)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int q = problem_of<false>(mprobs, nprob, t);
        const ModeProblem &mq = mprobs[q];
        const int64_t f = t - mq.fine_base;
        predeconv_cell<D, T>(mq.mp, mq.inv_phihat, (const T *)probs[q].in, f,
                             (T *)probs[q].grid + 2 * f);
    }
}

// Row form of the predeconvolution: workgroup (b, c) writes cells
// [256 c, 256 c + 256) of the fine rows b, b + gridDim.x, ... (a row: the
// cells of one vector with every axis above 0 fixed).  A lane forms its
// axis-0 mode index and factor once; each row's part takes 32-bit divisions
// only, and each cell is one vector store.  The amplitude is multiplied in
// the axis order of k_predeconv, so both kernels give the same bits.
template <int D, typename T>
__global__ __launch_bounds__(256) void k_predeconv_rows(ModeParams mp,
                                                        const double *__restrict__ inv_phihat,
                                                        const T *__restrict__ modes,
                                                        T *__restrict__ grid,
                                                        uint32_t rows_per_vector, uint32_t rows)
{
    using V2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
    const uint32_t n0 = (uint32_t)mp.fine[0], N0 = (uint32_t)mp.n_modes[0];
    const uint32_t x = blockIdx.y * 256u + threadIdx.x;
    if (x >= n0) return;
    // the lane's axis-0 part, the same for every row
    bool in = true;
    int64_t m = 0;
    if (x <= N0 - 1 - N0 / 2)
        m = x;
    else if (x >= n0 - N0 / 2)
        m = (int64_t)x - n0;
    else
        in = false;
    const int64_t k0 = mp.centered ? m + N0 / 2 : (m >= 0 ? m : m + N0);
    const double ph0 = inv_phihat[mp.phihat_off[0] + (m < 0 ? -m : m)];
    for (uint32_t r = blockIdx.x; r < rows; r += gridDim.x) {
        const uint32_t v = r / rows_per_vector;
        const uint32_t row = r - v * rows_per_vector;
        uint32_t rest = row;
        int64_t mode_row = 0, stride = N0;
        bool has = in;
        double amp = 1.0;
        amp *= ph0;
#pragma unroll
        for (int a = 1; a < D; ++a) {
            const uint32_t N = (uint32_t)mp.n_modes[a], n = (uint32_t)mp.fine[a];
            const uint32_t fa = rest % n;
            rest /= n;
            int64_t ma = 0;
            if (fa <= N - 1 - N / 2)
                ma = fa;
            else if (fa >= n - N / 2)
                ma = (int64_t)fa - n;
            else
                has = false;
            const int64_t ka = mp.centered ? ma + N / 2 : (ma >= 0 ? ma : ma + N);
            mode_row += ka * stride;
            stride *= N;
            amp *= inv_phihat[mp.phihat_off[a] + (ma < 0 ? -ma : ma)];
        }
        V2 val;
        val.x = (T)0;
        val.y = (T)0;
        if (has) {
            const T *src = modes + 2 * ((int64_t)v * mp.mode_total + mode_row + k0);
            const T ta = (T)amp;
            val.x = src[0] * ta;
            val.y = src[1] * ta;
        }
        *(V2 *)(grid + 2 * ((int64_t)v * mp.fine_total + (int64_t)row * n0 + x)) = val;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill(T *__restrict__ p, int64_t n)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x)
        p[t] = (T)0;
}

// ---------------------------------------------------------------------
// Type 3.
// ---------------------------------------------------------------------
template <int D, typename T>
__global__ __launch_bounds__(256) void k_type3_source(Type3Params tp, int64_t M, int64_t vectors,
                                                      const T *__restrict__ x,
                                                      const T *__restrict__ c,
                                                      T *__restrict__ xr, T *__restrict__ cp)
{
    const int64_t n_total = M * vectors;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_total;
         t += (int64_t)gridDim.x * blockDim.x) {
        int64_t j = t % M;
        double phase = 0.0;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            double xa = (double)x[j * D + a];
            if (xr && t < M) xr[j * D + a] = (T)((xa - tp.src_center[a]) * tp.inv_gamma[a]);
            phase += xa * tp.tgt_center[a];
        }
        double sn, cs;
        sincos(phase * tp.sign, &sn, &cs);
        double re = (double)c[2 * t], im = (double)c[2 * t + 1];
        cp[2 * t] = (T)(re * cs - im * sn);
        cp[2 * t + 1] = (T)(re * sn + im * cs);
    }
}

template <int D, typename T>
__global__ __launch_bounds__(256) void k_type3_target(Type3Params tp, int64_t N,
                                                      const T *__restrict__ s,
                                                      T *__restrict__ sr,
                                                      double *__restrict__ factor)
{
    const int64_t n_total = N;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n_total;
         k += (int64_t)gridDim.x * blockDim.x) {
        double phase = 0.0, phi = 1.0;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            double cen = (double)s[k * D + a] - tp.tgt_center[a];
            double resc = cen * tp.tgt_scale[a];
            sr[k * D + a] = (T)resc;
            phase += cen * tp.src_center[a];
            double ft = 0.0;
            for (int q = 0; q < tp.nquad; ++q) ft += tp.quad_w[q] * cos(resc * tp.quad_x[q]);
            phi *= ft;
        }
        double sn, cs;
        sincos(phase * tp.sign, &sn, &cs);
        factor[2 * k] = cs / phi;
        factor[2 * k + 1] = sn / phi;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_type3_final(int64_t N, int64_t vectors,
                                                     const T *__restrict__ interp,
                                                     const double *__restrict__ factor,
                                                     T *__restrict__ out,
                                                     const uint32_t *__restrict__ perm,
                                                     int64_t n_out)
{
    const int64_t n_total = N * vectors;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = t / N, k = t - v * N;
        double fr = factor[2 * k], fi = factor[2 * k + 1];
        double re = (double)interp[2 * t], im = (double)interp[2 * t + 1];
        const int64_t o = perm ? v * n_out + perm[k] : t;
        out[2 * o] = (T)(re * fr - im * fi);
        out[2 * o + 1] = (T)(re * fi + im * fr);
    }
}

// Box of a target coordinate vector (clamped to the boxes).
template <int D, typename T>
__device__ inline uint32_t target_box(const Type3Boxes &bx, const T *sk)
{
    int64_t b = 0, stride = 1;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        double q = floor(((double)sk[a] - bx.lo[a]) / bx.width[a]);
        q = q < 0.0 ? 0.0 : (q > (double)(bx.k[a] - 1) ? (double)(bx.k[a] - 1) : q);
        if (!(q == q)) q = 0.0;  // NaN coordinates land in box 0
        b += (int64_t)q * stride;
        stride *= bx.k[a];
    }
    return (uint32_t)b;
}

__global__ void k_clear_u32(uint32_t *p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

template <int D, typename T>
__global__ __launch_bounds__(256) void k_type3_box_count(Type3Boxes bx, int64_t N,
                                                         const T *__restrict__ s,
                                                         uint32_t *__restrict__ box,
                                                         uint32_t *__restrict__ counts)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const uint32_t b = target_box<D, T>(bx, s + k * D);
    box[k] = b;
    atomicAdd(counts + b, 1u);
}

template <int D, typename T>
__global__ __launch_bounds__(256) void k_type3_box_scatter(int64_t N, const T *__restrict__ s,
                                                           const uint32_t *__restrict__ box,
                                                           const uint32_t *__restrict__ offsets,
                                                           uint32_t *__restrict__ cursor,
                                                           uint32_t *__restrict__ perm,
                                                           T *__restrict__ s_sorted)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const uint32_t b = box[k];
    const int64_t pos = (int64_t)offsets[b] + atomicAdd(cursor + b, 1u);
    perm[pos] = (uint32_t)k;
#pragma unroll
    for (int a = 0; a < D; ++a) s_sorted[pos * D + a] = s[k * D + a];
}

// ---------------------------------------------------------------------
// Complex matrix transpose through LDS: 32 x 32 tiles, read along rows and
// written along the columns of the output, each workgroup taking tiles with
// a grid stride; blockIdx.y = vector.
// ---------------------------------------------------------------------
template <typename V2>
__global__ __launch_bounds__(256) void k_transpose(const V2 *__restrict__ in, V2 *__restrict__ out,
                                                   int64_t rows, int64_t cols, int64_t tiles_c,
                                                   int64_t tiles)
{
    __shared__ V2 t[32][33];
    const int64_t off = (int64_t)blockIdx.y * rows * cols;
    const V2 *src = in + off;
    V2 *dst = out + off;
    const int x = threadIdx.x & 31, y = threadIdx.x >> 5;
    for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const int64_t r0 = (tile / tiles_c) * 32, c0 = (tile % tiles_c) * 32;
        for (int j = y; j < 32; j += 8) {
            const int64_t r = r0 + j, c = c0 + x;
            if (r < rows && c < cols) t[j][x] = src[r * cols + c];
        }
        __syncthreads();
        for (int j = y; j < 32; j += 8) {
            const int64_t c = c0 + j, r = r0 + x;
            if (r < rows && c < cols) dst[c * rows + r] = t[x][j];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------
// FP64 direct sums (reference src/direct.rs:9-131, src/type3.rs:609-673):
// one lane per output, the inputs streamed through LDS in tiles.
// ---------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(256) void k_direct(int kind, DirectParams dp,
                                                const double *__restrict__ x,
                                                const double *__restrict__ in,
                                                const int64_t *__restrict__ sel,
                                                const double *__restrict__ s,
                                                double *__restrict__ out)
{
    __shared__ double tx[256 * D];
    __shared__ double tc[256 * 2];
    int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool has = o < dp.nout;
    int64_t target = has ? (sel ? sel[o] : o) : 0;
    // what this lane holds fixed: a mode (type 1), a point (type 2) or a
    // target frequency (type 3)
    double fixed[D];
    if (kind == 1) {
        int64_t rest = target;
        for (int a = 0; a < D; ++a) {
            int64_t N = dp.n_modes[a];
            int64_t ka = rest % N;
            rest /= N;
            fixed[a] = (double)(dp.centered ? ka - N / 2 : (ka < (N + 1) / 2 ? ka : ka - N));
        }
    } else if (kind == 2) {
        for (int a = 0; a < D; ++a) fixed[a] = has ? x[target * D + a] : 0.0;
    } else {
        for (int a = 0; a < D; ++a) fixed[a] = has ? s[target * D + a] : 0.0;
    }
    double re = 0.0, im = 0.0;
    int64_t count = kind == 2 ? dp.mode_total : dp.M;
    for (int64_t base = 0; base < count; base += 256) {
        int64_t i = base + threadIdx.x;
        if (i < count) {
            if (kind == 2) {
                int64_t rest = i;
                for (int a = 0; a < D; ++a) {
                    int64_t N = dp.n_modes[a];
                    int64_t ka = rest % N;
                    rest /= N;
                    tx[threadIdx.x * D + a] =
                        (double)(dp.centered ? ka - N / 2 : (ka < (N + 1) / 2 ? ka : ka - N));
                }
            } else {
                for (int a = 0; a < D; ++a) tx[threadIdx.x * D + a] = x[i * D + a];
            }
            tc[2 * threadIdx.x] = in[2 * i];
            tc[2 * threadIdx.x + 1] = in[2 * i + 1];
        }
        __syncthreads();
        int64_t m = count - base < 256 ? count - base : 256;
        for (int q = 0; q < m; ++q) {
            double ph = 0.0;
            for (int a = 0; a < D; ++a) ph += fixed[a] * tx[q * D + a];
            double sn, cs;
            sincos(dp.sign * ph, &sn, &cs);
            double cr = tc[2 * q], ci = tc[2 * q + 1];
            re += cr * cs - ci * sn;
            im += cr * sn + ci * cs;
        }
        __syncthreads();
    }
    if (has) {
        out[2 * o] = re;
        out[2 * o + 1] = im;
    }
}

}  // namespace

template <int D, typename T>
bool deconvolve(const ModeParams &mp, const double *inv_phihat, const T *grid, T *modes,
                int64_t vectors, hipStream_t stream)
{
    int64_t n = mp.mode_total * vectors;
    if (n == 0) return true;
    const int64_t rows_per_vector = mp.mode_total / mp.n_modes[0];
    const int64_t rows = rows_per_vector * vectors;
    const int64_t chunks = (mp.n_modes[0] + 255) / 256;
    bool fits32 = rows <= 0x7fffffff && chunks <= 65535;
    for (int a = 0; a < D; ++a) fits32 = fits32 && mp.fine[a] <= 0x7fffffff;
    if (fits32) {
        // a few thousand workgroups, each taking rows with a stride
        const int64_t per = std::max<int64_t>(1, 8192 / chunks);
        const unsigned gx = (unsigned)std::min<int64_t>(rows, per);
        k_deconv_rows<D, T><<<dim3(gx, (unsigned)chunks), 256, 0, stream>>>(
            mp, inv_phihat, grid, modes, (uint32_t)rows_per_vector, (uint32_t)rows);
    } else {
        k_deconv<D, T><<<grid_for(n, 256), 256, 0, stream>>>(mp, inv_phihat, grid, modes,
                                                              vectors);
    }
    return hip_ok(hipGetLastError(), "deconvolution launch");
}

template <int D, typename T>
bool predeconvolve(const ModeParams &mp, const double *inv_phihat, const T *modes, T *grid,
                   int64_t vectors, hipStream_t stream)
{
    int64_t n = mp.fine_total * vectors;
    if (n == 0) return true;
    const int64_t rows_per_vector = mp.fine_total / mp.fine[0];
    const int64_t rows = rows_per_vector * vectors;
    const int64_t chunks = (mp.fine[0] + 255) / 256;
    bool fits32 = rows <= 0x7fffffff && chunks <= 65535;
    for (int a = 0; a < D; ++a) fits32 = fits32 && mp.fine[a] <= 0x7fffffff;
    if (fits32) {
        // a few thousand workgroups, each taking rows with a stride
        const int64_t per = std::max<int64_t>(1, 8192 / chunks);
        const unsigned gx = (unsigned)std::min<int64_t>(rows, per);
        k_predeconv_rows<D, T><<<dim3(gx, (unsigned)chunks), 256, 0, stream>>>(
            mp, inv_phihat, modes, grid, (uint32_t)rows_per_vector, (uint32_t)rows);
    } else {
        k_predeconv<D, T><<<grid_for(n, 256), 256, 0, stream>>>(mp, inv_phihat, modes, grid,
                                                                 vectors);
    }
    return hip_ok(hipGetLastError(), "predeconvolution launch");
}

template <typename T>
bool fill_zero(T *p, int64_t n, hipStream_t stream)
{
    if (n <= 0) return true;
    k_fill<T><<<grid_for(n, 256), 256, 0, stream>>>(p, n);
    return hip_ok(hipGetLastError(), "fill launch");
}

template <int D, typename T>
bool type3_source(const Type3Params &tp, int64_t M, int64_t vectors, const T *x, const T *c, T *xr,
                  T *cp, hipStream_t stream)
{
    if (M == 0) return true;
    k_type3_source<D, T><<<grid_for(M * vectors, 256), 256, 0, stream>>>(tp, M, vectors, x, c, xr,
                                                                          cp);
    return hip_ok(hipGetLastError(), "type-3 source launch");
}

template <int D, typename T>
bool type3_target(const Type3Params &tp, int64_t N, const T *s, T *sr, double *factor,
                  hipStream_t stream)
{
    if (N == 0) return true;
    k_type3_target<D, T><<<grid_for(N, 256), 256, 0, stream>>>(tp, N, s, sr, factor);
    return hip_ok(hipGetLastError(), "type-3 target launch");
}

template <typename T>
bool type3_final(int64_t N, int64_t vectors, const T *interp, const double *factor, T *out,
                 hipStream_t stream, const uint32_t *perm, int64_t n_out)
{
    if (N == 0) return true;
    k_type3_final<T><<<grid_for(N * vectors, 256), 256, 0, stream>>>(N, vectors, interp, factor,
                                                                     out, perm, n_out);
    return hip_ok(hipGetLastError(), "type-3 final launch");
}

template <int D, typename T>
bool type3_box_count(const Type3Boxes &bx, int64_t N, const T *s, uint32_t *box,
                     uint32_t *counts, hipStream_t stream)
{
    k_clear_u32<<<grid_for(bx.count, 256), 256, 0, stream>>>(counts, bx.count);
    if (N > 0)
        k_type3_box_count<D, T><<<grid_for(N, 256), 256, 0, stream>>>(bx, N, s, box, counts);
    return hip_ok(hipGetLastError(), "type-3 box count launch");
}

template <int D, typename T>
bool type3_box_scatter(const Type3Boxes &bx, int64_t N, const T *s, const uint32_t *box,
                       const uint32_t *offsets, uint32_t *cursor, uint32_t *perm, T *s_sorted,
                       hipStream_t stream)
{
    k_clear_u32<<<grid_for(bx.count, 256), 256, 0, stream>>>(cursor, bx.count);
    if (N > 0)
        k_type3_box_scatter<D, T><<<grid_for(N, 256), 256, 0, stream>>>(N, s, box, offsets, cursor,
                                                                       perm, s_sorted);
    return hip_ok(hipGetLastError(), "type-3 box scatter launch");
}

template <typename T>
bool transpose_complex(const T *in, T *out, int64_t rows, int64_t cols, int64_t vectors,
                       hipStream_t stream)
{
    using V2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
    if (rows <= 0 || cols <= 0 || vectors <= 0) return true;
    if (vectors > 65535) {
        set_error("transpose: at most 65535 vectors");
        return false;
    }
    const int64_t tiles_c = (cols + 31) / 32;
    const int64_t tiles = ((rows + 31) / 32) * tiles_c;
    const unsigned gx = (unsigned)std::min<int64_t>(tiles, 16384);
    k_transpose<V2><<<dim3(gx, (unsigned)vectors), 256, 0, stream>>>(
        (const V2 *)in, (V2 *)out, rows, cols, tiles_c, tiles);
    return hip_ok(hipGetLastError(), "transpose launch");
}
template bool transpose_complex<float>(const float *, float *, int64_t, int64_t, int64_t,
                                       hipStream_t);
template bool transpose_complex<double>(const double *, double *, int64_t, int64_t, int64_t,
                                        hipStream_t);

template <int D, typename T>
bool deconvolve_many(const BatchProblem *probs, const ModeProblem *mprobs, int nprob,
                     int64_t total_modes, hipStream_t stream)
{
    if (total_modes == 0) return true;
    k_deconv_many<D, T><<<grid_for(total_modes, 256), 256, 0, stream>>>(probs, mprobs, nprob,
                                                                         total_modes);
    return hip_ok(hipGetLastError(), "batched deconvolution launch");
}

template <int D, typename T>
bool predeconvolve_many(const BatchProblem *probs, const ModeProblem *mprobs, int nprob,
                        int64_t total_cells, hipStream_t stream)
{
    if (total_cells == 0) return true;
    k_predeconv_many<D, T><<<grid_for(total_cells, 256), 256, 0, stream>>>(probs, mprobs, nprob,
                                                                            total_cells);
    return hip_ok(hipGetLastError(), "batched predeconvolution launch");
}
#define MNUFFT_MODES_MANY_INSTANTIATE(D, T)                                                       \
    template bool deconvolve_many<D, T>(const BatchProblem *, const ModeProblem *, int, int64_t,  \
                                        hipStream_t);                                             \
    template bool predeconvolve_many<D, T>(const BatchProblem *, const ModeProblem *, int,        \
                                           int64_t, hipStream_t);
MNUFFT_MODES_MANY_INSTANTIATE(1, float)
MNUFFT_MODES_MANY_INSTANTIATE(2, float)
MNUFFT_MODES_MANY_INSTANTIATE(3, float)
MNUFFT_MODES_MANY_INSTANTIATE(1, double)
MNUFFT_MODES_MANY_INSTANTIATE(2, double)
MNUFFT_MODES_MANY_INSTANTIATE(3, double)

bool direct_sum(int kind, int d, const DirectParams &dp, const double *x, const double *in,
                const int64_t *sel, const double *s, double *out, hipStream_t stream)
{
    // one lane per output, no grid stride
    if (dp.nout > ((int64_t)1 << 31)) {
        set_error("direct sums take at most 2^31 outputs per call");
        return false;
    }
    const unsigned g = (unsigned)std::max<int64_t>(1, (dp.nout + 255) / 256);
    switch (d) {
    case 1: k_direct<1><<<g, 256, 0, stream>>>(kind, dp, x, in, sel, s, out); break;
    case 2: k_direct<2><<<g, 256, 0, stream>>>(kind, dp, x, in, sel, s, out); break;
    case 3: k_direct<3><<<g, 256, 0, stream>>>(kind, dp, x, in, sel, s, out); break;
    case 4: k_direct<4><<<g, 256, 0, stream>>>(kind, dp, x, in, sel, s, out); break;
    case 5: k_direct<5><<<g, 256, 0, stream>>>(kind, dp, x, in, sel, s, out); break;
    case 6: k_direct<6><<<g, 256, 0, stream>>>(kind, dp, x, in, sel, s, out); break;
    case 7: k_direct<7><<<g, 256, 0, stream>>>(kind, dp, x, in, sel, s, out); break;
    case 8: k_direct<8><<<g, 256, 0, stream>>>(kind, dp, x, in, sel, s, out); break;
    default: set_error("direct sums take 1 to 8 dimensions"); return false;
    }
    return hip_ok(hipGetLastError(), "direct-sum launch");
}

#define MNUFFT_MODES_INSTANTIATE(D, T)                                                            \
    template bool deconvolve<D, T>(const ModeParams &, const double *, const T *, T *, int64_t,    \
                                   hipStream_t);                                                  \
    template bool predeconvolve<D, T>(const ModeParams &, const double *, const T *, T *, int64_t,  \
                                      hipStream_t);                                               \
    template bool type3_source<D, T>(const Type3Params &, int64_t, int64_t, const T *, const T *,  \
                                      T *, T *, hipStream_t);                                     \
    template bool type3_target<D, T>(const Type3Params &, int64_t, const T *, T *, double *,       \
                                      hipStream_t);                                               \
    template bool type3_box_count<D, T>(const Type3Boxes &, int64_t, const T *, uint32_t *,        \
                                         uint32_t *, hipStream_t);                                \
    template bool type3_box_scatter<D, T>(const Type3Boxes &, int64_t, const T *,                  \
                                           const uint32_t *, const uint32_t *, uint32_t *,        \
                                           uint32_t *, T *, hipStream_t);
MNUFFT_MODES_INSTANTIATE(1, float)
MNUFFT_MODES_INSTANTIATE(2, float)
MNUFFT_MODES_INSTANTIATE(3, float)
MNUFFT_MODES_INSTANTIATE(1, double)
MNUFFT_MODES_INSTANTIATE(2, double)
MNUFFT_MODES_INSTANTIATE(3, double)
MNUFFT_MODES_INSTANTIATE(4, float)
MNUFFT_MODES_INSTANTIATE(5, float)
MNUFFT_MODES_INSTANTIATE(6, float)
MNUFFT_MODES_INSTANTIATE(7, float)
MNUFFT_MODES_INSTANTIATE(8, float)
MNUFFT_MODES_INSTANTIATE(4, double)
MNUFFT_MODES_INSTANTIATE(5, double)
MNUFFT_MODES_INSTANTIATE(6, double)
MNUFFT_MODES_INSTANTIATE(7, double)
MNUFFT_MODES_INSTANTIATE(8, double)
template bool fill_zero<float>(float *, int64_t, hipStream_t);
template bool fill_zero<double>(double *, int64_t, hipStream_t);
template bool fill_zero<uint32_t>(uint32_t *, int64_t, hipStream_t);
template bool type3_final<float>(int64_t, int64_t, const float *, const double *, float *,
                                 hipStream_t, const uint32_t *, int64_t);
template bool type3_final<double>(int64_t, int64_t, const double *, const double *, double *,
                                  hipStream_t, const uint32_t *, int64_t);

}  // namespace mnufft
