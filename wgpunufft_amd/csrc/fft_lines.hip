// Line-transform passes of the 3D fine-grid FFT (see fft_lines.hpp).
//
// A workgroup takes `lines` lines of one axis into LDS (element k of line c
// at LDS index c n + k when the axis is the contiguous one, k lines + c
// otherwise, so both loads are coalesced: whole lines, or rows of adjacent
// columns), runs the Stockham passes of its radix plan in place (every
// thread reads its butterflies into registers, the workgroup waits, every
// thread writes them), and stores the lines back.  Pass (radix R, p =
// product of the earlier radices), butterfly i of a line:
//
//   k = i mod p;  u_t = x[i + t n/R] w_n^{t k n/(p R)};  u = DFT_R(u);
//   x'[(i - k) R + k + t p] = u_t,   t = 0 .. R-1,
//
// which leaves the transform in natural order after the last pass.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <vector>

#include "fft_lines.hpp"

namespace mnufft {

namespace {

constexpr int kThreads = 512;
constexpr int kMaxGroupLines = 128;

// Dynamic LDS per workgroup: lines + twiddles, at most 80 KB (two
// workgroups per CU) where the device allows it, else 64 KB; the static
// register arrays are sized for the largest.  Strided axes take `columns`
// adjacent columns (128-byte rows) when they fit.
template <typename T>
struct LdsElems;
template <>
struct LdsElems<float> {
    static constexpr int value = 9728;  // 80 KB less the per-line arrays
    static constexpr int base = 7680;   // 64 KB less the per-line arrays
    static constexpr int columns = 16;
};
template <>
struct LdsElems<double> {
    static constexpr int value = 4864;
    static constexpr int base = 3840;
    static constexpr int columns = 8;
};
// static LDS: line bases, mode rows and two factors per line
constexpr size_t kLineLds = kMaxGroupLines * (2 * sizeof(int64_t) + 2 * sizeof(double));
constexpr size_t kLdsWide = 80 * 1024 - kLineLds;

template <typename T>
struct Cx {
    T re, im;
};

template <typename T>
__device__ inline Cx<T> cmul(Cx<T> a, Cx<T> b)
{
    return Cx<T>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ inline Cx<T> cadd(Cx<T> a, Cx<T> b)
{
    return Cx<T>{a.re + b.re, a.im + b.im};
}
template <typename T>
__device__ inline Cx<T> csub(Cx<T> a, Cx<T> b)
{
    return Cx<T>{a.re - b.re, a.im - b.im};
}
// i s a for real s
template <typename T>
__device__ inline Cx<T> irot(Cx<T> a, T s)
{
    return Cx<T>{-s * a.im, s * a.re};
}

// q = a / b, r = a mod b for 0 <= a < 2^20, from a float reciprocal with
// one correction step.
__device__ inline void divmod_small(int a, int b, float inv_b, int &q, int &r)
{
    q = (int)((float)a * inv_b);
    r = a - q * b;
    if (r < 0) {
        --q;
        r += b;
    } else if (r >= b) {
        ++q;
        r -= b;
    }
}

// Small DFTs, X_s = sum_t u_t exp(DIR 2 pi i s t / R).
template <typename T, int DIR>
__device__ inline void dft2(Cx<T> *u)
{
    const Cx<T> a = u[0], b = u[1];
    u[0] = cadd(a, b);
    u[1] = csub(a, b);
}

template <typename T, int DIR>
__device__ inline void dft3(Cx<T> *u)
{
    const T s = (T)DIR * (T)0.86602540378443864676;  // sin(2 pi / 3)
    const Cx<T> t1 = cadd(u[1], u[2]);
    const Cx<T> d = irot(csub(u[1], u[2]), s);
    const Cx<T> m = Cx<T>{u[0].re - (T)0.5 * t1.re, u[0].im - (T)0.5 * t1.im};
    u[0] = cadd(u[0], t1);
    u[1] = cadd(m, d);
    u[2] = csub(m, d);
}

template <typename T, int DIR>
__device__ inline void dft4(Cx<T> *u)
{
    const Cx<T> a = cadd(u[0], u[2]), b = csub(u[0], u[2]);
    const Cx<T> c = cadd(u[1], u[3]);
    const Cx<T> d = irot(csub(u[1], u[3]), (T)DIR);
    u[0] = cadd(a, c);
    u[2] = csub(a, c);
    u[1] = cadd(b, d);
    u[3] = csub(b, d);
}

template <typename T, int DIR>
__device__ inline void dft5(Cx<T> *u)
{
    const T c1 = (T)0.30901699437494742410;           // cos(2 pi / 5)
    const T c2 = (T)-0.80901699437494742410;          // cos(4 pi / 5)
    const T s1 = (T)DIR * (T)0.95105651629515357212;  // sin(2 pi / 5)
    const T s2 = (T)DIR * (T)0.58778525229247312917;  // sin(4 pi / 5)
    const Cx<T> a1 = cadd(u[1], u[4]), b1 = csub(u[1], u[4]);
    const Cx<T> a2 = cadd(u[2], u[3]), b2 = csub(u[2], u[3]);
    const Cx<T> u0 = u[0];
    const Cx<T> m1 = Cx<T>{u0.re + c1 * a1.re + c2 * a2.re, u0.im + c1 * a1.im + c2 * a2.im};
    const Cx<T> m2 = Cx<T>{u0.re + c2 * a1.re + c1 * a2.re, u0.im + c2 * a1.im + c1 * a2.im};
    // i (s1 b1 + s2 b2) and i (s2 b1 - s1 b2)
    const Cx<T> e1 = Cx<T>{-(s1 * b1.im + s2 * b2.im), s1 * b1.re + s2 * b2.re};
    const Cx<T> e2 = Cx<T>{-(s2 * b1.im - s1 * b2.im), s2 * b1.re - s1 * b2.re};
    u[0] = Cx<T>{u0.re + a1.re + a2.re, u0.im + a1.im + a2.im};
    u[1] = cadd(m1, e1);
    u[4] = csub(m1, e1);
    u[2] = cadd(m2, e2);
    u[3] = csub(m2, e2);
}

// Even and odd halves as two DFT4s, the odd half turned by w^s, w = exp(DIR 2 pi i / 8).
template <typename T, int DIR>
__device__ inline void dft8(Cx<T> *u)
{
    const T h = (T)0.70710678118654752440;  // cos(pi / 4)
    const T d = (T)DIR;
    Cx<T> e[4] = {u[0], u[2], u[4], u[6]};
    Cx<T> o[4] = {u[1], u[3], u[5], u[7]};
    dft4<T, DIR>(e);
    dft4<T, DIR>(o);
    const Cx<T> o1 = Cx<T>{h * (o[1].re - d * o[1].im), h * (o[1].im + d * o[1].re)};
    const Cx<T> o2 = irot(o[2], d);
    const Cx<T> o3 = Cx<T>{h * (-o[3].re - d * o[3].im), h * (-o[3].im + d * o[3].re)};
    u[0] = cadd(e[0], o[0]);
    u[4] = csub(e[0], o[0]);
    u[1] = cadd(e[1], o1);
    u[5] = csub(e[1], o1);
    u[2] = cadd(e[2], o2);
    u[6] = csub(e[2], o2);
    u[3] = cadd(e[3], o3);
    u[7] = csub(e[3], o3);
}

// cos and sin of 2 pi k / R, k = 1 .. (R - 1) / 2, for the prime radices 7, 11, 13.
__host__ __device__ constexpr double prime_cos(int R, int k)
{
    constexpr double c7[3] = {0.6234898018587336, -0.22252093395631434, -0.900968867902419};
    constexpr double c11[5] = {0.8412535328311812, 0.41541501300188644, -0.142314838273285,
                               -0.654860733945285, -0.9594929736144974};
    constexpr double c13[6] = {0.8854560256532099, 0.5680647467311559, 0.120536680255323,
                               -0.35460488704253545, -0.7485107481711012, -0.970941817426052};
    return R == 7 ? c7[k - 1] : (R == 11 ? c11[k - 1] : c13[k - 1]);
}
__host__ __device__ constexpr double prime_sin(int R, int k)
{
    constexpr double s7[3] = {0.7818314824680298, 0.9749279121818236, 0.43388373911755823};
    constexpr double s11[5] = {0.5406408174555976, 0.9096319953545183, 0.9898214418809328,
                               0.7557495743542583, 0.28173255684142967};
    constexpr double s13[6] = {0.4647231720437685, 0.8229838658936564, 0.992708874098054,
                               0.9350162426854148, 0.6631226582407952, 0.23931566428755768};
    return R == 7 ? s7[k - 1] : (R == 11 ? s11[k - 1] : s13[k - 1]);
}

// Odd prime R from the pairs t, R - t:
//   X_s = u_0 + sum_t (u_t + u_{R-t}) cos(2 pi s t / R) + i DIR (u_t - u_{R-t}) sin(2 pi s t / R),
// t = 1 .. (R - 1) / 2, and X_{R-s} with the sine terms negated.
template <typename T, int DIR, int R>
__device__ inline void dft_prime(Cx<T> *u)
{
    constexpr int H = (R - 1) / 2;
    Cx<T> a[H], b[H];
#pragma unroll
    for (int t = 1; t <= H; ++t) {
        a[t - 1] = cadd(u[t], u[R - t]);
        b[t - 1] = csub(u[t], u[R - t]);
    }
    Cx<T> out[R];
    out[0] = u[0];
#pragma unroll
    for (int t = 0; t < H; ++t) out[0] = cadd(out[0], a[t]);
#pragma unroll
    for (int s = 1; s <= H; ++s) {
        Cx<T> m = u[0], e = Cx<T>{(T)0, (T)0};
#pragma unroll
        for (int t = 1; t <= H; ++t) {
            const int idx = (s * t) % R;
            const T c = (T)(idx <= H ? prime_cos(R, idx) : prime_cos(R, R - idx));
            const T sn = (T)(idx <= H ? prime_sin(R, idx) : -prime_sin(R, R - idx));
            m = Cx<T>{m.re + c * a[t - 1].re, m.im + c * a[t - 1].im};
            e = Cx<T>{e.re + sn * b[t - 1].re, e.im + sn * b[t - 1].im};
        }
        const Cx<T> ie = Cx<T>{-(T)DIR * e.im, (T)DIR * e.re};
        out[s] = cadd(m, ie);
        out[R - s] = csub(m, ie);
    }
#pragma unroll
    for (int t = 0; t < R; ++t) u[t] = out[t];
}

template <typename T, int DIR, int R>
__device__ inline void dft(Cx<T> *u)
{
    if constexpr (R == 7 || R == 11 || R == 13) dft_prime<T, DIR, R>(u);
    if constexpr (R == 2) dft2<T, DIR>(u);
    if constexpr (R == 3) dft3<T, DIR>(u);
    if constexpr (R == 4) dft4<T, DIR>(u);
    if constexpr (R == 5) dft5<T, DIR>(u);
    if constexpr (R == 8) dft8<T, DIR>(u);
}

// One Stockham pass over the workgroup's lines.  kFromGrid: the pass reads
// its butterflies' inputs from the grid (the first pass: p = 1, no
// twiddles; cells outside the read window are zeros); kToGrid: it writes its
// outputs to the grid (the last pass: p R = n; cells outside the write
// window are dropped).  Otherwise LDS in place: every thread reads its
// butterflies into registers, the workgroup waits, every thread writes.
constexpr int kFromGrid = 1, kToGrid = 2;

// Per-line data of a workgroup in LDS: grid base, mode row and the line's
// deconvolution factors (type 2: the u and v axes' 1/phi_hat; type 1: their
// product in axis order).
struct LineData {
    int64_t *base, *mrow;
    double *f1, *f2;
    void *modes;  // this vector's modes (IO != 0)
};

// Mode storage index of signed frequency m on an axis of N modes.
__device__ inline int64_t mode_slot(int64_t m, int64_t N, int centered)
{
    return centered ? m + N / 2 : (m >= 0 ? m : m + N);
}

// IO (the launch's mode reads or writes): 0 grid only; 1 the first pass
// reads the modes and predeconvolves them (type 2); 2 the last pass writes
// the deconvolved modes (type 1).
template <typename T, int DIR, int R, int LDS, int MODE, int IO>
__device__ inline void line_pass(Cx<T> *X, const Cx<T> *tw, const LineFftAxis &P,
                                 const LineFftPass &ps, Cx<T> *__restrict__ grid,
                                 const LineData &ld)
{
    const int64_t *base = ld.base;
    Cx<T> *modes = (Cx<T> *)ld.modes;
    constexpr int MAXB = (LDS / R + kThreads - 1) / kThreads;
    const int n = P.n, nb = n / R, total = P.lines * nb;
    const int ks = P.line_major ? 1 : P.lines, cs = P.line_major ? n : 1;
    const float inv_nb = 1.0f / (float)nb, inv_p = 1.0f / (float)ps.p;
    Cx<T> u[MAXB][R];
    int dst[MAXB], line[MAXB];
#pragma unroll
    for (int s = 0; s < MAXB; ++s) {
        const int b = (int)threadIdx.x + s * kThreads;
        dst[s] = -1;
        if (b < total) {
            int c, i;
            if (P.line_major) {
                divmod_small(b, nb, inv_nb, c, i);
            } else {
                i = b >> P.log_lines;
                c = b & (P.lines - 1);
            }
            int q, k;
            divmod_small(i, ps.p, inv_p, q, k);
            line[s] = c;
            if constexpr ((MODE & kFromGrid) != 0) {
                const int64_t bc = base[c];
                if constexpr (IO == 1) {
                    // the predeconvolution: mode times (1/phi_hat) in axis order
                    const int64_t mr = ld.mrow[c];
                    const double g1 = ld.f1[c], g2 = ld.f2[c];
#pragma unroll
                    for (int t = 0; t < R; ++t) {
                        const int e = i + t * nb;
                        u[s][t] = Cx<T>{(T)0, (T)0};
                        if (bc >= 0 && (e < P.in_hi || e >= n - P.in_lo)) {
                            const int64_t m = e < P.in_hi ? e : (int64_t)e - n;
                            double amp = 1.0;
                            amp *= P.inv_phihat[P.eph + (m < 0 ? -m : m)];
                            amp *= g1;
                            amp *= g2;
                            const Cx<T> src = modes[mr + mode_slot(m, P.eN, P.centered) * P.ems];
                            const T ta = (T)amp;
                            u[s][t] = Cx<T>{src.re * ta, src.im * ta};
                        }
                    }
                } else {
#pragma unroll
                    for (int t = 0; t < R; ++t) {
                        const int e = i + t * nb;
                        u[s][t] = Cx<T>{(T)0, (T)0};
                        if (bc >= 0 && (e < P.in_hi || e >= n - P.in_lo))
                            u[s][t] = grid[bc + (int64_t)e * P.estride];
                    }
                }
            } else {
#pragma unroll
                for (int t = 0; t < R; ++t) u[s][t] = X[(i + t * nb) * ks + c * cs];
#pragma unroll
                for (int t = 1; t < R; ++t) u[s][t] = cmul(u[s][t], tw[t * k * ps.tw_step]);
            }
            dft<T, DIR, R>(u[s]);
            dst[s] = q * ps.p * R + k;
        }
    }
    if constexpr ((MODE & kToGrid) != 0) {
#pragma unroll
        for (int s = 0; s < MAXB; ++s) {
            if (dst[s] >= 0) {
                const int64_t bc = base[line[s]];
                if constexpr (IO == 2) {
                    // the deconvolution: cell times (1/phi_hat) in axis order
                    const int64_t mr = ld.mrow[line[s]];
                    const double g = ld.f1[line[s]];
#pragma unroll
                    for (int t = 0; t < R; ++t) {
                        const int e = dst[s] + t * ps.p;
                        if (bc >= 0 && (e < P.out_hi || e >= n - P.out_lo)) {
                            const int64_t m = e < P.out_hi ? e : (int64_t)e - n;
                            double amp = g;
                            amp *= P.inv_phihat[P.eph + (m < 0 ? -m : m)];
                            const T ta = (T)amp;
                            modes[mr + mode_slot(m, P.eN, P.centered) * P.ems] =
                                Cx<T>{u[s][t].re * ta, u[s][t].im * ta};
                        }
                    }
                } else {
#pragma unroll
                    for (int t = 0; t < R; ++t) {
                        const int e = dst[s] + t * ps.p;
                        if (bc >= 0 && (e < P.out_hi || e >= n - P.out_lo))
                            grid[bc + (int64_t)e * P.estride] = u[s][t];
                    }
                }
            }
        }
    } else {
        if constexpr ((MODE & kFromGrid) == 0) __syncthreads();
#pragma unroll
        for (int s = 0; s < MAXB; ++s) {
            if (dst[s] >= 0) {
#pragma unroll
                for (int t = 0; t < R; ++t)
                    X[(dst[s] + t * ps.p) * ks + line[s] * cs] = u[s][t];
            }
        }
        __syncthreads();
    }
}

template <typename T, int DIR, int LDS, int MODE, int IO>
__device__ inline void line_pass_any(Cx<T> *X, const Cx<T> *tw, const LineFftAxis &P,
                                     const LineFftPass &ps, Cx<T> *grid, const LineData &base)
{
    switch (ps.radix) {
    case 2: line_pass<T, DIR, 2, LDS, MODE, IO>(X, tw, P, ps, grid, base); break;
    case 3: line_pass<T, DIR, 3, LDS, MODE, IO>(X, tw, P, ps, grid, base); break;
    case 4: line_pass<T, DIR, 4, LDS, MODE, IO>(X, tw, P, ps, grid, base); break;
    case 5: line_pass<T, DIR, 5, LDS, MODE, IO>(X, tw, P, ps, grid, base); break;
    case 7: line_pass<T, DIR, 7, LDS, MODE, IO>(X, tw, P, ps, grid, base); break;
    case 8: line_pass<T, DIR, 8, LDS, MODE, IO>(X, tw, P, ps, grid, base); break;
    case 11: line_pass<T, DIR, 11, LDS, MODE, IO>(X, tw, P, ps, grid, base); break;
    default: line_pass<T, DIR, 13, LDS, MODE, IO>(X, tw, P, ps, grid, base); break;
    }
}

// The transform of workgroup `group`'s lines of one axis pass (P) of the
// grid `grid` (one vector), X: (lines + 1) n elements of LDS.
template <typename T, int DIR, int IO>
__device__ __forceinline__ void fft_lines_body(const LineFftAxis &P, Cx<T> *__restrict__ grid,
                                               const Cx<T> *__restrict__ tw_global,
                                               int64_t group, Cx<T> *X, const LineData &ld)
{
    int64_t *base = ld.base;
    constexpr int LDS = LdsElems<T>::value;
    Cx<T> *tw = X + P.lines * P.n;
    const int n = P.n;
    for (int m = threadIdx.x; m < n; m += kThreads) tw[m] = tw_global[m];
    const int64_t nlines = P.nu * P.nv;
    const int64_t line0 = group * P.lines;
    for (int c = threadIdx.x; c < P.lines; c += kThreads) {
        const int64_t L = line0 + c;
        int64_t b = -1;
        if (L < nlines) {
            const int64_t u = L % P.nu, w = L / P.nu;
            const int64_t uu = u < P.u_hi ? u : u + P.u_gap;
            const int64_t vv = w < P.v_hi ? w : w + P.v_gap;
            b = uu * P.su + vv * P.sv;
            if constexpr (IO != 0) {
                // the line's mode row and factors (u, v: fine indices of the
                // kept windows of their axes)
                int64_t mr = 0;
                double fu = 1.0, fv = 1.0;
                if (P.has_u) {
                    const int64_t m = uu < (P.uN + 1) / 2 ? uu : uu - P.un;
                    mr += mode_slot(m, P.uN, P.centered) * P.ums;
                    fu = P.inv_phihat[P.uph + (m < 0 ? -m : m)];
                }
                if (P.has_v) {
                    const int64_t m = vv < (P.vN + 1) / 2 ? vv : vv - P.vn;
                    mr += mode_slot(m, P.vN, P.centered) * P.vms;
                    fv = P.inv_phihat[P.vph + (m < 0 ? -m : m)];
                }
                ld.mrow[c] = mr;
                if constexpr (IO == 1) {
                    ld.f1[c] = fu;
                    ld.f2[c] = fv;
                } else {
                    double amp = 1.0;
                    amp *= fu;
                    if (P.has_v) amp *= fv;
                    ld.f1[c] = amp;
                }
            }
        }
        base[c] = b;
    }
    __syncthreads();
    const int np = P.npass;
    if (np == 1) {
        line_pass_any<T, DIR, LDS, kFromGrid | kToGrid, IO>(X, tw, P, P.pass[0], grid, ld);
        return;
    }
    if (np == 0) return;
    // first pass from the grid, the middle ones in LDS grouped by radix (5s,
    // 3s, 7s, 11s, 13s, 8s, then a 4 or a 2; separate loops keep each
    // radix's registers to itself), the last one to the grid
    line_pass_any<T, DIR, LDS, kFromGrid, IO == 1 ? 1 : 0>(X, tw, P, P.pass[0], grid, ld);
    const int last = np - 1;
    int s = 1;
    for (; s < last && P.pass[s].radix == 5; ++s)
        line_pass<T, DIR, 5, LDS, 0, 0>(X, tw, P, P.pass[s], grid, ld);
    for (; s < last && P.pass[s].radix == 3; ++s)
        line_pass<T, DIR, 3, LDS, 0, 0>(X, tw, P, P.pass[s], grid, ld);
    for (; s < last && P.pass[s].radix == 7; ++s)
        line_pass<T, DIR, 7, LDS, 0, 0>(X, tw, P, P.pass[s], grid, ld);
    for (; s < last && P.pass[s].radix == 11; ++s)
        line_pass<T, DIR, 11, LDS, 0, 0>(X, tw, P, P.pass[s], grid, ld);
    for (; s < last && P.pass[s].radix == 13; ++s)
        line_pass<T, DIR, 13, LDS, 0, 0>(X, tw, P, P.pass[s], grid, ld);
    for (; s < last && P.pass[s].radix == 8; ++s)
        line_pass<T, DIR, 8, LDS, 0, 0>(X, tw, P, P.pass[s], grid, ld);
    for (; s < last && P.pass[s].radix == 4; ++s)
        line_pass<T, DIR, 4, LDS, 0, 0>(X, tw, P, P.pass[s], grid, ld);
    for (; s < last && P.pass[s].radix == 2; ++s)
        line_pass<T, DIR, 2, LDS, 0, 0>(X, tw, P, P.pass[s], grid, ld);
    line_pass_any<T, DIR, LDS, kToGrid, IO == 2 ? 2 : 0>(X, tw, P, P.pass[last], grid, ld);
}

// One plan's axis pass: workgroup (group, vector); IO / modes_all: this
// launch's mode reads or writes (all vectors' modes).
template <typename T, int DIR, int IO>
__global__ __launch_bounds__(kThreads) void k_fft_lines(LineFftAxis P,
                                                        void *modes_all, T *__restrict__ data,
                                                        const Cx<T> *__restrict__ tw_global)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    Cx<T> *grid = (Cx<T> *)data + (int64_t)blockIdx.y * P.vec_stride;
    if constexpr (IO == 0) {
        __shared__ int64_t base[kMaxGroupLines];
        fft_lines_body<T, DIR, 0>(P, grid, tw_global, blockIdx.x, (Cx<T> *)lds_raw,
                                  LineData{base, nullptr, nullptr, nullptr, nullptr});
    } else {
        __shared__ int64_t base[kMaxGroupLines], mrow[kMaxGroupLines];
        __shared__ double f1[kMaxGroupLines], f2[kMaxGroupLines];
        Cx<T> *modes = (Cx<T> *)modes_all + (int64_t)blockIdx.y * P.mode_total;
        fft_lines_body<T, DIR, IO>(P, grid, tw_global, blockIdx.x, (Cx<T> *)lds_raw,
                                   LineData{base, mrow, f1, f2, modes});
    }
}

// The same axis pass of many plans (one vector each) in one launch: job q
// takes workgroups [group0[q], group0[q + 1]).
template <typename T, int DIR>
__global__ __launch_bounds__(kThreads) void k_fft_lines_many(const LineFftJob *__restrict__ jobs,
                                                             int njobs)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ int64_t base[kMaxGroupLines];
    const int64_t b = blockIdx.x;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].group0 <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    const LineFftJob &J = jobs[lo];
    fft_lines_body<T, DIR, 0>(J.ax, (Cx<T> *)J.grid, (const Cx<T> *)J.tw, b - J.group0,
                              (Cx<T> *)lds_raw, LineData{base, nullptr, nullptr, nullptr, nullptr});
}

// Radix plan: 5s, 3s, 7s, 11s, 13s, 8s, then a 4 or a 2; false for other
// prime factors.
bool radix_plan(int n, LineFftAxis *ax)
{
    std::vector<int> r;
    int m = n;
    while (m % 5 == 0) {
        r.push_back(5);
        m /= 5;
    }
    while (m % 3 == 0) {
        r.push_back(3);
        m /= 3;
    }
    for (int q : {7, 11, 13})
        while (m % q == 0) {
            r.push_back(q);
            m /= q;
        }
    while (m % 8 == 0) {
        r.push_back(8);
        m /= 8;
    }
    while (m % 4 == 0) {
        r.push_back(4);
        m /= 4;
    }
    while (m % 2 == 0) {
        r.push_back(2);
        m /= 2;
    }
    if (m != 1 || (int)r.size() > kLineFftMaxPasses) return false;
    ax->n = n;
    ax->npass = (int)r.size();
    int p = 1;
    for (size_t i = 0; i < r.size(); ++i) {
        ax->pass[i] = LineFftPass{r[i], p, n / (p * r[i])};
        p *= r[i];
    }
    ax->inv_n = 1.0f / (float)n;
    return true;
}

// Lines per workgroup: the lines and the twiddle table fit the LDS array;
// along the contiguous axis a line per thread at least, else few enough
// lines for about 2048 workgroups (`count` lines in all), small groups
// sharing a CU.
// Whether the line kernels may take the wide LDS budget (set once).
bool wide_lds()
{
    static std::once_flag once;
    static bool ok = false;
    std::call_once(once, [] {
        const void *kernels[] = {
            (const void *)k_fft_lines<float, 1, 0>,      (const void *)k_fft_lines<float, 1, 1>,
            (const void *)k_fft_lines<float, 1, 2>,      (const void *)k_fft_lines<float, -1, 0>,
            (const void *)k_fft_lines<float, -1, 1>,     (const void *)k_fft_lines<float, -1, 2>,
            (const void *)k_fft_lines<double, 1, 0>,     (const void *)k_fft_lines<double, 1, 1>,
            (const void *)k_fft_lines<double, 1, 2>,     (const void *)k_fft_lines<double, -1, 0>,
            (const void *)k_fft_lines<double, -1, 1>,    (const void *)k_fft_lines<double, -1, 2>,
            (const void *)k_fft_lines_many<float, 1>,    (const void *)k_fft_lines_many<float, -1>,
            (const void *)k_fft_lines_many<double, 1>,   (const void *)k_fft_lines_many<double, -1>};
        ok = true;
        for (const void *k : kernels)
            ok = ok && hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)kLdsWide) == hipSuccess;
        (void)hipGetLastError();
    });
    return ok;
}

// `wide`: the 80 KB budget, else 64 KB (the support test uses 64 KB).
bool group_lines(int n, bool line_major, bool f64, int64_t count, bool wide, LineFftAxis *ax)
{
    const int lds = f64 ? (wide ? LdsElems<double>::value : LdsElems<double>::base)
                        : (wide ? LdsElems<float>::value : LdsElems<float>::base);
    if (2 * n > lds) return false;
    ax->line_major = line_major ? 1 : 0;
    if (line_major) {
        const int64_t fill = std::max<int64_t>((kThreads + n - 1) / n, count / 2048);
        int c = (int)std::min<int64_t>(kMaxGroupLines, fill);
        while (c > 1 && (c + 1) * n > lds) --c;
        ax->lines = c;
        ax->log_lines = 0;
    } else {
        int c = f64 ? LdsElems<double>::columns : LdsElems<float>::columns;
        while (c > 1 && (c + 1) * n > lds) c /= 2;
        ax->lines = c;
        ax->log_lines = 0;
        while ((1 << ax->log_lines) < c) ++ax->log_lines;
    }
    return ax->lines >= 1;
}

// Keep only the window of N modes of an axis of n cells.
void keep_u(LineFftAxis &a, int64_t n, int64_t N)
{
    a.nu = N;
    a.u_hi = (N + 1) / 2;
    a.u_gap = n - N;
}
void keep_v(LineFftAxis &a, int64_t n, int64_t N)
{
    a.nv = N;
    a.v_hi = (N + 1) / 2;
    a.v_gap = n - N;
}

}  // namespace

bool LineFft::supported(int dim, const int64_t *fine, bool f64)
{
    if (dim != 2 && dim != 3) return false;
    for (int a = 0; a < dim; ++a) {
        LineFftAxis ax;
        if (fine[a] < 1 || fine[a] > 65536 || !radix_plan((int)fine[a], &ax) ||
            !group_lines((int)fine[a], a == 0, f64, 1, wide_lds(), &ax))
            return false;
    }
    return true;
}

int64_t LineFft::row_length_limit(bool f64)
{
    return f64 ? LdsElems<double>::value / 5 : LdsElems<float>::value / 9;
}

LineFft::~LineFft()
{
    for (void *&p : tw_)
        if (p) (void)hipFree(p);
}

bool LineFft::init(int dim, const int64_t *fine, const int64_t *n_modes, int type, bool f64,
                   int isign, const double *inv_phihat, const int64_t *phihat_off,
                   bool centered)
{
    if (!supported(dim, fine, f64)) {
        set_error("line FFT needs a 2D or 3D grid of 2,3,5,7,11,13-smooth lengths within the "
                  "LDS lines");
        return false;
    }
    naxes_ = dim;
    type_ = type;
    f64_ = f64;
    dir_ = isign >= 0 ? 1 : -1;
    const int64_t n0 = fine[0], n1 = fine[1], n2 = dim == 3 ? fine[2] : 1;
    fine_total_ = n0 * n1 * n2;
    const int64_t plane = n0 * n1;
    for (int a = 0; a < dim; ++a) {
        LineFftAxis &x = ax_[a];
        x = LineFftAxis{};
        radix_plan((int)fine[a], &x);
        x.in_hi = x.out_hi = (int)fine[a];
        x.in_lo = x.out_lo = 0;
        x.vec_stride = fine_total_;
    }
    // x: lines (y, z); y: columns (x, z); z: columns (x, y); a 2D grid has no z
    ax_[0].estride = 1;
    ax_[0].nu = ax_[0].u_hi = n1;
    ax_[0].su = n0;
    ax_[0].nv = ax_[0].v_hi = n2;
    ax_[0].sv = plane;
    ax_[1].estride = n0;
    ax_[1].nu = ax_[1].u_hi = n0;
    ax_[1].su = 1;
    ax_[1].nv = ax_[1].v_hi = n2;
    ax_[1].sv = plane;
    if (dim == 3) {
        ax_[2].estride = plane;
        ax_[2].nu = ax_[2].u_hi = n0;
        ax_[2].su = 1;
        ax_[2].nv = ax_[2].v_hi = n1;
        ax_[2].sv = n0;
    }
    const int64_t N0 = n_modes[0], N1 = n_modes[1], N2 = dim == 3 ? n_modes[2] : 1;
    auto window = [](int &hi, int &lo, int64_t N) {
        hi = (int)((N + 1) / 2);
        lo = (int)(N / 2);
    };
    if (type == 2) {
        keep_u(ax_[0], n1, N1);
        window(ax_[0].in_hi, ax_[0].in_lo, N0);
        window(ax_[1].in_hi, ax_[1].in_lo, N1);
        if (dim == 3) {
            keep_v(ax_[0], n2, N2);
            keep_v(ax_[1], n2, N2);
            window(ax_[2].in_hi, ax_[2].in_lo, N2);
        }
    } else {
        window(ax_[0].out_hi, ax_[0].out_lo, N0);
        keep_u(ax_[1], n0, N0);
        window(ax_[1].out_hi, ax_[1].out_lo, N1);
        if (dim == 3) {
            keep_u(ax_[2], n0, N0);
            keep_v(ax_[2], n1, N1);
            window(ax_[2].out_hi, ax_[2].out_lo, N2);
        }
    }
    const bool wide = wide_lds();
    for (int a = 0; a < dim; ++a)
        group_lines((int)fine[a], a == 0, f64, ax_[a].nu * ax_[a].nv, wide, &ax_[a]);
    // the pass that reads (type 2: x) or writes (type 1: the last axis) modes
    LineFftAxis &mx = type == 2 ? ax_[0] : ax_[dim - 1];
    mx.centered = centered ? 1 : 0;
    mx.inv_phihat = inv_phihat;
    int64_t mstride[3] = {1, 1, 1};
    int64_t stride = 1;
    for (int a = 0; a < dim; ++a) {
        mstride[a] = stride;
        stride *= n_modes[a];
    }
    mx.mode_total = stride;
    // type 2: x lines over (y, z); type 1: last-axis lines over (x, y)
    const int ua = type == 2 ? 1 : 0, va = dim == 3 ? (type == 2 ? 2 : 1) : -1;
    const int ea = type == 2 ? 0 : dim - 1;
    mx.has_u = 1;
    mx.uN = n_modes[ua];
    mx.un = fine[ua];
    mx.ums = mstride[ua];
    mx.uph = phihat_off[ua];
    mx.has_v = va >= 0 ? 1 : 0;
    if (va >= 0) {
        mx.vN = n_modes[va];
        mx.vn = fine[va];
        mx.vms = mstride[va];
        mx.vph = phihat_off[va];
    }
    mx.eN = n_modes[ea];
    mx.ems = mstride[ea];
    mx.eph = phihat_off[ea];
    // twiddles w_n^m = exp(dir 2 pi i m / n), from long double
    for (int a = 0; a < dim; ++a) {
        const int n = (int)fine[a];
        const size_t eb = f64 ? 2 * sizeof(double) : 2 * sizeof(float);
        std::vector<double> hd(f64 ? 2 * (size_t)n : 0);
        std::vector<float> hf(f64 ? 0 : 2 * (size_t)n);
        for (int m = 0; m < n; ++m) {
            const long double ang =
                2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
            const long double c = cosl(ang), s = (long double)dir_ * sinl(ang);
            if (f64) {
                hd[2 * m] = (double)c;
                hd[2 * m + 1] = (double)s;
            } else {
                hf[2 * m] = (float)c;
                hf[2 * m + 1] = (float)s;
            }
        }
        if (tw_[a]) (void)hipFree(tw_[a]);
        tw_[a] = nullptr;
        if (!hip_ok(hipMalloc(&tw_[a], eb * (size_t)n), "line FFT twiddles") ||
            !hip_ok(hipMemcpy(tw_[a], f64 ? (const void *)hd.data() : (const void *)hf.data(),
                              eb * (size_t)n, hipMemcpyHostToDevice),
                    "line FFT twiddle upload"))
            return false;
    }
    return true;
}

namespace {

struct JobChunk {
    static constexpr int N = (int)((4096 - 64) / sizeof(LineFftJob));
    LineFftJob j[N];
    int n;
};

__global__ void k_line_jobs(JobChunk chunk, LineFftJob *dst)
{
    if ((int)threadIdx.x < chunk.n) dst[threadIdx.x] = chunk.j[threadIdx.x];
}

}  // namespace

bool LineFft::run_many(const std::vector<const LineFft *> &ffts, const std::vector<void *> &grids,
                       LineFftJob **jobs, int64_t *cap, hipStream_t stream)
{
    static_assert(JobChunk::N >= 1 && sizeof(JobChunk) <= 4096, "job chunk");
    if (ffts.empty()) return true;
    const int naxes = ffts[0]->naxes_;
    const bool f64 = ffts[0]->f64_;
    for (const LineFft *f : ffts)
        if (f->naxes_ != naxes || f->f64_ != f64) {
            set_error("batched line FFT: plans of one rank and precision");
            return false;
        }
    // jobs grouped by (axis, sign), each group's workgroups numbered from 0
    struct Launch {
        int dir;
        size_t first, count;
        int64_t groups;
        size_t lds;
    };
    std::vector<LineFftJob> all;
    std::vector<Launch> launches;
    for (int a = 0; a < naxes; ++a)
        for (int dir : {1, -1}) {
            Launch L{dir, all.size(), 0, 0, 0};
            for (size_t i = 0; i < ffts.size(); ++i) {
                const LineFft &f = *ffts[i];
                if (f.dir_ != dir) continue;
                const LineFftAxis &x = f.ax_[a];
                const int64_t g = (x.nu * x.nv + x.lines - 1) / x.lines;
                if (g == 0) continue;
                all.push_back(LineFftJob{x, grids[i], f.tw_[a], L.groups});
                L.groups += g;
                L.lds = std::max(L.lds, (size_t)(x.lines + 1) * (size_t)x.n * (f64 ? 16 : 8));
            }
            L.count = all.size() - L.first;
            if (L.count) launches.push_back(L);
        }
    const int64_t n = (int64_t)all.size();
    if (n > *cap) {
        if (*jobs) (void)hipFree(*jobs);
        *jobs = nullptr;
        *cap = 0;
        if (!hip_ok(hipMalloc(jobs, sizeof(LineFftJob) * n), "line FFT job table")) return false;
        *cap = n;
    }
    for (int64_t i0 = 0; i0 < n; i0 += JobChunk::N) {
        JobChunk chunk;
        chunk.n = (int)std::min<int64_t>(JobChunk::N, n - i0);
        for (int k = 0; k < chunk.n; ++k) chunk.j[k] = all[(size_t)(i0 + k)];
        k_line_jobs<<<1, 64, 0, stream>>>(chunk, *jobs + i0);
    }
    for (const Launch &L : launches) {
        if (L.groups > 0x7fffffff) {
            set_error("batched line FFT: too many workgroups");
            return false;
        }
        const dim3 g((unsigned)L.groups);
        const LineFftJob *J = *jobs + L.first;
        if (f64) {
            if (L.dir > 0)
                k_fft_lines_many<double, 1><<<g, kThreads, L.lds, stream>>>(J, (int)L.count);
            else
                k_fft_lines_many<double, -1><<<g, kThreads, L.lds, stream>>>(J, (int)L.count);
        } else {
            if (L.dir > 0)
                k_fft_lines_many<float, 1><<<g, kThreads, L.lds, stream>>>(J, (int)L.count);
            else
                k_fft_lines_many<float, -1><<<g, kThreads, L.lds, stream>>>(J, (int)L.count);
        }
    }
    return hip_ok(hipGetLastError(), "batched line FFT launch");
}

namespace {

template <typename T>
void launch_axis(int dir, int io, dim3 g, size_t lds, hipStream_t stream, const LineFftAxis *hx,
                 void *mv, T *d, const Cx<T> *t)
{
    if (dir > 0) {
        if (io == 1)
            k_fft_lines<T, 1, 1><<<g, kThreads, lds, stream>>>(*hx, mv, d, t);
        else if (io == 2)
            k_fft_lines<T, 1, 2><<<g, kThreads, lds, stream>>>(*hx, mv, d, t);
        else
            k_fft_lines<T, 1, 0><<<g, kThreads, lds, stream>>>(*hx, mv, d, t);
    } else {
        if (io == 1)
            k_fft_lines<T, -1, 1><<<g, kThreads, lds, stream>>>(*hx, mv, d, t);
        else if (io == 2)
            k_fft_lines<T, -1, 2><<<g, kThreads, lds, stream>>>(*hx, mv, d, t);
        else
            k_fft_lines<T, -1, 0><<<g, kThreads, lds, stream>>>(*hx, mv, d, t);
    }
}

}  // namespace

bool LineFft::run(void *data, int64_t vectors, hipStream_t stream, void *modes) const
{
    if (vectors <= 0) return true;
    const size_t eb = f64_ ? 16 : 8;
    for (int64_t v0 = 0; v0 < vectors; v0 += 65535) {
        const unsigned vc = (unsigned)std::min<int64_t>(65535, vectors - v0);
        for (int a = 0; a < naxes_; ++a) {
            const LineFftAxis &x = ax_[a];
            const bool io = modes && a == (type_ == 2 ? 0 : naxes_ - 1);
            const int mode_io = io ? (type_ == 2 ? 1 : 2) : 0;
            void *mv = io ? (void *)((char *)modes + (size_t)v0 * (size_t)x.mode_total * eb)
                          : nullptr;
            const int64_t groups = (x.nu * x.nv + x.lines - 1) / x.lines;
            if (groups == 0) continue;
            const dim3 g((unsigned)groups, vc);
            const size_t lds = (size_t)(x.lines + 1) * (size_t)x.n * eb;
            const LineFftAxis *hx = &ax_[a];
            if (f64_) {
                double *d = (double *)data + 2 * v0 * fine_total_;
                launch_axis<double>(dir_, mode_io, g, lds, stream, hx, mv, d,
                                    (const Cx<double> *)tw_[a]);
            } else {
                float *d = (float *)data + 2 * v0 * fine_total_;
                launch_axis<float>(dir_, mode_io, g, lds, stream, hx, mv, d,
                                   (const Cx<float> *)tw_[a]);
            }
        }
    }
    return hip_ok(hipGetLastError(), "line FFT launch");
}

}  // namespace mnufft
