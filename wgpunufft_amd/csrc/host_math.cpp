// Host planning math: ES kernel parameters and Fourier coefficients, fine
// grid sizes, type-3 grid lengths, F32 feasibility and the reference's spread
// layout / bin-geometry choice.  Pure f64 host code run at plan creation.
#include <cmath>
#include <cstring>
#include <mutex>

#include "nufft_internal.hpp"

namespace mnufft {

namespace {
thread_local std::string g_last_error;
}  // namespace

void set_error(const std::string &message) { g_last_error = message; }
const char *last_error() { return g_last_error.c_str(); }

bool hip_ok(hipError_t e, const char *what)
{
    if (e == hipSuccess) return true;
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return false;
}

bool point_count_ok(int64_t count, const char *what)
{
    if (count >= 0 && count <= kMaxPoints) return true;
    set_error(std::string(what) + " count " + std::to_string(count) +
              " is outside the 32-bit point index space (0.." + std::to_string(kMaxPoints) + ")");
    return false;
}

// EsKernel::evaluate (reference src/kernel.rs:81-89).
double EsKernel::evaluate(double distance) const
{
    double a = std::fabs(distance);
    if (!std::isfinite(a) || a >= half_width()) return 0.0;
    double s = 2.0 * a / width;
    return std::exp(beta * (std::sqrt(1.0 - s * s) - 1.0));
}

// EsKernel::for_tolerance (reference src/kernel.rs:21-57): w from the
// tolerance, beta from the fixed sigma=2 table or the general formula.
bool es_kernel_for_tolerance(double eps, double sigma, EsKernel *out)
{
    if (!std::isfinite(eps) || eps <= 0.0 || eps >= 1.0) {
        set_error("tolerance must lie strictly between zero and one");
        return false;
    }
    if (!std::isfinite(sigma) || sigma <= 1.0) {
        set_error("upsampling factor must be finite and greater than one");
        return false;
    }
    double wf = sigma == 2.0 ? std::ceil(std::log10(10.0 / eps))
                             : std::log(1.0 / eps) / (M_PI * std::sqrt(1.0 - 1.0 / sigma));
    int64_t w = (int64_t)std::ceil(wf);
    if (w < 2) w = 2;
    if (w > kMaxWidth) {
        set_error("kernel width " + std::to_string(w) + " exceeds the maximum of 16");
        return false;
    }
    double beta;
    if (sigma == 2.0) {
        double r = w == 2 ? 2.20 : w == 3 ? 2.26 : w == 4 ? 2.38 : 2.30;
        beta = r * (double)w;
    } else {
        beta = 0.97 * M_PI * (double)w * (1.0 - 1.0 / (2.0 * sigma));
    }
    out->width = (int)w;
    out->beta = beta;
    out->sigma = sigma;
    out->eps = eps;
    return true;
}

// Lengths rocFFT runs with its own radix kernels: prime factors 2..13.
bool fft_length_supported(int64_t n)
{
    if (n <= 0) return false;
    for (int p : {2, 3, 5, 7, 11, 13})
        while (n % p == 0) n /= p;
    return n == 1;
}

// select_fine_grid_size (reference src/kernel.rs:278-333).
int64_t select_fine_grid_size(int64_t n_modes, double sigma, int width)
{
    if (n_modes <= 0) return -1;
    int64_t scaled = sigma == 2.0 ? n_modes * 2 : (int64_t)std::ceil(sigma * (double)n_modes);
    int64_t minimum = std::max<int64_t>(scaled, 2 * (int64_t)width);
    int64_t c = minimum % 2 == 0 ? minimum : minimum + 1;
    while (c <= (int64_t)UINT32_MAX) {
        if (fft_length_supported(c)) return c;
        c += 2;
    }
    return -1;
}

// Gauss-Legendre nodes and weights by Newton iteration on the Legendre
// recurrence (reference src/kernel.rs:367-407).
void gauss_legendre(int order, double *nodes, double *weights)
{
    auto pair = [order](double x, double *p, double *q) {
        double prev = 1.0;
        if (order == 0) {
            *p = prev;
            *q = 0.0;
            return;
        }
        double cur = x;
        for (int k = 2; k <= order; ++k) {
            double next = ((double)(2 * k - 1) * x * cur - (double)(k - 1) * prev) / (double)k;
            prev = cur;
            cur = next;
        }
        *p = cur;
        *q = prev;
    };
    int pairs = (order + 1) / 2;
    for (int i = 0; i < pairs; ++i) {
        double root = std::cos(M_PI * ((double)i + 0.75) / ((double)order + 0.5));
        double p, q;
        for (;;) {
            pair(root, &p, &q);
            double der = (double)order * (root * p - q) / (root * root - 1.0);
            double next = root - p / der;
            bool done = std::fabs(next - root) <= 4.0 * 2.220446049250313e-16;
            root = next;
            if (done) break;
        }
        pair(root, &p, &q);
        double der = (double)order * (root * p - q) / (root * root - 1.0);
        double w = 2.0 / ((1.0 - root * root) * der * der);
        nodes[i] = -root;
        nodes[order - 1 - i] = root;
        weights[i] = w;
        weights[order - 1 - i] = w;
    }
}

// phi_hat[k], k = 0..n/2, of the kernel on a grid whose coordinate zero is
// -pi, including the (-1)^k phase (reference src/kernel.rs:153-198).
std::vector<double> centered_fourier_coefficients(const EsKernel &k, int64_t n)
{
    double hw = k.half_width();
    int q = (int)std::floor(2.0 + 3.0 * hw);
    std::vector<double> nodes(2 * q), weights(2 * q), qn(q), wk(q);
    gauss_legendre(2 * q, nodes.data(), weights.data());
    for (int i = 0; i < q; ++i) {
        qn[i] = hw * nodes[i];
        wk[i] = hw * weights[i] * k.evaluate(qn[i]);
    }
    std::vector<double> out(n / 2 + 1);
    for (int64_t m = 0; m <= n / 2; ++m) {
        double f = 2.0 * M_PI * (double)m / (double)n;
        double v = 0.0;
        for (int i = 0; i < q; ++i) v += 2.0 * wk[i] * std::cos(f * qn[i]);
        out[m] = (m % 2) ? -v : v;
    }
    return out;
}

// Continuous Fourier transform of the kernel at an arbitrary frequency
// (reference src/type3.rs:688-706).
double es_kernel_fourier_transform(const EsKernel &k, double frequency)
{
    double hw = k.half_width();
    int pn = k.width + 2;
    std::vector<double> nodes(2 * pn), weights(2 * pn);
    gauss_legendre(2 * pn, nodes.data(), weights.data());
    double s = 0.0;
    for (int i = pn; i < 2 * pn; ++i) {
        double d = hw * nodes[i];
        s += 2.0 * hw * weights[i] * k.evaluate(d) * std::cos(frequency * d);
    }
    return s;
}

// Smallest even 2^a 3^b 5^c at least `minimum` (reference
// src/type3.rs:763-797).
int64_t next_even_235(int64_t minimum)
{
    if (minimum < 2) minimum = 2;
    const int64_t cap = (int64_t)1 << 50;
    int64_t best = -1;
    for (int64_t p2 = 2; p2 <= cap; p2 *= 2)
        for (int64_t p3 = p2; p3 <= cap; p3 *= 3)
            for (int64_t c = p3; c <= cap; c *= 5)
                if (c >= minimum) {
                    if (best < 0 || c < best) best = c;
                    break;
                }
    return best;
}

bool validate_f32_rounding(const std::vector<int64_t> &n_modes, double eps, double sigma,
                           Precision precision, const std::vector<std::vector<double>> &coeffs)
{
    if (precision != Precision::F32) return true;
    double estimate = 1.1920928955078125e-07 / 2.0;
    for (size_t a = 0; a < n_modes.size() && a < coeffs.size(); ++a) {
        const auto &axis = coeffs[a];
        if (axis.empty()) continue;
        double center = axis[0];
        int64_t modes = n_modes[a], half = modes / 2;
        double mean = 0.0;
        for (int64_t i = 0; i < modes; ++i) {
            int64_t m = i > half ? i - half : half - i;
            double v = m < (int64_t)axis.size() ? std::fabs(axis[m]) : 0.0;
            mean += v > 0.0 ? (center / v) * (center / v) : INFINITY;
        }
        mean /= (double)modes;
        estimate *= std::sqrt(mean);
    }
    if (estimate > 10.0 * eps && estimate > 1.0e-5) {
        set_error("an F32 plan at eps " + std::to_string(eps) + " and sigma " +
                  std::to_string(sigma) + " has an estimated rounding error of " +
                  std::to_string(estimate) + "; use F64");
        return false;
    }
    return true;
}

// ----------------------------------------------------------------------
// Spread layout (reference src/gpu_nd_spread.rs:194-459) for the WebGPU
// default limits the reference plans with (256 invocations).
// ----------------------------------------------------------------------
namespace {

int64_t pow2_at_most(int64_t v)
{
    int64_t p = 1;
    while (v > 0 && p * 2 <= v) p *= 2;
    return p;
}

int64_t next_pow2(int64_t v)
{
    int64_t p = 1;
    while (p < v) p *= 2;
    return p;
}

int64_t target_workgroup_size(int d, Precision p)
{
    if (d == 1) return 64;
    if (d == 2 && p == Precision::F32) return 32;
    if ((d == 2 || d == 3) && p == Precision::F64) return 64;
    if (p == Precision::F32) return 64;
    return 256;
}

int64_t max_run(int d, Precision p)
{
    if (d == 1 || (d == 2 && p == Precision::Df64)) return 2;
    if (d == 2) return 4;
    if (d == 3 || p != Precision::F32) return 8;
    return 16;
}

}  // namespace

SpreadLayout build_spread_layout(int d, const int64_t *fine, int width, Precision precision)
{
    SpreadLayout L;
    L.dims = d;
    L.width = width;
    for (int a = 0; a < d; ++a) L.fine[a] = fine[a];
    int64_t target = std::min<int64_t>(target_workgroup_size(d, precision), 256);
    auto edge_limit = [&](int a) { return std::max<int64_t>(fine[a] + 1 - width, 1); };
    int64_t run = pow2_at_most(std::min(max_run(d, precision), edge_limit(0)));
    int64_t block[kMaxDims] = {1, 1, 1, 1, 1, 1, 1, 1};
    int64_t inv = 1;
    for (;;) {
        int best = -1;
        for (int a = 1; a < d; ++a) {
            if (block[a] * 2 <= std::min<int64_t>(edge_limit(a), 16) && inv * 2 <= target) {
                if (best < 0 || block[a] < block[best] ||
                    (block[a] == block[best] && fine[a] > fine[best]))
                    best = a;
            }
        }
        if (best < 0) break;
        block[best] *= 2;
        inv *= 2;
    }
    int64_t g0 = 1;
    while (inv * 2 <= target && run * g0 * 2 <= edge_limit(0)) {
        g0 *= 2;
        inv *= 2;
    }
    block[0] = run * g0;
    int64_t wg = g0;
    for (int a = 1; a < d; ++a) wg *= block[a];
    L.run = run;
    L.groups0 = g0;
    L.workgroup_size = wg;
    for (int a = 0; a < d; ++a) L.block[a] = block[a];

    auto mean_reach = [&](const int64_t *side, int a) {
        int64_t reach = block[a] + width - 1;
        double v = (double)(reach - 1) / (double)side[a] + 1.0;
        double cap = (double)((fine[a] + side[a] - 1) / side[a]);
        return std::min(v, cap);
    };
    auto geometry = [&](const int64_t *side) {
        BinGeometry g{};
        for (int a = 0; a < kMaxDims; ++a) {
            g.side[a] = 1;
            g.bins[a] = 1;
            g.reach_bins[a] = 1;
            g.mean_reach[a] = 1.0;
        }
        for (int a = 0; a < d; ++a) {
            int64_t count = (fine[a] + side[a] - 1) / side[a];
            int64_t reach = block[a] + width - 1;
            bool whole = reach + side[a] > fine[a];
            g.side[a] = side[a];
            g.bins[a] = count;
            g.full[a] = whole;
            g.reach_bins[a] = whole ? count : std::min(count, (reach + side[a] - 1) / side[a] + 2);
            g.mean_reach[a] = mean_reach(side, a);
        }
        return g;
    };

    int64_t side[kMaxDims] = {1, 1, 1, 1, 1, 1, 1, 1};
    if (!(d == 2 && precision == Precision::F32)) side[0] = 2;
    for (;;) {
        double rows = 1.0;
        for (int a = 1; a < d; ++a) rows *= mean_reach(side, a);
        double bc = 1.0;
        for (int a = 0; a < d; ++a) bc *= (double)((fine[a] + side[a] - 1) / side[a]);
        if (rows <= 1024.0 && bc <= (double)(1 << 23)) break;
        int widest = -1;
        double wv = 0.0;
        for (int a = 1; a < d; ++a) {
            if (side[a] < fine[a]) {
                double v = mean_reach(side, a);
                if (widest < 0 || v >= wv) {
                    widest = a;  // the last of equal maxima, as Iterator::max_by
                    wv = v;
                }
            }
        }
        if (widest < 0) {
            if (bc > (double)(1 << 23) && side[0] < fine[0])
                widest = 0;
            else
                break;
        }
        side[widest] *= 2;
    }
    BinGeometry finest = geometry(side);
    for (;;) {
        double rows = 1.0;
        for (int a = 1; a < d; ++a) rows *= (double)finest.reach_bins[a];
        if (rows <= (double)(1 << 20)) break;
        int best = -1;
        for (int a = 1; a < d; ++a)
            if (finest.reach_bins[a] > 1 &&
                (best < 0 || finest.reach_bins[a] >= finest.reach_bins[best]))
                best = a;
        if (best < 0) break;
        side[best] *= 2;
        finest = geometry(side);
    }
    L.geometries.push_back(finest);
    for (int others = 1; others <= 3; ++others) {
        for (int first = 0; first <= others; ++first) {
            int64_t s[kMaxDims] = {1, 1, 1, 1, 1, 1, 1, 1};
            for (int a = 0; a < d; ++a)
                s[a] = std::min<int64_t>(side[a] << (a == 0 ? first : others), next_pow2(fine[a]));
            bool dup = false;
            for (const auto &g : L.geometries) {
                bool same = true;
                for (int a = 0; a < d; ++a) same = same && g.side[a] == s[a];
                dup = dup || same;
            }
            if (!dup) L.geometries.push_back(geometry(s));
        }
    }
    int64_t rows0 = 1;
    for (int a = 1; a < d; ++a) rows0 *= L.geometries[0].reach_bins[a];
    rows0 = std::max<int64_t>(rows0, 1);
    L.row_batch = std::max<int64_t>(std::min({(int64_t)128, wg, next_pow2(rows0)}), 1);
    for (int a = 0; a < d; ++a) L.blocks[a] = (fine[a] + block[a] - 1) / block[a];
    return L;
}

int SpreadLayout::choose_geometry(int64_t point_count) const
{
    double cells = 1.0, nblocks = 1.0;
    for (int a = 0; a < dims; ++a) {
        cells *= (double)fine[a];
        nblocks *= (double)blocks[a];
    }
    double density = (double)point_count / cells;
    double dd = (double)dims;
    double record_cost = (double)workgroup_size * (dd + 1.0) * (dd + 1.0) / 1024.0;
    std::vector<double> staged(geometries.size()), cost(geometries.size());
    double least = INFINITY;
    for (size_t g = 0; g < geometries.size(); ++g) {
        const auto &G = geometries[g];
        double rows = 1.0;
        for (int a = 1; a < dims; ++a) rows *= G.mean_reach[a];
        double batches = std::ceil(rows / (double)row_batch);
        double st = 1.0, bins = 1.0;
        for (int a = 0; a < dims; ++a) {
            st *= std::min(G.mean_reach[a] * (double)G.side[a], (double)fine[a]);
            bins *= (double)G.bins[a];
        }
        cost[g] = nblocks * (100.0 * batches + rows + record_cost * density * st) + 0.1 * bins;
        staged[g] = st;
        least = std::min(least, cost[g]);
    }
    int best = -1;
    for (size_t g = 0; g < geometries.size(); ++g)
        if (cost[g] <= least * 1.1 && (best < 0 || staged[g] < staged[best])) best = (int)g;
    return best < 0 ? 0 : best;
}

}  // namespace mnufft
