/*
 * CPU oracle for the NUFFT hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C with f64 (and binary128 where exactness
 * matters), the reference algorithm of the GPU execution path so that the
 * HIP product can be checked against it.  Only tests/, the smoke check and
 * the CPU-baseline leg of bench.py may load it; the product never does.
 *
 * Each function cites the reference file:line it follows (paths relative to
 * the reference checkout).  Pinning of every function is listed in the
 * header comment of oracle/oracle.py.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>


#define OR_MAX_DIMS 8
#define OR_MAX_GEOMETRIES 16

/* ------------------------------------------------------------------ */
/* ES kernel: src/kernel.rs:21-89                                       */
/* ------------------------------------------------------------------ */

/* EsKernel::for_tolerance, src/kernel.rs:21-57.  Returns 0 or -1. */
int or_es_params(double eps, double sigma, int *width, double *beta)
{
    if (!isfinite(eps) || eps <= 0.0 || eps >= 1.0) return -1;
    if (!isfinite(sigma) || sigma <= 1.0) return -1;
    double wf;
    if (sigma == 2.0)
        wf = ceil(log10(10.0 / eps));
    else
        wf = log(1.0 / eps) / (M_PI * sqrt(1.0 - 1.0 / sigma));
    long w = (long)ceil(wf);
    if (w < 2) w = 2;
    if (w > 16) return -1;
    double b;
    if (sigma == 2.0) {
        double r = (w == 2) ? 2.20 : (w == 3) ? 2.26 : (w == 4) ? 2.38 : 2.30;
        b = r * (double)w;
    } else {
        b = 0.97 * M_PI * (double)w * (1.0 - 1.0 / (2.0 * sigma));
    }
    *width = (int)w;
    *beta = b;
    return 0;
}

/* EsKernel::evaluate, src/kernel.rs:81-89. */
double or_es_eval(int width, double beta, double distance)
{
    double a = fabs(distance);
    if (!isfinite(a) || a >= 0.5 * width) return 0.0;
    double s = 2.0 * a / (double)width;
    return exp(beta * (sqrt(1.0 - s * s) - 1.0));
}

/* gauss_legendre / legendre_pair, src/kernel.rs:367-407. */
static void legendre_pair(int order, double x, double *p, double *prev)
{
    double previous = 1.0;
    if (order == 0) { *p = previous; *prev = 0.0; return; }
    double current = x;
    for (int degree = 2; degree <= order; ++degree) {
        double next = ((double)(2 * degree - 1) * x * current - (double)(degree - 1) * previous) /
                      (double)degree;
        previous = current;
        current = next;
    }
    *p = current;
    *prev = previous;
}

void or_gauss_legendre(int order, double *nodes, double *weights)
{
    int pairs = (order + 1) / 2;
    for (int i = 0; i < pairs; ++i) {
        double root = cos(M_PI * ((double)i + 0.75) / ((double)order + 0.5));
        double p, q;
        for (;;) {
            legendre_pair(order, root, &p, &q);
            double der = (double)order * (root * p - q) / (root * root - 1.0);
            double next = root - p / der;
            if (fabs(next - root) <= 4.0 * 2.220446049250313e-16) { root = next; break; }
            root = next;
        }
        legendre_pair(order, root, &p, &q);
        double der = (double)order * (root * p - q) / (root * root - 1.0);
        double w = 2.0 / ((1.0 - root * root) * der * der);
        nodes[i] = -root;
        nodes[order - 1 - i] = root;
        weights[i] = w;
        weights[order - 1 - i] = w;
    }
}

/* centered_fourier_coefficients_using, src/kernel.rs:153-198: phi_hat[k] for
 * k = 0..n/2 including the (-1)^k phase. */
int or_phihat(int width, double beta, long n, double *out)
{
    if (n <= 0 || (n % 2) != 0) return -1;
    double hw = 0.5 * width;
    int q = (int)floor(2.0 + 3.0 * hw);
    double nodes[128], weights[128];
    or_gauss_legendre(2 * q, nodes, weights);
    double qn[64], wk[64];
    for (int i = 0; i < q; ++i) {
        qn[i] = hw * nodes[i];
        wk[i] = hw * weights[i] * or_es_eval(width, beta, qn[i]);
    }
    for (long k = 0; k <= n / 2; ++k) {
        double f = 2.0 * M_PI * (double)k / (double)n;
        double v = 0.0;
        for (int i = 0; i < q; ++i) v += 2.0 * wk[i] * cos(f * qn[i]);
        if (k % 2) v = -v;
        out[k] = v;
    }
    return 0;
}

/* es_kernel_fourier_transform, src/type3.rs:688-706. */
double or_es_ft(int width, double beta, double freq)
{
    double hw = 0.5 * width;
    int pn = width + 2;
    double nodes[64], weights[64];
    or_gauss_legendre(2 * pn, nodes, weights);
    double s = 0.0;
    for (int i = pn; i < 2 * pn; ++i) {
        double d = hw * nodes[i];
        s += 2.0 * hw * weights[i] * or_es_eval(width, beta, d) * cos(freq * d);
    }
    return s;
}

/* polynomial_fit_chebyshev + horner_table, src/kernel.rs:93-130, 242-274.
 * coef is coefficient-major [c * width + panel]; returns the row count. */
int or_horner_table(int width, double beta, double eps, double *coef)
{
    int fitted = width + 3;
    if (fitted > 19) fitted = 19;
    double table[19 * 16];
    for (int panel = 0; panel < width; ++panel) {
        double shift = (double)(2 * panel + 1) - (double)width;
        int cnt = fitted;
        double nodes[19], div[19];
        for (int i = 0; i < cnt; ++i) {
            nodes[i] = cos(((double)(2 * i + 1) * M_PI) / (double)(2 * cnt));
            div[i] = or_es_eval(width, beta, (nodes[i] + shift) * 0.5);
        }
        for (int order = 1; order < cnt; ++order)
            for (int i = cnt - 1; i >= order; --i)
                div[i] = (div[i] - div[i - 1]) / (nodes[i] - nodes[i - order]);
        double mono[19] = {0};
        double basis[20] = {1.0};
        int blen = 1;
        mono[0] = div[0];
        for (int order = 1; order < cnt; ++order) {
            double node = nodes[order - 1];
            double nb[20] = {0};
            for (int p = 0; p < blen; ++p) {
                nb[p] -= node * basis[p];
                nb[p + 1] += basis[p];
            }
            blen += 1;
            memcpy(basis, nb, sizeof(double) * blen);
            for (int p = 0; p < blen; ++p) mono[p] += div[order] * basis[p];
        }
        for (int c = 0; c < cnt; ++c) table[c * width + panel] = mono[cnt - 1 - c];
    }
    int minimum = width > 8 ? width - 4 : 4;
    double thr = 0.05 * eps;
    int lead = 0;
    while (fitted - lead > minimum) {
        int all = 1;
        for (int p = 0; p < width; ++p)
            if (!(fabs(table[lead * width + p]) < thr)) { all = 0; break; }
        if (!all) break;
        ++lead;
    }
    int rows = fitted - lead;
    memcpy(coef, table + lead * width, sizeof(double) * rows * width);
    return rows;
}

/* EsHornerTable::evaluate, src/kernel.rs:222-237. */
double or_horner_eval(int width, int rows, const double *coef, double distance)
{
    double hw = 0.5 * width;
    if (!isfinite(distance) || fabs(distance) >= hw) return 0.0;
    long panel = (long)ceil(distance + hw) - 1;
    if (panel < 0) panel = 0;
    if (panel > width - 1) panel = width - 1;
    double local = 2.0 * (distance - (double)panel) + (double)(width - 1);
    double v = 0.0;
    for (int c = 0; c < rows; ++c) v = v * local + coef[c * width + panel];
    return v;
}

/* ------------------------------------------------------------------ */
/* Fine-grid size: src/kernel.rs:278-333                                */
/* ------------------------------------------------------------------ */

/* Directly factorable lengths of the FFT library: every prime factor in
 * {2,3,5,7,11,13}.  The reference defers to wgpu-fft's
 * factor_supported_length, which is not vendored; its own tests fix
 * 34 -> 36, 14 -> 14 and 22 -> 22 (tests/cpu_foundations.rs:503-544). */
int or_fft_length_supported(long n)
{
    if (n <= 0) return 0;
    static const int primes[] = {2, 3, 5, 7, 11, 13};
    for (int i = 0; i < 6; ++i)
        while (n % primes[i] == 0) n /= primes[i];
    return n == 1;
}

long or_fine_size(long modes, double sigma, int width)
{
    if (modes <= 0) return -1;
    long scaled = (sigma == 2.0) ? modes * 2 : (long)ceil(sigma * (double)modes);
    long minimum = scaled > 2L * width ? scaled : 2L * width;
    long c = (minimum % 2 == 0) ? minimum : minimum + 1;
    while (c <= 4294967295L) {
        if (or_fft_length_supported(c)) return c;
        c += 2;
    }
    return -1;
}

/* next_even_235, src/type3.rs:763-797 (smallest even 2^a 3^b 5^c >= min). */
long or_next_even_235(long minimum)
{
    if (minimum < 2) minimum = 2;
    long best = -1;
    for (long p2 = 2; p2 <= (1L << 40); p2 *= 2)
        for (long p3 = p2; p3 <= (1L << 40); p3 *= 3)
            for (long c = p3; c <= (1L << 40); c *= 5)
                if (c >= minimum) {
                    if (best < 0 || c < best) best = c;
                    break;
                }
    return best;
}

/* ------------------------------------------------------------------ */
/* Spread layout and bin geometry: src/gpu_nd_spread.rs:194-459         */
/* ------------------------------------------------------------------ */

typedef struct {
    long side[OR_MAX_DIMS];
    long bins[OR_MAX_DIMS];
    int full[OR_MAX_DIMS];
    long reach_bins[OR_MAX_DIMS];
    double mean_reach[OR_MAX_DIMS];
} or_geometry;

typedef struct {
    int dims;
    int width;
    long fine[OR_MAX_DIMS];
    long run, groups0, workgroup_size, row_batch;
    long block[OR_MAX_DIMS];
    long blocks[OR_MAX_DIMS];
    int ngeom;
    or_geometry geom[OR_MAX_GEOMETRIES];
} or_layout;

static long pow2_at_most(long v)
{
    if (v <= 0) return 1;
    long p = 1;
    while (p * 2 <= v) p *= 2;
    return p;
}

static long next_pow2(long v)
{
    long p = 1;
    while (p < v) p *= 2;
    return p;
}

/* target_workgroup_size / max_run, src/gpu_nd_spread.rs:562-584.
 * precision: 0 = F32, 1 = F64, 2 = Df64. */
static long target_wg(int d, int prec)
{
    if (d == 1) return 64;
    if (d == 2 && prec == 0) return 32;
    if ((d == 2 || d == 3) && prec == 1) return 64;
    if (prec == 0) return 64;
    return 256;
}

static long max_run_of(int d, int prec)
{
    if (d == 1 || (d == 2 && prec == 2)) return 2;
    if (d == 2) return 4;
    if (d == 3 || prec == 1 || prec == 2) return 8;
    return 16;
}

static double mean_reach_bins(const or_layout *L, const long *side, int axis)
{
    long reach = L->block[axis] + L->width - 1;
    double v = (double)(reach - 1) / (double)side[axis] + 1.0;
    double cap = (double)((L->fine[axis] + side[axis] - 1) / side[axis]);
    return v < cap ? v : cap;
}

static void reach_geometry(const or_layout *L, const long *side, or_geometry *g)
{
    for (int a = 0; a < L->dims; ++a) {
        long count = (L->fine[a] + side[a] - 1) / side[a];
        long reach = L->block[a] + L->width - 1;
        int whole = reach + side[a] > L->fine[a];
        g->side[a] = side[a];
        g->bins[a] = count;
        g->full[a] = whole;
        long rb = (reach + side[a] - 1) / side[a] + 2;
        g->reach_bins[a] = whole ? count : (count < rb ? count : rb);
        g->mean_reach[a] = mean_reach_bins(L, side, a);
    }
}

/* NdSpreadLayout::for_grid, src/gpu_nd_spread.rs:194-407, for the WebGPU
 * default limits (256 invocations per workgroup). */
int or_layout_build(int d, const long *fine, int width, int prec, or_layout *L)
{
    memset(L, 0, sizeof(*L));
    L->dims = d;
    L->width = width;
    for (int a = 0; a < d; ++a) L->fine[a] = fine[a];
    long target = target_wg(d, prec);
    if (target > 256) target = 256;
#define EDGE_LIMIT(a) ((fine[a] + 1 - width) > 1 ? (fine[a] + 1 - width) : 1)
    long mr = max_run_of(d, prec);
    long el0 = EDGE_LIMIT(0);
    long run = pow2_at_most(mr < el0 ? mr : el0);
    long block[OR_MAX_DIMS];
    for (int a = 0; a < OR_MAX_DIMS; ++a) block[a] = 1;
    long inv = 1;
    for (;;) {
        int best = -1;
        for (int a = 1; a < d; ++a) {
            long lim = EDGE_LIMIT(a) < 16 ? EDGE_LIMIT(a) : 16;
            if (block[a] * 2 <= lim && inv * 2 <= target) {
                /* min_by_key (block, usize::MAX - fine): the first minimum */
                if (best < 0 || block[a] < block[best] ||
                    (block[a] == block[best] && fine[a] > fine[best]))
                    best = a;
            }
        }
        if (best < 0) break;
        block[best] *= 2;
        inv *= 2;
    }
    long g0 = 1;
    while (inv * 2 <= target && run * g0 * 2 <= el0) {
        g0 *= 2;
        inv *= 2;
    }
    block[0] = run * g0;
    long wg = g0;
    for (int a = 1; a < d; ++a) wg *= block[a];
    L->run = run;
    L->groups0 = g0;
    L->workgroup_size = wg;
    for (int a = 0; a < d; ++a) L->block[a] = block[a];

    long side[OR_MAX_DIMS];
    for (int a = 0; a < OR_MAX_DIMS; ++a) side[a] = 1;
    if (!(d == 2 && prec == 0)) side[0] = 2;
    for (;;) {
        double rows = 1.0;
        for (int a = 1; a < d; ++a) rows *= mean_reach_bins(L, side, a);
        double bc = 1.0;
        long bin_count = 1;
        int overflow = 0;
        for (int a = 0; a < d; ++a) {
            long c = (fine[a] + side[a] - 1) / side[a];
            bc *= (double)c;
            bin_count *= c;
        }
        if (bc > 9.0e18) overflow = 1;
        if (!overflow && rows <= 1024.0 && bin_count <= (1L << 23)) break;
        int widest = -1;
        double wv = 0.0;
        for (int a = 1; a < d; ++a) {
            if (side[a] < fine[a]) {
                double v = mean_reach_bins(L, side, a);
                /* max_by on total_cmp: the last maximum wins */
                if (widest < 0 || v >= wv) { widest = a; wv = v; }
            }
        }
        if (widest < 0) {
            if ((overflow || bin_count > (1L << 23)) && side[0] < fine[0]) widest = 0;
            else break;
        }
        side[widest] *= 2;
    }
    or_geometry finest;
    reach_geometry(L, side, &finest);
    for (;;) {
        double rows = 1.0;
        long rowsl = 1;
        int ovf = 0;
        for (int a = 1; a < d; ++a) {
            rows *= (double)finest.reach_bins[a];
            if (rowsl > (1L << 40)) ovf = 1;
            rowsl *= finest.reach_bins[a];
        }
        if (!ovf && rowsl <= (1L << 20)) break;
        int best = -1;
        for (int a = 1; a < d; ++a)
            if (finest.reach_bins[a] > 1 &&
                (best < 0 || finest.reach_bins[a] >= finest.reach_bins[best]))
                best = a;
        if (best < 0) break;
        side[best] *= 2;
        reach_geometry(L, side, &finest);
    }
    L->geom[0] = finest;
    L->ngeom = 1;
    for (int others = 1; others <= 3; ++others) {
        for (int first = 0; first <= others; ++first) {
            long s[OR_MAX_DIMS];
            for (int a = 0; a < d; ++a) {
                int dbl = (a == 0) ? first : others;
                long v = side[a] << dbl;
                long cap = next_pow2(fine[a]);
                s[a] = v < cap ? v : cap;
            }
            int dup = 0;
            for (int g = 0; g < L->ngeom; ++g) {
                int same = 1;
                for (int a = 0; a < d; ++a)
                    if (L->geom[g].side[a] != s[a]) same = 0;
                if (same) dup = 1;
            }
            if (!dup) reach_geometry(L, s, &L->geom[L->ngeom++]);
        }
    }
    long rows0 = 1;
    for (int a = 1; a < d; ++a) rows0 *= L->geom[0].reach_bins[a];
    if (rows0 < 1) rows0 = 1;
    long rb = 128;
    if (wg < rb) rb = wg;
    long np = next_pow2(rows0);
    if (np < rb) rb = np;
    if (rb < 1) rb = 1;
    L->row_batch = rb;
    for (int a = 0; a < d; ++a) L->blocks[a] = (fine[a] + block[a] - 1) / block[a];
#undef EDGE_LIMIT
    return 0;
}

/* NdSpreadLayout::choose_geometry, src/gpu_nd_spread.rs:424-459. */
int or_choose_geometry(const or_layout *L, long point_count)
{
    double cells = 1.0;
    double blocks = 1.0;
    for (int a = 0; a < L->dims; ++a) {
        cells *= (double)L->fine[a];
        blocks *= (double)L->blocks[a];
    }
    double density = (double)point_count / cells;
    double dd = (double)L->dims;
    double record_cost = (double)L->workgroup_size * (dd + 1.0) * (dd + 1.0) / 1024.0;
    double staged_v[OR_MAX_GEOMETRIES], cost_v[OR_MAX_GEOMETRIES];
    double least = INFINITY;
    for (int g = 0; g < L->ngeom; ++g) {
        const or_geometry *G = &L->geom[g];
        double rows = 1.0;
        for (int a = 1; a < L->dims; ++a) rows *= G->mean_reach[a];
        double batches = ceil(rows / (double)L->row_batch);
        double staged = 1.0;
        double bins = 1.0;
        for (int a = 0; a < L->dims; ++a) {
            double v = G->mean_reach[a] * (double)G->side[a];
            double f = (double)L->fine[a];
            staged *= v < f ? v : f;
            bins *= (double)G->bins[a];
        }
        double cost = blocks * (100.0 * batches + 1.0 * rows + record_cost * density * staged) +
                      0.1 * bins;
        staged_v[g] = staged;
        cost_v[g] = cost;
        if (cost < least) least = cost;
    }
    int best = -1;
    for (int g = 0; g < L->ngeom; ++g) {
        if (cost_v[g] <= least * 1.1) {
            if (best < 0 || staged_v[g] < staged_v[best]) best = g;
        }
    }
    return best < 0 ? 0 : best;
}

/* Flattened accessors for ctypes. */
int or_layout_query(int d, const long *fine, int width, int prec, long point_count,
                    long *out /* >= 128 longs */)
{
    or_layout L;
    if (d < 1 || d > OR_MAX_DIMS) return -1;
    if (or_layout_build(d, fine, width, prec, &L)) return -1;
    int g = or_choose_geometry(&L, point_count);
    for (int i = 0; i < 128; ++i) out[i] = 0;
    out[0] = L.run;
    out[1] = L.workgroup_size;
    out[2] = L.row_batch;
    out[3] = L.ngeom;
    out[4] = g;
    for (int a = 0; a < OR_MAX_DIMS; ++a) {
        out[8 + a] = a < d ? L.block[a] : 1;
        out[16 + a] = a < d ? L.blocks[a] : 1;
    }
    for (int i = 0; i < L.ngeom && i < 12; ++i)
        for (int a = 0; a < OR_MAX_DIMS; ++a) out[32 + 8 * i + a] = a < d ? L.geom[i].side[a] : 1;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Fold, support start and bin key: src/gpu_nd_wgsl.rs:198-310,          */
/* src/gpu_nd_bins.rs:474-554                                           */
/* ------------------------------------------------------------------ */

/* Bits of 4*pi as f32, the fold's reach (src/gpu_nd_wgsl.rs:219-224). */
static const uint32_t FOLD_REACH_BITS = 0x41490fdbu;

/* The scale n/(2 pi) split into an f32 head and tail, as the reference's
 * position constants (src/gpu_nd_wgsl.rs:226-228). */
static void scale_split(long n, float *hi, float *lo)
{
    double s = (double)n / 6.283185307179586;
    *hi = (float)s;
    *lo = (float)(s - (double)*hi);
}

/* The double-precision fold the HIP kernels compute: both partial products
 * are exact in f64, their sum and the origin are rounded once each. */
static double fold_f64(float x, float hi, float lo, double n)
{
    double t1 = (double)x * (double)hi;
    double t2 = (double)x * (double)lo;
    double pos = (t1 + t2) + 0.5 * n;
    if (pos < 0.0) pos += n;
    if (pos < 0.0) pos += n;
    if (pos >= n) pos -= n;
    if (pos >= n) pos -= n;
    return pos;
}

/* The same fold in exact arithmetic (binary128 holds every intermediate
 * exactly for the grid lengths used here). */
static __float128 fold_exact(float x, float hi, float lo, long n)
{
    __float128 pos = (__float128)x * (__float128)hi + (__float128)x * (__float128)lo +
                     (__float128)(n / 2);
    __float128 nq = (__float128)n;
    if (pos < 0) pos += nq;
    if (pos < 0) pos += nq;
    if (pos >= nq) pos -= nq;
    if (pos >= nq) pos -= nq;
    return pos;
}

static int in_reach_f32(float x)
{
    uint32_t b;
    memcpy(&b, &x, 4);
    return (b & 0x7fffffffu) <= FOLD_REACH_BITS;
}

/* Support start, wrapped start, bin key and prepared offset of every point
 * of an F32 point set (point-major coordinates).  Also counts the
 * "ambiguous" points: those whose f64 start differs from the exact one, and
 * those whose exact pos - w/2 lies within DF64_WINDOW * n of an integer,
 * where the reference's double-float fold (a few df64 operations, each
 * within about 2^-48 of the exact result relative to n) could round the
 * start either way.  A point set with no ambiguous point has the same
 * starts under the f64 fold, the exact fold and any fold within that
 * window of it. */
#define DF64_WINDOW 0x1p-44
long or_bin_keys_f32(int d, const long *fine, int width, const long *side, long M,
                     const float *coords, uint32_t *keys, int32_t *starts, double *offsets)
{
    float hi[OR_MAX_DIMS], lo[OR_MAX_DIMS];
    long bins[OR_MAX_DIMS];
    for (int a = 0; a < d; ++a) {
        scale_split(fine[a], &hi[a], &lo[a]);
        bins[a] = (fine[a] + side[a] - 1) / side[a];
    }
    double hw = 0.5 * (double)width;
    long ambiguous = 0;
    for (long i = 0; i < M; ++i) {
        int reach = 1;
        for (int a = 0; a < d; ++a) reach &= in_reach_f32(coords[i * d + a]);
        uint64_t key = 0, stride = 1;
        for (int a = 0; a < d; ++a) {
            long start = 0;
            double off = 1024.0;
            if (reach) {
                double pos = fold_f64(coords[i * d + a], hi[a], lo[a], (double)fine[a]);
                start = (long)ceil(pos - hw);
                off = (double)start - pos;
                __float128 pe = fold_exact(coords[i * d + a], hi[a], lo[a], fine[a]);
                __float128 t = pe - (__float128)hw;
                long se = (long)t; /* truncation, then up to the ceiling */
                if ((__float128)se < t) se += 1;
                /* distance of t to the nearest integer (se or se - 1) */
                __float128 up = (__float128)se - t, down = t - (__float128)(se - 1);
                __float128 dist = up < down ? up : down;
                __float128 window = (__float128)DF64_WINDOW * (__float128)fine[a];
                if (se != start || dist <= window) ++ambiguous;
            }
            long ws = start < 0 ? start + fine[a] : start;
            if (ws < 0) ws = 0;
            if (ws > fine[a] - 1) ws = fine[a] - 1;
            if (starts) starts[i * d + a] = (int32_t)start;
            if (offsets) offsets[i * d + a] = off;
            long sh = 0;
            while ((1L << sh) < side[a]) ++sh;
            key += (uint64_t)(ws >> sh) * stride;
            stride *= (uint64_t)bins[a];
        }
        keys[i] = (uint32_t)key;
    }
    return ambiguous;
}

/* The permutation contract of src/gpu_bin_sort.rs:1-7 and
 * src/gpu_nd_bins.rs:23-29: a stable counting sort of point indices by bin
 * key (ties in ascending original index), offsets = exclusive scan of the
 * bin counts followed by the total. */
void or_bin_sort(long M, const uint32_t *keys, long nbins, uint32_t *sorted, uint32_t *offsets)
{
    memset(offsets, 0, sizeof(uint32_t) * (nbins + 1));
    for (long i = 0; i < M; ++i) offsets[keys[i]]++;
    uint32_t run = 0;
    for (long b = 0; b <= nbins; ++b) {
        uint32_t c = offsets[b];
        offsets[b] = run;
        run += c;
    }
    uint32_t *cursor = malloc(sizeof(uint32_t) * (nbins + 1));
    memcpy(cursor, offsets, sizeof(uint32_t) * (nbins + 1));
    for (long i = 0; i < M; ++i) sorted[cursor[keys[i]]++] = (uint32_t)i;
    free(cursor);
}

/* ------------------------------------------------------------------ */
/* Direct FP64 sums: src/direct.rs:9-131, src/type3.rs:609-673          */
/* ------------------------------------------------------------------ */

/* mode_for_storage_index, src/direct.rs:120-131. */
static long mode_of(int modeord, long index, long length)
{
    if (modeord == 0) return index - length / 2;
    return index < (length + 1) / 2 ? index : index - length;
}

typedef struct {
    int kind; /* 1, 2 or 3 */
    int d;    /* 1..8 */
    long nmodes[8];
    int modeord;
    double sign;
    long M;
    const double *x;   /* point-major */
    const double *in;  /* interleaved complex */
    double *out;
    long nout;         /* outputs to compute (a selection for type 1/3) */
    const long *sel;   /* type 1: linear mode indices; type 2: point indices */
    const double *s;   /* type 3 targets, point-major */
    int nthreads, tid;
} direct_job;

static void *direct_worker(void *arg)
{
    direct_job *J = (direct_job *)arg;
    for (long o = J->tid; o < J->nout; o += J->nthreads) {
        double re = 0.0, im = 0.0;
        if (J->kind == 1) {
            long lin = J->sel ? J->sel[o] : o;
            long k[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            long rem = lin;
            for (int a = 0; a < J->d; ++a) {
                k[a] = mode_of(J->modeord, rem % J->nmodes[a], J->nmodes[a]);
                rem /= J->nmodes[a];
            }
            for (long j = 0; j < J->M; ++j) {
                double ph = 0.0;
                for (int a = 0; a < J->d; ++a) ph += (double)k[a] * J->x[j * J->d + a];
                double sn, cs;
                sincos(J->sign * ph, &sn, &cs);
                double cr = J->in[2 * j], ci = J->in[2 * j + 1];
                re += cr * cs - ci * sn;
                im += cr * sn + ci * cs;
            }
        } else if (J->kind == 2) {
            long j = J->sel ? J->sel[o] : o;
            long total = 1;
            for (int a = 0; a < J->d; ++a) total *= J->nmodes[a];
            for (long lin = 0; lin < total; ++lin) {
                long rem = lin;
                double ph = 0.0;
                for (int a = 0; a < J->d; ++a) {
                    long k = mode_of(J->modeord, rem % J->nmodes[a], J->nmodes[a]);
                    rem /= J->nmodes[a];
                    ph += (double)k * J->x[j * J->d + a];
                }
                double sn, cs;
                sincos(J->sign * ph, &sn, &cs);
                double cr = J->in[2 * lin], ci = J->in[2 * lin + 1];
                re += cr * cs - ci * sn;
                im += cr * sn + ci * cs;
            }
        } else {
            long t = J->sel ? J->sel[o] : o;
            for (long j = 0; j < J->M; ++j) {
                double ph = 0.0;
                for (int a = 0; a < J->d; ++a) ph += J->x[j * J->d + a] * J->s[t * J->d + a];
                double sn, cs;
                sincos(J->sign * ph, &sn, &cs);
                double cr = J->in[2 * j], ci = J->in[2 * j + 1];
                re += cr * cs - ci * sn;
                im += cr * sn + ci * cs;
            }
        }
        J->out[2 * o] = re;
        J->out[2 * o + 1] = im;
    }
    return NULL;
}

/* Type-1 / type-2 / type-3 direct sums for one transform in 1 to 8
 * dimensions (src/direct.rs:9-131; src/config.rs:9: eight at most).  `sel` (may be
 * NULL) selects the outputs: linear mode indices (type 1), point indices
 * (type 2) or target indices (type 3).  isign >= 0 is the positive sign
 * (src/config.rs:25-31). */
int or_direct(int kind, int d, const long *nmodes, int modeord, int isign, long M,
              const double *x, const double *in, long nout, const long *sel, const double *s,
              double *out, int nthreads)
{
    if (d < 1 || d > 8) return -1;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    direct_job jobs[256];
    for (int t = 0; t < nthreads; ++t) {
        direct_job *J = &jobs[t];
        J->kind = kind;
        J->d = d;
        for (int a = 0; a < 8; ++a) J->nmodes[a] = (nmodes && a < d) ? nmodes[a] : 1;
        J->modeord = modeord;
        J->sign = isign >= 0 ? 1.0 : -1.0;
        J->M = M;
        J->x = x;
        J->in = in;
        J->out = out;
        J->nout = nout;
        J->sel = sel;
        J->s = s;
        J->nthreads = nthreads;
        J->tid = t;
        pthread_create(&th[t], NULL, direct_worker, J);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    return 0;
}

/* Type-3 axis metadata: src/type3.rs:64-86 (centre snap), 708-761. */
int or_type3_axis(double src_lo, double src_hi, double tgt_lo, double tgt_hi, double sigma,
                  int width, double *out /* X, C, S, D, nf, h, gamma */)
{
    double X = (src_hi - src_lo) * 0.5, C = (src_hi + src_lo) * 0.5;
    if (fabs(C) < 0.1 * X) { X += fabs(C); C = 0.0; }
    double S = (tgt_hi - tgt_lo) * 0.5, D = (tgt_hi + tgt_lo) * 0.5;
    if (fabs(D) < 0.1 * S) { S += fabs(D); D = 0.0; }
    double x = X, s = S;
    if (x == 0.0) {
        if (s == 0.0) { x = 1.0; s = 1.0; }
        else x = 1.0 / s;
    } else if (s < 1.0 / x) {
        s = 1.0 / x;
    }
    double requested = 2.0 * sigma * s * x / M_PI + (double)(width + 1);
    long minimum = (long)requested;
    if (minimum < 2L * width) minimum = 2L * width;
    long nf = or_next_even_235(minimum);
    out[0] = X;
    out[1] = C;
    out[2] = S;
    out[3] = D;
    out[4] = (double)nf;
    out[5] = 2.0 * M_PI / (double)nf;
    out[6] = (double)nf / (2.0 * sigma * s);
    return 0;
}
