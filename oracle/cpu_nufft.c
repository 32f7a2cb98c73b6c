/*
 * CPU port of the reference's CpuNufftPlan -- the CPU baseline of bench.py
 * (TEST INFRASTRUCTURE: only tests/ and bench.py's cpu_baseline leg load it).
 *
 * Follows src/cpu.rs and the files of src/cpu/ for F32 plans:
 *  - support start and the local panel coordinate t of every axis exactly as
 *    FineGrid::support (src/cpu/grid.rs:47-79); weights from the piecewise
 *    Horner table evaluated in f32 (src/cpu/grid.rs:85-129);
 *  - points sorted into bins keyed by (slab of the last-axis start, whether
 *    the support crosses into the next slab, tile of the other axes with
 *    32-cell tiles on axis 0 and 16 on the others), input order inside a bin
 *    (src/cpu/points.rs:12-14, 21-84, 107-236);
 *  - slab-parallel spreading: each slab task zeroes its cells, then adds the
 *    previous slab's crossing points and its own points in bin order, so the
 *    sums do not depend on the thread count (src/cpu/spread.rs:107-163);
 *  - slabs of min(16, max(1, n_last/64)) planes, at least w-1 thick
 *    (src/cpu.rs:31-36, 873-889);
 *  - an unnormalized C2C FFT of the fine grid (+i for isign >= 0), then
 *    deconvolution and truncation (src/cpu/modes.rs:68-107);
 *  - type 2: predeconvolution, FFT, per-point interpolation.
 * The FFT is this file's own (radix-2 plus a generic mixed-radix fallback):
 * the reference's CPU FFT lives in the un-vendored wgpu-fft.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

int or_es_params(double eps, double sigma, int *width, double *beta);
int or_phihat(int width, double beta, long n, double *out);
int or_horner_table(int width, double beta, double eps, double *coef);
long or_fine_size(long modes, double sigma, int width);

typedef struct {
    int d, w, rows, isign, threads;
    long nm[3], n[3], cells, modes;
    double scale[3], hw;
    float horner[19 * 16];
    float *amp[3]; /* 1/phi_hat per |k| */
} cplan;

static int cplan_init(cplan *p, int d, const long *nm, double eps, int isign, int threads)
{
    memset(p, 0, sizeof(*p));
    double beta;
    if (or_es_params(eps, 2.0, &p->w, &beta)) return -1;
    double coef[19 * 16];
    p->rows = or_horner_table(p->w, beta, eps, coef);
    for (int i = 0; i < p->rows * p->w; ++i) p->horner[i] = (float)coef[i];
    p->d = d;
    p->isign = isign;
    p->threads = threads < 1 ? 1 : threads;
    p->hw = 0.5 * p->w;
    p->cells = 1;
    p->modes = 1;
    for (int a = 0; a < 3; ++a) {
        p->nm[a] = a < d ? nm[a] : 1;
        p->n[a] = a < d ? or_fine_size(nm[a], 2.0, p->w) : 1;
        p->scale[a] = (double)p->n[a] / 6.283185307179586;
        p->cells *= p->n[a];
        p->modes *= p->nm[a];
        if (a < d) {
            double *ph = malloc(sizeof(double) * (p->n[a] / 2 + 1));
            or_phihat(p->w, beta, p->n[a], ph);
            p->amp[a] = malloc(sizeof(float) * (p->n[a] / 2 + 1));
            for (long k = 0; k <= p->n[a] / 2; ++k) p->amp[a][k] = (float)(1.0 / ph[k]);
            free(ph);
        }
    }
    return 0;
}

static void cplan_free(cplan *p)
{
    for (int a = 0; a < 3; ++a) free(p->amp[a]);
}

/* FineGrid::support + Weights::evaluate in f32. */
static inline long support(const cplan *p, int axis, double x, float *wts)
{
    long n = p->n[axis];
    double nf = (double)n;
    double pos = x * p->scale[axis] + (double)(n / 2);
    if (pos < 0.0) pos += nf;
    if (pos < 0.0) pos += nf;
    if (pos >= nf) pos -= nf;
    if (pos >= nf) pos -= nf;
    double start = ceil(pos - p->hw);
    float t = (float)(2.0 * (start - pos) + (double)(p->w - 1));
    long s = (long)start;
    if (s < 0) s += n;
    for (int j = 0; j < p->w; ++j) {
        float v = 0.0f;
        for (int r = 0; r < p->rows; ++r) v = v * t + p->horner[r * p->w + j];
        wts[j] = v;
    }
    return s;
}

/* ---------------- FFT (float complex, in place, strided lines) ---------- */
typedef struct { float re, im; } cf;

static void fft_line(cf *a, long n, int sign, cf *tmp)
{
    if ((n & (n - 1)) == 0) {
        for (long i = 1, j = 0; i < n; ++i) {
            long bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) { cf t = a[i]; a[i] = a[j]; a[j] = t; }
        }
        for (long len = 2; len <= n; len <<= 1) {
            double ang = 2.0 * M_PI / (double)len * sign;
            for (long k = 0; k < len / 2; ++k) {
                float wr = (float)cos(ang * k), wi = (float)sin(ang * k);
                for (long i = k; i < n; i += len) {
                    cf u = a[i], v = a[i + len / 2];
                    cf t = {v.re * wr - v.im * wi, v.re * wi + v.im * wr};
                    a[i].re = u.re + t.re;
                    a[i].im = u.im + t.im;
                    a[i + len / 2].re = u.re - t.re;
                    a[i + len / 2].im = u.im - t.im;
                }
            }
        }
        return;
    }
    /* generic fallback: direct DFT in double (small non-power-of-two grids) */
    for (long k = 0; k < n; ++k) {
        double sr = 0, si = 0;
        for (long j = 0; j < n; ++j) {
            double ang = 2.0 * M_PI * (double)((k * j) % n) / (double)n * sign;
            sr += a[j].re * cos(ang) - a[j].im * sin(ang);
            si += a[j].re * sin(ang) + a[j].im * cos(ang);
        }
        tmp[k].re = (float)sr;
        tmp[k].im = (float)si;
    }
    memcpy(a, tmp, sizeof(cf) * n);
}

typedef struct {
    cplan *p;
    cf *grid;
    int axis, sign;
    long next, total;
    pthread_mutex_t *mu;
} fft_job;

static void *fft_worker(void *arg)
{
    fft_job *J = (fft_job *)arg;
    cplan *p = J->p;
    long n = p->n[J->axis];
    long stride = 1;
    for (int a = 0; a < J->axis; ++a) stride *= p->n[a];
    cf *line = malloc(sizeof(cf) * n), *tmp = malloc(sizeof(cf) * n);
    for (;;) {
        pthread_mutex_lock(J->mu);
        long first = J->next;
        J->next += 64;
        pthread_mutex_unlock(J->mu);
        if (first >= J->total) break;
        long last = first + 64 < J->total ? first + 64 : J->total;
        for (long l = first; l < last; ++l) {
            long lo = l % stride, hi = l / stride;
            cf *base = J->grid + hi * stride * n + lo;
            for (long i = 0; i < n; ++i) line[i] = base[i * stride];
            fft_line(line, n, J->sign, tmp);
            for (long i = 0; i < n; ++i) base[i * stride] = line[i];
        }
    }
    free(line);
    free(tmp);
    return NULL;
}

static void fft_grid(cplan *p, cf *grid)
{
    int sign = p->isign >= 0 ? 1 : -1;
    pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
    for (int axis = 0; axis < p->d; ++axis) {
        fft_job J = {p, grid, axis, sign, 0, p->cells / p->n[axis], &mu};
        pthread_t th[256];
        int nt = p->threads > 256 ? 256 : p->threads;
        for (int t = 0; t < nt; ++t) pthread_create(&th[t], NULL, fft_worker, &J);
        for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
    }
}

/* ---------------- sort into slab / crossing / tile bins ------------------ */
typedef struct {
    long count;                 /* slabs */
    long *bounds;               /* count + 1 */
    int *of_cell;               /* slab of each last-axis cell */
    uint32_t *order;            /* points in bin order */
    long *slab_start, *crossing_start;
} sorted_points;

static void sort_points(cplan *p, long M, const float *x, sorted_points *S)
{
    int d = p->d;
    long last_n = p->n[d - 1];
    long thick = d == 1 ? 4096 : (last_n / 64 < 1 ? 1 : (last_n / 64 > 16 ? 16 : last_n / 64));
    if (thick < p->w - 1) thick = p->w - 1;
    if (thick < 1) thick = 1;
    long count = last_n / thick;
    if (count < 1) count = 1;
    S->count = count;
    S->bounds = malloc(sizeof(long) * (count + 1));
    for (long s = 0; s <= count; ++s) S->bounds[s] = s * last_n / count;
    S->of_cell = malloc(sizeof(int) * last_n);
    for (long s = 0; s < count; ++s)
        for (long c = S->bounds[s]; c < S->bounds[s + 1]; ++c) S->of_cell[c] = (int)s;
    long tiles[3] = {1, 1, 1}, tpg = 1;
    for (int a = 0; a < d - 1; ++a) {
        long edge = a == 0 ? 32 : 16;
        tiles[a] = (p->n[a] + edge - 1) / edge;
        tpg *= tiles[a];
    }
    long nb = count * 2 * tpg;
    uint32_t *keys = malloc(sizeof(uint32_t) * (M > 0 ? M : 1));
    long *hist = calloc(nb + 1, sizeof(long));
    float wts[16];
    for (long j = 0; j < M; ++j) {
        long tile = 0;
        for (int a = d - 2; a >= 0; --a) {
            long st = support(p, a, x[j * d + a], wts);
            tile = tile * tiles[a] + st / (a == 0 ? 32 : 16);
        }
        long sl = support(p, d - 1, x[j * d + d - 1], wts);
        long slab = S->of_cell[sl];
        int crosses = count > 1 && sl + p->w > S->bounds[slab + 1];
        keys[j] = (uint32_t)((2 * slab + crosses) * tpg + tile);
        hist[keys[j]]++;
    }
    long run = 0;
    for (long b = 0; b <= nb; ++b) {
        long c = hist[b];
        hist[b] = run;
        run += c;
    }
    S->slab_start = malloc(sizeof(long) * (count + 1));
    S->crossing_start = malloc(sizeof(long) * count);
    for (long s = 0; s <= count; ++s) S->slab_start[s] = hist[2 * s * tpg];
    for (long s = 0; s < count; ++s) S->crossing_start[s] = hist[(2 * s + 1) * tpg];
    S->order = malloc(sizeof(uint32_t) * (M > 0 ? M : 1));
    for (long j = 0; j < M; ++j) S->order[hist[keys[j]]++] = (uint32_t)j;
    free(keys);
    free(hist);
}

static void sorted_free(sorted_points *S)
{
    free(S->bounds);
    free(S->of_cell);
    free(S->order);
    free(S->slab_start);
    free(S->crossing_start);
}

/* ---------------- spread ------------------------------------------------- */
typedef struct {
    cplan *p;
    const float *x, *c;
    sorted_points *S;
    cf *grid;
    long next;
    pthread_mutex_t *mu;
} spread_job;

static void spread_point(cplan *p, const float *xj, const float *cj, long low, long high, cf *grid)
{
    int d = p->d, w = p->w;
    float wt[3][16];
    long st[3] = {0, 0, 0};
    for (int a = 0; a < d; ++a) st[a] = support(p, a, xj[a], wt[a]);
    float cre = cj[0], cim = cj[1];
    float vre[16], vim[16];
    for (int i = 0; i < w; ++i) {
        vre[i] = wt[0][i] * cre;
        vim[i] = wt[0][i] * cim;
    }
    long n0 = p->n[0];
    if (d == 1) {
        for (int i = 0; i < w; ++i) {
            long cell = st[0] + i;
            if (cell >= n0) cell -= n0;
            if (cell >= low && cell < high) {
                grid[cell].re += vre[i];
                grid[cell].im += vim[i];
            }
        }
        return;
    }
    long nl = p->n[d - 1];
    long plane = p->cells / nl;
    for (int dz = 0; dz < w; ++dz) {
        long cz = st[d - 1] + dz;
        if (cz >= nl) cz -= nl;
        if (cz < low || cz >= high) continue;
        float wz = wt[d - 1][dz];
        int ny = d == 3 ? w : 1;
        for (int dy = 0; dy < ny; ++dy) {
            long row = cz * plane;
            float s = wz;
            if (d == 3) {
                long cy = st[1] + dy;
                if (cy >= p->n[1]) cy -= p->n[1];
                row += cy * n0;
                s *= wt[1][dy];
            }
            for (int i = 0; i < w; ++i) {
                long xx = st[0] + i;
                if (xx >= n0) xx -= n0;
                grid[row + xx].re += vre[i] * s;
                grid[row + xx].im += vim[i] * s;
            }
        }
    }
}

static void *spread_worker(void *arg)
{
    spread_job *J = (spread_job *)arg;
    cplan *p = J->p;
    sorted_points *S = J->S;
    int d = p->d;
    long plane = p->cells / p->n[d - 1];
    for (;;) {
        pthread_mutex_lock(J->mu);
        long slab = J->next++;
        pthread_mutex_unlock(J->mu);
        if (slab >= S->count) break;
        long low = S->bounds[slab], high = S->bounds[slab + 1];
        memset(J->grid + low * plane, 0, sizeof(cf) * (high - low) * plane);
        long prev = (slab + S->count - 1) % S->count;
        if (prev != slab)
            for (long r = S->crossing_start[prev]; r < S->slab_start[prev + 1]; ++r) {
                uint32_t j = S->order[r];
                spread_point(p, J->x + (long)j * d, J->c + 2L * j, low, high, J->grid);
            }
        for (long r = S->slab_start[slab]; r < S->slab_start[slab + 1]; ++r) {
            uint32_t j = S->order[r];
            spread_point(p, J->x + (long)j * d, J->c + 2L * j, low, high, J->grid);
        }
    }
    return NULL;
}

static void run_threads(int n, void *(*fn)(void *), void *arg)
{
    pthread_t th[256];
    if (n > 256) n = 256;
    for (int t = 0; t < n; ++t) pthread_create(&th[t], NULL, fn, arg);
    for (int t = 0; t < n; ++t) pthread_join(th[t], NULL);
}

static long mode_fine_index(long k, long N, long n, long *m_out)
{
    long m = k - N / 2; /* centered order */
    *m_out = m < 0 ? -m : m;
    return m < 0 ? m + n : m;
}

/* Wall time of the last type-1 call's point sort (prepare_points of the
 * reference's CpuNufftPlan, src/cpu.rs); the rest of that call is its
 * execute_prepared. */
static double g_last_prepare_s = 0.0;

double cpu_last_prepare_seconds(void) { return g_last_prepare_s; }

static double now_seconds(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* One type-1 execute: sort + spread + FFT + deconvolution. */
int cpu_type1(int d, const long *nm, double eps, int isign, long M, const float *x,
              const float *c, float *out, int threads)
{
    cplan p;
    if (cplan_init(&p, d, nm, eps, isign, threads)) return -1;
    sorted_points S;
    const double t0 = now_seconds();
    sort_points(&p, M, x, &S);
    g_last_prepare_s = now_seconds() - t0;
    cf *grid = malloc(sizeof(cf) * p.cells);
    pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
    spread_job J = {&p, x, c, &S, grid, 0, &mu};
    run_threads(p.threads, spread_worker, &J);
    fft_grid(&p, grid);
    cf *o = (cf *)out;
    for (long k = 0; k < p.modes; ++k) {
        long rest = k, lin = 0, stride = 1;
        float a = 1.0f;
        for (int ax = 0; ax < d; ++ax) {
            long m, f = mode_fine_index(rest % p.nm[ax], p.nm[ax], p.n[ax], &m);
            rest /= p.nm[ax];
            lin += f * stride;
            stride *= p.n[ax];
            a *= p.amp[ax][m];
        }
        o[k].re = grid[lin].re * a;
        o[k].im = grid[lin].im * a;
    }
    free(grid);
    sorted_free(&S);
    cplan_free(&p);
    return 0;
}

typedef struct {
    cplan *p;
    const float *x;
    const cf *grid;
    float *out;
    long M, next;
    pthread_mutex_t *mu;
} interp_job;

static void *interp_worker(void *arg)
{
    interp_job *J = (interp_job *)arg;
    cplan *p = J->p;
    int d = p->d, w = p->w;
    for (;;) {
        pthread_mutex_lock(J->mu);
        long first = J->next;
        J->next += 1024;
        pthread_mutex_unlock(J->mu);
        if (first >= J->M) break;
        long last = first + 1024 < J->M ? first + 1024 : J->M;
        for (long j = first; j < last; ++j) {
            float wt[3][16];
            long st[3] = {0, 0, 0};
            for (int a = 0; a < d; ++a) st[a] = support(p, a, J->x[j * d + a], wt[a]);
            float sre = 0, sim = 0;
            long n0 = p->n[0];
            int ny = d >= 2 ? w : 1, nz = d == 3 ? w : 1;
            for (int dz = 0; dz < nz; ++dz) {
                long cz = d == 3 ? (st[2] + dz) % p->n[2] : 0;
                for (int dy = 0; dy < ny; ++dy) {
                    long cy = d >= 2 ? (st[1] + dy) % p->n[1] : 0;
                    float s = (d >= 2 ? wt[1][dy] : 1.0f) * (d == 3 ? wt[2][dz] : 1.0f);
                    const cf *row = J->grid + (cz * p->n[1] + cy) * n0;
                    float rr = 0, ri = 0;
                    for (int i = 0; i < w; ++i) {
                        long xx = st[0] + i;
                        if (xx >= n0) xx -= n0;
                        rr += row[xx].re * wt[0][i];
                        ri += row[xx].im * wt[0][i];
                    }
                    sre += rr * s;
                    sim += ri * s;
                }
            }
            J->out[2 * j] = sre;
            J->out[2 * j + 1] = sim;
        }
    }
    return NULL;
}

/* One type-2 execute: predeconvolution + FFT + interpolation. */
int cpu_type2(int d, const long *nm, double eps, int isign, long M, const float *x,
              const float *f, float *out, int threads)
{
    cplan p;
    if (cplan_init(&p, d, nm, eps, isign, threads)) return -1;
    cf *grid = calloc(p.cells, sizeof(cf));
    const cf *fi = (const cf *)f;
    for (long k = 0; k < p.modes; ++k) {
        long rest = k, lin = 0, stride = 1;
        float a = 1.0f;
        for (int ax = 0; ax < d; ++ax) {
            long m, fidx = mode_fine_index(rest % p.nm[ax], p.nm[ax], p.n[ax], &m);
            rest /= p.nm[ax];
            lin += fidx * stride;
            stride *= p.n[ax];
            a *= p.amp[ax][m];
        }
        grid[lin].re = fi[k].re * a;
        grid[lin].im = fi[k].im * a;
    }
    fft_grid(&p, grid);
    pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
    interp_job J = {&p, x, grid, out, M, 0, &mu};
    run_threads(p.threads, interp_worker, &J);
    free(grid);
    cplan_free(&p);
    return 0;
}

int cpu_threads(void)
{
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n < 1 ? 1 : (int)n;
}
