// Device-resident NUFFT plans: the MI355X counterpart of the reference's
// NufftPlan (src/plan.rs:403-1229) and NufftType3Plan (src/type3.rs:346-539).
//
// A plan owns its scratch (bins, prepared records, fine grid, rocFFT work
// area) and enqueues every stage on the caller's HIP stream; executions of
// one plan must stay in stream order (src/plan.rs:452-456).
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

#include "bins.hpp"
#include "interp.hpp"
#include "modes.hpp"
#include "plan.hpp"
#include "spread.hpp"

namespace mnufft {

namespace {

std::once_flag g_rocfft_once;
bool g_rocfft_ok = false;

bool rocfft_ready()
{
    std::call_once(g_rocfft_once, [] { g_rocfft_ok = rocfft_setup() == rocfft_status_success; });
    if (!g_rocfft_ok) set_error("rocFFT setup failed");
    return g_rocfft_ok;
}

int64_t pow2_ceil(int64_t v)
{
    int64_t p = 1;
    while (p < v) p *= 2;
    return p;
}

int log2i(int64_t v)
{
    int s = 0;
    while (((int64_t)1 << s) < v) ++s;
    return s;
}

}  // namespace

SpreadParams make_spread_params(int d, const int64_t *fine, int width, bool f64)
{
    SpreadParams sp{};
    // cells per lane along axis 0: 8 in 3D, 4 in 1D, in 2D 2 for F32 (blocks
    // of 8 x 16 cells, the reference's 2D F32 block; measured 1.1-1.6x
    // faster than 16 x 16) and 4 for F64, from 4D on 16 (F32) or 8 (F64) as
    // the reference's max_run (src/gpu_nd_spread.rs:575-582); never more
    // than n0 + 1 - w, so a run sees every support offset at most once
    int run = d == 3 ? 8 : (d >= 4 ? (f64 ? 8 : 16) : (d == 2 && !f64 ? 2 : 4));
    while (run > 2 && run > fine[0] + 1 - width) run /= 2;
    int64_t lanes_other = 1;
    int64_t block[kMaxDims] = {1, 1, 1, 1, 1, 1, 1, 1};
    if (d == 3) {
        block[1] = 8;
        block[2] = 8;
    } else if (d == 2) {
        block[1] = 16;
    } else if (d >= 4) {
        // one cell per lane along the other axes: their block edges doubled
        // in turn up to 64 lanes, the smallest edge first (ties: the longer
        // axis), each at most 16 cells and n + 1 - w, as the reference's
        // layout does for 64-invocation workgroups (194-260)
        int64_t inv = 1;
        for (;;) {
            int best = -1;
            for (int a = 1; a < d; ++a) {
                const int64_t lim = std::min<int64_t>(std::max<int64_t>(fine[a] + 1 - width, 1), 16);
                if (block[a] * 2 <= lim && inv * 2 <= 64 &&
                    (best < 0 || block[a] < block[best] ||
                     (block[a] == block[best] && fine[a] > fine[best])))
                    best = a;
            }
            if (best < 0) break;
            block[best] *= 2;
            inv *= 2;
        }
    }
    for (int a = 1; a < d; ++a) lanes_other *= block[a];
    int groups0 = (int)(64 / lanes_other);
    block[0] = (int64_t)run * groups0;
    sp.run = run;
    sp.groups0 = groups0;
    sp.block_count = 1;
    for (int a = 0; a < kMaxDims; ++a) {
        sp.block[a] = a < d ? block[a] : 1;
        sp.nblocks[a] = a < d ? (fine[a] + block[a] - 1) / block[a] : 1;
        sp.block_count *= sp.nblocks[a];
    }
    return sp;
}

SpreadParams make_tile_params(int d, const int64_t *fine, const int64_t *block)
{
    SpreadParams q{};
    q.run = 0;
    q.groups0 = 0;
    q.block_count = 1;
    for (int a = 0; a < kMaxDims; ++a) {
        q.block[a] = a < d ? block[a] : 1;
        q.nblocks[a] = a < d ? (fine[a] + q.block[a] - 1) / q.block[a] : 1;
        q.block_count *= q.nblocks[a];
    }
    return q;
}

bool make_sweep_params(int d, const int64_t *fine, int width, int64_t tx, int max_width,
                       SpreadParams *sp)
{
    if (d != 3 || width > max_width) return false;
    // every axis at least 16 + width cells (the matrix-core tiles), axis 0
    // at least tx + width + 1 (the plain row tiles' x reach within the grid)
    if (fine[0] < 16 + width + 1 || fine[1] < 16 + width || fine[2] < 16 + width) return false;
    if (fine[0] < tx + width + 1) return false;
    const int64_t block[kMaxDims] = {tx, 8, 8, 1, 1, 1, 1, 1};
    *sp = make_tile_params(d, fine, block);
    return true;
}

bool make_tile2_params(int d, const int64_t *fine, int width, SpreadParams *sp)
{
    constexpr int64_t tile = 32;
    if (d != 2 || width > 8) return false;
    if (fine[0] < tile + width + 1 || fine[1] < tile + width) return false;
    SpreadParams q{};
    q.run = 0;
    q.groups0 = 0;
    q.block_count = 1;
    for (int a = 0; a < kMaxDims; ++a) {
        q.block[a] = a < d ? tile : 1;
        q.nblocks[a] = a < d ? (fine[a] + tile - 1) / tile : 1;
        q.block_count *= q.nblocks[a];
    }
    *sp = q;
    return true;
}

FftPlan::~FftPlan()
{
    if (info) rocfft_execution_info_destroy((rocfft_execution_info)info);
    if (plan) rocfft_plan_destroy((rocfft_plan)plan);
    if (work) (void)hipFree(work);
}

Plan::~Plan()
{
    for (void *p : {(void *)d_inv_phihat, grid, (void *)rec_start, rec_off, (void *)rec_perm,
                    c_sorted, t3_xr, t3_cp, t3_sr, t3_factor, t3_interp, t3_box, t3_perm,
                    t3_s_sorted, t3_counts, (void *)item_off, (void *)batch_table, (void *)batch_modes,
                    (void *)batch_spread, fft_scratch, (void *)line_jobs})
        if (p) (void)hipFree(p);
    for (auto &sl : slots)
        for (auto &e : sl.ev)
            if (e) (void)hipEventDestroy(e);
}

// ----------------------------------------------------------------------
// Plan creation.
// ----------------------------------------------------------------------
bool Plan::init_common()
{
    if (!es_kernel_for_tolerance(eps, sigma, &kernel)) return false;
    mode_total = 1;
    fine_total = 1;
    for (int a = 0; a < dim; ++a) {
        if (!fine_given) {
            fine[a] = select_fine_grid_size(n_modes[a], sigma, kernel.width);
            if (fine[a] < 0) {
                set_error("no supported fine-grid length for axis " + std::to_string(a));
                return false;
            }
        }
        mode_total *= n_modes[a];
        fine_total *= fine[a];
    }
    std::vector<std::vector<double>> coeffs;
    std::vector<double> table;
    for (int a = 0; a < dim; ++a) {
        coeffs.push_back(centered_fourier_coefficients(kernel, fine[a]));
        phihat_off[a] = (int64_t)table.size();
        for (double v : coeffs.back()) table.push_back(1.0 / v);
    }
    if (!fine_given &&
        !validate_f32_rounding(std::vector<int64_t>(n_modes, n_modes + dim), eps, sigma, precision,
                               coeffs))
        return false;
    if (!hip_ok(hipMalloc(&d_inv_phihat, sizeof(double) * table.size()), "amplitude table"))
        return false;
    if (!hip_ok(hipMemcpy(d_inv_phihat, table.data(), sizeof(double) * table.size(),
                          hipMemcpyHostToDevice),
                "amplitude upload"))
        return false;
    layout = build_spread_layout(dim, fine, kernel.width,
                                 precision == Precision::F32 ? Precision::F32 : Precision::F64);
    sp = make_spread_params(dim, fine, kernel.width, precision != Precision::F32);
    sp_generic = sp;
    // plain row tiles of 8 x 8 x 8 cells (FP64 plans up to width 16)
    sweep = make_sweep_params(dim, fine, kernel.width, 8,
                              precision == Precision::F32 ? 8 : kMaxRowsWidth, &sp);
    tiles2 = make_tile2_params(dim, fine, kernel.width, &tile2_sp);
    // plain 2D row tiles of 16 x 16 cells: widths 2..8, every axis at
    // least 16 + width cells (axis 0 one more)
    rows2 = dim == 2 && kernel.width >= 2 && kernel.width <= 8 &&
            fine[0] >= 16 + kernel.width + 1 &&
            fine[1] >= 16 + kernel.width;
    if (rows2) {
        const int64_t block[kMaxDims] = {16, 16, 1, 1, 1, 1, 1, 1};
        rows2_sp = make_tile_params(2, fine, block);
    }
    if (sweep && kernel.width <= 8) {
        // matrix-core tiles: 16 x 16 cells, 8 deep along z
        const int64_t block[kMaxDims] = {16, 16, 8, 1, 1, 1, 1, 1};
        tile3_sp = make_tile_params(3, fine, block);
    }
    return true;
}

GridParams Plan::grid_params(const int64_t *side) const
{
    GridParams g{};
    g.dims = dim;
    g.width = kernel.width;
    g.half_width = kernel.half_width();
    g.beta = kernel.beta;
    g.nbins = 1;
    g.fine_total = fine_total;
    for (int a = 0; a < kMaxDims; ++a) {
        AxisParams &x = g.ax[a];
        if (a >= dim) {
            x.n = 1;
            x.bins = 1;
            x.stride = g.nbins;
            continue;
        }
        x.n = fine[a];
        x.shift = log2i(side[a]);
        x.bins = (fine[a] + side[a] - 1) / side[a];
        x.stride = g.nbins;
        g.nbins *= x.bins;
        double s = (double)fine[a] / 6.283185307179586;
        x.scale_hi = (float)s;
        x.scale_lo = (float)(s - (double)x.scale_hi);
        x.scale = s;
        x.origin = (double)(fine[a] / 2);
    }
    return g;
}

GridParams Plan::spread_geometry(int64_t M, int *chosen) const
{
    if (!rows_geometry()) return type1_geometry(M, chosen);
    // one-cell bins along the other axes; along axis 0 the smallest power of
    // two from 2 that keeps the bin count within 2^24
    int64_t side[kMaxDims] = {2, 1, 1, 1, 1, 1, 1, 1};
    int64_t lines = 1;
    for (int a = 1; a < dim; ++a) lines *= fine[a];
    while (side[0] < fine[0] && ((fine[0] + side[0] - 1) / side[0]) * lines > ((int64_t)1 << 24))
        side[0] *= 2;
    // a tuned side (a power of two): coarser only, never beyond the axis
    if (rows_side0 > side[0]) side[0] = std::min<int64_t>(rows_side0, pow2_ceil(fine[0]));
    if (chosen) *chosen = -1;
    return grid_params(side);
}

GridParams Plan::type1_geometry(int64_t M, int *chosen) const
{
    int gi = layout.choose_geometry(M);
    if (chosen) *chosen = gi;
    return grid_params(layout.geometries[gi].side);
}

bool Plan::tiled_interp() const
{
    if (dim != 3 || kernel.width > 8) return false;
    for (int a = 0; a < 3; ++a)
        if (fine[a] < 8 + kernel.width - 1) return false;
    return true;
}

bool Plan::tiled2_interp() const
{
    if (dim != 2 || kernel.width > 8) return false;
    for (int a = 0; a < 2; ++a)
        if (fine[a] < 8 + kernel.width - 1 || fine[a] < 8) return false;
    return true;
}

GridParams Plan::type2_geometry() const
{
    int64_t side[kMaxDims] = {1, 1, 1, 1, 1, 1, 1, 1};
    for (int a = 0; a < dim; ++a) {
        int64_t s = dim >= 4 ? 4 : dim == 3 ? (tiled_interp() ? 8 : 4) : dim == 2 ? 8 : 32;
        side[a] = std::min<int64_t>(s, pow2_ceil(fine[a]));
    }
    return grid_params(side);
}

ModeParams Plan::mode_params() const
{
    ModeParams mp{};
    for (int a = 0; a < kMaxDims; ++a) {
        mp.n_modes[a] = a < dim ? n_modes[a] : 1;
        mp.fine[a] = a < dim ? fine[a] : 1;
        mp.phihat_off[a] = a < dim ? phihat_off[a] : 0;
    }
    mp.mode_total = mode_total;
    mp.fine_total = fine_total;
    mp.centered = mode_order == ModeOrder::Centered;
    return mp;
}

size_t Plan::real_bytes() const { return precision == Precision::F32 ? 4 : 8; }

bool Plan::ensure_grid(int64_t vectors)
{
    size_t need = (size_t)vectors * (size_t)fine_total * 2 * real_bytes();
    if (need <= grid_bytes) return true;
    if (grid) (void)hipFree(grid);
    grid = nullptr;
    grid_bytes = 0;
    if (!hip_ok(hipMalloc(&grid, need), "fine grid")) return false;
    grid_bytes = need;
    return true;
}

bool Plan::ensure_records(int64_t M)
{
    if (M <= rec_cap) return true;
    if (rec_start) (void)hipFree(rec_start);
    if (rec_off) (void)hipFree(rec_off);
    if (rec_perm) (void)hipFree(rec_perm);
    rec_start = nullptr;
    rec_off = nullptr;
    rec_perm = nullptr;
    int64_t cap = M + M / 4 + 64;
    const size_t quads = dim <= 3 ? 1 : 2;  // rec_quads of the plan's rank
    if (!hip_ok(hipMalloc(&rec_start, sizeof(int4) * quads * cap), "prepared starts"))
        return false;
    if (!hip_ok(hipMalloc(&rec_off, 4 * quads * real_bytes() * cap), "prepared offsets"))
        return false;
    if (!hip_ok(hipMalloc(&rec_perm, sizeof(uint32_t) * cap), "record permutation")) return false;
    rec_cap = cap;
    return true;
}

bool Plan::ensure_items(int64_t n)
{
    if (n <= item_cap) return true;
    if (item_off) (void)hipFree(item_off);
    item_off = nullptr;
    if (!hip_ok(hipMalloc(&item_off, sizeof(uint32_t) * n), "interpolation items")) return false;
    item_cap = n;
    return true;
}

bool Plan::ensure_strengths(int64_t M, int64_t vectors)
{
    size_t need = (size_t)std::max<int64_t>(M, 1) * vectors * 2 * real_bytes();
    if (need <= c_sorted_bytes) return true;
    if (c_sorted) (void)hipFree(c_sorted);
    c_sorted = nullptr;
    if (!hip_ok(hipMalloc(&c_sorted, need), "sorted strengths")) return false;
    c_sorted_bytes = need;
    return true;
}

namespace {

// An in-place batched C2C plan over `rank` contiguous axes (fastest first)
// with its execution info and work buffer.
std::unique_ptr<FftPlan> create_fft(int rank, const int64_t *len, int64_t batch, int isign,
                                    Precision precision)
{
    if (!rocfft_ready()) return nullptr;
    auto fp = std::make_unique<FftPlan>();
    size_t lengths[3];
    for (int a = 0; a < rank; ++a) lengths[a] = (size_t)len[a];
    rocfft_transform_type tt = isign >= 0 ? rocfft_transform_type_complex_inverse
                                          : rocfft_transform_type_complex_forward;
    rocfft_precision pr =
        precision == Precision::F32 ? rocfft_precision_single : rocfft_precision_double;
    rocfft_plan p = nullptr;
    if (rocfft_plan_create(&p, rocfft_placement_inplace, tt, pr, (size_t)rank, lengths,
                           (size_t)batch, nullptr) != rocfft_status_success) {
        set_error("rocFFT plan creation failed");
        return nullptr;
    }
    fp->plan = p;
    size_t ws = 0;
    rocfft_plan_get_work_buffer_size(p, &ws);
    rocfft_execution_info info = nullptr;
    rocfft_execution_info_create(&info);
    fp->info = info;
    if (ws) {
        if (!hip_ok(hipMalloc(&fp->work, ws), "rocFFT work buffer")) return nullptr;
        rocfft_execution_info_set_work_buffer(info, fp->work, ws);
    }
    return fp;
}

bool run_fft(FftPlan *fp, void *data, hipStream_t stream)
{
    rocfft_execution_info info = (rocfft_execution_info)fp->info;
    rocfft_execution_info_set_stream(info, stream);
    void *buf[1] = {data};
    if (rocfft_execute((rocfft_plan)fp->plan, buf, nullptr, info) != rocfft_status_success) {
        set_error("rocFFT execution failed");
        return false;
    }
    return true;
}

// Axis groups of the rank 4-8 FFT, fastest axes first: an even number of
// groups, so the transposes (one per group) leave the result in the grid.
int fft_axis_groups(int d, int *len)
{
    static const int groups[5][4] = {{3, 1}, {3, 2}, {3, 3}, {2, 2, 2, 1}, {2, 2, 2, 2}};
    const int n = d <= 6 ? 2 : 4;
    for (int g = 0; g < n; ++g) len[g] = groups[d - 4][g];
    return n;
}

}  // namespace

FftPlan *Plan::fft_for(int64_t vectors)
{
    auto it = ffts.find(vectors);
    if (it != ffts.end()) return it->second.get();
    auto fp = create_fft(dim, fine, vectors, isign, precision);
    if (!fp) return nullptr;
    FftPlan *raw = fp.get();
    ffts[vectors] = std::move(fp);
    return raw;
}

FftPlan *Plan::fft_group_plan(int rank, const int64_t *lengths, int64_t batch)
{
    std::vector<int64_t> key{rank};
    for (int a = 0; a < rank; ++a) key.push_back(lengths[a]);
    key.push_back(batch);
    auto it = fft_groups.find(key);
    if (it != fft_groups.end()) return it->second.get();
    auto fp = create_fft(rank, lengths, batch, isign, precision);
    if (!fp) return nullptr;
    FftPlan *raw = fp.get();
    fft_groups[key] = std::move(fp);
    return raw;
}

bool Plan::fft_nd(int64_t vectors, void *data, hipStream_t stream, bool run)
{
    const size_t need = (size_t)vectors * (size_t)fine_total * 2 * real_bytes();
    if (need > fft_scratch_bytes) {
        if (fft_scratch) (void)hipFree(fft_scratch);
        fft_scratch = nullptr;
        fft_scratch_bytes = 0;
        if (!hip_ok(hipMalloc(&fft_scratch, need), "FFT transpose buffer")) return false;
        fft_scratch_bytes = need;
    }
    int len[4];
    const int ng = fft_axis_groups(dim, len);
    int order[kMaxDims];
    for (int a = 0; a < dim; ++a) order[a] = a;
    void *cur = data, *other = fft_scratch;
    const size_t elem = 2 * real_bytes();
    for (int gi = 0; gi < ng; ++gi) {
        const int r = len[gi];
        int64_t lengths[3], lead = 1;
        for (int a = 0; a < r; ++a) {
            lengths[a] = fine[order[a]];
            lead *= lengths[a];
        }
        const int64_t rest = fine_total / lead;
        // the lead axes of every (rest index, vector), at most 2^31 complex
        // values per rocFFT execution
        const int64_t total = rest * vectors;
        const int64_t per = std::max<int64_t>(1, ((int64_t)1 << 31) / lead);
        for (int64_t b0 = 0; b0 < total; b0 += per) {
            FftPlan *fp = fft_group_plan(r, lengths, std::min<int64_t>(per, total - b0));
            if (!fp) return false;
            if (run && !run_fft(fp, (char *)cur + (size_t)b0 * (size_t)lead * elem, stream))
                return false;
        }
        if (run) {
            const bool ok = precision == Precision::F32
                                ? transpose_complex<float>((const float *)cur, (float *)other,
                                                           rest, lead, vectors, stream)
                                : transpose_complex<double>((const double *)cur, (double *)other,
                                                            rest, lead, vectors, stream);
            if (!ok) return false;
        }
        std::swap(cur, other);
        int rotated[kMaxDims];
        for (int a = 0; a < dim; ++a) rotated[a] = order[(a + r) % dim];
        for (int a = 0; a < dim; ++a) order[a] = rotated[a];
    }
    return true;
}

// The rocFFT plans (and work buffers) an FFT of `vectors` vectors uses,
// created right after the grid so that later scratch (the dense-bin windows)
// does not change where they land.
bool Plan::use_line_fft() const
{
    return fft_engine == 0 && (type == 1 || type == 2) &&
           LineFft::supported(dim, fine, precision != Precision::F32);
}

bool Plan::ensure_fft(int64_t vectors)
{
    if (use_line_fft()) {
        if (line_fft) return true;
        auto lf = std::make_unique<LineFft>();
        if (!lf->init(dim, fine, n_modes, type, precision != Precision::F32, isign, d_inv_phihat,
                      phihat_off, mode_order == ModeOrder::Centered))
            return false;
        line_fft = std::move(lf);
        return true;
    }
    if (dim >= 4) return fft_nd(vectors, nullptr, nullptr, false);
    const int64_t per_exec = std::max<int64_t>(1, ((int64_t)1 << 31) / fine_total);
    for (int64_t v0 = 0; v0 < vectors; v0 += per_exec)
        if (!fft_for(std::min<int64_t>(per_exec, vectors - v0))) return false;
    return true;
}

bool Plan::fft(int64_t vectors, void *data, hipStream_t stream)
{
    if (use_line_fft()) {
        if (!ensure_fft(vectors)) return false;
        return line_fft->run(data, vectors, stream);
    }
    if (dim >= 4) return fft_nd(vectors, data, stream, true);
    // Batches of at most 2^31 complex values per rocFFT execution: larger
    // batched executions were seen to corrupt their last transform.
    const int64_t per_exec = std::max<int64_t>(1, ((int64_t)1 << 31) / fine_total);
    const size_t vbytes = (size_t)fine_total * 2 * real_bytes();
    for (int64_t v0 = 0; v0 < vectors; v0 += per_exec) {
        const int64_t vc = std::min<int64_t>(per_exec, vectors - v0);
        FftPlan *fp = fft_for(vc);
        if (!fp || !run_fft(fp, (char *)data + (size_t)v0 * vbytes, stream)) return false;
    }
    return true;
}

bool Plan::begin_call()
{
    if (!timing) return true;
    slot = history ? slots_filled++ : 0;
    if (!history) slots_filled = 1;
    while ((int64_t)slots.size() <= slot) {
        slots.emplace_back();
        for (auto &e : slots.back().ev)
            if (!hip_ok(hipEventCreate(&e), "timing event")) return false;
    }
    for (auto &u : slots[slot].used) u = false;
    return true;
}

void Plan::mark(int stage, bool end, hipStream_t stream)
{
    if (!timing || slot >= (int64_t)slots.size()) return;
    (void)hipEventRecord(slots[slot].ev[2 * stage + (end ? 1 : 0)], stream);
    slots[slot].used[stage] = true;
}

void Plan::reset_marks() { (void)begin_call(); }

// ----------------------------------------------------------------------
// Type 1 and type 2.
// ----------------------------------------------------------------------
template <int D, typename T>
bool Plan::set_points_t(int64_t M, const T *pts, hipStream_t stream)
{
    if (!point_count_ok(M, "point")) return false;
    reset_marks();
    points = pts;
    point_count = M;
    points_set = false;
    if (type == 1) {
        gp = spread_geometry(M, &geometry_index);
        gp_rows = rows_geometry();
        binned_side0 = rows_side0;
        binned_order = point_order;
        binned_dense = sizeof(T) == 4 ? dense_threshold(gp) : 0u;
        mark(kStageSort, false, stream);
        if (!sorter.sort<D, T>(M, pts, gp, point_order, stream)) return false;
        mark(kStageSort, true, stream);
        if (!ensure_records(M)) return false;
        mark(kStagePrepare, false, stream);
        if (binned_dense && !dense.classify(gp, spread_params(), sorter.offsets(), stream))
            return false;
        if (!sorter.prepare<D, T>(M, pts, gp, point_order, rec_start, (T *)rec_off, rec_perm,
                                  stream, binned_dense ? dense.before : nullptr))
            return false;
        mark(kStagePrepare, true, stream);
    } else {
        gp = type2_geometry();
        mark(kStageSort, false, stream);
        if (!sorter.sort<D, T>(M, pts, gp, BinOrder::Grouped, stream)) return false;
        mark(kStageSort, true, stream);
    }
    points_set = true;
    return true;
}

template <int D, typename T>
bool Plan::execute_t(int64_t active, const T *in, T *out, hipStream_t stream, bool fresh_marks)
{
    if (fresh_marks) reset_marks();
    const int64_t M = point_count;
    if (!ensure_grid(active) || (!spread_only && !ensure_fft(active))) return false;
    ModeParams mp = mode_params();
    if (type == 1) {
        if (M == 0) return fill_zero<T>(out, 2 * mode_total * active, stream);
        // a spread method that needs other bins, or another in-bin order,
        // re-bins the kept points
        if ((gp_rows != rows_geometry() || binned_order != point_order ||
             binned_side0 != rows_side0 ||
             binned_dense != (sizeof(T) == 4 ? dense_threshold(gp) : 0u)) &&
            !set_points_t<D, T>(M, (const T *)points, stream))
            return false;
        if (!ensure_strengths(M, active)) return false;
        mark(kStageSpread, false, stream);
        bool timed = timing && slot < (int64_t)slots.size();
        spreader.kernel_start = timed ? slots[slot].ev[2 * kStageSpreadKernel] : nullptr;
        spreader.kernel_end = timed ? slots[slot].ev[2 * kStageSpreadKernel + 1] : nullptr;
        if (timed) slots[slot].used[kStageSpreadKernel] = true;
        // dense-bin records read their strengths in point order
        if (!gather_strengths<T>(M, active, rec_perm, in, (T *)c_sorted, stream,
                                 binned_dense ? dense.before + gp.nbins : nullptr))
            return false;
        if constexpr (sizeof(T) == 4) dense.strengths = in;
        if (!spreader.spread<D, T>(gp, spread_params(),
                                   binned_dense ? dense.light : sorter.offsets(), rec_start,
                                   (const T *)rec_off, (const T *)c_sorted, M, active, (T *)grid,
                                   stream, spread_method, binned_dense ? &dense : nullptr))
            return false;
        mark(kStageSpread, true, stream);
        if (spread_only) return true;
        mark(kStageFft, false, stream);
        if (use_line_fft()) {
            // the last line pass writes the deconvolved modes
            if (!ensure_fft(active) || !line_fft->run(grid, active, stream, out)) return false;
            mark(kStageFft, true, stream);
            return true;
        }
        if (!fft(active, grid, stream)) return false;
        mark(kStageFft, true, stream);
        mark(kStageDeconv, false, stream);
        if (!deconvolve<D, T>(mp, d_inv_phihat, (const T *)grid, out, active, stream))
            return false;
        mark(kStageDeconv, true, stream);
    } else {
        if (M == 0) return true;
        if (use_line_fft()) {
            // the first line pass reads the modes and predeconvolves them
            mark(kStageFft, false, stream);
            if (!ensure_fft(active) || !line_fft->run(grid, active, stream, (void *)in))
                return false;
            mark(kStageFft, true, stream);
        } else {
            mark(kStageDeconv, false, stream);
            if (!predeconvolve<D, T>(mp, d_inv_phihat, in, (T *)grid, active, stream))
                return false;
            mark(kStageDeconv, true, stream);
            mark(kStageFft, false, stream);
            if (!fft(active, grid, stream)) return false;
            mark(kStageFft, true, stream);
        }
        mark(kStageInterp, false, stream);
        if (D == 3 && tiled_interp()) {
            // bin item offsets and the scratch of their scan
            if (!ensure_items(gp.nbins + 1 + scan_tiles(gp.nbins + 1) + 1)) return false;
            if (!interpolate_tiled3<T>(M, active, sorter.sorted(), sorter.offsets(), item_off,
                                       (const T *)points, gp, (const T *)grid, out, stream))
                return false;
        } else if (D == 2 && tiled2_interp()) {
            if (!interpolate_tiled2<T>(M, active, sorter.sorted(), sorter.offsets(),
                                       (const T *)points, gp, (const T *)grid, out, stream))
                return false;
        } else if (!interpolate<D, T>(M, active, sorter.sorted(), (const T *)points, gp,
                                      (const T *)grid, out, stream)) {
            return false;
        }
        mark(kStageInterp, true, stream);
    }
    return true;
}

template <int D, typename T>
bool Plan::bin_sort_t(int64_t M, const T *pts, uint32_t *sorted, uint32_t *offsets,
                      hipStream_t stream)
{
    if (!point_count_ok(M, "point")) return false;
    GridParams g = type1_geometry(M, nullptr);
    reset_marks();
    mark(kStageSort, false, stream);
    if (!sorter.sort<D, T>(M, pts, g, BinOrder::Index, stream)) return false;
    if (!ensure_records(M)) return false;
    if (!sorter.prepare<D, T>(M, pts, g, BinOrder::Index, rec_start, (T *)rec_off, sorted, stream))
        return false;
    mark(kStageSort, true, stream);
    if (offsets &&
        !hip_ok(hipMemcpyAsync(offsets, sorter.offsets(), sizeof(uint32_t) * (g.nbins + 1),
                               hipMemcpyDeviceToDevice, stream),
                "offsets copy"))
        return false;
    return true;
}

template <int D, typename T>
bool Plan::bin_sort_static_t(int64_t M, const T *pts, uint32_t *sorted, uint32_t *offsets,
                             hipStream_t stream)
{
    if (!point_count_ok(M, "point")) return false;
    const GridParams g = type1_geometry(M, nullptr);
    void *base = static_sort_scratch();
    if (!base) return false;
    // a sorter bound to the static region: its buffers are views, never
    // allocated or freed
    BinSorter s;
    if (!s.bind_external(base, static_sort_bytes(), g.nbins, M, sorted, offsets)) return false;
    return s.sort<D, T>(M, pts, g, BinOrder::Index, stream);
}

// ----------------------------------------------------------------------
// Type 3: rescale + pre-phase, raw spread onto the outer grid, target
// rescale + correction, inner type 2 on the outer grid, final multiply
// (src/gpu_type3.rs:267-430).
// ----------------------------------------------------------------------
bool Plan::init_type3(const double *src_bounds, const double *tgt_bounds, const int *min_boxes)
{
    if (!es_kernel_for_tolerance(eps, sigma, &kernel)) return false;
    std::memset(&t3, 0, sizeof(t3));
    t3.sign = isign >= 0 ? 1.0 : -1.0;
    for (int a = 0; a < 2 * dim; ++a) {
        t3_bounds[a] = src_bounds[a];
        t3_bounds[2 * kMaxDims + a] = tgt_bounds[a];
    }
    // source half-width and centre, target half-width and centre per axis
    // (src/type3.rs:708-797)
    double X[kMaxDims] = {}, C[kMaxDims] = {}, S[kMaxDims] = {}, Dc[kMaxDims] = {};
    double tlo[kMaxDims] = {}, thi[kMaxDims] = {};
    for (int a = 0; a < dim; ++a) {
        double slo = src_bounds[2 * a], shi = src_bounds[2 * a + 1];
        tlo[a] = tgt_bounds[2 * a];
        thi[a] = tgt_bounds[2 * a + 1];
        if (!(std::isfinite(slo) && std::isfinite(shi) && std::isfinite(tlo[a]) &&
              std::isfinite(thi[a])) ||
            slo > shi || tlo[a] > thi[a]) {
            set_error("type-3 bounds must be finite with lower <= upper");
            return false;
        }
        X[a] = (shi - slo) * 0.5;
        C[a] = (shi + slo) * 0.5;
        if (std::fabs(C[a]) < 0.1 * X[a]) {
            X[a] += std::fabs(C[a]);
            C[a] = 0.0;
        }
        S[a] = (thi[a] - tlo[a]) * 0.5;
        Dc[a] = (thi[a] + tlo[a]) * 0.5;
        if (std::fabs(Dc[a]) < 0.1 * S[a]) {
            S[a] += std::fabs(Dc[a]);
            Dc[a] = 0.0;
        }
    }
    // outer grid length, gamma and h of an axis for a target half-width
    auto axis_grid = [&](int a, double half, int64_t *nf_out, double *gamma_out, double *h_out) {
        double x = X[a], s = half;
        if (x == 0.0) {
            if (s == 0.0) {
                x = 1.0;
                s = 1.0;
            } else {
                x = 1.0 / s;
            }
        } else {
            s = std::max(s, 1.0 / x);
        }
        double requested = 2.0 * sigma * s * x / M_PI + (double)(kernel.width + 1);
        int64_t minimum = std::max<int64_t>((int64_t)requested, 2 * (int64_t)kernel.width);
        int64_t nf = next_even_235(minimum);
        *nf_out = nf;
        *h_out = 2.0 * M_PI / (double)nf;
        *gamma_out = (double)nf / (2.0 * sigma * s);
    };
    // Target boxes: the whole interval per axis unless the inner fine grid
    // would exceed kType3InnerCells cells or the byte budget; then the axis
    // with the longest inner grid is cut into one more box, until it fits.
    int k[kMaxDims] = {1, 1, 1, 1, 1, 1, 1, 1};
    for (int a = 0; a < dim; ++a)
        if (min_boxes && min_boxes[a] > 1) k[a] = std::min(min_boxes[a], kType3MaxBoxes);
    int64_t outer_n[kMaxDims] = {1, 1, 1, 1, 1, 1, 1, 1};
    double gam[kMaxDims] = {1, 1, 1, 1, 1, 1, 1, 1}, hh[kMaxDims] = {1, 1, 1, 1, 1, 1, 1, 1};
    for (;;) {
        int64_t fine_len[kMaxDims] = {1, 1, 1, 1, 1, 1, 1, 1};
        double cells = 1.0;
        for (int a = 0; a < dim; ++a) {
            const double half = k[a] == 1 ? S[a] : (thi[a] - tlo[a]) / (2.0 * k[a]);
            axis_grid(a, half, &outer_n[a], &gam[a], &hh[a]);
            fine_len[a] = select_fine_grid_size(outer_n[a], sigma, kernel.width);
            if (fine_len[a] < 0) {
                set_error("no supported inner fine-grid length for type-3 axis " +
                          std::to_string(a));
                return false;
            }
            cells *= (double)fine_len[a];
        }
        const double bytes = cells * 2.0 * (double)real_bytes() * (double)batch;
        // 3D inner grids on the line FFT: axes short enough for 64-byte rows
        // in the strided passes (a long axis costs more than extra boxes)
        bool rows_ok = true;
        if (dim == 3 && LineFft::supported(dim, fine_len, precision != Precision::F32))
            for (int a = 0; a < dim; ++a)
                rows_ok = rows_ok && fine_len[a] <= LineFft::row_length_limit(
                                                        precision != Precision::F32);
        const bool fits = cells <= (double)kType3InnerCells && bytes <= (double)t3_inner_budget;
        if (fits && rows_ok) break;
        // the longest splittable axis (once the grid fits: one beyond the row limit)
        int split = -1;
        for (int a = 0; a < dim; ++a)
            if (thi[a] > tlo[a] && k[a] < kType3MaxBoxes &&
                (!fits || fine_len[a] > LineFft::row_length_limit(precision != Precision::F32)) &&
                (split < 0 || fine_len[a] > fine_len[split]))
                split = a;
        if (split < 0 && fits) break;
        if (split < 0) {
            set_error("type-3 inner grid too large even with the target interval cut into boxes");
            return false;
        }
        ++k[split];
    }
    t3_boxes = Type3Boxes{};
    t3_boxes.count = 1;
    for (int a = 0; a < kMaxDims; ++a) {
        t3_boxes.k[a] = a < dim ? k[a] : 1;
        t3_boxes.lo[a] = a < dim ? tlo[a] : 0.0;
        t3_boxes.width[a] =
            a < dim && thi[a] > tlo[a] ? (thi[a] - tlo[a]) / (double)k[a] : 1.0;
        t3_boxes.count *= t3_boxes.k[a];
        t3_center_full[a] = a < dim ? Dc[a] : 0.0;
    }
    for (int a = 0; a < dim; ++a) {
        t3.src_center[a] = C[a];
        t3.inv_gamma[a] = 1.0 / gam[a];
        t3.tgt_center[a] = k[a] == 1 ? Dc[a] : tlo[a] + 0.5 * t3_boxes.width[a];
        t3.tgt_scale[a] = hh[a] * gam[a];
    }
    int pn = kernel.width + 2;
    std::vector<double> nodes(2 * pn), weights(2 * pn);
    gauss_legendre(2 * pn, nodes.data(), weights.data());
    double hw = kernel.half_width();
    t3.nquad = pn;
    for (int i = 0; i < pn; ++i) {
        double d = hw * nodes[pn + i];
        t3.quad_x[i] = d;
        t3.quad_w[i] = 2.0 * hw * weights[pn + i] * kernel.evaluate(d);
    }
    // Raw spread onto the outer grid: a type-1 plan whose fine grid is the
    // outer grid itself.
    outer = std::make_unique<Plan>();
    outer->type = 1;
    outer->dim = dim;
    outer->isign = isign;
    outer->batch = batch;
    outer->eps = eps;
    outer->sigma = sigma;
    outer->precision = precision;
    outer->mode_order = ModeOrder::Centered;
    outer->fine_given = true;
    outer->spread_only = true;
    outer->timing = false;
    for (int a = 0; a < dim; ++a) {
        outer->n_modes[a] = outer_n[a] / 2;
        outer->fine[a] = outer_n[a];
    }
    if (!outer->init_common()) return false;
    inner = std::make_unique<Plan>();
    inner->type = 2;
    inner->dim = dim;
    inner->isign = isign;
    inner->batch = batch;
    inner->eps = eps;
    inner->sigma = sigma;
    inner->precision = precision;
    inner->mode_order = ModeOrder::Centered;
    inner->timing = false;
    for (int a = 0; a < dim; ++a) inner->n_modes[a] = outer_n[a];
    if (!inner->init_common()) return false;
    for (int a = 0; a < dim; ++a) type3_outer[a] = outer_n[a];
    // (src/gpu_type3.rs:98-110: F32 grids the dense spreader supports)
    type3_small = precision == Precision::F32 && small_spread_supported(dim, outer_n, kernel.width);
    return true;
}

template <int D, typename T>
bool Plan::spread_type3_sources(int64_t active, int64_t M, const T *cp, bool set_points,
                                hipStream_t stream)
{
    if constexpr (D <= 2 && sizeof(T) == 4) {
        if (type3_small && M > 0 && outer->spread_method == SpreadMethod::Plain) {
            // many sources on a small outer grid: point groups in their
            // original order, partial grids added in group order
            if (!outer->ensure_grid(active)) return false;
            const int64_t side[kMaxDims] = {1, 1, 1, 1, 1, 1, 1, 1};
            return small.spread(outer->grid_params(side), M, (const float *)t3_xr,
                                (const float *)cp, active, (float *)outer->grid, stream);
        }
    }
    if (set_points && !outer->set_points_t<D, T>(M, (const T *)t3_xr, stream)) return false;
    if (M == 0) {
        if (!outer->ensure_grid(active)) return false;
        return fill_zero<T>((T *)outer->grid, 2 * outer->fine_total * active, stream);
    }
    return outer->execute_t<D, T>(active, cp, nullptr, stream, false);
}

template <int D, typename T>
bool Plan::execute_type3_t(int64_t active, int64_t M, const T *x, const T *c, int64_t N,
                           const T *s, T *out, hipStream_t stream)
{
    if (!point_count_ok(M, "source") || !point_count_ok(N, "target")) return false;
    reset_marks();
    if (N == 0) return true;
    size_t rb = real_bytes();
    auto grow = [&](void **p, size_t *cap, size_t need, const char *what) {
        if (need <= *cap) return true;
        if (*p) (void)hipFree(*p);
        *p = nullptr;
        *cap = 0;
        if (!hip_ok(hipMalloc(p, need), what)) return false;
        *cap = need;
        return true;
    };
    if (!grow(&t3_xr, &t3_xr_bytes, std::max<int64_t>(M, 1) * D * rb, "type-3 sources") ||
        !grow(&t3_cp, &t3_cp_bytes, std::max<int64_t>(M, 1) * active * 2 * rb, "type-3 strengths") ||
        !grow(&t3_sr, &t3_sr_bytes, N * D * rb, "type-3 targets") ||
        !grow(&t3_factor, &t3_factor_bytes, N * 2 * sizeof(double), "type-3 factors") ||
        !grow(&t3_interp, &t3_interp_bytes, N * active * 2 * rb, "type-3 interpolated"))
        return false;
    mark(kStageType3, false, stream);
    if (t3_boxes.count == 1) {
        // one box: rescale + pre-phase, raw spread onto the outer grid,
        // target rescale + correction, inner type 2, final multiply
        // (src/gpu_type3.rs:267-430)
        if (!type3_source<D, T>(t3, M, active, x, c, (T *)t3_xr, (T *)t3_cp, stream)) return false;
        if (!spread_type3_sources<D, T>(active, M, (const T *)t3_cp, true, stream)) return false;
        if (!type3_target<D, T>(t3, N, s, (T *)t3_sr, (double *)t3_factor, stream)) return false;
        if (!inner->set_points_t<D, T>(N, (const T *)t3_sr, stream)) return false;
        if (!inner->execute_t<D, T>(active, (const T *)outer->grid, (T *)t3_interp, stream, false))
            return false;
        if (!type3_final<T>(N, active, (const T *)t3_interp, (const double *)t3_factor, out,
                            stream))
            return false;
        mark(kStageType3, true, stream);
        return true;
    }
    // Target boxes: the targets are grouped by box (counted on the device;
    // the counts are read back once to size each box's inner problem), and
    // every box runs the pipeline above with its centre as target centre.
    // The sources' rescaled positions and their binning on the outer grid
    // are shared by every box; only the pre-phase of the strengths differs.
    const int64_t nbox = t3_boxes.count;
    if (!grow(&t3_box, &t3_box_bytes, N * sizeof(uint32_t), "type-3 target boxes") ||
        !grow(&t3_perm, &t3_perm_bytes, N * sizeof(uint32_t), "type-3 target order") ||
        !grow(&t3_s_sorted, &t3_s_sorted_bytes, N * D * rb, "type-3 grouped targets") ||
        !grow(&t3_counts, &t3_counts_bytes, 3 * nbox * sizeof(uint32_t), "type-3 box counts"))
        return false;
    uint32_t *counts = (uint32_t *)t3_counts, *offsets = counts + nbox, *cursor = counts + 2 * nbox;
    if (!type3_box_count<D, T>(t3_boxes, N, s, (uint32_t *)t3_box, counts, stream)) return false;
    t3_host_counts.assign(2 * nbox, 0u);
    if (!hip_ok(hipMemcpyAsync(t3_host_counts.data(), counts, nbox * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, stream),
                "type-3 box counts read") ||
        !hip_ok(hipStreamSynchronize(stream), "type-3 box counts wait"))
        return false;
    uint32_t run = 0;
    for (int64_t b = 0; b < nbox; ++b) {
        t3_host_counts[nbox + b] = run;
        run += t3_host_counts[b];
    }
    if (!hip_ok(hipMemcpy(offsets, t3_host_counts.data() + nbox, nbox * sizeof(uint32_t),
                          hipMemcpyHostToDevice),
                "type-3 box offsets"))
        return false;
    if (!type3_box_scatter<D, T>(t3_boxes, N, s, (const uint32_t *)t3_box, offsets, cursor,
                                 (uint32_t *)t3_perm, (T *)t3_s_sorted, stream))
        return false;
    Type3Params tp = t3;
    bool first = true;
    for (int64_t b = 0; b < nbox; ++b) {
        const int64_t nb = t3_host_counts[b], off = t3_host_counts[nbox + b];
        if (nb == 0) continue;
        int64_t rest = b;
        for (int a = 0; a < D; ++a) {
            const int64_t i = rest % t3_boxes.k[a];
            rest /= t3_boxes.k[a];
            tp.tgt_center[a] = t3_boxes.k[a] == 1
                                   ? t3_center_full[a]
                                   : t3_boxes.lo[a] + ((double)i + 0.5) * t3_boxes.width[a];
        }
        if (!type3_source<D, T>(tp, M, active, x, c, first ? (T *)t3_xr : nullptr, (T *)t3_cp,
                                stream))
            return false;
        // later boxes spread the same records: the outer work plan is kept
        outer->spreader.reuse_plan = !first;
        if (!spread_type3_sources<D, T>(active, M, (const T *)t3_cp, first, stream)) return false;
        outer->spreader.reuse_plan = false;
        first = false;
        const T *sb = (const T *)t3_s_sorted + off * D;
        if (!type3_target<D, T>(tp, nb, sb, (T *)t3_sr, (double *)t3_factor, stream))
            return false;
        if (!inner->set_points_t<D, T>(nb, (const T *)t3_sr, stream)) return false;
        if (!inner->execute_t<D, T>(active, (const T *)outer->grid, (T *)t3_interp, stream, false))
            return false;
        if (!type3_final<T>(nb, active, (const T *)t3_interp, (const double *)t3_factor, out,
                            stream, (const uint32_t *)t3_perm + off, N))
            return false;
    }
    mark(kStageType3, true, stream);
    return true;
}

// ----------------------------------------------------------------------
// Dispatch on rank and precision.
// ----------------------------------------------------------------------
#define MNUFFT_DISPATCH(PLAN, CALL_F, CALL_D)                                                     \
    do {                                                                                          \
        bool f32 = (PLAN).precision == Precision::F32;                                            \
        switch ((PLAN).dim) {                                                                     \
        case 1: return f32 ? CALL_F(1, float) : CALL_D(1, double);                                \
        case 2: return f32 ? CALL_F(2, float) : CALL_D(2, double);                                \
        case 3: return f32 ? CALL_F(3, float) : CALL_D(3, double);                                \
        case 4: return f32 ? CALL_F(4, float) : CALL_D(4, double);                                \
        case 5: return f32 ? CALL_F(5, float) : CALL_D(5, double);                                \
        case 6: return f32 ? CALL_F(6, float) : CALL_D(6, double);                                \
        case 7: return f32 ? CALL_F(7, float) : CALL_D(7, double);                                \
        case 8: return f32 ? CALL_F(8, float) : CALL_D(8, double);                                \
        default: set_error("GPU plans take 1 to 8 dimensions"); return false;                     \
        }                                                                                         \
    } while (0)

bool Plan::set_points(int64_t M, const void *pts, hipStream_t stream)
{
#define SP(D, T) set_points_t<D, T>(M, (const T *)pts, stream)
    MNUFFT_DISPATCH(*this, SP, SP);
#undef SP
}

bool Plan::execute(int64_t active, const void *in, void *out, hipStream_t stream)
{
#define EX(D, T) execute_t<D, T>(active, (const T *)in, (T *)out, stream, true)
    MNUFFT_DISPATCH(*this, EX, EX);
#undef EX
}

bool Plan::bin_sort(int64_t M, const void *pts, uint32_t *sorted, uint32_t *offsets,
                    hipStream_t stream)
{
#define BS(D, T) bin_sort_t<D, T>(M, (const T *)pts, sorted, offsets, stream)
    MNUFFT_DISPATCH(*this, BS, BS);
#undef BS
}

bool Plan::bin_sort_static(int64_t M, const void *pts, uint32_t *sorted, uint32_t *offsets,
                           hipStream_t stream)
{
#define BS(D, T) bin_sort_static_t<D, T>(M, (const T *)pts, sorted, offsets, stream)
    MNUFFT_DISPATCH(*this, BS, BS);
#undef BS
}

bool Plan::execute_type3(int64_t active, int64_t M, const void *x, const void *c, int64_t N,
                         const void *s, void *out, hipStream_t stream)
{
#define T3(D, T) execute_type3_t<D, T>(active, M, (const T *)x, (const T *)c, N, (const T *)s, \
                                       (T *)out, stream)
    MNUFFT_DISPATCH(*this, T3, T3);
#undef T3
}

bool Plan::stage_times(double *ms, int n, int64_t *calls)
{
    for (int i = 0; i < n; ++i) ms[i] = -1.0;
    int64_t filled = std::min<int64_t>(slots_filled, (int64_t)slots.size());
    for (int64_t sl = 0; sl < filled; ++sl) {
        for (int st = 0; st < kStageCount && st < n; ++st) {
            if (!slots[sl].used[st]) continue;
            if (!hip_ok(hipEventSynchronize(slots[sl].ev[2 * st + 1]), "stage event")) return false;
            float t = 0.0f;
            if (hipEventElapsedTime(&t, slots[sl].ev[2 * st], slots[sl].ev[2 * st + 1]) ==
                hipSuccess)
                ms[st] = (ms[st] < 0 ? 0.0 : ms[st]) + t;
        }
    }
    if (calls) *calls = filled;
    if (history) slots_filled = 0;
    return true;
}

// ----------------------------------------------------------------------
// Many type-2 plans as one batch.
// ----------------------------------------------------------------------
namespace {

// Problem tables are written by kernels whose arguments carry up to
// TableChunk<E>::N entries (kernel arguments are copied at launch, so the
// sequence can be captured in a graph), within 4 KiB of arguments.
template <typename E>
struct TableChunk {
    static constexpr int N = (int)std::min<size_t>(64, (4096 - 64) / sizeof(E));
    E p[N];
    int n;
};

template <typename E>
__global__ void k_batch_table(TableChunk<E> chunk, E *dst)
{
    if ((int)threadIdx.x < chunk.n) dst[threadIdx.x] = chunk.p[threadIdx.x];
}

// dst[0..n) = v on the device, *cap grown (allocation) when n exceeds it.
template <typename E>
bool upload_table(const std::vector<E> &v, E **dst, int64_t *cap, hipStream_t stream)
{
    static_assert(TableChunk<E>::N >= 1 && sizeof(TableChunk<E>) <= 4096, "table chunk");
    const int64_t n = (int64_t)v.size();
    if (n > *cap) {
        if (*dst) (void)hipFree(*dst);
        *dst = nullptr;
        *cap = 0;
        if (!hip_ok(hipMalloc(dst, sizeof(E) * n), "batch table")) return false;
        *cap = n;
    }
    for (int64_t i0 = 0; i0 < n; i0 += TableChunk<E>::N) {
        TableChunk<E> chunk;
        chunk.n = (int)std::min<int64_t>(TableChunk<E>::N, n - i0);
        for (int k = 0; k < chunk.n; ++k) chunk.p[k] = v[(size_t)(i0 + k)];
        k_batch_table<E><<<1, 64, 0, stream>>>(chunk, *dst + i0);
    }
    return hip_ok(hipGetLastError(), "batch table launch");
}

// FFT of one vector of each plan in `ps`: on the line FFT one launch per
// axis pass and transform sign over all of them, else plan by plan.
bool fft_batch(Plan &head, const std::vector<Plan *> &ps, hipStream_t stream)
{
    bool lines = !ps.empty();
    for (Plan *p : ps) lines = lines && p->use_line_fft();
    if (!lines) {
        for (Plan *p : ps)
            if (!p->fft(1, p->grid, stream)) return false;
        return true;
    }
    std::vector<const LineFft *> f;
    std::vector<void *> g;
    for (Plan *p : ps) {
        if (!p->ensure_fft(1)) return false;
        f.push_back(p->line_fft.get());
        g.push_back(p->grid);
    }
    return LineFft::run_many(f, g, &head.line_jobs, &head.line_jobs_cap, stream);
}

template <int D>
bool execute_type2_many_t(int64_t count, Plan *const *plans, const int64_t *M,
                          const void *const *points, const void *const *inputs,
                          void *const *outputs, hipStream_t stream)
{
    Plan &head = *plans[0];
    // problems whose own execute uses the 2D tiled interpolation first,
    // grouped by kernel width (one launch per width), then the others (the
    // gather interpolation, one launch): every problem gets the kernel its
    // own encode uses
    std::vector<int64_t> order((size_t)count);
    for (int64_t i = 0; i < count; ++i) order[(size_t)i] = i;
    auto tiled_w = [&](int64_t i) {
        const bool tiled = (D == 2 && plans[i]->tiled2_interp()) || (D == 3 && plans[i]->tiled_interp());
        return tiled ? plans[i]->kernel.width : 1000;
    };
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return tiled_w(a) < tiled_w(b); });
    int maxw = 0;
    int64_t first_gather = count;
    for (int64_t r = 0; r < count; ++r) {
        const int64_t i = order[(size_t)r];
        if (tiled_w(i) == 1000) {
            first_gather = std::min(first_gather, r);
            maxw = std::max(maxw, plans[i]->kernel.width);
        }
    }
    const int G = maxw <= 4 ? 4 : (maxw <= 8 ? 8 : 16);
    const int64_t per_item = 256 / G;
    std::vector<BatchProblem> table((size_t)count);
    std::vector<ModeProblem> mtable((size_t)count);
    std::vector<int> widths;
    std::vector<int64_t> bin_lo, bin_hi, class_points;
    int64_t m_off = 0, items = 0, gather_items = 0, cells = 0;
    uint64_t bins = 0;
    for (int64_t r = 0; r < count; ++r) {
        const int64_t i = order[(size_t)r];
        Plan &p = *plans[i];
        p.points_set = false;
        p.points = nullptr;
        p.point_count = 0;
        if (!p.ensure_grid(1) || !p.ensure_fft(1)) return false;
        ModeProblem &mq = mtable[(size_t)r];
        mq.mp = p.mode_params();
        mq.inv_phihat = p.d_inv_phihat;
        mq.mode_base = 0;
        mq.fine_base = cells;
        mq.zero = 0;
        cells += p.fine_total;
        BatchProblem &b = table[(size_t)r];
        b.coords = points[i];
        b.out = outputs[i];
        b.grid = p.grid;
        b.in = inputs[i];
        b.m_off = m_off;
        b.M = M[i];
        b.item_base = items;
        b.bin_base = bins;
        b.g = p.type2_geometry();
        if (r < first_gather) {
            const int w = p.kernel.width;
            if (widths.empty() || widths.back() != w) {
                widths.push_back(w);
                bin_lo.push_back((int64_t)bins);
                bin_hi.push_back((int64_t)bins);
                class_points.push_back(0);
            }
            bin_hi.back() = (int64_t)(bins + (uint64_t)b.g.nbins);
            class_points.back() += M[i];
        }
        if (r == first_gather) gather_items = items;
        m_off += M[i];
        items += (M[i] + per_item - 1) / per_item;
        bins += (uint64_t)b.g.nbins;
    }
    if (first_gather == count) gather_items = items;
    if (m_off > 0xffffffffll) {
        set_error("batched type 2: more than 2^32 points over the plans");
        return false;
    }
    if (!upload_table(table, &head.batch_table, &head.batch_table_cap, stream) ||
        !upload_table(mtable, &head.batch_modes, &head.batch_modes_cap, stream))
        return false;
    // every plan's grid from its modes in one launch, then its FFT
    if (!predeconvolve_many<D, float>(head.batch_table, head.batch_modes, (int)count, cells,
                                      stream))
        return false;
    {
        std::vector<Plan *> ps;
        for (int64_t r = 0; r < count; ++r) ps.push_back(plans[order[(size_t)r]]);
        if (!fft_batch(head, ps, stream)) return false;
    }
    if (m_off == 0) return true;
    if (!head.batch_sorter) head.batch_sorter = std::make_unique<BinSorter>();
    BinSorter &bs = *head.batch_sorter;
    if (!bs.sort_many<D, float>(m_off, head.batch_table, (int)count, bins, stream)) return false;
    if (!widths.empty()) {
        if constexpr (D == 3) {
            // work-item offsets of the combined bins and their scan scratch
            const int64_t nb = (int64_t)bins;
            if (!head.ensure_items(nb + 1 + scan_tiles(nb + 1) + 1) ||
                !interpolate_tiled3_many(bs.sorted(), bs.offsets(), head.item_off, nb,
                                         head.batch_table, (int)count, widths.data(),
                                         bin_lo.data(), bin_hi.data(), class_points.data(),
                                         (int)widths.size(), stream))
                return false;
        } else if (!interpolate_tiled2_many(bs.sorted(), bs.offsets(), head.batch_table,
                                            (int)count, widths.data(), bin_lo.data(),
                                            bin_hi.data(), (int)widths.size(), stream)) {
            return false;
        }
    }
    return interpolate_many<D, float>(bs.sorted(), head.batch_table, (int)count, gather_items,
                                      items, std::max(maxw, 2), stream);
}

}  // namespace

bool execute_type2_many(int64_t count, Plan *const *plans, const int64_t *M,
                        const void *const *points, const void *const *inputs,
                        void *const *outputs, hipStream_t stream)
{
    if (count < 1 || !plans || !M || !points || !inputs || !outputs) {
        set_error("batched type 2: invalid arguments");
        return false;
    }
    const int dim = plans[0]->dim;
    for (int64_t i = 0; i < count; ++i) {
        const Plan *p = plans[i];
        if (!p || p->type != 2 || p->dim != dim || p->precision != Precision::F32 || dim < 1 ||
            dim > 3) {
            set_error("batched type 2 takes F32 type-2 plans of one dimension, 1D to 3D");
            return false;
        }
        if (M[i] < 0 || (M[i] > 0 && (!points[i] || !outputs[i])) || !inputs[i]) {
            set_error("batched type 2: invalid point count or null buffer");
            return false;
        }
        for (int64_t k = 0; k < i; ++k)
            if (plans[k] == p) {
                set_error("batched type 2: a plan appears twice");
                return false;
            }
    }
    if (dim == 1) return execute_type2_many_t<1>(count, plans, M, points, inputs, outputs, stream);
    if (dim == 2) return execute_type2_many_t<2>(count, plans, M, points, inputs, outputs, stream);
    return execute_type2_many_t<3>(count, plans, M, points, inputs, outputs, stream);
}

namespace {

template <int D, typename T>
bool execute_type1_many_t(int64_t count, Plan *const *plans, const int64_t *M,
                          const void *const *points, const void *const *inputs,
                          void *const *outputs, hipStream_t stream)
{
    Plan &head = *plans[0];
    // problems grouped by run length (one spread launch per run; the row
    // tiles, run 0, one launch per kernel width), in their order otherwise
    auto class_key = [&](int64_t i) {
        const int run = plans[i]->spread_params().run;
        return run == 0 ? plans[i]->kernel.width : 64 * run;
    };
    std::vector<int64_t> order((size_t)count);
    for (int64_t i = 0; i < count; ++i) order[(size_t)i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return class_key(a) < class_key(b); });
    std::vector<BatchProblem> table((size_t)count);
    std::vector<SpreadProblem> stable((size_t)count);
    std::vector<ModeProblem> mtable((size_t)count);
    std::vector<SpreadClass> classes;
    int64_t m_off = 0, modes = 0;
    uint64_t bins = 0;
    for (int64_t r = 0; r < count; ++r) {
        const int64_t i = order[(size_t)r];
        Plan &p = *plans[i];
        p.points_set = false;
        p.points = nullptr;
        p.point_count = 0;
        if (!p.ensure_grid(1) || !p.ensure_fft(1)) return false;
        const SpreadParams &sp = p.spread_params();
        BatchProblem &b = table[(size_t)r];
        b.coords = points[i];
        b.out = outputs[i];
        b.grid = p.grid;
        b.in = inputs[i];
        b.m_off = m_off;
        b.M = M[i];
        b.item_base = 0;
        b.bin_base = bins;
        b.g = p.spread_geometry(M[i], nullptr);
        if (classes.empty() || classes.back().run != sp.run ||
            (sp.run == 0 && classes.back().width != p.kernel.width))
            classes.push_back(SpreadClass{(int)r, (int)r, sp.run, p.kernel.width, 0, 0});
        SpreadClass &k = classes.back();
        SpreadProblem &q = stable[(size_t)r];
        q.sp = sp;
        q.block_base = k.blocks;
        const int64_t cells =
            sp.run == 0 ? sp.block[0] * sp.block[1] * sp.block[2] : (int64_t)64 * sp.run;
        const int64_t per_part = cells * 2 * (int64_t)sizeof(T);
        q.part_capacity = std::max<int64_t>(p.spreader.budget_bytes / per_part, 64);
        q.light_limit = p.spreader.light_limit;
        q.min_part = p.spreader.min_part;
        // partial slots the problem's heavy parts can take: at most its
        // capacity, and at most a part per heavy block plus its records in
        // reach (each record counted by at most `reach` blocks) over the
        // smallest part, with a factor two of margin
        double reach = 1.0;
        for (int a = 0; a < D; ++a) {
            const int64_t B = sp.block[a], n = p.fine[a];
            const int64_t side = (int64_t)1 << b.g.ax[a].shift;
            const int64_t L = std::min<int64_t>(B + p.kernel.width - 1, n);
            const int64_t all = sp.nblocks[a];
            reach *= (double)(L + side > n
                                  ? all
                                  : std::min<int64_t>(all, (side + p.kernel.width + 2 * B - 3) / B + 1));
        }
        const double bound = (double)sp.block_count +
                             2.0 * reach * (double)M[i] / (double)std::max<uint32_t>(q.min_part, 1);
        k.slots += std::min<int64_t>(q.part_capacity, (int64_t)std::min(bound, 9.0e15));
        k.blocks += sp.block_count;
        k.p1 = (int)r + 1;
        ModeProblem &mq = mtable[(size_t)r];
        mq.mp = p.mode_params();
        mq.inv_phihat = p.d_inv_phihat;
        mq.mode_base = modes;
        mq.fine_base = 0;
        mq.zero = M[i] == 0 ? 1 : 0;
        m_off += M[i];
        bins += (uint64_t)b.g.nbins;
        modes += p.mode_total;
    }
    if (m_off > 0xffffffffll) {
        set_error("batched type 1: more than 2^32 points over the plans");
        return false;
    }
    if (!upload_table(table, &head.batch_table, &head.batch_table_cap, stream) ||
        !upload_table(stable, &head.batch_spread, &head.batch_spread_cap, stream) ||
        !upload_table(mtable, &head.batch_modes, &head.batch_modes_cap, stream))
        return false;
    if (m_off > 0) {
        if (!head.batch_sorter) head.batch_sorter = std::make_unique<BinSorter>();
        BinSorter &bs = *head.batch_sorter;
        if (!bs.sort_many<D, T>(m_off, head.batch_table, (int)count, bins, stream) ||
            !head.ensure_records(m_off) || !head.ensure_strengths(m_off, 1) ||
            !prepare_many<D, T>(m_off, bs.sorted(), head.batch_table, (int)count,
                                head.rec_start, (T *)head.rec_off, head.rec_perm, stream) ||
            !gather_strengths_many<T>(m_off, head.rec_perm, head.batch_table, (int)count,
                                      (T *)head.c_sorted, stream) ||
            !head.spreader.spread_many<D, T>(head.batch_table, head.batch_spread, classes,
                                             bs.offsets(), head.rec_start,
                                             (const T *)head.rec_off,
                                             (const T *)head.c_sorted, stream))
            return false;
        std::vector<Plan *> ps;
        for (int64_t r = 0; r < count; ++r)
            if (M[order[(size_t)r]] > 0) ps.push_back(plans[order[(size_t)r]]);
        if (!fft_batch(head, ps, stream)) return false;
    }
    return deconvolve_many<D, T>(head.batch_table, head.batch_modes, (int)count, modes, stream);
}

}  // namespace

bool type1_many_valid(int64_t count, Plan *const *plans, const int64_t *M,
                      const void *const *points, const void *const *inputs,
                      void *const *outputs)
{
    if (count < 1 || !plans || !M || !points || !inputs || !outputs) {
        set_error("batched type 1: invalid arguments");
        return false;
    }
    const int dim = plans[0] ? plans[0]->dim : 0;
    const Precision prec = plans[0] ? plans[0]->precision : Precision::F32;
    for (int64_t i = 0; i < count; ++i) {
        const Plan *p = plans[i];
        if (!p || p->type != 1 || p->dim != dim || p->precision != prec || dim < 1 || dim > 3) {
            set_error("batched type 1 takes type-1 plans of one rank (1D to 3D) and precision");
            return false;
        }
        if (p->spread_method != SpreadMethod::Plain || p->point_order != BinOrder::Index) {
            set_error("batched type 1 takes plans of the plain spread (output-stationary "
                      "blocks or row tiles) in index order");
            return false;
        }
        if (M[i] < 0 || (M[i] > 0 && !points[i]) || (M[i] > 0 && !inputs[i]) || !outputs[i]) {
            set_error("batched type 1: invalid point count or null buffer");
            return false;
        }
        for (int64_t k = 0; k < i; ++k)
            if (plans[k] == p) {
                set_error("batched type 1: a plan appears twice");
                return false;
            }
    }
    return true;
}

bool execute_type1_many(int64_t count, Plan *const *plans, const int64_t *M,
                        const void *const *points, const void *const *inputs,
                        void *const *outputs, hipStream_t stream)
{
    if (!type1_many_valid(count, plans, M, points, inputs, outputs)) return false;
    const int dim = plans[0]->dim;
    const bool f32 = plans[0]->precision == Precision::F32;
    switch (dim) {
    case 1:
        return f32 ? execute_type1_many_t<1, float>(count, plans, M, points, inputs, outputs, stream)
                   : execute_type1_many_t<1, double>(count, plans, M, points, inputs, outputs, stream);
    case 2:
        return f32 ? execute_type1_many_t<2, float>(count, plans, M, points, inputs, outputs, stream)
                   : execute_type1_many_t<2, double>(count, plans, M, points, inputs, outputs, stream);
    default:
        return f32 ? execute_type1_many_t<3, float>(count, plans, M, points, inputs, outputs, stream)
                   : execute_type1_many_t<3, double>(count, plans, M, points, inputs, outputs, stream);
    }
}

}  // namespace mnufft
