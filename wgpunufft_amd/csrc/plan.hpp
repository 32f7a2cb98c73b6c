// Device-resident plan objects behind the C ABI.
#pragma once

#include <map>
#include <memory>

#include "bins.hpp"
#include "modes.hpp"
#include "fft_lines.hpp"
#include "nufft_internal.hpp"
#include "spread.hpp"

namespace mnufft {

struct FftPlan {
    ~FftPlan();
    void *plan = nullptr;  // rocfft_plan
    void *info = nullptr;  // rocfft_execution_info
    void *work = nullptr;
};

struct Plan {
    ~Plan();

    // configuration (NufftConfig, src/config.rs:64-74)
    int type = 1;
    int dim = 1;
    int64_t n_modes[kMaxDims] = {1, 1, 1, 1, 1, 1, 1, 1};
    int isign = 1;
    int64_t batch = 1;
    double eps = 1e-6;
    double sigma = 2.0;
    Precision precision = Precision::F32;
    ModeOrder mode_order = ModeOrder::Centered;
    bool fine_given = false;   // fine[] set by the caller (type-3 outer grid)
    bool spread_only = false;  // stop after spreading (type-3 outer grid)
    bool timing = true;

    // planned quantities
    EsKernel kernel;
    int64_t fine[kMaxDims] = {1, 1, 1, 1, 1, 1, 1, 1};
    int64_t fine_total = 1, mode_total = 1;
    int64_t phihat_off[kMaxDims] = {};
    double *d_inv_phihat = nullptr;
    SpreadLayout layout;
    SpreadParams sp{};         // plain spread: the row tiles where `sweep`, else sp_generic
    SpreadParams sp_generic{};  // output-stationary blocks of the reference layout
    bool sweep = false;  // 3D grid qualifies for the row tiles (plain and matrix-core)
    // Tuning the autotuner sets (and mnufft_set_bin_tuning): the plain 2D/3D
    // spread on row tiles (else the output-stationary blocks), and the side
    // of the row geometry's bins along axis 0 (0: the smallest power of two
    // from 2 keeping the bin count within 2^24).
    bool plain_rows = true;
    int64_t rows_side0 = 0;
    int64_t binned_side0 = 0;  // axis-0 bin side of the current binning
    SpreadParams tile2_sp{};
    bool tiles2 = false;  // 2D grid qualifies for the matrix-core 32 x 32 tiles
    SpreadParams rows2_sp{};
    bool rows2 = false;   // 2D grid qualifies for the plain 16 x 16 row tiles
    SpreadParams tile3_sp{};  // 3D matrix-core tiles (16 x 16 x 8; row-sweep grids)
    SpreadMethod spread_method = SpreadMethod::Plain;
    // Type-1 order of the points inside a bin: ascending original index (the
    // reference's), or coordinate bits then index (shuffle invariant).
    BinOrder point_order = BinOrder::Index;
    BinOrder binned_order = BinOrder::Index;  // order of the current binning
    bool gp_rows = false;  // points binned in one-cell rows (row tiles)
    // Dense bins of the 3D matrix-core spread (records of dense bins kept
    // apart); binned_dense is the threshold of the current binning, 0 when
    // the binning has none.
    DenseBins dense;
    uint32_t binned_dense = 0;
    uint32_t dense_threshold(const GridParams &g) const
    {
        if (type != 1 || dim != 3 || !sweep || spread_method == SpreadMethod::Plain ||
            precision != Precision::F32 || dense.threshold == 0 || DenseBins::x_blocks(g) == 0)
            return 0;
        return dense.threshold;
    }
    // Type-1 spreads on row tiles: the 3D plain row tiles or matrix-core
    // tiles, or a 2D matrix-core method.
    bool rows_geometry() const
    {
        if (spread_method == SpreadMethod::Plain) return (sweep || rows2) && plain_rows;
        return sweep || tiles2;
    }
    const SpreadParams &spread_params() const
    {
        if (spread_method == SpreadMethod::Plain) {
            if (!plain_rows) return sp_generic;
            return sweep ? sp : (rows2 ? rows2_sp : sp_generic);
        }
        return dim == 2 ? tile2_sp : tile3_sp;
    }

    // points
    const void *points = nullptr;
    int64_t point_count = 0;
    bool points_set = false;
    GridParams gp{};
    int geometry_index = 0;

    // scratch
    BinSorter sorter;
    Spreader spreader;
    void *grid = nullptr;
    size_t grid_bytes = 0;
    int4 *rec_start = nullptr;
    void *rec_off = nullptr;
    uint32_t *rec_perm = nullptr;  // original index of every record slot (type 1)
    int64_t rec_cap = 0;
    void *c_sorted = nullptr;
    size_t c_sorted_bytes = 0;
    uint32_t *item_off = nullptr;  // work items of the tiled interpolation
    int64_t item_cap = 0;
    std::map<int64_t, std::unique_ptr<FftPlan>> ffts;
    // 3D grids: the pruned line-transform FFT (fft_engine 0, where the grid
    // qualifies) or rocFFT's 3D plan (fft_engine 1).
    int fft_engine = 0;
    std::unique_ptr<LineFft> line_fft;
    bool use_line_fft() const;
    // Ranks 4-8 (rocFFT plans take at most three axes): batched FFTs over
    // groups of the fastest axes, each followed by a transpose that moves
    // the transformed axes behind the others; plans keyed by {group rank,
    // lengths, batch}, and the transpose's second buffer.
    std::map<std::vector<int64_t>, std::unique_ptr<FftPlan>> fft_groups;
    void *fft_scratch = nullptr;
    size_t fft_scratch_bytes = 0;

    // type 3
    Type3Params t3{};
    int64_t type3_outer[kMaxDims] = {1, 1, 1, 1, 1, 1, 1, 1};
    std::unique_ptr<Plan> outer, inner;
    // F32 1D / 2D type-3 plans whose outer grid is small spread their
    // sources with the small-grid spreader (no point sort)
    bool type3_small = false;
    SmallSpreader small;
    // Target boxes (one box per axis unless the inner grid of the whole
    // target interval would exceed kType3InnerCells cells or
    // t3_inner_budget bytes): every box is a type-3 problem of its own with
    // the box centre as target centre, sharing the outer and inner plans.
    Type3Boxes t3_boxes{};
    double t3_center_full[kMaxDims] = {};  // target centre of an unsplit axis
    double t3_bounds[4 * kMaxDims] = {};   // source then target bounds
    int64_t t3_inner_budget = (int64_t)16 << 30;
    void *t3_box = nullptr, *t3_perm = nullptr, *t3_s_sorted = nullptr, *t3_counts = nullptr;
    size_t t3_box_bytes = 0, t3_perm_bytes = 0, t3_s_sorted_bytes = 0, t3_counts_bytes = 0;
    std::vector<uint32_t> t3_host_counts;
    void *t3_xr = nullptr, *t3_cp = nullptr, *t3_sr = nullptr, *t3_factor = nullptr,
         *t3_interp = nullptr;
    size_t t3_xr_bytes = 0, t3_cp_bytes = 0, t3_sr_bytes = 0, t3_factor_bytes = 0,
           t3_interp_bytes = 0;

    // stage timing: one slot of events per call, or (history mode) one per
    // call since the last read, so callers can sum stage durations over a
    // timed region without synchronizing inside it
    struct TimingSlot {
        hipEvent_t ev[2 * kStageCount] = {};
        bool used[kStageCount] = {};
    };
    std::vector<TimingSlot> slots;
    int64_t slot = 0;
    int64_t slots_filled = 0;
    bool history = false;
    bool begin_call();

    bool init_common();
    // min_boxes (may be null): at least this many target boxes per axis.
    bool init_type3(const double *src_bounds, const double *tgt_bounds,
                    const int *min_boxes = nullptr);
    GridParams grid_params(const int64_t *side) const;
    GridParams type1_geometry(int64_t M, int *chosen) const;
    GridParams spread_geometry(int64_t M, int *chosen) const;
    GridParams type2_geometry() const;
    ModeParams mode_params() const;
    size_t real_bytes() const;
    bool ensure_grid(int64_t vectors);
    bool ensure_records(int64_t M);
    bool ensure_strengths(int64_t M, int64_t vectors);
    bool ensure_items(int64_t n);
    bool tiled_interp() const;
    bool tiled2_interp() const;
    FftPlan *fft_for(int64_t vectors);
    bool ensure_fft(int64_t vectors);
    bool fft(int64_t vectors, void *data, hipStream_t stream);
    FftPlan *fft_group_plan(int rank, const int64_t *lengths, int64_t batch);
    // run = false: create the plans and the scratch only
    bool fft_nd(int64_t vectors, void *data, hipStream_t stream, bool run);
    void mark(int stage, bool end, hipStream_t stream);
    void reset_marks();

    bool set_points(int64_t M, const void *pts, hipStream_t stream);
    bool execute(int64_t active, const void *in, void *out, hipStream_t stream);
    // Scratch of a multi-plan type-2 execution led by this plan: the
    // combined bin sort and the device table of the problems.
    std::unique_ptr<BinSorter> batch_sorter;
    BatchProblem *batch_table = nullptr;
    int64_t batch_table_cap = 0;
    ModeProblem *batch_modes = nullptr;     // mode stages of the problems
    SpreadProblem *batch_spread = nullptr;  // spread blocks of the problems (type 1)
    int64_t batch_modes_cap = 0, batch_spread_cap = 0;
    LineFftJob *line_jobs = nullptr;  // batched line-FFT job table
    int64_t line_jobs_cap = 0;
    bool bin_sort(int64_t M, const void *pts, uint32_t *sorted, uint32_t *offsets,
                  hipStream_t stream);
    // The same permutation and offsets as bin_sort, with every intermediate
    // in the library's static sort scratch: enqueues kernels only (no
    // allocation, no timing marks).
    bool bin_sort_static(int64_t M, const void *pts, uint32_t *sorted, uint32_t *offsets,
                         hipStream_t stream);
    bool execute_type3(int64_t active, int64_t M, const void *x, const void *c, int64_t N,
                       const void *s, void *out, hipStream_t stream);
    bool stage_times(double *ms, int n, int64_t *calls);

    template <int D, typename T>
    bool set_points_t(int64_t M, const T *pts, hipStream_t stream);
    template <int D, typename T>
    bool execute_t(int64_t active, const T *in, T *out, hipStream_t stream, bool fresh_marks);
    template <int D, typename T>
    bool bin_sort_t(int64_t M, const T *pts, uint32_t *sorted, uint32_t *offsets,
                    hipStream_t stream);
    template <int D, typename T>
    bool bin_sort_static_t(int64_t M, const T *pts, uint32_t *sorted, uint32_t *offsets,
                           hipStream_t stream);
    // Spreads the type-3 sources (rescaled positions in t3_xr, pre-phased
    // strengths cp) onto the outer grid; set_points: bin them first.
    template <int D, typename T>
    bool spread_type3_sources(int64_t active, int64_t M, const T *cp, bool set_points,
                              hipStream_t stream);
    template <int D, typename T>
    bool execute_type3_t(int64_t active, int64_t M, const T *x, const T *c, int64_t N,
                         const T *s, T *out, hipStream_t stream);
};

// Type-2 encodes of `count` plans (1D or 2D, F32, one vector each) as one
// batch: plans[i] interpolates inputs[i] (modes) at points[i] (M[i] points)
// into outputs[i].  One bin sort over all the points and one interpolation
// launch; each plan predeconvolves and transforms its own grid.  Outputs
// equal the plans' own encodes bit for bit.  plans[0] keeps the batch
// scratch; the plans keep no points afterwards.
bool execute_type2_many(int64_t count, Plan *const *plans, const int64_t *M,
                        const void *const *points, const void *const *inputs,
                        void *const *outputs, hipStream_t stream);

// Type-1 encodes of `count` plans (one rank and precision, 1D to 3D, one
// vector each) as one batch: plans[i] spreads inputs[i] (strengths) at
// points[i] (M[i] points) and writes outputs[i] (modes).  One bin sort over
// all the points, one prepare, one strength gather, per run length one
// spread launch and one heavy-part reduction over every problem's blocks
// (each problem's work plan made with its own plan's limits), each plan's
// FFT, one deconvolution launch.  The plans must spread on the
// output-stationary blocks in index order (the plain spread; 3D plans with
// the row tiles switched off); outputs then equal the plans' own encodes
// bit for bit.  plans[0] keeps the batch scratch; the plans keep no points.
// The argument checks of execute_type1_many (last error set when false).
bool type1_many_valid(int64_t count, Plan *const *plans, const int64_t *M,
                      const void *const *points, const void *const *inputs,
                      void *const *outputs);
bool execute_type1_many(int64_t count, Plan *const *plans, const int64_t *M,
                        const void *const *points, const void *const *inputs,
                        void *const *outputs, hipStream_t stream);

}  // namespace mnufft
