// Deterministic output-stationary type-1 spreader.
#pragma once

#include <vector>

#include "nufft_internal.hpp"

namespace mnufft {

// Dense bins of the 3D matrix-core spread (spread_dense.hip).  classify()
// runs on sorted points: a bin of at least `threshold` records (the threshold
// doubled, up to seven times, until the bins' window parts fit part_cap) is
// dense, provided the dense bins then hold at least 1/8 of the records.  Its records are then kept after the light records (prepare() with
// `before`), the tile kernels see the light records only (offsets `light`),
// and spread() sums each dense bin's window of (side + w - 1) cells per axis
// in parts of 1024 records, adds the parts in order and adds the windows to
// the cells of the touched tiles in a fixed bin order.
struct DenseBins {
    ~DenseBins();
    void release();
    // Column blocks of 8 x cells of the bins' windows (1 or 2), or 0 when
    // the grid does not qualify (3D, bins dividing every axis, windows
    // within 16 x 8 x 8 cells).
    static int x_blocks(const GridParams &g);
    static int64_t window_floats(const GridParams &g) { return 1024 * (int64_t)x_blocks(g); }
    // (sp: the spread's tiles, whose list the dense windows reach is kept)
    bool classify(const GridParams &g, const SpreadParams &sp, const uint32_t *offsets,
                  hipStream_t stream);
    // add == false: the windows of the dense bins (their parts summed, then
    // added in part order), before the tile kernels; add == true: the
    // windows added to the cells of the tiles they reach, after them.  The
    // dense records' strengths are read from the caller's vectors (point
    // order) through the original index of each record.
    bool spread(SpreadMethod method, const GridParams &g, const SpreadParams &sp, int64_t M,
                const int4 *rec_start, const float *rec_off, const float *c,
                const uint32_t *cmax, int64_t vectors, float *grid, bool add,
                hipStream_t stream);

    uint32_t threshold = 32;   // records of the smallest dense bin (0: no dense bins)
    const float *strengths = nullptr;  // the execution's strength vectors, point order
    int64_t part_cap = 16384;  // window parts (1024 records each) at most
    uint32_t *before = nullptr;  // dense records before each bin (nbins + 1)
    uint32_t *light = nullptr;   // light-record offsets (nbins + 1)
    uint32_t *info = nullptr;    // 1 + first window part of a dense bin, 0 for a light bin
    uint2 *parts = nullptr;      // (bin, part) of every window part
    uint32_t *ctrl = nullptr;    // parts per threshold level, threshold, parts
    uint32_t *sums = nullptr;    // scan scratch
    float *win = nullptr;        // window parts, part_cap per vector
    uint32_t *tile_flag = nullptr;  // per tile: a dense window reaches it
    uint32_t *touched = nullptr;    // count, then those tiles (any order)
    int64_t bins_cap = 0, parts_alloc = 0, win_bytes = 0, tiles_cap = 0;
};

// Output-stationary blocks of one problem of a multi-plan type-1 spread:
// its blocks take indices [block_base, block_base + sp.block_count) of the
// combined block list of its class (problems of one run length).
// Its work plan uses its own plan's limits (light limit, smallest part,
// partial slots), so its heavy blocks split exactly as in its own spread.
struct SpreadProblem {
    SpreadParams sp;
    int64_t block_base;
    int64_t part_capacity;
    uint32_t light_limit, min_part;
};

// A class of consecutive problems [p0, p1) of one run length (or of the row
// tiles and one kernel width): its blocks
// and a bound on the partial slots its heavy parts take.
struct SpreadClass {
    int p0, p1;
    int run;    // 0: the 2D or 3D row tiles, all of kernel width `width`
    int width;
    int64_t blocks;
    int64_t slots;
};

// Device buffers of the spreader's work plan, reused across executions.
struct Spreader {
    ~Spreader();
    void release();
    bool ensure(int64_t nblocks, int64_t partial_need, int64_t vectors);

    // Spreads the prepared records (in bin order) of `vectors` strength
    // vectors onto `grid` (vectors * fine_total complex values).  Every cell
    // is written exactly once.
    // A matrix-core method needs FP32 and row tiles (run == 0).
    // With `dense`, offsets are the light-record offsets and the dense bins'
    // windows are added after the tile kernels.
    template <int D, typename T>
    bool spread(const GridParams &g, const SpreadParams &sp, const uint32_t *offsets,
                const int4 *rec_start, const T *rec_off, const T *c_sorted, int64_t M,
                int64_t vectors, T *grid, hipStream_t stream,
                SpreadMethod method = SpreadMethod::Plain, DenseBins *dense = nullptr);

    // Many problems (one vector each; table `probs` / `sprobs` on the
    // device, classes on the host): each class's blocks get one work plan
    // over all its problems, one spread launch and one heavy-part
    // reduction; every problem's grid is written cell by cell as its own
    // spread writes it.  offsets: the combined bin offsets of the batched
    // sort; records and strengths in combined slot order.
    template <int D, typename T>
    bool spread_many(const BatchProblem *probs, const SpreadProblem *sprobs,
                     const std::vector<SpreadClass> &classes, const uint32_t *offsets,
                     const int4 *rec_start, const T *rec_off, const T *c_sorted,
                     hipStream_t stream);

    // Optional events recorded right before the first and after the last
    // spread kernel launch.
    hipEvent_t kernel_start = nullptr, kernel_end = nullptr;

    // One-shot request: the next spread keeps the work plan of the previous
    // one (same records, offsets and parameters; only the strengths
    // differ), when the plan buffers still hold it.
    bool reuse_plan = false;
    uint64_t blocks_epoch = 0, planned_epoch = ~0ull;
    int64_t planned_blocks = -1, planned_capacity = -1;
    uint32_t planned_light = 0, planned_min = 0;
    const uint32_t *planned_offsets = nullptr;

    uint32_t *totals = nullptr;      // records in reach per block
    uint32_t *parts = nullptr;       // work items per block
    uint32_t *item_off = nullptr;    // first work item per block
    uint32_t *heavy_base = nullptr;  // first partial slot per heavy block
    uint32_t *heavy_list = nullptr;  // heavy blocks
    uint32_t *ctrl = nullptr;        // items, heavy parts, heavy blocks, part size
    uint32_t *vmax = nullptr;        // largest strength magnitude per vector (float bits)
    uint32_t *pctrl = nullptr;       // many problems: work-plan counters, 8 per problem
    int64_t *pbase = nullptr;        // many problems: item, slot, heavy-block bases
    int64_t problems_cap = 0;
    void *partial = nullptr;
    int64_t blocks_cap = 0;
    int64_t partial_bytes = 0;
    int64_t vmax_cap = 0;
    int64_t part_capacity = 0;
    // Records in reach above which a block is split into parts, and the
    // smallest part.
    uint32_t light_limit = 4096;
    uint32_t min_part = 1024;
    // Partial-sum slots of one vector, and the scratch a chunk of vectors
    // may take (the chunk is at least one vector).
    int64_t budget_bytes = (int64_t)256 << 20;
    int64_t chunk_budget_bytes = (int64_t)1 << 30;
};

// Work plan of the tiles (16 x 16 x 8 in 3D, 32^2 in 2D) handed to the matrix-core
// tile kernel.
struct TileWork {
    const uint32_t *totals, *parts, *item_off, *heavy_base, *ctrl, *cmax;
    unsigned items_grid;
};

// Tile kernel of the matrix-core spread (spread_mfma.hip); partial sums of
// heavy parts go to `partial` in the layout k_reduce_tiles reads.
bool launch_mfma_tiles(SpreadMethod method, int dim, const GridParams &g, const SpreadParams &sp,
                       const uint32_t *offsets, const int4 *rec_start, const float *rec_off,
                       const float *c_sorted, int64_t M, int64_t vectors, const TileWork &w,
                       float *grid, float *partial, hipStream_t stream);

// out[v] (cleared by the caller) receives the largest |re| or |im| of the n
// floats at c + v * n as float bits, for v < vectors.
bool launch_absmax(const float *c, int64_t n, int64_t vectors, uint32_t *out,
                   hipStream_t stream);

// Spread of many points onto a small F32 grid (the outer grid of a 1D or 2D
// type-3 plan): points cut into equal groups in their original order, each
// group's partial grid written by output-stationary lanes, the partial grids
// added in group order (spread_small.hip; reference
// src/gpu_dense_spread.rs:1-48).  Bitwise repeatable, no point sort.
constexpr int64_t kSmallSpreadMaxCells = 1 << 16;
constexpr int64_t kSmallGroupPoints = 256;
constexpr int64_t kSmallMaxGroups = 1024;

// 1D or 2D, every axis at least 2 * width long, at most
// kSmallSpreadMaxCells cells (src/gpu_dense_spread.rs:65-77).
bool small_spread_supported(int dim, const int64_t *fine, int width);
int64_t small_spread_groups(int64_t M);

struct SmallSpreader {
    ~SmallSpreader();
    // grid (vectors * fine_total complex values) = the spread of the M
    // points (point-major, F32) with strengths c (vectors * M complex).
    bool spread(const GridParams &g, int64_t M, const float *pts, const float *c,
                int64_t vectors, float *grid, hipStream_t stream);
    void *partial = nullptr;
    int64_t partial_bytes = 0;
    // partial grids of one chunk of vectors at most (one vector at least)
    int64_t budget_bytes = (int64_t)1 << 30;
};

// Output-stationary blocks for a fine grid: 64 lanes, `run` cells per lane
// along axis 0 (run + width - 1 <= n0 keeps every cell's support offset
// unique within a run).
SpreadParams make_spread_params(int d, const int64_t *fine, int width, bool f64);

// Tiles of the given cells per axis (run = 0 marks tiles).
SpreadParams make_tile_params(int d, const int64_t *fine, const int64_t *block);

// Widest kernel of the plain row tiles (FP64 plans; FP32 plans stop at 8).
constexpr int kMaxRowsWidth = 16;

// Plain row tiles of tx x 8 x 8 cells (tx = 8; 3D, width <= max_width, every
// axis at least 16 + width long, axis 0 at least 16 + width + 1, the
// matrix-core tiles' bound).  Returns false when the grid does not qualify.
bool make_sweep_params(int d, const int64_t *fine, int width, int64_t tx, int max_width,
                       SpreadParams *sp);

// Tile kernel of the plain row-tile spread (spread_rows.hip): 8 x 8 x 8
// tiles, partial sums of heavy parts in the layout k_reduce_tiles reads.
// The 2D row tiles (16 x 16 cells) of the problems of one batch class, all
// of kernel width `width` (one vector each); the work plan arrays and the
// item / slot bases as the batched output-stationary spread takes them.
template <typename T>
bool launch_rows2_many(int width, const BatchProblem *probs, const SpreadProblem *sprobs,
                       int nprob, unsigned items_grid, const uint32_t *offsets,
                       const int4 *rec_start, const T *rec_off, const T *c_sorted,
                       const uint32_t *totals, const uint32_t *parts, const uint32_t *item_off,
                       const uint32_t *heavy_base, const uint32_t *pctrl, const int64_t *pbase,
                       T *partial, hipStream_t stream);

// The same for the 3D row tiles (8 x 8 x 8 cells).
template <typename T>
bool launch_rows3_many(int width, const BatchProblem *probs, const SpreadProblem *sprobs,
                       int nprob, unsigned items_grid, const uint32_t *offsets,
                       const int4 *rec_start, const T *rec_off, const T *c_sorted,
                       const uint32_t *totals, const uint32_t *parts, const uint32_t *item_off,
                       const uint32_t *heavy_base, const uint32_t *pctrl, const int64_t *pbase,
                       T *partial, hipStream_t stream);

template <typename T>
bool launch_rows_tiles(const GridParams &g, const SpreadParams &sp, const uint32_t *offsets,
                       const int4 *rec_start, const T *rec_off, const T *c_sorted, int64_t M,
                       int64_t vectors, const TileWork &w, T *grid, T *partial,
                       hipStream_t stream);

// 32 x 32 tiles of the 2D matrix-core spread (width <= 8, every axis at least
// 32 + width long), run = 0.  Returns false when the grid does not qualify.
bool make_tile2_params(int d, const int64_t *fine, int width, SpreadParams *sp);

}  // namespace mnufft
