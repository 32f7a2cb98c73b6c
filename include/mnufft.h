/*
 * mnufft: device-resident NUFFT plans on AMD Instinct MI355X (gfx950).
 *
 * This is the device-pointer face of the engine: the counterpart of the
 * reference's GPU plan objects, whose methods record work into a caller's
 * command encoder on caller-owned GPU buffers.  Here every call enqueues
 * work on a caller's HIP stream (NULL: the default stream) on caller-owned
 * device pointers, and returns once the work is enqueued.
 *
 * Conventions (as the reference's plan surface):
 *  - coordinates are point-major, points[j * dim + axis], radians in
 *    [-3 pi, 3 pi] for types 1 and 2 (periodic modulo 2 pi);
 *  - complex values are interleaved (re, im): float2 for F32 plans, double2
 *    otherwise;
 *  - modes are axis-0 fastest, centered (index k holds frequency
 *    k - floor(N/2)) or FFT order;
 *  - batches are transform-major: vector 0 whole, then vector 1, ...
 *
 * Every function returns 0 or one of the MNUFFT_ERROR_* codes and then
 * mnufft_last_error() describes the failure (thread-local, valid until the
 * next call on the thread).
 */
#ifndef MNUFFT_H
#define MNUFFT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MNUFFT_SUCCESS 0
#define MNUFFT_ERROR_INVALID_ARGUMENT 1
#define MNUFFT_ERROR_PLAN 2
#define MNUFFT_ERROR_GPU_UNAVAILABLE 3
#define MNUFFT_ERROR_GPU 4
#define MNUFFT_ERROR_INTERNAL 5

/* Arithmetic of a plan (same values as WGPU_NUFFT_PRECISION_*); DF64 runs
 * native f64 on MI355X. */
#define MNUFFT_PRECISION_F64 1
#define MNUFFT_PRECISION_DF64 2
#define MNUFFT_PRECISION_F32 3

/* Stages reported by mnufft_stage_times, in this order. */
#define MNUFFT_STAGE_SORT 0
#define MNUFFT_STAGE_PREPARE 1
#define MNUFFT_STAGE_SPREAD 2
#define MNUFFT_STAGE_FFT 3
#define MNUFFT_STAGE_DECONV 4
#define MNUFFT_STAGE_INTERP 5
#define MNUFFT_STAGE_TYPE3 6
#define MNUFFT_STAGE_SPREAD_KERNEL 7 /* the spread kernel alone, inside SPREAD */
#define MNUFFT_STAGE_COUNT 8

typedef struct mnufft_plan_s *mnufft_plan;

/*
 * A type-1 or type-2 plan for `batch` vectors over one point set.
 * Replaces NufftConfig (src/config.rs:64-230) with NufftPlan::type1_gpu /
 * NufftPlan::type2_gpu (src/plan.rs:457-463, 606-651): dim 1..8 axes
 * (MAX_GPU_NUFFT_DIMENSIONS, src/config.rs:7-12; ranks 4-8 take the
 * rank-generic spread, interpolation and an FFT composed of rocFFT
 * transforms over axis groups), n_modes per axis (axis 0 first), isign >= 0
 * positive, eps the tolerance, precision one of MNUFFT_PRECISION_*,
 * mode_order 0 centered / 1 FFT, sigma the upsampling factor (0: 2).
 */
int mnufft_makeplan(int32_t type, int32_t dim, const int64_t *n_modes, int32_t isign,
                    int64_t batch, double eps, int32_t precision, int32_t mode_order, double sigma,
                    mnufft_plan *plan);

/*
 * A type-3 plan: NufftType3Config + NufftType3Plan::new (src/type3.rs:90-344),
 * dim 1..8.  source_bounds / target_bounds hold (lower, upper) per axis; every point
 * and target must lie inside them (a caller guarantee, src/type3.rs:346-353).
 */
int mnufft_makeplan_type3(int32_t dim, int32_t isign, int64_t batch, double eps, int32_t precision,
                          double sigma, const double *source_bounds, const double *target_bounds,
                          mnufft_plan *plan);

/* Releases a plan and its device scratch; NULL is ignored. */
void mnufft_destroy(mnufft_plan plan);

/*
 * Bins M points for later executions: NufftPlan::set_points_gpu
 * (src/plan.rs:874-925).  The plan keeps the device pointer; the points must
 * stay unchanged while executions use them.
 */
int mnufft_set_points(mnufft_plan plan, int64_t M, const void *d_points, void *stream);

/*
 * Runs active_batch (1..batch) transforms on the points last set:
 * NufftPlan::execute_type1_gpu_batch / execute_type2_gpu_batch
 * (src/plan.rs:931-977).  Type 1 reads strengths (M x active_batch) and
 * writes modes; type 2 reads modes and writes values at the points.  Zero
 * points give an all-zero type-1 output.
 */
int mnufft_execute(mnufft_plan plan, int64_t active_batch, const void *d_in, void *d_out,
                   void *stream);

/* set_points + execute: NufftPlan::encode_type{1,2}_gpu_batch
 * (src/plan.rs:808-872, 1090-1177). */
int mnufft_encode(mnufft_plan plan, int64_t active_batch, int64_t M, const void *d_points,
                  const void *d_in, void *d_out, void *stream);

/*
 * Type-2 encodes of `count` independent plans as one batch (the batching
 * form of the reference's plan surface, src/plan.rs:606 type2_gpu per plan;
 * no single reference entry point does this): plans[i] interpolates its
 * modes d_in[i] at the M[i] points d_points[i] into d_out[i], as
 * mnufft_encode(plans[i], 1, M[i], d_points[i], d_in[i], d_out[i]) would,
 * bit for bit.  The points of all plans are binned by one sort and
 * interpolated by one launch; one launch predeconvolves every plan's grid,
 * and each plan transforms its own.  F32 type-2 plans of one dimension, 1D
 * to 3D, each at most once; one vector per plan.  The plans keep no points
 * afterwards (set them again before a plain execute); plans[0] keeps the
 * batch scratch.
 */
int mnufft_execute_type2_many(int64_t count, const mnufft_plan *plans, const int64_t *M,
                              const void *const *d_points, const void *const *d_in,
                              void *const *d_out, void *stream);

/*
 * Type-1 encodes of `count` independent plans as one batch (the reference
 * runs one type1_gpu plan per problem, src/plan.rs:457-463, 808-872; no
 * single reference entry point does this): plans[i] spreads its strengths
 * d_in[i] at the M[i] points d_points[i] and writes its modes d_out[i], as
 * mnufft_encode(plans[i], 1, M[i], d_points[i], d_in[i], d_out[i]) would,
 * bit for bit.  One bin sort over all the points, one prepare and strength
 * gather, one spread launch and one heavy-part reduction per class of plans
 * (the output-stationary blocks of one run length, or the 2D / 3D row
 * tiles of one kernel width; each plan's tiles split by its own work plan),
 * each plan's FFT, one deconvolution launch.  Type-1 plans of one rank (1D
 * to 3D) and precision, each at most once, one vector per plan, on the
 * plain spread in index point order; else
 * MNUFFT_ERROR_INVALID_ARGUMENT.  The plans keep no points afterwards;
 * plans[0] keeps the batch scratch.
 */
int mnufft_execute_type1_many(int64_t count, const mnufft_plan *plans, const int64_t *M,
                              const void *const *d_points, const void *const *d_in,
                              void *const *d_out, void *stream);

/*
 * Type 3: NufftType3Plan::encode_gpu_batch (src/type3.rs:473-539).  Reads M
 * sources and their strengths (M x active_batch), N targets, writes the N x
 * active_batch outputs.
 */
int mnufft_execute_type3(mnufft_plan plan, int64_t active_batch, int64_t M, const void *d_sources,
                         const void *d_strengths, int64_t N, const void *d_targets, void *d_out,
                         void *stream);

/*
 * Target boxes of a type-3 plan.  A target interval whose inner fine grid
 * would exceed 2^30 cells (or 16 GiB) is cut into boxes along its longest
 * axes; every box is a type-3 problem of its own with the box centre as
 * target centre (the same outer and inner grids for every box), so the
 * interval [-200, 200]^3 over sources in [-5, 5]^3 runs on one GPU.  No
 * counterpart in the reference, which plans the whole interval at once
 * (src/type3.rs:708-797).  min_boxes (may be NULL, dim values 1..64)
 * re-plans with at least that many boxes per axis; boxes (may be NULL, dim
 * values) receives the boxes per axis in use.
 */
int mnufft_type3_boxes(mnufft_plan plan, const int32_t *min_boxes, int32_t *boxes);

/*
 * The reference's stable bin permutation of M points for a type-1 plan:
 * the count / scan / scatter / large-bin sort sequence of
 * NdPointBins::encode with NdBinOrder::Stable (src/gpu_nd_bins.rs:336-471,
 * src/gpu_bin_sort.rs:1-56).  d_sorted receives M point indices (ascending
 * index inside each bin), d_offsets (may be NULL) the nbins + 1 bin offsets;
 * *nbins receives the bin count of the geometry chosen for M.
 */
int mnufft_bin_sort(mnufft_plan plan, int64_t M, const void *d_points, uint32_t *d_sorted,
                    uint32_t *d_offsets, int64_t *nbins, void *stream);

/*
 * mnufft_bin_sort without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same permutation and offsets as mnufft_bin_sort, bit for bit, but every
 * intermediate (keys, values, histograms, counters, gap lists) lives in one
 * static device array of the library (32 MiB per GPU the library runs on,
 * held until the process ends); d_offsets is required (nbins + 1 values).
 * At most MNUFFT_NOALLOC_MAX_POINTS points and MNUFFT_NOALLOC_MAX_BINS bins
 * (MNUFFT_ERROR_INVALID_ARGUMENT beyond).  Every call on one GPU shares that
 * array, so calls that overlap on one GPU (two streams, two threads) race;
 * calls in stream order are safe.
 */
#define MNUFFT_NOALLOC_MAX_POINTS 2097152
#define MNUFFT_NOALLOC_MAX_BINS 16777216
int mnufft_bin_sort_noalloc(mnufft_plan plan, int64_t M, const void *d_points,
                            uint32_t *d_sorted, uint32_t *d_offsets, int64_t *nbins,
                            void *stream);

/* Bin count of the geometry a type-1 plan chooses for M points (sizes the
 * offsets buffer of mnufft_bin_sort). */
int mnufft_bin_count(mnufft_plan plan, int64_t M, int64_t *nbins);

/*
 * The binning the executes of a type-1 plan use, as left by its last
 * set_points (or by an execute that re-binned the kept points): d_sorted
 * receives the M point indices in record-slot order (the order the spread
 * sums them in), d_offsets (may be NULL) the nbins + 1 bin offsets, bin_side
 * (may be NULL, max(3, dim) values) the bin side per axis.  For the index point order
 * this is the reference's stable bin permutation of that geometry (the
 * contract of mnufft_bin_sort).  Reads nothing but the plan's own scratch;
 * MNUFFT_ERROR_INVALID_ARGUMENT when no points are binned or dense bins
 * hold records apart.  Call mnufft_bin_count-sized buffers: *nbins receives
 * the bin count first.
 */
int mnufft_points_permutation(mnufft_plan plan, uint32_t *d_sorted, uint32_t *d_offsets,
                              int64_t *nbins, int64_t *bin_side, void *stream);

/*
 * Plan facts, for tests and benchmarks: info[0] kernel width, [1..3] fine
 * lengths, [4..6] bin sides of the geometry last chosen (type 1) or used
 * (type 2), [7] its bin count, [8] geometry index, [9..11] the tile or
 * block edge of the type-1 spread the plan's method and tuning select,
 * [12] its run (0: tiles), [13..15] type-3 outer grid lengths (axes 0-2 only for
 * plans of more axes).  Writes min(n, 16).
 */
int mnufft_plan_info(mnufft_plan plan, int64_t *info, int32_t n);

/* Kernel beta of the plan. */
double mnufft_plan_beta(mnufft_plan plan);

/*
 * Milliseconds per stage of the plan's last set_points / execute, measured
 * with HIP events on the stream of the call (-1 for stages that did not
 * run); waits for those events.  The stage indices are MNUFFT_STAGE_*.
 */
int mnufft_stage_times(mnufft_plan plan, double *ms, int32_t n);

/*
 * With on != 0, every later call of the plan records its stage events into a
 * fresh slot, and mnufft_stage_times returns the sums over all calls since
 * the previous read (then starts a new history): a caller can time many
 * executions without synchronizing between them.  on = 0 restores
 * last-call timing.
 */
int mnufft_timing_history(mnufft_plan plan, int32_t on);

/*
 * Arithmetic of the plan's type-1 spread (no counterpart in the reference,
 * whose spread is plain FP32/FP64 arithmetic): MNUFFT_SPREAD_PLAIN (the
 * default), or the 16 x 16 x 8-tile contraction on the matrix cores with FP16
 * operands (MNUFFT_SPREAD_MFMA_F16, about 1e-3 relative), FP16 hi/lo splits
 * (MNUFFT_SPREAD_MFMA_F16X3, about 1e-6) or bfloat16 hi/lo splits
 * (MNUFFT_SPREAD_MFMA_BF16X3, about 1e-5), accumulated in FP32.  The matrix
 * methods need an FP32 3D plan whose fine grid has at least 16 + width cells
 * per axis and a width <= 8; otherwise MNUFFT_ERROR_INVALID_ARGUMENT.  A
 * type-3 plan applies the method to its outer (source) grid.
 */
#define MNUFFT_SPREAD_PLAIN 0
#define MNUFFT_SPREAD_MFMA_F16 1
#define MNUFFT_SPREAD_MFMA_F16X3 2
#define MNUFFT_SPREAD_MFMA_BF16X3 3
int mnufft_set_spread_method(mnufft_plan plan, int32_t method);

/*
 * Order in which a type-1 plan sums the points of one bin.  The default,
 * MNUFFT_POINT_ORDER_INDEX, is ascending original index, the reference's
 * order (src/gpu_bin_sort.rs:1-7: consumers restore original point order):
 * results are bitwise repeatable and a batch equals its vectors one by one.
 * MNUFFT_POINT_ORDER_COORDINATES (no counterpart in the reference) sums in
 * ascending coordinate bits, then index, so the output also does not depend
 * on the order in which the points are given (bitwise), at the cost of an
 * in-bin sort in set_points.  A change after set_points re-bins the kept
 * points at the next execution.  Type-2 plans: MNUFFT_ERROR_INVALID_ARGUMENT.
 */
#define MNUFFT_POINT_ORDER_INDEX 0
#define MNUFFT_POINT_ORDER_COORDINATES 1
int mnufft_set_point_order(mnufft_plan plan, int32_t order);

/*
 * Work-plan knobs of the type-1 spread (no counterpart in the reference):
 * a tile whose reach holds more than light_limit records is split into
 * parts of at least min_part records, with partial sums in slots of at most
 * part_budget_bytes per vector (the light limit is raised until the heavy
 * tiles fit half the slots).  Zero keeps a knob's current value (defaults
 * 4096, 1024, 256 MiB).  Results stay bitwise repeatable for fixed knobs;
 * other knobs change the summation order of heavy tiles only.
 */
int mnufft_set_spread_tuning(mnufft_plan plan, int64_t light_limit, int64_t min_part,
                             int64_t part_budget_bytes);

/*
 * Bins and tiles of the type-1 spread (no counterpart in the reference):
 * rows_side0, the side along axis 0 of the bins of the row geometry (the
 * plain and matrix-core tiles; 0 = the smallest power of two from 2 keeping
 * the bin count within 2^24, else a power of two up to 64, never finer than
 * the automatic side); plain_rows, the plain spread on row tiles (3D: 8 x 8
 * x 8 cells, 2D: 16 x 16; 1, the default) or on the output-stationary
 * blocks of the reference's layout (0).  Applies from the next set_points (an execute
 * re-bins kept points); mnufft_autotune sets both.
 */
int mnufft_set_bin_tuning(mnufft_plan plan, int64_t rows_side0, int32_t plain_rows);

/*
 * FFT of the fine grid (no counterpart in the reference, whose FFT is its
 * wgpu-fft dependency): MNUFFT_FFT_LINES (default) runs 2D and 3D grids
 * whose axes are at most 4864 (FP32) / 2432 (FP64) cells long as
 * one pass of line transforms in LDS per axis, pruned to the lines that
 * carry modes (batches of plans: one launch per pass over all of them); MNUFFT_FFT_ROCFFT keeps rocFFT's plan.  Other grids always use
 * rocFFT.  A type-3 plan applies it to its inner type-2 plan.  Takes effect
 * at the next execute.
 */
#define MNUFFT_FFT_LINES 0
#define MNUFFT_FFT_ROCFFT 1
int mnufft_set_fft_engine(mnufft_plan plan, int32_t engine);

/* The engine the plan's next FFT runs on (MNUFFT_FFT_LINES or _ROCFFT). */
int mnufft_fft_engine(mnufft_plan plan, int32_t *engine);

/*
 * The plan's spread tuning: vals[0] spread method, [1] dense-bin threshold,
 * [2] rows_side0, [3] plain_rows, [4] light limit, [5] smallest part.
 * Writes min(n, 6) values.
 */
int mnufft_plan_tuning(mnufft_plan plan, int64_t *vals, int32_t n);

/*
 * Dense bins of the 3D matrix-core type-1 spread (no counterpart in the
 * reference, whose heavy blocks cover clusters): a bin holding at least
 * min_records points (default 32; doubled, up to seven times, while the
 * bins' windows exceed 16384 parts of 1024 points) is summed on its own over
 * its window of (bin side + w - 1) cells per axis, and the windows are added
 * to the tile sums in a fixed bin order.  0 turns dense bins off.  Applies
 * from the next set_points (an execute re-bins kept points); output stays
 * bitwise repeatable and, in coordinate order, independent of the point
 * order.  Grids whose bins do not fit 16 x 8 x 8-cell windows have none.
 */
int mnufft_set_dense_bins(mnufft_plan plan, int64_t min_records);

/*
 * Autotuner of a type-1 plan (no counterpart in the reference): times
 * `repeats` encodes of the caller's points and strengths (d_out receives
 * them) for every candidate -- the plain spread and the matrix-core methods
 * whose operand error stays below the plan's tolerance (FP16 hi/lo split;
 * bfloat16 split at tol >= 1e-4; single FP16 at tol >= 1e-3), in 3D each
 * with dense bins on (32) and off -- and keeps the fastest on the plan
 * (dense bins only when at least 3% faster than the same method without);
 * then, for the method kept, its bins and tiles one knob at a time, each
 * kept when at least 3% faster: the plain 3D spread on the reference's
 * output-stationary blocks instead of the row tiles, bins of 4 or 8 cells
 * along axis 0 of the row geometry, and a light limit of 1024 or 16384
 * records for the heavy-tile split (mnufft_plan_tuning reads the result).
 * choice (optional, 4 values): spread method, dense-bin threshold, time per
 * encode in ns, candidates tried.  Synchronizes the stream; the plan's
 * output stays bitwise repeatable for the chosen configuration.
 */
int mnufft_autotune(mnufft_plan plan, int64_t active_batch, int64_t M, const void *d_points,
                    const void *d_in, void *d_out, int32_t repeats, int64_t *choice,
                    void *stream);

/*
 * FP64 direct sums on the GPU: reference_type1_f64 / reference_type2_f64
 * (src/direct.rs:9-118) and reference_type3_f64 (src/type3.rs:609-673).
 * kind 1: out[o] = sum_j in[j] exp(i sign k_o.x_j) over the linear mode
 * indices d_sel (all modes if NULL); kind 2: out[o] = sum_k in[k]
 * exp(i sign k.x_o) at the points d_sel (all if NULL); kind 3: out[o] =
 * sum_j in[j] exp(i sign s_o.x_j) at the targets d_sel (all if NULL).
 */
int mnufft_direct_f64(int32_t kind, int32_t dim, const int64_t *n_modes, int32_t isign,
                      int32_t mode_order, int64_t M, const double *d_x, const double *d_in,
                      int64_t nout, const int64_t *d_sel, const double *d_s, double *d_out,
                      void *stream);

/*
 * Host-side planning facts (no GPU needed):
 *  - mnufft_kernel_params: ES width and beta of a tolerance
 *    (EsKernel::for_tolerance, src/kernel.rs:21-57);
 *  - mnufft_fine_grid_size: fine-grid length of an axis
 *    (select_fine_grid_size, src/kernel.rs:278-333), -1 if none;
 *  - mnufft_layout_query: the reference spread layout and the bin geometry
 *    it chooses for point_count points (NdSpreadLayout::for_grid and
 *    choose_geometry, src/gpu_nd_spread.rs:194-459), dim 1..8.  out (128
 *    entries): [0] run, [1] workgroup size, [2] row batch, [3] geometry
 *    count, [4] chosen geometry, [8 + a] block, [16 + a] blocks,
 *    [32 + 8g + a] the bin side of geometry g along axis a.
 */
int mnufft_kernel_params(double eps, double sigma, int32_t *width, double *beta);
int64_t mnufft_fine_grid_size(int64_t n_modes, double sigma, int32_t width);
/*
 *  - mnufft_tolerance_check: the plan-time tolerance checks of a plan of
 *    `dim` (1..8) axes -- the ES kernel of eps at sigma and, for
 *    MNUFFT_PRECISION_F32, the estimate of the F32 rounding error against
 *    the tolerance (validate_f32_rounding, src/config.rs:252-284).  0 when a
 *    plan of that shape could meet eps, else MNUFFT_ERROR_PLAN (also for a
 *    shape no fine grid fits).  mnufft_makeplan runs the same checks.
 */
int mnufft_tolerance_check(int32_t dim, const int64_t *n_modes, double eps, double sigma,
                           int32_t precision);
int mnufft_layout_query(int32_t dim, const int64_t *fine, int32_t width, int32_t precision,
                        int64_t point_count, int64_t *out);

/* The message of the last error on this thread. */
const char *mnufft_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MNUFFT_H */
