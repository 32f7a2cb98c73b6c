// Internal declarations of the MI355X NUFFT engine (host side and kernels).
//
// Layout conventions (the reference's plan surface, src/plan.rs and
// capi/include/wgpu_nufft.h:10-20):
//  * coordinates are point-major: x[j * dim + axis];
//  * complex values are interleaved (re, im);
//  * modes and fine grids are axis-0 fastest;
//  * batches are transform-major: the whole first vector, then the next.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace mnufft {

constexpr int kMaxDims = 8;
constexpr int kMaxWidth = 16;
// Prepared type-1 records of rank D: ranks 1-3 one int4 per record (the
// wrapped support starts along x, y, z and the original point index) and
// four offsets; ranks 4-8 two int4 (eight starts) and eight offsets.
template <int D>
constexpr int rec_quads()
{
    return D <= 3 ? 1 : 2;
}
// Point indices, bin offsets, record slots and permutations are 32-bit:
// a point set (or type-3 target set) holds at most this many points.
constexpr int64_t kMaxPoints = (int64_t)UINT32_MAX - 1;
constexpr int kMaxGeometries = 16;
// Inner fine grids of a type-3 plan stay within 32-bit cell indices per
// vector; larger target intervals are cut into boxes.
constexpr int64_t kType3InnerCells = (int64_t)1 << 30;
constexpr int kType3MaxBoxes = 64;  // per axis
// Bins above this many points are ordered by the large-bin sort; smaller
// ones are ranked inside the prepare pass (src/gpu_bin_sort.rs SMALL_BIN).
constexpr uint32_t kSmallBin = 64;

enum class Precision : int { F64 = 1, Df64 = 2, F32 = 3 };
enum class ModeOrder : int { Centered = 0, Fft = 1 };
// Arithmetic of the type-1 spread: plain FP32/FP64 FMAs, or the per-tile
// contraction on the matrix cores with FP16 inputs (one product), FP16
// hi/lo splits (three products) or bfloat16 hi/lo splits (three products),
// all accumulated in FP32.
enum class SpreadMethod : int { Plain = 0, MfmaF16 = 1, MfmaF16x3 = 2, MfmaBf16x3 = 3 };
// In-bin order of the sorted points.
enum class BinOrder : int {
    Grouped = 0,    // by support-start cell inside each bin (type 2)
    Index = 1,      // ascending original index: the reference permutation
    Canonical = 2,  // ascending coordinate bits, then index (type 1, opt-in)
};

// ---------------------------------------------------------------------
// Host math: ES kernel, fine grids, spread layout (es_kernel.cpp,
// spread_layout.cpp).
// ---------------------------------------------------------------------
struct EsKernel {
    int width = 0;
    double beta = 0.0;
    double sigma = 2.0;
    double eps = 1e-6;
    double half_width() const { return 0.5 * width; }
    double evaluate(double distance) const;
};

// Returns false (and sets the last error) for an unsupported tolerance.
bool es_kernel_for_tolerance(double eps, double sigma, EsKernel *out);
bool fft_length_supported(int64_t n);
int64_t select_fine_grid_size(int64_t n_modes, double sigma, int width);
void gauss_legendre(int order, double *nodes, double *weights);
std::vector<double> centered_fourier_coefficients(const EsKernel &k, int64_t n);
double es_kernel_fourier_transform(const EsKernel &k, double frequency);
int64_t next_even_235(int64_t minimum);
// Rejects F32 plans whose rounding misses the tolerance by far
// (src/config.rs:252-284).
bool validate_f32_rounding(const std::vector<int64_t> &n_modes, double eps, double sigma,
                           Precision precision, const std::vector<std::vector<double>> &coeffs);

struct BinGeometry {
    int64_t side[kMaxDims];
    int64_t bins[kMaxDims];
    bool full[kMaxDims];
    int64_t reach_bins[kMaxDims];
    double mean_reach[kMaxDims];
    int64_t bin_count(int d) const {
        int64_t c = 1;
        for (int a = 0; a < d; ++a) c *= bins[a];
        return c;
    }
};

// The reference's spread layout (src/gpu_nd_spread.rs:194-407), used to
// pick the bin geometry of a point count exactly as the reference does.
struct SpreadLayout {
    int dims = 0;
    int width = 0;
    int64_t fine[kMaxDims] = {1, 1, 1, 1, 1, 1, 1, 1};
    int64_t run = 1, groups0 = 1, workgroup_size = 1, row_batch = 1;
    int64_t block[kMaxDims] = {1, 1, 1, 1, 1, 1, 1, 1};
    int64_t blocks[kMaxDims] = {1, 1, 1, 1, 1, 1, 1, 1};
    std::vector<BinGeometry> geometries;
    int choose_geometry(int64_t point_count) const;
};
SpreadLayout build_spread_layout(int d, const int64_t *fine, int width, Precision precision);

// ---------------------------------------------------------------------
// Kernel parameter blocks (passed by value).
// ---------------------------------------------------------------------
struct AxisParams {
    int64_t n;        // fine-grid length
    int shift;        // log2 of the bin side
    int64_t bins;     // bins along the axis
    int64_t stride;   // bin-key stride
    float scale_hi;   // n / (2 pi) split into f32 head and tail (F32 fold)
    float scale_lo;
    double scale;     // n / (2 pi) (F64 fold)
    double origin;    // n / 2
};

struct GridParams {
    int dims;
    int width;
    double half_width;
    double beta;
    AxisParams ax[kMaxDims];
    int64_t nbins;
    int64_t fine_total;
};

// One problem of a multi-plan execution: its points take indices
// [m_off, m_off + M) of the concatenated point set and bins [bin_base,
// bin_base + g.nbins) of the combined bin key; its interpolation work items
// start at item_base.  Type 2 reads `grid` and writes `out` (values at the
// points); type 1 reads `in` (strengths), spreads onto `grid` and writes
// `out` (modes).
struct BatchProblem {
    const void *coords;
    void *out;
    const void *grid;
    const void *in;
    int64_t m_off, M, item_base;
    uint64_t bin_base;
    GridParams g;
};

// Spread blocks of the output-stationary spreader: `run` cells per lane
// along axis 0, 64 lanes per block.
struct SpreadParams {
    int64_t block[kMaxDims];
    int64_t nblocks[kMaxDims];
    int64_t block_count;
    int run;
    int groups0;
};

// ---------------------------------------------------------------------
// Errors: a thread-local message, as the reference C ABI keeps one
// (capi/src/error.rs:61-78).
// ---------------------------------------------------------------------
void set_error(const std::string &message);
const char *last_error();
bool hip_ok(hipError_t e, const char *what);
// False (with the last error set) for a count outside 0..kMaxPoints.
bool point_count_ok(int64_t count, const char *what);

// Stage names of the per-execution timing (src/gpu_profile.rs:12-42).
enum Stage : int {
    kStageSort = 0,        // count + scan + scatter + in-bin order
    kStagePrepare = 1,     // prepared records
    kStageSpread = 2,      // type-1 spread (classify, parts, reduce)
    kStageFft = 3,         // fine-grid FFT (rocFFT)
    kStageDeconv = 4,      // deconvolution / predeconvolution
    kStageInterp = 5,      // type-2 interpolation
    kStageType3 = 6,       // type-3 source/target/final stages
    kStageSpreadKernel = 7,  // the output-stationary spread kernel alone
    kStageCount = 8,
};

}  // namespace mnufft
