// Fine-grid FFT of 2D and 3D plans as passes of batched line transforms
// (one per axis), each line staged in LDS and transformed there by
// Stockham passes of radix 2, 3, 4, 5, 7, 8, 11 and 13 (the fine lengths'
// prime factors), with the passes pruned to the
// lines that carry data: a type-2 grid holds modes only in the kept window
// of every axis (the other cells are never read), and a type-1 grid is
// needed only at the kept window, so
//
//   type 2 (modes in, grid out): x over the kept (y, z) lines, reading the
//     kept x cells; y over every x of the kept z planes, reading kept y;
//     z over every (x, y) column, reading kept z;
//   type 1 (grid in, modes out): x over every line, writing kept x; y over
//     the kept x of every z plane, writing kept y; z over the kept (x, y)
//     columns, writing kept z.
//
// (2D: x over the kept rows then y over every column, or the reverse.)
// That is 1/4 + 1/2 + 1 of the full grid's lines instead of 3 full passes,
// and a type-2 grid needs no zeros between the modes.  The kept window of
// an axis of n cells holding N modes is [0, ceil(N/2)) and [n - floor(N/2), n)
// (the fine index of mode m is m mod n).  Unnormalized; isign > 0 is
// exp(+i k x).  Every axis at most 4864 (FP32) or 2432 (FP64) cells (one
// line and its twiddles within the 76 KB LDS budget).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "nufft_internal.hpp"

namespace mnufft {

constexpr int kLineFftMaxPasses = 24;

struct LineFftPass {
    int radix, p, tw_step;  // radix, product of the earlier radices, n / (p radix)
};

// One axis pass: lines enumerated as (u, v), u fastest, each mapped through
// its kept window (u' = u < u_hi ? u : u + u_gap), line base u' su + v' sv,
// element k at base + k estride.  Lines per workgroup: `lines` (strided
// axes: a power of two, adjacent u in one workgroup).
struct LineFftAxis {
    int n = 1, npass = 0;
    LineFftPass pass[kLineFftMaxPasses] = {};
    int lines = 1, log_lines = 0;
    int line_major = 0;  // the axis is the contiguous one
    int64_t estride = 1;
    int64_t nu = 1, nv = 1, u_hi = 1, u_gap = 0, v_hi = 1, v_gap = 0, su = 0, sv = 0;
    int in_hi = 1, in_lo = 0, out_hi = 1, out_lo = 0;  // elements read / written
    int64_t vec_stride = 0;
    float inv_n = 1.0f;
    // Modes in place of the grid (chosen per launch): the first pass reads
    // each cell from the modes, scaled by the deconvolution factors (type 2),
    // or the last pass writes each kept cell to the modes, scaled (type 1).
    // Per role (u, v, e = the line's elements): modes N, fine cells n,
    // mode-array stride and the axis's offset in the 1/phi_hat table.
    int centered = 0, has_u = 0, has_v = 0;
    int64_t uN = 1, un = 1, ums = 0, uph = 0;
    int64_t vN = 1, vn = 1, vms = 0, vph = 0;
    int64_t eN = 1, ems = 0, eph = 0;
    int64_t mode_total = 0;
    const double *inv_phihat = nullptr;
};

// One plan's axis pass inside a batched launch: its workgroups start at
// group0 of the launch.
struct LineFftJob {
    LineFftAxis ax;
    void *grid;
    const void *tw;
    int64_t group0;
};

class LineFft {
public:
    ~LineFft();
    // False (no error set) when the grid does not qualify.
    static bool supported(int dim, const int64_t *fine, bool f64);
    // Longest axis whose strided passes still move 64-byte rows (4 FP64 or
    // 8 FP32 adjacent columns per workgroup): 972 FP64, 1080 FP32 cells.
    static int64_t row_length_limit(bool f64);
    // type 1: output pruned to the modes; type 2: input pruned to them.
    // inv_phihat / phihat_off / centered: the deconvolution factors 1/phi_hat
    // per axis and the mode order, for the runs that read or write modes.
    bool init(int dim, const int64_t *fine, const int64_t *n_modes, int type, bool f64,
              int isign, const double *inv_phihat, const int64_t *phihat_off, bool centered);
    // With `modes`: type 2 reads the modes (vectors x mode_total) and scales
    // them as the predeconvolution does, type 1 writes the deconvolved modes;
    // the results equal predeconvolution + FFT (+ deconvolution) bit for bit.
    bool run(void *data, int64_t vectors, hipStream_t stream, void *modes = nullptr) const;
    // One vector of each of many plans (one rank and precision): every axis
    // pass of all of them in one launch per transform sign.  The job table
    // (*jobs, *cap entries) grows as needed and is filled by kernel launches
    // (no host copies), so the call can be captured in a graph.
    static bool run_many(const std::vector<const LineFft *> &ffts,
                         const std::vector<void *> &grids, LineFftJob **jobs, int64_t *cap,
                         hipStream_t stream);

private:
    int naxes_ = 0, type_ = 1;
    bool f64_ = false;
    int dir_ = 1;
    int64_t fine_total_ = 1;
    LineFftAxis ax_[3];
    void *tw_[3] = {nullptr, nullptr, nullptr};
};

}  // namespace mnufft
