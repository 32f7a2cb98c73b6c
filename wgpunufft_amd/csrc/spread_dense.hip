// Dense bins of the 3D matrix-core spread.
//
// Same contract as the tile spread (src/gpu_nd_spread.rs:1-36): every cell
// sums its contributions in a fixed order, with no float atomics, so the
// output is bitwise repeatable and, in coordinate order, independent of the
// input order of the points.  The reference meets clusters with heavy blocks
// split into parts (src/gpu_nd_spread.rs:473-476, 1781-1905); here a bin
// that holds many records is summed on its own instead, once, rather than
// by every tile its records reach.
//
// A bin of sides (sx, sy, sz) cells holds the records whose wrapped support
// start lies in it, so their supports fall inside the bin's window of
// (sx + w - 1) x (sy + w - 1) x (sz + w - 1) cells from the bin's first cell.
// The window's sum is one matrix product over the bin's records j:
//
//   Win[(y,z), (x,re|im)] = sum_j  Ky_j[y] Kz_j[z]  *  Kx_j[x] c_j
//                                  \_ A (64 x K) _/    \_ B (K x 16 NXB) _/
//
// (y, z < 8, x < 8 NXB), on v_mfma_f32_16x16x32 with the operand precision
// of the tile spread.  Bins of at least `threshold` records are dense; their
// records sit after the light ones in slot order and the tile kernels see the
// light records only.  A dense bin's records are summed in parts of
// kDensePart records (one wavefront each), the parts are added in part
// order, and every cell of a tile that a dense window reaches then adds the
// windows of its candidate bins in a fixed bin order to the value the tile
// kernels wrote.
#include <hip/hip_runtime.h>

#include "bins.hpp"
#include "mfma_ops.hpp"
#include "spread.hpp"
#include "spread_device.hpp"

namespace mnufft {

namespace {

using namespace mfma;

constexpr uint32_t kDensePart = 1024;  // records per window part
constexpr int kDenseLevels = 8;        // thresholds t, 2t, ..., 128t
constexpr int kDP = 36;                // LDS pitch of the weight rows (32 records + 4)
// ctrl words: [0, 8) window parts needed at each threshold level, [8] the
// threshold in use, [9] window parts handed out, [16, 24) records of the
// dense bins at each level.
constexpr int kCtrlThreshold = 8;
constexpr int kCtrlParts = 9;
constexpr int kCtrlRecords = 16;  // [16, 24): records of the dense bins at each level
constexpr int kDenseShare = 8;    // dense bins hold at least 1/8 of the records

inline unsigned grid_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

__global__ void k_dense_clear(uint32_t *ctrl)
{
    if (threadIdx.x < 32) ctrl[threadIdx.x] = 0u;
}

// Window parts (ctrl[i]) and records (ctrl[kCtrlRecords + i]) of the bins
// that would be dense at each threshold level.
__global__ __launch_bounds__(256) void k_dense_levels(const uint32_t *__restrict__ offsets,
                                                      int64_t nbins, uint32_t t0,
                                                      uint32_t *__restrict__ ctrl)
{
    uint32_t acc[kDenseLevels], rec[kDenseLevels];
#pragma unroll
    for (int i = 0; i < kDenseLevels; ++i) acc[i] = rec[i] = 0;
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nbins;
         b += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t n = offsets[b + 1] - offsets[b];
        if (n >= t0) {
            const uint32_t np = (n + kDensePart - 1) / kDensePart;
#pragma unroll
            for (int i = 0; i < kDenseLevels; ++i)
                if ((uint64_t)n >= ((uint64_t)t0 << i)) {
                    acc[i] += np;
                    rec[i] += n;
                }
        }
    }
#pragma unroll
    for (int i = 0; i < kDenseLevels; ++i) {
        uint32_t x = acc[i], y = rec[i];
        for (int o = 32; o > 0; o >>= 1) {
            x += __shfl_xor(x, o, 64);
            y += __shfl_xor(y, o, 64);
        }
        if ((threadIdx.x & 63) == 0 && x) {
            atomicAdd(&ctrl[i], x);
            atomicAdd(&ctrl[kCtrlRecords + i], y);
        }
    }
}

// Smallest threshold level whose window parts fit the capacity (all-or-
// nothing per level, so the choice depends on the bin counts only); none
// when the dense bins at that level hold less than 1/kDenseShare of the
// records (the window sums would not pay for their adds).
__device__ inline uint32_t dense_threshold(const uint32_t *ctrl, uint32_t t0, uint32_t cap,
                                           int64_t M)
{
    for (int i = 0; i < kDenseLevels; ++i)
        if (ctrl[i] <= cap) {
            if ((int64_t)ctrl[kCtrlRecords + i] * kDenseShare < M) return 0xffffffffu;
            return (uint32_t)min((uint64_t)t0 << i, (uint64_t)0xffffffffu);
        }
    return 0xffffffffu;
}

// cnt[b] = records of bin b when it is dense, else 0 (cnt[nbins] = 0).
__global__ __launch_bounds__(256) void k_dense_mark(const uint32_t *__restrict__ offsets,
                                                    int64_t nbins, uint32_t t0, uint32_t cap,
                                                    const uint32_t *__restrict__ ctrl,
                                                    uint32_t *__restrict__ cnt)
{
    const uint32_t t = dense_threshold(ctrl, t0, cap, offsets[nbins]);
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nbins) return;
    uint32_t c = 0;
    if (b < nbins) {
        const uint32_t n = offsets[b + 1] - offsets[b];
        c = n >= t ? n : 0u;
    }
    cnt[b] = c;
}

// Light-record offsets, and the window parts of every dense bin: info[b] =
// 1 + its first part (0 for a light bin), parts[first + p] = (b, p).
// Tiles of the spread (blocks of sp) that cells [lo, lo + ext) of one axis
// (periodic) touch: at most 3 (a wrap through a partial last tile).
__device__ inline int axis_tiles(int64_t lo, int ext, int64_t n, int64_t block, int64_t *t)
{
    int k = 0;
    for (int e = 0; e < ext; ++e) {
        int64_t c = lo + e;
        if (c >= n) c -= n;
        const int64_t q = c / block;
        bool seen = false;
        for (int i = 0; i < k; ++i) seen = seen || t[i] == q;
        if (!seen && k < 3) t[k++] = q;
    }
    return k;
}

__global__ __launch_bounds__(256) void k_dense_finish(const uint32_t *__restrict__ offsets,
                                                      const uint32_t *__restrict__ before,
                                                      int64_t nbins, uint32_t t0, uint32_t cap,
                                                      GridParams g, SpreadParams sp,
                                                      uint32_t *__restrict__ ctrl,
                                                      uint32_t *__restrict__ light,
                                                      uint32_t *__restrict__ info,
                                                      uint2 *__restrict__ parts,
                                                      uint32_t *__restrict__ tile_flag,
                                                      uint32_t *__restrict__ touched)
{
    const uint32_t t = dense_threshold(ctrl, t0, cap, offsets[nbins]);
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b == 0) ctrl[kCtrlThreshold] = t;
    if (b > nbins) return;
    light[b] = offsets[b] - before[b];
    if (b == nbins) return;
    const uint32_t n = offsets[b + 1] - offsets[b];
    uint32_t tag = 0;
    if (n >= t) {
        const uint32_t np = (n + kDensePart - 1) / kDensePart;
        const uint32_t first = atomicAdd(&ctrl[kCtrlParts], np);
        for (uint32_t p = 0; p < np; ++p) parts[first + p] = make_uint2((uint32_t)b, p);
        tag = first + 1;
        // list every tile the bin's window touches, once
        int64_t tt[3][3];
        int nt[3];
        int64_t rest = b;
        for (int a = 0; a < 3; ++a) {
            const int64_t bi = rest % g.ax[a].bins;
            rest /= g.ax[a].bins;
            nt[a] = axis_tiles(bi << g.ax[a].shift, (1 << g.ax[a].shift) + g.width - 1,
                               g.ax[a].n, sp.block[a], tt[a]);
        }
        for (int k2 = 0; k2 < nt[2]; ++k2)
            for (int k1 = 0; k1 < nt[1]; ++k1)
                for (int k0 = 0; k0 < nt[0]; ++k0) {
                    const int64_t tile =
                        tt[0][k0] + sp.nblocks[0] * (tt[1][k1] + sp.nblocks[1] * tt[2][k2]);
                    if (atomicExch(&tile_flag[tile], 1u) == 0u)
                        touched[1 + atomicAdd(&touched[0], 1u)] = (uint32_t)tile;
                }
    }
    info[b] = tag;
}

__global__ void k_dense_clear_tiles(uint32_t *tile_flag, int64_t n, uint32_t *touched)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) tile_flag[i] = 0u;
    if (i == 0) touched[0] = 0u;
}

// One lane's record of a k-block, its loads in flight until used: the
// record two k-blocks ahead, its strength (through the record's original
// index) one k-block ahead.
struct DenseRec {
    int4 st;
    float4 off;
    float2 c;
    bool valid;
};

__device__ inline DenseRec dense_fetch(int64_t slot, int64_t end, const int4 *__restrict__ rec_start,
                                       const float *__restrict__ rec_off)
{
    DenseRec r;
    r.valid = slot < end;
    r.st = make_int4(0, 0, 0, 0);
    r.off = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    r.c = make_float2(0.0f, 0.0f);
    if (r.valid) {
        r.st = rec_start[slot];
        r.off = *(const float4 *)(rec_off + 4 * slot);
    }
    return r;
}

__device__ inline void dense_fetch_strength(DenseRec &r, const float *__restrict__ cv)
{
    if (r.valid) r.c = *(const float2 *)(cv + 2 * (int64_t)(uint32_t)r.st.w);
}

// Window part sums: one wavefront per part (grid-stride), 32 records per
// k-block.  Lane l computes half of the weight rows of record l % 32 (rows:
// the window's x cells, then 8 y cells, then 8 z cells, each the weight of
// the record at that cell or zero off its support).
template <int MODE, int NXB>
__global__ __launch_bounds__(64) void k_dense_windows(
    GridParams g, const uint2 *__restrict__ parts, const uint32_t *__restrict__ ctrl,
    const uint32_t *__restrict__ before, int64_t M, const int4 *__restrict__ rec_start,
    const float *__restrict__ rec_off, const float *__restrict__ c,
    const uint32_t *__restrict__ cmax, float *__restrict__ win, int64_t win_cap)
{
    using Op = Operand<MODE>;
    using V = typename Op::V;
    constexpr int NX = 8 * NXB;      // window cells along x
    constexpr int E = NX + 16;       // weight rows per record
    constexpr int EH = E / 2;        // weight rows per lane
    constexpr int WF = 64 * 2 * NX;  // floats per window
    __shared__ __align__(16) float Wt[E * kDP];
    __shared__ __align__(16) float CC[2 * kDP];

    const int lane = threadIdx.x;
    const int j = lane & 31, h = lane >> 5;
    const int r16 = lane & 15, j0 = 8 * (lane >> 4);
    const int64_t v = blockIdx.y;
    float scale, unscale;
    strength_scale(cmax + v, &scale, &unscale);
    const int w = g.width;
    const float two_over_w = 2.0f / (float)w;
    const float beta_l2e = (float)(g.beta * 1.4426950408889634);
    const int64_t nparts = ctrl[kCtrlParts];
    const int64_t light_total = M - (int64_t)before[g.nbins];
    const float *cv = c + 2 * v * M;
    float *wv = win + v * win_cap * WF;

    for (int64_t part = blockIdx.x; part < nparts; part += gridDim.x) {
        const uint2 bp = parts[part];
        const int64_t b = bp.x;
        const uint32_t n = before[b + 1] - before[b];
        const uint32_t done = bp.y * kDensePart;
        const int64_t s0 = light_total + before[b] + done;
        const int64_t s1 = s0 + (n - done < kDensePart ? n - done : kDensePart);
        int org[3];
        {
            int64_t rest = b;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int64_t bi = rest % g.ax[a].bins;
                rest /= g.ax[a].bins;
                org[a] = (int)(bi << g.ax[a].shift);
            }
        }
        f32x4 acc[NXB][4];
#pragma unroll
        for (int xb = 0; xb < NXB; ++xb)
#pragma unroll
            for (int m = 0; m < 4; ++m) acc[xb][m] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        DenseRec nx = dense_fetch(s0 + j, s1, rec_start, rec_off);
        DenseRec nx2 = dense_fetch(s0 + 32 + j, s1, rec_start, rec_off);
        dense_fetch_strength(nx, cv);
        for (int64_t k0 = s0; k0 < s1; k0 += 32) {
            const DenseRec cur = nx;
            nx = nx2;
            dense_fetch_strength(nx, cv);
            nx2 = dense_fetch(k0 + 64 + j, s1, rec_start, rec_off);
            const int dx = cur.st.x - org[0], dy = cur.st.y - org[1], dz = cur.st.z - org[2];
#pragma unroll
            for (int i = 0; i < EH; ++i) {
                const int e = h * EH + i;
                const bool ax0 = e < NX, ax1 = !ax0 && e < NX + 8;
                const int u = ax0 ? e : (ax1 ? e - NX : e - NX - 8);
                const int d = ax0 ? dx : (ax1 ? dy : dz);
                const float o = ax0 ? cur.off.x : (ax1 ? cur.off.y : cur.off.z);
                const int t = u - d;
                // computed for every row, then selected (no branches)
                const float wt = es_weight_fast((float)t + o, two_over_w, beta_l2e);
                Wt[e * kDP + j] = cur.valid && (unsigned)t < (unsigned)w ? wt : 0.0f;
            }
            CC[h * kDP + j] = cur.valid ? (h == 0 ? cur.c.x : cur.c.y) * scale : 0.0f;
            wave_sync();
            V bh[NXB], bl[NXB];
#pragma unroll
            for (int xb = 0; xb < NXB; ++xb) {
                const f32x8 kx = load8(Wt + (xb * 8 + (r16 >> 1)) * kDP + j0);
                const f32x8 cc = load8(CC + (r16 & 1) * kDP + j0);
                Op::split(kx * cc, bh[xb], bl[xb]);
            }
            const f32x8 ky = load8(Wt + (NX + (r16 & 7)) * kDP + j0);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const f32x8 kz = load8(Wt + (NX + 8 + 2 * m + (r16 >> 3)) * kDP + j0);
                V ah, al;
                Op::split(ky * kz, ah, al);
#pragma unroll
                for (int xb = 0; xb < NXB; ++xb) {
                    acc[xb][m] = Op::mma16(ah, bh[xb], acc[xb][m]);
                    if constexpr (Op::kSplit) {
                        acc[xb][m] = Op::mma16(ah, bl[xb], acc[xb][m]);
                        acc[xb][m] = Op::mma16(al, bh[xb], acc[xb][m]);
                    }
                }
            }
            wave_sync();
        }
        // D element i of block m: row 16 m + 4 (lane / 16) + i = 8 z + y,
        // column r16 = 2 x + (re|im) within the x block
        float *wp = wv + part * WF;
#pragma unroll
        for (int xb = 0; xb < NXB; ++xb)
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int row = 16 * m + 4 * (lane >> 4) + i;
                    wp[(row * NX + xb * 8) * 2 + r16] = acc[xb][m][i] * unscale;
                }
    }
}

// A dense bin's parts added in part order into its first part: one
// workgroup per bin (the entry of its first part), one float4 per thread,
// the parts' loads issued eight at a time.
template <int NXB>
__global__ __launch_bounds__(256) void k_dense_reduce(const uint2 *__restrict__ parts,
                                                      const uint32_t *__restrict__ ctrl,
                                                      const uint32_t *__restrict__ before,
                                                      float *__restrict__ win, int64_t win_cap)
{
    constexpr int WF = 64 * 2 * 8 * NXB;
    const int64_t v = blockIdx.y;
    const int64_t nparts = ctrl[kCtrlParts];
    // the last window of every vector stays zero: the add kernel reads it for
    // candidate bins that are not dense
    if (blockIdx.x == 0)
        for (int e = 4 * threadIdx.x; e < WF; e += 4 * 256)
            *(float4 *)(win + (v * win_cap + win_cap - 1) * WF + e) =
                make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int64_t part = blockIdx.x; part < nparts; part += gridDim.x) {
        const uint2 bp = parts[part];
        if (bp.y != 0) continue;
        const uint32_t n = before[bp.x + 1] - before[bp.x];
        const uint32_t np = (n + kDensePart - 1) / kDensePart;
        if (np < 2) continue;
        float *w0 = win + (v * win_cap + part) * WF;
        for (int e = 4 * threadIdx.x; e < WF; e += 4 * 256) {
            float4 s = *(const float4 *)(w0 + e);
            for (uint32_t q0 = 1; q0 < np; q0 += 8) {
                float4 t[8];
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (q0 + k < np) t[k] = *(const float4 *)(w0 + (int64_t)(q0 + k) * WF + e);
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (q0 + k < np) {
                        s.x += t[k].x;
                        s.y += t[k].y;
                        s.z += t[k].z;
                        s.w += t[k].w;
                    }
            }
            *(float4 *)(w0 + e) = s;
        }
    }
}

// Candidate bins of a cell along one axis: the k-th has its window's first
// cell k bins below the cell's bin, at offset (c mod side) + k side.
struct AxisCand {
    int64_t bin0;  // the cell's bin
    int off0;      // the cell's offset in its bin
    int count;     // candidates whose window reaches the cell
};

__device__ inline AxisCand axis_cand(int64_t c, const AxisParams &a, int ext)
{
    AxisCand r;
    r.bin0 = c >> a.shift;
    r.off0 = (int)(c & ((1 << a.shift) - 1));
    const int side = 1 << a.shift;
    r.count = (ext - r.off0 + side - 1) / side;
    return r;
}

// Every cell of a tile reached by a dense window adds the windows of its
// candidate bins, z, then y, then x candidates, each from the cell's own bin
// down.  One workgroup per (listed tile, z cell), one thread per (x, y)
// cell (tiles of at most 16 x 16 cells per z).  The window tags of the
// slice's candidate bins are staged in LDS, with an ordered list of the
// dense ones.  A short list is walked by every cell (a window that does not
// reach the cell is skipped); otherwise every cell reads all its candidates,
// a light one from the zero window.  Both visit a cell's dense candidates
// in the same order and add exact zeros for the rest, so they give the same
// bits.
constexpr int kAddBx = 12, kAddBy = 24, kAddBz = 8;  // staged candidate bins per axis
constexpr int kAddStaged = kAddBx * kAddBy * kAddBz;
constexpr int kAddCy = 8;      // y candidates of a cell at most
constexpr int kAddList = 128;  // dense candidates walked by every cell at most

template <int NXB>
__global__ __launch_bounds__(256) void k_dense_add(GridParams g, SpreadParams sp,
                                                   const uint32_t *__restrict__ touched,
                                                   const uint32_t *__restrict__ info,
                                                   const float *__restrict__ win,
                                                   int64_t win_cap, float *__restrict__ grid)
{
    constexpr int NX = 8 * NXB;
    constexpr int WF = 64 * 2 * NX;
    constexpr int CX = NXB == 1 ? 4 : 3;  // x candidates of a cell at most
    __shared__ uint32_t tags[kAddStaged + CX];
    __shared__ uint32_t list[kAddStaged];  // (kz << 16 | ky << 8 | kx) of the dense candidates
    __shared__ uint32_t wsum[4];
    const int64_t v = blockIdx.y;
    const int w = g.width;
    int ext[3], sh[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        sh[a] = g.ax[a].shift;
        ext[a] = (1 << sh[a]) + w - 1;
    }
    const float *wv = win + v * win_cap * WF;
    float *gv = grid + 2 * v * g.fine_total;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int64_t units = (int64_t)touched[0] * sp.block[2];
    for (int64_t unit = blockIdx.x; unit < units; unit += gridDim.x) {
        const int64_t blk = touched[1 + unit / sp.block[2]];
        const int64_t bx = blk % sp.nblocks[0];
        const int64_t by = (blk / sp.nblocks[0]) % sp.nblocks[1];
        const int64_t bz = blk / (sp.nblocks[0] * sp.nblocks[1]);
        const int64_t x0 = bx * sp.block[0], y0 = by * sp.block[1];
        const int64_t cz = bz * sp.block[2] + unit % sp.block[2];
        if (cz >= g.ax[2].n) continue;
        // candidate bins of the slice: per axis from the bin of the slice's
        // last cell inside the grid down to the lowest bin whose window
        // reaches the first cell
        const int64_t hx = (std::min<int64_t>(x0 + sp.block[0], g.ax[0].n) - 1) >> sh[0];
        const int64_t hy = (std::min<int64_t>(y0 + sp.block[1], g.ax[1].n) - 1) >> sh[1];
        const int64_t hz = cz >> sh[2];
        const int nbx = (int)(hx - ((x0 - ext[0] + 1) >> sh[0]) + 1);
        const int nby = (int)(hy - ((y0 - ext[1] + 1) >> sh[1]) + 1);
        const int nbz = (int)(hz - ((cz - ext[2] + 1) >> sh[2]) + 1);
        const int ns = nbz * nby * nbx;
        // staged t = kx + nbx (ky + nby kz) is bin (hx - kx, hy - ky, hz - kz);
        // each thread stages a run of consecutive t, then the dense ones are
        // listed in t order (a block-wide scan of the runs' counts)
        const int run = (ns + 255) / 256;
        __syncthreads();
        uint32_t mine = 0;
        for (int r = 0; r < run; ++r) {
            const int t = threadIdx.x * run + r;
            if (t >= ns) break;
            const int kx = t % nbx, ky = (t / nbx) % nby, kz = t / (nbx * nby);
            int64_t ix = hx - kx, iy = hy - ky, iz = hz - kz;
            if (ix < 0) ix += g.ax[0].bins;
            if (iy < 0) iy += g.ax[1].bins;
            if (iz < 0) iz += g.ax[2].bins;
            const uint32_t tag = info[ix + g.ax[0].bins * (iy + g.ax[1].bins * iz)];
            tags[(kz * kAddBy + ky) * kAddBx + kx] = tag ? tag - 1 : (uint32_t)(win_cap - 1);
            mine += tag ? 1u : 0u;
        }
        uint32_t x = mine;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wid] = x;
        __syncthreads();
        uint32_t pos = x - mine, total = 0;
        for (int k = 0; k < 4; ++k) {
            pos += k < wid ? wsum[k] : 0u;
            total += wsum[k];
        }
        for (int r = 0; r < run; ++r) {
            const int t = threadIdx.x * run + r;
            if (t >= ns) break;
            const int kx = t % nbx, ky = (t / nbx) % nby, kz = t / (nbx * nby);
            if (tags[(kz * kAddBy + ky) * kAddBx + kx] != (uint32_t)(win_cap - 1))
                list[pos++] = (uint32_t)(kz << 16 | ky << 8 | kx);
        }
        __syncthreads();
        if (total == 0) continue;
        const int64_t cx = x0 + (threadIdx.x % sp.block[0]);
        const int64_t cy = y0 + (threadIdx.x / sp.block[0]);
        if (threadIdx.x >= sp.block[0] * sp.block[1] || cx >= g.ax[0].n || cy >= g.ax[1].n)
            continue;
        float sre = 0.0f, sim = 0.0f;
        if (total <= (uint32_t)kAddList) {
            // the cell's offset in the window of list entry (kx, ky, kz)
            const int bx0 = (int)(cx - (hx << sh[0])), by0 = (int)(cy - (hy << sh[1])),
                      bz0 = (int)(cz - (hz << sh[2]));
            for (uint32_t e0 = 0; e0 < total; e0 += 8) {
                float2 t[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const uint32_t e = e0 + k < total ? e0 + k : total - 1;
                    const uint32_t c = list[e];
                    const int kx = c & 0xff, ky = (c >> 8) & 0xff, kz = c >> 16;
                    const int ox = bx0 + (kx << sh[0]), oy = by0 + (ky << sh[1]),
                              oz = bz0 + (kz << sh[2]);
                    const bool in = e0 + k < total && (unsigned)ox < (unsigned)ext[0] &&
                                    (unsigned)oy < (unsigned)ext[1] &&
                                    (unsigned)oz < (unsigned)ext[2];
                    t[k] = make_float2(0.0f, 0.0f);
                    if (in)
                        t[k] = *(const float2 *)(wv + (int64_t)tags[(kz * kAddBy + ky) * kAddBx +
                                                                   kx] * WF +
                                                 ((oz * 8 + oy) * NX + ox) * 2);
                }
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    sre += t[k].x;
                    sim += t[k].y;
                }
            }
        } else {
            const AxisCand qx = axis_cand(cx, g.ax[0], ext[0]);
            const AxisCand qy = axis_cand(cy, g.ax[1], ext[1]);
            const AxisCand qz = axis_cand(cz, g.ax[2], ext[2]);
            const int sx = (int)(hx - qx.bin0), sy = (int)(hy - qy.bin0);
            int ox[CX];
#pragma unroll
            for (int kx = 0; kx < CX; ++kx) ox[kx] = (qx.off0 + (kx << sh[0])) * 2;
            // per z candidate, the loads of all (y, x) candidates in flight
            // together, then added in order
            for (int kz = 0; kz < qz.count; ++kz) {
                const int oz = qz.off0 + (kz << sh[2]);
                float2 t[kAddCy][CX];
#pragma unroll
                for (int ky = 0; ky < kAddCy; ++ky) {
                    const bool yin = ky < qy.count;
                    const int oy = qy.off0 + (ky << sh[1]);
                    const uint32_t *trow =
                        tags + (kz * kAddBy + sy + (yin ? ky : 0)) * kAddBx + sx;
                    const float *wrow = wv + (yin ? (oz * 8 + oy) * NX * 2 : 0);
#pragma unroll
                    for (int kx = 0; kx < CX; ++kx) {
                        const bool in = yin && kx < qx.count;
                        const int64_t wi = in ? trow[kx] : win_cap - 1;
                        t[ky][kx] = *(const float2 *)(wrow + wi * WF + (in ? ox[kx] : 0));
                    }
                }
#pragma unroll
                for (int ky = 0; ky < kAddCy; ++ky)
#pragma unroll
                    for (int kx = 0; kx < CX; ++kx) {
                        sre += t[ky][kx].x;
                        sim += t[ky][kx].y;
                    }
            }
        }
        float2 *cell = (float2 *)(gv + 2 * (cx + g.ax[0].n * (cy + g.ax[1].n * cz)));
        float2 val = *cell;
        val.x += sre;
        val.y += sim;
        *cell = val;
    }
}

template <int MODE, int NXB>
void launch_dense(const GridParams &g, const SpreadParams &sp, const DenseBins &d, int64_t M,
                  const int4 *rec_start, const float *rec_off, const float *c,
                  const uint32_t *cmax, int64_t vectors, const uint32_t *touched, float *grid,
                  bool add, hipStream_t stream)
{
    // one workgroup per window part while they fit the grid
    const unsigned pg = (unsigned)std::min<int64_t>(d.part_cap, 65535);
    // part_cap windows per vector and a zero window
    const int64_t win_cap = d.part_cap + 1;
    if (!add) {
        k_dense_windows<MODE, NXB><<<dim3(pg, (unsigned)vectors), 64, 0, stream>>>(
            g, d.parts, d.ctrl, d.before, M, rec_start, rec_off, c, cmax, d.win, win_cap);
        k_dense_reduce<NXB><<<dim3(std::min<unsigned>(pg, 8192), (unsigned)vectors), 256, 0,
                              stream>>>(d.parts, d.ctrl, d.before, d.win, win_cap);
        return;
    }
    const unsigned ag = (unsigned)std::min<int64_t>(sp.block_count * sp.block[2], 8192);
    k_dense_add<NXB><<<dim3(ag, (unsigned)vectors), 256, 0, stream>>>(g, sp, touched, d.info,
                                                                     d.win, win_cap, grid);
}

template <int NXB>
bool launch_dense_mode(SpreadMethod method, const GridParams &g, const SpreadParams &sp,
                       const DenseBins &d, int64_t M, const int4 *rec_start,
                       const float *rec_off, const float *c, const uint32_t *cmax,
                       int64_t vectors, const uint32_t *touched, float *grid, bool add,
                       hipStream_t stream)
{
    switch (method) {
    case SpreadMethod::MfmaF16:
        launch_dense<(int)SpreadMethod::MfmaF16, NXB>(g, sp, d, M, rec_start, rec_off, c,
                                                      cmax, vectors, touched, grid, add,
                                                      stream);
        return true;
    case SpreadMethod::MfmaF16x3:
        launch_dense<(int)SpreadMethod::MfmaF16x3, NXB>(g, sp, d, M, rec_start, rec_off,
                                                        c, cmax, vectors, touched, grid,
                                                        add, stream);
        return true;
    case SpreadMethod::MfmaBf16x3:
        launch_dense<(int)SpreadMethod::MfmaBf16x3, NXB>(g, sp, d, M, rec_start, rec_off,
                                                         c, cmax, vectors, touched, grid,
                                                         add, stream);
        return true;
    default:
        return false;
    }
}

}  // namespace

DenseBins::~DenseBins() { release(); }

void DenseBins::release()
{
    for (void *p : {(void *)before, (void *)light, (void *)info, (void *)parts, (void *)ctrl,
                    (void *)sums, (void *)win, (void *)tile_flag, (void *)touched})
        if (p) (void)hipFree(p);
    before = light = info = ctrl = sums = tile_flag = touched = nullptr;
    tiles_cap = 0;
    parts = nullptr;
    win = nullptr;
    bins_cap = parts_alloc = 0;
    win_bytes = 0;
}

int DenseBins::x_blocks(const GridParams &g)
{
    if (g.dims != 3) return 0;
    for (int a = 0; a < 3; ++a)
        if (g.ax[a].shift > 4 || g.ax[a].n % ((int64_t)1 << g.ax[a].shift) != 0) return 0;
    const int w = g.width;
    if ((1 << g.ax[1].shift) + w - 1 > 8 || (1 << g.ax[2].shift) + w - 1 > 8) return 0;
    // candidate bins of a 16 x 16 x 1 slice of a tile, staged by the add
    // kernel: (15 >> s) + ceil((ext - 1) / side) + 1 per axis (1 cell in z)
    const int cap[3] = {kAddBx, kAddBy, kAddBz};
    const int span[3] = {15, 15, 0};
    for (int a = 0; a < 3; ++a) {
        const int side = 1 << g.ax[a].shift, ext = side + w - 1;
        if ((span[a] >> g.ax[a].shift) + (ext - 1 + side - 1) / side + 1 > cap[a]) return 0;
    }
    const int ex = (1 << g.ax[0].shift) + w - 1;
    const int ey = (1 << g.ax[1].shift) + w - 1;
    const int nxb = ex <= 8 ? 1 : (ex <= 16 ? 2 : 0);
    // x and y candidates of a cell within the add kernel's unrolled loops
    if ((ex + (1 << g.ax[0].shift) - 1) >> g.ax[0].shift > (nxb == 1 ? 4 : 3) ||
        (ey + (1 << g.ax[1].shift) - 1) >> g.ax[1].shift > kAddCy)
        return 0;
    return nxb;
}

bool DenseBins::classify(const GridParams &g, const SpreadParams &sp, const uint32_t *offsets,
                         hipStream_t stream)
{
    const int64_t nbins = g.nbins;
    if (sp.block_count + 1 > tiles_cap) {
        for (uint32_t **p : {&tile_flag, &touched}) {
            if (*p) (void)hipFree(*p);
            *p = nullptr;
            if (!hip_ok(hipMalloc(p, sizeof(uint32_t) * (sp.block_count + 1)), "dense tiles"))
                return false;
        }
        tiles_cap = sp.block_count + 1;
    }
    if (nbins + 1 > bins_cap) {
        for (uint32_t **p : {&before, &light, &info}) {
            if (*p) (void)hipFree(*p);
            *p = nullptr;
            if (!hip_ok(hipMalloc(p, sizeof(uint32_t) * (nbins + 1)), "dense bin arrays"))
                return false;
        }
        if (sums) (void)hipFree(sums);
        sums = nullptr;
        if (!hip_ok(hipMalloc(&sums, sizeof(uint32_t) * (scan_tiles(nbins + 1) + 1)),
                    "dense bin scan"))
            return false;
        bins_cap = nbins + 1;
    }
    if (!ctrl && !hip_ok(hipMalloc(&ctrl, sizeof(uint32_t) * 32), "dense bin control"))
        return false;
    if (part_cap > parts_alloc) {
        if (parts) (void)hipFree(parts);
        parts = nullptr;
        if (!hip_ok(hipMalloc(&parts, sizeof(uint2) * part_cap), "dense bin parts")) return false;
        parts_alloc = part_cap;
    }
    const uint32_t t0 = threshold;
    const uint32_t cap = (uint32_t)part_cap;
    k_dense_clear<<<1, 64, 0, stream>>>(ctrl);
    k_dense_clear_tiles<<<grid_for(sp.block_count, 256), 256, 0, stream>>>(
        tile_flag, sp.block_count, touched);
    k_dense_levels<<<std::min<unsigned>(grid_for(nbins, 256), 2048), 256, 0, stream>>>(
        offsets, nbins, t0, ctrl);
    // cnt lives in `info` until the scan has read it
    k_dense_mark<<<grid_for(nbins + 1, 256), 256, 0, stream>>>(offsets, nbins, t0, cap, ctrl,
                                                               info);
    scan_u32(info, nbins + 1, sums, before, stream);
    k_dense_finish<<<grid_for(nbins + 1, 256), 256, 0, stream>>>(
        offsets, before, nbins, t0, cap, g, sp, ctrl, light, info, parts, tile_flag, touched);
    return hip_ok(hipGetLastError(), "dense bin launch");
}

bool DenseBins::spread(SpreadMethod method, const GridParams &g, const SpreadParams &sp,
                       int64_t M, const int4 *rec_start, const float *rec_off,
                       const float *c, const uint32_t *cmax, int64_t vectors, float *grid,
                       bool add, hipStream_t stream)
{
    const int nxb = x_blocks(g);
    if (nxb == 0) {
        set_error("dense bins need a 3D grid whose bins fit 16 x 8 x 8 windows");
        return false;
    }
    const int64_t need = (part_cap + 1) * window_floats(g) * (int64_t)sizeof(float) * vectors;
    if (need > win_bytes) {
        if (win) (void)hipFree(win);
        win = nullptr;
        if (!hip_ok(hipMalloc(&win, need), "dense windows")) return false;
        win_bytes = need;
    }
    const bool ok = nxb == 1 ? launch_dense_mode<1>(method, g, sp, *this, M, rec_start, rec_off,
                                                    c, cmax, vectors, touched, grid, add, stream)
                             : launch_dense_mode<2>(method, g, sp, *this, M, rec_start, rec_off,
                                                    c, cmax, vectors, touched, grid, add, stream);
    if (!ok) {
        set_error("dense bins need a matrix-core spread method");
        return false;
    }
    return hip_ok(hipGetLastError(), "dense window launch");
}

}  // namespace mnufft
