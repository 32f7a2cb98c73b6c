// Deterministic output-stationary type-1 spreading on gfx950.
//
// Reference behaviour (src/gpu_nd_spread.rs:1-36, shader 1336-1695,
// classify 1704-1771, plan 1781-1905, reduce 1914-2056): every block of
// fine-grid cells is owned by one workgroup whose lanes keep their cells'
// sums in registers; the workgroup walks the bins that can reach the block in
// a fixed order and streams their records through LDS, so every cell sums its
// contributions in a fixed order and is written exactly once -- no float
// atomics.  Blocks whose reach holds more records than a light limit are
// split into parts of consecutive records; partial sums land in a scratch
// and are added in part order.
//
// MI355X design: a block is one 64-lane wavefront (lane = one run of `R`
// cells along axis 0 and one cell along every other axis); records are
// staged 64 at a time, compacted with a ballot so the accumulate loop only
// sees records whose support touches the block, with each record's kernel
// weights computed once by the lane that staged it.  Light blocks and heavy
// parts form one flat list of work items consumed by a grid-stride loop,
// which keeps dense clusters spread over all 256 CUs.
#include <hip/hip_runtime.h>

#include "fold.hpp"
#include "spread.hpp"
#include "spread_device.hpp"

namespace mnufft {

namespace {

constexpr int kLanes = 64;
constexpr int kKw = 16;  // weights per axis staged per record

// Records in reach of every block: one wavefront per block, lanes over the
// block's bin segments.
template <int D>
__global__ __launch_bounds__(256) void k_classify(GridParams g, SpreadParams sp,
                                                  const uint32_t *__restrict__ offsets,
                                                  uint32_t *__restrict__ totals)
{
    const int lane = threadIdx.x & 63;
    const int64_t b = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (b >= sp.block_count) return;
    BlockReach<D> r = block_reach<D>(b, g, sp);
    uint64_t total = 0;
    for (int64_t t = lane; t < r.nseg; t += 64) {
        int64_t f, l;
        r.segment(t, g, offsets, &f, &l);
        total += (uint64_t)(l - f);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) total += __shfl_xor(total, o, 64);
    if (lane == 0) totals[b] = (uint32_t)(total > 0xffffffffull ? 0xffffffffull : total);
}

// One workgroup plans the work items: a light block is one item, a heavy
// block as many parts as its records need.  ctrl[0] = items, ctrl[1] = heavy
// parts, ctrl[2] = heavy blocks, ctrl[3] = records per part.  (One
// instantiation per kernel: each keeps its LDS variables to itself.)
template <int Kernel>
__device__ inline void plan_blocks(int64_t nblocks, const uint32_t *__restrict__ totals,
                                   uint32_t light_limit, uint32_t min_part,
                                   int64_t part_capacity, uint32_t *__restrict__ parts,
                                   uint32_t *__restrict__ item_off,
                                   uint32_t *__restrict__ heavy_base,
                                   uint32_t *__restrict__ heavy_list,
                                   uint32_t *__restrict__ ctrl)
{
    __shared__ uint64_t s_heavy_records;
    __shared__ uint32_t s_heavy_blocks;
    __shared__ uint32_t s_part;
    __shared__ uint32_t s_wsum[16];
    __shared__ uint32_t s_carry[2];
    if (threadIdx.x == 0) {
        s_heavy_records = 0;
        s_heavy_blocks = 0;
        s_carry[0] = s_carry[1] = 0;
    }
    __syncthreads();
    // Raise the light limit until the heavy blocks fill at most half of the
    // partial-sum slots, so every heavy block keeps room for several parts.
    __shared__ uint32_t s_limit;
    if (threadIdx.x == 0) s_limit = light_limit;
    __syncthreads();
    for (;;) {
        const uint32_t limit = s_limit;
        uint64_t mine = 0;
        uint32_t mine_blocks = 0;
        for (int64_t b = threadIdx.x; b < nblocks; b += blockDim.x) {
            uint32_t t = totals[b];
            if (t > limit) {
                mine += t;
                ++mine_blocks;
            }
        }
        atomicAdd((unsigned long long *)&s_heavy_records, (unsigned long long)mine);
        atomicAdd(&s_heavy_blocks, mine_blocks);
        __syncthreads();
        const bool fits = (int64_t)s_heavy_blocks * 2 <= part_capacity || limit >= 0x7fffffffu;
        __syncthreads();
        if (fits) break;
        if (threadIdx.x == 0) {
            s_limit = limit * 2;
            s_heavy_records = 0;
            s_heavy_blocks = 0;
        }
        __syncthreads();
    }
    light_limit = s_limit;
    if (threadIdx.x == 0) {
        uint64_t spare = part_capacity > (int64_t)s_heavy_blocks
                             ? (uint64_t)(part_capacity - s_heavy_blocks)
                             : 1ull;
        uint64_t part = (s_heavy_records + spare - 1) / spare;
        if (part < min_part) part = min_part;
        s_part = (uint32_t)(part > 0xffffffffull ? 0xffffffffull : part);
    }
    __syncthreads();
    uint32_t part = s_part;
    // Ordered exclusive scans of items and heavy parts, 1024 blocks a round.
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int64_t base = 0; base < nblocks; base += blockDim.x) {
        int64_t b = base + threadIdx.x;
        uint32_t np = 0, hp = 0, hb = 0;
        if (b < nblocks) {
            uint32_t t = totals[b];
            np = 1;
            if (t > light_limit) {
                np = (uint32_t)((t + (uint64_t)part - 1) / part);
                // a block that fits one part is written directly (np == 1
                // means light to the tile kernels and the reduction)
                if (np > 1) {
                    hp = np;
                    hb = 1;
                } else {
                    np = 1;
                }
            }
            parts[b] = np;
        }
        // pack (items, heavy parts) into one 64-bit scan via two passes
        for (int pass = 0; pass < 2; ++pass) {
            uint32_t v = pass == 0 ? np : hp;
            uint32_t x = v;
            for (int o = 1; o < 64; o <<= 1) {
                uint32_t y = __shfl_up(x, o, 64);
                if (lane >= o) x += y;
            }
            if (lane == 63) s_wsum[wid] = x;
            __syncthreads();
            uint32_t pre = 0, all = 0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
                if (w < wid) pre += s_wsum[w];
                all += s_wsum[w];
            }
            uint32_t ex = s_carry[pass] + pre + x - v;
            if (b < nblocks) {
                if (pass == 0)
                    item_off[b] = ex;
                else
                    heavy_base[b] = ex;
            }
            __syncthreads();
            if (threadIdx.x == 0) s_carry[pass] += all;
            __syncthreads();
        }
        if (hb) {
            uint32_t slot = atomicAdd(&ctrl[2], 1u);  // ctrl[2] cleared by the host
            heavy_list[slot] = (uint32_t)b;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ctrl[0] = s_carry[0];
        ctrl[1] = s_carry[1];
        ctrl[3] = part;
    }
}

__global__ __launch_bounds__(1024) void k_plan(int64_t nblocks, const uint32_t *__restrict__ totals,
                                               uint32_t light_limit, uint32_t min_part,
                                               int64_t part_capacity,
                                               uint32_t *__restrict__ parts,
                                               uint32_t *__restrict__ item_off,
                                               uint32_t *__restrict__ heavy_base,
                                               uint32_t *__restrict__ heavy_list,
                                               uint32_t *__restrict__ ctrl)
{
    plan_blocks<0>(nblocks, totals, light_limit, min_part, part_capacity, parts, item_off,
                   heavy_base, heavy_list, ctrl);
}

// Many problems: workgroup q plans problem q's blocks alone, with its own
// plan's limits, exactly as that plan's own spread does (its heavy list
// holds its own block indices).  ctrl: 8 words per problem, cleared first.
__global__ __launch_bounds__(1024) void k_plan_many(const SpreadProblem *__restrict__ sprobs,
                                                    uint32_t *__restrict__ totals,
                                                    uint32_t *__restrict__ parts,
                                                    uint32_t *__restrict__ item_off,
                                                    uint32_t *__restrict__ heavy_base,
                                                    uint32_t *__restrict__ heavy_list,
                                                    uint32_t *__restrict__ ctrl)
{
    const SpreadProblem &q = sprobs[blockIdx.x];
    const int64_t bb = q.block_base;
    plan_blocks<1>(q.sp.block_count, totals + bb, q.light_limit, q.min_part, q.part_capacity,
                   parts + bb, item_off + bb, heavy_base + bb, heavy_list + bb,
                   ctrl + 8 * (int64_t)blockIdx.x);
}

// Bases of the problems' work in the combined lists: base[4q] first item,
// [4q + 1] first partial slot, [4q + 2] first heavy block; base[4n..4n+1]
// the totals of items and heavy blocks.
__global__ void k_plan_bases(int nprob, const uint32_t *__restrict__ ctrl,
                             int64_t *__restrict__ base)
{
    if (threadIdx.x != 0) return;
    int64_t items = 0, slots = 0, heavy = 0;
    for (int q = 0; q < nprob; ++q) {
        base[4 * q] = items;
        base[4 * q + 1] = slots;
        base[4 * q + 2] = heavy;
        items += ctrl[8 * q];
        slots += ctrl[8 * q + 1];
        heavy += ctrl[8 * q + 2];
    }
    base[4 * nprob] = items;
    base[4 * nprob + 1] = heavy;
}

// A staged record: wrapped starts, strength, the axis-0 weights padded for
// runs that begin before or after the support, the weights of the other axes.
template <typename T, int R, int D>
struct Staged {
    int s[D];
    T cre, cim;
    T kxp[2 * R + kKw];
    T kr[D > 1 ? D - 1 : 1][kKw];
};

// Wrapped starts of record `slot` (layout of rec_quads<D>).
template <int D>
__device__ inline void load_starts(const int4 *__restrict__ rec_start, int64_t slot, int *ss)
{
    const int4 a = rec_start[rec_quads<D>() * slot];
    ss[0] = a.x;
    if constexpr (D > 1) ss[1] = a.y;
    if constexpr (D > 2) ss[2] = a.z;
    if constexpr (D > 3) {
        ss[3] = a.w;
        const int4 b = rec_start[2 * slot + 1];
        if constexpr (D > 4) ss[4] = b.x;
        if constexpr (D > 5) ss[5] = b.y;
        if constexpr (D > 6) ss[6] = b.z;
        if constexpr (D > 7) ss[7] = b.w;
    }
}

// LDS of one spreading workgroup.
template <int D, typename T, int R>
struct SpreadShared {
    Staged<T, R, D> st[kLanes];
    int64_t seg_first[kLanes];
    int64_t seg_excl[kLanes + 1];
};

// Records [rlo, rhi) of the reach of block b (grid g, blocks sp), strengths
// cv in record-slot order: the lane's cells summed in record order, written
// to the block's cells of grid_v (np == 1) or to the lane's cells of the
// partial slot part_v (a heavy part).
template <int D, typename T, int R>
__device__ __forceinline__ void spread_block(SpreadShared<D, T, R> &sh, const GridParams &g,
                                    const SpreadParams &sp, int64_t b, int64_t rlo, int64_t rhi,
                                    uint32_t np, const uint32_t *__restrict__ offsets,
                                    const int4 *__restrict__ rec_start,
                                    const T *__restrict__ rec_off, const T *__restrict__ cv,
                                    T *__restrict__ grid_v, T *__restrict__ part_v)
{
    constexpr int RS = 4 * rec_quads<D>();  // offsets per record
    Staged<T, R, D> *st = sh.st;
    int64_t *seg_first = sh.seg_first;
    int64_t *seg_excl = sh.seg_excl;
    const int lane = threadIdx.x;
    const int w = g.width;
    const T beta = (T)g.beta;

    // lane -> cells: run of R cells along axis 0 and one cell per other axis
    int lane_x = lane % sp.groups0;
    int rest_lane = lane / sp.groups0;
    int lane_c[D];
    lane_c[0] = lane_x * R;
#pragma unroll
    for (int a = 1; a < D; ++a) {
        lane_c[a] = rest_lane % (int)sp.block[a];
        rest_lane /= (int)sp.block[a];
    }

    BlockReach<D> r = block_reach<D>(b, g, sp);
    int64_t cell[D];
#pragma unroll
    for (int a = 0; a < D; ++a) cell[a] = r.b0[a] + lane_c[a];

    T acc_re[R], acc_im[R];
#pragma unroll
    for (int i = 0; i < R; ++i) acc_re[i] = acc_im[i] = (T)0;

    int64_t cursor = 0;
    for (int64_t sb = 0; sb < r.nseg && cursor < rhi; sb += kLanes) {
        int64_t t = sb + lane;
        int64_t f = 0, l = 0;
        if (t < r.nseg) r.segment(t, g, offsets, &f, &l);
        int64_t cnt = l - f;
        // inclusive wave scan of segment lengths
        int64_t x = cnt;
#pragma unroll
        for (int o = 1; o < kLanes; o <<= 1) {
            int64_t y = __shfl_up(x, o, kLanes);
            if (lane >= o) x += y;
        }
        int64_t batch_total = __shfl(x, kLanes - 1, kLanes);
        seg_first[lane] = f;
        seg_excl[lane] = x - cnt;
        if (lane == 0) seg_excl[kLanes] = batch_total;
        __syncthreads();
        int64_t klo = rlo > cursor ? rlo - cursor : 0;
        int64_t khi = rhi - cursor < batch_total ? rhi - cursor : batch_total;
        for (int64_t k0 = klo; k0 < khi; k0 += kLanes) {
            int64_t k = k0 + lane;
            bool touch = false;
            int64_t slot = 0;
            if (k < khi) {
                int s_lo = 0, s_hi = kLanes - 1;
                while (s_lo < s_hi) {
                    int mid = (s_lo + s_hi + 1) >> 1;
                    if (seg_excl[mid] <= k)
                        s_lo = mid;
                    else
                        s_hi = mid - 1;
                }
                slot = seg_first[s_lo] + (k - seg_excl[s_lo]);
                int ss[D];
                load_starts<D>(rec_start, slot, ss);
                touch = true;
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    int64_t n = g.ax[a].n;
                    int64_t e = r.b0[a] - ss[a];
                    if (e < 0) e += n;
                    touch = touch && (e < w || e > n - sp.block[a]);
                }
            }
            // compact the touching records into LDS
            uint64_t mask = __ballot(touch);
            int pos = __popcll(mask & ((1ull << lane) - 1ull));
            int nstaged = __popcll(mask);
            if (touch) {
                Staged<T, R, D> &S = st[pos];
                int ss[D];
                load_starts<D>(rec_start, slot, ss);
#pragma unroll
                for (int a = 0; a < D; ++a) S.s[a] = ss[a];
                S.cre = cv[2 * slot];
                S.cim = cv[2 * slot + 1];
                T ox = rec_off[RS * slot + 0];
#pragma unroll
                for (int j = 0; j < 2 * R + kKw; ++j) {
                    int tt = j - (R - 1);
                    S.kxp[j] = (tt >= 0 && tt < w) ? es_weight<T>((T)tt + ox, w, beta) : (T)0;
                }
#pragma unroll
                for (int a = 1; a < D; ++a) {
                    T oa = rec_off[RS * slot + a];
                    for (int j = 0; j < w; ++j) S.kr[a - 1][j] = es_weight<T>((T)j + oa, w, beta);
                }
            }
            __syncthreads();
            for (int q = 0; q < nstaged; ++q) {
                const Staged<T, R, D> &S = st[q];
                // product of the other axes' weights, axis 1 first
                T wyz = (T)1;
                bool on = true;
#pragma unroll
                for (int a = 1; a < D; ++a) {
                    int64_t da = cell[a] - S.s[a];
                    if (da < 0) da += g.ax[a].n;
                    on = on && da < w;
                    if (on) wyz = wyz * S.kr[a - 1][da];
                }
                if (!on) continue;
                int64_t nx = g.ax[0].n;
                int64_t e = cell[0] - S.s[0];
                if (e < 0) e += nx;
                int base;
                if (e < w)
                    base = (int)e + R - 1;
                else if (e > nx - R)
                    base = (int)(e - nx) + R - 1;
                else
                    continue;
                T vre = S.cre * wyz, vim = S.cim * wyz;
#pragma unroll
                for (int i = 0; i < R; ++i) {
                    T wx = S.kxp[base + i];
                    acc_re[i] += wx * vre;
                    acc_im[i] += wx * vim;
                }
            }
            __syncthreads();
        }
        cursor += batch_total;
        __syncthreads();
    }

    // write the lane's cells
    if (np == 1) {
        bool inside = true;
        int64_t lin = 0, stride = 1;
#pragma unroll
        for (int a = 1; a < D; ++a) {
            inside = inside && cell[a] < g.ax[a].n;
            stride *= g.ax[a - 1].n;
            lin += cell[a] * stride;
        }
        if (inside) {
            T *out = grid_v + 2 * lin;
#pragma unroll
            for (int i = 0; i < R; ++i) {
                int64_t xc = cell[0] + i;
                if (xc < g.ax[0].n) {
                    out[2 * xc] = acc_re[i];
                    out[2 * xc + 1] = acc_im[i];
                }
            }
        }
    } else {
        T *out = part_v + 2 * ((int64_t)lane * R);
#pragma unroll
        for (int i = 0; i < R; ++i) {
            out[2 * i] = acc_re[i];
            out[2 * i + 1] = acc_im[i];
        }
    }
    __syncthreads();
}

// Record range of work item `item` of the plan: block b, part p of np.
__device__ inline void item_records(int64_t item, int64_t nblocks,
                                    const uint32_t *__restrict__ item_off,
                                    const uint32_t *__restrict__ parts,
                                    const uint32_t *__restrict__ totals, uint32_t part_len_cap,
                                    int64_t *b, int64_t *p, uint32_t *np, int64_t *rlo,
                                    int64_t *rhi)
{
    // block of this item: last b with item_off[b] <= item
    int64_t lo = 0, hi = nblocks - 1;
    while (lo < hi) {
        int64_t mid = (lo + hi + 1) >> 1;
        if ((int64_t)item_off[mid] <= item)
            lo = mid;
        else
            hi = mid - 1;
    }
    *b = lo;
    *p = item - item_off[lo];
    *np = parts[lo];
    const int64_t total = totals[lo];
    *rlo = 0;
    *rhi = total;
    if (*np > 1) {
        const int64_t len = part_len_cap;
        *rlo = *p * len;
        *rhi = *rlo + len < total ? *rlo + len : total;
    }
}

// One work item per loop turn: records [rlo, rhi) of block b's reach, for
// vector v = blockIdx.y; writes the block's cells (light) or a partial slot
// (heavy part).
template <int D, typename T, int R>
__global__ __launch_bounds__(kLanes) void k_spread(GridParams g, SpreadParams sp,
                                                   const uint32_t *__restrict__ offsets,
                                                   const int4 *__restrict__ rec_start,
                                                   const T *__restrict__ rec_off,
                                                   const T *__restrict__ c_sorted, int64_t M,
                                                   const uint32_t *__restrict__ totals,
                                                   const uint32_t *__restrict__ parts,
                                                   const uint32_t *__restrict__ item_off,
                                                   const uint32_t *__restrict__ heavy_base,
                                                   const uint32_t *__restrict__ ctrl,
                                                   T *__restrict__ grid, T *__restrict__ partial,
                                                   int64_t block_cells)
{
    __shared__ SpreadShared<D, T, R> sh;
    const int64_t v = blockIdx.y;
    const int64_t n_items = ctrl[0];
    const uint32_t part_len_cap = ctrl[3];
    const T *cv = c_sorted + 2 * v * M;
    for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        int64_t b, p, rlo, rhi;
        uint32_t np;
        item_records(item, sp.block_count, item_off, parts, totals, part_len_cap, &b, &p, &np,
                     &rlo, &rhi);
        const int64_t slot = (int64_t)heavy_base[b] + p;
        spread_block<D, T, R>(sh, g, sp, b, rlo, rhi, np, offsets, rec_start, rec_off, cv,
                              grid + 2 * v * g.fine_total,
                              partial + 2 * (slot * gridDim.y + v) * block_cells);
    }
}

// Sums of the partial slots of heavy blocks.  Sixteen lanes share a cell,
// lane l adding parts l, l + 16, ...; the sixteen lane sums are then added in
// a fixed pairwise order, so a cell's sum does not depend on timing.  A
// workgroup of 256 takes 16 consecutive (heavy block, cell) units.
constexpr int kPartLanes = 16;

template <typename T>
struct PartSums {
    T re[kPartLanes][kPartLanes + 1];
    T im[kPartLanes][kPartLanes + 1];
};

template <typename T>
__device__ inline void part_lane_sum(const T *__restrict__ partial, int64_t base, int64_t np,
                                     int64_t vectors, int64_t v, int64_t cells, int64_t c, int pl,
                                     T &re, T &im)
{
    for (int64_t q = pl; q < np; q += kPartLanes) {
        const T *src = partial + 2 * (((base + q) * vectors + v) * cells + c);
        re += src[0];
        im += src[1];
    }
}

template <typename T>
__device__ inline void combine_lanes(const PartSums<T> &s, int u, T &re, T &im)
{
    T r[kPartLanes], i[kPartLanes];
#pragma unroll
    for (int k = 0; k < kPartLanes; ++k) {
        r[k] = s.re[k][u];
        i[k] = s.im[k][u];
    }
#pragma unroll
    for (int w = kPartLanes / 2; w >= 1; w >>= 1) {
#pragma unroll
        for (int k = 0; k < w; ++k) {
            r[k] += r[k + w];
            i[k] += i[k + w];
        }
    }
    re = r[0];
    im = i[0];
}

// Cell c (= lane * R + i) of block b: its sum written to grid_v.
template <int D, typename T, int R>
__device__ inline void write_block_cell(const GridParams &g, const SpreadParams &sp, int64_t b,
                                        int64_t c, T re, T im, T *__restrict__ grid_v)
{
    int lane = (int)(c / R), i = (int)(c % R);
    int lane_x = lane % sp.groups0;
    int rl = lane / sp.groups0;
    int64_t rest = b;
    int64_t cell[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        int64_t bi = rest % sp.nblocks[a];
        rest /= sp.nblocks[a];
        cell[a] = bi * sp.block[a];
    }
    cell[0] += lane_x * R + i;
#pragma unroll
    for (int a = 1; a < D; ++a) {
        cell[a] += rl % (int)sp.block[a];
        rl /= (int)sp.block[a];
    }
    bool inside = true;
    int64_t lin = 0, stride = 1;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        inside = inside && cell[a] < g.ax[a].n;
        lin += cell[a] * stride;
        stride *= g.ax[a].n;
    }
    if (inside) {
        T *out = grid_v + 2 * lin;
        out[0] = re;
        out[1] = im;
    }
}

template <int D, typename T, int R>
__global__ __launch_bounds__(256) void k_reduce(GridParams g, SpreadParams sp,
                                                const uint32_t *__restrict__ parts,
                                                const uint32_t *__restrict__ heavy_base,
                                                const uint32_t *__restrict__ heavy_list,
                                                const uint32_t *__restrict__ ctrl,
                                                const T *__restrict__ partial, T *__restrict__ grid,
                                                int64_t block_cells)
{
    __shared__ PartSums<T> ps;
    const int64_t v = blockIdx.y;
    const int64_t vectors = gridDim.y;
    const int64_t nheavy = ctrl[2];
    const int64_t work = nheavy * block_cells;
    const int u = threadIdx.x % kPartLanes, pl = threadIdx.x / kPartLanes;
    for (int64_t t0 = (int64_t)blockIdx.x * kPartLanes; t0 < work;
         t0 += (int64_t)gridDim.x * kPartLanes) {
        const int64_t t = t0 + u;
        int64_t hb = 0, c = 0, b = 0;
        T sre = (T)0, sim = (T)0;
        if (t < work) {
            hb = t / block_cells;
            c = t - hb * block_cells;  // lane * R + i
            b = heavy_list[hb];
            part_lane_sum(partial, (int64_t)heavy_base[b], (int64_t)parts[b], vectors, v,
                          block_cells, c, pl, sre, sim);
        }
        ps.re[pl][u] = sre;
        ps.im[pl][u] = sim;
        __syncthreads();
        if (pl == 0 && t < work) {
            combine_lanes(ps, u, sre, sim);
            write_block_cell<D, T, R>(g, sp, b, c, sre, sim, grid + 2 * v * g.fine_total);
        }
        __syncthreads();
    }
}

// Problems of a multi-plan type-1 spread: the problem of class block gb
// (last problem whose block_base <= gb).
__device__ inline int problem_of_block(const SpreadProblem *__restrict__ sprobs, int nprob,
                                       int64_t gb)
{
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sprobs[mid].block_base <= gb)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// k_classify over the blocks of every problem of a class.
template <int D>
__global__ __launch_bounds__(256) void k_classify_many(const BatchProblem *__restrict__ probs,
                                                       const SpreadProblem *__restrict__ sprobs,
                                                       int nprob, int64_t nblocks,
                                                       const uint32_t *__restrict__ offsets,
                                                       uint32_t *__restrict__ totals)
{
    const int lane = threadIdx.x & 63;
    const int64_t gb = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (gb >= nblocks) return;
    const int q = problem_of_block(sprobs, nprob, gb);
    const BatchProblem &pb = probs[q];
    const SpreadProblem &spb = sprobs[q];
    BlockReach<D> r = block_reach<D>(gb - spb.block_base, pb.g, spb.sp);
    const uint32_t *off = offsets + pb.bin_base;
    uint64_t total = 0;
    for (int64_t t = lane; t < r.nseg; t += 64) {
        int64_t f, l;
        r.segment(t, pb.g, off, &f, &l);
        total += (uint64_t)(l - f);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) total += __shfl_xor(total, o, 64);
    if (lane == 0) totals[gb] = (uint32_t)(total > 0xffffffffull ? 0xffffffffull : total);
}

// k_spread over the work items of every problem of a class (one vector):
// item -> problem (item bases) -> the problem's block and part, summed as
// the problem's own spread sums it.
template <int D, typename T, int R>
__global__ __launch_bounds__(kLanes) void k_spread_many(
    const BatchProblem *__restrict__ probs, const SpreadProblem *__restrict__ sprobs, int nprob,
    const uint32_t *__restrict__ offsets, const int4 *__restrict__ rec_start,
    const T *__restrict__ rec_off, const T *__restrict__ c_sorted,
    const uint32_t *__restrict__ totals, const uint32_t *__restrict__ parts,
    const uint32_t *__restrict__ item_off, const uint32_t *__restrict__ heavy_base,
    const uint32_t *__restrict__ pctrl, const int64_t *__restrict__ pbase,
    T *__restrict__ partial, int64_t block_cells)
{
    __shared__ SpreadShared<D, T, R> sh;
    const int64_t n_items = pbase[4 * nprob];
    for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const int q = batch_problem_of(pbase, nprob, 0, item);
        const BatchProblem &pb = probs[q];
        const SpreadProblem &spb = sprobs[q];
        const int64_t bb = spb.block_base;
        int64_t b, p, rlo, rhi;
        uint32_t np;
        item_records(item - pbase[4 * q], spb.sp.block_count, item_off + bb, parts + bb,
                     totals + bb, pctrl[8 * q + 3], &b, &p, &np, &rlo, &rhi);
        const int64_t slot = pbase[4 * q + 1] + (int64_t)heavy_base[bb + b] + p;
        spread_block<D, T, R>(sh, pb.g, spb.sp, b, rlo, rhi, np, offsets + pb.bin_base, rec_start,
                              rec_off, c_sorted, (T *)pb.grid, partial + 2 * slot * block_cells);
    }
}

// k_reduce over the heavy blocks of every problem of a class (one vector).
template <int D, typename T, int R>
__global__ __launch_bounds__(256) void k_reduce_many(
    const BatchProblem *__restrict__ probs, const SpreadProblem *__restrict__ sprobs, int nprob,
    const uint32_t *__restrict__ parts, const uint32_t *__restrict__ heavy_base,
    const uint32_t *__restrict__ heavy_list, const int64_t *__restrict__ pbase,
    const T *__restrict__ partial, int64_t block_cells)
{
    __shared__ PartSums<T> ps;
    const int64_t work = pbase[4 * nprob + 1] * block_cells;
    const int u = threadIdx.x % kPartLanes, pl = threadIdx.x / kPartLanes;
    for (int64_t t0 = (int64_t)blockIdx.x * kPartLanes; t0 < work;
         t0 += (int64_t)gridDim.x * kPartLanes) {
        const int64_t t = t0 + u;
        int64_t c = 0, b = 0;
        int q = 0;
        T sre = (T)0, sim = (T)0;
        if (t < work) {
            const int64_t hb = t / block_cells;
            c = t - hb * block_cells;
            q = batch_problem_of(pbase, nprob, 2, hb);
            const int64_t bb = sprobs[q].block_base;
            b = heavy_list[bb + (hb - pbase[4 * q + 2])];
            part_lane_sum(partial, pbase[4 * q + 1] + (int64_t)heavy_base[bb + b],
                          (int64_t)parts[bb + b], 1, 0, block_cells, c, pl, sre, sim);
        }
        ps.re[pl][u] = sre;
        ps.im[pl][u] = sim;
        __syncthreads();
        if (pl == 0 && t < work) {
            combine_lanes(ps, u, sre, sim);
            write_block_cell<D, T, R>(probs[q].g, sprobs[q].sp, b, c, sre, sim,
                                      (T *)probs[q].grid);
        }
        __syncthreads();
    }
}

// k_reduce_many for the row tiles (TX x TY x TZ cells) of a batch class.
template <typename T, int TX, int TY, int TZ>
__global__ __launch_bounds__(256) void k_reduce_tiles_many(
    const BatchProblem *__restrict__ probs, const SpreadProblem *__restrict__ sprobs, int nprob,
    const uint32_t *__restrict__ parts, const uint32_t *__restrict__ heavy_base,
    const uint32_t *__restrict__ heavy_list, const int64_t *__restrict__ pbase,
    const T *__restrict__ partial)
{
    constexpr int64_t cells = (int64_t)TX * TY * TZ;
    __shared__ PartSums<T> ps;
    const int64_t work = pbase[4 * nprob + 1] * cells;
    const int u = threadIdx.x % kPartLanes, pl = threadIdx.x / kPartLanes;
    for (int64_t t0 = (int64_t)blockIdx.x * kPartLanes; t0 < work;
         t0 += (int64_t)gridDim.x * kPartLanes) {
        const int64_t t = t0 + u;
        int64_t c = 0, b = 0;
        int q = 0;
        T sre = (T)0, sim = (T)0;
        if (t < work) {
            const int64_t hb = t / cells;
            c = t - hb * cells;
            q = batch_problem_of(pbase, nprob, 2, hb);
            const int64_t bb = sprobs[q].block_base;
            b = heavy_list[bb + (hb - pbase[4 * q + 2])];
            part_lane_sum(partial, pbase[4 * q + 1] + (int64_t)heavy_base[bb + b],
                          (int64_t)parts[bb + b], 1, 0, cells, c, pl, sre, sim);
        }
        ps.re[pl][u] = sre;
        ps.im[pl][u] = sim;
        __syncthreads();
        if (pl == 0 && t < work) {
            combine_lanes(ps, u, sre, sim);
            const GridParams &g = probs[q].g;
            const SpreadParams &sp = sprobs[q].sp;
            const int64_t nb0 = sp.nblocks[0], nb1 = sp.nblocks[1];
            const int64_t gx = (b % nb0) * TX + c % TX;
            const int64_t gy = ((b / nb0) % nb1) * TY + (c / TX) % TY;
            const int64_t gz = (b / (nb0 * nb1)) * TZ + c / (TX * TY);
            const int64_t n0 = g.ax[0].n, n1 = g.ax[1].n;
            if (gx < n0 && gy < n1 && gz < g.ax[2].n) {
                T *o = (T *)probs[q].grid + 2 * (gx + n0 * (gy + n1 * gz));
                o[0] = sre;
                o[1] = sim;
            }
        }
        __syncthreads();
    }
}

// Heavy tiles of the row tiles and the matrix-core tiles: partial sums as in
// k_reduce.
template <typename T, int TX, int TY, int TZ>
__global__ __launch_bounds__(256) void k_reduce_tiles(GridParams g, SpreadParams sp,
                                                      const uint32_t *__restrict__ parts,
                                                      const uint32_t *__restrict__ heavy_base,
                                                      const uint32_t *__restrict__ heavy_list,
                                                      const uint32_t *__restrict__ ctrl,
                                                      const T *__restrict__ partial,
                                                      T *__restrict__ grid)
{
    constexpr int64_t cells = (int64_t)TX * TY * TZ;
    __shared__ PartSums<T> ps;
    const int64_t v = blockIdx.y;
    const int64_t vectors = gridDim.y;
    const int64_t work = (int64_t)ctrl[2] * cells;
    const int u = threadIdx.x % kPartLanes, pl = threadIdx.x / kPartLanes;
    for (int64_t t0 = (int64_t)blockIdx.x * kPartLanes; t0 < work;
         t0 += (int64_t)gridDim.x * kPartLanes) {
        const int64_t t = t0 + u;
        int64_t hb = 0, c = 0, b = 0;
        T sre = (T)0, sim = (T)0;
        if (t < work) {
            hb = t / cells;
            c = t - hb * cells;
            b = heavy_list[hb];
            part_lane_sum(partial, (int64_t)heavy_base[b], (int64_t)parts[b], vectors, v, cells,
                          c, pl, sre, sim);
        }
        ps.re[pl][u] = sre;
        ps.im[pl][u] = sim;
        __syncthreads();
        if (pl == 0 && t < work) {
            combine_lanes(ps, u, sre, sim);
            const int64_t xx = c % TX, yy = (c / TX) % TY, zz = c / (TX * TY);
            const int64_t bx = b % sp.nblocks[0];
            const int64_t by = (b / sp.nblocks[0]) % sp.nblocks[1];
            const int64_t bz = b / (sp.nblocks[0] * sp.nblocks[1]);
            const int64_t gx = bx * TX + xx, gy = by * TY + yy, gz = bz * TZ + zz;
            const int64_t n0 = g.ax[0].n, n1 = g.ax[1].n;
            if (gx < n0 && gy < n1 && gz < g.ax[2].n) {
                T *o = grid + 2 * (v * g.fine_total + gx + n0 * (gy + n1 * gz));
                o[0] = sre;
                o[1] = sim;
            }
        }
        __syncthreads();
    }
}

inline unsigned grid_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

__global__ void k_fill_ctrl(uint32_t *ctrl)
{
    if (threadIdx.x < 8) ctrl[threadIdx.x] = 0;
}

// Tile kernels and heavy-part reductions for one chunk of vectors (gridDim.y
// = chunk; c_sorted and grid point at the chunk's first vector).  The work
// plan (totals, parts, item offsets) is shared by every chunk.
inline dim3 items_grid(int64_t nb, int64_t vectors)
{
    // enough workgroups to fill the GPU several times over; each loops over
    // items with a stride
    return dim3((unsigned)std::min<int64_t>(nb + 16384, 65536), (unsigned)vectors);
}

template <int D, typename T, int R>
bool chunk_spread(Spreader &s, const GridParams &g, const SpreadParams &sp,
                  const uint32_t *offsets, const int4 *rec_start, const T *rec_off,
                  const T *c_sorted, int64_t M, int64_t vectors, T *grid, hipStream_t stream)
{
    const int64_t cells = (int64_t)kLanes * R;
    k_spread<D, T, R><<<items_grid(sp.block_count, vectors), kLanes, 0, stream>>>(
        g, sp, offsets, rec_start, rec_off, c_sorted, M, s.totals, s.parts, s.item_off,
        s.heavy_base, s.ctrl, grid, (T *)s.partial, cells);
    if (s.kernel_end) (void)hipEventRecord(s.kernel_end, stream);
    k_reduce<D, T, R><<<dim3(2048, (unsigned)vectors), 256, 0, stream>>>(
        g, sp, s.parts, s.heavy_base, s.heavy_list, s.ctrl, (const T *)s.partial, grid, cells);
    return hip_ok(hipGetLastError(), "spread launch");
}

template <typename T>
bool chunk_rows(Spreader &s, const GridParams &g, const SpreadParams &sp, const uint32_t *offsets,
                const int4 *rec_start, const T *rec_off, const T *c_sorted, int64_t M,
                int64_t vectors, T *grid, hipStream_t stream)
{
    const dim3 ig = items_grid(sp.block_count, vectors);
    TileWork w{s.totals, s.parts, s.item_off, s.heavy_base, s.ctrl, nullptr, ig.x};
    if (!launch_rows_tiles<T>(g, sp, offsets, rec_start, rec_off, c_sorted, M, vectors, w, grid,
                              (T *)s.partial, stream))
        return false;
    if (s.kernel_end) (void)hipEventRecord(s.kernel_end, stream);
    const dim3 rg(2048, (unsigned)vectors);
    if (g.dims == 2)
        k_reduce_tiles<T, 16, 16, 1><<<rg, 256, 0, stream>>>(g, sp, s.parts, s.heavy_base,
                                                             s.heavy_list, s.ctrl,
                                                             (const T *)s.partial, grid);
    else
        k_reduce_tiles<T, 8, 8, 8><<<rg, 256, 0, stream>>>(g, sp, s.parts, s.heavy_base,
                                                           s.heavy_list, s.ctrl,
                                                           (const T *)s.partial, grid);
    return hip_ok(hipGetLastError(), "row-tile spread launch");
}

template <int D>
bool chunk_mfma(Spreader &s, SpreadMethod method, const GridParams &g, const SpreadParams &sp,
                const uint32_t *offsets, const int4 *rec_start, const float *rec_off,
                const float *c_sorted, int64_t M, int64_t vectors, const uint32_t *vmax,
                float *grid, bool mark_end, hipStream_t stream)
{
    constexpr int TE = D == 3 ? 16 : 32;  // tile edge along x and y
    const dim3 ig = items_grid(sp.block_count, vectors);
    TileWork w{s.totals, s.parts, s.item_off, s.heavy_base, s.ctrl, vmax, ig.x};
    if (!launch_mfma_tiles(method, D, g, sp, offsets, rec_start, rec_off, c_sorted, M, vectors, w,
                           grid, (float *)s.partial, stream))
        return false;
    if (mark_end && s.kernel_end) (void)hipEventRecord(s.kernel_end, stream);
    const dim3 rg(2048, (unsigned)vectors);
    if (D == 2)
        k_reduce_tiles<float, TE, TE, 1><<<rg, 256, 0, stream>>>(
            g, sp, s.parts, s.heavy_base, s.heavy_list, s.ctrl, (const float *)s.partial, grid);
    else
        k_reduce_tiles<float, TE, TE, 8><<<rg, 256, 0, stream>>>(
            g, sp, s.parts, s.heavy_base, s.heavy_list, s.ctrl, (const float *)s.partial, grid);
    return hip_ok(hipGetLastError(), "matrix-core spread launch");
}

__global__ void k_fill_u32s(uint32_t *p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

}  // namespace

Spreader::~Spreader() { release(); }

void Spreader::release()
{
    for (void *p : {(void *)totals, (void *)parts, (void *)item_off, (void *)heavy_base,
                    (void *)heavy_list, (void *)ctrl, (void *)vmax, (void *)pctrl, (void *)pbase,
                    partial})
        if (p) (void)hipFree(p);
    totals = parts = item_off = heavy_base = heavy_list = ctrl = vmax = pctrl = nullptr;
    pbase = nullptr;
    problems_cap = 0;
    partial = nullptr;
    blocks_cap = 0;
    partial_bytes = 0;
    vmax_cap = 0;
}

bool Spreader::ensure(int64_t nblocks, int64_t partial_need, int64_t vectors)
{
    if (vectors > vmax_cap) {
        if (vmax) (void)hipFree(vmax);
        vmax = nullptr;
        if (!hip_ok(hipMalloc(&vmax, sizeof(uint32_t) * vectors), "strength ranges")) return false;
        vmax_cap = vectors;
    }
    if (nblocks > blocks_cap) {
        for (uint32_t **p : {&totals, &parts, &item_off, &heavy_base, &heavy_list}) {
            if (*p) (void)hipFree(*p);
            *p = nullptr;
            if (!hip_ok(hipMalloc(p, sizeof(uint32_t) * (nblocks + 1)), "spread block arrays"))
                return false;
        }
        if (!ctrl && !hip_ok(hipMalloc(&ctrl, sizeof(uint32_t) * 64), "spread control"))
            return false;
        blocks_cap = nblocks;
        ++blocks_epoch;
    }
    if (partial_need > partial_bytes) {
        if (partial) (void)hipFree(partial);
        partial = nullptr;
        if (!hip_ok(hipMalloc(&partial, partial_need), "spread partial sums")) return false;
        partial_bytes = partial_need;
    }
    return true;
}

template <int D, typename T>
bool Spreader::spread(const GridParams &g, const SpreadParams &sp, const uint32_t *offsets,
                      const int4 *rec_start, const T *rec_off, const T *c_sorted, int64_t M,
                      int64_t vectors, T *grid, hipStream_t stream, SpreadMethod method,
                      DenseBins *dense)
{
    const bool tiles = sp.run == 0;
    const bool mfma = method != SpreadMethod::Plain;
    if (vectors > 65535) {
        set_error("at most 65535 vectors per execution");
        return false;
    }
    if (mfma && !(D == 3 && sizeof(T) == 4 && tiles)) {
        // the 2D matrix-core tiles are switched off: an open GPU memory fault
        set_error("matrix-core spread needs an FP32 3D plan on row tiles");
        return false;
    }
    if (!mfma && tiles && D != 3 && D != 2) {
        set_error("plain row tiles need a 2D or 3D plan");
        return false;
    }
    if (!mfma && tiles &&
        (g.width < 2 || g.width > (D == 3 && sizeof(T) == 8 ? kMaxRowsWidth : 8))) {
        set_error("plain row tiles need a kernel width of 2..8 (2..16 for 3D FP64)");
        return false;
    }
    if (!mfma && !tiles && sp.run != 2 && sp.run != 4 && sp.run != 8 &&
        !(D >= 4 && sp.run == 16)) {
        set_error("unsupported spread run");
        return false;
    }
    // The work plan depends on the records only, never on the number of
    // vectors: part slots are sized for one vector, and the vectors are
    // spread in chunks that share the plan, so a batch sums every vector
    // exactly as a single-vector call does.
    const int64_t cells =
        tiles ? sp.block[0] * sp.block[1] * sp.block[2] : (int64_t)kLanes * sp.run;
    const int64_t per_part = cells * 2 * (int64_t)sizeof(T);
    part_capacity = std::max<int64_t>(budget_bytes / per_part, 64);
    if (dense && !(mfma && D == 3 && sizeof(T) == 4 && tiles && DenseBins::x_blocks(g))) {
        set_error("dense bins need the 3D matrix-core spread");
        return false;
    }
    int64_t chunk =
        std::max<int64_t>(1, std::min<int64_t>(vectors, chunk_budget_bytes /
                                                            (part_capacity * per_part)));
    if (dense)
        chunk = std::max<int64_t>(
            1, std::min<int64_t>(chunk, chunk_budget_bytes / (dense->part_cap *
                                                              DenseBins::window_floats(g) *
                                                              (int64_t)sizeof(float))));
    if (!ensure(sp.block_count, part_capacity * per_part * chunk, vectors)) return false;

    const int64_t nb = sp.block_count;
    const bool reuse = reuse_plan && planned_epoch == blocks_epoch && planned_blocks == nb &&
                       planned_capacity == part_capacity && planned_light == light_limit &&
                       planned_min == min_part && planned_offsets == offsets;
    reuse_plan = false;
    if (!reuse) {
        k_classify<D><<<grid_for(nb * 64, 256), 256, 0, stream>>>(g, sp, offsets, totals);
        k_fill_ctrl<<<1, 64, 0, stream>>>(ctrl);
        k_plan<<<1, 1024, 0, stream>>>(nb, totals, light_limit, min_part, part_capacity, parts,
                                       item_off, heavy_base, heavy_list, ctrl);
        planned_epoch = blocks_epoch;
        planned_blocks = nb;
        planned_capacity = part_capacity;
        planned_light = light_limit;
        planned_min = min_part;
        planned_offsets = offsets;
    }
    if constexpr (sizeof(T) == 4) {
        if (mfma) {
            k_fill_u32s<<<grid_for(vectors, 256), 256, 0, stream>>>(vmax, vectors);
            // the largest magnitude of the caller's vectors (with dense bins
            // only the light records are gathered)
            if (!launch_absmax(dense ? dense->strengths : c_sorted, 2 * M, vectors, vmax,
                               stream))
                return false;
        }
    }
    if (kernel_start) (void)hipEventRecord(kernel_start, stream);
    for (int64_t v0 = 0; v0 < vectors; v0 += chunk) {
        const int64_t vc = std::min<int64_t>(chunk, vectors - v0);
        const T *cs = c_sorted + 2 * v0 * M;
        T *gr = grid + 2 * v0 * g.fine_total;
        bool ok = false;
        if (mfma) {
            if constexpr ((D == 2 || D == 3) && sizeof(T) == 4) {
                // dense windows first, so the grid the tile kernels write is
                // still in the caches for the window adds and the FFT
                ok = !dense || dense->spread(method, g, sp, M, rec_start, rec_off,
                                             dense->strengths + 2 * v0 * M, vmax + v0, vc, gr,
                                             false, stream);
                ok = ok && chunk_mfma<D>(*this, method, g, sp, offsets, rec_start, rec_off, cs,
                                         M, vc, vmax + v0, gr, dense == nullptr, stream);
                if (ok && dense)
                    ok = dense->spread(method, g, sp, M, rec_start, rec_off,
                                       dense->strengths + 2 * v0 * M, vmax + v0, vc, gr, true,
                                       stream);
                if (ok && dense && kernel_end) (void)hipEventRecord(kernel_end, stream);
            }
        } else if (tiles) {
            if constexpr (D == 2 || D == 3)
                ok = chunk_rows<T>(*this, g, sp, offsets, rec_start, rec_off, cs, M, vc, gr,
                                   stream);
        } else {
            switch (sp.run) {
            case 2: ok = chunk_spread<D, T, 2>(*this, g, sp, offsets, rec_start, rec_off, cs, M, vc, gr, stream); break;
            case 4: ok = chunk_spread<D, T, 4>(*this, g, sp, offsets, rec_start, rec_off, cs, M, vc, gr, stream); break;
            case 8: ok = chunk_spread<D, T, 8>(*this, g, sp, offsets, rec_start, rec_off, cs, M, vc, gr, stream); break;
            case 16:
                if constexpr (D >= 4)
                    ok = chunk_spread<D, T, 16>(*this, g, sp, offsets, rec_start, rec_off, cs, M, vc, gr, stream);
                break;
            default: break;
            }
        }
        if (!ok) return false;
    }
    return true;
}

template <int D, typename T>
bool Spreader::spread_many(const BatchProblem *probs, const SpreadProblem *sprobs,
                           const std::vector<SpreadClass> &classes, const uint32_t *offsets,
                           const int4 *rec_start, const T *rec_off, const T *c_sorted,
                           hipStream_t stream)
{
    planned_epoch = ~0ull;  // the plan buffers take the batch's plans
    reuse_plan = false;
    // scratch for the largest class first: nothing is freed between launches
    int64_t most_blocks = 1, most_partial = 0, most_problems = 1;
    // the row tiles: 16 x 16 cells (2D), 8 x 8 x 8 (3D)
    constexpr int64_t kTileCells = D == 2 ? 16 * 16 : 8 * 8 * 8;
    constexpr int kTileWidthMax = D == 3 && sizeof(T) == 8 ? kMaxRowsWidth : 8;
    for (const SpreadClass &k : classes) {
        const int R = k.run;
        const bool tiles = R == 0;
        if (tiles ? !((D == 2 || D == 3) && k.width >= 2 && k.width <= kTileWidthMax)
                  : !(R == 2 || R == 4 || R == 8 || (D >= 4 && R == 16))) {
            set_error(tiles ? "batched spread: row tiles need 2D or 3D plans of width 2..8 "
                              "(2..16 for 3D FP64)"
                            : "batched spread: unsupported run");
            return false;
        }
        const int64_t per_part =
            (tiles ? kTileCells : (int64_t)kLanes * R) * 2 * (int64_t)sizeof(T);
        most_blocks = std::max(most_blocks, k.blocks);
        most_partial = std::max(most_partial, std::max<int64_t>(k.slots, 1) * per_part);
        most_problems = std::max<int64_t>(most_problems, k.p1 - k.p0);
    }
    if (!ensure(most_blocks, most_partial, 1)) return false;
    if (most_problems > problems_cap) {
        if (pctrl) (void)hipFree(pctrl);
        if (pbase) (void)hipFree(pbase);
        pctrl = nullptr;
        pbase = nullptr;
        problems_cap = 0;
        if (!hip_ok(hipMalloc(&pctrl, sizeof(uint32_t) * 8 * most_problems),
                    "batched spread counters") ||
            !hip_ok(hipMalloc(&pbase, sizeof(int64_t) * (4 * most_problems + 4)),
                    "batched spread bases"))
            return false;
        problems_cap = most_problems;
    }
    for (const SpreadClass &k : classes) {
        const int R = k.run;
        const int64_t nb = k.blocks;
        const int n = k.p1 - k.p0;
        if (nb == 0 || n == 0) continue;
        const int64_t cells = R == 0 ? kTileCells : (int64_t)kLanes * R;
        const BatchProblem *pb = probs + k.p0;
        const SpreadProblem *sb = sprobs + k.p0;
        k_classify_many<D><<<grid_for(nb * 64, 256), 256, 0, stream>>>(pb, sb, n, nb, offsets,
                                                                      totals);
        k_fill_u32s<<<grid_for(8 * n, 256), 256, 0, stream>>>(pctrl, 8 * n);
        k_plan_many<<<(unsigned)n, 1024, 0, stream>>>(sb, totals, parts, item_off, heavy_base,
                                                      heavy_list, pctrl);
        k_plan_bases<<<1, 64, 0, stream>>>(n, pctrl, pbase);
        const dim3 ig = items_grid(nb, 1);
        const dim3 rg(2048, 1);
#define MNUFFT_SPREAD_MANY(RR)                                                                    \
    do {                                                                                          \
        k_spread_many<D, T, RR><<<ig, kLanes, 0, stream>>>(pb, sb, n, offsets, rec_start,         \
                                                           rec_off, c_sorted, totals, parts,      \
                                                           item_off, heavy_base, pctrl, pbase,    \
                                                           (T *)partial, cells);                  \
        k_reduce_many<D, T, RR><<<rg, 256, 0, stream>>>(pb, sb, n, parts, heavy_base, heavy_list, \
                                                        pbase, (const T *)partial, cells);        \
    } while (0)
        switch (R) {
        case 0:
            if constexpr (D == 2) {
                if (!launch_rows2_many<T>(k.width, pb, sb, n, ig.x, offsets, rec_start, rec_off,
                                          c_sorted, totals, parts, item_off, heavy_base, pctrl,
                                          pbase, (T *)partial, stream))
                    return false;
                k_reduce_tiles_many<T, 16, 16, 1><<<rg, 256, 0, stream>>>(
                    pb, sb, n, parts, heavy_base, heavy_list, pbase, (const T *)partial);
            } else if constexpr (D == 3) {
                if (!launch_rows3_many<T>(k.width, pb, sb, n, ig.x, offsets, rec_start, rec_off,
                                          c_sorted, totals, parts, item_off, heavy_base, pctrl,
                                          pbase, (T *)partial, stream))
                    return false;
                k_reduce_tiles_many<T, 8, 8, 8><<<rg, 256, 0, stream>>>(
                    pb, sb, n, parts, heavy_base, heavy_list, pbase, (const T *)partial);
            }
            break;
        case 2: MNUFFT_SPREAD_MANY(2); break;
        case 4: MNUFFT_SPREAD_MANY(4); break;
        case 8: MNUFFT_SPREAD_MANY(8); break;
        default:
            if constexpr (D >= 4) MNUFFT_SPREAD_MANY(16);
            break;
        }
#undef MNUFFT_SPREAD_MANY
        if (!hip_ok(hipGetLastError(), "batched spread launch")) return false;
    }
    return true;
}

#define MNUFFT_SPREAD_INSTANTIATE(D, T)                                                           \
    template bool Spreader::spread<D, T>(const GridParams &, const SpreadParams &,                 \
                                         const uint32_t *, const int4 *, const T *, const T *,     \
                                         int64_t, int64_t, T *, hipStream_t, SpreadMethod,         \
                                         DenseBins *);
#define MNUFFT_SPREAD_MANY_INSTANTIATE(D, T)                                                      \
    template bool Spreader::spread_many<D, T>(const BatchProblem *, const SpreadProblem *,        \
                                              const std::vector<SpreadClass> &, const uint32_t *, \
                                              const int4 *, const T *, const T *, hipStream_t);
MNUFFT_SPREAD_MANY_INSTANTIATE(1, float)
MNUFFT_SPREAD_MANY_INSTANTIATE(2, float)
MNUFFT_SPREAD_MANY_INSTANTIATE(3, float)
MNUFFT_SPREAD_MANY_INSTANTIATE(1, double)
MNUFFT_SPREAD_MANY_INSTANTIATE(2, double)
MNUFFT_SPREAD_MANY_INSTANTIATE(3, double)
MNUFFT_SPREAD_INSTANTIATE(1, float)
MNUFFT_SPREAD_INSTANTIATE(2, float)
MNUFFT_SPREAD_INSTANTIATE(3, float)
MNUFFT_SPREAD_INSTANTIATE(1, double)
MNUFFT_SPREAD_INSTANTIATE(2, double)
MNUFFT_SPREAD_INSTANTIATE(3, double)
MNUFFT_SPREAD_INSTANTIATE(4, float)
MNUFFT_SPREAD_INSTANTIATE(5, float)
MNUFFT_SPREAD_INSTANTIATE(6, float)
MNUFFT_SPREAD_INSTANTIATE(7, float)
MNUFFT_SPREAD_INSTANTIATE(8, float)
MNUFFT_SPREAD_INSTANTIATE(4, double)
MNUFFT_SPREAD_INSTANTIATE(5, double)
MNUFFT_SPREAD_INSTANTIATE(6, double)
MNUFFT_SPREAD_INSTANTIATE(7, double)
MNUFFT_SPREAD_INSTANTIATE(8, double)

}  // namespace mnufft
