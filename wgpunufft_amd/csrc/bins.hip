// Point binning for the MI355X engine: radix sort by bin key -> bin offsets
// -> in-bin order -> prepared records.
//
// Reference behaviour (src/gpu_nd_bins.rs:1-33, src/gpu_bin_sort.rs:1-56):
// points are counted into bins of support starts with integer atomics, the
// counts are scanned into bin offsets (plus the total), point indices are
// scattered into their bins, and the order inside every bin is then made
// deterministic.  Index order restores ascending original index (the
// reference permutation); canonical order sorts by coordinate bits so the
// type-1 sums do not depend on the input order of the points.
//
// MI355X design: no global atomics on the bin counts.  The points are
// radix-sorted by bin key (least significant 8-bit digit first, per-tile
// digit histograms, a scan of the digit-major histogram table, tile-ranked
// scatters staged through LDS); the bin offsets then come from the sorted
// keys, every offset written by the one position where its bin starts; bins
// above 64 points are sorted into canonical order by a tiled merge sort over
// all workgroups (LDS bitonic runs, merge-path rounds).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <type_traits>

#include "bins.hpp"
#include "fold.hpp"

namespace mnufft {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;  // 4096
constexpr int kSortThreads = 256;
constexpr int64_t kSelfScanTiles = 64;  // radix tiles up to which a scatter scans for itself
constexpr int kRunThreads = 1024;  // LDS bitonic runs: one thread per pair
constexpr int kSortRun = 2048;

// ---------------------------------------------------------------------
// Exclusive scan of n u32 counts (reduce, top-level scan, down-sweep).
// ---------------------------------------------------------------------
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t *total)
{
    __shared__ uint32_t warp_sums[kScanThreads / 64];
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) warp_sums[wid] = x;
    __syncthreads();
    uint32_t base = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) {
        uint32_t s = warp_sums[w];
        if (w < wid) base += s;
        all += s;
    }
    __syncthreads();
    *total = all;
    return base + x - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t *__restrict__ in,
                                                              int64_t n,
                                                              uint32_t *__restrict__ tile_sums)
{
    int64_t base = (int64_t)blockIdx.x * kScanTile;
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        int64_t k = base + (int64_t)i * kScanThreads + threadIdx.x;
        if (k < n) s += in[k];
    }
    uint32_t total;
    block_exclusive_scan(s, &total);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Second half of the scan: every tile adds the sums of the tiles before it
// (summed here, a few thousand at most) to its own exclusive scan.  out may
// alias in: every entry is read and written by the same thread.
__global__ __launch_bounds__(kScanThreads) void k_scan_down(const uint32_t *in, int64_t n,
                                                            const uint32_t *__restrict__ tile_sums,
                                                            uint32_t *out)
{
    uint32_t before = 0;
    for (uint32_t t = threadIdx.x; t < blockIdx.x; t += kScanThreads) before += tile_sums[t];
    uint32_t carry;
    block_exclusive_scan(before, &carry);
    int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        int64_t k = base + i;
        v[i] = k < n ? in[k] : 0u;
        s += v[i];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(s, &total) + carry;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        int64_t k = base + i;
        if (k < n) out[k] = run;
        run += v[i];
    }
}

// ---------------------------------------------------------------------
// Least-significant-digit radix sort of the points by bin key, 8 bits per
// pass, stable, with no global atomics: every pass is a per-tile digit
// histogram, an exclusive scan of the digit-major histogram table, and a
// scatter in which every tile ranks its keys per digit in LDS and writes
// them out digit by digit (contiguous runs per digit).  Pass 0 computes the
// bin keys from the coordinates; the value of a point is its index, so
// stability leaves every bin in ascending index order (the reference
// permutation).
// ---------------------------------------------------------------------
constexpr int kRadixThreads = 256;
constexpr int kRadixWaves = kRadixThreads / 64;
constexpr int kRadixItems = 16;
constexpr int kRadixTile = kRadixThreads * kRadixItems;  // 4096 keys
constexpr int kDigits = 256;
// Sort keys below the bin key: none, the in-bin start cell, the quantized
// in-bin position.
constexpr int kKeyBin = 0, kKeyCell = 1, kKeyPosition = 2;

// Lanes of the wavefront whose digit equals this lane's (valid lanes only).
__device__ inline uint64_t digit_peers(uint32_t d, bool valid)
{
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

// Point j of a tile: item i of lane `lane` in wavefront w (wavefront-major,
// then item, then lane), the order in which ranks are handed out.
__device__ inline int64_t radix_pos(int64_t t0, int w, int i, int lane)
{
    return t0 + (int64_t)w * (kRadixItems * 64) + (int64_t)i * 64 + lane;
}

// Adds the wavefront's digits of one item to its LDS counters; returns each
// lane's rank among the wavefront's earlier keys of the same digit.
__device__ inline uint32_t wave_rank_digit(uint32_t d, bool valid, int lane,
                                           uint32_t *__restrict__ wcnt)
{
    const uint64_t peers = digit_peers(d, valid);
    const uint32_t below = (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
    // LDS instructions of a wavefront execute in order: every lane reads the
    // counter before the digit's first lane writes it back
    const uint32_t base = wcnt[d];
    __builtin_amdgcn_wave_barrier();
    __asm__ volatile("" ::: "memory");
    if (valid && below == 0) wcnt[d] = base + (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
    __asm__ volatile("" ::: "memory");
    return base + below;
}

__device__ inline void radix_clear(uint32_t (*wcnt)[kDigits])
{
    for (int i = threadIdx.x; i < kRadixWaves * kDigits; i += kRadixThreads)
        (&wcnt[0][0])[i] = 0u;
    __syncthreads();
}

// The tile's count of every digit into the digit-major table (histogram
// passes count with LDS atomics; the scatter ranks with wavefront matches,
// which keeps it stable).
__device__ inline void radix_store_hist(uint32_t (*wcnt)[kDigits], uint32_t *__restrict__ hist,
                                        int64_t ntiles)
{
    __syncthreads();
    const int d = threadIdx.x;
    uint32_t s = 0;
#pragma unroll
    for (int w = 0; w < kRadixWaves; ++w) s += wcnt[w][d];
    hist[(int64_t)d * ntiles + blockIdx.x] = s;
}

// Sort key of a point below its bin key (sub_bits per axis, axis 0 lowest):
// grouped order, the support-start cell inside the bin (sub_bits = the bin
// side's bits); canonical order, the position inside the bin quantized to
// 2^sub_bits steps per axis, a function of the coordinates that splits dense
// bins into small groups before the coordinate-bit order is applied.
template <int D, typename T>
__device__ inline uint64_t sort_key_of(const T *x, const GridParams &g, int mode, int sub_bits)
{
    bool reach = true;
#pragma unroll
    for (int a = 0; a < D; ++a) reach = reach && FoldTraits<T>::in_reach(x[a]);
    uint32_t key = 0;
    uint64_t sub = 0;
    int at = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        int64_t s;
        double off;
        support_start<T>(x[a], reach, g, a, &s, &off);
        const uint32_t ws = (uint32_t)wrapped_start(s, g.ax[a].n);
        key += (ws >> g.ax[a].shift) * (uint32_t)g.ax[a].stride;
        if (mode == kKeyCell) {
            sub |= (uint64_t)(ws & ((1u << g.ax[a].shift) - 1u)) << at;
        } else if (mode == kKeyPosition) {
            // position in the bin in units of the bin side, in (0, 1]
            const double side = (double)(1u << g.ax[a].shift);
            const double frac = 1.0 - off - g.half_width;  // position - (start - 1) - w/2
            const double u = ((double)(ws & ((1u << g.ax[a].shift) - 1u)) + frac) / side;
            double q = floor(u * (double)(1u << sub_bits));
            q = q < 0.0 ? 0.0 : (q > (double)((1u << sub_bits) - 1u) ? (double)((1u << sub_bits) - 1u) : q);
            sub |= (uint64_t)q << at;
        }
        at += mode == kKeyCell ? g.ax[a].shift : sub_bits;
    }
    return mode == kKeyBin ? (uint64_t)key : ((uint64_t)key << at) | sub;
}

// Pass 0 histogram: sort key of every point (written out) and its low
// digit.
template <int D, typename T, typename K>
__global__ __launch_bounds__(kRadixThreads) void k_radix_keys(int64_t M,
                                                              const T *__restrict__ coords,
                                                              GridParams g, int mode,
                                                              int sub_bits, K *__restrict__ keys,
                                                              uint32_t *__restrict__ hist,
                                                              int64_t ntiles)
{
    __shared__ uint32_t wcnt[kRadixWaves][kDigits];
    radix_clear(wcnt);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t t0 = (int64_t)blockIdx.x * kRadixTile;
    // coordinates of 8 items in flight at a time
    constexpr int kGroup = 8;
    for (int i0 = 0; i0 < kRadixItems; i0 += kGroup) {
        T x[kGroup][D];
#pragma unroll
        for (int i = 0; i < kGroup; ++i) {
            const int64_t j = radix_pos(t0, w, i0 + i, lane);
#pragma unroll
            for (int a = 0; a < D; ++a) x[i][a] = j < M ? coords[j * D + a] : (T)0;
        }
#pragma unroll
        for (int i = 0; i < kGroup; ++i) {
            const int64_t j = radix_pos(t0, w, i0 + i, lane);
            const bool valid = j < M;
            if (valid) {
                const K key = (K)sort_key_of<D, T>(x[i], g, mode, sub_bits);
                keys[j] = key;
                atomicAdd(&wcnt[w][(uint32_t)key & (kDigits - 1)], 1u);
            }
        }
    }
    radix_store_hist(wcnt, hist, ntiles);
}

// Pass 0 of a multi-problem sort: the key of point j (of the concatenated
// point set) is its problem's first bin plus its bin in the problem's grid.
template <int D, typename T>
__global__ __launch_bounds__(kRadixThreads) void k_radix_keys_many(
    int64_t M, const BatchProblem *__restrict__ probs, int nprob, uint32_t *__restrict__ keys,
    uint32_t *__restrict__ hist, int64_t ntiles)
{
    __shared__ uint32_t wcnt[kRadixWaves][kDigits];
    radix_clear(wcnt);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t t0 = (int64_t)blockIdx.x * kRadixTile;
    // the problem of the lane's first point by a binary search, later
    // points (increasing j) by stepping forward
    int lo = 0;
    {
        const int64_t j = radix_pos(t0, w, 0, lane);
        int hi = nprob - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (probs[mid].m_off <= j)
                lo = mid;
            else
                hi = mid - 1;
        }
    }
    for (int i = 0; i < kRadixItems; ++i) {
        const int64_t j = radix_pos(t0, w, i, lane);
        if (j >= M) continue;
        while (lo + 1 < nprob && probs[lo + 1].m_off <= j) ++lo;
        const BatchProblem &pb = probs[lo];
        const T *c = (const T *)pb.coords + (j - pb.m_off) * D;
        T x[D];
#pragma unroll
        for (int a = 0; a < D; ++a) x[a] = c[a];
        const uint32_t key = (uint32_t)(pb.bin_base + sort_key_of<D, T>(x, pb.g, kKeyBin, 0));
        keys[j] = key;
        atomicAdd(&wcnt[w][key & (kDigits - 1)], 1u);
    }
    radix_store_hist(wcnt, hist, ntiles);
}

// The 32-bit word of a key that holds the digit at `shift`, and the digit.
template <typename K>
__device__ inline uint32_t key_digit(K k, int shift)
{
    return (uint32_t)(k >> shift) & (kDigits - 1);
}

// Histogram of the digit at `shift` for passes after the first.
template <typename K>
__global__ __launch_bounds__(kRadixThreads) void k_radix_hist(int64_t M,
                                                              const K *__restrict__ keys,
                                                              int shift,
                                                              uint32_t *__restrict__ hist,
                                                              int64_t ntiles)
{
    __shared__ uint32_t wcnt[kRadixWaves][kDigits];
    radix_clear(wcnt);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t t0 = (int64_t)blockIdx.x * kRadixTile;
    // only the 32-bit word that holds the digit is read
    const uint32_t *kw = (const uint32_t *)keys;
    constexpr int kWords = sizeof(K) / 4;
    const int word = shift >> 5, wshift = shift & 31;
    uint32_t k[kRadixItems];
#pragma unroll
    for (int i = 0; i < kRadixItems; ++i) {
        const int64_t j = radix_pos(t0, w, i, lane);
        k[i] = j < M ? kw[j * kWords + word] : 0u;
    }
#pragma unroll
    for (int i = 0; i < kRadixItems; ++i)
        if (radix_pos(t0, w, i, lane) < M) atomicAdd(&wcnt[w][(k[i] >> wshift) & (kDigits - 1)], 1u);
    radix_store_hist(wcnt, hist, ntiles);
}

// Scatter of one pass: hist holds the exclusive scan of the digit-major
// table, i.e. where each (digit, tile) run starts.  vals_in == nullptr means
// the value of point j is j.
template <typename K>
__global__ __launch_bounds__(kRadixThreads) void k_radix_scatter(
    int64_t M, const K *__restrict__ keys_in, const uint32_t *__restrict__ vals_in, int shift,
    const uint32_t *__restrict__ hist, int64_t ntiles, K *__restrict__ keys_out,
    uint32_t *__restrict__ vals_out, int self_scan)
{
    __shared__ uint32_t wcnt[kRadixWaves][kDigits];
    __shared__ uint32_t dstart[kDigits], gbase[kDigits];
    __shared__ K skey[kRadixTile];
    __shared__ uint32_t sval[kRadixTile];
    radix_clear(wcnt);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t t0 = (int64_t)blockIdx.x * kRadixTile;
    K k[kRadixItems];
    uint32_t v[kRadixItems], r[kRadixItems];
#pragma unroll
    for (int i = 0; i < kRadixItems; ++i) {
        const int64_t j = radix_pos(t0, w, i, lane);
        const bool valid = j < M;
        k[i] = valid ? keys_in[j] : (K)0;
        v[i] = valid ? (vals_in ? vals_in[j] : (uint32_t)j) : 0u;
    }
#pragma unroll
    for (int i = 0; i < kRadixItems; ++i)
        r[i] = wave_rank_digit(key_digit(k[i], shift), radix_pos(t0, w, i, lane) < M, lane,
                               wcnt[w]);
    __syncthreads();
    {
        // per digit: earlier wavefronts' counts, the tile's digit start and
        // the run's global start
        const int d = threadIdx.x;
        uint32_t s = 0;
#pragma unroll
        for (int q = 0; q < kRadixWaves; ++q) {
            const uint32_t c = wcnt[q][d];
            wcnt[q][d] = s;
            s += c;
        }
        uint32_t total;
        dstart[d] = block_exclusive_scan(s, &total);
        if (self_scan) {
            // few tiles: the run's start from the unscanned table, every
            // digit's total over all tiles plus this digit's earlier tiles
            const uint32_t *hd = hist + (int64_t)d * ntiles;
            uint32_t tot = 0, mine = 0;
            for (int64_t t0 = 0; t0 < ntiles; t0 += 8) {
                uint32_t h[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) h[q] = t0 + q < ntiles ? hd[t0 + q] : 0u;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    tot += h[q];
                    mine += t0 + q < (int64_t)blockIdx.x ? h[q] : 0u;
                }
            }
            uint32_t all;
            gbase[d] = block_exclusive_scan(tot, &all) + mine;
        } else {
            gbase[d] = hist[(int64_t)d * ntiles + blockIdx.x];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kRadixItems; ++i) {
        if (radix_pos(t0, w, i, lane) >= M) continue;
        const uint32_t d = key_digit(k[i], shift);
        const uint32_t p = dstart[d] + wcnt[w][d] + r[i];
        skey[p] = k[i];
        sval[p] = v[i];
    }
    __syncthreads();
    const int64_t n_here = M - t0 < kRadixTile ? M - t0 : kRadixTile;
#pragma unroll
    for (int i = 0; i < kRadixItems; ++i) {
        const int s = i * kRadixThreads + threadIdx.x;
        if (s >= n_here) continue;
        const K key = skey[s];
        const uint32_t d = key_digit(key, shift);
        const int64_t o = (int64_t)gbase[d] + (s - (int64_t)dstart[d]);
        keys_out[o] = key;
        vals_out[o] = sval[s];
    }
}

// Bin offsets from the sorted keys: position j (0 <= j <= M, j = M standing
// for a key of nbins) is the offset of every bin in (bin[j-1], bin[j]] (bin =
// sort key >> cell_shift), which gives each of the nbins + 1 offsets exactly
// one writer.  Gaps of up to kGapInline bins are written here; longer ones
// are listed (kGapMid: one workgroup each, longer: split over every
// workgroup) for k_fill_gaps.  With a segment list, every group of more than
// kSmallBin equal sort keys is listed once, by its first position, as
// [first, end) (end found by a binary search over the sorted keys).
constexpr int64_t kGapInline = 64;
constexpr int64_t kGapMid = 4096;

template <typename K>
__global__ __launch_bounds__(256) void k_bin_offsets(int64_t M, const K *__restrict__ keys,
                                                     int cell_shift, int64_t nbins,
                                                     uint32_t *__restrict__ offsets,
                                                     uint4 *__restrict__ gaps_mid,
                                                     uint4 *__restrict__ gaps_long,
                                                     uint32_t *__restrict__ ctrl,
                                                     uint2 *__restrict__ segs)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const uint64_t below = (1ull << lane) - 1ull;
    int64_t lo = 0, hi = -1;  // bins [lo, hi] take offset j
    bool large = false;
    uint32_t seg_end = 0;
    if (j <= M) {
        const int64_t kp = j > 0 ? (int64_t)(keys[j - 1] >> cell_shift) : -1;
        const int64_t kj = j < M ? (int64_t)(keys[j] >> cell_shift) : nbins;
        if (kj != kp) {
            lo = kp + 1;
            hi = kj;
        }
        if (segs && j + (int64_t)kSmallBin < M) {
            const K k = keys[j];
            if ((j == 0 || keys[j - 1] != k) && keys[j + kSmallBin] == k) {
                large = true;
                int64_t a = j + kSmallBin, b = M;  // first key above k in (a, b]
                while (b - a > 1) {
                    const int64_t mid = (a + b) >> 1;
                    if (keys[mid] == k)
                        a = mid;
                    else
                        b = mid;
                }
                seg_end = (uint32_t)b;
            }
        }
    }
    const int64_t len = hi - lo + 1;
    if (len > 0 && len <= kGapInline)
        for (int64_t b = lo; b <= hi; ++b) offsets[b] = (uint32_t)j;
    // one atomic per wavefront and list
    const bool mid = len > kGapInline && len <= kGapMid;
    const bool lng = len > kGapMid;
    const uint64_t mm = __ballot(mid), ml = __ballot(lng), mg = __ballot(large);
    uint32_t bm = 0, bl = 0, bg = 0;
    if (lane == 0) {
        if (mm) bm = atomicAdd(ctrl + 2, (uint32_t)__popcll(mm));
        if (ml) bl = atomicAdd(ctrl + 3, (uint32_t)__popcll(ml));
        if (mg) bg = atomicAdd(ctrl + 0, (uint32_t)__popcll(mg));
    }
    bm = __shfl(bm, 0, 64);
    bl = __shfl(bl, 0, 64);
    bg = __shfl(bg, 0, 64);
    const uint4 gap = make_uint4((uint32_t)lo, (uint32_t)hi, (uint32_t)j, 0u);
    if (mid) gaps_mid[bm + __popcll(mm & below)] = gap;
    if (lng) gaps_long[bl + __popcll(ml & below)] = gap;
    if (large) segs[bg + __popcll(mg & below)] = make_uint2((uint32_t)j, seg_end);
}

// Listed gaps: a mid gap per workgroup (grid-stride); every long gap cut in
// chunks of kGapMid bins spread round-robin over the workgroups.
__global__ __launch_bounds__(256) void k_fill_gaps(const uint4 *__restrict__ gaps_mid,
                                                   const uint4 *__restrict__ gaps_long,
                                                   const uint32_t *__restrict__ ctrl,
                                                   uint32_t *__restrict__ offsets)
{
    const uint32_t n_mid = ctrl[2], n_long = ctrl[3];
    for (uint32_t q = blockIdx.x; q < n_mid; q += gridDim.x) {
        const uint4 gp = gaps_mid[q];
        for (int64_t b = (int64_t)gp.x + threadIdx.x; b <= (int64_t)gp.y; b += blockDim.x)
            offsets[b] = gp.z;
    }
    for (uint32_t q = 0; q < n_long; ++q) {
        const uint4 gp = gaps_long[q];
        const int64_t len = (int64_t)gp.y - gp.x + 1;
        const int64_t chunks = (len + kGapMid - 1) / kGapMid;
        for (int64_t c = (blockIdx.x + (int64_t)gridDim.x - q % gridDim.x) % gridDim.x; c < chunks;
             c += gridDim.x) {
            const int64_t c0 = (int64_t)gp.x + c * kGapMid;
            const int64_t c1 = c0 + kGapMid - 1 < (int64_t)gp.y ? c0 + kGapMid - 1 : (int64_t)gp.y;
            for (int64_t b = c0 + threadIdx.x; b <= c1; b += blockDim.x) offsets[b] = gp.z;
        }
    }
}

// ---------------------------------------------------------------------
// Ordering keys.  Index order compares indices only; canonical order
// compares the coordinate bits (a strict order on distinct coordinates),
// then the index.
// ---------------------------------------------------------------------
struct SortElem {
    uint32_t k0, k1, k2, idx;
};

__device__ inline bool elem_less(const SortElem &a, const SortElem &b)
{
    if (a.k0 != b.k0) return a.k0 < b.k0;
    if (a.k1 != b.k1) return a.k1 < b.k1;
    if (a.k2 != b.k2) return a.k2 < b.k2;
    return a.idx < b.idx;
}

__device__ inline uint32_t mix32(uint64_t v)
{
    v ^= v >> 33;
    v *= 0xff51afd7ed558ccdull;
    v ^= v >> 33;
    v *= 0xc4ceb9fe1a85ec53ull;
    v ^= v >> 33;
    return (uint32_t)v;
}

// Ordering element of a point from its coordinates.
template <int D, typename T>
__device__ inline SortElem elem_of(const T *x, uint32_t idx)
{
    SortElem e{0u, 0u, 0u, idx};
    if constexpr (sizeof(T) == 4) {
        e.k0 = __float_as_uint((float)x[0]);
        if constexpr (D > 1) e.k1 = __float_as_uint((float)x[1]);
        if constexpr (D > 2) e.k2 = __float_as_uint((float)x[2]);
    } else {
        const uint64_t b0 = (uint64_t)__double_as_longlong((double)x[0]);
        e.k0 = (uint32_t)(b0 >> 32);
        e.k1 = (uint32_t)b0;
        uint64_t rest = 0;
        if constexpr (D > 1) rest = (uint64_t)__double_as_longlong((double)x[1]);
        if constexpr (D > 2)
            rest = rest * 0x9e3779b97f4a7c15ull ^ (uint64_t)__double_as_longlong((double)x[2]);
        e.k2 = D > 1 ? mix32(rest) : 0u;
    }
    return e;
}

// Ordering element of point idx of point-major coordinates.
template <int D, typename T>
__device__ inline SortElem elem_at(const T *coords, uint32_t idx)
{
    T x[D];
#pragma unroll
    for (int a = 0; a < D; ++a) x[a] = coords[(int64_t)idx * D + a];
    return elem_of<D, T>(x, idx);
}

// ---------------------------------------------------------------------
// Large groups (more than kSmallBin points of one sort key): a merge sort
// spread over every workgroup.  The listed segments are cut into tiles of
// kSortRun elements;
//  1. k_large_tiles scans the tile counts of the listed segments;
//  2. k_large_runs bitonic-sorts every tile of elements in LDS, in place;
//  3. k_large_merge, one launch per round of run length kSortRun << r,
//     gives every output tile of a segment still being merged its two input
//     ranges by a merge-path search, stages them in LDS and merges there;
//  4. k_large_final copies segments whose last round wrote the second
//     buffer back, so every segment ends up sorted in the first one.
// Element keys are unique (the index breaks ties), so every merge is exact.
// ---------------------------------------------------------------------
// Bitonic sort of s[0..n2) (n2 a power of two): every thread takes one
// compare-exchange pair per step (lower index i with bit j clear, partner
// i + j), so no lane idles.
__device__ inline void lds_bitonic(SortElem *s, int n2)
{
    const int half = n2 >> 1;
    for (int k = 2; k <= n2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = threadIdx.x; t < half; t += blockDim.x) {
                const int i = ((t & ~(j - 1)) << 1) | (t & (j - 1));
                const int p = i + j;
                const bool up = (i & k) == 0;
                const SortElem a = s[i], b = s[p];
                if (elem_less(b, a) == up) {
                    s[i] = b;
                    s[p] = a;
                }
            }
            __syncthreads();
        }
    }
}

// Elements of a taken from the first `diag` outputs of merging runs
// a[0..na) and b[0..nb) (ties cannot occur: keys are unique).
__device__ inline int64_t merge_split(const SortElem *a, int64_t na, const SortElem *b,
                                      int64_t nb, int64_t diag)
{
    int64_t lo = diag > nb ? diag - nb : 0;
    int64_t hi = diag < na ? diag : na;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (elem_less(a[mid], b[diag - 1 - mid]))
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__device__ inline int merge_rounds(int64_t n)
{
    int r = 0;
    for (int64_t len = kSortRun; len < n; len <<= 1) ++r;
    return r;
}

// tile_off[i] = first tile of listed segment i; tile_off[nlarge] = all
// tiles.
__global__ __launch_bounds__(1024) void k_large_tiles(const uint2 *__restrict__ segs,
                                                      const uint32_t *__restrict__ large_count,
                                                      uint32_t *__restrict__ tile_off,
                                                      uint32_t *__restrict__ large_max)
{
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry;
    const uint32_t nlarge = *large_count;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (uint32_t base = 0; base < nlarge; base += blockDim.x) {
        uint32_t i = base + threadIdx.x;
        uint32_t t = 0;
        if (i < nlarge) {
            const uint2 sg = segs[i];
            t = (sg.y - sg.x + kSortRun - 1) / kSortRun;
        }
        uint32_t x = t;
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wid] = x;
        __syncthreads();
        uint32_t pre = 0, all = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            if (w < wid) pre += wsum[w];
            all += wsum[w];
        }
        if (i < nlarge) tile_off[i] = carry + pre + x - t;
        __syncthreads();
        if (threadIdx.x == 0) carry += all;
        __syncthreads();
    }
    if (threadIdx.x == 0) tile_off[nlarge] = carry;
    // the largest listed segment, for the merge rounds to skip themselves
    uint32_t mx = 0;
    for (uint32_t i = threadIdx.x; i < nlarge; i += blockDim.x) {
        const uint2 sg = segs[i];
        const uint32_t n = sg.y - sg.x;
        mx = n > mx ? n : mx;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t y = __shfl_xor(mx, o, 64);
        mx = y > mx ? y : mx;
    }
    if (lane == 0 && mx) atomicMax(large_max, mx);
}

__device__ inline uint32_t tile_bin(const uint32_t *tile_off, uint32_t nlarge, uint32_t t)
{
    uint32_t lo = 0, hi = nlarge - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (tile_off[mid] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Each tile's ordering elements are built from the sorted indices and the
// coordinates, sorted in LDS and stored at the tile's slots.
template <int D, typename T>
__global__ __launch_bounds__(kRunThreads) void k_large_runs(
    const uint2 *__restrict__ segs, const uint32_t *__restrict__ large_count,
    const uint32_t *__restrict__ tile_off, const uint32_t *__restrict__ sorted,
    const T *__restrict__ coords, SortElem *__restrict__ elems)
{
    __shared__ SortElem s[kSortRun];
    const uint32_t nlarge = *large_count;
    if (nlarge == 0) return;
    const uint32_t ntiles = tile_off[nlarge];
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t i = tile_bin(tile_off, nlarge, t);
        const uint2 sg = segs[i];
        const int64_t lo = sg.x;
        const int64_t n = (int64_t)sg.y - lo;
        const int64_t r0 = (int64_t)(t - tile_off[i]) * kSortRun;
        const int m = (int)(n - r0 < kSortRun ? n - r0 : kSortRun);
        int n2 = 1;
        while (n2 < m) n2 <<= 1;
        for (int k = threadIdx.x; k < n2; k += kRunThreads)
            s[k] = k < m ? elem_at<D, T>(coords, sorted[lo + r0 + k])
                         : SortElem{0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu};
        __syncthreads();
        lds_bitonic(s, n2);
        for (int k = threadIdx.x; k < m; k += kRunThreads) elems[lo + r0 + k] = s[k];
        __syncthreads();
    }
}

__global__ __launch_bounds__(kSortThreads) void k_large_merge(
    const uint2 *__restrict__ segs, const uint32_t *__restrict__ large_count,
    const uint32_t *__restrict__ large_max, const uint32_t *__restrict__ tile_off, int round,
    const SortElem *__restrict__ src, SortElem *__restrict__ dst)
{
    __shared__ SortElem s[kSortRun];
    __shared__ int64_t split[2];
    const uint32_t nlarge = *large_count;
    const int64_t len = (int64_t)kSortRun << round;
    if (nlarge == 0 || len >= (int64_t)*large_max) return;  // every segment merged already
    const uint32_t ntiles = tile_off[nlarge];
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t i = tile_bin(tile_off, nlarge, t);
        const uint2 sg = segs[i];
        const int64_t lo = sg.x;
        const int64_t n = (int64_t)sg.y - lo;
        if (len >= n) continue;  // merged in an earlier round (uniform per tile)
        const int64_t out0 = (int64_t)(t - tile_off[i]) * kSortRun;
        const int64_t out1 = out0 + kSortRun < n ? out0 + kSortRun : n;
        const int64_t ps = (out0 / (2 * len)) * (2 * len);
        const int64_t na = n - ps < len ? n - ps : len;
        int64_t nb = n - ps - na;
        nb = nb < 0 ? 0 : (nb > len ? len : nb);
        const SortElem *a = src + lo + ps;
        const SortElem *b = a + na;
        const int64_t d0 = out0 - ps, d1 = out1 - ps;
        if (threadIdx.x < 2) split[threadIdx.x] = merge_split(a, na, b, nb, threadIdx.x ? d1 : d0);
        __syncthreads();
        const int64_t ia0 = split[0], ia1 = split[1];
        const int la = (int)(ia1 - ia0), lb = (int)((d1 - ia1) - (d0 - ia0));
        for (int k = threadIdx.x; k < la + lb; k += kSortThreads)
            s[k] = k < la ? a[ia0 + k] : b[(d0 - ia0) + (k - la)];
        __syncthreads();
        constexpr int per = kSortRun / kSortThreads;
        const int o0 = threadIdx.x * per;
        const int total = la + lb;
        if (o0 < total) {
            const int o1 = o0 + per < total ? o0 + per : total;
            int ia = (int)merge_split(s, la, s + la, lb, o0);
            int ib = o0 - ia;
            for (int o = o0; o < o1; ++o) {
                bool take_a = ib >= lb || (ia < la && elem_less(s[ia], s[la + ib]));
                dst[lo + out0 + o] = take_a ? s[ia++] : s[la + ib++];
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kSortThreads) void k_large_final(
    const uint2 *__restrict__ segs, const uint32_t *__restrict__ large_count,
    const uint32_t *__restrict__ tile_off, SortElem *__restrict__ elems,
    const SortElem *__restrict__ other)
{
    const uint32_t nlarge = *large_count;
    if (nlarge == 0) return;
    const uint32_t ntiles = tile_off[nlarge];
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t i = tile_bin(tile_off, nlarge, t);
        const uint2 sg = segs[i];
        const int64_t lo = sg.x;
        const int64_t n = (int64_t)sg.y - lo;
        if (n <= kSortRun || !(merge_rounds(n) & 1)) continue;
        const int64_t out0 = (int64_t)(t - tile_off[i]) * kSortRun;
        const int64_t out1 = out0 + kSortRun < n ? out0 + kSortRun : n;
        for (int64_t o = out0 + threadIdx.x; o < out1; o += kSortThreads)
            elems[lo + o] = other[lo + o];
    }
}

// ---------------------------------------------------------------------
// Prepare: final slot of every point (rank inside small bins), its
// wrapped support starts and offsets, and the original index.
// ---------------------------------------------------------------------
template <int D, typename T>
__device__ inline T elem_coord(const SortElem &e, const T *coords, int a, bool from_bits)
{
    if constexpr (sizeof(T) == 4) {
        if (from_bits) return __uint_as_float(a == 0 ? e.k0 : (a == 1 ? e.k1 : e.k2));
    }
    return coords[(int64_t)e.idx * D + a];
}

// Final slot of the record at bin-order slot s of bin b when dense bins are
// split off: light records first (slot minus the dense records before the
// bin), then the dense bins' records in bin order.
__device__ inline int64_t dense_slot(int64_t s, int64_t b, int64_t M, int64_t nbins,
                                     const uint32_t *__restrict__ offsets,
                                     const uint32_t *__restrict__ dense_before)
{
    const uint32_t d0 = dense_before[b];
    if (dense_before[b + 1] == d0) return s - d0;
    return M - (int64_t)dense_before[nbins] + d0 + (s - (int64_t)offsets[b]);
}

template <int D>
__device__ inline int64_t bin_of_starts(const int *st, const GridParams &g)
{
    int64_t b = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) b += (int64_t)(st[a] >> g.ax[a].shift) * g.ax[a].stride;
    return b;
}

// Record s of rank D (layout of rec_quads<D>): starts and offsets of every
// axis; ranks 1-3 keep the point index in the fourth start word.
template <int D, typename T>
__device__ inline void store_record(int4 *__restrict__ rec_start, T *__restrict__ rec_off,
                                    int64_t s, const int *st, const T *off, uint32_t idx)
{
    if constexpr (D <= 3) {
        rec_start[s] = make_int4(st[0], st[1], st[2], (int)idx);
        if constexpr (sizeof(T) == 4) {
            *(float4 *)(rec_off + s * 4) = make_float4(off[0], off[1], off[2], off[3]);
        } else {
            *(double2 *)(rec_off + s * 4) = make_double2(off[0], off[1]);
            *(double2 *)(rec_off + s * 4 + 2) = make_double2(off[2], off[3]);
        }
    } else {
        rec_start[2 * s] = make_int4(st[0], st[1], st[2], st[3]);
        rec_start[2 * s + 1] = make_int4(st[4], st[5], st[6], st[7]);
        T *o = rec_off + s * 8;
        if constexpr (sizeof(T) == 4) {
            *(float4 *)o = make_float4(off[0], off[1], off[2], off[3]);
            *(float4 *)(o + 4) = make_float4(off[4], off[5], off[6], off[7]);
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                *(double2 *)(o + 2 * q) = make_double2(off[2 * q], off[2 * q + 1]);
        }
    }
}

// Grouped and index orders: slot s keeps the point the radix sort put there.
// The coordinates are gathered once (all loads in flight before any store)
// and each record is written with two 16-byte stores.
template <int D, typename T>
__global__ __launch_bounds__(256) void k_prepare_grouped(int64_t M,
                                                         const uint32_t *__restrict__ sorted,
                                                         const T *__restrict__ coords,
                                                         GridParams g, int4 *__restrict__ rec_start,
                                                         T *__restrict__ rec_off,
                                                         uint32_t *__restrict__ out_sorted,
                                                         const uint32_t *__restrict__ offsets,
                                                         const uint32_t *__restrict__ dense_before)
{
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    const uint32_t idx = sorted[s];
    T x[D];
#pragma unroll
    for (int a = 0; a < D; ++a) x[a] = coords[(int64_t)idx * D + a];
    bool reach = true;
#pragma unroll
    for (int a = 0; a < D; ++a) reach = reach && FoldTraits<T>::in_reach(x[a]);
    int st[kMaxDims] = {};
    T off[kMaxDims] = {};
#pragma unroll
    for (int a = 0; a < D; ++a) {
        int64_t start;
        double o;
        support_start<T>(x[a], reach, g, a, &start, &o);
        st[a] = (int)wrapped_start(start, g.ax[a].n);
        off[a] = (T)o;
    }
    if (dense_before) {
        const int64_t sorted_slot = s;
        s = dense_slot(sorted_slot, bin_of_starts<D>(st, g), M, g.nbins, offsets, dense_before);
    }
    store_record<D, T>(rec_start, rec_off, s, st, off, idx);
    if (out_sorted) out_sorted[s] = idx;
}


// Canonical order: the point at slot s of the radix order (groups of equal
// sort keys, in index order) takes the final slot lo + rank inside its group
// (rank by coordinate bits, then index; the group found from the sorted keys).
// Groups of more than kSmallBin points were sorted by the large-group sort:
// their elements are read from there and keep their slots.  Then the wrapped
// support starts and offsets of the record.
template <int D, typename T>
__global__ __launch_bounds__(256) void k_prepare(int64_t M, const uint32_t *__restrict__ sorted,
                                                 const uint64_t *__restrict__ skeys,
                                                 const SortElem *__restrict__ large_elems,
                                                 const T *__restrict__ coords, GridParams g,
                                                 int4 *__restrict__ rec_start,
                                                 T *__restrict__ rec_off,
                                                 uint32_t *__restrict__ out_sorted,
                                                 const uint32_t *__restrict__ offsets,
                                                 const uint32_t *__restrict__ dense_before)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    const uint64_t key = skeys[s];
    // the group of equal keys around s, looked for up to kSmallBin + 1 on
    // either side
    int64_t lo = s, hi = s + 1;
    while (lo > 0 && s - lo <= (int64_t)kSmallBin && skeys[lo - 1] == key) --lo;
    while (hi < M && hi - s <= (int64_t)kSmallBin && skeys[hi] == key) ++hi;
    const bool large = hi - lo > (int64_t)kSmallBin;
    uint32_t idx;
    T x[D];
    int64_t dst = s;
    if (large) {
        const SortElem e = large_elems[s];
        idx = e.idx;
#pragma unroll
        for (int a = 0; a < D; ++a) x[a] = elem_coord<D, T>(e, coords, a, sizeof(T) == 4);
    } else {
        idx = sorted[s];
#pragma unroll
        for (int a = 0; a < D; ++a) x[a] = coords[(int64_t)idx * D + a];
        if (hi - lo > 1) {
            const SortElem me = elem_of<D, T>(x, idx);
            int64_t rank = 0;
            for (int64_t m = lo; m < hi; ++m) {
                if (m == s) continue;
                const uint32_t im = sorted[m];
                T y[D];
#pragma unroll
                for (int a = 0; a < D; ++a) y[a] = coords[(int64_t)im * D + a];
                rank += elem_less(elem_of<D, T>(y, im), me) ? 1 : 0;
            }
            dst = lo + rank;
        }
    }
    bool reach = true;
#pragma unroll
    for (int a = 0; a < D; ++a) reach = reach && FoldTraits<T>::in_reach(x[a]);
    int st[kMaxDims] = {};
    T off[kMaxDims] = {};
#pragma unroll
    for (int a = 0; a < D; ++a) {
        int64_t start;
        double o;
        support_start<T>(x[a], reach, g, a, &start, &o);
        st[a] = (int)wrapped_start(start, g.ax[a].n);
        off[a] = (T)o;
    }
    if (dense_before)
        dst = dense_slot(dst, bin_of_starts<D>(st, g), M, g.nbins, offsets, dense_before);
    store_record<D, T>(rec_start, rec_off, dst, st, off, idx);
    if (out_sorted) out_sorted[dst] = idx;
}

// Strengths of every vector gathered into record order (the first
// M - *tail slots when tail is given).
template <typename T>
__global__ __launch_bounds__(256) void k_gather_strengths(int64_t M, int64_t vectors,
                                                          const uint32_t *__restrict__ perm,
                                                          const T *__restrict__ c,
                                                          T *__restrict__ c_sorted,
                                                          const uint32_t *__restrict__ tail)
{
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M - (tail ? (int64_t)*tail : 0)) return;
    using V2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
    const int64_t idx = perm[s];
    // one vector load and store per complex value
    for (int64_t v = 0; v < vectors; ++v)
        *(V2 *)(c_sorted + 2 * (v * M + s)) = *(const V2 *)(c + 2 * (v * M + idx));
}

// Problem of concatenated slot s: the last problem whose m_off <= s
// (problems without points share the next one's m_off).
__device__ inline int problem_of_slot(const BatchProblem *__restrict__ probs, int nprob,
                                      int64_t s)
{
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (probs[mid].m_off <= s)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

template <int D, typename T>
__global__ __launch_bounds__(256) void k_prepare_many(int64_t M,
                                                      const uint32_t *__restrict__ sorted,
                                                      const BatchProblem *__restrict__ probs,
                                                      int nprob, int4 *__restrict__ rec_start,
                                                      T *__restrict__ rec_off,
                                                      uint32_t *__restrict__ rec_perm)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    const BatchProblem &pb = probs[problem_of_slot(probs, nprob, s)];
    const uint32_t idx = (uint32_t)((int64_t)sorted[s] - pb.m_off);
    const T *coords = (const T *)pb.coords;
    T x[D];
#pragma unroll
    for (int a = 0; a < D; ++a) x[a] = coords[(int64_t)idx * D + a];
    bool reach = true;
#pragma unroll
    for (int a = 0; a < D; ++a) reach = reach && FoldTraits<T>::in_reach(x[a]);
    int st[kMaxDims] = {};
    T off[kMaxDims] = {};
#pragma unroll
    for (int a = 0; a < D; ++a) {
        int64_t start;
        double o;
        support_start<T>(x[a], reach, pb.g, a, &start, &o);
        st[a] = (int)wrapped_start(start, pb.g.ax[a].n);
        off[a] = (T)o;
    }
    store_record<D, T>(rec_start, rec_off, s, st, off, idx);
    rec_perm[s] = idx;
}

template <typename T>
__global__ __launch_bounds__(256) void k_gather_strengths_many(
    int64_t M, const uint32_t *__restrict__ perm, const BatchProblem *__restrict__ probs,
    int nprob, T *__restrict__ c_sorted)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    using V2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
    const BatchProblem &pb = probs[problem_of_slot(probs, nprob, s)];
    *(V2 *)(c_sorted + 2 * s) = *((const V2 *)pb.in + perm[s]);
}

inline unsigned grid_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

}  // namespace

template <int D, typename T>
bool prepare_many(int64_t M, const uint32_t *sorted, const BatchProblem *probs, int nprob,
                  int4 *rec_start, T *rec_off, uint32_t *rec_perm, hipStream_t stream)
{
    if (M == 0) return true;
    k_prepare_many<D, T><<<grid_for(M, 256), 256, 0, stream>>>(M, sorted, probs, nprob,
                                                               rec_start, rec_off, rec_perm);
    return hip_ok(hipGetLastError(), "batched prepare launch");
}

template <typename T>
bool gather_strengths_many(int64_t M, const uint32_t *rec_perm, const BatchProblem *probs,
                           int nprob, T *c_sorted, hipStream_t stream)
{
    if (M == 0) return true;
    k_gather_strengths_many<T><<<grid_for(M, 256), 256, 0, stream>>>(M, rec_perm, probs, nprob,
                                                                     c_sorted);
    return hip_ok(hipGetLastError(), "batched strength gather launch");
}
#define MNUFFT_PREPARE_MANY_INSTANTIATE(D, T)                                                     \
    template bool prepare_many<D, T>(int64_t, const uint32_t *, const BatchProblem *, int,        \
                                     int4 *, T *, uint32_t *, hipStream_t);
MNUFFT_PREPARE_MANY_INSTANTIATE(1, float)
MNUFFT_PREPARE_MANY_INSTANTIATE(2, float)
MNUFFT_PREPARE_MANY_INSTANTIATE(3, float)
MNUFFT_PREPARE_MANY_INSTANTIATE(1, double)
MNUFFT_PREPARE_MANY_INSTANTIATE(2, double)
MNUFFT_PREPARE_MANY_INSTANTIATE(3, double)
template bool gather_strengths_many<float>(int64_t, const uint32_t *, const BatchProblem *, int,
                                           float *, hipStream_t);
template bool gather_strengths_many<double>(int64_t, const uint32_t *, const BatchProblem *, int,
                                            double *, hipStream_t);

// ---------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------
BinSorter::~BinSorter() { release(); }

void BinSorter::release()
{
    if (external_) {
        // bound to a caller's region: nothing here to free
        ctrl_ = offsets_ = keys_ = keys_b_ = vals_ = sorted_ = hist_ = hist_sums_ = nullptr;
        large_list_ = tile_off_ = nullptr;
        gaps_ = nullptr;
        scratch_a_ = scratch_b_ = nullptr;
        bins_cap_ = points_cap_ = sort_scratch_cap_ = 0;
        external_ = false;
        return;
    }
    for (uint32_t *p : {ctrl_, offsets_, keys_, keys_b_, vals_, sorted_, hist_, hist_sums_,
                        large_list_, tile_off_})
        if (p) (void)hipFree(p);
    for (void *p : {(void *)gaps_, scratch_a_, scratch_b_})
        if (p) (void)hipFree(p);
    ctrl_ = offsets_ = nullptr;
    keys_ = keys_b_ = vals_ = sorted_ = hist_ = hist_sums_ = large_list_ = tile_off_ = nullptr;
    gaps_ = nullptr;
    scratch_a_ = scratch_b_ = nullptr;
    bins_cap_ = points_cap_ = sort_scratch_cap_ = 0;
}

namespace {

bool alloc_u32(uint32_t **p, int64_t n, const char *what)
{
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    return hip_ok(hipMalloc(p, sizeof(uint32_t) * (size_t)(n > 0 ? n : 1)), what);
}

int64_t radix_tiles(int64_t M) { return (M + kRadixTile - 1) / kRadixTile; }

}  // namespace

namespace {

// Region sizes of an external binding, in 256-byte units so every buffer
// starts aligned.
struct ExternalLayout {
    size_t ctrl, keys, keys_b, vals, hist, hist_sums, gaps, total;
    int64_t n_mid, n_long;
};

size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

ExternalLayout external_layout(int64_t nbins, int64_t M)
{
    ExternalLayout l{};
    const int64_t m = M > 0 ? M : 1;
    const int64_t hist = kDigits * radix_tiles(m);
    l.n_mid = (nbins + 1) / (kGapInline + 1) + 2;
    l.n_long = (nbins + 1) / (kGapMid + 1) + 2;
    l.ctrl = round256(4 * sizeof(uint32_t));
    l.keys = l.keys_b = l.vals = round256(sizeof(uint32_t) * (size_t)m);
    l.hist = round256(sizeof(uint32_t) * (size_t)hist);
    l.hist_sums = round256(sizeof(uint32_t) * (size_t)((hist + kScanTile - 1) / kScanTile + 1));
    l.gaps = round256(sizeof(uint4) * (size_t)(l.n_mid + l.n_long));
    l.total = l.ctrl + l.keys + l.keys_b + l.vals + l.hist + l.hist_sums + l.gaps;
    return l;
}

}  // namespace

size_t BinSorter::external_bytes(int64_t nbins, int64_t M)
{
    return external_layout(nbins, M).total;
}

bool BinSorter::bind_external(void *base, size_t bytes, int64_t nbins, int64_t M,
                              uint32_t *sorted_out, uint32_t *offsets_out)
{
    release();
    const ExternalLayout l = external_layout(nbins, M);
    if (!base || !sorted_out || !offsets_out || l.total > bytes) {
        set_error("the static sort scratch holds " + std::to_string(bytes) + " bytes; " +
                  std::to_string(M) + " points in " + std::to_string(nbins) + " bins need " +
                  std::to_string(l.total));
        return false;
    }
    char *p = (char *)base;
    auto take = [&p](size_t n) {
        char *q = p;
        p += n;
        return q;
    };
    ctrl_ = (uint32_t *)take(l.ctrl);
    keys_ = (uint32_t *)take(l.keys);
    keys_b_ = (uint32_t *)take(l.keys_b);
    vals_ = (uint32_t *)take(l.vals);
    hist_ = (uint32_t *)take(l.hist);
    hist_sums_ = (uint32_t *)take(l.hist_sums);
    gaps_ = (uint4 *)take(l.gaps);
    gaps_mid_cap_ = l.n_mid;
    sorted_ = sorted_out;
    offsets_ = offsets_out;
    bins_cap_ = nbins;
    points_cap_ = M > 0 ? M : 1;
    sort_scratch_cap_ = 0;
    external_ = true;
    return true;
}

bool BinSorter::ensure(int64_t nbins, int64_t M, bool need_sort_scratch)
{
    if (external_) {
        // a bound region is never grown: the sort must fit what was bound
        if (nbins > bins_cap_ || M > points_cap_ || need_sort_scratch) {
            set_error("sort outside the bound static scratch (index order, bound sizes only)");
            return false;
        }
        return true;
    }
    if (nbins > bins_cap_) {
        if (!alloc_u32(&offsets_, nbins + 1, "bin offsets")) return false;
        // gap lists: at most one gap of more than kGapInline (kGapMid) bins
        // per kGapInline + 1 (kGapMid + 1) offsets
        const int64_t n_mid = (nbins + 1) / (kGapInline + 1) + 2;
        const int64_t n_long = (nbins + 1) / (kGapMid + 1) + 2;
        if (gaps_) (void)hipFree(gaps_);
        gaps_ = nullptr;
        if (!hip_ok(hipMalloc(&gaps_, sizeof(uint4) * (n_mid + n_long)), "bin gap lists"))
            return false;
        gaps_mid_cap_ = n_mid;
        bins_cap_ = nbins;
    }
    if (!ctrl_ && !alloc_u32(&ctrl_, 4, "bin sort counters")) return false;
    if (M > points_cap_) {
        const int64_t cap = M + M / 4 + 1024;
        // 64-bit sort keys (canonical order), else 32-bit
        if (!alloc_u32(&keys_, 2 * cap, "point keys")) return false;
        if (!alloc_u32(&keys_b_, 2 * cap, "point keys")) return false;
        if (!alloc_u32(&vals_, cap, "point values")) return false;
        if (!alloc_u32(&sorted_, cap, "sorted points")) return false;
        const int64_t hist = kDigits * radix_tiles(cap);
        if (!alloc_u32(&hist_, hist, "radix histograms")) return false;
        if (!alloc_u32(&hist_sums_, (hist + kScanTile - 1) / kScanTile + 1, "radix scan tiles"))
            return false;
        // (first, end) of every group above kSmallBin points
        if (!alloc_u32(&large_list_, 2 * (cap / (kSmallBin + 1) + 2), "large-group list"))
            return false;
        if (!alloc_u32(&tile_off_, cap / (kSmallBin + 1) + 3, "large-bin tiles")) return false;
        points_cap_ = cap;
    }
    if (need_sort_scratch && sort_scratch_cap_ < points_cap_) {
        if (scratch_a_) (void)hipFree(scratch_a_);
        if (scratch_b_) (void)hipFree(scratch_b_);
        scratch_a_ = scratch_b_ = nullptr;
        if (!hip_ok(hipMalloc(&scratch_a_, 16 * points_cap_), "large-bin scratch")) return false;
        if (!hip_ok(hipMalloc(&scratch_b_, 16 * points_cap_), "large-bin scratch")) return false;
        sort_scratch_cap_ = points_cap_;
    }
    return true;
}

__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Exclusive scan of n entries into out (which may be in): tile sums, then
// every tile's scan plus the sum of the tiles before it.
int64_t scan_tiles(int64_t n) { return (n + kScanTile - 1) / kScanTile; }

void scan_u32(const uint32_t *in, int64_t n, uint32_t *sums, uint32_t *out, hipStream_t stream)
{
    const int64_t tiles = scan_tiles(n);
    k_scan_reduce<<<(unsigned)tiles, kScanThreads, 0, stream>>>(in, n, sums);
    k_scan_down<<<(unsigned)tiles, kScanThreads, 0, stream>>>(in, n, sums, out);
}

// LSD radix passes over `bits` key bits: sort keys end in sorted_keys_,
// point indices in sorted_.
template <int D, typename T, typename K>
void BinSorter::radix_passes(int64_t M, const T *coords, const GridParams &g, int mode,
                             int sub_bits, int bits, hipStream_t stream)
{
    radix_loop<K>(M, bits, stream, [&](int64_t ntiles, K *keys) {
        k_radix_keys<D, T, K><<<(unsigned)ntiles, kRadixThreads, 0, stream>>>(
            M, coords, g, mode, sub_bits, keys, hist_, ntiles);
    });
}

template <typename K, typename F>
void BinSorter::radix_loop(int64_t M, int bits, hipStream_t stream, F keys0)
{
    const int passes = bits > 8 ? (bits + 7) / 8 : 1;
    const int64_t ntiles = radix_tiles(M);
    const int64_t nhist = kDigits * ntiles;
    K *Kb[2] = {(K *)keys_, (K *)keys_b_};
    uint32_t *V[2];
    V[passes & 1] = sorted_;
    V[(passes + 1) & 1] = vals_;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        if (p == 0)
            keys0(ntiles, Kb[0]);
        else
            k_radix_hist<K><<<(unsigned)ntiles, kRadixThreads, 0, stream>>>(M, Kb[p & 1], shift,
                                                                            hist_, ntiles);
        // up to kSelfScanTiles tiles the scatter finds its runs' starts in the
        // unscanned table itself
        const bool self_scan = ntiles <= kSelfScanTiles;
        if (!self_scan) scan_u32(hist_, nhist, hist_sums_, hist_, stream);
        k_radix_scatter<K><<<(unsigned)ntiles, kRadixThreads, 0, stream>>>(
            M, Kb[p & 1], p == 0 ? nullptr : V[p & 1], shift, hist_, ntiles, Kb[(p + 1) & 1],
            V[(p + 1) & 1], self_scan ? 1 : 0);
    }
    sorted_keys_ = Kb[passes & 1];
}

template <int D, typename T>
bool BinSorter::sort(int64_t M, const T *coords, const GridParams &g, BinOrder order,
                     hipStream_t stream)
{
    nbins_ = g.nbins;
    M_ = M;
    if (!ensure(g.nbins, M, order == BinOrder::Canonical)) return false;
    // Clear the counters (large bins, largest large bin, mid and long gaps)
    // with a fill kernel (no memset nodes, so callers may capture the
    // sequence in a graph).
    k_fill_u32<<<1, 64, 0, stream>>>(ctrl_, 4, 0u);
    if (M == 0) {
        k_fill_u32<<<grid_for(g.nbins + 1, 256), 256, 0, stream>>>(offsets_, g.nbins + 1, 0u);
        return hip_ok(hipGetLastError(), "bin clear");
    }
    uint32_t *large_count = ctrl_;
    uint32_t *large_max = ctrl_ + 1;

    // radix passes over the bits of the largest key, values ending in sorted_.
    // Below the bin key, the grouped order also sorts by the support-start
    // cell inside the bin (neighbouring points then read neighbouring grid
    // windows) and the canonical order by the position inside the bin
    // quantized to 2^5 steps per axis (64-bit keys; groups of equal keys then
    // hold a few points even in dense clusters, ranked by coordinate bits in
    // prepare); the index order sorts by bin only, so stability leaves every
    // bin in index order.
    int bits = 0;
    while (bits < 32 && ((uint64_t)1 << bits) < (uint64_t)g.nbins) ++bits;
    int mode = kKeyBin, sub_bits = 0, cell_shift = 0;
    if (order == BinOrder::Grouped && bits + cell_bits<D>(g) <= 32) {
        mode = kKeyCell;
        cell_shift = cell_bits<D>(g);
    } else if (order == BinOrder::Canonical) {
        sub_bits = std::min(5, (64 - bits) / D);
        mode = kKeyPosition;
        cell_shift = D * sub_bits;
    }
    bits += cell_shift;
    const bool wide = order == BinOrder::Canonical;
    if (wide)
        radix_passes<D, T, uint64_t>(M, coords, g, mode, sub_bits, bits, stream);
    else
        radix_passes<D, T, uint32_t>(M, coords, g, mode, sub_bits, bits, stream);
    uint2 *segs = (uint2 *)large_list_;
    uint4 *gaps_long = gaps_ + gaps_mid_cap_;
    if (wide)
        k_bin_offsets<uint64_t><<<grid_for(M + 1, 256), 256, 0, stream>>>(
            M, (const uint64_t *)sorted_keys_, cell_shift, g.nbins, offsets_, gaps_, gaps_long,
            ctrl_, segs);
    else
        k_bin_offsets<uint32_t><<<grid_for(M + 1, 256), 256, 0, stream>>>(
            M, (const uint32_t *)sorted_keys_, cell_shift, g.nbins, offsets_, gaps_, gaps_long,
            ctrl_, nullptr);
    k_fill_gaps<<<1024, 256, 0, stream>>>(gaps_, gaps_long, ctrl_, offsets_);

    if (order == BinOrder::Canonical) {
        unsigned grid = (unsigned)std::min<int64_t>(2048, M / kSortRun + 64);
        SortElem *A = (SortElem *)scratch_a_, *B = (SortElem *)scratch_b_;
        k_large_tiles<<<1, 1024, 0, stream>>>(segs, large_count, tile_off_, large_max);
        k_large_runs<D, T><<<grid, kRunThreads, 0, stream>>>(segs, large_count, tile_off_,
                                                             sorted_, coords, A);
        // rounds for the largest possible group; groups done earlier skip
        int rounds = 0;
        for (int64_t len = kSortRun; len < M; len <<= 1) ++rounds;
        for (int r = 0; r < rounds; ++r)
            k_large_merge<<<grid, kSortThreads, 0, stream>>>(segs, large_count, large_max,
                                                             tile_off_, r, (r & 1) ? B : A,
                                                             (r & 1) ? A : B);
        if (rounds > 0)
            k_large_final<<<grid, kSortThreads, 0, stream>>>(segs, large_count, tile_off_, A, B);
    }
    return hip_ok(hipGetLastError(), "bin sort launch");
}

template <int D, typename T>
bool BinSorter::sort_many(int64_t M, const BatchProblem *probs, int nprob, uint64_t nbins,
                          hipStream_t stream)
{
    nbins_ = 0;
    M_ = M;
    if (M == 0) return true;
    if (nbins >= 0xffffffffull) {
        set_error("batched bin sort: more than 2^32 bins over the problems");
        return false;
    }
    if (!ensure((int64_t)nbins, M, false)) return false;
    nbins_ = (int64_t)nbins;
    k_fill_u32<<<1, 64, 0, stream>>>(ctrl_, 4, 0u);
    int bits = 0;
    while (bits < 32 && ((uint64_t)1 << bits) < nbins) ++bits;
    radix_loop<uint32_t>(M, bits, stream, [&](int64_t ntiles, uint32_t *keys) {
        k_radix_keys_many<D, T><<<(unsigned)ntiles, kRadixThreads, 0, stream>>>(
            M, probs, nprob, keys, hist_, ntiles);
    });
    // offsets over the combined bin list
    k_bin_offsets<uint32_t><<<grid_for(M + 1, 256), 256, 0, stream>>>(
        M, (const uint32_t *)sorted_keys_, 0, (int64_t)nbins, offsets_, gaps_,
        gaps_ + gaps_mid_cap_, ctrl_, nullptr);
    k_fill_gaps<<<1024, 256, 0, stream>>>(gaps_, gaps_ + gaps_mid_cap_, ctrl_, offsets_);
    return hip_ok(hipGetLastError(), "batched bin sort launch");
}

template <int D, typename T>
bool BinSorter::prepare(int64_t M, const T *coords, const GridParams &g, BinOrder order,
                        int4 *rec_start, T *rec_off, uint32_t *out_sorted, hipStream_t stream,
                        const uint32_t *dense_before)
{
    if (M == 0) return true;
    // the radix sort leaves every bin in index order: the grouped and the
    // index orders read the sorted indices as they are
    if (order != BinOrder::Canonical)
        k_prepare_grouped<D, T><<<grid_for(M, 256), 256, 0, stream>>>(M, sorted_, coords, g,
                                                                      rec_start, rec_off,
                                                                      out_sorted, offsets_,
                                                                      dense_before);
    else
        k_prepare<D, T><<<grid_for(M, 256), 256, 0, stream>>>(
            M, sorted_, (const uint64_t *)sorted_keys_, (const SortElem *)scratch_a_, coords, g,
            rec_start, rec_off, out_sorted, offsets_, dense_before);
    return hip_ok(hipGetLastError(), "prepare launch");
}

template <typename T>
bool gather_strengths(int64_t M, int64_t vectors, const uint32_t *perm, const T *c, T *c_sorted,
                      hipStream_t stream, const uint32_t *tail)
{
    if (M == 0) return true;
    k_gather_strengths<T><<<grid_for(M, 256), 256, 0, stream>>>(M, vectors, perm, c, c_sorted,
                                                                tail);
    return hip_ok(hipGetLastError(), "strength gather launch");
}

// ---------------------------------------------------------------------
// Static sort scratch of the entry point that allocates nothing: a
// file-scope device array, one copy per GPU the library runs on (held until
// the process ends).  Host code never passes `scratch` itself to a kernel:
// its name is a host-side stand-in, and the device address comes from
// hipGetSymbolAddress, once per device.
// ---------------------------------------------------------------------
namespace {
constexpr size_t kStaticSortBytes = (size_t)32 << 20;
static __device__ __align__(256) uint32_t scratch[kStaticSortBytes / sizeof(uint32_t)];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

size_t static_sort_bytes() { return kStaticSortBytes; }

void *static_sort_scratch()
{
    int dev = 0;
    if (!hip_ok(hipGetDevice(&dev), "current device")) return nullptr;
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device ordinal beyond the static scratch table");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        if (!hip_ok(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "static sort scratch lookup"))
            return nullptr;
        if (!ptr) {
            set_error("static sort scratch lookup returned no address");
            return nullptr;
        }
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

#define MNUFFT_BINS_INSTANTIATE(D, T)                                                             \
    template bool BinSorter::sort<D, T>(int64_t, const T *, const GridParams &, BinOrder,          \
                                        hipStream_t);                                             \
    template bool BinSorter::prepare<D, T>(int64_t, const T *, const GridParams &, BinOrder,       \
                                           int4 *, T *, uint32_t *, hipStream_t,                 \
                                           const uint32_t *);
#define MNUFFT_SORT_MANY_INSTANTIATE(D, T)                                                        \
    template bool BinSorter::sort_many<D, T>(int64_t, const BatchProblem *, int, uint64_t,        \
                                             hipStream_t);
MNUFFT_SORT_MANY_INSTANTIATE(1, float)
MNUFFT_SORT_MANY_INSTANTIATE(2, float)
MNUFFT_SORT_MANY_INSTANTIATE(3, float)
MNUFFT_SORT_MANY_INSTANTIATE(1, double)
MNUFFT_SORT_MANY_INSTANTIATE(2, double)
MNUFFT_SORT_MANY_INSTANTIATE(3, double)
MNUFFT_BINS_INSTANTIATE(1, float)
MNUFFT_BINS_INSTANTIATE(2, float)
MNUFFT_BINS_INSTANTIATE(3, float)
MNUFFT_BINS_INSTANTIATE(1, double)
MNUFFT_BINS_INSTANTIATE(2, double)
MNUFFT_BINS_INSTANTIATE(3, double)
MNUFFT_BINS_INSTANTIATE(4, float)
MNUFFT_BINS_INSTANTIATE(5, float)
MNUFFT_BINS_INSTANTIATE(6, float)
MNUFFT_BINS_INSTANTIATE(7, float)
MNUFFT_BINS_INSTANTIATE(8, float)
MNUFFT_BINS_INSTANTIATE(4, double)
MNUFFT_BINS_INSTANTIATE(5, double)
MNUFFT_BINS_INSTANTIATE(6, double)
MNUFFT_BINS_INSTANTIATE(7, double)
MNUFFT_BINS_INSTANTIATE(8, double)
template bool gather_strengths<float>(int64_t, int64_t, const uint32_t *, const float *,
                                      float *, hipStream_t, const uint32_t *);
template bool gather_strengths<double>(int64_t, int64_t, const uint32_t *, const double *,
                                       double *, hipStream_t, const uint32_t *);

}  // namespace mnufft
