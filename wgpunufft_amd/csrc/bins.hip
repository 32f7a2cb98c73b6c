// Point binning for the MI355X engine: count -> exclusive scan -> scatter ->
// in-bin order -> prepared records.
//
// Reference behaviour (src/gpu_nd_bins.rs:1-33, src/gpu_bin_sort.rs:1-56):
// points are counted into bins of support starts with integer atomics, the
// counts are scanned into bin offsets (plus the total), point indices are
// scattered into their bins, and the order inside every bin is then made
// deterministic.  Index order restores ascending original index (the
// reference permutation); canonical order sorts by coordinate bits so the
// type-1 sums do not depend on the input order of the points.
//
// MI355X design: the count pass aggregates each workgroup's 1024 points in an
// LDS hash table before one global atomic per distinct bin (the bin arrays,
// a few tens of MB, stay in the 256 MB Infinity Cache); a three-kernel
// reduce-then-scan over 4096-element tiles; bins above 64 points are sorted
// by a tiled merge sort over all workgroups (LDS bitonic runs, merge-path
// rounds).
#include <hip/hip_runtime.h>

#include "bins.hpp"
#include "fold.hpp"

namespace mnufft {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;  // 4096
constexpr int kSortThreads = 256;
constexpr int kRunThreads = 1024;  // LDS bitonic runs: one thread per pair
constexpr int kSortRun = 2048;

// ---------------------------------------------------------------------
// Count: bin key and atomic rank of every point; bins reaching
// kSmallBin + 1 points are listed once for the large-bin sort.
// ---------------------------------------------------------------------
// Each workgroup counts kCountPoints consecutive points: their keys are
// first aggregated in an LDS hash table (LDS atomics give every point its
// rank among the workgroup's points of the same bin), then every distinct
// bin takes one global atomic.  A dense cluster, whose points crowd a few
// bins, then costs a few global atomics per workgroup instead of one per
// point on the same address.
constexpr int kCountThreads = 256;
constexpr int kCountPerThread = 4;
constexpr int kCountPoints = kCountThreads * kCountPerThread;
constexpr int kHashSize = 2 * kCountPoints;
constexpr uint32_t kEmpty = 0xffffffffu;

template <int D, typename T>
__global__ __launch_bounds__(kCountThreads) void k_count(int64_t M, const T *__restrict__ coords,
                                                         GridParams g,
                                                         uint32_t *__restrict__ counts,
                                                         uint32_t *__restrict__ keys,
                                                         uint32_t *__restrict__ ranks,
                                                         uint32_t *__restrict__ large_list,
                                                         uint32_t *__restrict__ large_count)
{
    __shared__ uint32_t ht_key[kHashSize];
    __shared__ uint32_t ht_cnt[kHashSize];
    __shared__ uint32_t ht_base[kHashSize];
    for (int h = threadIdx.x; h < kHashSize; h += kCountThreads) {
        ht_key[h] = kEmpty;
        ht_cnt[h] = 0u;
    }
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kCountPoints;
    const int lane = threadIdx.x & 63;
    const uint64_t lanes_below = (1ull << lane) - 1ull;
    uint32_t my_key[kCountPerThread], my_slot[kCountPerThread], my_rank[kCountPerThread];
#pragma unroll
    for (int i = 0; i < kCountPerThread; ++i) {
        int64_t j = base + (int64_t)i * kCountThreads + threadIdx.x;
        my_slot[i] = kEmpty;
        const bool valid = j < M;
        const uint32_t key = valid ? (uint32_t)point_bin_key<D, T>(coords, j, g) : kEmpty;
        my_key[i] = key;
        bool pending = valid;
        // Lanes of the wavefront that share a key (a dense cluster) take one
        // LDS atomic per group instead of one each on the same address;
        // peeling stops at the first group of one lane (spread-out keys).
        for (int it = 0; it < 8; ++it) {
            const uint64_t pm = __ballot(pending);
            if (pm == 0) break;
            const int leader = __ffsll((long long)pm) - 1;
            const uint32_t lk = __shfl(key, leader, 64);
            const bool mine = pending && key == lk;
            const uint64_t gm = __ballot(mine);
            const int gcount = __popcll(gm);
            uint32_t h = 0, rank0 = 0;
            if (lane == leader) {
                h = (lk * 2654435761u) & (kHashSize - 1);
                for (;;) {
                    uint32_t old = atomicCAS(&ht_key[h], kEmpty, lk);
                    if (old == kEmpty || old == lk) break;
                    h = (h + 1) & (kHashSize - 1);
                }
                rank0 = atomicAdd(&ht_cnt[h], (uint32_t)gcount);
            }
            h = __shfl(h, leader, 64);
            rank0 = __shfl(rank0, leader, 64);
            if (mine) {
                my_slot[i] = h;
                my_rank[i] = rank0 + (uint32_t)__popcll(gm & lanes_below);
                pending = false;
            }
            if (gcount < 2) break;
        }
        if (pending) {
            uint32_t h = (key * 2654435761u) & (kHashSize - 1);
            for (;;) {
                uint32_t old = atomicCAS(&ht_key[h], kEmpty, key);
                if (old == kEmpty || old == key) break;
                h = (h + 1) & (kHashSize - 1);
            }
            my_slot[i] = h;
            my_rank[i] = atomicAdd(&ht_cnt[h], 1u);
        }
    }
    __syncthreads();
    // every slot's global atomic issued before any result is awaited
    constexpr int kSlotsPerThread = kHashSize / kCountThreads;
    uint32_t hk[kSlotsPerThread], hn[kSlotsPerThread], ho[kSlotsPerThread];
#pragma unroll
    for (int q = 0; q < kSlotsPerThread; ++q) {
        const int h = threadIdx.x + q * kCountThreads;
        hk[q] = ht_key[h];
        hn[q] = ht_cnt[h];
    }
#pragma unroll
    for (int q = 0; q < kSlotsPerThread; ++q)
        ho[q] = hk[q] != kEmpty ? atomicAdd(&counts[hk[q]], hn[q]) : 0u;
#pragma unroll
    for (int q = 0; q < kSlotsPerThread; ++q) {
        if (hk[q] == kEmpty) continue;
        ht_base[threadIdx.x + q * kCountThreads] = ho[q];
        if (ho[q] <= kSmallBin && ho[q] + hn[q] > kSmallBin) {
            uint32_t slot = atomicAdd(large_count, 1u);
            large_list[slot] = hk[q];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kCountPerThread; ++i) {
        int64_t j = base + (int64_t)i * kCountThreads + threadIdx.x;
        if (j >= M) continue;
        keys[j] = my_key[i];
        ranks[j] = ht_base[my_slot[i]] + my_rank[i];
    }
}

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

__global__ __launch_bounds__(kScanThreads) void k_scan_top(uint32_t *__restrict__ tile_sums,
                                                           int64_t ntiles)
{
    uint32_t carry = 0;
    for (int64_t base = 0; base < ntiles; base += kScanThreads) {
        int64_t k = base + threadIdx.x;
        uint32_t v = k < ntiles ? tile_sums[k] : 0u;
        uint32_t total;
        uint32_t ex = block_exclusive_scan(v, &total);
        if (k < ntiles) tile_sums[k] = carry + ex;
        carry += total;
    }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_down(const uint32_t *__restrict__ in,
                                                            int64_t n,
                                                            const uint32_t *__restrict__ tile_sums,
                                                            uint32_t *__restrict__ out)
{
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
    uint32_t run = block_exclusive_scan(s, &total) + tile_sums[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        int64_t k = base + i;
        if (k < n) out[k] = run;
        run += v[i];
    }
}

// ---------------------------------------------------------------------
// Scatter: sorted[offset[key] + rank] = point.
// ---------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_scatter(int64_t M, const uint32_t *__restrict__ keys,
                                                 const uint32_t *__restrict__ ranks,
                                                 const uint32_t *__restrict__ offsets,
                                                 uint32_t *__restrict__ sorted)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    sorted[offsets[keys[j]] + ranks[j]] = (uint32_t)j;
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

template <int D, typename T>
__device__ inline SortElem make_elem(const T *coords, uint32_t idx, int order)
{
    SortElem e{0u, 0u, 0u, idx};
    if (order != (int)BinOrder::Canonical) return e;
    if constexpr (sizeof(T) == 4) {
        e.k0 = __float_as_uint((float)coords[(int64_t)idx * D + 0]);
        if (D > 1) e.k1 = __float_as_uint((float)coords[(int64_t)idx * D + (D > 1 ? 1 : 0)]);
        if (D > 2) e.k2 = __float_as_uint((float)coords[(int64_t)idx * D + (D > 2 ? 2 : 0)]);
    } else {
        uint64_t b0 = (uint64_t)__double_as_longlong((double)coords[(int64_t)idx * D + 0]);
        e.k0 = (uint32_t)(b0 >> 32);
        e.k1 = (uint32_t)b0;
        uint64_t rest = 0;
        if (D > 1) rest = (uint64_t)__double_as_longlong((double)coords[(int64_t)idx * D + (D > 1 ? 1 : 0)]);
        if (D > 2)
            rest = rest * 0x9e3779b97f4a7c15ull ^
                   (uint64_t)__double_as_longlong((double)coords[(int64_t)idx * D + (D > 2 ? 2 : 0)]);
        e.k2 = D > 1 ? mix32(rest) : 0u;
    }
    return e;
}

// Ordered sorts: every point's ordering element (coordinate bits or index,
// and the index) scattered into its bin's slot range, read sequentially.
template <int D, typename T>
__global__ __launch_bounds__(256) void k_scatter_elems(int64_t M, const T *__restrict__ coords,
                                                       const uint32_t *__restrict__ keys,
                                                       const uint32_t *__restrict__ ranks,
                                                       const uint32_t *__restrict__ offsets,
                                                       int order, SortElem *__restrict__ elems)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    elems[offsets[keys[j]] + ranks[j]] = make_elem<D, T>(coords, (uint32_t)j, order);
}

// ---------------------------------------------------------------------
// Large bins (more than kSmallBin points): a merge sort spread over every
// workgroup.  The listed bins are cut into tiles of kSortRun elements;
//  1. k_large_tiles scans the tile counts of the listed bins;
//  2. k_large_runs bitonic-sorts every tile of elements in LDS, in place;
//  3. k_large_merge, one launch per round of run length kSortRun << r,
//     gives every output tile of a bin still being merged its two input
//     ranges by a merge-path search, stages them in LDS and merges there;
//  4. k_large_final copies bins whose last round wrote the second buffer
//     back, so the elements of every bin end up sorted in the first one.
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

// tile_off[i] = first tile of listed bin i; tile_off[nlarge] = all tiles.
__global__ __launch_bounds__(1024) void k_large_tiles(const uint32_t *__restrict__ large_list,
                                                      const uint32_t *__restrict__ large_count,
                                                      const uint32_t *__restrict__ offsets,
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
            uint32_t bin = large_list[i];
            uint32_t n = offsets[bin + 1] - offsets[bin];
            t = (n + kSortRun - 1) / kSortRun;
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
    // the largest listed bin, for the merge rounds to skip themselves
    uint32_t mx = 0;
    for (uint32_t i = threadIdx.x; i < nlarge; i += blockDim.x) {
        const uint32_t bin = large_list[i];
        const uint32_t n = offsets[bin + 1] - offsets[bin];
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

__global__ __launch_bounds__(kRunThreads) void k_large_runs(
    const uint32_t *__restrict__ large_list, const uint32_t *__restrict__ large_count,
    const uint32_t *__restrict__ tile_off, const uint32_t *__restrict__ offsets,
    SortElem *__restrict__ elems)
{
    __shared__ SortElem s[kSortRun];
    const uint32_t nlarge = *large_count;
    if (nlarge == 0) return;
    const uint32_t ntiles = tile_off[nlarge];
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t i = tile_bin(tile_off, nlarge, t);
        const uint32_t bin = large_list[i];
        const int64_t lo = offsets[bin];
        const int64_t n = (int64_t)offsets[bin + 1] - lo;
        const int64_t r0 = (int64_t)(t - tile_off[i]) * kSortRun;
        const int m = (int)(n - r0 < kSortRun ? n - r0 : kSortRun);
        int n2 = 1;
        while (n2 < m) n2 <<= 1;
        for (int k = threadIdx.x; k < n2; k += kRunThreads)
            s[k] = k < m ? elems[lo + r0 + k]
                         : SortElem{0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu};
        __syncthreads();
        lds_bitonic(s, n2);
        for (int k = threadIdx.x; k < m; k += kRunThreads) elems[lo + r0 + k] = s[k];
        __syncthreads();
    }
}

__global__ __launch_bounds__(kSortThreads) void k_large_merge(
    const uint32_t *__restrict__ large_list, const uint32_t *__restrict__ large_count,
    const uint32_t *__restrict__ large_max, const uint32_t *__restrict__ tile_off,
    const uint32_t *__restrict__ offsets, int round,
    const SortElem *__restrict__ src, SortElem *__restrict__ dst)
{
    __shared__ SortElem s[kSortRun];
    __shared__ int64_t split[2];
    const uint32_t nlarge = *large_count;
    const int64_t len = (int64_t)kSortRun << round;
    if (nlarge == 0 || len >= (int64_t)*large_max) return;  // every bin merged already
    const uint32_t ntiles = tile_off[nlarge];
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t i = tile_bin(tile_off, nlarge, t);
        const uint32_t bin = large_list[i];
        const int64_t lo = offsets[bin];
        const int64_t n = (int64_t)offsets[bin + 1] - lo;
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
    const uint32_t *__restrict__ large_list, const uint32_t *__restrict__ large_count,
    const uint32_t *__restrict__ tile_off, const uint32_t *__restrict__ offsets,
    SortElem *__restrict__ elems, const SortElem *__restrict__ other)
{
    const uint32_t nlarge = *large_count;
    if (nlarge == 0) return;
    const uint32_t ntiles = tile_off[nlarge];
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t i = tile_bin(tile_off, nlarge, t);
        const uint32_t bin = large_list[i];
        const int64_t lo = offsets[bin];
        const int64_t n = (int64_t)offsets[bin + 1] - lo;
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

// Grouped order: slot s keeps the point the scatter put there.
template <int D, typename T>
__global__ __launch_bounds__(256) void k_prepare_grouped(int64_t M,
                                                         const uint32_t *__restrict__ sorted,
                                                         const T *__restrict__ coords,
                                                         GridParams g, int4 *__restrict__ rec_start,
                                                         T *__restrict__ rec_off,
                                                         uint32_t *__restrict__ out_sorted)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    const uint32_t idx = sorted[s];
    const bool reach = point_in_reach<D, T>(coords, idx);
    int st[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < D; ++a) {
        int64_t start;
        double off;
        support_start<T>(coords[(int64_t)idx * D + a], reach, g, a, &start, &off);
        st[a] = (int)wrapped_start(start, g.ax[a].n);
        rec_off[s * 4 + a] = (T)off;
    }
    rec_start[s] = make_int4(st[0], st[1], st[2], (int)idx);
    if (out_sorted) out_sorted[s] = idx;
}

// Ordered sorts: the element at slot s (large bins already sorted there;
// small bins ranked here among their contiguous elements) gives the final
// slot, the wrapped support starts and offsets.  F32 canonical elements hold
// the coordinate bits, so nothing is gathered.
template <int D, typename T>
__global__ __launch_bounds__(256) void k_prepare(int64_t M, const SortElem *__restrict__ elems,
                                                 const uint32_t *__restrict__ offsets,
                                                 const T *__restrict__ coords, GridParams g,
                                                 int order, int4 *__restrict__ rec_start,
                                                 T *__restrict__ rec_off,
                                                 uint32_t *__restrict__ out_sorted)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    const SortElem me = elems[s];
    const bool from_bits = sizeof(T) == 4 && order == (int)BinOrder::Canonical;
    T x[3];
#pragma unroll
    for (int a = 0; a < D; ++a) x[a] = elem_coord<D, T>(me, coords, a, from_bits);
    bool reach = true;
#pragma unroll
    for (int a = 0; a < D; ++a) reach = reach && FoldTraits<T>::in_reach(x[a]);
    int st[3] = {0, 0, 0};
    T off3[3];
    uint64_t key = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        int64_t start;
        double off;
        support_start<T>(x[a], reach, g, a, &start, &off);
        const int64_t ws = wrapped_start(start, g.ax[a].n);
        st[a] = (int)ws;
        off3[a] = (T)off;
        key += (uint64_t)(ws >> g.ax[a].shift) * (uint64_t)g.ax[a].stride;
    }
    int64_t dst = s;
    const int64_t lo = offsets[key], hi = offsets[key + 1];
    if (hi - lo <= kSmallBin && hi - lo > 1) {
        int64_t rank = 0;
        for (int64_t m = lo; m < hi; ++m) {
            const SortElem e = elems[m];
            rank += (m != s && elem_less(e, me)) ? 1 : 0;
        }
        dst = lo + rank;
    }
#pragma unroll
    for (int a = 0; a < D; ++a) rec_off[dst * 4 + a] = off3[a];
    rec_start[dst] = make_int4(st[0], st[1], st[2], (int)me.idx);
    if (out_sorted) out_sorted[dst] = me.idx;
}

// Strengths of every vector gathered into record order.
template <typename T>
__global__ __launch_bounds__(256) void k_gather_strengths(int64_t M, int64_t vectors,
                                                          const int4 *__restrict__ rec_start,
                                                          const T *__restrict__ c,
                                                          T *__restrict__ c_sorted)
{
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    int64_t idx = (uint32_t)rec_start[s].w;
    for (int64_t v = 0; v < vectors; ++v) {
        const T *src = c + 2 * (v * M + idx);
        T *dst = c_sorted + 2 * (v * M + s);
        dst[0] = src[0];
        dst[1] = src[1];
    }
}

inline unsigned grid_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

}  // namespace

// ---------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------
BinSorter::~BinSorter() { release(); }

void BinSorter::release()
{
    for (void *p : {(void *)counts_, (void *)offsets_, (void *)keys_, (void *)ranks_,
                    (void *)sorted_, (void *)large_list_, (void *)tile_off_, (void *)tile_sums_,
                    (void *)scratch_a_, (void *)scratch_b_})
        if (p) (void)hipFree(p);
    counts_ = offsets_ = keys_ = ranks_ = sorted_ = large_list_ = tile_off_ = tile_sums_ = nullptr;
    scratch_a_ = scratch_b_ = nullptr;
    bins_cap_ = points_cap_ = sort_scratch_cap_ = 0;
}

bool BinSorter::ensure(int64_t nbins, int64_t M, bool need_sort_scratch)
{
    if (nbins > bins_cap_) {
        if (counts_) (void)hipFree(counts_);
        if (offsets_) (void)hipFree(offsets_);
        if (tile_sums_) (void)hipFree(tile_sums_);
        counts_ = offsets_ = tile_sums_ = nullptr;
        // counts: nbins + trailing zero, then the large-bin count and the
        // largest large bin
        if (!hip_ok(hipMalloc(&counts_, sizeof(uint32_t) * (nbins + 3)), "bin counts")) return false;
        if (!hip_ok(hipMalloc(&offsets_, sizeof(uint32_t) * (nbins + 1)), "bin offsets")) return false;
        int64_t tiles = (nbins + 1 + kScanTile - 1) / kScanTile;
        if (!hip_ok(hipMalloc(&tile_sums_, sizeof(uint32_t) * (tiles + 1)), "scan tiles")) return false;
        bins_cap_ = nbins;
    }
    if (M > points_cap_) {
        for (uint32_t **p : {&keys_, &ranks_, &sorted_, &large_list_, &tile_off_}) {
            if (*p) (void)hipFree(*p);
            *p = nullptr;
        }
        int64_t cap = M + M / 4 + 1024;
        if (!hip_ok(hipMalloc(&keys_, sizeof(uint32_t) * cap), "point keys")) return false;
        if (!hip_ok(hipMalloc(&ranks_, sizeof(uint32_t) * cap), "point ranks")) return false;
        if (!hip_ok(hipMalloc(&sorted_, sizeof(uint32_t) * cap), "sorted points")) return false;
        if (!hip_ok(hipMalloc(&large_list_, sizeof(uint32_t) * (cap / (kSmallBin + 1) + 2)),
                    "large-bin list"))
            return false;
        if (!hip_ok(hipMalloc(&tile_off_, sizeof(uint32_t) * (cap / (kSmallBin + 1) + 3)),
                    "large-bin tiles"))
            return false;
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

template <int D, typename T>
bool BinSorter::sort(int64_t M, const T *coords, const GridParams &g, BinOrder order,
                     hipStream_t stream)
{
    nbins_ = g.nbins;
    M_ = M;
    if (!ensure(g.nbins, M, order != BinOrder::Grouped)) return false;
    // Clear counts, trailing zero and the large-bin count with a fill kernel
    // (no memset nodes, so callers may capture the sequence in a graph).
    k_fill_u32<<<grid_for(g.nbins + 3, 256), 256, 0, stream>>>(counts_, g.nbins + 3, 0u);
    if (M == 0) {
        k_fill_u32<<<grid_for(g.nbins + 1, 256), 256, 0, stream>>>(offsets_, g.nbins + 1, 0u);
        return hip_ok(hipGetLastError(), "bin clear");
    }
    uint32_t *large_count = counts_ + g.nbins + 1;
    uint32_t *large_max = counts_ + g.nbins + 2;
    k_count<D, T><<<grid_for(M, kCountPoints), kCountThreads, 0, stream>>>(
        M, coords, g, counts_, keys_, ranks_, large_list_, large_count);
    int64_t n = g.nbins + 1;
    int64_t tiles = (n + kScanTile - 1) / kScanTile;
    k_scan_reduce<<<(unsigned)tiles, kScanThreads, 0, stream>>>(counts_, n, tile_sums_);
    k_scan_top<<<1, kScanThreads, 0, stream>>>(tile_sums_, tiles);
    k_scan_down<<<(unsigned)tiles, kScanThreads, 0, stream>>>(counts_, n, tile_sums_, offsets_);
    if (order == BinOrder::Grouped) {
        k_scatter<<<grid_for(M, 256), 256, 0, stream>>>(M, keys_, ranks_, offsets_, sorted_);
    } else {
        unsigned grid = (unsigned)std::min<int64_t>(2048, M / kSortRun + 64);
        SortElem *A = (SortElem *)scratch_a_, *B = (SortElem *)scratch_b_;
        k_scatter_elems<D, T><<<grid_for(M, 256), 256, 0, stream>>>(M, coords, keys_, ranks_,
                                                                    offsets_, (int)order, A);
        k_large_tiles<<<1, 1024, 0, stream>>>(large_list_, large_count, offsets_, tile_off_,
                                              large_max);
        k_large_runs<<<grid, kRunThreads, 0, stream>>>(large_list_, large_count, tile_off_,
                                                       offsets_, A);
        // rounds for the largest possible bin; bins done earlier skip
        int rounds = 0;
        for (int64_t len = kSortRun; len < M; len <<= 1) ++rounds;
        for (int r = 0; r < rounds; ++r)
            k_large_merge<<<grid, kSortThreads, 0, stream>>>(large_list_, large_count, large_max,
                                                             tile_off_,
                                                             offsets_, r, (r & 1) ? B : A,
                                                             (r & 1) ? A : B);
        if (rounds > 0)
            k_large_final<<<grid, kSortThreads, 0, stream>>>(large_list_, large_count, tile_off_,
                                                             offsets_, A, B);
    }
    return hip_ok(hipGetLastError(), "bin sort launch");
}

template <int D, typename T>
bool BinSorter::prepare(int64_t M, const T *coords, const GridParams &g, BinOrder order,
                        int4 *rec_start, T *rec_off, uint32_t *out_sorted, hipStream_t stream)
{
    if (M == 0) return true;
    if (order == BinOrder::Grouped)
        k_prepare_grouped<D, T><<<grid_for(M, 256), 256, 0, stream>>>(M, sorted_, coords, g,
                                                                      rec_start, rec_off,
                                                                      out_sorted);
    else
        k_prepare<D, T><<<grid_for(M, 256), 256, 0, stream>>>(
            M, (const SortElem *)scratch_a_, offsets_, coords, g, (int)order, rec_start, rec_off,
            out_sorted);
    return hip_ok(hipGetLastError(), "prepare launch");
}

template <typename T>
bool gather_strengths(int64_t M, int64_t vectors, const int4 *rec_start, const T *c, T *c_sorted,
                      hipStream_t stream)
{
    if (M == 0) return true;
    k_gather_strengths<T><<<grid_for(M, 256), 256, 0, stream>>>(M, vectors, rec_start, c,
                                                                c_sorted);
    return hip_ok(hipGetLastError(), "strength gather launch");
}

#define MNUFFT_BINS_INSTANTIATE(D, T)                                                             \
    template bool BinSorter::sort<D, T>(int64_t, const T *, const GridParams &, BinOrder,          \
                                        hipStream_t);                                             \
    template bool BinSorter::prepare<D, T>(int64_t, const T *, const GridParams &, BinOrder,       \
                                           int4 *, T *, uint32_t *, hipStream_t);
MNUFFT_BINS_INSTANTIATE(1, float)
MNUFFT_BINS_INSTANTIATE(2, float)
MNUFFT_BINS_INSTANTIATE(3, float)
MNUFFT_BINS_INSTANTIATE(1, double)
MNUFFT_BINS_INSTANTIATE(2, double)
MNUFFT_BINS_INSTANTIATE(3, double)
template bool gather_strengths<float>(int64_t, int64_t, const int4 *, const float *, float *,
                                      hipStream_t);
template bool gather_strengths<double>(int64_t, int64_t, const int4 *, const double *, double *,
                                       hipStream_t);

}  // namespace mnufft
