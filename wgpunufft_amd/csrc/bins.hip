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
// MI355X design: one 64-lane wavefront per 64 points for the count and
// scatter passes (atomics land in the 256 MB Infinity Cache for bin arrays of
// a few tens of MB); a three-kernel reduce-then-scan over 4096-element tiles;
// bins above 64 points get a workgroup each that sorts 2048-element runs in
// LDS and merges them along merge paths.
#include <hip/hip_runtime.h>

#include "bins.hpp"
#include "fold.hpp"

namespace mnufft {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;  // 4096
constexpr int kSortThreads = 256;
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
    uint32_t my_key[kCountPerThread], my_slot[kCountPerThread], my_rank[kCountPerThread];
#pragma unroll
    for (int i = 0; i < kCountPerThread; ++i) {
        int64_t j = base + (int64_t)i * kCountThreads + threadIdx.x;
        my_slot[i] = kEmpty;
        if (j >= M) continue;
        uint32_t key = (uint32_t)point_bin_key<D, T>(coords, j, g);
        my_key[i] = key;
        uint32_t h = (key * 2654435761u) & (kHashSize - 1);
        for (;;) {
            uint32_t old = atomicCAS(&ht_key[h], kEmpty, key);
            if (old == kEmpty || old == key) break;
            h = (h + 1) & (kHashSize - 1);
        }
        my_slot[i] = h;
        my_rank[i] = atomicAdd(&ht_cnt[h], 1u);
    }
    __syncthreads();
    for (int h = threadIdx.x; h < kHashSize; h += kCountThreads) {
        uint32_t key = ht_key[h];
        if (key == kEmpty) continue;
        uint32_t n = ht_cnt[h];
        uint32_t old = atomicAdd(&counts[key], n);
        ht_base[h] = old;
        if (old <= kSmallBin && old + n > kSmallBin) {
            uint32_t slot = atomicAdd(large_count, 1u);
            large_list[slot] = key;
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

// ---------------------------------------------------------------------
// Large bins: one workgroup per listed bin.  Runs of kSortRun elements are
// bitonic-sorted in LDS; longer bins are then merged pairwise through a
// global scratch along merge paths.
// ---------------------------------------------------------------------
__device__ inline void lds_bitonic(SortElem *s, int n2)
{
    for (int k = 2; k <= n2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n2; i += kSortThreads) {
                int p = i ^ j;
                if (p > i) {
                    bool up = (i & k) == 0;
                    SortElem a = s[i], b = s[p];
                    if (elem_less(b, a) == up) {
                        s[i] = b;
                        s[p] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Index of the merge-path split on diagonal `diag` of runs a[0..na), b[0..nb).
__device__ inline int64_t merge_split(const SortElem *a, int64_t na, const SortElem *b,
                                      int64_t nb, int64_t diag)
{
    int64_t lo = diag > nb ? diag - nb : 0;
    int64_t hi = diag < na ? diag : na;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        // take a[mid] before b[diag-1-mid] iff a[mid] < b[diag-1-mid]
        if (elem_less(a[mid], b[diag - 1 - mid]))
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

template <int D, typename T>
__global__ __launch_bounds__(kSortThreads) void k_large_sort(
    const uint32_t *__restrict__ large_list, const uint32_t *__restrict__ large_count,
    const uint32_t *__restrict__ offsets, const T *__restrict__ coords, int order,
    uint32_t *__restrict__ sorted, SortElem *__restrict__ scratch_a,
    SortElem *__restrict__ scratch_b)
{
    __shared__ SortElem s[kSortRun];
    uint32_t nlarge = *large_count;
    for (uint32_t lb = blockIdx.x; lb < nlarge; lb += gridDim.x) {
        uint32_t bin = large_list[lb];
        int64_t lo = offsets[bin], hi = offsets[bin + 1];
        int64_t n = hi - lo;
        SortElem *A = scratch_a + lo;
        SortElem *B = scratch_b + lo;
        // Sorted runs of kSortRun elements.
        for (int64_t r0 = 0; r0 < n; r0 += kSortRun) {
            int m = (int)(n - r0 < kSortRun ? n - r0 : kSortRun);
            int n2 = 1;
            while (n2 < m) n2 <<= 1;
            for (int i = threadIdx.x; i < n2; i += kSortThreads) {
                if (i < m)
                    s[i] = make_elem<D, T>(coords, sorted[lo + r0 + i], order);
                else
                    s[i] = SortElem{0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu};
            }
            __syncthreads();
            lds_bitonic(s, n2);
            for (int i = threadIdx.x; i < m; i += kSortThreads) {
                if (n <= kSortRun)
                    sorted[lo + i] = s[i].idx;
                else
                    A[r0 + i] = s[i];
            }
            __syncthreads();
        }
        if (n <= kSortRun) continue;
        // Pairwise merges of runs of length `len`.
        for (int64_t len = kSortRun; len < n; len <<= 1) {
            for (int64_t p0 = 0; p0 < n; p0 += 2 * len) {
                int64_t na = n - p0 < len ? n - p0 : len;
                int64_t nb = n - p0 - na;
                if (nb < 0) nb = 0;
                if (nb > len) nb = len;
                const SortElem *a = A + p0;
                const SortElem *b = a + na;
                int64_t total = na + nb;
                int64_t per = (total + kSortThreads - 1) / kSortThreads;
                int64_t d0 = per * threadIdx.x;
                if (d0 < total) {
                    int64_t d1 = d0 + per < total ? d0 + per : total;
                    int64_t ia = merge_split(a, na, b, nb, d0);
                    int64_t ib = d0 - ia;
                    for (int64_t o = d0; o < d1; ++o) {
                        bool take_a = ib >= nb || (ia < na && !elem_less(b[ib], a[ia]));
                        B[p0 + o] = take_a ? a[ia++] : b[ib++];
                    }
                }
            }
            __syncthreads();
            SortElem *t = A;
            A = B;
            B = t;
        }
        for (int64_t i = threadIdx.x; i < n; i += kSortThreads) sorted[lo + i] = A[i].idx;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------
// Prepare: final slot of every point (rank inside small bins), its
// wrapped support starts and offsets, and the original index.
// ---------------------------------------------------------------------
template <int D, typename T>
__global__ __launch_bounds__(256) void k_prepare(int64_t M, const uint32_t *__restrict__ sorted,
                                                 const uint32_t *__restrict__ keys,
                                                 const uint32_t *__restrict__ offsets,
                                                 const T *__restrict__ coords, GridParams g,
                                                 int order, int4 *__restrict__ rec_start,
                                                 T *__restrict__ rec_off,
                                                 uint32_t *__restrict__ out_sorted)
{
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    uint32_t idx = sorted[s];
    int64_t dst = s;
    if (order != (int)BinOrder::Grouped) {
        uint32_t key = keys[idx];
        int64_t lo = offsets[key], hi = offsets[key + 1];
        if (hi - lo <= kSmallBin && hi - lo > 1) {
            SortElem me = make_elem<D, T>(coords, idx, order);
            int64_t rank = 0;
            for (int64_t m = lo; m < hi; ++m) {
                uint32_t other = sorted[m];
                if (other == idx) continue;
                SortElem e = make_elem<D, T>(coords, other, order);
                rank += elem_less(e, me) ? 1 : 0;
            }
            dst = lo + rank;
        }
    }
    bool reach = point_in_reach<D, T>(coords, idx);
    int st[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < D; ++a) {
        int64_t start;
        double off;
        support_start<T>(coords[(int64_t)idx * D + a], reach, g, a, &start, &off);
        st[a] = (int)wrapped_start(start, g.ax[a].n);
        rec_off[dst * 4 + a] = (T)off;
    }
    rec_start[dst] = make_int4(st[0], st[1], st[2], (int)idx);
    if (out_sorted) out_sorted[dst] = idx;
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
                    (void *)sorted_, (void *)large_list_, (void *)tile_sums_,
                    (void *)scratch_a_, (void *)scratch_b_})
        if (p) (void)hipFree(p);
    counts_ = offsets_ = keys_ = ranks_ = sorted_ = large_list_ = tile_sums_ = nullptr;
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
        // counts: nbins + trailing zero, then the large-bin count word
        if (!hip_ok(hipMalloc(&counts_, sizeof(uint32_t) * (nbins + 2)), "bin counts")) return false;
        if (!hip_ok(hipMalloc(&offsets_, sizeof(uint32_t) * (nbins + 1)), "bin offsets")) return false;
        int64_t tiles = (nbins + 1 + kScanTile - 1) / kScanTile;
        if (!hip_ok(hipMalloc(&tile_sums_, sizeof(uint32_t) * (tiles + 1)), "scan tiles")) return false;
        bins_cap_ = nbins;
    }
    if (M > points_cap_) {
        for (uint32_t **p : {&keys_, &ranks_, &sorted_, &large_list_}) {
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
    k_fill_u32<<<grid_for(g.nbins + 2, 256), 256, 0, stream>>>(counts_, g.nbins + 2, 0u);
    if (M == 0) {
        k_fill_u32<<<grid_for(g.nbins + 1, 256), 256, 0, stream>>>(offsets_, g.nbins + 1, 0u);
        return hip_ok(hipGetLastError(), "bin clear");
    }
    uint32_t *large_count = counts_ + g.nbins + 1;
    k_count<D, T><<<grid_for(M, kCountPoints), kCountThreads, 0, stream>>>(
        M, coords, g, counts_, keys_, ranks_, large_list_, large_count);
    int64_t n = g.nbins + 1;
    int64_t tiles = (n + kScanTile - 1) / kScanTile;
    k_scan_reduce<<<(unsigned)tiles, kScanThreads, 0, stream>>>(counts_, n, tile_sums_);
    k_scan_top<<<1, kScanThreads, 0, stream>>>(tile_sums_, tiles);
    k_scan_down<<<(unsigned)tiles, kScanThreads, 0, stream>>>(counts_, n, tile_sums_, offsets_);
    k_scatter<<<grid_for(M, 256), 256, 0, stream>>>(M, keys_, ranks_, offsets_, sorted_);
    if (order != BinOrder::Grouped) {
        unsigned grid = (unsigned)std::min<int64_t>(1024, M / (kSmallBin + 1) + 1);
        k_large_sort<D, T><<<grid, kSortThreads, 0, stream>>>(
            large_list_, large_count, offsets_, coords, (int)order, sorted_,
            (SortElem *)scratch_a_, (SortElem *)scratch_b_);
    }
    return hip_ok(hipGetLastError(), "bin sort launch");
}

template <int D, typename T>
bool BinSorter::prepare(int64_t M, const T *coords, const GridParams &g, BinOrder order,
                        int4 *rec_start, T *rec_off, uint32_t *out_sorted, hipStream_t stream)
{
    if (M == 0) return true;
    k_prepare<D, T><<<grid_for(M, 256), 256, 0, stream>>>(M, sorted_, keys_, offsets_, coords, g,
                                                          (int)order, rec_start, rec_off,
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
