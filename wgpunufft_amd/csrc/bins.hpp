// Point binning (count, scan, scatter, in-bin order, prepare).
#pragma once

#include "nufft_internal.hpp"

namespace mnufft {

// Device buffers of one plan's binning, grown on demand and reused by every
// execution in stream order.
class BinSorter {
public:
    ~BinSorter();
    void release();

    // Bins the M points (point-major coordinates on the device) and orders
    // each bin; afterwards offsets() holds the nbins + 1 bin offsets (the
    // last one is M) and, for the grouped order, sorted() the permutation.
    // Ordered sorts keep each bin's ordering elements in scratch for
    // prepare(), which writes the permutation.
    template <int D, typename T>
    bool sort(int64_t M, const T *coords, const GridParams &g, BinOrder order,
              hipStream_t stream);

    // Prepared records in final slot order: rec_start[slot] = (wrapped start
    // per axis, original index), rec_off[4 * slot + axis] = start - position.
    // out_sorted (optional) receives the final permutation.  With
    // dense_before (dense records before each bin, nbins + 1 values; a bin
    // with any is dense) the light bins' records come first, in bin order,
    // then the dense bins' records, in bin order.
    template <int D, typename T>
    bool prepare(int64_t M, const T *coords, const GridParams &g, BinOrder order,
                 int4 *rec_start, T *rec_off, uint32_t *out_sorted, hipStream_t stream,
                 const uint32_t *dense_before = nullptr);

    // One sort over the concatenated points of nprob problems (device
    // table): sorted() then holds indices of the concatenated set, grouped
    // by problem and, inside a problem, by bin, and offsets() the nbins + 1
    // offsets of the combined bin list (nbins: all problems' bins).
    template <int D, typename T>
    bool sort_many(int64_t M, const BatchProblem *probs, int nprob, uint64_t nbins,
                   hipStream_t stream);

    // Index-order sorts in a caller-owned device region instead of buffers
    // of their own: the keys, values, histograms, counters and gap lists are
    // carved from [base, base + bytes) (nothing is allocated or freed), the
    // final permutation goes to sorted_out (M values) and the offsets to
    // offsets_out (nbins + 1 values).  False, with the last error set, when
    // the region is too small for M points and nbins bins.
    bool bind_external(void *base, size_t bytes, int64_t nbins, int64_t M, uint32_t *sorted_out,
                       uint32_t *offsets_out);
    // Bytes bind_external needs for M points and nbins bins.
    static size_t external_bytes(int64_t nbins, int64_t M);

    const uint32_t *sorted() const { return sorted_; }
    const uint32_t *offsets() const { return offsets_; }
    int64_t nbins() const { return nbins_; }

private:
    bool ensure(int64_t nbins, int64_t M, bool need_sort_scratch);
    template <int D, typename T, typename K>
    void radix_passes(int64_t M, const T *coords, const GridParams &g, int mode, int sub_bits,
                      int bits, hipStream_t stream);
    // LSD passes over `bits` key bits; keys0(ntiles, keys) launches pass 0
    // (writes the keys and the first digit's per-tile histogram).
    template <typename K, typename F>
    void radix_loop(int64_t M, int bits, hipStream_t stream, F keys0);

    uint32_t *ctrl_ = nullptr;     // large bins, largest large bin, mid gaps, long gaps
    uint32_t *offsets_ = nullptr;
    uint4 *gaps_ = nullptr;        // gap lists of the offsets (mid, then long)
    int64_t gaps_mid_cap_ = 0;
    uint32_t *keys_ = nullptr;     // radix key buffers (ping-pong)
    uint32_t *keys_b_ = nullptr;
    uint32_t *vals_ = nullptr;     // radix value buffer paired with sorted_
    uint32_t *sorted_ = nullptr;
    uint32_t *hist_ = nullptr;     // digit-major per-tile digit counts
    const void *sorted_keys_ = nullptr;  // sort keys in slot order (keys_ or keys_b_)
    uint32_t *hist_sums_ = nullptr;
    uint32_t *large_list_ = nullptr;  // (first, end) of the groups above kSmallBin
    uint32_t *tile_off_ = nullptr;
    void *scratch_a_ = nullptr;
    void *scratch_b_ = nullptr;
    int64_t bins_cap_ = 0, points_cap_ = 0, sort_scratch_cap_ = 0;
    bool external_ = false;  // buffers bound by bind_external (never freed here)
    int64_t nbins_ = 0, M_ = 0;
};

// Records of a multi-plan type-1 batch, from the permutation of
// BinSorter::sort_many (indices of the concatenated point set; a problem's
// points take the slots [m_off, m_off + M)): rec_start / rec_off as
// prepare() writes them with the problem's geometry, rec_perm[slot] = the
// point's index inside its problem.
template <int D, typename T>
bool prepare_many(int64_t M, const uint32_t *sorted, const BatchProblem *probs, int nprob,
                  int4 *rec_start, T *rec_off, uint32_t *rec_perm, hipStream_t stream);
// c_sorted[s] = the problem's strength (probs[q].in) of the point at slot s
// (one vector per problem).
template <typename T>
bool gather_strengths_many(int64_t M, const uint32_t *rec_perm, const BatchProblem *probs,
                           int nprob, T *c_sorted, hipStream_t stream);

// Device address, on the current device, of the library's static sort
// scratch (a file-scope device array of static_sort_bytes() bytes), looked
// up once per device; nullptr (last error set) when the lookup fails.
void *static_sort_scratch();
size_t static_sort_bytes();

// Exclusive scan of n u32 values (wrapping) into out (may alias in); sums
// holds scan_tiles(n) values of scratch.
void scan_u32(const uint32_t *in, int64_t n, uint32_t *sums, uint32_t *out, hipStream_t stream);
int64_t scan_tiles(int64_t n);

// c_sorted[v][s] = c[v][perm[s]] for every vector v, for the first M - *tail
// slots when tail (a device value) is given, else all M.
template <typename T>
bool gather_strengths(int64_t M, int64_t vectors, const uint32_t *perm, const T *c, T *c_sorted,
                      hipStream_t stream, const uint32_t *tail = nullptr);

}  // namespace mnufft
