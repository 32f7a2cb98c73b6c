// Host replay of the addressing of the matrix-core tile spreads and their
// heavy-part reduction (k_spread_mfma2 / k_spread_mfma3 and
// k_reduce_tiles), for a given bin geometry and bin offsets.  It runs the
// engine's own index code (tile_index.hpp: axis reach, reach rows, record
// slots, work items, partial-sum indices) on the host, restates the work
// plan of k_classify / k_plan, and checks every global index those kernels
// form against the size of the buffer it addresses, and every 32-bit
// intermediate against its range.  Test infrastructure only.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../wgpunufft_amd/csrc/tile_index.hpp"

using namespace mnufft;

namespace {

struct Report {
    int64_t violations = 0, first_code = 0, first_item = -1;
    int64_t items = 0, heavy_blocks = 0, heavy_parts = 0, part_len = 0;
    int64_t max_slot = -1, max_partial = -1, partial_floats = 0, records = 0, max_total = 0;
    int64_t max_offset_index = -1;
    void fail(int code, int64_t item)
    {
        if (violations++ == 0) {
            first_code = code;
            first_item = item;
        }
    }
};

enum Code : int {
    kOffsetIndex = 1,   // offsets read past nbins
    kTotals = 2,        // classify total differs from the staged rows' total
    kSlot = 3,          // record slot >= M
    kPartial = 4,       // partial-sum index >= the partial buffer
    kWide = 5,          // a 32-bit quantity overflows
    kItem = 6,          // work item decodes outside its tile's parts
    kGrid = 7,          // grid index outside the fine grid
};

GridParams make_grid(int dim, int width, const int64_t *fine, const int64_t *side)
{
    GridParams g{};
    g.dims = dim;
    g.width = width;
    g.nbins = 1;
    int64_t fine_total = 1;
    for (int a = 0; a < kMaxDims; ++a) {
        AxisParams &x = g.ax[a];
        if (a >= dim) {
            x.n = 1;
            x.bins = 1;
            x.shift = 0;
            x.stride = g.nbins;
            continue;
        }
        x.n = fine[a];
        int s = 0;
        while (((int64_t)1 << s) < side[a]) ++s;
        x.shift = s;
        x.bins = (fine[a] + side[a] - 1) / side[a];
        x.stride = g.nbins;
        g.nbins *= x.bins;
        fine_total *= fine[a];
    }
    g.fine_total = fine_total;
    return g;
}

// k_plan restated: the light limit doubled until the heavy blocks fit half
// the part slots, the part length from the heavy records over the spare
// slots, then the ordered scans of items and heavy parts.
void plan_items(const std::vector<uint32_t> &totals, uint32_t light_limit, uint32_t min_part,
                int64_t part_capacity, std::vector<uint32_t> &parts,
                std::vector<uint32_t> &item_off, std::vector<uint32_t> &heavy_base,
                std::vector<uint32_t> &heavy_list, uint32_t ctrl[4], Report &rep)
{
    const int64_t nb = (int64_t)totals.size();
    uint32_t limit = light_limit;
    uint64_t heavy_records = 0;
    uint32_t heavy_blocks = 0;
    for (;;) {
        heavy_records = 0;
        heavy_blocks = 0;
        for (uint32_t t : totals)
            if (t > limit) {
                heavy_records += t;
                ++heavy_blocks;
            }
        if ((int64_t)heavy_blocks * 2 <= part_capacity || limit >= 0x7fffffffu) break;
        limit *= 2;
    }
    uint64_t spare = part_capacity > (int64_t)heavy_blocks
                         ? (uint64_t)(part_capacity - heavy_blocks)
                         : 1ull;
    uint64_t part = (heavy_records + spare - 1) / spare;
    if (part < min_part) part = min_part;
    const uint32_t part32 = (uint32_t)(part > 0xffffffffull ? 0xffffffffull : part);
    parts.assign(nb, 0);
    item_off.assign(nb + 1, 0);
    heavy_base.assign(nb + 1, 0);
    heavy_list.clear();
    uint64_t items = 0, hparts = 0;
    for (int64_t b = 0; b < nb; ++b) {
        uint32_t np = 1, hp = 0;
        if (totals[b] > limit) {
            np = (uint32_t)((totals[b] + (uint64_t)part32 - 1) / part32);
            if (np > 1)
                hp = np;
            else
                np = 1;
        }
        parts[b] = np;
        item_off[b] = (uint32_t)items;
        heavy_base[b] = (uint32_t)hparts;
        if (items > 0xffffffffull || hparts > 0xffffffffull) rep.fail(kWide, -1);
        items += np;
        hparts += hp;
        if (hp) heavy_list.push_back((uint32_t)b);
    }
    ctrl[0] = (uint32_t)items;
    ctrl[1] = (uint32_t)hparts;
    ctrl[2] = (uint32_t)heavy_list.size();
    ctrl[3] = part32;
    if (items > 0xffffffffull) rep.fail(kWide, -1);
}

template <int D, int TX, int TY, int TZ, int W>
void replay(const GridParams &g, const SpreadParams &sp, const uint32_t *offsets, int64_t M,
            uint32_t light_limit, uint32_t min_part, int64_t budget_bytes, int64_t vectors,
            Report &rep)
{
    constexpr int NRY = TY + W - 1;
    constexpr int NROWS = D == 3 ? NRY * (TZ + W - 1) : NRY;
    const int64_t nb = sp.block_count;
    // k_classify: records in reach of every tile
    std::vector<uint32_t> totals((size_t)nb);
    for (int64_t b = 0; b < nb; ++b) {
        BlockReach<D> r = block_reach<D>(b, g, sp);
        uint64_t total = 0;
        for (int64_t t = 0; t < r.nseg; ++t) {
            int64_t f, l;
            r.segment(t, g, offsets, &f, &l);
            total += (uint64_t)(l - f);
        }
        if (total > 0xffffffffull) rep.fail(kWide, -1);
        totals[(size_t)b] = (uint32_t)std::min<uint64_t>(total, 0xffffffffull);
        rep.max_total = std::max<int64_t>(rep.max_total, (int64_t)total);
    }
    // the spread's part slots (Spreader::spread): one vector's slots, the
    // vectors in chunks that keep the scratch within 1 GiB
    const int64_t cells = (int64_t)TX * TY * TZ;
    const int64_t per_part = cells * 2 * (int64_t)sizeof(float);
    const int64_t part_capacity = std::max<int64_t>(budget_bytes / per_part, 64);
    const int64_t chunk =
        std::max<int64_t>(1, std::min<int64_t>(vectors, ((int64_t)1 << 30) /
                                                            (part_capacity * per_part)));
    rep.partial_floats = part_capacity * per_part * chunk / (int64_t)sizeof(float);
    std::vector<uint32_t> parts, item_off, heavy_base, heavy_list;
    uint32_t ctrl[4];
    plan_items(totals, light_limit, min_part, part_capacity, parts, item_off, heavy_base,
               heavy_list, ctrl, rep);
    rep.items = ctrl[0];
    rep.heavy_parts = ctrl[1];
    rep.heavy_blocks = ctrl[2];
    rep.part_len = ctrl[3];
    // tile kernel: every work item of the plan, every record it visits
    TileRows<NROWS> rows;
    for (int64_t item = 0; item < (int64_t)ctrl[0]; ++item) {
        const TileItem ti = tile_item(item, nb, item_off.data(), parts.data(), totals.data(),
                                      ctrl[3]);
        if (ti.b < 0 || ti.b >= nb || ti.p < 0 || ti.p >= (int64_t)ti.np || ti.rlo > ti.rhi) {
            rep.fail(kItem, item);
            continue;
        }
        const int64_t b = ti.b;
        const int64_t x0 = (b % sp.nblocks[0]) * TX;
        const int64_t y0 = ((b / sp.nblocks[0]) % sp.nblocks[1]) * TY;
        const int64_t z0 = D == 3 ? (b / (sp.nblocks[0] * sp.nblocks[1])) * TZ : 0;
        const AxisReach rx = axis_reach(x0, TX, W, g.ax[0]);
        uint64_t carry = 0;
        for (int r = 0; r < NROWS; ++r) {
            const int64_t row = tile_row_base<D, NRY, W>(r, y0, z0, g);
            int64_t idx[4] = {row + rx.lo0, row + rx.lo0 + rx.len0, row + rx.lo1,
                              row + rx.lo1 + rx.len1};
            const int nidx = rx.len1 > 0 ? 4 : 2;
            for (int i = 0; i < nidx; ++i) {
                rep.max_offset_index = std::max(rep.max_offset_index, idx[i]);
                if (idx[i] < 0 || idx[i] > g.nbins) rep.fail(kOffsetIndex, item);
            }
            if (rep.violations) return;
            const uint32_t f0 = offsets[idx[0]], c0 = offsets[idx[1]];
            const uint32_t f1 = rx.len1 > 0 ? offsets[idx[2]] : 0u;
            const uint32_t c1 = rx.len1 > 0 ? offsets[idx[3]] : 0u;
            rows.f0[r] = f0;
            rows.n0[r] = c0 - f0;
            rows.f1[r] = f1;
            rows.excl[r] = (uint32_t)carry;
            carry += (uint64_t)(c0 - f0) + (uint64_t)(c1 - f1);
            if (carry > 0xffffffffull) rep.fail(kWide, item);
        }
        rows.excl[NROWS] = (uint32_t)carry;
        if ((int64_t)carry != (int64_t)totals[(size_t)b] || ti.rhi > (int64_t)carry)
            rep.fail(kTotals, item);
        for (int64_t k = ti.rlo; k < ti.rhi; ++k) {
            int q = 0;
            const uint32_t slot = tile_slot<NROWS>(rows, (uint32_t)k, &q);
            if (q < 0 || q >= NROWS || (int64_t)slot >= M) {
                rep.fail(kSlot, item);
                break;
            }
            rep.max_slot = std::max<int64_t>(rep.max_slot, slot);
        }
        rep.records += ti.rhi - ti.rlo;
        // the item's output: light tiles write the grid (guarded per cell),
        // heavy parts their slot of every vector of the chunk
        for (int64_t v = 0; v < chunk; ++v) {
            if (ti.np == 1) {
                const int64_t gx = std::min<int64_t>(x0 + TX - 1, g.ax[0].n - 1);
                const int64_t gy = std::min<int64_t>(y0 + TY - 1, g.ax[1].n - 1);
                const int64_t gz = std::min<int64_t>(z0 + TZ - 1, g.ax[2].n - 1);
                const int64_t gi =
                    2 * (v * g.fine_total + gx + g.ax[0].n * (gy + g.ax[1].n * gz)) + 1;
                if (gi >= 2 * chunk * g.fine_total) rep.fail(kGrid, item);
            } else {
                const int64_t pslot = (int64_t)heavy_base[(size_t)b] + ti.p;
                const int64_t pi = partial_index<TX, TY, TZ>(pslot, chunk, v, TX - 1, TY - 1,
                                                             TZ - 1) + 1;
                rep.max_partial = std::max(rep.max_partial, pi);
                if (pi >= rep.partial_floats) rep.fail(kPartial, item);
            }
        }
        if (rep.violations) return;
    }
    // k_reduce_tiles: the parts of every heavy tile, every vector of a chunk
    for (size_t hb = 0; hb < heavy_list.size(); ++hb) {
        const int64_t b = heavy_list[hb];
        if (b < 0 || b >= nb) {
            rep.fail(kItem, -1);
            return;
        }
        for (int64_t q = 0; q < (int64_t)parts[(size_t)b]; ++q) {
            const int64_t pi = partial_index<TX, TY, TZ>((int64_t)heavy_base[(size_t)b] + q, chunk,
                                                         chunk - 1, TX - 1, TY - 1, TZ - 1) + 1;
            if (pi >= rep.partial_floats) rep.fail(kPartial, -1);
        }
    }
}

template <int D, int TX, int TY, int TZ>
bool replay_width(int W, const GridParams &g, const SpreadParams &sp, const uint32_t *offsets,
                  int64_t M, uint32_t light_limit, uint32_t min_part, int64_t budget,
                  int64_t vectors, Report &rep)
{
    switch (W) {
#define REPLAY_W(w)                                                                           \
    case w:                                                                                   \
        replay<D, TX, TY, TZ, w>(g, sp, offsets, M, light_limit, min_part, budget, vectors, rep); \
        return true;
        REPLAY_W(2) REPLAY_W(3) REPLAY_W(4) REPLAY_W(5) REPLAY_W(6) REPLAY_W(7) REPLAY_W(8)
#undef REPLAY_W
    default:
        return false;
    }
}

}  // namespace

extern "C" {

// dim 2: 32 x 32 tiles (k_spread_mfma2), or 16 x 16 (k_spread_rows2) with
// edge2 16; dim 3: 16 x 16 x 8 tiles (k_spread_mfma3).  offsets: nbins + 1
// bin offsets of the row geometry `side`.  report (16 values): violations,
// first violation code, its work item, items, heavy tiles, heavy parts, part
// length, largest slot, largest partial index, partial floats, records
// visited, largest tile total, largest offsets index, nbins.  Returns -1 for
// an unsupported shape.
int tile_replay_edge(int dim, int edge2, int width, const int64_t *fine, const int64_t *side,
                     const uint32_t *offsets, int64_t M, uint32_t light_limit, uint32_t min_part,
                     int64_t budget_bytes, int64_t vectors, int64_t *report)
{
    if (dim == 2 && edge2 != 16 && edge2 != 32) return -1;
    GridParams g = make_grid(dim, width, fine, side);
    SpreadParams sp{};
    const int64_t tile[3] = {dim == 2 ? edge2 : 16, dim == 2 ? edge2 : 16, dim == 2 ? 1 : 8};
    sp.block_count = 1;
    for (int a = 0; a < kMaxDims; ++a) {
        sp.block[a] = a < dim ? tile[a] : 1;
        sp.nblocks[a] = a < dim ? (fine[a] + tile[a] - 1) / tile[a] : 1;
        sp.block_count *= sp.nblocks[a];
    }
    Report rep;
    bool ok = dim == 2 && edge2 == 32
                  ? replay_width<2, 32, 32, 1>(width, g, sp, offsets, M, light_limit, min_part,
                                               budget_bytes, vectors, rep)
              : dim == 2 ? replay_width<2, 16, 16, 1>(width, g, sp, offsets, M, light_limit,
                                                      min_part, budget_bytes, vectors, rep)
              : dim == 3 ? replay_width<3, 16, 16, 8>(width, g, sp, offsets, M, light_limit,
                                                      min_part, budget_bytes, vectors, rep)
                         : false;
    if (!ok) return -1;
    const int64_t v[14] = {rep.violations,  rep.first_code,    rep.first_item, rep.items,
                           rep.heavy_blocks, rep.heavy_parts,  rep.part_len,   rep.max_slot,
                           rep.max_partial, rep.partial_floats, rep.records,    rep.max_total,
                           rep.max_offset_index, g.nbins};
    for (int i = 0; i < 14; ++i) report[i] = v[i];
    return 0;
}

int tile_replay(int dim, int width, const int64_t *fine, const int64_t *side,
                const uint32_t *offsets, int64_t M, uint32_t light_limit, uint32_t min_part,
                int64_t budget_bytes, int64_t vectors, int64_t *report)
{
    return tile_replay_edge(dim, 32, width, fine, side, offsets, M, light_limit, min_part,
                            budget_bytes, vectors, report);
}

}  // extern "C"
