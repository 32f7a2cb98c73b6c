"""The partial-sum scratch of a batched type-1 spread is sized on the host,
before any kernel runs, from a bound per problem: its blocks plus twice
(reach x M) / smallest part, where reach bounds how many blocks count one
record (wgpunufft_amd/csrc/plan.hip, execute_type1_many_t).  This CPU check
restates the device's block reach (tile_index.hpp axis_reach), counts every
block's records in reach for clustered and dense point sets from the
oracle's bin keys, replays the work plan's part count (spread.hip
plan_blocks: light-limit escalation, part length, parts per heavy block)
and asserts the bound holds, so the batched spread never writes beyond its
scratch.  A planted undercount of the reach must be caught.
"""
import numpy as np
import pytest

from oracle import oracle as O
from wgpunufft_amd import synthetic as S


def _axis_reach(b0, B, w, n, side, bins):
    """Bins of one axis whose starts reach the cells [b0, b0 + B)."""
    L = min(B + w - 1, n)
    if L + side > n:
        return np.arange(bins)
    r0 = b0 - (w - 1)
    if r0 < 0:
        r0 += n
    sh = side.bit_length() - 1
    if r0 + L <= n:
        return np.arange(r0 >> sh, ((r0 + L - 1) >> sh) + 1)
    first = np.arange(r0 >> sh, bins)
    return np.concatenate([first, np.arange(0, ((r0 + L - n - 1) >> sh) + 1)])


def _blocks(d, fine, w):
    """Output-stationary blocks of an F32 plan (plan.hip make_spread_params)."""
    run = {1: 4, 2: 2, 3: 8}[d]
    while run > 2 and run > fine[0] + 1 - w:
        run //= 2
    block = [1] * d
    if d == 3:
        block[1] = block[2] = 8
    elif d == 2:
        block[1] = 16
    lanes = int(np.prod(block[1:])) if d > 1 else 1
    block[0] = run * (64 // lanes)
    return block


def _reach_factor(d, fine, w, block, side):
    f = 1
    for a in range(d):
        B, n, s = block[a], fine[a], side[a]
        nb = -(-n // B)
        L = min(B + w - 1, n)
        f *= nb if L + s > n else min(nb, (s + w + 2 * B - 3) // B + 1)
    return f


def _totals(d, fine, w, block, side, counts):
    nblk = [-(-fine[a] // block[a]) for a in range(d)]
    bins = [-(-fine[a] // side[a]) for a in range(d)]
    reach = [[_axis_reach(i * block[a], block[a], w, fine[a], side[a], bins[a])
              for i in range(nblk[a])] for a in range(d)]
    out = []
    for idx in np.ndindex(*nblk[::-1]):
        idx = idx[::-1]  # axis 0 fastest
        sel = np.ix_(*[reach[a][idx[a]] for a in range(d)][::-1])
        out.append(int(counts[sel].sum()))
    return np.array(out, dtype=np.int64)


def _parts(totals, light_limit=4096, min_part=1024, capacity=None):
    """plan_blocks: heavy parts of the blocks' totals."""
    limit = light_limit
    while True:
        heavy = totals[totals > limit]
        if 2 * heavy.size <= capacity or limit >= 0x7FFFFFFF:
            break
        limit *= 2
    spare = capacity - heavy.size if capacity > heavy.size else 1
    part = max(-(-int(heavy.sum()) // spare), min_part)
    np_ = -(-heavy // part)
    return int(np_[np_ > 1].sum())


CASES = [
    # (modes, M, eps, kind)
    ([128, 128], 200_000, 1e-6, "clustered"),
    ([72, 40], 90_000, 1e-3, "clustered"),
    ([64, 64], 2_000_000, 1e-5, "uniform"),
    ([1000], 100_000, 1e-5, "clustered"),
    ([24, 20, 16], 30_000, 1e-6, "clustered"),
    ([16, 16, 16], 400_000, 1e-6, "uniform"),
]


@pytest.mark.parametrize("modes,M,eps,kind", CASES)
def test_batched_partial_slots_bound(modes, M, eps, kind):
    d = len(modes)
    w, _ = O.es_params(eps)
    fine = [O.fine_size(n, 2.0, w) for n in modes]
    if kind == "clustered":
        x = S.clustered_coords(S.config_seed(8, M), M, [0.3, -1.1, 2.0][:d], 0.05)
    else:
        x = S.uniform_coords(S.config_seed(8, M), M, d)
    lay = O.layout(fine, w, 0, M)
    side = lay["geometries"][lay["chosen"]]
    keys, _, _, _ = O.bin_keys_f32(fine, w, side, x)
    bins = [-(-fine[a] // side[a]) for a in range(d)]
    counts = np.bincount(keys, minlength=int(np.prod(bins))).reshape(bins[::-1])
    block = _blocks(d, fine, w)
    totals = _totals(d, fine, w, block, side, counts)
    reach = _reach_factor(d, fine, w, block, side)
    assert totals.sum() <= reach * M, (totals.sum(), reach, M)
    per_part = 64 * {1: 4, 2: 2, 3: 8}[d] * 2 * 4
    capacity = max((256 << 20) // per_part, 64)
    parts = _parts(totals, capacity=capacity)
    bound = min(capacity, int(totals.size + 2.0 * reach * M / 1024))
    print(f"BOUND d={d} {modes} M={M} {kind}: sum totals {totals.sum()} <= {reach} x M, "
          f"parts {parts} <= slots {bound}")
    assert parts <= bound
    if kind == "clustered":
        assert parts > 0  # the heavy path is exercised


def test_planted_reach_undercount_is_caught():
    """A reach that counts every record in one block only (the mistake of
    ignoring the halo of w - 1 cells) is exceeded by the clustered 2D set's
    records in reach: the check above is not vacuous."""
    modes, M, eps = [128, 128], 200_000, 1e-6
    w, _ = O.es_params(eps)
    fine = [O.fine_size(n, 2.0, w) for n in modes]
    x = S.clustered_coords(S.config_seed(8, M), M, [0.3, -1.1], 0.05)
    lay = O.layout(fine, w, 0, M)
    side = lay["geometries"][lay["chosen"]]
    keys, _, _, _ = O.bin_keys_f32(fine, w, side, x)
    bins = [-(-fine[a] // side[a]) for a in range(2)]
    counts = np.bincount(keys, minlength=int(np.prod(bins))).reshape(bins[::-1])
    block = _blocks(2, fine, w)
    totals = _totals(2, fine, w, block, side, counts)
    assert totals.sum() <= _reach_factor(2, fine, w, block, side) * M
    assert totals.sum() > 1 * M
