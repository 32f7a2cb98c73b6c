"""CPU replay of the matrix-core tile spreads' addressing at the benchmark
sizes.

The 2D matrix-core spread of the 1024^2-mode, 5e7-point workload once ended
in a GPU memory aperture violation.  This replay runs the engine's own index
code (wgpunufft_amd/csrc/tile_index.hpp) on the host over the bin offsets of
that workload (the oracle's stable counting sort of the row-geometry bin
keys): tile totals, the work plan (light tiles, heavy parts), every record
slot of every work item, every partial-sum index of the tile kernel and of
the heavy-part reduction, and every 32-bit intermediate, each checked
against the buffer it addresses.  The same replay covers the 3D tiles of
the 1e7-point half-clustered workload.  A planted fault in the oracle's
offsets must be caught.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from wgpunufft_amd import synthetic as S

HERE = os.path.dirname(os.path.abspath(__file__))
REPLAY = os.path.join(HERE, "replay")
LIB = os.path.join(REPLAY, "_build", "libtile_replay.so")
FIELDS = ["violations", "code", "item", "items", "heavy_tiles", "heavy_parts", "part_len",
          "max_slot", "max_partial", "partial_floats", "records", "max_total",
          "max_offset_index", "nbins"]


@pytest.fixture(scope="module")
def replay_lib():
    subprocess.run(["make", "-s", "-C", REPLAY], check=True)
    L = ctypes.CDLL(LIB)
    L.tile_replay.restype = ctypes.c_int
    L.tile_replay_edge.restype = ctypes.c_int
    L.tile_replay_edge.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                   ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int64,
                                   ctypes.c_int64, ctypes.c_void_p]
    return L


def _run(L, dim, width, fine, side, offsets, M, vectors=1, edge2=32):
    """Light limit 4096 records, parts of at least 1024, 256 MiB of partial
    sums: the spreader's defaults (spread.hpp)."""
    f = np.asarray(fine, dtype=np.int64)
    s = np.asarray(side, dtype=np.int64)
    off = np.ascontiguousarray(offsets, dtype=np.uint32)
    rep = np.zeros(16, dtype=np.int64)
    rc = L.tile_replay_edge(dim, edge2, width, f.ctypes.data, s.ctypes.data, off.ctypes.data, M,
                            4096, 1024, 256 << 20, vectors, rep.ctypes.data)
    assert rc == 0
    return dict(zip(FIELDS, rep.tolist()))


def _row_side(fine):
    """Bins of the row tiles (plan.hip spread_geometry): one cell along the
    other axes, along axis 0 the smallest power of two from 2 keeping the
    bin count within 2^24."""
    side0 = 2
    lines = int(np.prod(fine[1:]))
    while side0 < fine[0] and ((fine[0] + side0 - 1) // side0) * lines > (1 << 24):
        side0 *= 2
    return [side0] + [1] * (len(fine) - 1)


def _offsets(fine, width, side, x):
    keys, _, _, amb = O.bin_keys_f32(fine, width, side, x)
    # (the addressing replay needs the oracle's own keys only; points in the
    # double-float window of a start boundary are a few per 10^8)
    assert amb <= max(1, x.size // 10_000_000)
    nbins = int(np.prod([(n + s - 1) // s for n, s in zip(fine, side)]))
    return O.bin_sort(keys, nbins)[1]


def test_replay_2d_tiles_at_5e7_points(replay_lib):
    """1024^2 modes (fine 2048^2, w 7), 5e7 uniform points: every index in
    bounds; every tile is split into parts (the heavy path)."""
    M, fine, w = 50_000_000, [2048, 2048], 7
    x = S.uniform_coords(S.config_seed(3, 1), M, 2)
    side = _row_side(fine)
    off = _offsets(fine, w, side, x)
    del x
    r = _run(replay_lib, 2, w, fine, side, off, M)
    print("REPLAY 2D", r)
    assert r["violations"] == 0, r
    assert r["heavy_tiles"] == 64 * 64 and r["heavy_parts"] <= 32768
    assert r["max_partial"] < r["partial_floats"] and r["max_slot"] < M
    # tol 1e-3 (w 4) on the same points
    r4 = _run(replay_lib, 2, 4, fine, side, _offsets(fine, 4, side,
                                                     S.uniform_coords(S.config_seed(3, 1), M, 2)),
              M)
    assert r4["violations"] == 0, r4


def test_replay_2d_plain_row_tiles_at_5e7_points(replay_lib):
    """The plain 2D row tiles (16 x 16 cells, k_spread_rows2 and
    k_reduce_tiles<T,16,16,1>) on the same 5e7 uniform points and on a
    half-clustered set: every index in bounds."""
    M, fine, w = 50_000_000, [2048, 2048], 7
    side = _row_side(fine)
    x = S.uniform_coords(S.config_seed(3, 1), M, 2)
    r = _run(replay_lib, 2, w, fine, side, _offsets(fine, w, side, x), M, edge2=16)
    print("REPLAY 2D rows", r)
    assert r["violations"] == 0, r
    assert r["heavy_tiles"] == 128 * 128
    assert r["max_partial"] < r["partial_floats"] and r["max_slot"] < M
    x[: M // 2] = S.clustered_coords(S.config_seed(3, 2), M // 2, (0.3, -1.1), 0.05)
    r = _run(replay_lib, 2, w, fine, side, _offsets(fine, w, side, x), M, edge2=16)
    print("REPLAY 2D rows clustered", r)
    assert r["violations"] == 0, r


def test_replay_3d_tiles_at_1e7_clustered_points(replay_lib):
    x = S.config2_points()
    M, fine, w = x.shape[0], [256, 256, 256], 7
    side = _row_side(fine)
    off = _offsets(fine, w, side, x)
    r = _run(replay_lib, 3, w, fine, side, off, M, vectors=1)
    print("REPLAY 3D", r)
    assert r["violations"] == 0, r
    assert r["heavy_tiles"] > 0


def test_replay_catches_a_planted_offset_fault(replay_lib):
    """The replay is a real check: offsets with one bin's end pushed past M
    (a sort that overran) are reported."""
    M, fine, w = 200_000, [256, 256], 7
    x = S.uniform_coords(S.config_seed(3, 11), M, 2)
    side = _row_side(fine)
    off = _offsets(fine, w, side, x).copy()
    ok = _run(replay_lib, 2, w, fine, side, off, M)
    assert ok["violations"] == 0
    off[len(off) // 2:] += 5000
    bad = _run(replay_lib, 2, w, fine, side, off, M)
    assert bad["violations"] > 0 and bad["code"] in (2, 3)
