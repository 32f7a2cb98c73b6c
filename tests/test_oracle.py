"""CPU pinning of the oracle against the reference's own known answers and
against mathematics that fixes each function independently."""
import json
import math
import os

import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O

# known answers the reference's tests assert, kept as data
with open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")) as _f:
    GOLDEN = json.load(_f)


def test_es_parameter_table():
    # tests/cpu_foundations.rs:482-501
    table = GOLDEN["es_parameters"]
    assert len(table["rows"]) == 6
    for eps, w, beta in table["rows"]:
        ww, bb = O.es_params(eps)
        assert ww == w and abs(bb - beta) < table["beta_abs_tol"]
        assert O.es_eval(w, bb, 0.0) == 1.0
        assert O.es_eval(w, bb, w / 2) == 0.0 and O.es_eval(w, bb, -w / 2) == 0.0
        assert abs(O.es_eval(w, bb, 0.31) - O.es_eval(w, bb, -0.31)) < 1e-15


def test_custom_sigma_formulas():
    # tests/cpu_foundations.rs:546-560
    sigma, eps = 1.25, 1e-6
    w, b = O.es_params(eps, sigma)
    ew = math.ceil(math.log(1 / eps) / (math.pi * math.sqrt(1 - 1 / sigma)))
    assert w == ew
    assert abs(b - 0.97 * math.pi * ew * (1 - 1 / (2 * sigma))) < 1e-13
    assert O.fine_size(17, sigma, w) == 22


def test_fine_grid_known_lengths():
    # tests/cpu_foundations.rs:503-560, tests/gpu_nufft_3d.rs:472
    for row in GOLDEN["fine_grid_lengths"]["rows"]:
        w, _ = O.es_params(row["eps"], row["sigma"])
        assert O.fine_size(row["modes"], row["sigma"], w) == row["length"], row
    w, _ = O.es_params(1e-6)
    assert O.fine_size(17, 2.0, w) == 36
    assert O.fine_size(1, 2.0, w) == 14
    assert O.fine_size(128, 2.0, w) == 256
    assert O.fine_size(256, 2.0, w) == 512
    for modes in range(1, 128):
        n = O.fine_size(modes, 2.0, w)
        assert n % 2 == 0 and n >= 2 * modes and n >= 2 * w


def test_horner_table_accuracy_and_endpoints():
    # src/kernel.rs:413-458: |horner - exact| <= 0.1 eps over the support
    for eps in [1e-1, 1e-2, 1e-4, 1e-6, 1e-9, 1e-12]:
        w, b = O.es_params(eps)
        rows, coef = O.horner_table(w, b, eps)
        assert max(4, w - 4) <= rows <= min(w + 3, 19)
        hw = w / 2
        xs = np.linspace(-hw, hw, 4001)
        err = max(abs(O.horner_eval(w, rows, coef, d) - O.es_eval(w, b, d)) for d in xs)
        assert err <= 0.1 * eps, (eps, err)
        for d in [-hw, hw, float("inf"), float("-inf"), float("nan")]:
            assert O.horner_eval(w, rows, coef, d) == 0.0


@pytest.mark.parametrize("eps", [1e-6, 1e-12])
def test_phihat_against_midpoint_integration(eps):
    # src/kernel.rs:460-488: independent dense midpoint integration.  The
    # exact kernel's edge singularity (a square root) limits both rules to
    # about 1e-9, far below any plausible mistake (a sign, a factor 2, a
    # wrong frequency).
    w, b = O.es_params(eps)
    n = 128
    ph = O.phihat(w, b, n)
    samples = 200000
    dx = w / samples
    xs = -w / 2 + (np.arange(samples) + 0.5) * dx
    for mode in [0, 1, 7, 31, 64]:
        freq = 2 * np.pi * mode / n
        ker = np.exp(b * (np.sqrt(np.maximum(0.0, 1 - (2 * xs / w) ** 2)) - 1))
        ker[np.abs(xs) >= w / 2] = 0
        expected = np.sum(ker * np.cos(freq * xs)) * dx * (-1 if mode % 2 else 1)
        assert abs(ph[mode] - expected) <= 1e-8


def test_es_ft_matches_phihat_at_grid_frequencies():
    # the two quadratures of the same integral agree (type-3 correction vs
    # the grid coefficients without the (-1)^k phase) over the frequencies
    # type 3 evaluates, |s'| <= pi / sigma
    w, b = O.es_params(1e-6)
    n = 64
    ph = O.phihat(w, b, n)
    for k in range(0, 17):
        ft = O.es_ft(w, b, 2 * np.pi * k / n)
        assert abs(ft - ph[k] * (-1) ** k) <= 1e-8 * max(1.0, abs(ft))


def test_layout_reference_cases():
    # src/gpu_nd_spread.rs:2173-2275
    w, _ = O.es_params(1e-6)
    for row in GOLDEN["spread_layouts"]["rows"]:
        L = O.layout(row["fine"], w, row["precision"])
        assert (L["run"], L["workgroup_size"], L["block"]) == (
            row["run"], row["workgroup_size"], row["block"]), row
    L = O.layout([512, 512], w, 0)
    assert (L["run"], L["workgroup_size"], L["block"]) == (4, 32, [8, 16])
    assert O.layout([2000, 2000], w, 0)["geometries"][0] == [1, 1]
    assert O.layout([200, 200, 200], w, 0)["geometries"][0] == [2, 1, 1]
    assert O.layout([400, 400, 400], w, 0)["geometries"][0] == [2, 2, 2]
    sides = O.layout([200, 200, 200], w, 0)["geometries"]
    for s in ([2, 2, 2], [4, 4, 4], [16, 8, 8]):
        assert s in sides
    chosen = lambda M: (lambda L: L["geometries"][L["chosen"]])(O.layout([200] * 3, w, 0, M))
    dense, sparse, sparsest = chosen(10_000_000), chosen(100_000), chosen(1_000)
    assert all(s <= 2 for s in dense[1:])
    assert all(s >= 4 for s in sparse[1:])
    assert all(a >= b for a, b in zip(sparsest, sparse))
    # configuration geometries
    L = O.layout([256, 256, 256], w, 0, 10_000_000)
    assert L["geometries"][L["chosen"]] == [2, 1, 1] and L["block"] == [8, 8, 8]
    assert O.layout([256, 256], w, 0, 100_000)["geometries"][0] == [1, 1]


def test_layout_reference_cases_ranks_4_to_8():
    # src/gpu_nd_spread.rs:2122-2171, 2210-2220, 2244-2283
    G = GOLDEN["spread_layouts_nd"]
    for row in G["layouts"]:
        w, _ = O.es_params(row["eps"])
        L = O.layout(row["fine"], w, row["precision"])
        if "run" in row:
            assert (L["run"], L["workgroup_size"]) == (row["run"], row["workgroup_size"]), row
        if "block" in row:
            assert L["block"] == row["block"], row
        if "finest_side" in row:
            assert L["geometries"][0] == row["finest_side"], row
        if "max_side" in row:
            assert all(s <= row["max_side"] for g in L["geometries"] for s in g), row
    for row in G["finest_geometry_for_many_points"]:
        w, _ = O.es_params(row["eps"])
        assert O.layout(row["fine"], w, row["precision"], row["points"])["chosen"] == row["chosen"]
    for shape, eps in G["unique_support_starts"]["shapes"]:
        w, _ = O.es_params(eps)
        for prec in G["unique_support_starts"]["precisions"]:
            L = O.layout(shape, w, prec)
            assert all(b + w - 1 <= n for b, n in zip(L["block"], shape)), (shape, prec)
            assert L["workgroup_size"] <= 256


def test_ambiguity_counter_covers_the_double_float_window():
    """A coordinate whose folded position minus w/2 is exactly an integer
    (x = 0 on a 256-cell axis with an even width: 128 - 4 = 124) is counted
    as ambiguous although the f64 and exact folds agree: a double-float fold
    could land on either side of the boundary.  An odd width puts the same
    point half a cell from any boundary."""
    x = np.zeros((1, 1), dtype=np.float32)
    _, starts, _, amb = O.bin_keys_f32([256], 8, [1], x)
    assert starts[0, 0] == 124 and amb == 1
    _, _, _, amb = O.bin_keys_f32([256], 7, [1], x)
    assert amb == 0


def test_bin_keys_and_stable_sort_definition():
    rng = np.random.default_rng(7)
    x = rng.uniform(-np.pi, np.pi, size=(5000, 3)).astype(np.float32)
    fine, w, side = [36, 30, 28], 7, [2, 1, 1]
    keys, starts, offs, amb = O.bin_keys_f32(fine, w, side, x)
    assert amb == 0
    # independent float64 restatement of fold + start + key
    xd = x.astype(np.float64)
    n = np.array(fine, dtype=np.float64)
    pos = xd * (n / (2 * np.pi)) + np.floor(n / 2)
    pos = np.where(pos < 0, pos + n, pos)
    pos = np.where(pos >= n, pos - n, pos)
    st = np.ceil(pos - w / 2).astype(np.int64)
    assert np.array_equal(st, starts)
    ws = np.where(st < 0, st + np.array(fine), st)
    bins = [(f + s - 1) // s for f, s in zip(fine, side)]
    k = (ws[:, 0] >> 1) + bins[0] * (ws[:, 1] + bins[1] * ws[:, 2])
    assert np.array_equal(k.astype(np.uint32), keys)
    assert np.all(offs <= 0) and np.all(offs > -1.0 - w / 2 - 1e-9)
    nb = int(np.prod(bins))
    srt, off = O.bin_sort(keys, nb)
    assert np.array_equal(srt, np.argsort(keys, kind="stable").astype(np.uint32))
    assert off[-1] == len(x) and np.array_equal(np.diff(off), np.bincount(keys, minlength=nb))


def test_out_of_reach_points_land_in_first_bin():
    x = np.array([[np.nan, 0.0], [0.0, np.inf], [1e30, 0.0], [5 * np.pi, 0.1]], np.float32)
    keys, starts, offs, _ = O.bin_keys_f32([32, 32], 5, [1, 1], x)
    # 5 pi is within the fold's reach (4 pi)? no: 5 pi > 4 pi
    assert np.array_equal(keys, [0, 0, 0, 0])
    assert np.all(offs == 1024.0)


def _numpy_direct_type1(x, c, n_modes, isign, modeord):
    d = x.shape[1]
    grids = []
    for a, N in enumerate(n_modes):
        k = np.arange(N)
        grids.append(k - N // 2 if modeord == 0 else np.where(k < (N + 1) // 2, k, k - N))
    mesh = np.meshgrid(*grids, indexing="ij")
    ks = np.stack([m.ravel(order="F") for m in mesh], axis=1)  # axis 0 fastest
    return np.exp(1j * isign * ks @ x.T) @ c


def test_direct_sums_analytic_cases():
    # tests/cpu_foundations.rs:327-351
    c = np.array([1.5 - 0.25j, -0.5 + 2.0j])
    out = O.direct_type1(np.zeros((2, 1)), c, [5])
    assert np.allclose(out, 1.0 + 1.75j, atol=1e-14)
    f = np.zeros(5, complex)
    f[4] = 1.0  # centered k = +2
    for s in (1, -1):
        v = O.direct_type2(np.array([[0.37]]), f, [5], isign=s)
        assert abs(v[0] - np.exp(1j * s * 2 * 0.37)) < 1e-14


@pytest.mark.parametrize("modeord", [0, 1])
def test_direct_sums_against_numpy_closed_form(modeord):
    rng = np.random.default_rng(3)
    x = rng.uniform(-np.pi, np.pi, size=(50, 3))
    c = rng.normal(size=50) + 1j * rng.normal(size=50)
    n_modes = [5, 4, 3]
    ref = _numpy_direct_type1(x, c, n_modes, -1, modeord)
    got = O.direct_type1(x, c, n_modes, isign=-1, modeord=modeord)
    assert rel_l2(got, ref) < 1e-13
    # type 2 is the adjoint: <f, A c> = <A^H f, c>
    f = rng.normal(size=60) + 1j * rng.normal(size=60)
    t2 = O.direct_type2(x, f, n_modes, isign=1, modeord=modeord)
    lhs = np.vdot(f, got)
    rhs = np.vdot(t2, c)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
    s = rng.normal(size=(7, 3)) * 3
    t3 = O.direct_type3(x, c, s, isign=1)
    assert rel_l2(t3, np.exp(1j * s @ x.T) @ c) < 1e-13


@pytest.mark.parametrize("n_modes", [[3, 2, 2, 3], [2, 3, 2, 2, 2], [2] * 8])
def test_direct_sums_ranks_4_to_8(n_modes):
    """Ranks 4-8 (src/direct.rs:498-607; axis 0 fastest as
    tests/cpu_foundations.rs:402-413 states for 2D): the closed form per
    mode, and type 2 as the adjoint of type 1."""
    rng = np.random.default_rng(len(n_modes))
    d = len(n_modes)
    x = rng.uniform(-3 * np.pi, 3 * np.pi, size=(23, d))
    c = rng.normal(size=23) + 1j * rng.normal(size=23)
    for modeord in (0, 1):
        ref = _numpy_direct_type1(x, c, n_modes, 1, modeord)
        got = O.direct_type1(x, c, n_modes, isign=1, modeord=modeord)
        assert rel_l2(got, ref) < 1e-13
        f = rng.normal(size=ref.size) + 1j * rng.normal(size=ref.size)
        # type 2 of the other sign is the adjoint: <f, A c> = <A^H f, c>
        t2 = O.direct_type2(x, f, n_modes, isign=-1, modeord=modeord)
        assert abs(np.vdot(f, got) - np.vdot(t2, c)) < 1e-10 * abs(np.vdot(f, got))
        # and the sign flips under conjugation
        assert rel_l2(np.conj(t2), O.direct_type2(x, np.conj(f), n_modes, isign=1,
                                                  modeord=modeord)) < 1e-13


def test_type3_axis_metadata():
    # src/type3.rs:708-797; config 5 gives 1296 per axis
    m = O.type3_axis(-5, 5, -200, 200, 2.0, 13)
    assert m["nf"] == 1296 and m["C"] == 0.0 and m["D"] == 0.0
    n = m["nf"]
    while n % 2 == 0:
        n //= 2
    while n % 3 == 0:
        n //= 3
    while n % 5 == 0:
        n //= 5
    assert n == 1
    # centre snap is strict at 10 %
    m = O.type3_axis(1.0, 3.0, 0, 1, 2.0, 7)
    assert m["C"] == 2.0
    assert O.next_even_235(1287) == 1296


def test_cpu_port_matches_direct_sums():
    """The CPU baseline (port of CpuNufftPlan) meets the reference's CPU
    tolerance against the direct sums, in 2D and 3D, both types."""
    from oracle import cpu_nufft as C
    from wgpunufft_amd import synthetic as S
    for nm in ([16, 12, 10], [30, 20]):
        d = len(nm)
        x = S.uniform_coords(S.config_seed(1, 900 + d), 2000, d)
        c = S.complex_normal(S.config_seed(1, 901), 2000)
        fm = S.complex_normal(S.config_seed(1, 902), int(np.prod(nm)))
        for isign in (1, -1):
            f = C.type1(x, c, nm, 1e-6, isign)
            ref = O.direct_type1(x.astype(np.float64), c.astype(np.complex128), nm, isign)
            assert rel_l2(f, ref) <= 10e-6
            v = C.type2(x, fm, nm, 1e-6, isign)
            ref2 = O.direct_type2(x.astype(np.float64), fm.astype(np.complex128), nm, isign)
            assert rel_l2(v, ref2) <= 10e-6
    # thread-count independence, bitwise (tests/cpu_nufft.rs:314)
    x = S.uniform_coords(S.config_seed(1, 903), 5000, 3)
    c = S.complex_normal(S.config_seed(1, 904), 5000)
    a = C.type1(x, c, [20, 20, 20], 1e-6, 1, 1)
    b = C.type1(x, c, [20, 20, 20], 1e-6, 1, 4)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
