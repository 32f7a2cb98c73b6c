"""GPU parity at the benchmark workloads' own sizes.

Every benchmark configuration runs here at its stated size (config 5's
targets aside, whose full interval has its own test) and is checked against
the CPU oracle: the full direct sums where they finish in seconds (2D,
1e5 points), seeded samples of modes or points elsewhere, the reference bin
permutation bit for bit, and bitwise repeats and point shuffles as the
workloads ask (tests/gpu_nufft_3d.rs:71-155, 264-274, 616-663 of the
reference run the same kinds of checks at smaller sizes).

Tolerances: relative l2 against FP64 direct sums <= 20 eps in 2D and
<= 32 eps in 3D (tests/gpu_nufft_2d.rs:15, tests/gpu_nufft_3d.rs:13).
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _setup():
    import torch
    import wgpunufft_amd as W
    from oracle import oracle as O
    from wgpunufft_amd import synthetic as S
    return torch, W, O, S


def _say(msg):
    print(msg, flush=True)


def _bits(t):
    import torch
    return t.view(torch.float32) if t.dtype == torch.complex64 else t.view(torch.float64)


def _sample(S, seed, n, total):
    return np.sort(S.splitmix64(seed, n) % np.uint64(total)).astype(np.int64)


# ---------------------------------------------------------------------------
# Config 1: 2D type-1, 128 x 128 modes, M = 1e5 uniform, tol 1e-6.
# ---------------------------------------------------------------------------
def test_config1_full_oracle_and_bitwise_repeat():
    torch, W, O, S = _setup()
    _say("CASE config1: 2D type-1 128^2, M=1e5, tol 1e-6, full direct sums")
    M, n_modes, eps = 100_000, [128, 128], 1e-6
    x = S.uniform_coords(S.config_seed(1, 1), M, 2)
    c = S.complex_normal(S.config_seed(1, 2), M)
    xt, ct = torch.from_numpy(x).cuda(), torch.from_numpy(c).cuda()
    plan = W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, eps))
    outs = [plan.encode_gpu(xt, ct).clone() for _ in range(5)]
    torch.cuda.synchronize()
    for o in outs[1:]:
        assert torch.equal(_bits(o), _bits(outs[0]))
    ref = O.direct_type1(x.astype(np.float64), c.astype(np.complex128), n_modes)
    err = rel_l2(outs[0].cpu().numpy(), ref)
    _say(f"config1 type-1 rel l2 {err:.3e} (all 16384 modes)")
    assert err <= 20 * eps
    # and the type-2 direction on the same points, all points
    fm = S.complex_normal(S.config_seed(1, 3), 128 * 128)
    p2 = W.NufftPlan.type2_gpu(W.NufftConfig(n_modes, eps))
    v = p2.encode_gpu(xt, torch.from_numpy(fm).cuda())
    torch.cuda.synchronize()
    ref2 = O.direct_type2(x.astype(np.float64), fm.astype(np.complex128), n_modes)
    err2 = rel_l2(v.cpu().numpy(), ref2)
    _say(f"config1 type-2 rel l2 {err2:.3e} (all 1e5 points)")
    assert err2 <= 20 * eps


# ---------------------------------------------------------------------------
# Config 2: 3D type-1 / type-2, 128^3 modes, M = 1e7 (half clustered).
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def config2():
    torch, W, O, S = _setup()
    _say("CASE config2 data: 1e7 points (half uniform, half clustered), 128^3 modes")
    x = S.config2_points()
    M = x.shape[0]
    c = S.complex_normal(S.config_seed(2, 3), M)
    fm = S.complex_normal(S.config_seed(2, 4), 128 ** 3)
    d = {"x": x, "c": c, "fm": fm, "M": M,
         "xt": torch.from_numpy(x).cuda(), "ct": torch.from_numpy(c).cuda(),
         "fmt": torch.from_numpy(fm).cuda()}
    yield d
    d.clear()
    torch.cuda.empty_cache()


def test_config2_bin_sort_is_the_reference_permutation(config2):
    """mnufft_bin_sort and the permutation the executes of the benchmarked
    plan use (index order, plain and matrix-core spread, dense bins off)
    equal the oracle's stable counting sort of all 1e7 points bit for bit."""
    torch, W, O, S = _setup()
    _say("CASE config2 bin sort: 1e7 points vs oracle counting sort")
    x, M = config2["x"], config2["M"]
    plan = W.NufftPlan.type1_gpu(W.NufftConfig([128] * 3, 1e-6))
    info = plan.info()
    fine, w = info["fine"], info["width"]
    lay = O.layout(fine, w, 0, M)
    side = lay["geometries"][lay["chosen"]]
    assert side == [2, 1, 1]
    keys, _, _, amb = O.bin_keys_f32(fine, w, side, x)
    # the engine's f64 fold equals the oracle's for every point (checked
    # below); the few points within the double-float window of a start
    # boundary (the reference folds in df64) are reported: parity with the
    # reference's own rounding is unpinned for them
    _say(f"config2 bin keys: {amb} of {x.shape[0]} points in the df64 window")
    assert amb <= max(1, x.size // 10_000_000)
    nbins = int(np.prod([(f + s - 1) // s for f, s in zip(fine, side)]))
    ref_sorted, ref_off = O.bin_sort(keys, nbins)
    srt, off = plan.bin_sort(config2["xt"])
    torch.cuda.synchronize()
    assert np.array_equal(off.cpu().numpy().astype(np.uint32), ref_off)
    assert np.array_equal(srt.cpu().numpy().astype(np.uint32), ref_sorted)
    _say(f"config2 bin sort bit-exact: {nbins} bins, largest {int(np.diff(ref_off).max())}")
    for method in ("plain", "mfma_f16x3"):
        plan.set_spread_method(method).set_dense_bins(0).set_point_order("index")
        plan.set_points_gpu(config2["xt"])
        psrt, poff, pside = plan.points_permutation()
        torch.cuda.synchronize()
        assert pside == side, (method, pside)
        assert np.array_equal(poff.cpu().numpy().astype(np.uint32), ref_off), method
        assert np.array_equal(psrt.cpu().numpy().astype(np.uint32), ref_sorted), method
        _say(f"config2 {method} encode permutation bit-exact")


@pytest.mark.parametrize("method", ["plain", "mfma_f16x3"])
def test_config2_type1_sampled_oracle_repeat_and_shuffles(config2, method):
    """Type 1: error on 200 seeded modes, 5 bitwise repeats (index order),
    and in coordinate order the same bits under 3 seeded shuffles of the
    points (with dense bins for the matrix-core spread, as benchmarked)."""
    torch, W, O, S = _setup()
    _say(f"CASE config2 type-1 {method}: repeats, shuffles, sampled oracle")
    eps, n_modes, M = 1e-6, [128] * 3, config2["M"]
    xt, ct = config2["xt"], config2["ct"]
    plan = W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, eps)).set_spread_method(method)
    outs = [plan.encode_gpu(xt, ct).clone() for _ in range(5)]
    torch.cuda.synchronize()
    for o in outs[1:]:
        assert torch.equal(_bits(o), _bits(outs[0]))
    sel = _sample(S, S.config_seed(2, 7), 200, 128 ** 3)
    ref = O.direct_type1(config2["x"].astype(np.float64), config2["c"].astype(np.complex128),
                         n_modes, modes=sel)
    err = rel_l2(outs[0].cpu().numpy()[sel], ref)
    _say(f"config2 type-1 {method} rel l2 {err:.3e} (200 modes)")
    assert err <= 32 * eps
    plan.set_point_order("coordinates")
    base = plan.encode_gpu(xt, ct).clone()
    for k in range(3):
        perm = torch.from_numpy(S.shuffle_perm(S.config_seed(2, 20 + k), M)).cuda()
        o = plan.encode_gpu(xt[perm].contiguous(), ct[perm].contiguous())
        torch.cuda.synchronize()
        assert torch.equal(_bits(o), _bits(base)), f"shuffle {k}"
    err_c = rel_l2(base.cpu().numpy()[sel], ref)
    _say(f"config2 type-1 {method} coordinate order: 3 shuffles bitwise, rel l2 {err_c:.3e}")
    assert err_c <= 32 * eps


def test_config2_type2_sampled_oracle_repeat_and_shuffles(config2):
    torch, W, O, S = _setup()
    _say("CASE config2 type-2: repeats, shuffles, sampled oracle")
    eps, n_modes, M = 1e-6, [128] * 3, config2["M"]
    xt, fmt = config2["xt"], config2["fmt"]
    plan = W.NufftPlan.type2_gpu(W.NufftConfig(n_modes, eps))
    outs = [plan.encode_gpu(xt, fmt).clone() for _ in range(5)]
    torch.cuda.synchronize()
    for o in outs[1:]:
        assert torch.equal(_bits(o), _bits(outs[0]))
    sel = _sample(S, S.config_seed(2, 8), 2000, M)
    ref = O.direct_type2(config2["x"].astype(np.float64), config2["fm"].astype(np.complex128),
                         n_modes, points=sel)
    err = rel_l2(outs[0].cpu().numpy()[sel], ref)
    _say(f"config2 type-2 rel l2 {err:.3e} (2000 points)")
    assert err <= 32 * eps
    for k in range(3):
        perm = torch.from_numpy(S.shuffle_perm(S.config_seed(2, 30 + k), M)).cuda()
        o = plan.encode_gpu(xt[perm].contiguous(), fmt)
        torch.cuda.synchronize()
        assert torch.equal(_bits(o), _bits(outs[0][perm])), f"shuffle {k}"
    _say("config2 type-2: 3 shuffles bitwise (un-permuted)")


# ---------------------------------------------------------------------------
# Config 3: 2D type-1, 1024^2 modes, M = 5e7 uniform (plain FP32 spreads).
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def config3():
    torch, W, O, S = _setup()
    _say("CASE config3 data: 5e7 uniform points in 2D")
    M = 50_000_000
    x = S.uniform_coords(S.config_seed(3, 1), M, 2)
    c = S.complex_normal(S.config_seed(3, 2), M)
    sel = _sample(S, S.config_seed(3, 7), 40, 1024 * 1024)
    d = {"xt": torch.from_numpy(x).cuda(), "ct": torch.from_numpy(c).cuda(), "sel": sel,
         "ref": O.direct_type1(x.astype(np.float64), c.astype(np.complex128), [1024, 1024],
                               modes=sel)}
    yield d
    d.clear()
    torch.cuda.empty_cache()


@pytest.mark.parametrize("eps", [1e-6, 1e-3])
def test_config3_plain_sampled_oracle(config3, eps):
    torch, W, O, S = _setup()
    _say(f"CASE config3 plain tol {eps:g}: 2D 1024^2, M=5e7")
    xt, ct = config3["xt"], config3["ct"]
    plan = W.NufftPlan.type1_gpu(W.NufftConfig([1024, 1024], eps))
    a = plan.encode_gpu(xt, ct).clone()
    b = plan.encode_gpu(xt, ct)
    torch.cuda.synchronize()
    assert torch.equal(_bits(a), _bits(b))
    err = rel_l2(a.cpu().numpy()[config3["sel"]], config3["ref"])
    _say(f"config3 plain tol {eps:g} rel l2 {err:.3e} (40 modes)")
    assert err <= 20 * eps


# ---------------------------------------------------------------------------
# Config 4a: 64 independent 2D type-2 problems as one batch.
# ---------------------------------------------------------------------------
def test_config4a_batch_matches_oracle_and_per_plan_encodes():
    torch, W, O, S = _setup()
    _say("CASE config4a: 64 2D type-2 problems, one batch")
    plans, xs, fs, hx, hf, dims = [], [], [], [], [], []
    for n0, n1, M, sx, sf in S.batch_problem_specs():
        x = S.uniform_coords(sx, M, 2)
        f = S.complex_normal(sf, n0 * n1)
        plans.append(W.NufftPlan.type2_gpu(W.NufftConfig([n0, n1], 1e-5)))
        xs.append(torch.from_numpy(x).cuda())
        fs.append(torch.from_numpy(f).cuda())
        hx.append(x)
        hf.append(f)
        dims.append([n0, n1])
    outs = W.encode_type2_many(plans, xs, fs)
    torch.cuda.synchronize()
    worst = 0.0
    for i, (p, x, f, o) in enumerate(zip(plans, xs, fs, outs)):
        own = p.encode_gpu(x, f)
        torch.cuda.synchronize()
        assert torch.equal(_bits(own), _bits(o)), f"problem {i}"
        M = x.shape[0]
        sel = _sample(S, S.config_seed(4, 300 + i), 64, M)
        ref = O.direct_type2(hx[i].astype(np.float64), hf[i].astype(np.complex128), dims[i],
                             points=sel)
        err = rel_l2(o.cpu().numpy()[sel], ref)
        worst = max(worst, err)
        assert err <= 20 * 1e-5, f"problem {i}: {err}"
    _say(f"config4a: 64 problems equal their own encodes bitwise; worst rel l2 {worst:.3e}")


def test_config4a_points_as_type1_batch_matches_oracle_and_per_plan_encodes():
    """The same 64 point sets as type-1 problems (strengths -> modes) in one
    batch: each output equals its plan's own encode bit for bit and the
    FP64 direct sums on sampled modes (2D tolerance 20 eps)."""
    torch, W, O, S = _setup()
    _say("CASE config4a points as 64 2D type-1 problems, one batch")
    plans, xs, cs, hx, hc, dims = [], [], [], [], [], []
    for i, (n0, n1, M, sx, _) in enumerate(S.batch_problem_specs()):
        x = S.uniform_coords(sx, M, 2)
        c = S.complex_normal(S.config_seed(4, 300 + i), M)
        plans.append(W.NufftPlan.type1_gpu(W.NufftConfig([n0, n1], 1e-5)))
        xs.append(torch.from_numpy(x).cuda())
        cs.append(torch.from_numpy(c).cuda())
        hx.append(x)
        hc.append(c)
        dims.append([n0, n1])
    outs = W.encode_type1_many(plans, xs, cs)
    torch.cuda.synchronize()
    worst = 0.0
    for i, (p, x, c, o) in enumerate(zip(plans, xs, cs, outs)):
        own = p.encode_gpu(x, c)
        torch.cuda.synchronize()
        assert torch.equal(_bits(own), _bits(o)), f"problem {i}"
        if i % 8 == 0:
            nm = dims[i][0] * dims[i][1]
            sel = _sample(S, S.config_seed(4, 400 + i), 48, nm)
            ref = O.direct_type1(hx[i].astype(np.float64), hc[i].astype(np.complex128), dims[i],
                                 modes=sel)
            err = rel_l2(o.cpu().numpy()[sel], ref)
            worst = max(worst, err)
            assert err <= 20 * 1e-5, f"problem {i}: {err}"
    _say(f"config4a type 1: 64 problems equal their own encodes bitwise; worst rel l2 {worst:.3e}")


# ---------------------------------------------------------------------------
# Config 4b: 32 vectors over 8192 radial spokes x 512 samples, 256^3 modes.
# ---------------------------------------------------------------------------
def test_config4b_radial_batch_sampled_oracle():
    torch, W, O, S = _setup()
    _say("CASE config4b: 3D type-1 256^3, 32 vectors, radial spokes")
    x = S.radial_spokes()
    M, B, n_modes, eps = x.shape[0], 32, [256] * 3, 1e-6
    N = 256 ** 3
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0x4E554654_00000004)
    ct = torch.randn(M * B, dtype=torch.complex64, device="cuda", generator=gen)
    xt = torch.from_numpy(x).cuda()
    plan = W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, eps).with_batch(B))
    plan.set_spread_method("mfma_f16x3")
    out = plan.encode_gpu(xt, ct)
    again = plan.execute_gpu(ct)
    torch.cuda.synchronize()
    assert torch.equal(_bits(out), _bits(again))
    sel = _sample(S, S.config_seed(4, 9), 60, N)
    xd = x.astype(np.float64)
    for v in (0, B - 1):
        cv = ct[v * M:(v + 1) * M].cpu().numpy().astype(np.complex128)
        ref = O.direct_type1(xd, cv, n_modes, modes=sel)
        err = rel_l2(out[v * N:(v + 1) * N].cpu().numpy()[sel], ref)
        _say(f"config4b vector {v} rel l2 {err:.3e} (60 modes)")
        assert err <= 32 * eps
    del out, again, plan
    torch.cuda.empty_cache()


# ---------------------------------------------------------------------------
# Config 5: 3D type-3, 1e6 sources in [-5,5]^3, 1e6 targets in [-200,200]^3.
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("prec,eps", [("f64", 1e-12), ("f64", 1e-10), ("f32", 1e-6)])
def test_config5_full_interval_sampled_oracle(prec, eps):
    """The target interval's inner grid (2592^3 cells) does not fit one GPU:
    the plan cuts it into target boxes by itself; every target is computed
    and 1,000 seeded targets are checked against the FP64 direct sums."""
    torch, W, O, S = _setup()
    _say(f"CASE config5 {prec} tol {eps:g}: 1e6 x 1e6, targets in [-200,200]^3")
    M = N = 1_000_000
    f64 = prec == "f64"
    x = S.uniform_coords(S.config_seed(5, 1), M, 3, lo=-5.0, hi=5.0, dtype=np.float64)
    s = S.uniform_coords(S.config_seed(5, 2), N, 3, lo=-200.0, hi=200.0, dtype=np.float64)
    c = S.complex_normal(S.config_seed(5, 3), M, dtype=np.complex128)
    sel = _sample(S, S.config_seed(5, 4), 1000, N)
    if not f64:
        # the F32 plan sees the coordinates and strengths rounded to F32
        x = x.astype(np.float32).astype(np.float64)
        s = s.astype(np.float32).astype(np.float64)
        c = c.astype(np.complex64).astype(np.complex128)
    ref = O.direct_type3(x, c, s, targets=sel)
    dt, ct_ = (torch.float64, torch.complex128) if f64 else (torch.float32, torch.complex64)
    xt = torch.from_numpy(x).cuda().to(dt)
    st = torch.from_numpy(s).cuda().to(dt)
    ct = torch.from_numpy(c).cuda().to(ct_)
    plan = W.NufftType3Plan(3, [(-5, 5)] * 3, [(-200, 200)] * 3, eps=eps,
                            precision=W.F64 if f64 else W.F32)
    boxes = plan.boxes()
    assert int(np.prod(boxes)) > 1, boxes
    out = plan.encode_gpu(xt, ct, st)
    torch.cuda.synchronize()
    err = rel_l2(out.cpu().numpy()[sel], ref)
    _say(f"config5 {prec} tol {eps:g}: boxes {boxes}, outer {plan.info()['outer']}, "
         f"rel l2 {err:.3e} (1000 targets)")
    # F32: the phases s.x reach 3,000 rad, so FP32 target positions alone
    # limit the F32 transform to about 1e-5 here; it is reported beside the
    # FP64 legs and held to a sanity bound
    assert err <= (100 * eps if f64 else 1e-3)
    del out, plan
    torch.cuda.empty_cache()


# ---------------------------------------------------------------------------
# The GPU FP64 direct sums the benchmark's error figures use.
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n_modes,order", [([23], 0), ([24], 1), ([8, 7, 6], 0),
                                           ([8, 7, 6], 1), ([5, 6, 4], 0), ([5, 4, 3, 3], 1),
                                           ([3, 2, 3, 2, 2, 3], 0), ([2] * 8, 1)])
def test_gpu_direct_sums_match_oracle_1d_3d(n_modes, order):
    torch, W, O, S = _setup()
    d = len(n_modes)
    _say(f"CASE direct sums d={d} modes={n_modes} order={order}")
    M = 400
    x = S.uniform_coords(S.config_seed(3, 10 + d), M, d, dtype=np.float64)
    c = S.complex_normal(S.config_seed(3, 20 + d), M, dtype=np.complex128)
    nm = int(np.prod(n_modes))
    fm = S.complex_normal(S.config_seed(3, 30 + d), nm, dtype=np.complex128)
    xt = torch.from_numpy(x).cuda()
    for isign in (1, -1):
        g1 = W.reference_type1_f64(n_modes, xt, torch.from_numpy(c).cuda(), isign, order)
        r1 = O.direct_type1(x, c, n_modes, isign, order)
        assert rel_l2(g1.cpu().numpy(), r1) <= 1e-13
        g2 = W.reference_type2_f64(n_modes, xt, torch.from_numpy(fm).cuda(), isign, order)
        r2 = O.direct_type2(x, fm, n_modes, isign, order)
        assert rel_l2(g2.cpu().numpy(), r2) <= 1e-13
    sel = np.array([0, nm // 3, nm - 1], dtype=np.int64)
    g1s = W.reference_type1_f64(n_modes, xt, torch.from_numpy(c).cuda(), 1, order,
                                modes=torch.from_numpy(sel).cuda())
    assert rel_l2(g1s.cpu().numpy(), O.direct_type1(x, c, n_modes, 1, order)[sel]) <= 1e-13


def test_oversize_point_counts_are_plan_errors():
    """Counts beyond the 32-bit point index space are plan errors at every
    device-pointer entry point, before any work or allocation (the buffers
    passed are far smaller than the counts)."""
    import ctypes
    torch, W, O, S = _setup()
    _say("CASE oversize point counts")
    L = W.lib()
    big = 1 << 32
    x = torch.zeros(16, 3, device="cuda")
    c = torch.zeros(16, dtype=torch.complex64, device="cuda")
    plan = W.NufftPlan.type1_gpu(W.NufftConfig([8, 8, 8], 1e-6))
    out = torch.zeros(512, dtype=torch.complex64, device="cuda")
    vp = ctypes.c_void_p
    h = plan._handle
    assert L.mnufft_set_points(h, big, vp(x.data_ptr()), None) == 2
    assert L.mnufft_autotune(h, 1, big, vp(x.data_ptr()), vp(c.data_ptr()),
                             vp(out.data_ptr()), 1, None, None) == 2
    srt = torch.zeros(16, dtype=torch.int32, device="cuda")
    assert L.mnufft_bin_sort(h, big, vp(x.data_ptr()), vp(srt.data_ptr()), None, None,
                             None) == 2
    t3 = W.NufftType3Plan(3, [(-1, 1)] * 3, [(-2, 2)] * 3, eps=1e-6)
    assert L.mnufft_execute_type3(t3._handle, 1, big, vp(x.data_ptr()), vp(c.data_ptr()), 4,
                                  vp(x.data_ptr()), vp(out.data_ptr()), None) == 2
    assert L.mnufft_execute_type3(t3._handle, 1, 4, vp(x.data_ptr()), vp(c.data_ptr()), big,
                                  vp(x.data_ptr()), vp(out.data_ptr()), None) == 2
    assert b"32-bit" in L.mnufft_last_error()
