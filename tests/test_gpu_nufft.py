"""GPU parity of the type-1 / type-2 paths against the CPU oracle.

Tolerances follow the reference's GPU suites: relative l2 error against FP64
direct sums <= 20*eps in 2D and <= 32*eps in 3D (tests/gpu_nufft_2d.rs:15,
tests/gpu_nufft_3d.rs:13), 1D <= 8*eps per class (tests/gpu_nufft.rs:14-18),
F64 <= 20*eps at 1e-8 / 40*eps at 1e-12 (tests/gpu_nufft_precision.rs:16-20).
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL_FACTOR = {1: 8.0, 2: 20.0, 3: 32.0}


def _setup():
    import torch
    import wgpunufft_amd as W
    from oracle import oracle as O
    from wgpunufft_amd import synthetic as S
    return torch, W, O, S


def _points(S, kind, M, d, seed):
    if kind == "uniform":
        return S.uniform_coords(seed, M, d)
    if kind == "clustered":
        center = [0.3, -1.1, 2.0][:d]
        return S.clustered_coords(seed, M, center, 0.05)
    if kind == "boundary":
        x = S.uniform_coords(seed, M, d)
        x[: M // 4] = np.float32(-np.pi)
        x[M // 4: M // 2] = np.nextafter(np.float32(np.pi), np.float32(0))
        return x
    if kind == "duplicates":
        x = S.uniform_coords(seed, M // 4, d)
        return np.concatenate([x, x, x, x])
    raise ValueError(kind)


CASES_T1 = [
    # (n_modes, eps, isign, modeord, point kind, M)
    ([16, 12, 10], 1e-6, 1, 0, "uniform", 4000),
    ([16, 12, 10], 1e-6, -1, 1, "clustered", 4000),
    ([10, 10, 10], 1e-3, 1, 0, "boundary", 3000),
    ([12, 10, 8], 1e-5, -1, 0, "duplicates", 3000),
    ([40, 36], 1e-6, 1, 0, "uniform", 6000),
    ([32, 30], 1e-4, -1, 1, "clustered", 6000),
    ([128], 1e-6, 1, 0, "uniform", 5000),
    ([77], 1e-2, -1, 1, "boundary", 2000),
]


@pytest.mark.parametrize("n_modes,eps,isign,modeord,kind,M", CASES_T1)
def test_type1_f32_matches_direct_oracle(n_modes, eps, isign, modeord, kind, M):
    torch, W, O, S = _setup()
    d = len(n_modes)
    x = _points(S, kind, M, d, S.config_seed(1, 100 + M))
    c = S.complex_normal(S.config_seed(1, 200 + M), M)
    cfg = W.NufftConfig(n_modes, eps).with_isign(isign).with_mode_order(modeord)
    plan = W.NufftPlan.type1_gpu(cfg)
    f = plan.encode_gpu(torch.from_numpy(x).cuda(), torch.from_numpy(c).cuda())
    torch.cuda.synchronize()
    ref = O.direct_type1(x.astype(np.float64), c.astype(np.complex128), n_modes, isign, modeord)
    err = rel_l2(f.cpu().numpy(), ref)
    print(f"NUFFT_T1 d={d} eps={eps} kind={kind} err={err:.3e}")
    assert err <= TOL_FACTOR[d] * eps


@pytest.mark.parametrize("n_modes,eps,isign,modeord,kind,M", CASES_T1)
def test_type2_f32_matches_direct_oracle(n_modes, eps, isign, modeord, kind, M):
    torch, W, O, S = _setup()
    d = len(n_modes)
    x = _points(S, kind, M, d, S.config_seed(1, 300 + M))
    f = S.complex_normal(S.config_seed(1, 400 + M), int(np.prod(n_modes)))
    cfg = W.NufftConfig(n_modes, eps).with_isign(isign).with_mode_order(modeord)
    plan = W.NufftPlan.type2_gpu(cfg)
    out = plan.encode_gpu(torch.from_numpy(x).cuda(), torch.from_numpy(f).cuda())
    torch.cuda.synchronize()
    ref = O.direct_type2(x.astype(np.float64), f.astype(np.complex128), n_modes, isign, modeord)
    err = rel_l2(out.cpu().numpy(), ref)
    print(f"NUFFT_T2 d={d} eps={eps} kind={kind} err={err:.3e}")
    assert err <= TOL_FACTOR[d] * eps


@pytest.mark.parametrize("n_modes,eps,factor", [([12, 10, 8], 1e-8, 20.0),
                                                ([24, 20], 1e-12, 40.0),
                                                ([64], 1e-10, 20.0)])
def test_f64_plans_match_direct_oracle(n_modes, eps, factor):
    torch, W, O, S = _setup()
    d = len(n_modes)
    M = 3000
    x = S.uniform_coords(S.config_seed(5, 1), M, d, dtype=np.float64)
    c = S.complex_normal(S.config_seed(5, 2), M, dtype=np.complex128)
    fm = S.complex_normal(S.config_seed(5, 3), int(np.prod(n_modes)), dtype=np.complex128)
    cfg = W.NufftConfig(n_modes, eps).with_precision(W.F64)
    p1 = W.NufftPlan.type1_gpu(cfg)
    p2 = W.NufftPlan.type2_gpu(cfg)
    xt = torch.from_numpy(x).cuda()
    f = p1.encode_gpu(xt, torch.from_numpy(c).cuda())
    v = p2.encode_gpu(xt, torch.from_numpy(fm).cuda())
    torch.cuda.synchronize()
    e1 = rel_l2(f.cpu().numpy(), O.direct_type1(x, c, n_modes))
    e2 = rel_l2(v.cpu().numpy(), O.direct_type2(x, fm, n_modes))
    print(f"NUFFT_F64 d={d} eps={eps} t1={e1:.3e} t2={e2:.3e}")
    assert e1 <= factor * eps and e2 <= factor * eps


def test_single_mode_closed_form():
    """A single mode k = (1, 2, 3) gives exactly c e^{i(x + 2y + 3z)}
    (tests/gpu_nufft_3d.rs:466-514, at a smaller grid)."""
    torch, W, O, S = _setup()
    n_modes = [16, 16, 16]
    M = 2000
    x = S.uniform_coords(S.config_seed(2, 50), M, 3)
    fm = np.zeros(n_modes[::-1], dtype=np.complex64)
    fm[3 + 8, 2 + 8, 1 + 8] = 1.0 + 0.0j  # centered storage index k + N/2
    plan = W.NufftPlan.type2_gpu(W.NufftConfig(n_modes, 1e-6))
    out = plan.encode_gpu(torch.from_numpy(x).cuda(), torch.from_numpy(fm.ravel()).cuda())
    torch.cuda.synchronize()
    xd = x.astype(np.float64)
    exact = np.exp(1j * (xd[:, 0] + 2 * xd[:, 1] + 3 * xd[:, 2]))
    assert rel_l2(out.cpu().numpy(), exact) <= 32e-6


def test_type1_bitwise_repeat_and_shuffle_invariance():
    """Index order (default, the reference's): bitwise repeatable.  Coordinate
    order: also independent of the input order of the points, bit for bit; a
    switch after set_points re-bins the kept points."""
    torch, W, O, S = _setup()
    n_modes = [24, 20, 16]
    M = 20000
    x = S.clustered_coords(S.config_seed(2, 60), M, (0.3, -1.1, 2.0), 0.05)
    x[: M // 2] = S.uniform_coords(S.config_seed(2, 61), M // 2, 3)
    c = S.complex_normal(S.config_seed(2, 62), M)
    xt, ct = torch.from_numpy(x).cuda(), torch.from_numpy(c).cuda()
    plan = W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, 1e-6))
    runs = [plan.encode_gpu(xt, ct).cpu().numpy() for _ in range(3)]
    for r in runs[1:]:
        assert np.array_equal(r.view(np.uint32), runs[0].view(np.uint32))
    plan.set_points_gpu(xt)
    plan.set_point_order("coordinates")
    canon = plan.execute_gpu(ct).cpu().numpy()
    assert rel_l2(canon, runs[0]) <= 1e-6
    for s in range(2):
        p = S.shuffle_perm(S.config_seed(2, 70 + s), M)
        f = plan.encode_gpu(torch.from_numpy(x[p].copy()).cuda(), torch.from_numpy(c[p].copy()).cuda())
        assert np.array_equal(f.cpu().numpy().view(np.uint32), canon.view(np.uint32))
    plan.set_point_order("index")
    again = plan.encode_gpu(xt, ct).cpu().numpy()
    assert np.array_equal(again.view(np.uint32), runs[0].view(np.uint32))
    with pytest.raises(Exception):
        W.NufftPlan.type2_gpu(W.NufftConfig(n_modes, 1e-6)).set_point_order("coordinates")


def test_set_points_then_execute_equals_encode():
    torch, W, O, S = _setup()
    n_modes = [20, 18]
    M = 5000
    x = torch.from_numpy(S.uniform_coords(S.config_seed(1, 80), M, 2)).cuda()
    c1 = torch.from_numpy(S.complex_normal(S.config_seed(1, 81), M)).cuda()
    c2 = torch.from_numpy(S.complex_normal(S.config_seed(1, 82), M)).cuda()
    plan = W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, 1e-6))
    a = plan.encode_gpu(x, c1).cpu().numpy()
    b = plan.encode_gpu(x, c2).cpu().numpy()
    plan.set_points_gpu(x)
    a2 = plan.execute_gpu(c1).cpu().numpy()
    b2 = plan.execute_gpu(c2).cpu().numpy()
    assert np.array_equal(a.view(np.uint32), a2.view(np.uint32))
    assert np.array_equal(b.view(np.uint32), b2.view(np.uint32))


def test_batched_equals_loop_of_single_bitwise():
    """tests/gpu_nufft_batch.rs:96-133: a batch equals its vectors one by one."""
    torch, W, O, S = _setup()
    n_modes = [16, 14, 12]
    M = 6000
    B = 3
    x = torch.from_numpy(S.uniform_coords(S.config_seed(4, 1), M, 3)).cuda()
    c = S.complex_normal(S.config_seed(4, 2), M * B)
    fm = S.complex_normal(S.config_seed(4, 3), int(np.prod(n_modes)) * B)
    cfg = W.NufftConfig(n_modes, 1e-6).with_batch(B)
    p1 = W.NufftPlan.type1_gpu(cfg)
    p2 = W.NufftPlan.type2_gpu(cfg)
    fb = p1.encode_gpu(x, torch.from_numpy(c).cuda()).cpu().numpy()
    vb = p2.encode_gpu(x, torch.from_numpy(fm).cuda()).cpu().numpy()
    N = int(np.prod(n_modes))
    for v in range(B):
        fs = p1.encode_gpu(x, torch.from_numpy(c[v * M:(v + 1) * M].copy()).cuda(),
                           active_batch=1).cpu().numpy()
        vs = p2.encode_gpu(x, torch.from_numpy(fm[v * N:(v + 1) * N].copy()).cuda(),
                           active_batch=1).cpu().numpy()
        assert np.array_equal(fs.view(np.uint32), fb[v * N:(v + 1) * N].view(np.uint32))
        assert np.array_equal(vs.view(np.uint32), vb[v * M:(v + 1) * M].view(np.uint32))


def test_zero_points_give_zero_output():
    torch, W, O, S = _setup()
    plan = W.NufftPlan.type1_gpu(W.NufftConfig([8, 8], 1e-6))
    x = torch.zeros((0, 2), dtype=torch.float32, device="cuda")
    c = torch.zeros(1, dtype=torch.complex64, device="cuda")
    out = torch.full((64,), 7.0 + 7.0j, dtype=torch.complex64, device="cuda")
    plan.set_points_gpu(x)
    plan.execute_gpu(c, out=out, active_batch=1)
    assert torch.count_nonzero(out).item() == 0


def test_invalid_coordinates_do_not_disturb_other_points():
    """tests/gpu_invalid_points.rs:20-44: NaN, inf, 1e30 and 5 pi points."""
    torch, W, O, S = _setup()
    n_modes = [12, 10]
    M = 2000
    x = S.uniform_coords(S.config_seed(1, 90), M, 2)
    fm = S.complex_normal(S.config_seed(1, 91), int(np.prod(n_modes)))
    bad = x.copy()
    bad[0, 0] = np.nan
    bad[1, 1] = np.inf
    bad[2, 0] = 1e30
    bad[3, 1] = 5 * np.pi
    plan = W.NufftPlan.type2_gpu(W.NufftConfig(n_modes, 1e-6))
    good = plan.encode_gpu(torch.from_numpy(x).cuda(), torch.from_numpy(fm).cuda()).cpu().numpy()
    got = plan.encode_gpu(torch.from_numpy(bad).cuda(), torch.from_numpy(fm).cuda()).cpu().numpy()
    assert np.array_equal(got[4:].view(np.uint32), good[4:].view(np.uint32))
    c = S.complex_normal(S.config_seed(1, 92), M)
    c[:4] = 0
    p1 = W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, 1e-6))
    ref = O.direct_type1(x[4:].astype(np.float64), c[4:].astype(np.complex128), n_modes)
    f = p1.encode_gpu(torch.from_numpy(bad).cuda(), torch.from_numpy(c).cuda()).cpu().numpy()
    assert rel_l2(f, ref) <= 20e-6


@pytest.mark.parametrize("shape,kind,M", [([32, 32, 32], "uniform", 20000),
                                          ([32, 32, 32], "clustered", 20000),
                                          ([64, 64], "clustered", 30000),
                                          ([256], "uniform", 3000)])
def test_bin_sort_matches_reference_permutation(shape, kind, M):
    """Bit-exact stable bin permutation and offsets against the oracle's
    counting sort, with large bins (> 64 and > 2048 points) in the clustered
    cases."""
    torch, W, O, S = _setup()
    d = len(shape)
    x = _points(S, kind, M, d, S.config_seed(2, 500 + M))
    plan = W.NufftPlan.type1_gpu(W.NufftConfig(shape, 1e-6))
    info = plan.info()
    fine = info["fine"][:d]
    w = info["width"]
    lay = O.layout(fine, w, 0, M)
    side = lay["geometries"][lay["chosen"]]
    keys, _, _, amb = O.bin_keys_f32(fine, w, side, x)
    nbins = int(np.prod([(f + s - 1) // s for f, s in zip(fine, side)]))
    ref_sorted, ref_off = O.bin_sort(keys, nbins)
    assert amb == 0
    srt, off = plan.bin_sort(torch.from_numpy(x).cuda())
    torch.cuda.synchronize()
    assert off.numel() == nbins + 1
    assert np.array_equal(off.cpu().numpy().astype(np.uint32), ref_off)
    assert np.array_equal(srt.cpu().numpy().astype(np.uint32), ref_sorted)
    if kind == "clustered":
        assert np.diff(ref_off).max() > 64


def test_clustered_heavy_blocks_linearity():
    """tests/gpu_clustered_points.rs:160-225: a dense cluster (heavy blocks
    split into parts) spread in two halves sums to the whole."""
    torch, W, O, S = _setup()
    n_modes = [32, 32, 32]
    M = 300000
    x = S.clustered_coords(S.config_seed(2, 600), M, (0.3, -1.1, 2.0), 0.001)
    c = S.complex_normal(S.config_seed(2, 601), M)
    plan = W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, 1e-6))
    xt = torch.from_numpy(x).cuda()
    whole = plan.encode_gpu(xt, torch.from_numpy(c).cuda()).cpu().numpy()
    c1 = c.copy()
    c1[M // 2:] = 0
    c2 = c - c1
    a = plan.encode_gpu(xt, torch.from_numpy(c1).cuda()).cpu().numpy()
    b = plan.encode_gpu(xt, torch.from_numpy(c2).cuda()).cpu().numpy()
    assert rel_l2(a + b, whole) <= 1e-5
    sel = np.arange(0, int(np.prod(n_modes)), 97)
    ref = O.direct_type1(x.astype(np.float64), c.astype(np.complex128), n_modes, modes=sel)
    assert rel_l2(whole[sel], ref) <= 32e-6


def test_type3_matches_direct_oracle():
    torch, W, O, S = _setup()
    M, N = 2000, 1500
    x = S.uniform_coords(S.config_seed(5, 10), M, 3, lo=-5.0, hi=5.0, dtype=np.float64)
    s = S.uniform_coords(S.config_seed(5, 11), N, 3, lo=-20.0, hi=20.0, dtype=np.float64)
    c = S.complex_normal(S.config_seed(5, 12), M, dtype=np.complex128)
    for eps in (1e-6, 1e-10):
        plan = W.NufftType3Plan(3, [(-5, 5)] * 3, [(-20, 20)] * 3, eps=eps, precision=W.F64)
        out = plan.encode_gpu(torch.from_numpy(x).cuda(), torch.from_numpy(c).cuda(),
                              torch.from_numpy(s).cuda())
        torch.cuda.synchronize()
        ref = O.direct_type3(x, c, s)
        err = rel_l2(out.cpu().numpy(), ref)
        print(f"NUFFT_T3 eps={eps} err={err:.3e}")
        assert err <= 100 * eps


@pytest.mark.parametrize("dim,prec,eps,boxes", [
    (3, "f64", 1e-9, [2, 2, 2]), (3, "f64", 1e-6, [3, 1, 2]), (3, "f32", 1e-5, [2, 3, 1]),
    (2, "f32", 1e-5, [3, 2]), (1, "f64", 1e-10, [4])])
def test_type3_target_boxes_match_oracle(dim, prec, eps, boxes):
    """A type-3 plan whose target interval is cut into boxes (each box its
    own type-3 problem around the box centre) gives the direct sums to the
    plan's tolerance, like the uncut plan, for targets in every box and on
    box edges."""
    torch, W, O, S = _setup()
    M, N = 3000, 2000
    f64 = prec == "f64"
    dt = np.float64 if f64 else np.float32
    x = S.uniform_coords(S.config_seed(5, 70 + dim), M, dim, lo=-5.0, hi=5.0, dtype=dt)
    s = S.uniform_coords(S.config_seed(5, 71 + dim), N, dim, lo=-30.0, hi=30.0, dtype=dt)
    s[:3] = np.array([-30.0, 0.0, 29.999])[:, None]  # ends of the interval and a box edge
    c = S.complex_normal(S.config_seed(5, 72 + dim), M,
                         dtype=np.complex128 if f64 else np.complex64)
    ref = O.direct_type3(x.astype(np.float64), c.astype(np.complex128), s.astype(np.float64))
    xt, ct, st = (torch.from_numpy(a).cuda() for a in (x, c, s))
    outs = {}
    for name, kb in (("whole", None), ("boxes", boxes)):
        plan = W.NufftType3Plan(dim, [(-5, 5)] * dim, [(-30, 30)] * dim, eps=eps,
                                precision=W.F64 if f64 else W.F32)
        if kb is not None:
            assert plan.boxes(kb) == kb
        else:
            assert plan.boxes() == [1] * dim
        outs[name] = plan.encode_gpu(xt, ct, st).cpu().numpy()
        err = rel_l2(outs[name], ref)
        print(f"T3 {name} d={dim} {prec} eps={eps} boxes={kb}: err={err:.3e}", flush=True)
        assert err <= 100 * eps


@pytest.mark.parametrize("dim,M,tgt", [(1, 300_000, 30.0), (2, 120_000, 12.0), (2, 700, 5.0)])
def test_type3_small_outer_grid_f32(dim, M, tgt):
    """F32 1D / 2D type-3 plans whose outer grid is small (at most 65,536
    cells) spread their sources by point groups in input order with the
    group partial grids added in group order (src/gpu_dense_spread.rs):
    direct-sum tolerance, bitwise repeat, and a batch of two equal to the
    vectors one by one."""
    torch, W, O, S = _setup()
    eps, N, B = 1e-5, 400, 2
    x = S.uniform_coords(S.config_seed(5, 40 + dim), M, dim, lo=-5.0, hi=5.0)
    s = S.uniform_coords(S.config_seed(5, 50 + dim), N, dim, lo=-tgt, hi=tgt)
    c = S.complex_normal(S.config_seed(5, 60 + dim), M * B)
    plan = W.NufftType3Plan(dim, [(-5, 5)] * dim, [(-tgt, tgt)] * dim, eps=eps,
                            precision=W.F32, batch=B)
    outer = plan.info()["outer"][:dim]
    assert int(np.prod(outer)) <= 65536, outer
    xt, st = torch.from_numpy(x).cuda(), torch.from_numpy(s).cuda()
    ct = torch.from_numpy(c).cuda()
    out = plan.encode_gpu(xt, ct, st).clone()
    again = plan.encode_gpu(xt, ct, st)
    torch.cuda.synchronize()
    assert torch.equal(out.view(torch.float32), again.view(torch.float32))
    for v in range(B):
        one = plan.encode_gpu(xt, ct[v * M:(v + 1) * M].contiguous(), st, active_batch=1)
        torch.cuda.synchronize()
        assert torch.equal(one.view(torch.float32), out[v * N:(v + 1) * N].view(torch.float32))
        ref = O.direct_type3(x.astype(np.float64), c[v * M:(v + 1) * M].astype(np.complex128),
                             s.astype(np.float64))
        err = rel_l2(out[v * N:(v + 1) * N].cpu().numpy(), ref)
        print(f"T3 small grid d={dim} M={M} outer={outer} vector {v}: err={err:.3e}", flush=True)
        assert err <= 100 * eps


def test_gpu_direct_sums_match_oracle():
    torch, W, O, S = _setup()
    M = 500
    x = S.uniform_coords(S.config_seed(3, 1), M, 2, dtype=np.float64)
    c = S.complex_normal(S.config_seed(3, 2), M, dtype=np.complex128)
    n_modes = [9, 8]
    got = W.reference_type1_f64(n_modes, torch.from_numpy(x).cuda(), torch.from_numpy(c).cuda())
    ref = O.direct_type1(x, c, n_modes)
    assert rel_l2(got.cpu().numpy(), ref) <= 1e-13
    fm = S.complex_normal(S.config_seed(3, 3), 72, dtype=np.complex128)
    got2 = W.reference_type2_f64(n_modes, torch.from_numpy(x).cuda(), torch.from_numpy(fm).cuda())
    assert rel_l2(got2.cpu().numpy(), O.direct_type2(x, fm, n_modes)) <= 1e-13


def test_f64_row_sweep_and_tiled_interp():
    """F64 3D plans at a width <= 8 run the row-sweep spreader and the
    LDS-tiled interpolation (fine grids of at least 24 cells per axis)."""
    torch, W, O, S = _setup()
    n_modes = [16, 14, 12]
    M = 4000
    x = S.clustered_coords(S.config_seed(5, 20), M, (0.3, -1.1, 2.0), 0.2).astype(np.float64)
    c = S.complex_normal(S.config_seed(5, 21), M, dtype=np.complex128)
    fm = S.complex_normal(S.config_seed(5, 22), int(np.prod(n_modes)), dtype=np.complex128)
    cfg = W.NufftConfig(n_modes, 1e-6).with_precision(W.F64)
    p1 = W.NufftPlan.type1_gpu(cfg)
    p2 = W.NufftPlan.type2_gpu(cfg)
    xt = torch.from_numpy(x).cuda()
    f = p1.encode_gpu(xt, torch.from_numpy(c).cuda()).cpu().numpy()
    vv = p2.encode_gpu(xt, torch.from_numpy(fm).cuda()).cpu().numpy()
    assert p2.info()["bin_side"] == [8, 8, 8]
    e1 = rel_l2(f, O.direct_type1(x, c, n_modes))
    e2 = rel_l2(vv, O.direct_type2(x, fm, n_modes))
    print(f"NUFFT_F64_SWEEP t1={e1:.3e} t2={e2:.3e}")
    assert e1 <= 32e-6 and e2 <= 32e-6


def test_bin_sort_multi_round_merges():
    """A dense cluster puts tens of thousands of points in single bins (many
    radix tiles per bin); the permutation stays the reference's, bit for
    bit."""
    torch, W, O, S = _setup()
    shape = [32, 32, 32]
    M = 200000
    x = S.clustered_coords(S.config_seed(2, 700), M, (0.3, -1.1, 2.0), 0.002)
    plan = W.NufftPlan.type1_gpu(W.NufftConfig(shape, 1e-6))
    info = plan.info()
    fine = info["fine"]
    w = info["width"]
    lay = O.layout(fine, w, 0, M)
    side = lay["geometries"][lay["chosen"]]
    keys, _, _, amb = O.bin_keys_f32(fine, w, side, x)
    nbins = int(np.prod([(f + s - 1) // s for f, s in zip(fine, side)]))
    ref_sorted, ref_off = O.bin_sort(keys, nbins)
    assert amb == 0 and np.diff(ref_off).max() > 4 * 2048
    srt, off = plan.bin_sort(torch.from_numpy(x).cuda())
    torch.cuda.synchronize()
    assert np.array_equal(off.cpu().numpy().astype(np.uint32), ref_off)
    assert np.array_equal(srt.cpu().numpy().astype(np.uint32), ref_sorted)


def test_type1_dense_cluster_shuffle_invariance():
    """Canonical in-bin order with bins of tens of thousands of points (the
    large-bin merge sort runs several rounds): a shuffled input gives the
    same type-1 output, bit for bit, and it matches the direct sums."""
    torch, W, O, S = _setup()
    n_modes = [32, 32, 32]
    M = 200000
    x = S.clustered_coords(S.config_seed(2, 710), M, (0.3, -1.1, 2.0), 0.002)
    c = S.complex_normal(S.config_seed(2, 711), M)
    plan = W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, 1e-6)).set_point_order("coordinates")
    f0 = plan.encode_gpu(torch.from_numpy(x).cuda(), torch.from_numpy(c).cuda()).cpu().numpy()
    p = S.shuffle_perm(S.config_seed(2, 712), M)
    f1 = plan.encode_gpu(torch.from_numpy(x[p].copy()).cuda(),
                         torch.from_numpy(c[p].copy()).cuda()).cpu().numpy()
    assert np.array_equal(f0.view(np.uint32), f1.view(np.uint32))
    N = int(np.prod(n_modes))
    sel = np.arange(0, N, 211)
    ref = O.direct_type1(x.astype(np.float64), c.astype(np.complex128), n_modes, modes=sel)
    assert rel_l2(f0[sel], ref) <= 32e-6


@pytest.mark.parametrize("case", ["tight_cluster", "duplicates", "f64", "2d"])
def test_coordinate_order_large_groups_and_ranks(case):
    """Coordinate order where equal sort keys form large groups: a cluster
    far tighter than one position step (groups of thousands: the
    large-group merge sort with several rounds) and points repeated 200
    times with equal strengths (groups of 200); plus F64 and 2D plans.  A
    shuffled input gives the same output bit for bit, and it matches the
    direct sums."""
    torch, W, O, S = _setup()
    d = 2 if case == "2d" else 3
    n_modes = [40, 36] if d == 2 else [24, 20, 16]
    f64 = case == "f64"
    if case == "tight_cluster":
        M = 60000
        x = S.clustered_coords(S.config_seed(2, 720), M, (0.3, -1.1, 2.0), 1e-4)
        # distinct coordinates: exact duplicates keep index order (not
        # permutation independent when their strengths differ)
        assert np.unique(x, axis=0).shape[0] == M
        c = S.complex_normal(S.config_seed(2, 721), M)
    elif case == "duplicates":
        base = S.clustered_coords(S.config_seed(2, 722), 300, (0.3, -1.1, 2.0), 0.05)
        cb = S.complex_normal(S.config_seed(2, 723), 300)
        x = np.repeat(base, 200, axis=0)
        c = np.repeat(cb, 200)
        M = x.shape[0]
    else:
        M = 30000
        x = S.clustered_coords(S.config_seed(2, 724), M, (0.3, -1.1, 2.0)[:d], 0.02)
        x[: M // 2] = S.uniform_coords(S.config_seed(2, 725), M // 2, d)
        c = S.complex_normal(S.config_seed(2, 726), M)
    if f64:
        x = x.astype(np.float64)
        c = c.astype(np.complex128)
    cfg = W.NufftConfig(n_modes, 1e-9 if f64 else 1e-6)
    if f64:
        cfg = cfg.with_precision(W.F64)
    plan = W.NufftPlan.type1_gpu(cfg).set_point_order("coordinates")
    f0 = plan.encode_gpu(torch.from_numpy(x).cuda(), torch.from_numpy(c).cuda()).cpu().numpy()
    p = S.shuffle_perm(S.config_seed(2, 727), M)
    f1 = plan.encode_gpu(torch.from_numpy(x[p].copy()).cuda(),
                         torch.from_numpy(c[p].copy()).cuda()).cpu().numpy()
    view = np.uint64 if f64 else np.uint32
    assert np.array_equal(f0.view(view), f1.view(view))
    N = int(np.prod(n_modes))
    sel = np.arange(0, N, 97)
    ref = O.direct_type1(x.astype(np.float64), c.astype(np.complex128), n_modes, modes=sel)
    err = rel_l2(f0[sel], ref)
    print(f"COORD {case} err={err:.3e}")
    assert err <= (1e-7 if f64 else 32e-6)


def test_type2_batched_tiled_interp_matches_oracle():
    torch, W, O, S = _setup()
    n_modes = [20, 18, 16]
    M, B = 5000, 2
    x = S.uniform_coords(S.config_seed(4, 30), M, 3)
    fm = S.complex_normal(S.config_seed(4, 31), int(np.prod(n_modes)) * B)
    plan = W.NufftPlan.type2_gpu(W.NufftConfig(n_modes, 1e-6).with_batch(B))
    out = plan.encode_gpu(torch.from_numpy(x).cuda(), torch.from_numpy(fm).cuda()).cpu().numpy()
    N = int(np.prod(n_modes))
    for v in range(B):
        ref = O.direct_type2(x.astype(np.float64), fm[v * N:(v + 1) * N].astype(np.complex128),
                             n_modes)
        assert rel_l2(out[v * M:(v + 1) * M], ref) <= 32e-6


def test_spread_tuning_small_budget_matches_default():
    """A tiny partial-sum budget forces the light limit up and makes parts
    so large that some heavy tiles fit one part (written directly); the
    transform stays the same to rounding, for the plain and matrix-core
    spreads."""
    torch, W, O, S = _setup()
    n_modes = [32, 32, 32]
    M = 200000
    x = S.clustered_coords(S.config_seed(2, 1100), M, (0.3, -1.1, 2.0), 0.01)
    x[: M // 4] = S.uniform_coords(S.config_seed(2, 1101), M // 4, 3)
    c = S.complex_normal(S.config_seed(2, 1102), M)
    xt, ct = torch.from_numpy(x).cuda(), torch.from_numpy(c).cuda()
    sel = np.arange(0, int(np.prod(n_modes)), 101)
    ref = O.direct_type1(x.astype(np.float64), c.astype(np.complex128), n_modes, modes=sel)
    for method in ("plain", "mfma_f16x3"):
        base = W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, 1e-6)).set_spread_method(method)
        tuned = W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, 1e-6)).set_spread_method(method)
        tuned.set_spread_tuning(light_limit=256, min_part=64, part_budget_bytes=1)
        a = base.encode_gpu(xt, ct).cpu().numpy()
        b = tuned.encode_gpu(xt, ct).cpu().numpy()
        b2 = tuned.encode_gpu(xt, ct).cpu().numpy()
        assert np.array_equal(b.view(np.uint32), b2.view(np.uint32))
        assert rel_l2(b, a) <= 1e-6
        assert rel_l2(b[sel], ref) <= 32e-6


def test_batch_beyond_2_32_values():
    """32 vectors over a 512^3 fine grid hold 2^32 complex values: every
    vector, the last included, is the single-vector transform (the rocFFT
    executions and elementwise launches are split below 2^32 lanes /
    values); identical inputs give bitwise identical vectors."""
    torch, W, O, S = _setup()
    n_modes = [256, 256, 256]
    N = int(np.prod(n_modes))
    M, B = 20000, 32
    x = torch.from_numpy(S.uniform_coords(S.config_seed(4, 1200), M, 3)).cuda()
    c = torch.from_numpy(S.complex_normal(S.config_seed(4, 1201), M)).cuda()
    single = W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, 1e-6))
    f1 = single.encode_gpu(x, c)
    plan = W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, 1e-6).with_batch(B))
    fb = plan.encode_gpu(x, c.repeat(B)).view(B, N)
    for v in (B // 2 - 1, B // 2, B - 1):
        assert torch.equal(fb[v], fb[0]), v
    assert rel_l2(fb[B - 1].cpu().numpy(), f1.cpu().numpy()) <= 1e-6
    del plan, fb
    fm = torch.from_numpy(S.complex_normal(S.config_seed(4, 1202), N)).cuda()
    single2 = W.NufftPlan.type2_gpu(W.NufftConfig(n_modes, 1e-6))
    c1 = single2.encode_gpu(x, fm)
    plan2 = W.NufftPlan.type2_gpu(W.NufftConfig(n_modes, 1e-6).with_batch(B))
    cb = plan2.encode_gpu(x, fm.repeat(B)).view(B, M)
    for v in (B // 2 - 1, B // 2, B - 1):
        assert torch.equal(cb[v], cb[0]), v
    assert rel_l2(cb[B - 1].cpu().numpy(), c1.cpu().numpy()) <= 1e-6
    sel = torch.arange(0, M, 97, device="cuda")
    r2 = W.reference_type2_f64(n_modes, x.double(), fm.to(torch.complex128), points=sel)
    assert rel_l2(cb[B - 1][sel].cpu().numpy(), r2.cpu().numpy()) <= 32e-6
