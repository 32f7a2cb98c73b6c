"""Batched type-1 encodes of many independent plans (the batching form of the
reference's plan surface: one type1_gpu plan per problem,
src/plan.rs:457-463, 808-872): one bin sort over all points, one prepare,
one spread launch over every plan's blocks, each plan's FFT, one
deconvolution launch.  Every output must equal the plan's own encode bit for
bit (each plan's blocks are split by its own work plan, so even heavy blocks
sum in the same order), repeat bit for bit, and match the FP64 direct sums
to the reference's tolerances (tests/gpu_nufft.rs:14-18, gpu_nufft_2d.rs:15,
gpu_nufft_3d.rs:13: 8, 20, 32 eps).
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = {1: 8.0, 2: 20.0, 3: 32.0}


def _setup():
    import torch
    import wgpunufft_amd as W
    from oracle import oracle as O
    from wgpunufft_amd import synthetic as S
    return torch, W, O, S


def _bits(t):
    import torch
    t = t.contiguous()
    if t.dtype == torch.complex128:
        return t.view(torch.float64).view(torch.int64)
    return t.view(torch.float32).view(torch.int32)


def _make(W, S, torch, problems, seed0, precision=None, plain_rows=True):
    """Plans for a batch: their defaults (2D and 3D: the row tiles where the
    grid qualifies), or the output-stationary blocks unless plain_rows."""
    plans, xs, cs = [], [], []
    f64 = precision is not None and precision != W.F32
    for i, (n_modes, M, eps, isign, order, kind) in enumerate(problems):
        cfg = W.NufftConfig(n_modes, eps).with_isign(isign).with_mode_order(order)
        if precision is not None:
            cfg = cfg.with_precision(precision)
        p = W.NufftPlan.type1_gpu(cfg)
        if len(n_modes) >= 2 and not plain_rows:
            p.set_bin_tuning(0, plain_rows=False)
        plans.append(p)
        d = len(n_modes)
        dt = np.float64 if f64 else np.float32
        if kind == "clustered":
            x = S.clustered_coords(S.config_seed(7, seed0 + i), M, [0.3, -1.1, 2.0][:d], 0.05,
                                   dtype=dt)
        else:
            x = S.uniform_coords(S.config_seed(7, seed0 + i), M, d, dtype=dt)
        c = S.complex_normal(S.config_seed(7, seed0 + 100 + i), M,
                             dtype=np.complex128 if f64 else np.complex64)
        xs.append(torch.from_numpy(x).cuda())
        cs.append(torch.from_numpy(c).cuda())
    return plans, xs, cs


def _check(torch, W, O, problems, plans, xs, cs, sample=(0,)):
    single = [p.encode_gpu(x, c).clone() for p, x, c in zip(plans, xs, cs)]
    batch = W.encode_type1_many(plans, xs, cs)
    again = W.encode_type1_many(plans, xs, cs)
    torch.cuda.synchronize()
    for i, (a, b, c) in enumerate(zip(single, batch, again)):
        assert torch.equal(_bits(a), _bits(b)), f"problem {i} differs from its own encode"
        assert torch.equal(_bits(b), _bits(c)), f"problem {i} not repeatable"
    for i in sample:
        n_modes, M, eps, isign, order, kind = problems[i]
        nm = int(np.prod(n_modes))
        modes = np.random.default_rng(i).choice(nm, min(nm, 700), replace=False)
        ref = O.direct_type1(xs[i].cpu().numpy().astype(np.float64),
                             cs[i].cpu().numpy().astype(np.complex128), n_modes, isign=isign,
                             modeord=order, modes=modes)
        e = rel_l2(batch[i].cpu().numpy()[modes], ref)
        print(f"BATCH T1 problem {i} {n_modes} M={M} eps={eps} {kind}: err={e:.3e}", flush=True)
        assert e <= TOL[len(n_modes)] * eps


# (n_modes, M, eps, isign, mode order, point kind)
PROBLEMS_2D = [
    ([64, 48], 5000, 1e-5, 1, 0, "uniform"),
    ([100, 70], 0, 1e-5, -1, 1, "uniform"),
    ([33, 90], 123, 1e-3, 1, 1, "uniform"),
    ([128, 128], 200000, 1e-6, -1, 0, "clustered"),
    ([16, 20], 777, 1e-5, 1, 0, "uniform"),
    ([72, 40], 90000, 1e-3, -1, 0, "clustered"),
    ([50, 50], 1, 1e-6, 1, 1, "uniform"),
    ([90, 33], 15000, 1e-5, 1, 0, "uniform"),
]


def test_batch_type1_2d_equals_each_plan_and_direct_sums():
    """Uneven modes and point counts, a problem without points, one point,
    two dense clusters whose blocks are split into heavy parts."""
    torch, W, O, S = _setup()
    plans, xs, cs = _make(W, S, torch, PROBLEMS_2D, 5000)
    _check(torch, W, O, PROBLEMS_2D, plans, xs, cs, sample=(0, 3, 5, 7))


def test_batch_type1_2d_row_tiles_and_blocks_in_one_batch():
    """2D plans on the row tiles (three kernel widths: three tile classes)
    and on the output-stationary blocks in one batch: each output equals its
    plan's own encode bit for bit, and the direct sums."""
    torch, W, O, S = _setup()
    plans, xs, cs = _make(W, S, torch, PROBLEMS_2D, 5050)
    for i, p in enumerate(plans):
        if i % 3 == 1:
            p.set_bin_tuning(0, plain_rows=False)
    tiles = [i for i, p in enumerate(plans) if p.info()["run"] == 0]
    assert len({plans[i].info()["width"] for i in tiles}) == 3 and len(tiles) < len(plans)
    _check(torch, W, O, PROBLEMS_2D, plans, xs, cs, sample=(1, 3, 5))


def test_batch_type1_1d_3d_and_f64():
    torch, W, O, S = _setup()
    p1 = [([200], 3000, 1e-6, 1, 0, "uniform"), ([64], 500, 1e-3, -1, 1, "uniform"),
          ([1000], 100000, 1e-5, 1, 0, "clustered")]
    plans, xs, cs = _make(W, S, torch, p1, 5100)
    _check(torch, W, O, p1, plans, xs, cs, sample=(0, 2))
    p3 = [([24, 20, 16], 30000, 1e-6, 1, 0, "clustered"), ([12, 12, 12], 2000, 1e-3, -1, 1,
                                                           "uniform"),
          ([32, 28, 24], 0, 1e-5, 1, 0, "uniform"), ([16, 16, 16], 8000, 1e-5, -1, 0, "uniform")]
    plans, xs, cs = _make(W, S, torch, p3, 5200)
    assert sum(p.info()["run"] == 0 for p in plans) >= 2
    _check(torch, W, O, p3, plans, xs, cs, sample=(0, 3))
    # the same problems on the output-stationary blocks, and mixed
    plans, xs, cs = _make(W, S, torch, p3, 5200, plain_rows=False)
    _check(torch, W, O, p3, plans, xs, cs, sample=(0,))
    plans, xs, cs = _make(W, S, torch, p3, 5200)
    plans[1].set_bin_tuning(0, plain_rows=False)
    _check(torch, W, O, p3, plans, xs, cs, sample=(1,))
    p3d = [([20, 18, 16], 20000, 1e-9, 1, 0, "clustered"), ([16, 16, 12], 3000, 1e-12, -1, 1,
                                                           "uniform")]
    plans, xs, cs = _make(W, S, torch, p3d, 5250, precision=W.F64)
    _check(torch, W, O, p3d, plans, xs, cs, sample=(0, 1))
    p64 = [([40, 36], 4000, 1e-9, 1, 0, "uniform"), ([24, 20], 30000, 1e-12, -1, 1, "clustered")]
    plans, xs, cs = _make(W, S, torch, p64, 5300, precision=W.F64)
    _check(torch, W, O, p64, plans, xs, cs, sample=(0, 1))


def test_batch_type1_refuses_other_spreads():
    torch, W, O, S = _setup()
    x = torch.from_numpy(S.uniform_coords(S.config_seed(7, 5400), 100, 3)).cuda()
    c = torch.from_numpy(S.complex_normal(S.config_seed(7, 5401), 100)).cuda()
    coord = W.NufftPlan.type1_gpu(W.NufftConfig([32, 32, 32], 1e-6))
    coord.set_point_order("coordinates")
    with pytest.raises(W.NufftError):
        W.encode_type1_many([coord], [x], [c])
    mfma = W.NufftPlan.type1_gpu(W.NufftConfig([32, 32, 32], 1e-6))
    mfma.set_spread_method("mfma_f16x3")
    with pytest.raises(W.NufftError):
        W.encode_type1_many([mfma], [x], [c])
    p2 = W.NufftPlan.type1_gpu(W.NufftConfig([20, 20], 1e-6))
    x2 = torch.from_numpy(S.uniform_coords(S.config_seed(7, 5402), 100, 2)).cuda()
    p2b = W.NufftPlan.type1_gpu(W.NufftConfig([20, 20], 1e-6).with_precision(W.F64))
    with pytest.raises(W.NufftError):
        W.encode_type1_many([p2, p2b], [x2, x2.double()], [c, c.to(torch.complex128)])
    with pytest.raises(W.NufftError):
        W.encode_type1_many([p2, p2], [x2, x2], [c, c])


def test_batch_type1_captured_in_a_graph():
    """The whole batch captured once; replays on new points and strengths
    written into the captured buffers equal eager batches bit for bit."""
    torch, W, O, S = _setup()
    problems = PROBLEMS_2D[:6]
    plans, xs, cs = _make(W, S, torch, problems, 5500)
    outs = [torch.empty(p.config.mode_count(), dtype=torch.complex64, device="cuda")
            for p in plans]
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            W.encode_type1_many(plans, xs, cs, outs)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        W.encode_type1_many(plans, xs, cs, outs)
    torch.cuda.synchronize()
    for r in range(3):
        for i, (x, c) in enumerate(zip(xs, cs)):
            if x.shape[0]:
                x.copy_(torch.from_numpy(
                    S.uniform_coords(S.config_seed(7, 5600 + 10 * r + i), x.shape[0], 2)))
                c.copy_(torch.from_numpy(
                    S.complex_normal(S.config_seed(7, 5700 + 10 * r + i), c.numel())))
        graph.replay()
        torch.cuda.synchronize()
        got = [o.clone() for o in outs]
        eager = W.encode_type1_many(plans, xs, cs)
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(zip(got, eager)):
            assert torch.equal(_bits(a), _bits(b)), f"replay {r} problem {i}"
        print(f"BATCH T1 graph replay {r} bitwise equal", flush=True)
