"""Batched type-2 encodes of many independent plans (the batching form of
the reference's plan surface: one type2_gpu plan per problem,
src/plan.rs:606): one predeconvolution launch, one bin sort over all points
and one interpolation launch per kernel class (1D, 2D and 3D).
Every output must equal the plan's own encode bit for bit (each output is a
fixed-order sum over its support, independent of the point order).  It must
match the FP64 direct sums to the 2D tolerance (tests/gpu_nufft_2d.rs:
<= 20*eps).  A graph-captured batch must replay like the eager one.
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


def _bits(t):
    import torch
    return t.contiguous().view(torch.float32).view(torch.int32)


# (n_modes, M, eps, isign, mode order)
PROBLEMS_2D = [
    ([64, 48], 5000, 1e-5, 1, 0),
    ([100, 70], 0, 1e-5, -1, 1),
    ([33, 90], 123, 1e-3, 1, 1),
    ([128, 128], 20000, 1e-6, -1, 0),
    ([16, 20], 777, 1e-5, 1, 0),
    ([72, 40], 9000, 1e-3, -1, 0),
    ([50, 50], 1, 1e-6, 1, 1),
    ([90, 33], 15000, 1e-5, 1, 0),
]


def _make(W, S, torch, problems, seed0):
    plans, xs, fs = [], [], []
    for i, (n_modes, M, eps, isign, order) in enumerate(problems):
        cfg = W.NufftConfig(n_modes, eps).with_isign(isign).with_mode_order(order)
        plans.append(W.NufftPlan.type2_gpu(cfg))
        x = S.uniform_coords(S.config_seed(4, seed0 + i), M, len(n_modes))
        f = S.complex_normal(S.config_seed(4, seed0 + 100 + i), int(np.prod(n_modes)))
        xs.append(torch.from_numpy(x).cuda())
        fs.append(torch.from_numpy(f).cuda())
    return plans, xs, fs


def test_batch_2d_equals_each_plan_and_direct_sums():
    torch, W, O, S = _setup()
    plans, xs, fs = _make(W, S, torch, PROBLEMS_2D, 3000)
    single = [p.encode_gpu(x, f).clone() for p, x, f in zip(plans, xs, fs)]
    batch = W.encode_type2_many(plans, xs, fs)
    again = W.encode_type2_many(plans, xs, fs)
    torch.cuda.synchronize()
    for i, (a, b, c) in enumerate(zip(single, batch, again)):
        assert torch.equal(_bits(a), _bits(b)), f"problem {i} differs from its own encode"
        assert torch.equal(_bits(b), _bits(c)), f"problem {i} not repeatable"
    for i in (0, 3, 7):
        n_modes, M, eps, isign, order = PROBLEMS_2D[i]
        sel = np.arange(0, M, 7)
        ref = O.direct_type2(xs[i].cpu().numpy().astype(np.float64),
                             fs[i].cpu().numpy().astype(np.complex128), n_modes, isign=isign,
                             modeord=order, points=sel)
        e = rel_l2(batch[i].cpu().numpy()[sel], ref)
        print(f"BATCH problem {i} {n_modes} M={M} eps={eps}: err={e:.3e}", flush=True)
        assert e <= 20 * eps


def test_batch_1d_and_argument_errors():
    torch, W, O, S = _setup()
    problems = [([200], 3000, 1e-6, 1, 0), ([64], 500, 1e-3, -1, 1), ([1000], 10000, 1e-5, 1, 0)]
    plans, xs, fs = _make(W, S, torch, problems, 3100)
    single = [p.encode_gpu(x, f).clone() for p, x, f in zip(plans, xs, fs)]
    batch = W.encode_type2_many(plans, xs, fs)
    torch.cuda.synchronize()
    for a, b in zip(single, batch):
        assert torch.equal(_bits(a), _bits(b))
    # a type-1 plan, mixed dimensions, a repeated plan, a 3D plan
    p1 = W.NufftPlan.type1_gpu(W.NufftConfig([200], 1e-6))
    with pytest.raises(ValueError):
        W.encode_type2_many([plans[0], p1], xs[:2], fs[:2])
    p2d = W.NufftPlan.type2_gpu(W.NufftConfig([20, 20], 1e-6))
    x2 = torch.from_numpy(S.uniform_coords(S.config_seed(4, 3190), 100, 2)).cuda()
    f2 = torch.from_numpy(S.complex_normal(S.config_seed(4, 3191), 400)).cuda()
    with pytest.raises(W.NufftError):
        W.encode_type2_many([plans[0], p2d], [xs[0], x2], [fs[0], f2])
    with pytest.raises(W.NufftError):
        W.encode_type2_many([plans[0], plans[0]], [xs[0], xs[0]], [fs[0], fs[0]])
    # an F64 plan
    p64 = W.NufftPlan.type2_gpu(W.NufftConfig([20, 20], 1e-9).with_precision(W.F64))
    with pytest.raises(W.NufftError):
        W.encode_type2_many([plans[0], p64], [xs[0], x2.double()],
                            [fs[0], f2.to(torch.complex128)])


# 3D problems: LDS-tiled ones (w <= 8, every axis at least 8 + w - 1 cells)
# grouped by width, then the gather ones (small grids, w > 8)
PROBLEMS_3D = [
    ([32, 28, 24], 20000, 1e-6, 1, 0),
    ([16, 20, 12], 3000, 1e-3, -1, 1),
    ([24, 24, 24], 0, 1e-5, 1, 0),
    ([6, 5, 4], 800, 1e-6, -1, 0),
    ([20, 16, 18], 5000, 1e-7, 1, 1),
    ([40, 36, 32], 60000, 1e-5, -1, 0),
]


def test_batch_3d_equals_each_plan_and_direct_sums():
    torch, W, O, S = _setup()
    plans, xs, fs = _make(W, S, torch, PROBLEMS_3D, 3600)
    single = [p.encode_gpu(x, f).clone() for p, x, f in zip(plans, xs, fs)]
    batch = W.encode_type2_many(plans, xs, fs)
    again = W.encode_type2_many(plans, xs, fs)
    torch.cuda.synchronize()
    for i, (a, b, c) in enumerate(zip(single, batch, again)):
        assert torch.equal(_bits(a), _bits(b)), f"problem {i} differs from its own encode"
        assert torch.equal(_bits(b), _bits(c)), f"problem {i} not repeatable"
    for i in (0, 1, 3, 4, 5):
        n_modes, M, eps, isign, order = PROBLEMS_3D[i]
        sel = np.arange(0, M, 11)
        ref = O.direct_type2(xs[i].cpu().numpy().astype(np.float64),
                             fs[i].cpu().numpy().astype(np.complex128), n_modes, isign=isign,
                             modeord=order, points=sel)
        e = rel_l2(batch[i].cpu().numpy()[sel], ref)
        print(f"BATCH 3D problem {i} {n_modes} M={M} eps={eps}: err={e:.3e}", flush=True)
        assert e <= 32 * max(eps, 1e-6)


def test_batch_captured_in_a_graph():
    """The batch (per-plan predeconvolution and FFT, the combined sort and
    the one interpolation launch) captured once; replays on new modes and
    points written into the captured buffers equal eager batches."""
    torch, W, O, S = _setup()
    plans, xs, fs = _make(W, S, torch, PROBLEMS_2D[:5], 3200)
    outs = [torch.empty(x.shape[0], dtype=torch.complex64, device="cuda") for x in xs]
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            W.encode_type2_many(plans, xs, fs, outs)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        W.encode_type2_many(plans, xs, fs, outs)
    torch.cuda.synchronize()
    for r in range(3):
        for i, (x, f) in enumerate(zip(xs, fs)):
            if x.shape[0]:
                x.copy_(torch.from_numpy(
                    S.uniform_coords(S.config_seed(4, 3300 + 10 * r + i), x.shape[0], 2)))
            f.copy_(torch.from_numpy(
                S.complex_normal(S.config_seed(4, 3400 + 10 * r + i), f.numel())))
        graph.replay()
        torch.cuda.synchronize()
        got = [o.clone() for o in outs]
        eager = W.encode_type2_many(plans, xs, fs)
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(zip(got, eager)):
            assert torch.equal(_bits(a), _bits(b)), f"replay {r} problem {i}"
        print(f"BATCH graph replay {r} bitwise equal", flush=True)
