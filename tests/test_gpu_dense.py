"""Dense bins of the 3D matrix-core type-1 spread.

A bin holding many points is summed once over its window, and the windows
are added to the tile sums in a fixed bin order.  The transform must meet
the reference's 3D tolerance against FP64 direct sums (<= 32*eps,
tests/gpu_nufft_3d.rs:13).  It must stay within the operand error of the
tile-only spread (dense bins off).  It must keep the tile path's properties:
bitwise repeat, shuffle invariance in coordinate order, and batch == loop
(tests/gpu_nufft_batch.rs:96-133, tests/gpu_clustered_points.rs:160-225).
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

# method -> (plan tolerance, bound on |dense - tile-only| / |tile-only|)
METHODS = {
    "mfma_f16x3": (1e-6, 2e-6),
    "mfma_bf16x3": (1e-6, 6e-5),
    "mfma_f16": (1e-3, 2e-3),
}


def _setup():
    import torch
    import wgpunufft_amd as W
    from oracle import oracle as O
    from wgpunufft_amd import synthetic as S
    return torch, W, O, S


def _points(S, M, seed, center=(0.3, -1.1, 2.0), sigma=0.05):
    """A third uniform, the rest in a wrapped Gaussian cluster."""
    x = S.clustered_coords(seed, M, center, sigma)
    x[: M // 3] = S.uniform_coords(seed + 1, M // 3, 3)
    return x


def _bits(t):
    import torch
    return t.contiguous().view(torch.float32).view(torch.int32)


CASES = [
    # (method, n_modes, order, center, M)
    ("mfma_f16x3", [24, 20, 16], "index", (0.3, -1.1, 2.0), 60000),
    ("mfma_f16x3", [24, 20, 16], "coordinates", (0.3, -1.1, 2.0), 60000),
    # a cluster across the periodic edges of every axis: wrapped windows
    ("mfma_f16x3", [20, 24, 18], "coordinates", (3.12, -3.12, 3.1), 50000),
    ("mfma_bf16x3", [24, 20, 16], "index", (0.3, -1.1, 2.0), 40000),
    ("mfma_f16", [24, 20, 16], "coordinates", (-2.0, 0.5, 1.0), 40000),
]


@pytest.mark.parametrize("method,n_modes,order,center,M", CASES)
def test_dense_bins_match_direct_oracle(method, n_modes, order, center, M):
    torch, W, O, S = _setup()
    eps, bound = METHODS[method]
    x = _points(S, M, S.config_seed(2, 1200 + M), center)
    c = S.complex_normal(S.config_seed(2, 1201 + M), M)
    cfg = W.NufftConfig(n_modes, eps)
    dense = W.NufftPlan.type1_gpu(cfg).set_spread_method(method).set_point_order(order)
    tiles = (W.NufftPlan.type1_gpu(cfg).set_spread_method(method).set_point_order(order)
             .set_dense_bins(0))
    xt, ct = torch.from_numpy(x).cuda(), torch.from_numpy(c).cuda()
    fd = dense.encode_gpu(xt, ct)
    ft = tiles.encode_gpu(xt, ct)
    torch.cuda.synchronize()
    exact = O.direct_type1(x.astype(np.float64), c.astype(np.complex128), n_modes)
    e_d = rel_l2(fd.cpu().numpy(), exact)
    e_t = rel_l2(ft.cpu().numpy(), exact)
    d = rel_l2(fd.cpu().numpy(), ft.cpu().numpy())
    print(f"dense {method} {n_modes} {order} {center}: err={e_d:.3e} tiles={e_t:.3e} "
          f"vs_tiles={d:.3e}", flush=True)
    assert e_d <= 32 * eps
    assert d <= bound
    # a different summation order ran (the dense windows), not the tile path
    assert not torch.equal(_bits(fd), _bits(ft))


def test_dense_bins_bitwise_repeat_shuffle_and_batch():
    """Bitwise repeat, shuffle invariance (coordinate order) and batch ==
    loop with dense bins; a threshold change re-bins the kept points."""
    torch, W, O, S = _setup()
    M, B, n_modes = 80000, 3, [24, 24, 24]
    N = int(np.prod(n_modes))
    x = _points(S, M, S.config_seed(2, 1300))
    c = S.complex_normal(S.config_seed(2, 1301), M * B)
    cfg = W.NufftConfig(n_modes, 1e-6).with_batch(B)
    p = W.NufftPlan.type1_gpu(cfg).set_spread_method("mfma_f16x3").set_point_order("coordinates")
    xt, ct = torch.from_numpy(x).cuda(), torch.from_numpy(c).cuda()
    f1 = p.encode_gpu(xt, ct)
    f2 = p.encode_gpu(xt, ct)
    perm = S.shuffle_perm(S.config_seed(2, 1302), M)
    cp = np.concatenate([c[v * M:(v + 1) * M][perm] for v in range(B)])
    f3 = p.encode_gpu(torch.from_numpy(x[perm].copy()).cuda(), torch.from_numpy(cp).cuda())
    torch.cuda.synchronize()
    assert torch.equal(_bits(f1), _bits(f2))
    assert torch.equal(_bits(f1), _bits(f3))
    for v in range(B):
        fv = p.encode_gpu(xt, ct[v * M:(v + 1) * M].contiguous(), active_batch=1)
        torch.cuda.synchronize()
        assert torch.equal(_bits(fv), _bits(f1[v * N:(v + 1) * N]))
    # every occupied bin dense: the threshold is raised until the windows fit
    p.set_dense_bins(1)
    p.set_points_gpu(xt)
    f4 = p.execute_gpu(ct)
    f5 = p.execute_gpu(ct)
    torch.cuda.synchronize()
    assert torch.equal(_bits(f4), _bits(f5))
    assert rel_l2(f4.cpu().numpy(), f1.cpu().numpy()) <= 2e-6
    # off again: an execute re-bins the kept points
    p.set_dense_bins(0)
    f6 = p.execute_gpu(ct)
    torch.cuda.synchronize()
    assert rel_l2(f6.cpu().numpy(), f1.cpu().numpy()) <= 2e-6
    assert not torch.equal(_bits(f6), _bits(f1))


def test_dense_bins_wide_windows():
    """A 256^3-mode plan bins 8 cells along x (windows 14 cells wide, two
    column blocks); dense vs tile-only spread and FP64 sums on sampled modes."""
    torch, W, O, S = _setup()
    M, n_modes = 200000, [256, 256, 256]
    x = _points(S, M, S.config_seed(2, 1400), sigma=0.02)
    c = S.complex_normal(S.config_seed(2, 1401), M)
    cfg = W.NufftConfig(n_modes, 1e-6)
    dense = W.NufftPlan.type1_gpu(cfg).set_spread_method("mfma_f16x3")
    tiles = W.NufftPlan.type1_gpu(cfg).set_spread_method("mfma_f16x3").set_dense_bins(0)
    xt, ct = torch.from_numpy(x).cuda(), torch.from_numpy(c).cuda()
    fd = dense.encode_gpu(xt, ct)
    ft = tiles.encode_gpu(xt, ct)
    torch.cuda.synchronize()
    d = rel_l2(fd.cpu().numpy(), ft.cpu().numpy())
    sel = np.sort(np.random.default_rng(7).choice(256 ** 3, 600, replace=False)).astype(np.int64)
    exact = O.direct_type1(x.astype(np.float64), c.astype(np.complex128), n_modes, modes=sel)
    e_d = rel_l2(fd.cpu().numpy()[sel], exact)
    print(f"dense 256^3: err={e_d:.3e} vs_tiles={d:.3e}", flush=True)
    assert e_d <= 32 * 1e-6
    assert d <= 2e-6
    assert not torch.equal(_bits(fd), _bits(ft))


def test_dense_bins_vector_chunks():
    """18 vectors: the dense windows of a vector take 64 MB, so the vectors
    are spread in two chunks; every vector equals its own execution bit for
    bit."""
    torch, W, O, S = _setup()
    M, B, n_modes = 20000, 18, [16, 16, 16]
    N = int(np.prod(n_modes))
    x = _points(S, M, S.config_seed(2, 1500))
    c = S.complex_normal(S.config_seed(2, 1501), M * B)
    p = (W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, 1e-6).with_batch(B))
         .set_spread_method("mfma_f16x3").set_point_order("coordinates"))
    xt, ct = torch.from_numpy(x).cuda(), torch.from_numpy(c).cuda()
    fb = p.encode_gpu(xt, ct)
    torch.cuda.synchronize()
    for v in (0, 15, 16, 17):
        fv = p.encode_gpu(xt, ct[v * M:(v + 1) * M].contiguous(), active_batch=1)
        torch.cuda.synchronize()
        assert torch.equal(_bits(fv), _bits(fb[v * N:(v + 1) * N])), f"vector {v}"
    tiles = (W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, 1e-6).with_batch(B))
             .set_spread_method("mfma_f16x3").set_point_order("coordinates").set_dense_bins(0))
    ft = tiles.encode_gpu(xt, ct)
    torch.cuda.synchronize()
    assert rel_l2(fb.cpu().numpy(), ft.cpu().numpy()) <= 2e-6
    assert not torch.equal(_bits(fb), _bits(ft))
