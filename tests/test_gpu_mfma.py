"""GPU parity of the matrix-core (MFMA) type-1 spread against the CPU oracle.

The reference spreads in plain FP32; the matrix-core methods must give the
same transform to the reference's 3D tolerance (relative l2 error against
FP64 direct sums <= 32*eps, tests/gpu_nufft_3d.rs:13), stay within their
stated operand error of the plain FP32 spread, and keep the plain path's
bitwise repeatability and batch == loop properties
(tests/gpu_nufft_batch.rs:96-133), and its shuffle invariance under the
coordinate point order.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

# method -> (tolerance of the plan, bound on |mfma - plain| / |plain|)
METHODS = {
    "mfma_f16": (1e-3, 2e-3),
    "mfma_f16x3": (1e-6, 2e-6),
    "mfma_bf16x3": (1e-6, 6e-5),
}


def _setup():
    import torch
    import wgpunufft_amd as W
    from oracle import oracle as O
    from wgpunufft_amd import synthetic as S
    return torch, W, O, S


def _pts(S, kind, M, seed, d=3):
    if kind == "uniform":
        return S.uniform_coords(seed, M, d)
    x = S.clustered_coords(seed, M, (0.3, -1.1, 2.0)[:d], 0.05)
    x[: M // 3] = S.uniform_coords(seed + 1, M // 3, d)
    return x


# The 2D matrix-core tiles are switched off in the engine: the 1024^2-mode,
# 5e7-point plan ended in a GPU memory fault whose cause is not found.
OFF_2D = pytest.mark.skip(reason="2D matrix-core spread switched off: open GPU memory fault "
                                 "at 1024^2 modes, 5e7 points")

CASES = [
    # (method, n_modes, isign, modeord, kind, M)
    ("mfma_f16x3", [16, 14, 12], 1, 0, "uniform", 4000),
    ("mfma_f16x3", [20, 18, 16], -1, 1, "clustered", 5000),
    ("mfma_bf16x3", [16, 14, 12], 1, 0, "clustered", 4000),
    ("mfma_f16", [16, 14, 12], -1, 0, "uniform", 4000),
    ("mfma_f16", [24, 20, 18], 1, 1, "clustered", 5000),
    pytest.param("mfma_f16x3", [40, 36], 1, 0, "uniform", 6000, marks=OFF_2D),
    pytest.param("mfma_f16x3", [70, 44], -1, 1, "clustered", 8000, marks=OFF_2D),
    pytest.param("mfma_bf16x3", [48, 40], 1, 0, "uniform", 6000, marks=OFF_2D),
    pytest.param("mfma_f16", [64, 50], -1, 0, "clustered", 8000, marks=OFF_2D),
]


@pytest.mark.parametrize("method,n_modes,isign,modeord,kind,M", CASES)
def test_mfma_type1_matches_direct_oracle(method, n_modes, isign, modeord, kind, M):
    torch, W, O, S = _setup()
    eps, bound = METHODS[method]
    x = _pts(S, kind, M, S.config_seed(2, 900 + M), len(n_modes))
    c = S.complex_normal(S.config_seed(2, 901 + M), M)
    cfg = W.NufftConfig(n_modes, eps).with_isign(isign).with_mode_order(modeord)
    plain = W.NufftPlan.type1_gpu(cfg)
    mf = W.NufftPlan.type1_gpu(cfg).set_spread_method(method)
    xt, ct = torch.from_numpy(x).cuda(), torch.from_numpy(c).cuda()
    fp = plain.encode_gpu(xt, ct).cpu().numpy()
    fm = mf.encode_gpu(xt, ct).cpu().numpy()
    exact = O.direct_type1(x.astype(np.float64), c.astype(np.complex128), n_modes, isign=isign,
                           modeord=modeord)
    e_m = rel_l2(fm, exact)
    e_p = rel_l2(fp, exact)
    d = rel_l2(fm, fp)
    print(f"MFMA {method} {n_modes} {kind}: err={e_m:.3e} plain={e_p:.3e} vs_plain={d:.3e}")
    assert e_m <= (32 if len(n_modes) == 3 else 20) * eps
    assert d <= bound


@OFF_2D
def test_mfma_2d_bitwise_batch_shuffle_and_rebinning():
    """2D tiles: bitwise repeat, shuffle invariance and batch == loop; a
    method switch after set_points re-bins the kept points."""
    torch, W, O, S = _setup()
    n_modes = [96, 80]
    M, B = 30000, 2
    x = _pts(S, "clustered", M, S.config_seed(1, 940), 2)
    c = S.complex_normal(S.config_seed(1, 941), M * B)
    xt = torch.from_numpy(x).cuda()
    plan = W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, 1e-6).with_batch(B))
    plan.set_point_order("coordinates")
    plan.set_points_gpu(xt)
    plain = plan.execute_gpu(torch.from_numpy(c).cuda()).cpu().numpy()
    plan.set_spread_method("mfma_f16x3")
    fb = plan.execute_gpu(torch.from_numpy(c).cuda()).cpu().numpy()
    assert rel_l2(fb, plain) <= 2e-6
    N = int(np.prod(n_modes))
    for v in range(B):
        fs = plan.encode_gpu(xt, torch.from_numpy(c[v * M:(v + 1) * M].copy()).cuda(),
                             active_batch=1).cpu().numpy()
        assert np.array_equal(fs.view(np.uint32), fb[v * N:(v + 1) * N].view(np.uint32))
    p = S.shuffle_perm(S.config_seed(1, 942), M)
    cp = np.concatenate([c[v * M:(v + 1) * M][p] for v in range(B)])
    f2 = plan.encode_gpu(torch.from_numpy(x[p].copy()).cuda(), torch.from_numpy(cp).cuda())
    assert np.array_equal(f2.cpu().numpy().view(np.uint32), fb.view(np.uint32))
    sel = np.arange(0, N, 37)
    ref = O.direct_type1(x.astype(np.float64), c[:M].astype(np.complex128), n_modes, modes=sel)
    assert rel_l2(fb[:N][sel], ref) <= 20e-6


def test_mfma_bitwise_repeat_and_shuffle_invariance():
    torch, W, O, S = _setup()
    n_modes = [24, 20, 16]
    M = 20000
    x = _pts(S, "clustered", M, S.config_seed(2, 960))
    c = S.complex_normal(S.config_seed(2, 962), M)
    plan = W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, 1e-6)).set_spread_method("mfma_f16x3")
    plan.set_point_order("coordinates")
    runs = [plan.encode_gpu(torch.from_numpy(x).cuda(), torch.from_numpy(c).cuda()).cpu().numpy()
            for _ in range(3)]
    for r in runs[1:]:
        assert np.array_equal(r.view(np.uint32), runs[0].view(np.uint32))
    for s in range(2):
        p = S.shuffle_perm(S.config_seed(2, 970 + s), M)
        f = plan.encode_gpu(torch.from_numpy(x[p].copy()).cuda(),
                            torch.from_numpy(c[p].copy()).cuda())
        assert np.array_equal(f.cpu().numpy().view(np.uint32), runs[0].view(np.uint32))


@pytest.mark.parametrize("dense_min", [0, 32])
def test_mfma_heavy_tiles_batch_equals_loop(dense_min):
    """A dense cluster splits tiles into parts (partial sums added in part
    order) or, with dense bins, is summed bin by bin; a batch of two vectors
    equals the vectors one by one, bit for bit, and matches the direct
    sums."""
    torch, W, O, S = _setup()
    n_modes = [32, 32, 32]
    M, B = 200000, 2
    x = S.clustered_coords(S.config_seed(2, 980), M, (0.3, -1.1, 2.0), 0.002)
    c = S.complex_normal(S.config_seed(2, 981), M * B)
    plan = W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, 1e-6).with_batch(B))
    plan.set_spread_method("mfma_f16x3").set_dense_bins(dense_min)
    xt = torch.from_numpy(x).cuda()
    fb = plan.encode_gpu(xt, torch.from_numpy(c).cuda()).cpu().numpy()
    N = int(np.prod(n_modes))
    for v in range(B):
        fs = plan.encode_gpu(xt, torch.from_numpy(c[v * M:(v + 1) * M].copy()).cuda(),
                             active_batch=1).cpu().numpy()
        assert np.array_equal(fs.view(np.uint32), fb[v * N:(v + 1) * N].view(np.uint32))
    sel = np.arange(0, N, 211)
    ref = O.direct_type1(x.astype(np.float64), c[:M].astype(np.complex128), n_modes, modes=sel)
    assert rel_l2(fb[:N][sel], ref) <= 32e-6


def test_mfma_strength_scale_is_exact():
    """Strengths are scaled by a power of two before the FP16 rounding:
    scaling the input by 2^k scales the output by exactly 2^k."""
    torch, W, O, S = _setup()
    n_modes = [16, 14, 12]
    M = 3000
    x = torch.from_numpy(S.uniform_coords(S.config_seed(2, 990), M, 3)).cuda()
    c = S.complex_normal(S.config_seed(2, 991), M)
    plan = W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, 1e-6)).set_spread_method("mfma_f16x3")
    a = plan.encode_gpu(x, torch.from_numpy(c).cuda()).cpu().numpy()
    for k in (40, -40):
        b = plan.encode_gpu(x, torch.from_numpy((c * np.float32(2.0 ** k)).astype(np.complex64))
                            .cuda()).cpu().numpy()
        assert np.array_equal((a * np.float32(2.0 ** k)).view(np.uint32), b.view(np.uint32))


def test_set_spread_method_rejects_unsupported_plans():
    torch, W, O, S = _setup()
    with pytest.raises(Exception):
        W.NufftPlan.type1_gpu(W.NufftConfig([16, 14, 12], 1e-6).with_precision(W.F64)) \
            .set_spread_method("mfma_f16x3")
    with pytest.raises(Exception):
        W.NufftPlan.type1_gpu(W.NufftConfig([12, 10], 1e-6)).set_spread_method("mfma_f16x3")
    # 2D matrix-core tiles are switched off at every size
    with pytest.raises(Exception):
        W.NufftPlan.type1_gpu(W.NufftConfig([64, 50], 1e-6)).set_spread_method("mfma_f16x3")
    with pytest.raises(Exception):
        W.NufftPlan.type1_gpu(W.NufftConfig([200], 1e-6)).set_spread_method("mfma_f16x3")
    with pytest.raises(Exception):
        W.NufftPlan.type2_gpu(W.NufftConfig([16, 14, 12], 1e-6)).set_spread_method("mfma_f16x3")
    with pytest.raises(Exception):
        W.NufftPlan.type1_gpu(W.NufftConfig([4, 4, 4], 1e-6)).set_spread_method("mfma_f16x3")
    W.NufftPlan.type1_gpu(W.NufftConfig([16, 14, 12], 1e-6)).set_spread_method("plain")
