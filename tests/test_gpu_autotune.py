"""The type-1 autotuner: it times the candidate spreads on the caller's data
and keeps the fastest one whose operand error stays within the plan's
tolerance.  The tuned plan must still meet the reference's 3D / 2D
tolerance against FP64 direct sums (tests/gpu_nufft_3d.rs:13: <= 32*eps;
2D <= 20*eps) and stay bitwise repeatable."""
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


# 2D plans offer the plain spread only: the 2D matrix-core tiles are switched
# off in the engine (an open GPU memory fault at 1024^2 modes, 5e7 points).
@pytest.mark.parametrize("n_modes,eps,allowed,min_cands", [
    ([32, 28, 24], 1e-6, {"plain", "mfma_f16x3"}, 2),
    ([32, 28, 24], 1e-3, {"plain", "mfma_f16x3", "mfma_bf16x3", "mfma_f16"}, 2),
    ([64, 48], 1e-6, {"plain"}, 1),
])
def test_autotune_picks_an_allowed_spread_and_meets_tolerance(n_modes, eps, allowed, min_cands):
    torch, W, O, S = _setup()
    d = len(n_modes)
    M = 150000
    x = S.clustered_coords(S.config_seed(2, 1600 + d), M, (0.3, -1.1, 2.0)[:d], 0.05)
    x[: M // 2] = S.uniform_coords(S.config_seed(2, 1601 + d), M // 2, d)
    c = S.complex_normal(S.config_seed(2, 1602 + d), M)
    plan = W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, eps)).set_point_order("coordinates")
    xt, ct = torch.from_numpy(x).cuda(), torch.from_numpy(c).cuda()
    choice = plan.autotune(xt, ct, repeats=2)
    print(f"AUTOTUNE {n_modes} eps={eps}: {choice}", flush=True)
    assert choice["spread_method"] in allowed
    assert choice["candidates"] >= min_cands
    f1 = plan.encode_gpu(xt, ct).clone()
    f2 = plan.execute_gpu(ct).clone()
    torch.cuda.synchronize()
    assert torch.equal(f1.view(torch.float32).view(torch.int32),
                       f2.view(torch.float32).view(torch.int32))
    N = int(np.prod(n_modes))
    sel = np.arange(0, N, 13)
    ref = O.direct_type1(x.astype(np.float64), c.astype(np.complex128), n_modes, modes=sel)
    e = rel_l2(f1.cpu().numpy()[sel], ref)
    print(f"AUTOTUNE error {e:.3e}", flush=True)
    assert e <= (32 if d == 3 else 20) * eps


def test_autotune_rejects_type2():
    torch, W, O, S = _setup()
    p = W.NufftPlan.type2_gpu(W.NufftConfig([16, 16, 16], 1e-6))
    x = torch.from_numpy(S.uniform_coords(S.config_seed(2, 1650), 100, 3)).cuda()
    c = torch.zeros(100, dtype=torch.complex64, device="cuda")
    with pytest.raises(ValueError):
        p.autotune(x, c)


@pytest.mark.parametrize("method,side0,plain_rows", [
    ("plain", 0, False), ("plain", 4, True), ("mfma_f16x3", 8, True), ("mfma_f16x3", 4, True)])
def test_bin_tuning_knobs_meet_tolerance(method, side0, plain_rows):
    """Every bins/tiles knob the autotuner may set gives the direct sums to
    the reference's 3D tolerance and stays bitwise repeatable; the plain
    output-stationary blocks and the row tiles agree to FP32 rounding."""
    torch, W, O, S = _setup()
    n_modes, eps, M = [32, 28, 24], 1e-6, 120000
    x = S.clustered_coords(S.config_seed(2, 1700), M, (0.3, -1.1, 2.0), 0.05)
    x[: M // 2] = S.uniform_coords(S.config_seed(2, 1701), M // 2, 3)
    c = S.complex_normal(S.config_seed(2, 1702), M)
    xt, ct = torch.from_numpy(x).cuda(), torch.from_numpy(c).cuda()
    plan = W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, eps)).set_spread_method(method)
    plan.set_bin_tuning(side0, plain_rows)
    t = plan.tuning()
    assert t["rows_side0"] == side0 and t["plain_rows"] == plain_rows
    f1 = plan.encode_gpu(xt, ct).clone()
    f2 = plan.execute_gpu(ct).clone()
    torch.cuda.synchronize()
    assert torch.equal(f1.view(torch.int32), f2.view(torch.int32))
    ref_plan = W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, eps))
    base = ref_plan.encode_gpu(xt, ct).cpu().numpy()
    got = f1.cpu().numpy()
    assert rel_l2(got, base) <= (2e-6 if method != "plain" else 1e-6)
    N = int(np.prod(n_modes))
    sel = np.arange(0, N, 17)
    ref = O.direct_type1(x.astype(np.float64), c.astype(np.complex128), n_modes, modes=sel)
    e = rel_l2(got[sel], ref)
    print(f"BIN TUNING {method} side0={side0} rows={plain_rows}: err={e:.3e}", flush=True)
    assert e <= 32 * eps
