"""Plain 2D type-1 spread on 16 x 16 row tiles (the 2D counterpart of the 3D
row tiles; the reference spreads on its output-stationary blocks,
src/gpu_nd_spread.rs:1583-1667): matches the FP64 direct sums to the
reference's 2D tolerance (tests/gpu_nufft_2d.rs:15: 20 eps), agrees with the
output-stationary blocks to rounding, repeats bit for bit, splits dense
clusters into heavy parts, and in coordinate order does not depend on the
order of the points.
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


@pytest.mark.parametrize("n_modes,eps,prec,kind,M", [
    ([64, 48], 1e-6, "f32", "uniform", 20000),
    ([100, 70], 1e-3, "f32", "clustered", 200000),
    ([40, 36], 1e-7, "f64", "clustered", 60000),
    ([33, 90], 1e-5, "f32", "uniform", 5000),
])
def test_2d_row_tiles_match_blocks_and_direct_sums(n_modes, eps, prec, kind, M):
    torch, W, O, S = _setup()
    f64 = prec == "f64"
    dt = np.float64 if f64 else np.float32
    if kind == "clustered":
        x = S.clustered_coords(S.config_seed(11, M), M, (0.3, -1.1), 0.05, dtype=dt)
        x[: M // 3] = S.uniform_coords(S.config_seed(11, M + 1), M // 3, 2, dtype=dt)
    else:
        x = S.uniform_coords(S.config_seed(11, M), M, 2, dtype=dt)
    c = S.complex_normal(S.config_seed(11, M + 2), M,
                         dtype=np.complex128 if f64 else np.complex64)
    cfg = W.NufftConfig(n_modes, eps)
    if f64:
        cfg = cfg.with_precision(W.F64)
    xt, ct = torch.from_numpy(x).cuda(), torch.from_numpy(c).cuda()
    rows = W.NufftPlan.type1_gpu(cfg)
    assert rows.tuning()["plain_rows"]
    blocks = W.NufftPlan.type1_gpu(cfg).set_bin_tuning(0, False)
    a = rows.encode_gpu(xt, ct).clone()
    again = rows.execute_gpu(ct).clone()
    b = blocks.encode_gpu(xt, ct).clone()
    torch.cuda.synchronize()
    iv = torch.int64 if f64 else torch.int32
    fv = torch.float64 if f64 else torch.float32
    assert torch.equal(a.view(fv).view(iv), again.view(fv).view(iv))
    an, bn = a.cpu().numpy(), b.cpu().numpy()
    ref = O.direct_type1(x.astype(np.float64), c.astype(np.complex128), n_modes)
    e_rows, e_blocks, d = rel_l2(an, ref), rel_l2(bn, ref), rel_l2(an, bn)
    print(f"ROWS2D {n_modes} {prec} eps={eps} {kind}: rows {e_rows:.3e} blocks {e_blocks:.3e} "
          f"rows-vs-blocks {d:.3e}", flush=True)
    assert e_rows <= 20 * eps
    assert d <= (1e-13 if f64 else 2e-6)


def test_2d_row_tiles_coordinate_order_shuffle_invariance():
    torch, W, O, S = _setup()
    M = 150000
    x = S.clustered_coords(S.config_seed(11, 7), M, (0.3, -1.1), 0.05)
    x[: M // 2] = S.uniform_coords(S.config_seed(11, 8), M // 2, 2)
    c = S.complex_normal(S.config_seed(11, 9), M)
    perm = np.random.default_rng(3).permutation(M)
    plan = W.NufftPlan.type1_gpu(W.NufftConfig([128, 96], 1e-6)).set_point_order("coordinates")
    a = plan.encode_gpu(torch.from_numpy(x).cuda(), torch.from_numpy(c).cuda()).clone()
    b = plan.encode_gpu(torch.from_numpy(x[perm].copy()).cuda(),
                        torch.from_numpy(c[perm].copy()).cuda()).clone()
    torch.cuda.synchronize()
    assert torch.equal(a.view(torch.float32).view(torch.int32),
                       b.view(torch.float32).view(torch.int32))
