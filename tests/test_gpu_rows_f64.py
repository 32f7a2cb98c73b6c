"""FP64 3D type-1 plans of kernel widths 9 to 16 on the plain 8x8x8 row
tiles (the reference spreads them on its output-stationary blocks,
src/gpu_nd_spread.rs:1583-1667, with the F64 layouts of
gpu_nd_spread.rs:194-407): results match the FP64 direct sums to the
reference's F64 tolerances (tests/gpu_nufft_3d.rs:13, 32 eps; 40 eps here as
in the 2D F64 case), agree with the output-stationary blocks to FP64
rounding and repeat bit for bit; the clustered sets put tens of thousands of
records in reach of a few tiles.
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


# grids of at least 16 + width + 1 cells along x and 16 + width along y, z
@pytest.mark.parametrize("eps,kind,n_modes", [(1e-8, "uniform", [16, 16, 16]),
                                              (1e-10, "clustered", [16, 16, 16]),
                                              (1e-12, "clustered", [16, 16, 16]),
                                              (1e-13, "uniform", [16, 16, 16]),
                                              (1e-15, "uniform", [20, 18, 18])])
def test_f64_row_tiles_wide_kernels(eps, kind, n_modes):
    torch, W, O, S = _setup()
    w, _ = O.es_params(eps)
    assert 9 <= w <= 16
    M = 30000 if kind == "clustered" else 3000
    if kind == "clustered":
        x = S.clustered_coords(S.config_seed(9, 100 + w), M, (0.3, -1.1, 2.0), 0.05,
                               dtype=np.float64)
        x[: M // 3] = S.uniform_coords(S.config_seed(9, 200 + w), M // 3, 3, dtype=np.float64)
    else:
        x = S.uniform_coords(S.config_seed(9, 300 + w), M, 3, dtype=np.float64)
    c = S.complex_normal(S.config_seed(9, 400 + w), M, dtype=np.complex128)
    xt, ct = torch.from_numpy(x).cuda(), torch.from_numpy(c).cuda()
    cfg = W.NufftConfig(n_modes, eps).with_precision(W.F64)
    rows = W.NufftPlan.type1_gpu(cfg)
    assert rows.tuning()["plain_rows"]
    blocks = W.NufftPlan.type1_gpu(cfg).set_bin_tuning(0, False)
    f_rows = rows.encode_gpu(xt, ct).clone()
    again = rows.execute_gpu(ct).clone()
    f_blocks = blocks.encode_gpu(xt, ct).clone()
    torch.cuda.synchronize()
    bits = lambda t: t.view(torch.float64).view(torch.int64)
    assert torch.equal(bits(f_rows), bits(again)), "row tiles not repeatable"
    a, b = f_rows.cpu().numpy(), f_blocks.cpu().numpy()
    # two kernels; where both bin geometries give the same record order per
    # cell they agree bit for bit, else to FP64 rounding
    d = rel_l2(a, b)
    ref = O.direct_type1(x, c, n_modes)
    e = rel_l2(a, ref)
    e_b = rel_l2(b, ref)
    print(f"ROWS F64 w={w} eps={eps} {kind}: rows err={e:.3e} blocks err={e_b:.3e} "
          f"rows-vs-blocks={d:.3e}", flush=True)
    assert d <= 1e-13
    assert e <= 40 * max(eps, 1e-14)
