"""The fine-grid FFT of 2D and 3D plans as pruned line-transform passes
(the default engine) against rocFFT's plan on the same plan and points, and
both against the FP64 direct sums to the reference's tolerances
(tests/gpu_nufft_2d.rs:15, gpu_nufft_3d.rs:13: 20 and 32 eps).  The reference's FFT is its wgpu-fft
dependency (Cargo.toml:34, not vendored), so the engines are pinned by the direct sums
and by each other: type 1 reads only the kept modes of the transformed grid,
type 2 reads only the modes' cells, and both agree with rocFFT to the FFT's
rounding (FP32 4e-7, FP64 1e-14 relative).
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


CASES = [
    # n_modes, eps, precision, isign, mode order, batch
    ([24, 20, 16], 1e-6, "f32", 1, 0, 1),
    ([64, 33], 1e-6, "f32", -1, 0, 1),
    ([21, 50], 1e-11, "f64", 1, 1, 2),
    ([17, 13, 9], 1e-5, "f32", -1, 1, 2),
    ([30, 24, 50], 1e-6, "f32", 1, 1, 1),
    ([16, 14, 12], 1e-12, "f64", -1, 0, 1),
    ([25, 9, 20], 1e-9, "f64", 1, 1, 3),
]


@pytest.mark.parametrize("n_modes,eps,prec,isign,order,batch", CASES)
def test_line_fft_matches_rocfft_and_direct_sums(n_modes, eps, prec, isign, order, batch):
    torch, W, O, S = _setup()
    f64 = prec == "f64"
    M = 4000
    d = len(n_modes)
    x = S.uniform_coords(S.config_seed(10, M + len(n_modes)), M, d,
                         dtype=np.float64 if f64 else np.float32)
    cdt = np.complex128 if f64 else np.complex64
    N = int(np.prod(n_modes))
    c = np.stack([S.complex_normal(S.config_seed(10, 50 + b), M, dtype=cdt) for b in range(batch)])
    fm = np.stack([S.complex_normal(S.config_seed(10, 80 + b), N, dtype=cdt)
                   for b in range(batch)])
    cfg = W.NufftConfig(n_modes, eps).with_isign(isign).with_mode_order(order).with_batch(batch)
    if f64:
        cfg = cfg.with_precision(W.F64)
    xt = torch.from_numpy(x).cuda()
    outs = {}
    for engine in ("lines", "rocfft"):
        p1 = W.NufftPlan.type1_gpu(cfg).set_fft_engine(engine)
        p2 = W.NufftPlan.type2_gpu(cfg).set_fft_engine(engine)
        assert p1.fft_engine() == engine and p2.fft_engine() == engine
        f = p1.encode_gpu(xt, torch.from_numpy(c.ravel()).cuda()).clone()
        v = p2.encode_gpu(xt, torch.from_numpy(fm.ravel()).cuda()).clone()
        torch.cuda.synchronize()
        outs[engine] = (f.cpu().numpy().reshape(batch, -1), v.cpu().numpy().reshape(batch, -1))
    lim = 1e-13 if f64 else 1e-6
    for b in range(batch):
        d1 = rel_l2(outs["lines"][0][b], outs["rocfft"][0][b])
        d2 = rel_l2(outs["lines"][1][b], outs["rocfft"][1][b])
        e1 = rel_l2(outs["lines"][0][b], O.direct_type1(x.astype(np.float64), c[b].astype(
            np.complex128), n_modes, isign=isign, modeord=order))
        e2 = rel_l2(outs["lines"][1][b], O.direct_type2(x.astype(np.float64), fm[b].astype(
            np.complex128), n_modes, isign, order))
        print(f"LINE FFT {n_modes} {prec} isign={isign} order={order} vec {b}: "
              f"t1 vs rocFFT {d1:.2e} t2 vs rocFFT {d2:.2e} err t1 {e1:.2e} t2 {e2:.2e}",
              flush=True)
        assert d1 <= lim and d2 <= lim
        assert e1 <= {2: 20, 3: 32}[d] * eps and e2 <= {2: 20, 3: 32}[d] * eps


def test_line_fft_type3_and_unsupported_grids():
    """A 3D type-3 plan runs its inner type-2 plan on the line FFT and agrees
    with rocFFT; 1D plans and a 3D grid with an axis beyond the LDS lines
    stay on rocFFT."""
    torch, W, O, S = _setup()
    M, N = 3000, 2000
    src = S.uniform_coords(S.config_seed(10, 90), M, 3, lo=-3.0, hi=3.0, dtype=np.float64)
    tgt = S.uniform_coords(S.config_seed(10, 91), N, 3, lo=-20.0, hi=20.0, dtype=np.float64)
    c = S.complex_normal(S.config_seed(10, 92), M, dtype=np.complex128)
    outs = {}
    for engine in ("lines", "rocfft"):
        p = W.NufftType3Plan(3, [(-3, 3)] * 3, [(-20, 20)] * 3, eps=1e-9, isign=1,
                             precision=W.F64)
        p.set_fft_engine(engine)
        out = p.encode_gpu(torch.from_numpy(src).cuda(), torch.from_numpy(c).cuda(),
                           torch.from_numpy(tgt).cuda())
        torch.cuda.synchronize()
        assert p.fft_engine() == engine
        outs[engine] = out.cpu().numpy()
    d = rel_l2(outs["lines"], outs["rocfft"])
    ref = O.direct_type3(src, c, tgt, 1)
    e = rel_l2(outs["lines"], ref)
    print(f"LINE FFT type 3: vs rocFFT {d:.2e}, err {e:.2e}", flush=True)
    assert d <= 1e-13 and e <= 100 * 1e-9
    p2 = W.NufftPlan.type1_gpu(W.NufftConfig([32, 32], 1e-6))
    assert p2.fft_engine() == "lines"
    assert W.NufftPlan.type1_gpu(W.NufftConfig([100], 1e-6)).fft_engine() == "rocfft"
    big = W.NufftPlan.type2_gpu(W.NufftConfig([2500, 8, 8], 1e-3))
    assert big.fft_engine() == "rocfft"
    with pytest.raises(ValueError):
        p2.set_fft_engine("other")
