"""GPU parity of the rank-generic (4D to 8D) type-1, type-2 and type-3 paths
against the CPU oracle's FP64 direct sums.

The cases follow the reference's rank-generic GPU suite
(tests/gpu_nufft_nd.rs:37-200): F32 at tol 1e-3 (relative l2 <= 8 eps) and
1e-6 (<= 16 eps), sigma 2 and 1.25, both signs and mode orders, batches,
coordinates beyond one period (2.5 pi, -2.75 pi) so the fold is exercised;
the adjoint identity (:490-560, residual <= 2e-5); the F64 (<= 40 eps at
1e-8) and double-float cases (the engine runs those in native F64; 48 eps at
1e-8, 16 eps at 1e-6) (:150-180, 380-490); set_points + execute equal to the
encode bit for bit (:206-300); and a rank above eight rejected (:186-199).
Beyond the reference: the stable bin permutation bit for bit against the
oracle in 4D and 5D, batches equal to loops of single vectors bit for bit,
a clustered 4D set that takes the heavy-block split, and 4D type 3.
"""
import ctypes

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


def _points(S, seed, M, d, dtype=np.float32):
    """Uniform points in [-pi, pi) with point 0 at 2.5 pi and point 1 at
    -2.75 pi on every axis (beyond one period)."""
    x = S.uniform_coords(seed, M, d, dtype=dtype)
    x[0] = 2.5 * np.pi
    x[1] = -2.75 * np.pi
    return x


RUN_CASES = [
    # (label, n_modes, eps, sigma, isign, mode_order, batch, tolerance factor)
    ("4d eps=1e-3 sigma=2 negative centered", [6, 5, 4, 4], 1e-3, 2.0, -1, 0, 1, 8.0),
    ("4d eps=1e-3 sigma=2 positive fft-order", [6, 5, 4, 4], 1e-3, 2.0, 1, 1, 1, 8.0),
    ("4d eps=1e-3 sigma=1.25 negative centered", [6, 5, 4, 4], 1e-3, 1.25, -1, 0, 1, 8.0),
    ("4d eps=1e-6 sigma=2 negative centered", [8, 7, 6, 5], 1e-6, 2.0, -1, 0, 1, 16.0),
    ("4d eps=1e-3 sigma=2 batch=3", [5, 4, 4, 3], 1e-3, 2.0, -1, 0, 3, 8.0),
    ("5d eps=1e-3 sigma=2 negative centered", [4, 4, 3, 3, 3], 1e-3, 2.0, -1, 0, 1, 8.0),
    ("6d eps=1e-3 sigma=2 negative centered", [4, 4, 3, 3, 3, 3], 1e-3, 2.0, -1, 0, 1, 8.0),
    ("7d eps=1e-3 sigma=2 positive fft-order batch=2", [3, 3, 3, 2, 2, 2, 2], 1e-3, 2.0, 1, 1,
     2, 8.0),
    ("8d eps=1e-3 sigma=2 negative centered", [3, 2, 2, 2, 2, 2, 2, 2], 1e-3, 2.0, -1, 0, 1,
     8.0),
]


@pytest.mark.parametrize("label,n_modes,eps,sigma,isign,modeord,batch,factor", RUN_CASES)
def test_nd_f32_matches_direct_oracle(label, n_modes, eps, sigma, isign, modeord, batch,
                                      factor):
    torch, W, O, S = _setup()
    d = len(n_modes)
    M = 37
    nm = int(np.prod(n_modes))
    print(f"CASE {label}", flush=True)
    x = _points(S, S.config_seed(6, d), M, d)
    f = S.complex_normal(S.config_seed(6, 10 + d), nm * batch)
    c = S.complex_normal(S.config_seed(6, 20 + d), M * batch)
    cfg = (W.NufftConfig(n_modes, eps).with_isign(isign).with_mode_order(modeord)
           .with_sigma(sigma).with_batch(batch))
    xt = torch.from_numpy(x).cuda()
    out2 = W.NufftPlan.type2_gpu(cfg).encode_gpu(xt, torch.from_numpy(f).cuda()).cpu().numpy()
    out1 = W.NufftPlan.type1_gpu(cfg).encode_gpu(xt, torch.from_numpy(c).cuda()).cpu().numpy()
    x64 = x.astype(np.float64)
    for v in range(batch):
        ref2 = O.direct_type2(x64, f[v * nm:(v + 1) * nm].astype(np.complex128), n_modes,
                              isign, modeord)
        ref1 = O.direct_type1(x64, c[v * M:(v + 1) * M].astype(np.complex128), n_modes, isign,
                              modeord)
        e2 = rel_l2(out2[v * M:(v + 1) * M], ref2)
        e1 = rel_l2(out1[v * nm:(v + 1) * nm], ref1)
        print(f"NUFFT_ND {label} vector {v}: type-2 {e2:.3e} type-1 {e1:.3e}", flush=True)
        assert e2 <= factor * eps and e1 <= factor * eps


def test_nd_adjoint_identity():
    """<y, A x> == <A^H y, x>: type 2 of one sign is the adjoint of type 1
    of the other (tests/gpu_nufft_nd.rs:490-560)."""
    torch, W, O, S = _setup()
    n_modes, eps, M = [6, 5, 4, 4], 1e-6, 29
    nm = int(np.prod(n_modes))
    x = S.uniform_coords(S.config_seed(6, 40), M, 4)
    fh = S.complex_normal(S.config_seed(6, 41), nm)
    y = S.complex_normal(S.config_seed(6, 42), M)
    xt = torch.from_numpy(x).cuda()
    t2 = W.NufftPlan.type2_gpu(W.NufftConfig(n_modes, eps).with_isign(-1)).encode_gpu(
        xt, torch.from_numpy(fh).cuda()).cpu().numpy().astype(np.complex128)
    t1 = W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, eps).with_isign(1)).encode_gpu(
        xt, torch.from_numpy(y).cuda()).cpu().numpy().astype(np.complex128)
    lhs = np.vdot(y.astype(np.complex128), t2)
    rhs = np.vdot(t1, fh.astype(np.complex128))
    res = abs(lhs - rhs) / max(abs(lhs), np.finfo(float).tiny)
    print(f"NUFFT_ND_ADJOINT 4d: residual={res:.3e}")
    assert res <= 2e-5


@pytest.mark.parametrize("label,n_modes,eps,precision,factor", [
    ("4d eps=1e-8 f64", [10, 9, 8, 8], 1e-8, "F64", 40.0),
    ("4d eps=1e-8 df64", [10, 9, 8, 8], 1e-8, "DF64", 48.0),
    ("6d eps=1e-6 df64", [4, 4, 3, 3, 3, 3], 1e-6, "DF64", 16.0),
])
def test_nd_high_precision_matches_direct_oracle(label, n_modes, eps, precision, factor):
    torch, W, O, S = _setup()
    d = len(n_modes)
    M = 37
    nm = int(np.prod(n_modes))
    x = _points(S, S.config_seed(6, 50 + d), M, d, dtype=np.float64)
    f = S.complex_normal(S.config_seed(6, 51 + d), nm, dtype=np.complex128)
    c = S.complex_normal(S.config_seed(6, 52 + d), M, dtype=np.complex128)
    cfg = W.NufftConfig(n_modes, eps).with_isign(-1).with_precision(getattr(W, precision))
    xt = torch.from_numpy(x).cuda()
    out2 = W.NufftPlan.type2_gpu(cfg).encode_gpu(xt, torch.from_numpy(f).cuda()).cpu().numpy()
    out1 = W.NufftPlan.type1_gpu(cfg).encode_gpu(xt, torch.from_numpy(c).cuda()).cpu().numpy()
    e2 = rel_l2(out2, O.direct_type2(x, f, n_modes, -1))
    e1 = rel_l2(out1, O.direct_type1(x, c, n_modes, -1))
    print(f"NUFFT_ND_ACCURACY {label}: type-2 {e2:.3e} type-1 {e1:.3e}")
    assert e2 <= factor * eps and e1 <= factor * eps


@pytest.mark.parametrize("n_modes", [[6, 5, 4, 4], [4, 3, 3, 3, 3, 3]])
def test_nd_set_points_then_execute_equals_encode(n_modes):
    """tests/gpu_nufft_nd.rs:206-300: executions after set_points equal the
    encode bit for bit, twice in a row."""
    torch, W, O, S = _setup()
    d = len(n_modes)
    M = 211
    nm = int(np.prod(n_modes))
    x = torch.from_numpy(S.uniform_coords(S.config_seed(6, 60 + d), M, d, lo=-3.0,
                                          hi=3.0)).cuda()
    f = torch.from_numpy(S.complex_normal(S.config_seed(6, 61 + d), nm)).cuda()
    c = torch.from_numpy(S.complex_normal(S.config_seed(6, 62 + d), M)).cuda()
    cfg = W.NufftConfig(n_modes, 1e-4).with_isign(-1)
    p2, p1 = W.NufftPlan.type2_gpu(cfg), W.NufftPlan.type1_gpu(cfg)
    e2 = p2.encode_gpu(x, f).cpu().numpy()
    e1 = p1.encode_gpu(x, c).cpu().numpy()
    p2.set_points_gpu(x)
    p1.set_points_gpu(x)
    for _ in range(2):
        assert np.array_equal(p2.execute_gpu(f).cpu().numpy().view(np.uint32), e2.view(np.uint32))
        assert np.array_equal(p1.execute_gpu(c).cpu().numpy().view(np.uint32), e1.view(np.uint32))


def test_rank_above_eight_is_rejected():
    torch, W, O, S = _setup()
    nm = (ctypes.c_int64 * 9)(*([2] * 9))
    h = ctypes.c_void_p()
    for kind in (1, 2):
        rc = W.lib().mnufft_makeplan(kind, 9, nm, 1, 1, 1e-3, W.F32, 0, 2.0, ctypes.byref(h))
        assert rc == 2 and not h.value
        assert b"1 to 8" in W.lib().mnufft_last_error()


@pytest.mark.parametrize("shape,kind,M", [([16, 16, 12, 12], "clustered", 20000),
                                          ([10, 10, 8, 8, 8], "uniform", 5000)])
def test_nd_bin_sort_matches_reference_permutation(shape, kind, M):
    """The stable counting sort of the reference's bin keys (geometry of its
    layout for M points), bit for bit, in 4D and 5D."""
    torch, W, O, S = _setup()
    d = len(shape)
    if kind == "clustered":
        x = S.clustered_coords(S.config_seed(6, 70 + d), M, [0.3, -1.1, 2.0, 0.5][:d] +
                               [0.0] * (d - 4), 0.05)
    else:
        x = S.uniform_coords(S.config_seed(6, 70 + d), M, d)
    plan = W.NufftPlan.type1_gpu(W.NufftConfig(shape, 1e-6))
    info = plan.info()
    w = info["width"]
    fine = [O.fine_size(n, 2.0, w) for n in shape]
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


def test_nd_clustered_heavy_blocks_batch_equals_loop():
    """A dense 4D cluster: blocks past the light limit are split into parts
    (the heavy path); the result meets the tolerance on sampled modes, a
    batch of 3 equals three single-vector encodes bit for bit, and a repeat
    is bitwise identical."""
    torch, W, O, S = _setup()
    n_modes, eps, M, B = [16, 16, 12, 12], 1e-6, 40000, 3
    nm = int(np.prod(n_modes))
    x = S.clustered_coords(S.config_seed(6, 80), M, [0.3, -1.1, 2.0, 0.5], 0.05)
    c = S.complex_normal(S.config_seed(6, 81), M * B)
    xt = torch.from_numpy(x).cuda()
    plan = W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, eps).with_batch(B))
    fb = plan.encode_gpu(xt, torch.from_numpy(c).cuda()).cpu().numpy()
    fr = plan.execute_gpu(torch.from_numpy(c).cuda()).cpu().numpy()
    assert np.array_equal(fb.view(np.uint32), fr.view(np.uint32))
    single = W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, eps))
    single.set_points_gpu(xt)
    modes = np.random.default_rng(6).choice(nm, 600, replace=False)
    for v in range(B):
        cv = c[v * M:(v + 1) * M]
        fv = single.execute_gpu(torch.from_numpy(cv).cuda()).cpu().numpy()
        assert np.array_equal(fv.view(np.uint32), fb[v * nm:(v + 1) * nm].view(np.uint32))
        ref = O.direct_type1(x.astype(np.float64), cv.astype(np.complex128), n_modes,
                             modes=modes)
        err = rel_l2(fv[modes], ref)
        print(f"NUFFT_ND clustered 4d vector {v}: err={err:.3e}")
        assert err <= 16 * eps


@pytest.mark.parametrize("prec,eps", [("F64", 1e-8), ("F32", 1e-4)])
def test_nd_type3_matches_direct_oracle(prec, eps):
    torch, W, O, S = _setup()
    d, M, N = 4, 400, 300
    f64 = prec == "F64"
    dt = np.float64 if f64 else np.float32
    x = S.uniform_coords(S.config_seed(6, 90), M, d, lo=-2.0, hi=2.0, dtype=dt)
    s = S.uniform_coords(S.config_seed(6, 91), N, d, lo=-5.0, hi=5.0, dtype=dt)
    c = S.complex_normal(S.config_seed(6, 92), M, dtype=np.complex128 if f64 else np.complex64)
    plan = W.NufftType3Plan(d, [(-2, 2)] * d, [(-5, 5)] * d, eps=eps,
                            precision=getattr(W, prec))
    out = plan.encode_gpu(torch.from_numpy(x).cuda(), torch.from_numpy(c).cuda(),
                          torch.from_numpy(s).cuda()).cpu().numpy()
    ref = O.direct_type3(x.astype(np.float64), c.astype(np.complex128), s.astype(np.float64))
    err = rel_l2(out, ref)
    print(f"NUFFT_ND_T3 4d {prec} eps={eps}: err={err:.3e}")
    assert err <= 100 * eps
