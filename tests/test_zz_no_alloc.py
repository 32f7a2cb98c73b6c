"""The bin sort that allocates nothing (mnufft_bin_sort_noalloc), called
through the C ABI: its permutation and bin offsets equal mnufft_bin_sort's
bit for bit on two calls in a row (the second with fewer points, so stale
contents of the shared static scratch would show), for 3D and 2D plans and
clustered points; sizes beyond its stated limit are refused with an error
before any GPU work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_POINTS = 2097152  # MNUFFT_NOALLOC_MAX_POINTS in include/mnufft.h


def _setup():
    import torch
    import wgpunufft_amd as W
    from wgpunufft_amd import synthetic as S
    return torch, W, S


def _noalloc(W, torch, plan, xt):
    M = xt.shape[0]
    nb = plan.bin_count(M)
    srt = torch.full((max(M, 1),), -1, dtype=torch.int32, device="cuda")
    off = torch.full((nb + 1,), -1, dtype=torch.int32, device="cuda")
    n = ctypes.c_int64(-1)
    rc = W.lib().mnufft_bin_sort_noalloc(plan._handle, M, ctypes.c_void_p(xt.data_ptr()),
                                         ctypes.c_void_p(srt.data_ptr()),
                                         ctypes.c_void_p(off.data_ptr()), ctypes.byref(n),
                                         None)
    assert rc == 0, W.lib().mnufft_last_error()
    assert n.value == nb
    return srt[:M], off


def _points(S, kind, seed, M, d):
    if kind == "uniform":
        return S.uniform_coords(seed, M, d)
    x = S.clustered_coords(seed, M, (0.3, -1.1, 2.0)[:d], 0.05)
    x[: M // 2] = S.uniform_coords(seed + 1, M // 2, d)
    return x


@pytest.mark.parametrize("n_modes,kind,sizes", [
    ([64, 64, 64], "uniform", (MAX_POINTS, 300_000)),
    ([128, 128, 128], "clustered", (1_500_000, 200_000)),
    ([200, 180], "clustered", (1_000_000, 50_000)),
])
def test_noalloc_sort_equals_bin_sort_on_two_calls(n_modes, kind, sizes):
    torch, W, S = _setup()
    d = len(n_modes)
    plan = W.NufftPlan.type1_gpu(W.NufftConfig(n_modes, 1e-6))
    for k, M in enumerate(sizes):
        print(f"CASE noalloc {n_modes} {kind} call {k} M={M}", flush=True)
        x = _points(S, kind, S.config_seed(2, 1700 + 10 * k + d), M, d)
        xt = torch.from_numpy(x).cuda()
        ref_srt, ref_off = plan.bin_sort(xt)
        srt, off = _noalloc(W, torch, plan, xt)
        torch.cuda.synchronize()
        assert torch.equal(srt, ref_srt), f"permutation differs on call {k}"
        assert torch.equal(off, ref_off), f"offsets differ on call {k}"
    # the plan's own (allocating) sort is undisturbed afterwards
    again, _ = plan.bin_sort(xt)
    torch.cuda.synchronize()
    assert torch.equal(again, ref_srt)


def test_noalloc_sort_refuses_sizes_beyond_its_limit():
    torch, W, S = _setup()
    plan = W.NufftPlan.type1_gpu(W.NufftConfig([64, 64, 64], 1e-6))
    x = torch.zeros(16, 3, device="cuda")
    buf = torch.zeros(16, dtype=torch.int32, device="cuda")
    vp = ctypes.c_void_p
    # the size check comes before any GPU work: the small buffers are never
    # touched
    rc = W.lib().mnufft_bin_sort_noalloc(plan._handle, MAX_POINTS + 1, vp(x.data_ptr()),
                                         vp(buf.data_ptr()), vp(buf.data_ptr()), None, None)
    assert rc == 1
    assert b"at most" in W.lib().mnufft_last_error()
    rc = W.lib().mnufft_bin_sort_noalloc(plan._handle, 16, vp(x.data_ptr()),
                                         vp(buf.data_ptr()), None, None, None)
    assert rc == 1
