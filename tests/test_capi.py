"""CPU checks of the C ABI library: it loads, exports every symbol the
headers declare, and its host-side planning math equals the oracle's."""
import ctypes
import os
import re

import numpy as np

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:mnufft|wgpu_nufftf?)\w*)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import wgpunufft_amd as W
    L = W.lib()
    names = _declared("mnufft.h") + _declared("wgpu_nufft.h")
    assert len(names) >= 55
    for name in names:
        assert hasattr(L, name), name


def test_wgpu_nufft_opts_layout_is_320_bytes():
    class Opts(ctypes.Structure):
        _fields_ = [("backend", ctypes.c_int32), ("precision", ctypes.c_int32),
                    ("mode_order", ctypes.c_int32), ("threads", ctypes.c_int32),
                    ("sigma", ctypes.c_double), ("adapter_index", ctypes.c_int32),
                    ("adapter_name", ctypes.c_char * 256), ("adapter_pci_bus_id", ctypes.c_char * 32)]
    assert ctypes.sizeof(Opts) == 320
    import wgpunufft_amd as W
    o = Opts()
    ctypes.memset(ctypes.byref(o), 0xFF, ctypes.sizeof(o))
    W.lib().wgpu_nufft_default_opts(ctypes.byref(o))
    assert bytes(o) == bytes(320)
    assert W.lib().wgpu_nufft_version() == b"0.3.2"


def test_no_gpu_paths_fail_with_codes_not_crashes():
    import torch
    import wgpunufft_amd as W
    if torch.cuda.is_available():
        return
    h = ctypes.c_void_p()
    nm = (ctypes.c_int64 * 2)(8, 8)
    rc = W.lib().wgpu_nufftf_makeplan(1, 2, nm, 1, ctypes.c_int64(1), ctypes.c_double(1e-6),
                                      ctypes.byref(h), None)
    assert rc == 3 and not h.value
    assert b"HIP" in W.lib().wgpu_nufft_last_error()


def test_product_host_math_equals_oracle():
    import wgpunufft_amd as W
    L = W.lib()
    for eps in [1e-1, 1e-3, 1e-6, 1e-9, 1e-12]:
        for sigma in [2.0, 1.25, 1.5]:
            w = ctypes.c_int32()
            b = ctypes.c_double()
            rc = L.mnufft_kernel_params(eps, sigma, ctypes.byref(w), ctypes.byref(b))
            try:
                ow, ob = O.es_params(eps, sigma)
            except ValueError:
                assert rc == 2  # width above 16 is rejected by both
                continue
            assert rc == 0
            assert (w.value, b.value) == (ow, ob)
            for modes in [1, 17, 100, 128, 1000]:
                assert L.mnufft_fine_grid_size(modes, sigma, w.value) == O.fine_size(modes, sigma, ow)
    rng = np.random.default_rng(1)
    shapes = [[256, 256, 256], [512, 512, 512], [200, 200, 200], [36, 30, 28], [256, 256],
              [2048, 2048], [14, 14, 14], [4096], [130, 62], [64, 64, 64, 64], [32, 28, 24, 20],
              [16, 16, 16, 16, 16], [20, 18, 16, 14, 12, 12], [12] * 7, [10] * 8]
    for shape in shapes:
        for eps in [1e-3, 1e-6, 1e-10]:
            w, _ = O.es_params(eps)
            for prec in [0, 1]:
                for M in [0, 1000, 100_000, 10_000_000, int(rng.integers(1, 10**8))]:
                    ref = O.layout(shape, w, prec, M)
                    out = (ctypes.c_int64 * 128)()
                    fine = (ctypes.c_int64 * len(shape))(*shape)
                    assert L.mnufft_layout_query(len(shape), fine, w, 3 if prec == 0 else 1, M,
                                                 out) == 0
                    assert out[0] == ref["run"] and out[1] == ref["workgroup_size"]
                    assert out[2] == ref["row_batch"] and out[4] == ref["chosen"]
                    d = len(shape)
                    assert [out[8 + a] for a in range(d)] == ref["block"]
                    assert [out[16 + a] for a in range(d)] == ref["blocks"]
                    got_geoms = [[out[32 + 8 * g + a] for a in range(d)] for g in range(out[3])]
                    assert got_geoms == ref["geometries"]


def _tolerance_check(n_modes, eps, sigma, precision):
    import wgpunufft_amd as W
    nm = (ctypes.c_int64 * len(n_modes))(*n_modes)
    return W.lib().mnufft_tolerance_check(len(n_modes), nm, eps, sigma, precision)


def test_f32_plans_reject_tolerances_their_rounding_misses_by_far():
    """Known answers of the reference's own test
    (src/config.rs:324-341, f32_plans_reject_tolerances_their_rounding_misses_by_far):
    at sigma 1.25 and eps 1e-6 F32 is rejected in 2D, 3D and 4D while F64 and
    Df64 are accepted; looser tolerances and sigma 2 stay reachable."""
    F64, DF64, F32 = 1, 2, 3
    for n_modes in ([256, 256], [64] * 3, [16] * 4):
        assert _tolerance_check(n_modes, 1e-6, 1.25, F32) == 2
        for prec in (F64, DF64):
            assert _tolerance_check(n_modes, 1e-6, 1.25, prec) == 0
    assert _tolerance_check([16] * 4, 1e-4, 1.25, F32) == 0
    assert _tolerance_check([6, 5, 4, 4], 1e-3, 1.25, F32) == 0
    for n_modes in ([1 << 20], [1024, 1024], [128] * 3, [16] * 6):
        assert _tolerance_check(n_modes, 1e-9, 2.0, F32) == 0


def test_tolerance_check_estimate_is_the_rms_gain():
    """The F32 estimate is eps_f32/2 times the per-axis RMS of
    phi_hat(0)/phi_hat(k) over the modes, rejected above 10 eps and 1e-5:
    an independent evaluation of that product (oracle phi_hat) brackets the
    rejection threshold -- the check accepts just above the estimate/10 and
    rejects just below it."""
    checked = 0
    for n_modes, sigma in (([256, 256], 1.25), ([512, 128], 1.25), ([1000, 1000], 1.25),
                           ([64] * 3, 1.25)):
        w, beta = O.es_params(1e-6, sigma)
        est = float(np.finfo(np.float32).eps) / 2.0
        for n in n_modes:
            coeffs = np.asarray(O.phihat(w, beta, O.fine_size(n, sigma, w)))
            k = np.abs(np.arange(n) - n // 2)
            est *= np.sqrt(np.mean((coeffs[0] / np.abs(coeffs[k])) ** 2))
        if est <= 1e-5:
            continue
        # eps = est/10 * (1 +- 1e-3): the ES kernel is the same at 1e-6 only
        # when eps keeps the width, so compare through a fixed-width window
        lo, hi = est / 10.0 * 0.999, est / 10.0 * 1.001
        if O.es_params(lo, sigma)[0] != w or O.es_params(hi, sigma)[0] != w:
            continue
        assert _tolerance_check(n_modes, hi, sigma, 3) == 0
        assert _tolerance_check(n_modes, lo, sigma, 3) == 2
        checked += 1
    assert checked >= 3


def test_oversize_point_counts_are_plan_errors_before_any_allocation():
    """Point counts beyond the 32-bit index space are rejected as plan
    errors (code 2) by the one-call host ABI before a plan or a device
    buffer exists (no arrays of that size are passed: the check comes
    first)."""
    import wgpunufft_amd as W
    L = W.lib()
    x = (ctypes.c_double * 4)()
    c = (ctypes.c_double * 8)()
    f = (ctypes.c_double * 64)()
    big = ctypes.c_int64(1 << 32)
    assert L.wgpu_nufft1d1(big, x, c, 1, ctypes.c_double(1e-6), ctypes.c_int64(8), f, None) == 2
    assert b"32-bit" in L.wgpu_nufft_last_error()
    xf = (ctypes.c_float * 4)()
    cf = (ctypes.c_float * 8)()
    ff = (ctypes.c_float * 64)()
    assert L.wgpu_nufftf1d3(ctypes.c_int64(4), xf, cf, 1, ctypes.c_double(1e-6), big, xf, ff,
                            None) == 2
