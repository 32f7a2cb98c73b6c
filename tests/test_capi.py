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
              [2048, 2048], [14, 14, 14], [4096], [130, 62]]
    for shape in shapes:
        for eps in [1e-3, 1e-6, 1e-10]:
            w, _ = O.es_params(eps)
            for prec in [0, 1]:
                for M in [0, 1000, 100_000, 10_000_000, int(rng.integers(1, 10**8))]:
                    ref = O.layout(shape, w, prec, M)
                    out = (ctypes.c_int64 * 64)()
                    fine = (ctypes.c_int64 * 3)(*shape)
                    assert L.mnufft_layout_query(len(shape), fine, w, 3 if prec == 0 else 1, M,
                                                 out) == 0
                    assert out[0] == ref["run"] and out[1] == ref["workgroup_size"]
                    assert out[2] == ref["row_batch"] and out[4] == ref["chosen"]
                    d = len(shape)
                    got_geoms = [[out[16 + 3 * g + a] for a in range(d)] for g in range(out[3])]
                    assert got_geoms == ref["geometries"]
