"""ctypes wrapper of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module; the product (``wgpunufft_amd``) never does.

What pins each oracle function (the checks live in ``tests/test_oracle.py``):

* ``es_params`` / ``es_eval``: the reference's ES table
  (tests/cpu_foundations.rs:482-501) and the sigma=1.25 formula case (:546+).
* ``horner_table``: accuracy bound <= 0.1*eps against ``es_eval`` and the
  endpoint contract (src/kernel.rs:413-458).
* ``phihat`` / ``es_ft``: independent midpoint integration of the kernel
  (src/kernel.rs:460-488) and agreement between the two quadratures.
* ``fine_size``: 17->36, 1->14, sigma 1.25: 17->22 (tests/cpu_foundations.rs:503-544)
  and 256 for 128 modes; beyond those values the FFT-library length set is
  "parity unpinned" (wgpu-fft is not vendored).
* ``layout``: the reference layout tests (src/gpu_nd_spread.rs:2123-2275).
* ``bin_keys_f32``: exact binary128 fold, with an "ambiguous" counter for
  points whose f64 start differs from the exact one or whose exact position
  lies within 2^-44 * n of a support-start boundary, the window in which the
  reference's double-float fold (un-vendored) could round either way; the
  counter is zero on every seeded point set the tests use, and at such ties
  parity is unpinned.
* ``bin_sort``: the permutation contract (stable counting sort, ties by
  index), checked against numpy's stable argsort.
* ``direct``: the analytic cases of tests/cpu_foundations.rs:327-351 and a
  numpy closed form written independently in the tests.
* ``type3_axis``: the 2,3,5-smoothness and config values of src/type3.rs:708-797.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_i64p = ctypes.POINTER(ctypes.c_long)
_dp = ctypes.POINTER(ctypes.c_double)


def build():
    """Compile the oracle with gcc (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_es_params.argtypes = [ctypes.c_double, ctypes.c_double,
                                   ctypes.POINTER(ctypes.c_int), _dp]
        L.or_es_eval.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.or_es_eval.restype = ctypes.c_double
        L.or_phihat.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_long, _dp]
        L.or_es_ft.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.or_es_ft.restype = ctypes.c_double
        L.or_horner_table.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp]
        L.or_horner_eval.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double]
        L.or_horner_eval.restype = ctypes.c_double
        L.or_fine_size.argtypes = [ctypes.c_long, ctypes.c_double, ctypes.c_int]
        L.or_fine_size.restype = ctypes.c_long
        L.or_next_even_235.argtypes = [ctypes.c_long]
        L.or_next_even_235.restype = ctypes.c_long
        L.or_layout_query.argtypes = [ctypes.c_int, _i64p, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_long, _i64p]
        L.or_bin_keys_f32.argtypes = [ctypes.c_int, _i64p, ctypes.c_int, _i64p, ctypes.c_long,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p]
        L.or_bin_keys_f32.restype = ctypes.c_long
        L.or_bin_sort.argtypes = [ctypes.c_long, ctypes.c_void_p, ctypes.c_long,
                                  ctypes.c_void_p, ctypes.c_void_p]
        L.or_direct.argtypes = [ctypes.c_int, ctypes.c_int, _i64p, ctypes.c_int, ctypes.c_int,
                                ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long,
                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.or_type3_axis.argtypes = [ctypes.c_double] * 5 + [ctypes.c_int, _dp]
        _lib = L
    return _lib


def _longs(values):
    arr = (ctypes.c_long * max(1, len(values)))(*values)
    return arr


def es_params(eps, sigma=2.0):
    w = ctypes.c_int()
    b = ctypes.c_double()
    if lib().or_es_params(eps, sigma, ctypes.byref(w), ctypes.byref(b)) != 0:
        raise ValueError("unsupported tolerance or sigma")
    return w.value, b.value


def es_eval(width, beta, d):
    return lib().or_es_eval(width, beta, d)


def phihat(width, beta, n):
    out = np.zeros(n // 2 + 1, dtype=np.float64)
    if lib().or_phihat(width, beta, n, out.ctypes.data_as(_dp)) != 0:
        raise ValueError("odd or empty grid")
    return out


def es_ft(width, beta, freq):
    return lib().or_es_ft(width, beta, freq)


def horner_table(width, beta, eps):
    coef = np.zeros(19 * 16, dtype=np.float64)
    rows = lib().or_horner_table(width, beta, eps, coef.ctypes.data_as(_dp))
    return rows, coef[: rows * width].copy()


def horner_eval(width, rows, coef, d):
    c = np.ascontiguousarray(coef, dtype=np.float64)
    return lib().or_horner_eval(width, rows, c.ctypes.data_as(_dp), d)


def fine_size(modes, sigma, width):
    return lib().or_fine_size(modes, sigma, width)


def next_even_235(n):
    return lib().or_next_even_235(n)


def layout(fine, width, precision=0, point_count=0):
    """Restated NdSpreadLayout for `fine` (axis 0 first).  precision: 0 F32,
    1 F64, 2 Df64.  Returns a dict with run, workgroup_size, row_batch,
    block, blocks, geometries (bin sides) and the geometry chosen for
    `point_count` points."""
    out = (ctypes.c_long * 128)()
    d = len(fine)
    if lib().or_layout_query(d, _longs(list(fine)), width, precision, point_count, out) != 0:
        raise ValueError("layout failed")
    ng = out[3]
    return {
        "run": out[0],
        "workgroup_size": out[1],
        "row_batch": out[2],
        "chosen": out[4],
        "block": [out[8 + a] for a in range(d)],
        "blocks": [out[16 + a] for a in range(d)],
        "geometries": [[out[32 + 8 * i + a] for a in range(d)] for i in range(ng)],
    }


def bin_keys_f32(fine, width, side, coords):
    """Bin keys, starts and offsets of F32 points (coords: (M, d) float32)."""
    coords = np.ascontiguousarray(coords, dtype=np.float32)
    M, d = coords.shape
    keys = np.zeros(M, dtype=np.uint32)
    starts = np.zeros((M, d), dtype=np.int32)
    offsets = np.zeros((M, d), dtype=np.float64)
    amb = lib().or_bin_keys_f32(d, _longs(list(fine)), width, _longs(list(side)), M,
                                coords.ctypes.data, keys.ctypes.data, starts.ctypes.data,
                                offsets.ctypes.data)
    return keys, starts, offsets, amb


def bin_sort(keys, nbins):
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    M = keys.shape[0]
    sorted_idx = np.zeros(M, dtype=np.uint32)
    offsets = np.zeros(nbins + 1, dtype=np.uint32)
    lib().or_bin_sort(M, keys.ctypes.data, nbins, sorted_idx.ctypes.data, offsets.ctypes.data)
    return sorted_idx, offsets


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


def direct_type1(x, c, n_modes, isign=1, modeord=0, modes=None):
    """f[k] = sum_j c_j exp(i*isign*k.x_j) for the linear mode indices `modes`
    (all modes if None).  x: (M, d) float64; c: (M,) complex128."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    c = np.ascontiguousarray(c, dtype=np.complex128)
    M, d = x.shape
    total = int(np.prod(n_modes))
    sel = None if modes is None else np.ascontiguousarray(modes, dtype=np.int64)
    nout = total if sel is None else sel.shape[0]
    out = np.zeros(nout, dtype=np.complex128)
    lib().or_direct(1, d, _longs(list(n_modes)), modeord, isign, M, x.ctypes.data,
                    c.ctypes.data, nout, None if sel is None else sel.ctypes.data, None,
                    out.ctypes.data, _threads())
    return out


def direct_type2(x, f, n_modes, isign=1, modeord=0, points=None):
    """c[j] = sum_k f_k exp(i*isign*k.x_j) for the point indices `points`."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.complex128)
    M, d = x.shape
    sel = None if points is None else np.ascontiguousarray(points, dtype=np.int64)
    nout = M if sel is None else sel.shape[0]
    out = np.zeros(nout, dtype=np.complex128)
    lib().or_direct(2, d, _longs(list(n_modes)), modeord, isign, M, x.ctypes.data,
                    f.ctypes.data, nout, None if sel is None else sel.ctypes.data, None,
                    out.ctypes.data, _threads())
    return out


def direct_type3(x, c, s, isign=1, targets=None):
    """f[k] = sum_j c_j exp(i*isign*s_k.x_j)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    s = np.ascontiguousarray(s, dtype=np.float64)
    c = np.ascontiguousarray(c, dtype=np.complex128)
    M, d = x.shape
    sel = None if targets is None else np.ascontiguousarray(targets, dtype=np.int64)
    nout = s.shape[0] if sel is None else sel.shape[0]
    out = np.zeros(nout, dtype=np.complex128)
    lib().or_direct(3, d, None, 0, isign, M, x.ctypes.data, c.ctypes.data, nout,
                    None if sel is None else sel.ctypes.data, s.ctypes.data, out.ctypes.data,
                    _threads())
    return out


def type3_axis(src_lo, src_hi, tgt_lo, tgt_hi, sigma, width):
    out = (ctypes.c_double * 7)()
    lib().or_type3_axis(src_lo, src_hi, tgt_lo, tgt_hi, sigma, width, out)
    keys = ["X", "C", "S", "D", "nf", "h", "gamma"]
    r = {k: out[i] for i, k in enumerate(keys)}
    r["nf"] = int(r["nf"])
    return r
