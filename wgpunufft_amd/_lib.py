"""Loader of libmnufft.so, the engine's C ABI (include/mnufft.h,
include/wgpu_nufft.h).

The library is built in-tree (``python -c "import __graft_entry__ as g;
g.build()"`` or ``make -C wgpunufft_amd/csrc``).  There is no fallback: a
missing library raises, so nothing silently runs anywhere but the HIP
kernels.
"""
import ctypes
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libmnufft.so")
CSRC = os.path.join(_HERE, "csrc")

_lib = None

_c_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p


def build(jobs=8):
    """Compile every HIP/C++ source of the engine for gfx950 (make)."""
    subprocess.run(["make", "-s", "-C", CSRC, f"-j{jobs}"], check=True)


# (name, restype, argtypes) of every exported entry point.
SIGNATURES = [
    ("mnufft_last_error", ctypes.c_char_p, []),
    ("mnufft_makeplan", ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, _c_i64p, ctypes.c_int32,
                                       ctypes.c_int64, ctypes.c_double, ctypes.c_int32,
                                       ctypes.c_int32, ctypes.c_double, ctypes.POINTER(_vp)]),
    ("mnufft_makeplan_type3", ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int64,
                                             ctypes.c_double, ctypes.c_int32, ctypes.c_double,
                                             ctypes.POINTER(ctypes.c_double),
                                             ctypes.POINTER(ctypes.c_double),
                                             ctypes.POINTER(_vp)]),
    ("mnufft_destroy", None, [_vp]),
    ("mnufft_set_points", ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp]),
    ("mnufft_execute", ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp]),
    ("mnufft_encode", ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int64, _vp, _vp, _vp, _vp]),
    ("mnufft_execute_type2_many", ctypes.c_int, [ctypes.c_int64, ctypes.POINTER(_vp),
                                                 ctypes.POINTER(ctypes.c_int64),
                                                 ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                                 ctypes.POINTER(_vp), _vp]),
    ("mnufft_execute_type1_many", ctypes.c_int, [ctypes.c_int64, ctypes.POINTER(_vp),
                                                 ctypes.POINTER(ctypes.c_int64),
                                                 ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                                 ctypes.POINTER(_vp), _vp]),
    ("mnufft_execute_type3", ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int64, _vp, _vp,
                                            ctypes.c_int64, _vp, _vp, _vp]),
    ("mnufft_bin_sort", ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp, _c_i64p, _vp]),
    ("mnufft_bin_sort_noalloc", ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp, _c_i64p,
                                               _vp]),
    ("mnufft_bin_count", ctypes.c_int, [_vp, ctypes.c_int64, _c_i64p]),
    ("mnufft_type3_boxes", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32),
                                          ctypes.POINTER(ctypes.c_int32)]),
    ("mnufft_points_permutation", ctypes.c_int, [_vp, _vp, _vp, _c_i64p, _c_i64p, _vp]),
    ("mnufft_plan_info", ctypes.c_int, [_vp, _c_i64p, ctypes.c_int32]),
    ("mnufft_plan_beta", ctypes.c_double, [_vp]),
    ("mnufft_stage_times", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double), ctypes.c_int32]),
    ("mnufft_timing_history", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("mnufft_set_spread_method", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("mnufft_set_point_order", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("mnufft_set_spread_tuning", ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int64,
                                                ctypes.c_int64]),
    ("mnufft_set_dense_bins", ctypes.c_int, [_vp, ctypes.c_int64]),
    ("mnufft_set_bin_tuning", ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int32]),
    ("mnufft_plan_tuning", ctypes.c_int, [_vp, _c_i64p, ctypes.c_int32]),
    ("mnufft_set_fft_engine", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("mnufft_fft_engine", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32)]),
    ("mnufft_autotune", ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int64, _vp, _vp, _vp,
                                       ctypes.c_int32, ctypes.POINTER(ctypes.c_int64), _vp]),
    ("mnufft_kernel_params", ctypes.c_int, [ctypes.c_double, ctypes.c_double,
                                            ctypes.POINTER(ctypes.c_int32),
                                            ctypes.POINTER(ctypes.c_double)]),
    ("mnufft_fine_grid_size", ctypes.c_int64, [ctypes.c_int64, ctypes.c_double, ctypes.c_int32]),
    ("mnufft_tolerance_check", ctypes.c_int, [ctypes.c_int32, _c_i64p, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_int32]),
    ("mnufft_layout_query", ctypes.c_int, [ctypes.c_int32, _c_i64p, ctypes.c_int32,
                                           ctypes.c_int32, ctypes.c_int64, _c_i64p]),
    ("mnufft_direct_f64", ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, _c_i64p, ctypes.c_int32,
                                         ctypes.c_int32, ctypes.c_int64, _vp, _vp, ctypes.c_int64,
                                         _vp, _vp, _vp, _vp]),
]

# Host-array C ABI of the reference (include/wgpu_nufft.h).
WGPU_NUFFT_SYMBOLS = [
    "wgpu_nufft_default_opts", "wgpu_nufft_list_adapters", "wgpu_nufft_last_error",
    "wgpu_nufft_version", "wgpu_nufft_gpu_name", "wgpu_nufft_shutdown",
    "wgpu_nufft_makeplan", "wgpu_nufft_setpts", "wgpu_nufft_setpts_nd", "wgpu_nufft_execute",
    "wgpu_nufft_destroy", "wgpu_nufft_plan_backend", "wgpu_nufft_plan_precision",
    "wgpu_nufft1d1", "wgpu_nufft1d2", "wgpu_nufft1d3", "wgpu_nufft2d1", "wgpu_nufft2d2",
    "wgpu_nufft2d3", "wgpu_nufft3d1", "wgpu_nufft3d2", "wgpu_nufft3d3",
    "wgpu_nufftf_makeplan", "wgpu_nufftf_setpts", "wgpu_nufftf_setpts_nd", "wgpu_nufftf_execute",
    "wgpu_nufftf_destroy", "wgpu_nufftf_plan_backend", "wgpu_nufftf_plan_precision",
    "wgpu_nufftf1d1", "wgpu_nufftf1d2", "wgpu_nufftf1d3", "wgpu_nufftf2d1", "wgpu_nufftf2d2",
    "wgpu_nufftf2d3", "wgpu_nufftf3d1", "wgpu_nufftf3d2", "wgpu_nufftf3d3",
]


def lib():
    """The loaded engine library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP engine first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L.wgpu_nufft_version.restype = ctypes.c_char_p
        L.wgpu_nufft_last_error.restype = ctypes.c_char_p
        L.wgpu_nufft_gpu_name.restype = ctypes.c_char_p
        _lib = L
    return _lib


class NufftError(RuntimeError):
    """A failed engine call: `code` is the C ABI return code."""

    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code


def check(rc):
    if rc != 0:
        msg = lib().mnufft_last_error()
        raise NufftError(rc, msg.decode() if msg else "")
    return rc
