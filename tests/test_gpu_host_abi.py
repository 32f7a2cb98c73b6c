"""The reference's host-array C ABI on the GPU, as its own C-ABI tests drive
it (capi/tests/c_api.rs:216-420, 691-697): the one-call transforms 1d1, 2d2,
3d1 and 2d3 in FP64, Df64 (native FP64 here) and F32; a reused 2D type-1
plan with a batch of three, FFT mode order, points given per axis and
point-major, coordinates in [-3 pi, 3 pi]; the single-precision calls.
Outputs are checked against FP64 direct sums (numpy, written here) to the
reference's tolerances."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F64, DF64, F32 = 1, 2, 3
BACKEND_GPU = 1
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_i64p = ctypes.POINTER(ctypes.c_int64)


class Opts(ctypes.Structure):
    _fields_ = [("backend", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("mode_order", ctypes.c_int32), ("threads", ctypes.c_int32),
                ("sigma", ctypes.c_double), ("adapter_index", ctypes.c_int32),
                ("adapter_name", ctypes.c_char * 256),
                ("adapter_pci_bus_id", ctypes.c_char * 32)]


def _lib():
    from wgpunufft_amd import _lib as L
    lib = L.lib()
    lib.wgpu_nufft_last_error.restype = ctypes.c_char_p
    return lib


def _opts(lib, precision, mode_order=0):
    o = Opts()
    lib.wgpu_nufft_default_opts(ctypes.byref(o))
    o.backend, o.precision, o.mode_order = BACKEND_GPU, precision, mode_order
    return o


def _check(lib, rc):
    assert rc == 0, lib.wgpu_nufft_last_error().decode()


class Rng:
    """An LCG stream of uniform values in [-scale, scale)."""

    def __init__(self, seed):
        self.s = seed

    def values(self, n, scale):
        out = np.empty(n)
        for i in range(n):
            self.s = (self.s * 6364136223846793005 + 1442695040888963407) % (1 << 64)
            out[i] = (2.0 * ((self.s >> 11) / float(1 << 53)) - 1.0) * scale
        return out


def _freqs(n_modes):
    """Centered frequencies of every mode, axis 0 fastest: (modes, d)."""
    grids = np.meshgrid(*[np.arange(n) - n // 2 for n in n_modes], indexing="ij")
    return np.stack([g.reshape(-1, order="F") for g in grids], axis=1).astype(np.float64)


def _cplx(a):
    return a[0::2] + 1j * a[1::2]


def _direct1(axes, c, isign, n_modes):
    k = _freqs(n_modes)
    x = np.stack(axes, axis=1)
    return np.exp(1j * isign * (k @ x.T)) @ _cplx(c)


def _direct2(axes, f, isign, n_modes):
    k = _freqs(n_modes)
    x = np.stack(axes, axis=1)
    return np.exp(1j * isign * (x @ k.T)) @ _cplx(f)


def _direct3(axes, c, isign, targets):
    s = np.stack(targets, axis=1)
    x = np.stack(axes, axis=1)
    return np.exp(1j * isign * (s @ x.T)) @ _cplx(c)


def _err(actual, expected):
    return float(np.linalg.norm(_cplx(actual) - expected) / np.linalg.norm(expected))


def _p(a, single):
    return a.ctypes.data_as(_fp if single else _dp)


@pytest.mark.parametrize("precision,eps,tol", [(0, 1e-9, 1e-8), (DF64, 1e-9, 1e-8),
                                               (F32, 1e-5, 1e-4)])
def test_one_call_suite(precision, eps, tol):
    lib = _lib()
    single = precision == F32
    dt = np.float32 if single else np.float64
    pre = "wgpu_nufftf" if single else "wgpu_nufft"
    rng = Rng(0x5EED + BACKEND_GPU * 7 + precision)
    m = 60
    x, y, z = (rng.values(m, np.pi).astype(dt) for _ in range(3))
    c = rng.values(2 * m, 1.0).astype(dt)
    o = _opts(lib, precision)
    wide = [a.astype(np.float64) for a in (x, y, z)]
    cw = c.astype(np.float64)
    I64, D = ctypes.c_int64, ctypes.c_double
    f = np.zeros(2 * 12, dt)
    _check(lib, getattr(lib, pre + "1d1")(I64(m), _p(x, single), _p(c, single), -1, D(eps),
                                          I64(12), _p(f, single), ctypes.byref(o)))
    e = _err(f.astype(np.float64), _direct1(wide[:1], cw, -1, [12]))
    assert e < tol, f"1d1 {e:.3e}"
    modes = rng.values(2 * 8 * 6, 1.0).astype(dt)
    vals = np.zeros(2 * m, dt)
    _check(lib, getattr(lib, pre + "2d2")(I64(m), _p(x, single), _p(y, single), _p(vals, single),
                                          -1, D(eps), I64(8), I64(6), _p(modes, single),
                                          ctypes.byref(o)))
    e = _err(vals.astype(np.float64), _direct2(wide[:2], modes.astype(np.float64), -1, [8, 6]))
    assert e < tol, f"2d2 {e:.3e}"
    f = np.zeros(2 * 6 * 5 * 4, dt)
    _check(lib, getattr(lib, pre + "3d1")(I64(m), _p(x, single), _p(y, single), _p(z, single),
                                          _p(c, single), 1, D(eps), I64(6), I64(5), I64(4),
                                          _p(f, single), ctypes.byref(o)))
    e = _err(f.astype(np.float64), _direct1(wide, cw, 1, [6, 5, 4]))
    assert e < tol, f"3d1 {e:.3e}"
    n = 40
    s = rng.values(n, 7.0).astype(dt)
    t = rng.values(n, 5.0).astype(dt)
    f = np.zeros(2 * n, dt)
    _check(lib, getattr(lib, pre + "2d3")(I64(m), _p(x, single), _p(y, single), _p(c, single), 1,
                                          D(eps), I64(n), _p(s, single), _p(t, single),
                                          _p(f, single), ctypes.byref(o)))
    e = _err(f.astype(np.float64),
             _direct3(wide[:2], cw, 1, [s.astype(np.float64), t.astype(np.float64)]))
    assert e < tol, f"2d3 {e:.3e}"
    print(f"one-call suite precision {precision}: ok", flush=True)


def test_plan_suite_batch_fft_order_and_two_point_sets():
    lib = _lib()
    rng = Rng(0xBEEF + BACKEND_GPU)
    n_modes, ntrans = [10, 7], 3
    nm = 70
    o = _opts(lib, 0, mode_order=1)
    plan = ctypes.c_void_p()
    _check(lib, lib.wgpu_nufft_makeplan(1, 2, (ctypes.c_int64 * 2)(10, 7), 1,
                                        ctypes.c_int64(ntrans), ctypes.c_double(1e-9),
                                        ctypes.byref(plan), ctypes.byref(o)))
    assert lib.wgpu_nufft_plan_backend(plan) == BACKEND_GPU
    try:
        for rnd in range(2):
            m = 30 + 11 * rnd
            x = rng.values(m, 3.0 * np.pi)
            y = rng.values(m, np.pi)
            c = rng.values(2 * m * ntrans, 1.0)
            if rnd == 0:
                _check(lib, lib.wgpu_nufft_setpts(plan, ctypes.c_int64(m), _p(x, False),
                                                  _p(y, False), None, ctypes.c_int64(0), None,
                                                  None, None))
            else:
                pts = np.stack([x, y], axis=1).reshape(-1).copy()
                _check(lib, lib.wgpu_nufft_setpts_nd(plan, ctypes.c_int64(m), _p(pts, False),
                                                     ctypes.c_int64(0), None))
            f = np.zeros(2 * nm * ntrans)
            _check(lib, lib.wgpu_nufft_execute(plan, _p(c, False), _p(f, False)))
            k = _freqs(n_modes)
            fft_index = (np.mod(k[:, 0], 10) + 10 * np.mod(k[:, 1], 7)).astype(np.int64)
            for tr in range(ntrans):
                centered = _direct1([x, y], c[2 * m * tr:2 * m * (tr + 1)], 1, n_modes)
                expected = np.zeros(nm, complex)
                expected[fft_index] = centered
                e = _err(f[2 * nm * tr:2 * nm * (tr + 1)], expected)
                assert e < 1e-8, f"round {rnd} transform {tr}: {e:.3e}"
    finally:
        lib.wgpu_nufft_destroy(plan)


def test_single_precision_plan_reports_f32():
    lib = _lib()
    o = _opts(lib, 0)
    plan = ctypes.c_void_p()
    _check(lib, lib.wgpu_nufftf_makeplan(2, 1, (ctypes.c_int64 * 1)(16), 1, ctypes.c_int64(1),
                                         ctypes.c_double(1e-5), ctypes.byref(plan),
                                         ctypes.byref(o)))
    try:
        assert lib.wgpu_nufftf_plan_precision(plan) == F32
    finally:
        lib.wgpu_nufftf_destroy(plan)


def test_errors_are_reported_with_messages():
    """capi/tests/c_api.rs:433-506 on the GPU backend: codes and the key
    phrase of each message."""
    lib = _lib()
    o = _opts(lib, 0)
    plan = ctypes.c_void_p()
    I64, D = ctypes.c_int64, ctypes.c_double
    msg = lambda: lib.wgpu_nufft_last_error().decode()
    assert lib.wgpu_nufft_makeplan(4, 1, (I64 * 1)(8), 1, I64(1), D(1e-6), ctypes.byref(plan),
                                   ctypes.byref(o)) == 1
    assert "type" in msg() and not plan.value
    assert lib.wgpu_nufft_makeplan(1, 1, (I64 * 1)(0), 1, I64(1), D(1e-6), ctypes.byref(plan),
                                   ctypes.byref(o)) == 1
    assert lib.wgpu_nufft_makeplan(1, 1, (I64 * 1)(8), 1, I64(1), D(1e-20), ctypes.byref(plan),
                                   ctypes.byref(o)) == 2
    _check(lib, lib.wgpu_nufft_makeplan(1, 1, (I64 * 1)(8), 1, I64(1), D(1e-6),
                                        ctypes.byref(plan), ctypes.byref(o)))
    try:
        f = np.zeros(16)
        c = np.zeros(2)
        assert lib.wgpu_nufft_execute(plan, _p(c, False), _p(f, False)) == 1
        assert "set the points" in msg(), msg()
        far = np.array([10.0])
        assert lib.wgpu_nufft_setpts(plan, I64(1), _p(far, False), None, None, I64(0), None,
                                     None, None) == 1
        assert "3*pi" in msg(), msg()
        assert lib.wgpu_nufft_setpts(plan, I64(1), None, None, None, I64(0), None, None,
                                     None) == 1
        # the single-precision functions refuse a double plan
        assert lib.wgpu_nufftf_execute(plan, None, None) == 1
        assert "array type" in msg(), msg()
    finally:
        lib.wgpu_nufft_destroy(plan)
    lib.wgpu_nufft_destroy(None)


def test_one_call_transforms_reuse_kept_plans():
    """Repeated one-call transforms reuse the plans they keep (the
    reference keeps four, capi/src/cache.rs:10-51, and its GPU test repeats
    the one-call suite, capi/tests/c_api.rs:698-702): new points on a kept
    plan give the direct sums, more than four specs evict the oldest,
    wgpu_nufft_shutdown drops them and later calls plan afresh."""
    import time
    lib = _lib()
    lib.wgpu_nufft_shutdown()
    o = _opts(lib, 0)
    I64, D = ctypes.c_int64, ctypes.c_double
    rng = Rng(0xCAC4E)

    def call(m, nm):
        x, y, z = (rng.values(m, np.pi) for _ in range(3))
        c = rng.values(2 * m, 1.0)
        f = np.zeros(2 * int(np.prod(nm)))
        _check(lib, lib.wgpu_nufft3d1(I64(m), _p(x, False), _p(y, False), _p(z, False),
                                      _p(c, False), 1, D(1e-9), I64(nm[0]), I64(nm[1]),
                                      I64(nm[2]), _p(f, False), ctypes.byref(o)))
        e = _err(f, _direct1([x, y, z], c, 1, nm))
        assert e < 1e-8, f"{nm} m={m}: {e:.3e}"

    t0 = time.perf_counter()
    call(50, [6, 5, 4])
    first = time.perf_counter() - t0
    t0 = time.perf_counter()
    for m in (40, 70, 55):  # the kept plan, new point sets of other sizes
        call(m, [6, 5, 4])
    again = (time.perf_counter() - t0) / 3
    print(f"one-call 3d1: first {first * 1e3:.1f} ms, kept plan {again * 1e3:.1f} ms", flush=True)
    for k in range(5):  # five specs: the first one is evicted on the way
        call(30, [4 + k, 5, 4])
    call(45, [6, 5, 4])
    lib.wgpu_nufft_shutdown()
    call(35, [6, 5, 4])
    lib.wgpu_nufft_shutdown()
