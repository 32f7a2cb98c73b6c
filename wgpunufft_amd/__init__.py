"""MI355X-native NUFFT engine with the plan surface of wgpuNUFFT.

Python mirror of the reference's GPU plan API (src/config.rs ``NufftConfig``,
src/plan.rs ``NufftPlan``, src/type3.rs ``NufftType3Plan``) over the engine's
C ABI.  Device buffers are torch tensors on the ROCm device: coordinates are
point-major ``(M, dim)``, complex values are ``complex64`` (F32 plans) or
``complex128`` (F64 plans), modes axis-0 fastest (a tensor of shape
``(batch, N_{d-1}, ..., N_0)`` or flat), batches transform-major.

Every call enqueues work on the current torch stream and returns; like the
reference, executions of one plan must stay in stream order.
"""
import ctypes

from ._lib import NufftError, build, check, lib

__all__ = [
    "NufftConfig", "NufftPlan", "NufftType3Plan", "NufftError", "F32", "F64", "DF64",
    "encode_type2_many",
    "build", "lib", "reference_type1_f64", "reference_type2_f64", "reference_type3_f64",
    "STAGES",
]

F64 = 1
DF64 = 2
F32 = 3
STAGES = ["sort", "prepare", "spread", "fft", "deconv", "interp", "type3", "spread_kernel"]


def _torch():
    import torch
    return torch


def _stream_ptr():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


class NufftConfig:
    """Mathematical configuration of a type-1/2 plan (src/config.rs:64-230):
    modes per axis (axis 0 first), tolerance, sign, mode order, sigma,
    batch and precision.  Builder methods return a new config."""

    def __init__(self, n_modes, eps=1e-6):
        self.n_modes = [int(n) for n in n_modes]
        self.eps = float(eps)
        self.isign = 1
        self.mode_order = 0
        self.sigma = 2.0
        self.batch = 1
        self.precision = F32

    def _copy(self, **kw):
        c = NufftConfig(self.n_modes, self.eps)
        c.__dict__.update(self.__dict__)
        c.__dict__.update(kw)
        return c

    def with_isign(self, isign):
        return self._copy(isign=1 if isign >= 0 else -1)

    def with_mode_order(self, order):
        return self._copy(mode_order=int(order))

    with_modeord = with_mode_order

    def with_sigma(self, sigma):
        return self._copy(sigma=float(sigma))

    def with_batch(self, batch):
        return self._copy(batch=int(batch))

    def with_precision(self, precision):
        return self._copy(precision=int(precision))

    @property
    def dimensions(self):
        return len(self.n_modes)

    def mode_count(self):
        n = 1
        for m in self.n_modes:
            n *= m
        return n


SPREAD_METHODS = {"plain": 0, "mfma_f16": 1, "mfma_f16x3": 2, "mfma_bf16x3": 3}
POINT_ORDERS = {"index": 0, "coordinates": 1}


class _PlanBase:
    _handle = None

    def __del__(self):
        if self._handle is not None:
            try:
                lib().mnufft_destroy(self._handle)
            except Exception:
                pass
            self._handle = None

    def info(self):
        v = (ctypes.c_int64 * 16)()
        check(lib().mnufft_plan_info(self._handle, v, 16))
        return {
            "width": v[0],
            "fine": list(v[1:4]),
            "bin_side": list(v[4:7]),
            "nbins": v[7],
            "geometry": v[8],
            "block": list(v[9:12]),
            "run": v[12],
            "outer": list(v[13:16]),
        }

    def beta(self):
        return lib().mnufft_plan_beta(self._handle)

    def set_spread_method(self, method):
        """Type-1 spread arithmetic: "plain" (default), or the 16 x 16 x 8-tile
        matrix-core contraction "mfma_f16" (FP16 operands, ~1e-3),
        "mfma_f16x3" (FP16 hi/lo splits, ~1e-6) or "mfma_bf16x3" (bfloat16
        hi/lo splits, ~1e-5); FP32 3D plans only."""
        code = SPREAD_METHODS[method] if isinstance(method, str) else int(method)
        check(lib().mnufft_set_spread_method(self._handle, code))
        return self

    def set_point_order(self, order):
        """Type-1 summation order inside a bin: "index" (default, the
        reference's ascending original index) or "coordinates" (coordinate
        bits, then index: the output then does not depend on the input order
        of the points, bit for bit)."""
        code = POINT_ORDERS[order] if isinstance(order, str) else int(order)
        check(lib().mnufft_set_point_order(self._handle, code))
        return self

    def set_spread_tuning(self, light_limit=0, min_part=0, part_budget_bytes=0):
        """Work-plan knobs of the type-1 spread (0 keeps the current value):
        records above which a tile is split into parts, the smallest part,
        and the partial-sum budget per vector in bytes."""
        check(lib().mnufft_set_spread_tuning(self._handle, int(light_limit), int(min_part),
                                             int(part_budget_bytes)))
        return self

    def set_dense_bins(self, min_records):
        """Dense bins of the 3D matrix-core type-1 spread: a bin of at least
        min_records points (default 32) is summed once over its window and
        the windows are added to the tile sums in a fixed bin order; 0 turns
        them off.  Takes effect at the next set_points or execute."""
        check(lib().mnufft_set_dense_bins(self._handle, int(min_records)))
        return self

    def set_bin_tuning(self, rows_side0=0, plain_rows=True):
        """Bins and tiles of the type-1 spread: the axis-0 side of the row
        geometry's bins (0: automatic, else a power of two up to 64) and the
        plain spread on row tiles (3D 8x8x8, 2D 16x16 cells; True) or on the
        reference's output-stationary blocks (False).  Takes effect at the next
        set_points or execute."""
        check(lib().mnufft_set_bin_tuning(self._handle, int(rows_side0), 1 if plain_rows else 0))
        return self

    def set_fft_engine(self, engine):
        """FFT of the fine grid: "lines" (default; 2D and 3D grids of axes up
        to 4864 FP32 / 2432 FP64 cells as pruned line-transform passes) or
        "rocfft".  Type-3 plans apply it to their inner plan."""
        codes = {"lines": 0, "rocfft": 1}
        if engine not in codes:
            raise ValueError(f"unknown FFT engine {engine!r}")
        check(lib().mnufft_set_fft_engine(self._handle, codes[engine]))
        return self

    def fft_engine(self):
        """The engine the next FFT runs on: "lines" or "rocfft"."""
        e = ctypes.c_int32(0)
        check(lib().mnufft_fft_engine(self._handle, ctypes.byref(e)))
        return "lines" if e.value == 0 else "rocfft"

    def tuning(self):
        """The spread tuning in use: method, dense-bin threshold, axis-0 bin
        side (0 automatic), plain row tiles, light limit, smallest part."""
        v = (ctypes.c_int64 * 6)()
        check(lib().mnufft_plan_tuning(self._handle, v, 6))
        names = {k: n for n, k in SPREAD_METHODS.items()}
        return {"spread_method": names.get(v[0], v[0]), "dense_bins": v[1], "rows_side0": v[2],
                "plain_rows": bool(v[3]), "light_limit": v[4], "min_part": v[5]}

    def autotune(self, points, values, out=None, repeats=3, active_batch=None):
        """Type 1: times `repeats` encodes of these points and strengths for
        the plain spread and the matrix-core methods within the plan's
        tolerance (3D: dense bins on and off) and keeps the fastest, then
        tries the bins and tiles of the method kept (plain output-stationary
        blocks, coarser axis-0 bins, the heavy-tile split); returns the
        choice.  `out` receives the encodes."""
        if self.kind != 1:
            raise ValueError("autotune takes a type-1 plan")
        self._check_points(points)
        M = points.shape[0]
        active = active_batch or max(1, values.numel() // max(M, 1))
        if values.dtype != self.complex_dtype or not values.is_contiguous() or \
                values.numel() < active * M:
            raise ValueError("values must hold the active batch of strengths")
        out = self._out(active, out, self.config.mode_count())
        choice = (ctypes.c_int64 * 4)()
        check(lib().mnufft_autotune(self._handle, active, M, _ptr(points), _ptr(values),
                                    _ptr(out), int(repeats), choice, _stream_ptr()))
        self._points = points
        names = {v: k for k, v in SPREAD_METHODS.items()}
        t = self.tuning()
        return {"spread_method": names.get(choice[0], choice[0]), "dense_bins": choice[1],
                "ms_per_encode": choice[2] / 1e6, "candidates": choice[3],
                "rows_side0": t["rows_side0"], "plain_rows": t["plain_rows"],
                "light_limit": t["light_limit"]}

    def timing_history(self, on=True):
        """Sum stage times over every call until the next stage_times()."""
        check(lib().mnufft_timing_history(self._handle, 1 if on else 0))

    def stage_times(self):
        """Milliseconds per stage of the last call, or summed over the calls
        since timing_history(True) (None where a stage did not run); waits
        for the stage events."""
        ms = (ctypes.c_double * len(STAGES))()
        check(lib().mnufft_stage_times(self._handle, ms, len(STAGES)))
        return {name: (ms[i] if ms[i] >= 0 else None) for i, name in enumerate(STAGES)}


class NufftPlan(_PlanBase):
    """A type-1 or type-2 GPU plan (src/plan.rs:403-1229)."""

    def __init__(self, kind, config):
        self.kind = kind
        self.config = config
        n = config.dimensions
        nm = (ctypes.c_int64 * n)(*config.n_modes)
        h = ctypes.c_void_p()
        check(lib().mnufft_makeplan(kind, n, nm, config.isign, config.batch, config.eps,
                                    config.precision, config.mode_order, config.sigma,
                                    ctypes.byref(h)))
        self._handle = h
        self._points = None

    @classmethod
    def type1_gpu(cls, config):
        return cls(1, config)

    @classmethod
    def type2_gpu(cls, config):
        return cls(2, config)

    @property
    def real_dtype(self):
        torch = _torch()
        return torch.float32 if self.config.precision == F32 else torch.float64

    @property
    def complex_dtype(self):
        torch = _torch()
        return torch.complex64 if self.config.precision == F32 else torch.complex128

    def _check_points(self, points):
        d = self.config.dimensions
        if points.dim() != 2 or points.shape[1] != d:
            raise ValueError(f"points must have shape (M, {d})")
        if points.dtype != self.real_dtype or not points.is_contiguous() or not points.is_cuda:
            raise ValueError("points must be a contiguous device tensor of the plan's real type")

    def set_points_gpu(self, points):
        """Bin the points for later executions; the plan keeps the tensor
        (src/plan.rs:874-925)."""
        self._check_points(points)
        self._points = points
        check(lib().mnufft_set_points(self._handle, points.shape[0], _ptr(points), _stream_ptr()))

    def _out(self, active, out, count):
        torch = _torch()
        if out is None:
            out = torch.empty(active * count, dtype=self.complex_dtype, device="cuda")
        if out.dtype != self.complex_dtype or not out.is_contiguous() or out.numel() < active * count:
            raise ValueError("output tensor has the wrong type, layout or size")
        return out

    def execute_gpu(self, values, out=None, active_batch=None):
        """Type 1: strengths (active x M) -> modes; type 2: modes -> values at
        the points (src/plan.rs:931-977, 1180-1229)."""
        if self._points is None:
            raise NufftError(1, "points must be set before executing")
        M = self._points.shape[0]
        modes = self.config.mode_count()
        per = M if self.kind == 1 else modes
        active = active_batch or max(1, values.numel() // max(per, 1))
        if values.dtype != self.complex_dtype or not values.is_contiguous():
            raise ValueError("values must be a contiguous tensor of the plan's complex type")
        if values.numel() < active * per:
            raise ValueError("too few input values for the active batch")
        out = self._out(active, out, modes if self.kind == 1 else M)
        check(lib().mnufft_execute(self._handle, active, _ptr(values), _ptr(out), _stream_ptr()))
        return out

    def encode_gpu(self, points, values, out=None, active_batch=None):
        """set_points_gpu + execute_gpu (src/plan.rs:808-872, 1090-1108)."""
        self.set_points_gpu(points)
        return self.execute_gpu(values, out, active_batch)

    execute_type1_gpu = execute_gpu
    execute_type2_gpu = execute_gpu
    encode_type1_gpu = encode_gpu
    encode_type2_gpu = encode_gpu

    def bin_count(self, M):
        n = ctypes.c_int64()
        check(lib().mnufft_bin_count(self._handle, M, ctypes.byref(n)))
        return n.value

    def bin_sort(self, points):
        """The reference's stable bin permutation (sorted indices, bin
        offsets) of `points` for this type-1 plan."""
        torch = _torch()
        self._check_points(points)
        M = points.shape[0]
        nb = self.bin_count(M)
        sorted_idx = torch.empty(max(M, 1), dtype=torch.int32, device="cuda")
        offsets = torch.empty(nb + 1, dtype=torch.int32, device="cuda")
        n = ctypes.c_int64()
        check(lib().mnufft_bin_sort(self._handle, M, _ptr(points), _ptr(sorted_idx),
                                    _ptr(offsets), ctypes.byref(n), _stream_ptr()))
        return sorted_idx[:M], offsets


    def points_permutation(self):
        """The binning the plan's executes use after its last set_points:
        (point indices in record-slot order, bin offsets, bin side per axis).
        For the index point order this is the reference's stable bin
        permutation of the plan's geometry."""
        torch = _torch()
        if self._points is None:
            raise NufftError(1, "points must be set first")
        M = self._points.shape[0]
        nb = ctypes.c_int64()
        side = (ctypes.c_int64 * 8)()
        check(lib().mnufft_points_permutation(self._handle, None, None, ctypes.byref(nb), side,
                                              _stream_ptr()))
        sorted_idx = torch.empty(max(M, 1), dtype=torch.int32, device="cuda")
        offsets = torch.empty(nb.value + 1, dtype=torch.int32, device="cuda")
        check(lib().mnufft_points_permutation(self._handle, _ptr(sorted_idx), _ptr(offsets),
                                              ctypes.byref(nb), side, _stream_ptr()))
        return sorted_idx[:M], offsets, list(side)[: self.config.dimensions]


class NufftType3Plan(_PlanBase):
    """A type-3 GPU plan over source / target intervals (src/type3.rs)."""

    def __init__(self, dim, source_bounds, target_bounds, eps=1e-6, isign=1, batch=1,
                 precision=F64, sigma=2.0):
        self.dim = dim
        self.batch = batch
        self.precision = precision
        sb = (ctypes.c_double * (2 * dim))(*[float(v) for pair in source_bounds for v in pair])
        tb = (ctypes.c_double * (2 * dim))(*[float(v) for pair in target_bounds for v in pair])
        h = ctypes.c_void_p()
        check(lib().mnufft_makeplan_type3(dim, isign, batch, eps, precision, sigma, sb, tb,
                                          ctypes.byref(h)))
        self._handle = h

    def boxes(self, min_boxes=None):
        """Target boxes per axis in use; min_boxes (one value per axis)
        re-plans with at least that many.  Target intervals whose inner grid
        would not fit are cut into boxes automatically."""
        out = (ctypes.c_int32 * 8)()
        arg = None
        if min_boxes is not None:
            arg = (ctypes.c_int32 * 8)(*(list(min_boxes) + [1] * 8)[:8])
        check(lib().mnufft_type3_boxes(self._handle, arg, out))
        return list(out[:self.dim])

    def encode_gpu(self, sources, strengths, targets, out=None, active_batch=None):
        torch = _torch()
        M = sources.shape[0]
        N = targets.shape[0]
        active = active_batch or max(1, strengths.numel() // max(M, 1))
        cdt = torch.complex64 if self.precision == F32 else torch.complex128
        if out is None:
            out = torch.empty(active * N, dtype=cdt, device="cuda")
        check(lib().mnufft_execute_type3(self._handle, active, M, _ptr(sources), _ptr(strengths),
                                         N, _ptr(targets), _ptr(out), _stream_ptr()))
        return out


def encode_type2_many(plans, points, values, outs=None):
    """Type-2 encodes of many independent plans as one batch: plans[i]
    interpolates values[i] (its modes) at points[i]; one bin sort over all
    the points and one interpolation launch, each plan's FFT on its own
    grid.  F32 type-2 plans of one dimension (1D to 3D), one vector each;
    outputs equal each plan's own encode_gpu bit for bit.  The plans keep no
    points afterwards."""
    n = len(plans)
    if not (len(points) == len(values) == n) or n == 0:
        raise ValueError("plans, points and values must have the same nonzero length")
    if outs is None:
        outs = [None] * n
    res = []
    for p, x, v, o in zip(plans, points, values, outs):
        if p.kind != 2:
            raise ValueError("batched encodes take type-2 plans")
        p._check_points(x)
        modes = p.config.mode_count()
        if v.dtype != p.complex_dtype or not v.is_contiguous() or v.numel() < modes:
            raise ValueError("values must be a contiguous tensor of the plan's modes")
        res.append(p._out(1, o, x.shape[0]))
    vp = ctypes.c_void_p
    hs = (vp * n)(*[p._handle for p in plans])
    ms = (ctypes.c_int64 * n)(*[int(x.shape[0]) for x in points])
    xs = (vp * n)(*[_ptr(x) for x in points])
    vs = (vp * n)(*[_ptr(v) for v in values])
    os_ = (vp * n)(*[_ptr(o) for o in res])
    check(lib().mnufft_execute_type2_many(n, hs, ms, xs, vs, os_, _stream_ptr()))
    for p in plans:
        p._points = None
    return res


def encode_type1_many(plans, points, values, outs=None):
    """Type-1 encodes of many independent plans as one batch: plans[i]
    spreads values[i] (its strengths) at points[i] into its modes; one bin
    sort over all the points, one spread launch per class of plans (run
    length of the blocks, or row tiles of one kernel width), each plan's
    FFT, one deconvolution launch.  Type-1 plans of one rank (1D to 3D) and
    precision, one vector each, on the plain spread in index order; outputs
    equal each
    plan's own encode_gpu bit for bit.  The plans keep no points
    afterwards."""
    n = len(plans)
    if not (len(points) == len(values) == n) or n == 0:
        raise ValueError("plans, points and values must have the same nonzero length")
    if outs is None:
        outs = [None] * n
    res = []
    for p, x, v, o in zip(plans, points, values, outs):
        if p.kind != 1:
            raise ValueError("batched type-1 encodes take type-1 plans")
        p._check_points(x)
        if v.dtype != p.complex_dtype or not v.is_contiguous() or v.numel() < x.shape[0]:
            raise ValueError("values must be a contiguous tensor of the plan's strengths")
        res.append(p._out(1, o, p.config.mode_count()))
    vp = ctypes.c_void_p
    hs = (vp * n)(*[p._handle for p in plans])
    ms = (ctypes.c_int64 * n)(*[int(x.shape[0]) for x in points])
    xs = (vp * n)(*[_ptr(x) for x in points])
    vs = (vp * n)(*[_ptr(v) for v in values])
    os_ = (vp * n)(*[_ptr(o) for o in res])
    check(lib().mnufft_execute_type1_many(n, hs, ms, xs, vs, os_, _stream_ptr()))
    for p in plans:
        p._points = None
    return res


def _direct(kind, dim, n_modes, isign, mode_order, x, values, nout, sel, s):
    torch = _torch()
    out = torch.empty(nout, dtype=torch.complex128, device="cuda")
    nm = (ctypes.c_int64 * 8)(*(list(n_modes or []) + [1] * 8)[:8])
    M = x.shape[0] if x is not None else 0
    check(lib().mnufft_direct_f64(kind, dim, nm, isign, mode_order, M, _ptr(x), _ptr(values),
                                  nout, _ptr(sel), _ptr(s), _ptr(out), _stream_ptr()))
    return out


def reference_type1_f64(n_modes, x, c, isign=1, mode_order=0, modes=None):
    """FP64 direct type-1 sums on the GPU (src/direct.rs:9-62): x (M, d)
    float64, c complex128; `modes` selects linear mode indices."""
    nout = modes.numel() if modes is not None else _prod(n_modes)
    return _direct(1, len(n_modes), n_modes, isign, mode_order, x, c, nout, modes, None)


def reference_type2_f64(n_modes, x, f, isign=1, mode_order=0, points=None):
    """FP64 direct type-2 sums on the GPU (src/direct.rs:65-118)."""
    nout = points.numel() if points is not None else x.shape[0]
    return _direct(2, len(n_modes), n_modes, isign, mode_order, x, f, nout, points, None)


def reference_type3_f64(x, c, s, isign=1, targets=None):
    """FP64 direct type-3 sums on the GPU (src/type3.rs:609-673)."""
    nout = targets.numel() if targets is not None else s.shape[0]
    return _direct(3, x.shape[1], None, isign, 0, x, c, nout, targets, s)


def _prod(v):
    p = 1
    for a in v:
        p *= int(a)
    return p
