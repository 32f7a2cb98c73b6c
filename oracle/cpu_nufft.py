"""ctypes wrapper of the CPU port of CpuNufftPlan (cpu_nufft.c) -- the CPU
baseline of bench.py; test infrastructure, never loaded by the product."""
import ctypes
import time

import numpy as np

from . import oracle as _oracle


def _lib():
    L = _oracle.lib()
    for name in ("cpu_type1", "cpu_type2"):
        fn = getattr(L, name)
        fn.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_long), ctypes.c_double, ctypes.c_int,
                       ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                       ctypes.c_int]
        fn.restype = ctypes.c_int
    L.cpu_threads.restype = ctypes.c_int
    L.cpu_last_prepare_seconds.restype = ctypes.c_double
    return L


def host_facts():
    """CPU model, sockets, cores, threads per core and online CPUs of this
    host (from /proc/cpuinfo and sysconf)."""
    import os
    model, phys, cores, siblings = None, set(), None, None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys.add(v)
                elif k == "cpu cores" and cores is None:
                    cores = int(v)
                elif k == "siblings" and siblings is None:
                    siblings = int(v)
    except OSError:
        pass
    tpc = siblings // cores if cores and siblings else None
    return {"cpu_model": model, "sockets": len(phys) or None, "cores_per_socket": cores,
            "threads_per_core": tpc, "online_cpus": os.cpu_count()}


def threads():
    """Host threads to use: every online CPU, capped by the job's CPU share
    (OMP_NUM_THREADS, 16 when unset)."""
    import os
    cap = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    return max(1, min(_lib().cpu_threads(), cap))


def _run(kind, x, vals, n_modes, eps, isign, nthreads):
    L = _lib()
    x = np.ascontiguousarray(x, dtype=np.float32)
    vals = np.ascontiguousarray(vals, dtype=np.complex64)
    M, d = x.shape
    nout = int(np.prod(n_modes)) if kind == 1 else M
    out = np.zeros(nout, dtype=np.complex64)
    nm = (ctypes.c_long * 3)(*(list(n_modes) + [1, 1])[:3])
    fn = L.cpu_type1 if kind == 1 else L.cpu_type2
    rc = fn(d, nm, eps, isign, M, x.ctypes.data, vals.ctypes.data, out.ctypes.data,
            nthreads or threads())
    if rc != 0:
        raise ValueError("cpu plan failed")
    return out


def type1(x, c, n_modes, eps=1e-6, isign=1, nthreads=0):
    return _run(1, x, c, n_modes, eps, isign, nthreads)


def type2(x, f, n_modes, eps=1e-6, isign=1, nthreads=0):
    return _run(2, x, f, n_modes, eps, isign, nthreads)


def _median_time(fn, reps):
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def time_type1(x, c, n_modes, eps, nthreads, reps=3):
    return _median_time(lambda: type1(x, c, n_modes, eps, 1, nthreads), reps)


def time_type1_stages(x, c, n_modes, eps, nthreads, reps=5):
    """Medians over `reps` type-1 executes of the whole execute (sort
    included), its point sort (prepare_points) and the rest
    (execute_prepared), in seconds."""
    L = _lib()
    ex, pre = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        type1(x, c, n_modes, eps, 1, nthreads)
        ex.append(time.perf_counter() - t0)
        pre.append(L.cpu_last_prepare_seconds())
    ep = [e - p for e, p in zip(ex, pre)]
    return {"execute_s": float(np.median(ex)), "prepare_s": float(np.median(pre)),
            "execute_prepared_s": float(np.median(ep))}


def time_type2(x, f, n_modes, eps, nthreads, reps=3):
    return _median_time(lambda: type2(x, f, n_modes, eps, 1, nthreads), reps)
