"""Launch-shape autotuner for the CG kernels.

For a given pattern and precision, times the CG SpMV kernel (``sn_cg_time_spmv``,
HIP events on the solver's stream) for every lanes-per-row candidate on the
actual matrix values and keeps the fastest.  Results are cached per
(rows, nonzeros, precision) for the life of the process.
"""
import ctypes

from . import _native as nat

_CACHE = {}


def tune_lanes(pattern, n, nnz, vals, mask, precision, candidates=(4, 8, 16), reps=30):
    """Best lanes-per-row for this matrix and the measured ms of each candidate."""
    key = (int(n), int(nnz), int(precision))
    if key in _CACHE:
        return _CACHE[key]
    lib = nat.load()
    times = {}
    for lpr in candidates:
        h = ctypes.c_void_p()
        nat.check(lib.sn_cg_create(ctypes.byref(h), pattern, int(precision), int(lpr), 2,
                                   nat.stream_handle()))
        try:
            ms = ctypes.c_double()
            nat.check(lib.sn_cg_time_spmv(h, nat.ptr(vals), nat.ptr(mask), 3, ctypes.byref(ms),
                                          nat.stream_handle()))          # warm-up
            nat.check(lib.sn_cg_time_spmv(h, nat.ptr(vals), nat.ptr(mask), reps,
                                          ctypes.byref(ms), nat.stream_handle()))
            times[lpr] = ms.value
        finally:
            lib.sn_cg_destroy(h)
    best = min(times, key=times.get)
    _CACHE[key] = (best, times)
    return best, times
