"""Launch-shape autotuner for the CG kernels.

* ``tune_lanes``: for a given pattern and precision, times the CG SpMV kernel
  (``sn_cg_time_spmv``, HIP events on the solver's stream) for every
  lanes-per-row candidate on the actual matrix values and keeps the fastest.
  Cached per (rows, nonzeros, precision).
* ``tune_element_grids``: for the mixed solver's element-operator CG, times
  whole CG iterations (element pass, DOF pass and the auxiliary-space
  preconditioner) on the driver's current Jacobian for candidate launch
  shapes (``sn_cg_set_grids``: element-pass workgroups, DOF-pass workgroups
  per CU; ``sn_ams_set_lanes``: lanes per row of the preconditioner's
  restriction, up and tail kernels) and keeps the fastest.  Cached per (elements, DOFs); drivers made
  later for a mesh of that size start with it.
Results live for the life of the process.
"""
import ctypes
import time

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


_GRIDS = {}


def element_grid_candidates(num_el, ams=True):
    """(grid_el, vec_per_cu, ams_lanes) triples: the default element grid (one
    workgroup per 256 elements) with 1 or 2 DOF-pass workgroups per CU and
    half the element grid with 1, each with 4, 8 or 16 preconditioner lanes
    (8 only without the preconditioner)."""
    base = (int(num_el) + 255) // 256
    half = base // 2 // 8 * 8
    shapes = [(0, 1), (0, 2)] + ([(half, 1)] if half >= 64 else [])
    return [(g, v, l) for g, v in shapes for l in ((4, 8, 16) if ams else (8,))]


def cached_element_grids(num_el, num_dofs):
    """The tuned (grid_el, vec_per_cu, ams_lanes) of a mesh of this size, or None."""
    hit = _GRIDS.get((int(num_el), int(num_dofs)))
    return hit[0] if hit else None


def _apply(lib, drv, shape):
    g, v, l = shape
    if drv.ams is not None:
        nat.check(lib.sn_ams_set_lanes(drv.ams.handle, l))
    # (also drops the captured graphs, so the next solve captures the new shape)
    nat.check(lib.sn_cg_set_grids(drv.cg_el.h, g, v, nat.stream_handle()))


def tune_element_grids(drv, candidates=None, iters=64, reps=2):
    """Fastest (grid_el, vec_per_cu, ams_lanes) of the mixed driver's
    element-operator CG and the measured ms per iteration of each candidate.
    ``drv`` must hold a Jacobian (residual() called) and its preconditioner;
    the CG runs ``iters`` iterations (tolerance 0) from A = 0 on b = -R,
    ``reps`` times per candidate after one untimed solve that also captures
    the graphs.  The winner is applied to drv's workspace."""
    import torch
    key = (int(drv.d['num_el']), int(drv.asm.num_dofs))
    lib = nat.load()
    cg = drv.cg_el
    if key in _GRIDS:
        _apply(lib, drv, _GRIDS[key][0])
        return _GRIDS[key]
    torch.neg(drv.R, out=drv.b)
    x = torch.zeros_like(drv.b)
    times = {}
    for shape in (candidates or element_grid_candidates(key[0], drv.ams is not None)):
        _apply(lib, drv, shape)
        try:
            cg.solve(drv.Mr, drv.mask, drv.b, x, rtol=0.0, atol=0.0, maxiter=iters)
            torch.cuda.current_stream().synchronize()
            t = time.perf_counter()
            for _ in range(reps):
                cg.solve(drv.Mr, drv.mask, drv.b, x, rtol=0.0, atol=0.0, maxiter=iters)
            torch.cuda.current_stream().synchronize()
        except RuntimeError:
            continue            # a shape whose solve failed is not a candidate
        times[shape] = (time.perf_counter() - t) * 1e3 / (reps * iters)
    if not times:
        raise RuntimeError('tune_element_grids: no launch shape completed a solve')
    best = min(times, key=times.get)
    _apply(lib, drv, best)
    _GRIDS[key] = (best, times)
    return _GRIDS[key]
