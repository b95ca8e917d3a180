"""Many magnet variants solved together on one GPU.

The variants' meshes (any sizes) are joined into one disjoint mesh: node,
element and edge ids are offset per variant, so the assembled Jacobian is
block diagonal with one block per variant and each variant's DOFs form one
contiguous row segment.  One element-kernel launch per (variant, domain), one
deterministic scatter for the whole batch, and one segment-aware CG
(``sn_cg_create_batched``) whose captured chunk graph advances every
variant's own CG (own alpha, beta, tolerance, stop flag) per replay.  Newton
runs in lock step: every variant gets its own forcing term and stop test; a
converged variant is frozen (its right-hand side is zero, so its CG segment
stops at once).  Residuals are FP64, the Jacobian/CG FP32 ("mixed") or FP64.
"""
import ctypes
import math
import time

import numpy as np

from . import _native as nat
from .curl_curl_assembler import CurlCurlAssembler
from .device import as_dev_i32, device
from .materials import NU0, ConstantReluctance
from .mesh import IRON_TAG, TetMesh, TetModel
from .sources import source_field_device


def join_meshes(meshes):
    """Disjoint union of TetMesh objects (tags of mesh v: 2v+1 iron, 2v+2 air)."""
    nodes, cells, tags, edges, gids, orient, bnd = [], [], [], [], [], [], []
    node_off = edge_off = 0
    edge_row0 = [0]
    for v, m in enumerate(meshes):
        nodes.append(m.nodes)
        cells.append(m.cells + node_off)
        tags.append(np.where(m.tags == IRON_TAG, 2 * v + 1, 2 * v + 2).astype(np.int32))
        edges.append(m.edges + node_off)
        gids.append(m.glob_ids + edge_off)
        orient.append(m.orient)
        bnd.append(m.boundary_edges + edge_off)
        node_off += m.num_nodes
        edge_off += m.num_edges
        edge_row0.append(edge_off)
    j = TetMesh.__new__(TetMesh)
    j.nodes = np.ascontiguousarray(np.concatenate(nodes))
    j.cells = np.ascontiguousarray(np.concatenate(cells), dtype=np.int32)
    j.tags = np.concatenate(tags)
    j.edges = np.ascontiguousarray(np.concatenate(edges), dtype=np.int32)
    j.glob_ids = np.ascontiguousarray(np.concatenate(gids), dtype=np.int32)
    j.orient = np.concatenate(orient).astype(np.int32)
    j.boundary_edges = np.concatenate(bnd).astype(np.int32)
    j.grid = None
    return j, np.array(edge_row0, dtype=np.int64)


class BatchedSolver:
    """Newton for a list of Dipole objects, all on the device at once."""

    def __init__(self, dipoles, quad_order=2, precision='mixed', lanes_per_row=8, chunk=32,
                 precond='jacobi'):
        import torch
        self.dipoles = list(dipoles)
        self.nvar = len(self.dipoles)
        self.quad_order = quad_order
        self.f32 = precision == 'mixed'
        self.precision = precision
        t0 = time.perf_counter()
        self.mesh, self.row0 = join_meshes([d.mesh for d in self.dipoles])
        self.model = TetModel(self.mesh)
        tags, mats = [], []
        for v, d in enumerate(self.dipoles):
            tags += [2 * v + 1, 2 * v + 2]
            mats += [d.reluctance, ConstantReluctance(NU0)]
        self.asm = CurlCurlAssembler(self.model.mesh, tags, mats)
        d = self.asm._device_state()
        self.n = self.asm.num_dofs
        mask = np.ones(self.n, dtype=np.uint8)
        mask[self.mesh.boundary_edges] = 0
        self.mask = torch.from_numpy(mask).to(device())
        self.row0_dev = as_dev_i32(self.row0)
        N, nnz = d['num_el'], d['nnz']
        jt = torch.float32 if self.f32 else torch.float64
        self.locJ = torch.empty(21 * N, dtype=jt, device=device())
        self.loc6 = torch.empty(6 * N, dtype=torch.float64, device=device())
        self.J = torch.empty(nnz, dtype=jt, device=device())
        self.R = torch.empty(self.n, dtype=torch.float64, device=device())
        self.b = torch.empty(self.n, dtype=torch.float64, device=device())
        self.cg = ctypes.c_void_p()
        self.cg_safe = ctypes.c_void_p()
        row0 = (ctypes.c_int * (self.nvar + 1))(*[int(r) for r in self.row0])
        nat.check(nat.load().sn_cg_create_batched(ctypes.byref(self.cg), d['pattern'],
                                                  32 if self.f32 else 64, lanes_per_row, chunk,
                                                  self.nvar, row0, nat.stream_handle()))
        if self.f32:
            # float64 Jacobian and CG for segments whose float32 CG stalls
            nat.check(nat.load().sn_cg_create_batched(ctypes.byref(self.cg_safe), d['pattern'],
                                                      64, lanes_per_row, chunk, self.nvar, row0,
                                                      nat.stream_handle()))
            self.locJ64 = torch.empty(21 * N, dtype=torch.float64, device=device())
            self.loc6b = torch.empty(6 * N, dtype=torch.float64, device=device())
            self.J64 = torch.empty(nnz, dtype=torch.float64, device=device())
        self.precond = precond
        self.nu_el = torch.empty(N, dtype=torch.float32, device=device()) \
            if precond == 'ams' else None
        self.ams = None
        self.t_setup = time.perf_counter() - t0
        self.src = None

    def _refresh_preconditioner(self):
        """Auxiliary-space preconditioner of the joined mesh (its cycle is
        block diagonal like the Jacobian) for the last assembled Jacobian."""
        if self.precond != 'ams':
            return
        if self.ams is None:
            from .ams import AmsPreconditioner
            self.ams = AmsPreconditioner(self.asm, self.mask.cpu().numpy().astype(bool),
                                         self.nu_el)
            for h in (self.cg, self.cg_safe):
                if h:
                    nat.check(nat.load().sn_cg_set_ams(h, self.ams.handle))
        else:
            self.ams.update(self.nu_el)

    def __del__(self):
        try:
            if nat._lib is not None:
                for h in (self.cg, self.cg_safe):
                    if h:
                        nat._lib.sn_cg_destroy(h)
        except Exception:
            pass

    # ------------------------------------------------------------ pieces
    def source_fields(self):
        """Biot-Savart of each variant's coils at its iron quadrature points."""
        pts = self.asm.get_quadrature_points_device(self.quad_order,
                                                    [2 * v for v in range(self.nvar)])
        src = []
        for v, d in enumerate(self.dipoles):
            src += [source_field_device(d.coils, pts[v]).reshape(-1), None]
        self.src = src
        return src

    def residual(self, x, active=None):
        """R(x) (FP64) and the Jacobian; per-variant ||R||.  Only the variants
        flagged in ``active`` are re-assembled (the others keep their last
        local values, so their entries of J and R are unchanged)."""
        lib = nat.load()
        d = self.asm._device_state()
        doms = None
        if active is not None:
            doms = [2 * v + k for v in range(self.nvar) if active[v] for k in (0, 1)]
        self.asm.newton_local_device(x, self.quad_order, self.src, 1, self.f32, self.locJ,
                                     self.loc6, domains=doms, nu_el=self.nu_el)
        nat.check(lib.sn_csr_gather_one(d['pattern'], int(self.f32), nat.ptr(self.locJ),
                                        nat.ptr(self.J), nat.stream_handle()))
        nat.check(lib.sn_dof_gather_f64(d['pattern'], nat.ptr(self.loc6), nat.ptr(self.R),
                                        nat.stream_handle()))
        out = (ctypes.c_double * self.nvar)()
        nat.check(lib.sn_seg_dot_f64(self.nvar, nat.ptr(self.row0_dev), nat.ptr(self.R),
                                     nat.ptr(self.R), nat.ptr(self.mask), out,
                                     nat.stream_handle()))
        return np.sqrt(np.array(out[:]))

    def _segment_mask(self, flags):
        """Per-DOF float64 factor: 1 on the rows of variants with flags[v]."""
        import torch
        f = torch.from_numpy(np.repeat(np.asarray(flags, dtype=np.float64),
                                       np.diff(self.row0))).to(device())
        return f

    # ------------------------------------------------------------ solve
    def solve(self, rtol=1e-8, max_newton=40, eta_max=0.1, eta_min=None, max_cg=-1, x0=None,
              log=None):
        """Newton for every variant to ||R_v|| <= rtol ||R_v(0)||."""
        import torch
        lib = nat.load()
        t0 = time.perf_counter()
        self.source_fields()
        if eta_min is None:
            eta_min = 1e-3 if self.f32 else 1e-6
        if max_cg < 0:
            max_cg = 50000
        z = torch.zeros(self.n, dtype=torch.float64, device=device())
        r0 = self.residual(z)
        x = z.clone() if x0 is None else x0.clone()
        x[self.mask == 0] = 0.0
        rn = self.residual(x)
        hist = [rn / r0]
        eta = np.full(self.nvar, eta_max)
        rn_prev = None
        conv = rn <= rtol * r0
        cg_its = []
        k = 0
        while not conv.all() and k < max_newton:
            if rn_prev is not None:
                eta = np.minimum(eta_max, np.maximum(0.9 * (rn / rn_prev) ** 2, eta_min))
            # the step that can reach the target asks for just that (with the
            # margin of NewtonDriver.solve)
            eta = np.maximum(eta, np.minimum(eta_max, 0.9 * rtol * r0 / np.maximum(rn, 1e-300)))
            active = self._segment_mask(~conv)
            torch.mul(self.R, active, out=self.b)
            self.b.neg_()
            self._refresh_preconditioner()
            rt = (ctypes.c_double * self.nvar)(*[float(e) for e in eta])
            its = (ctypes.c_int * self.nvar)()
            status = (ctypes.c_int * self.nvar)()
            if self.f32:
                # float32 CG; with the auxiliary-space preconditioner, forcing
                # terms below 5e-3 go to float64 directly (the rounding of J
                # stalls the float32 CG there)
                force64 = (eta < 4.9e-3) & ~conv if self.precond == 'ams' else \
                    np.zeros(self.nvar, dtype=bool)
                x_try = x.clone()
                b32 = self.b * self._segment_mask(~force64) if force64.any() else self.b
                cap = 600 if self.precond == 'ams' else 4000
                rc = lib.sn_cg_solve_batched(self.cg, nat.ptr(self.J), nat.ptr(self.mask),
                                             nat.ptr(b32), nat.ptr(x_try), rt, 0.0,
                                             min(max_cg, cap), its, None, status,
                                             nat.stream_handle())
                stalled = (np.array(status[:]) == nat.SN_ERR_NOT_CONVERGED) | force64
                if stalled.any():
                    # redo the stalled segments in float64 (Jacobian and CG)
                    self.asm.newton_local_device(x, self.quad_order, self.src, 1, False,
                                                 self.locJ64, self.loc6b)
                    nat.check(lib.sn_csr_gather_one(self.asm._device_state()['pattern'], 0,
                                                    nat.ptr(self.locJ64), nat.ptr(self.J64),
                                                    nat.stream_handle()))
                    b2 = self.b * self._segment_mask(stalled)
                    dx = torch.zeros_like(x)
                    its2 = (ctypes.c_int * self.nvar)()
                    rc = lib.sn_cg_solve_batched(self.cg_safe, nat.ptr(self.J64),
                                                 nat.ptr(self.mask), nat.ptr(b2), nat.ptr(dx), rt,
                                                 0.0, max_cg, its2, None, None,
                                                 nat.stream_handle())
                    keep = self._segment_mask(~stalled)
                    x_try = keep * x_try + (1.0 - keep) * (x + dx)
                    for v in np.nonzero(stalled)[0]:
                        its[v] += its2[v]
            else:
                dx = torch.zeros_like(x)
                rc = lib.sn_cg_solve_batched(self.cg, nat.ptr(self.J), nat.ptr(self.mask),
                                             nat.ptr(self.b), nat.ptr(dx), rt, 0.0, max_cg,
                                             its, None, None, nat.stream_handle())
                x_try = x + dx
            if rc not in (nat.SN_OK, nat.SN_ERR_NOT_CONVERGED):
                nat.check(rc)
            cg_its.append(list(its[:]))
            rn_try = self.residual(x_try, active=~conv)
            # per-variant backtracking on the residual norm
            dxs = x_try - x
            step = np.ones(self.nvar)
            for _ in range(6):
                bad = (rn_try >= rn) & ~conv
                if not bad.any():
                    break
                step[bad] *= 0.5
                x_try = x + self._segment_mask(step) * dxs
                rn_try = self.residual(x_try, active=bad)
            rn_prev, rn, x = rn, np.where(conv, rn, rn_try), x_try
            k += 1
            hist.append(rn / r0)
            conv = rn <= rtol * r0
            if log is not None:
                log('  batch newton %d: worst |R|/|R0| %.3e, converged %d/%d, cg max %d, %.3f s'
                    % (k, float(np.max(rn / r0)), int(conv.sum()), self.nvar,
                       max(cg_its[-1]), time.perf_counter() - t0))
        torch.cuda.synchronize()
        self.stats = {'residual_history': np.array(hist), 'cg_iterations': cg_its,
                      'newton_steps': k, 'converged': conv, 't_solve': time.perf_counter() - t0,
                      'precond': self.precond}
        return x

    def split(self, x):
        """Per-variant host vectors of a joined solution."""
        h = x.cpu().numpy()
        return [h[self.row0[v]:self.row0[v + 1]] for v in range(self.nvar)]


class ConcurrentSolver:
    """Many variants solved independently on several HIP streams at once.

    Each worker thread owns a stream (default priority) and solves the
    variants handed to it one after another with the single-variant Newton
    solver; the GPU overlaps the kernels of the streams, which fills the
    device where one small variant alone is launch- or latency-bound.  The
    host waits of one stream release the interpreter lock (they happen inside
    the C library), so the threads' host work overlaps too.

    ``warm_start='anchor'`` (variants on one mesh that differ in the coil
    current density ``p['J']`` only): the middle variant of the list is solved
    from zero first; every other variant then starts from its solution scaled
    by the ratio of the current densities, the solution of the linear part of
    the problem.  Each variant still runs Newton to ||R|| <= rtol ||R(0)|| of
    its own, R(0) its residual at zero.  The anchor is the variant of median
    current density; when it does not converge every variant starts from zero.

    ``warm_start='learned'`` with ``model`` (a trained WarmStart whose family
    topology every variant shares): the network predicts every variant's POD
    coefficients in one batched pass, then each worker expands its variants'
    predictions, takes one reduced Newton step in the leading modes and keeps
    the result if its residual is below the residual at zero (WarmStart.start);
    Newton then runs to the same tolerance."""

    def __init__(self, dipoles, workers=4, quad_order=2, precision='mixed', precond='ams',
                 lanes_per_row=8, warm_start=None, model=None, ams_options=None):
        import torch
        self.dipoles = list(dipoles)
        self.workers = max(1, int(workers))
        self.kw = dict(precision=precision, precond=precond, lanes_per_row=lanes_per_row)
        self.ams_options = dict(ams_options or {})
        # extra keyword arguments of every variant's solve_newton (forcing terms)
        self.newton_options = {}
        self.set_warm_start(warm_start, model)      # checks the variants first
        self.streams = [torch.cuda.Stream() for _ in range(self.workers)]
        self.solvers = [d.solver(quad_order=quad_order, verbose=False) for d in self.dipoles]
        if self.ams_options and precond == 'ams':
            # the options of the auxiliary-space preconditioner that each
            # variant's driver builds on its first Newton step
            for s in self.solvers:
                s.newton_driver(precision, lanes_per_row, 32, 'ams').ams_options = \
                    dict(self.ams_options)
        self._prepare_learned()

    def set_warm_start(self, warm_start, model=None):
        """Choose the start of the next solves: None (A = 0), 'anchor' or
        'learned' (with ``model``); the variants are checked for it."""
        if warm_start not in (None, 'anchor', 'learned'):
            raise ValueError("warm_start must be None, 'anchor' or 'learned'")
        precision, precond = self.kw['precision'], self.kw['precond']
        if warm_start == 'learned':
            if model is None:
                raise ValueError("learned warm start: pass the trained model (model=WarmStart)")
            if precision != 'mixed' or precond != 'ams':
                raise ValueError('learned warm start: the reduced Newton step needs the mixed '
                                 'solver with the auxiliary-space preconditioner')
            fm = model.family.mesh
            for d in self.dipoles:
                if d.mesh.cells.shape != fm.cells.shape or not np.array_equal(d.mesh.cells,
                                                                             fm.cells):
                    raise ValueError('learned warm start: every variant must be on the '
                                     "model's family topology")
        self.model = model
        if warm_start == 'anchor':
            m0, d0 = self.dipoles[0].mesh, self.dipoles[0]
            for d in self.dipoles:
                if (d.mesh.num_edges != m0.num_edges or d.mesh.cells.shape != m0.cells.shape
                        or not np.array_equal(d.mesh.cells, m0.cells)
                        or not np.array_equal(d.mesh.nodes, m0.nodes)):
                    raise ValueError('anchor warm start: every variant must share the mesh')
                if not d.p.get('J'):
                    raise ValueError('anchor warm start: every variant needs a current density')
                if _without_J(d.p) != _without_J(d0.p) or \
                        not _same_reluctance(d.reluctance, d0.reluctance):
                    raise ValueError('anchor warm start: the variants may differ in the coil '
                                     'current density J only')
        self.warm_start = warm_start
        self._prepare_learned()

    def _prepare_learned(self):
        """Set-up of the learned start on every variant's mesh (the modes'
        face circulations and the state-independent part of the reduced
        Jacobian), so that the solves do none of it."""
        if self.warm_start != 'learned' or not hasattr(self, 'solvers'):
            return
        for s in self.solvers:
            self.model.precompute_constant(s, lanes_per_row=self.kw['lanes_per_row'])

    def _run(self, w, idx, rtol, out, x0s=None, coeffs=None):
        """Worker w: solve on its stream the variants ``idx`` (a list, or a
        callable handing out the next variant, None when none is left)."""
        import torch
        if callable(idx):
            pull = idx
        else:
            it = iter([int(i) for i in idx])
            pull = lambda: next(it, None)    # noqa: E731
        with torch.cuda.stream(self.streams[w]):
            while True:
                i = pull()
                if i is None:
                    break
                s = self.solvers[i]
                s._src = None
                x0, info = None if x0s is None else x0s.get(int(i)), None
                if coeffs is not None:
                    from .warm_start import PARAM_NAMES
                    p = {k: self.dipoles[i].p[k] for k in PARAM_NAMES}
                    x0, info = self.model.start(s, p, coeffs=coeffs[i],
                                                lanes_per_row=self.kw['lanes_per_row'])
                x = s.solve_newton(self.kw['precision'], rtol=rtol, x0=x0,
                                   lanes_per_row=self.kw['lanes_per_row'],
                                   precond=self.kw['precond'], **self.newton_options)
                st = dict(s.stats)
                if info is not None:
                    st['warm_start'] = info
                out[i] = (x, st)
            torch.cuda.current_stream().synchronize()

    def solve(self, rtol=1e-8):
        """Solutions (device tensors) and stats of every variant, in order."""
        import threading
        import torch
        out = [None] * len(self.dipoles)
        # one queue, largest variants first: a worker takes the next variant
        # as soon as its last one has converged, so the streams stay busy
        # and finish together
        order = np.argsort([-d.num_tets for d in self.dipoles], kind='stable')
        x0s = None
        if self.warm_start == 'anchor' and len(self.dipoles) > 1:
            a = anchor_index([float(d.p['J']) for d in self.dipoles])
            torch.cuda.synchronize()
            self._run(0, [a], rtol, out)
            xa, ja = out[a][0], float(self.dipoles[a].p['J'])
            order = np.array([i for i in order if i != a])
            if anchor_usable(out[a][1], xa):
                x0s = {i: xa * (float(d.p['J']) / ja) for i, d in enumerate(self.dipoles)
                       if i != a}
            # else: the anchor failed, every other variant starts from zero
        coeffs = None
        if self.warm_start == 'learned':
            # one batched forward pass of the network for every variant (the
            # network's buffers are not shared across threads)
            from .warm_start import PARAM_NAMES
            coeffs = self.model.coefficients([{k: d.p[k] for k in PARAM_NAMES}
                                              for d in self.dipoles])
            self.model.to_device(self.solvers[0].curl_curl_factory)
        pending = [int(i) for i in order]
        lock = threading.Lock()

        def pull():
            with lock:
                return pending.pop(0) if pending else None
        torch.cuda.synchronize()
        threads = [threading.Thread(target=self._run, args=(w, pull, rtol, out, x0s, coeffs))
                   for w in range(self.workers)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        torch.cuda.synchronize()
        if any(o is None for o in out):
            raise RuntimeError('a worker failed; see the thread traceback above')
        return out


def _without_J(p):
    return {k: v for k, v in p.items() if k != 'J'}


def _same_reluctance(a, b):
    """Two reluctivity models with identical tables (or the same object)."""
    if a is b:
        return True
    from .materials import model_tables
    ta, tb = model_tables(a), model_tables(b)
    return len(ta) == len(tb) and all(
        (np.array_equal(np.asarray(u), np.asarray(v)) if not np.isscalar(u) else u == v)
        for u, v in zip(ta, tb))


def anchor_index(currents):
    """The variant of median current density (the lower median for an even
    count), the anchor whose solution seeds the others."""
    order = np.argsort(np.asarray(currents, dtype=float), kind='stable')
    return int(order[(len(order) - 1) // 2])


def anchor_usable(stats, x):
    """The anchor's solution seeds the other variants only if its Newton
    converged and every entry is finite; otherwise they start from zero."""
    import torch
    return bool(stats.get('converged')) and bool(torch.isfinite(x).all().item())


def _next_power_of_two(n):
    p = 1
    while p < n:
        p *= 2
    return p


class VariantBatchSolver:
    """Variants that share one mesh topology, solved in lock step by the
    variant-minor CG (``sn_vcg``): one set of kernel launches per CG iteration
    for all of them, every index stream read once for the batch.

    The variants may differ in their coils (J, placement), their materials
    and their node coordinates (moved copies of one mesh: the learned start's
    family), not in the mesh's cells or Dirichlet edges.  Each variant keeps
    its own residual, Jacobian, forcing terms and stop test (Newton to
    ||R_v|| <= rtol ||R_v(0)||, R_v(0) its residual at zero, as
    ``RedMVPSolver.solve_newton``); the float32 CG of a Newton step runs for
    all unconverged variants at once, each to its own forcing term.  The
    auxiliary-space preconditioner is shared: it is the preconditioner of the
    anchor variant (the middle one of the list), refreshed with the anchor's
    reluctivities at every step.  A variant whose float32 CG stops at its cap,
    or whose Newton stalls, is finished by its own ``solve_newton``.

    ``warm_start='learned'`` with ``model`` (a WarmStart of the variants'
    family topology): one batched forward pass of the network, then each
    variant's learned start (WarmStart.start) before the lock-step Newton.
    At most 8 variants; a batch of 3, 5, 6 or 7 pads the CG to 4 or 8 systems
    with empty right-hand sides."""

    def __init__(self, dipoles, quad_order=2, lanes_per_row=8, chunk=32, warm_start=None,
                 model=None, anchor=None, workers=8, batched_start=True):
        import torch
        self.dipoles = list(dipoles)
        V = len(self.dipoles)
        if not 1 <= V <= 8:
            raise ValueError('VariantBatchSolver: 1 to 8 variants (%d given)' % V)
        m0 = self.dipoles[0].mesh
        for d in self.dipoles[1:]:
            m = d.mesh
            if (m.num_edges != m0.num_edges or m.cells.shape != m0.cells.shape
                    or not np.array_equal(m.cells, m0.cells)
                    or not np.array_equal(m.glob_ids, m0.glob_ids)
                    or not np.array_equal(m.boundary_edges, m0.boundary_edges)):
                raise ValueError('VariantBatchSolver: every variant must share the mesh topology')
        self.nvar = V
        self.nv = _next_power_of_two(V)
        self.lanes_per_row, self.chunk = lanes_per_row, chunk
        self.solvers = [d.solver(quad_order=quad_order, verbose=False) for d in self.dipoles]
        self.anchor = (V - 1) // 2 if anchor is None else int(anchor)
        # the per-variant work around the batched CG (sources, learned start,
        # residuals) runs on one HIP stream per variant, from host threads:
        # variant v always on stream v % workers, so its tensors are
        # allocated, used and freed on one stream
        self.workers = max(1, min(int(workers), V))
        # learned starts: queued from this thread with the passes over the
        # modes shared by all variants and one host wait (_batched_starts;
        # 54.9 ms per batch of 8 at 1M tets), or each variant's own pipeline
        # on its stream from a worker thread (sources, WarmStart.start,
        # residual; 57.4 ms, same box)
        self.batched_start = bool(batched_start)
        self.streams = [torch.cuda.Stream() for _ in range(self.workers)]
        self._pool = None
        self._operands = None
        self.h = None
        self.ams = None
        self.stats = {}
        self.set_warm_start(warm_start, model)

    def set_warm_start(self, warm_start, model=None):
        if warm_start not in (None, 'learned'):
            raise ValueError("warm_start must be None or 'learned'")
        if warm_start == 'learned':
            if model is None:
                raise ValueError('learned warm start: pass the trained model (model=WarmStart)')
            fm = model.family.mesh
            for d in self.dipoles:
                if d.mesh.cells.shape != fm.cells.shape or not np.array_equal(d.mesh.cells,
                                                                             fm.cells):
                    raise ValueError('learned warm start: every variant must be on the '
                                     "model's family topology")
            for s in self.solvers:
                model.precompute_constant(s, lanes_per_row=self.lanes_per_row)
        self.warm_start, self.model = warm_start, model

    def _each(self, fn, idx):
        """fn(v) for the variants idx, each on its own stream (worker threads);
        returns once every stream has finished."""
        import torch
        from concurrent.futures import ThreadPoolExecutor
        idx = list(idx)
        if not idx:
            return
        torch.cuda.synchronize()
        if self._pool is None:
            # kept for the solver's lifetime: no thread start-up per call
            self._pool = ThreadPoolExecutor(max_workers=self.workers,
                                            thread_name_prefix='variant-batch')

        def work(w):
            with torch.cuda.stream(self.streams[w]):
                for v in idx:
                    if v % self.workers == w:
                        fn(v)
                torch.cuda.current_stream().synchronize()
        futures = [self._pool.submit(work, w) for w in range(self.workers)
                   if any(v % self.workers == w for v in idx)]
        errors = [f.exception() for f in futures]
        torch.cuda.synchronize()
        for e in errors:
            if e is not None:
                raise e          # the first error, in the calling thread

    def _driver(self, v):
        return self.solvers[v].newton_driver('mixed', self.lanes_per_row, self.chunk, 'ams')

    def _batched_starts(self, drvs, coeffs, r0, x0s, infos):
        """Sources, |R(0)| and starts of every variant, queued from this
        thread with one host wait at the end.  With the learned start:

        1. x = mean + U a of all variants in one pass over U (this stream);
        2. per variant on its own stream: Biot-Savart, R(0) (its norm left on
           the device), R and the Jacobian at x, H = U^T J U on the device;
        3. g = U^T R of all variants in one pass, the reduced systems solved
           on the device (sn_pod_reduced_solve_many), x += U d of all
           variants in one pass (2 and 3 once per reduced Newton step);
        4. the judged residuals, their norms and the |R(0)| read back
           together.

        The starts are WarmStart.start's up to the order of the correction's
        sum (added after the expansion here, inside it there).  From zero:
        the sources, then R at zero, which is R(0) and Newton's first
        residual."""
        import torch
        V, NV = self.nvar, self.nv
        ws = self.model
        lib = nat.load()
        n = self.solvers[0].curl_curl_factory.num_dofs
        if coeffs is None:
            rg, steps = 0, 0
            X = torch.zeros((NV, n), dtype=torch.float64, device=device())
        else:
            rg = ws.r_galerkin
            steps = ws.galerkin_steps if rg else 0
            A = np.zeros((NV, coeffs.shape[1]))
            A[:V] = coeffs
            X = ws.expand_many(A)
        # |R(0)|^2, |R(x0)|^2 of the judges, and |R|^2 at the reduced steps' x (unused)
        norm2 = torch.zeros(3 * V, dtype=torch.float64, device=device())
        H = torch.zeros((NV, rg, rg), dtype=torch.float64, device=device()) if rg else None
        main = torch.cuda.current_stream()
        for step in range(max(1, steps)):
            ready = main.record_event()
            done = []
            for v in range(V):
                st = self.streams[v % self.workers]
                st.wait_event(ready)
                with torch.cuda.stream(st):
                    drv, s = drvs[v], self.solvers[v]
                    if step == 0:
                        drv.retarget(s.source_fields_device(wait=False))
                        if drv._r0 is None and coeffs is not None:
                            drv.r0_queued(norm2[v])
                    if steps:
                        drv.residual(X[v], norm2[2 * V + v])
                        ws.reduced_jacobian_device(drv, H[v])
                    done.append(st.record_event())
            for e in done:
                main.wait_event(e)
            if not steps:
                break
            R = torch.zeros((NV, n), dtype=torch.float64, device=device())
            for v in range(V):
                R[v].copy_(drvs[v].R)
            G = torch.zeros((NV, rg), dtype=torch.float64, device=device())
            nat.check(lib.sn_pod_project_many_dev(n, rg, NV, nat.ptr(ws._U), nat.ptr(R),
                                                  nat.ptr(G), nat.stream_handle()))
            D = torch.zeros((NV, rg), dtype=torch.float32, device=device())
            nat.check(lib.sn_pod_reduced_solve_many(V, rg, nat.ptr(H), nat.ptr(G), nat.ptr(D),
                                                    nat.stream_handle()))
            X += ws.expand_many(D, mean=False)
        # one tensor per start (rows of X would share one version counter)
        xs = [X[v].clone() for v in range(V)]
        for v in range(V):
            drv = drvs[v]
            xs[v][drv.mask == 0] = 0.0
            drv.residual(xs[v], norm2[V + v])
        n2 = norm2.cpu().tolist()
        for v in range(V):
            drv = drvs[v]
            rn = math.sqrt(n2[V + v])
            if coeffs is None:
                drv._r0 = rn
            elif drv._r0 is None:
                drv._r0 = math.sqrt(n2[v])
            r0[v] = drv._r0
            drv._last = (xs[v], xs[v]._version, rn)     # R and the Jacobian at the start
            if coeffs is None:
                x0s[v] = xs[v]
                continue
            r = rn / r0[v]
            use = bool(np.isfinite(r) and r < 1.0)
            x0s[v] = xs[v] if use else None
            infos[v] = {'r_start_rel': float(r), 'used': use}

    def _setup(self):
        """The workspace and the shared preconditioner (first solve)."""
        if self.h is not None:
            return
        from .ams import AmsPreconditioner
        s0 = self.solvers[self.anchor]
        asm = s0.curl_curl_factory
        d = asm._device_state()
        self.h = ctypes.c_void_p()
        nat.check(nat.load().sn_vcg_create(ctypes.byref(self.h), d['pattern'], nat.ptr(d['gids']),
                                           nat.ptr(d['orient']), self.nv, self.chunk,
                                           nat.stream_handle()))
        drv = self._driver(self.anchor)
        free = s0.mask_device().cpu().numpy().astype(bool)
        self.ams = AmsPreconditioner(asm, free, drv.nu_el)
        nat.check(nat.load().sn_vcg_set_ams(self.h, self.ams.handle, nat.stream_handle()))

    def solve(self, rtol=1e-8, max_newton=40, eta_max=0.1, eta_min=1e-3, eta_margin=0.9,
              close_start=1e-2, max_cg=-1, log=None, plan_close=True, lag_preconditioner=True):
        """Solutions (device float64) and stats of every variant, in order.
        ``lag_preconditioner``: after the first Newton step, refresh the
        shared preconditioner only while the anchor variant's residual is
        above close_start |R(0)|."""
        import torch
        from .red_mvp_solver import next_forcing, planned_forcing
        V, NV = self.nvar, self.nv
        lib = nat.load()
        t0 = time.perf_counter()
        drvs = [self._driver(v) for v in range(V)]
        n = self.solvers[0].curl_curl_factory.num_dofs
        if max_cg < 0:
            max_cg = min(10 * n, 50000)
        # starts: the learned one (one batched forward pass of the network)
        # or zero
        x0s, infos = [None] * V, [None] * V
        coeffs = None
        if self.warm_start == 'learned':
            from .warm_start import PARAM_NAMES
            coeffs = self.model.coefficients([{k: d.p[k] for k in PARAM_NAMES}
                                              for d in self.dipoles])
        r0, xs, rn = [None] * V, [None] * V, [None] * V

        from .warm_start import PARAM_NAMES
        params = [{k: d.p[k] for k in PARAM_NAMES} for d in self.dipoles]

        def prepare(v):
            # R at the start (kept from the judge of the learned start)
            drv = drvs[v]
            last = getattr(drv, '_last', None)
            x0 = x0s[v]
            if x0 is not None and last is not None and last[0] is x0 and \
                    x0._version == last[1]:
                xs[v] = x0.clone()
                rn[v] = last[2]              # R and the Jacobian at x0 are still current
            else:
                x = torch.zeros(n, dtype=torch.float64, device=device()) if x0 is None else \
                    x0.clone()
                x[drv.mask == 0] = 0.0
                xs[v] = x
                rn[v] = drv.residual(x)

        def sources(v):
            # the variant's sources (Biot-Savart) and |R(0)|; unless the starts
            # are batched, its learned start and R there right after, on the
            # same stream
            drv, s = drvs[v], self.solvers[v]
            drv.retarget(s.source_fields_device(wait=False))
            r0[v] = drv.r0()
            if coeffs is not None:
                x0s[v], infos[v] = self.model.start(s, params[v], coeffs=coeffs[v],
                                                    lanes_per_row=self.lanes_per_row)
            prepare(v)
        if self.batched_start:
            self._batched_starts(drvs, coeffs, r0, x0s, infos)
            t_sources = time.perf_counter() - t0
            for v in range(V):
                prepare(v)
        else:
            self._each(sources, range(V))
            t_sources = time.perf_counter() - t0
        t_prepare = time.perf_counter() - t0
        t_cg = t_advance = 0.0
        self._setup()
        mask = drvs[0].mask
        hist = [[rn[v] / r0[v]] for v in range(V)]
        cg_its, etas = [[] for _ in range(V)], [[] for _ in range(V)]
        eta = [eta_max] * V
        for v in range(V):
            if x0s[v] is not None and 0.0 < rn[v] <= close_start * r0[v]:
                eta[v] = planned_forcing(rn[v] / r0[v], rtol, eta_min, eta_max)
        rn_prev = [None] * V
        conv = [rn[v] <= rtol * r0[v] for v in range(V)]
        live = [not c for c in conv]          # in the lock step
        stalls = [0] * V
        if self._operands is None:
            # kept for the solver's lifetime: the CG's captured graphs are
            # keyed by the operand pointers, so the same buffers let every
            # solve replay them instead of capturing them again
            self._operands = (
                torch.empty((6 * drvs[0].d['num_el'], NV), dtype=torch.float32, device=device()),
                torch.zeros((n, NV), dtype=torch.float64, device=device()),
                torch.empty((n, NV), dtype=torch.float64, device=device()))
        Mr, b, dx = self._operands
        norm2 = torch.zeros(V, dtype=torch.float64, device=device())
        its_h = (ctypes.c_int * NV)()
        rn_h = (ctypes.c_double * NV)()
        st_h = (ctypes.c_int * NV)()
        k = 0
        while any(live) and k < max_newton:
            rtol_h = (ctypes.c_double * NV)(*([1.0] * NV))
            cap = 300
            for v in range(V):
                if not live[v]:
                    continue
                eta[v] = next_forcing(rn[v], rn_prev[v], r0[v], rtol, eta[v], eta_min, eta_max,
                                      eta_margin, close_start if plan_close else None)
                etas[v].append(eta[v])
                rtol_h[v] = eta[v]
                cap = max(cap, 4 * (cg_its[v][-1] if cg_its[v] else 0))
            cap = min(cap, max_cg)
            # Mr and b = -R of the live variants packed variant-minor in one
            # pass (padding systems repeat the last variant's Mr, with b = 0)
            mr_src = (ctypes.c_void_p * NV)(*[drvs[min(v, V - 1)].Mr.data_ptr()
                                                for v in range(NV)])
            r_src = (ctypes.c_void_p * NV)(*[drvs[v].R.data_ptr() if v < V and live[v] else None
                                               for v in range(NV)])
            nat.check(lib.sn_vcg_pack(self.h, mr_src, r_src, nat.ptr(Mr), nat.ptr(b),
                                      nat.stream_handle()))
            a = self.anchor
            if k == 0 or not (rn[a] <= close_start * r0[a]) or not lag_preconditioner:
                # once the anchor is within close_start of the solution its
                # reluctivities barely move: the preconditioner of the step
                # before is kept (the same CG iterations at 1M tets, from the
                # learned start and from zero)
                self.ams.update(drvs[a].nu_el)
            dx.zero_()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            rc = lib.sn_vcg_solve(self.h, nat.ptr(Mr), nat.ptr(mask), nat.ptr(b), nat.ptr(dx),
                                  rtol_h, 0.0, int(cap), its_h, rn_h, st_h, nat.stream_handle())
            if rc not in (nat.SN_OK, nat.SN_ERR_NOT_CONVERGED, nat.SN_ERR_BREAKDOWN):
                nat.check(rc)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            t_cg += t2 - t1

            # the steps: every live variant's residual at x + dx queued on one
            # stream (memory-bound kernels that each fill the GPU), one read of
            # their norms, then line searches where a full step does not
            # reduce the residual
            steps = []
            for v in range(V):
                if not live[v]:
                    continue
                cg_its[v].append(int(its_h[v]))
                if st_h[v] != nat.SN_OK:
                    live[v] = False          # finished by its own solver below
                    continue
                x_try = xs[v] + dx[:, v]
                drvs[v].residual(x_try, norm2[v])
                steps.append((v, x_try))
            n2 = norm2.cpu().tolist() if steps else []
            for v, x_try in steps:
                drv = drvs[v]
                rn_try = math.sqrt(n2[v])
                lam, tries = 1.0, 0
                while not (rn_try < rn[v]) and tries < 6:
                    lam *= 0.5
                    x_try = xs[v] + lam * dx[:, v]
                    rn_try = drv.residual(x_try)
                    tries += 1
                stalled = rn_try >= rn[v] or (lam == 1.0 and rn_try > 0.9 * rn[v])
                stalls[v] = stalls[v] + 1 if stalled else 0
                rn_prev[v], rn[v], xs[v] = rn[v], rn_try, x_try
                hist[v].append(rn[v] / r0[v])
                conv[v] = rn[v] <= rtol * r0[v]
                if conv[v] or stalls[v] >= 3:
                    live[v] = False
            t_advance += time.perf_counter() - t2
            k += 1
            if log is not None:
                log('  lock-step newton %d: |R|/|R0| = %s, cg %s'
                    % (k, ['%.2e' % h[-1] for h in hist], [c[-1] if c else 0 for c in cg_its]))
        out = []
        for v in range(V):
            st = {'residual_history': hist[v], 'cg_iterations': cg_its[v], 'etas': etas[v],
                  'newton_steps': len(cg_its[v]), 'converged': bool(conv[v]), 'r0': r0[v],
                  'precision': 'mixed', 'precond': 'ams (shared, variant %d)' % self.anchor,
                  'fallback': False}
            if infos[v] is not None:
                st['warm_start'] = infos[v]
            x = xs[v]
            if not conv[v]:
                # the float32 CG stopped at its cap, or Newton stalled: this
                # variant's own solver (with its float64 fallback) finishes it
                s = self.solvers[v]
                x = s.solve_newton('mixed', rtol=rtol, x0=xs[v], precond='ams',
                                   lanes_per_row=self.lanes_per_row)
                st['fallback'] = True
                st['fallback_stats'] = {k2: s.stats.get(k2) for k2 in
                                        ('residual_history', 'cg_iterations', 'converged')}
                st['converged'] = bool(s.stats['converged'])
                st['cg_iterations'] = cg_its[v] + list(s.stats['cg_iterations'])
            out.append((x, st))
        torch.cuda.current_stream().synchronize()
        self.stats = {'t_solve': time.perf_counter() - t0, 'newton_rounds': k,
                      'anchor': self.anchor, 'systems': NV,
                      # sources, starts and the first residuals; the batched CG
                      # solves; the residuals and line searches after them
                      't_prepare': t_prepare, 't_sources': t_sources, 't_cg': t_cg,
                      't_advance': t_advance}
        return out

    def __del__(self):
        try:
            if self._pool is not None:
                self._pool.shutdown(wait=False)
                self._pool = None
            if self.h and nat._lib is not None:
                nat._lib.sn_vcg_destroy(self.h)
                self.h = None
        except Exception:
            pass
