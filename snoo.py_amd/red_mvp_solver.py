"""Reduced-vector-potential Newton solver (snoopy/red_mvp_solver.py:14-224).

``RedMVPSolver`` keeps the reference constructor, attributes and ``solve``
semantics: source fields from the coils at the iron quadrature points; Newton
with b = Jx - Kx + rhs; Dirichlet DOFs removed; Jacobi-PCG with scipy's rule
(atol = tolerance, rtol = 1e-5, x0 = x_n, maxiter = 10 n by default); the
iterate replaced by the CG solution; stop when max|h| < tolerance max|x|.
Every array stays on the device between steps; the host sees one scalar test
per Newton step and one stop flag per chunk of CG iterations.

``solve_newton`` is the configuration solver: inexact Newton on the FP64
residual R(x) = K(x)x - rhs(x) (matrix-free, per element), stopped on
||R(x)|| <= rtol ||R(0)||, with the linear solves in FP64 or (precision
"mixed") in FP32 with FP64 dot products.  Each mixed step is one step of
iterative refinement: the FP64 residual is recomputed at the new iterate.
"""
import ctypes
import math
import time

import numpy as np

from . import _native as nat
from .curl_curl_assembler import CurlCurlAssembler
from .device import as_dev_f64, device
from .materials import NU0, ConstantReluctance
from .mesh_tools import get_edge_boundary_dofs
from .sources import source_field_device


class RedMVPSolver:
    """Reference-compatible reduced vector potential solver."""

    def __init__(self, gmsh_model, coil_list, iron_tags, air_tags, reluctances,
                 max_newton_iterations=25, element_order=1, quad_order=6, verbose=True):
        self.iron_tags = list(iron_tags)
        self.air_tags = list(air_tags)
        self.coil_list = list(coil_list)
        self.domain_tags = self.iron_tags + self.air_tags
        material_properties = list(reluctances) + [ConstantReluctance(NU0)] * len(self.air_tags)
        self.max_newton_iterations = max_newton_iterations
        self.element_order = element_order
        self.quad_order = quad_order
        self.verbose = verbose
        self.model = gmsh_model
        self.dirichlet_dofs = get_edge_boundary_dofs(gmsh_model, ['Dirichlet Boundary'],
                                                     element_order)
        self.curl_curl_factory = CurlCurlAssembler(gmsh_model.mesh, self.domain_tags,
                                                   material_properties, element_order)
        self.num_dofs_mvp = self.curl_curl_factory.num_dofs
        self._src = None
        self._mask_dev = None
        self.stats = {}

    def get_global_ids(self):
        return self.curl_curl_factory.global_ids

    def make_boundary_mask(self):
        """Boolean mask, False at the Dirichlet DOFs (red_mvp_solver.py:105-119)."""
        mask = np.ones((self.num_dofs_mvp,), dtype=np.bool_)
        mask[np.asarray(self.dirichlet_dofs, dtype=np.int64) - 1] = False
        return mask

    # ------------------------------------------------------------------ setup
    def source_fields_device(self, wait=True):
        """B_s at the quadrature points of each iron domain; None for air
        (red_mvp_solver.py:149-164).  ``wait=False`` returns as soon as the
        kernels are queued on the current stream (stats['t_source'] is then
        the time to queue them)."""
        if self._src is None:
            t0 = time.perf_counter()
            n_iron = len(self.iron_tags)
            pts = self.curl_curl_factory.get_quadrature_points_device(self.quad_order,
                                                                      list(range(n_iron)))
            src = [source_field_device(self.coil_list, p).reshape(-1) for p in pts]
            src += [None] * len(self.air_tags)
            if wait:
                import torch
                torch.cuda.current_stream().synchronize()
            self.stats['t_source'] = time.perf_counter() - t0
            self.stats['num_targets'] = int(sum(p.shape[0] for p in pts))
            self._src = src
        return self._src

    def mask_device(self):
        if self._mask_dev is None:
            import torch
            self._mask_dev = torch.from_numpy(self.make_boundary_mask().astype(np.uint8)).to(device())
        return self._mask_dev

    def _log(self, *a):
        if self.verbose:
            print(*a)

    # ------------------------------------------------------------------ solve
    def solve(self, tolerance=1e-4, x_0=np.zeros((0,)), maxiter=-1, apply_gauge=False):
        """Newton iterations of the reference (red_mvp_solver.py:121-224)."""
        import torch
        lib = nat.load()
        asm = self.curl_curl_factory
        n = self.num_dofs_mvp
        if len(x_0) == 0:
            x_0 = np.zeros((n,))
        src = self.source_fields_device()
        mask = self.mask_device()
        if apply_gauge:
            from .mesh_tools import get_cotree_dofs
            tree, _ = get_cotree_dofs(asm.tet_mesh)
            mask = mask.clone()
            mask[torch.from_numpy(np.asarray(tree, dtype=np.int64)).to(device())] = 0
        d = asm._device_state()
        N = d['num_el']
        locJ = torch.empty(21 * N, dtype=torch.float64, device=device())
        loc6 = torch.empty(6 * N, dtype=torch.float64, device=device())
        J = torch.empty(d['nnz'], dtype=torch.float64, device=device())
        b = torch.empty(n, dtype=torch.float64, device=device())
        x_n = as_dev_f64(x_0).clone()
        cg = _CG(asm, 64)
        self._log('***********************')
        self._log('start Newton iterations')
        it_hist = []
        x_np1 = x_n
        for i in range(self.max_newton_iterations):
            asm.newton_local_device(x_n, self.quad_order, src, 0, False, locJ, loc6)
            nat.check(lib.sn_csr_gather_one(d['pattern'], 0, nat.ptr(locJ), nat.ptr(J),
                                            nat.stream_handle()))
            nat.check(lib.sn_dof_gather_f64(d['pattern'], nat.ptr(loc6), nat.ptr(b),
                                            nat.stream_handle()))
            if _absmax(b) < 1e-14:
                x_np1 = 0.0 * x_n
                self._log('  zero excitation!')
                break
            x_np1 = x_n.clone()
            iters, _, rc = cg.solve(J, mask, b, x_np1, rtol=1e-5, atol=tolerance, maxiter=maxiter)
            it_hist.append(iters)
            x_np1[mask == 0] = 0.0
            h = x_np1 - x_n
            hmax, xmax = _absmax(h), _absmax(x_np1)
            self._log(f"  step {i}")
            self._log(f"    max relative increment {hmax / xmax if xmax > 0 else float('inf'):.3e}")
            if hmax < tolerance * xmax:
                self._log('  tolerance reached!')
                break
            x_n = x_np1
        self.stats['cg_iterations'] = it_hist
        return x_np1.cpu().numpy()

    def solve_newton(self, precision='fp64', rtol=1e-8, x0=None, max_newton=40, eta_max=0.1,
                     eta_min=None, lanes_per_row=8, chunk=32, max_cg=-1, line_search=True,
                     log=None, precond='jacobi', tight='element', jacobian=None, eta_first=None,
                     eta_margin=0.9, plan_close=True):
        """Inexact Newton on R(x) = K(x)x - rhs(x) to ||R|| <= rtol ||R(0)||.

        Returns the device solution (float64) and fills ``self.stats`` with the
        residual history, per-step CG iteration counts and timings.
        ``jacobian`` and ``tight`` (mixed precision): the operator of the
        float32 CG, "element" (default) or "csr", and with "csr" the solver of
        the steps too tight for it, "element" or "fp64" (NewtonDriver)."""
        drv = self.newton_driver(precision, lanes_per_row, chunk, precond, tight, jacobian)
        drv.retarget(self.source_fields_device())
        return drv.solve(x0=x0, rtol=rtol, max_newton=max_newton, eta_max=eta_max,
                         eta_min=eta_min, max_cg=max_cg, line_search=line_search,
                         stats=self.stats, log=log, eta_first=eta_first, eta_margin=eta_margin,
                         plan_close=plan_close)

    def newton_driver(self, precision='fp64', lanes_per_row=8, chunk=32, precond='jacobi',
                      tight='element', jacobian=None):
        if jacobian is None:
            jacobian = 'element' if precision == 'mixed' else 'csr'
        if lanes_per_row == 0:
            lanes_per_row = self.tuned_lanes(precision) if jacobian == 'csr' else 8
        key = ('drv', precision, lanes_per_row, chunk, precond, tight, jacobian)
        if key not in self.__dict__.setdefault('_drivers', {}):
            self._drivers[key] = NewtonDriver(self.curl_curl_factory, self.source_fields_device(),
                                              self.mask_device(), self.quad_order, precision,
                                              lanes_per_row, chunk, precond, tight, jacobian)
        return self._drivers[key]


def _tuned_lanes(solver, precision):
    """Lanes per CSR row of the SpMV, timed on the Jacobian at x = 0."""
    import torch
    from .autotune import tune_lanes
    asm = solver.curl_curl_factory
    d = asm._device_state()
    drv = solver.newton_driver(precision, 8, 32, jacobian='csr')
    drv.retarget(solver.source_fields_device())
    drv.residual(torch.zeros(asm.num_dofs, dtype=torch.float64, device=device()))
    best, times = tune_lanes(d['pattern'], asm.num_dofs, d['nnz'], drv.J, drv.mask,
                             32 if precision == 'mixed' else 64)
    solver.stats['autotune_lanes'] = {'best': best, 'ms': times}
    return best


RedMVPSolver.tuned_lanes = _tuned_lanes


def _tuned_element_grids(solver, lanes_per_row=8):
    """Launch shapes of the mixed solver's element-operator CG, timed on
    the Jacobian at x = 0 with the auxiliary-space preconditioner; applied
    to this solver's driver and to drivers made later for meshes of the
    same size."""
    import torch
    from .autotune import tune_element_grids
    asm = solver.curl_curl_factory
    drv = solver.newton_driver('mixed', lanes_per_row, 32, 'ams')
    drv.retarget(solver.source_fields_device())
    drv.residual(torch.zeros(asm.num_dofs, dtype=torch.float64, device=device()))
    drv._refresh_preconditioner()
    best, times = tune_element_grids(drv)
    solver.stats['autotune_grids'] = {'best': list(best),
                                      'ms_per_iteration': {'%d/%d/%d' % k: v
                                                           for k, v in times.items()}}
    return best


RedMVPSolver.tuned_element_grids = _tuned_element_grids


def next_forcing(rn, rn_prev, r0, rtol, eta, eta_min, eta_max, eta_margin, plan=None):
    """Forcing term of the next Newton step: Eisenstat-Walker choice 2,
    0.9 (|R_k| / |R_k-1|)^2 within [eta_min, eta_max] (``eta`` itself on the
    first step, rn_prev None), and never below ``eta_margin`` times the
    distance still to go, rtol |R(0)| / |R_k| (capped at eta_max).  With
    ``plan``: once |R_k| <= plan |R(0)| (Newton in its linear regime) the
    step takes the planned forcing term instead of Eisenstat-Walker's."""
    if rn_prev is not None:
        if plan is not None and rn <= plan * r0:
            eta = planned_forcing(rn / r0, rtol, eta_min, eta_max, eta_margin)
        else:
            eta = min(eta_max, max(0.9 * (rn / rn_prev) ** 2, eta_min))
    return max(eta, min(eta_max, eta_margin * rtol * r0 / rn))


def planned_forcing(rel, rtol, eta_min, eta_max, margin=1.0):
    """Forcing term of a Newton step from a residual of ``rel`` |R(0)|: the
    fewest steps that reach ``rtol`` when each removes at most a factor
    1/eta_min, and the equal factor per step that does it (a start at
    1.6e-3 |R(0)|, rtol 1e-8, eta_min 1e-3: 2 steps of 2.5e-3); a single
    step asks for ``margin`` times the distance."""
    digits = math.log(rel / rtol)
    if digits <= 0.0:
        return eta_max
    steps = max(1, math.ceil(digits / math.log(1.0 / eta_min) - 1e-9))
    if steps == 1:
        return min(eta_max, margin * math.exp(-digits))
    return min(eta_max, max(eta_min, math.exp(-digits / steps)))


def _absmax(t):
    out = ctypes.c_double()
    nat.check(nat.load().sn_absmax_f64(int(t.numel()), nat.ptr(t), None, ctypes.byref(out),
                                       nat.stream_handle()))
    return out.value


class _CG:
    """Owning wrapper of an sn_cg workspace."""

    def __init__(self, asm, precision, lanes_per_row=8, chunk=32, element_operator=False):
        self.h = ctypes.c_void_p()
        nat.check(nat.load().sn_cg_create(ctypes.byref(self.h), asm.pattern, int(precision),
                                          int(lanes_per_row), int(chunk), nat.stream_handle()))
        if element_operator:
            # solves take the element matrices of sn_element_newton mode 2
            d = asm._device_state()
            nat.check(nat.load().sn_cg_set_element_operator(
                self.h, nat.ptr(d['gids']), nat.ptr(d['orient']), 1, nat.stream_handle()))

    def solve(self, vals, mask, b, x, rtol, atol, maxiter=-1, allow_cap=True):
        it = ctypes.c_int()
        rn = ctypes.c_double()
        rc = nat.load().sn_cg_solve(self.h, nat.ptr(vals), nat.ptr(mask), nat.ptr(b), nat.ptr(x),
                                    float(rtol), float(atol), int(maxiter), ctypes.byref(it),
                                    ctypes.byref(rn), nat.stream_handle())
        if rc == nat.SN_ERR_NOT_CONVERGED and allow_cap:
            # scipy returns its last iterate with info = maxiter; so does this
            return it.value, rn.value, rc
        nat.check(rc)
        return it.value, rn.value, rc

    def __del__(self):
        try:
            if self.h and nat._lib is not None:
                nat._lib.sn_cg_destroy(self.h)
        except Exception:
            pass


class NewtonDriver:
    """Device-resident inexact Newton for one mesh, precision and source field.

    precision "fp64": FP64 Jacobian, FP64 PCG solving J dx = -R.
    precision "mixed": FP32 Jacobian (element arithmetic in FP64, values
    rounded once), FP32 PCG with FP64 dot products, x += dx in FP64; R(x) is
    always FP64 (element-local K_e x_e - rhs_e, gathered in fixed order).
    precond "jacobi": the reference's diagonal preconditioner; "ams": the
    auxiliary-space preconditioner (ams.py), refreshed with the element
    reluctivities of every Jacobian.
    jacobian (mixed only): "element" (default): the Jacobian of every step as
    float32 element matrices in reference-curl form and the float32 CG on the
    element operator (sn_cg_set_element_operator), which keeps the gradient
    null space of J exact under rounding; "csr": the float32 CSR values of the
    assembled J (a rounding that breaks the null space and can stall the CG
    at tight forcing terms), with ``tight`` choosing the solver of the steps it
    cannot reach: "element", or "fp64" (float64 Jacobian and CG).  A float32
    step that stalls at its iteration cap is redone in float64.
    """

    def __init__(self, asm, src_dev, mask_dev, quad_order, precision='fp64', lanes_per_row=8,
                 chunk=32, precond='jacobi', tight='element', jacobian=None):
        import torch
        if precision not in ('fp64', 'mixed'):
            raise ValueError('precision must be "fp64" or "mixed"')
        if precond not in ('jacobi', 'ams'):
            raise ValueError('precond must be "jacobi" or "ams"')
        if tight not in ('element', 'fp64'):
            raise ValueError('tight must be "element" or "fp64"')
        if jacobian is None:
            jacobian = 'element' if precision == 'mixed' else 'csr'
        if jacobian not in ('csr', 'element'):
            raise ValueError('jacobian must be "csr" or "element"')
        if jacobian == 'element' and precision != 'mixed':
            raise ValueError('the element operator is a float32 operator (precision "mixed")')
        self.precond = precond
        self.tight = tight
        self.jacobian = jacobian
        self.el_all = jacobian == 'element'
        self.lanes_per_row, self.chunk = lanes_per_row, chunk
        self.asm, self.src, self.mask, self.quad_order = asm, src_dev, mask_dev, quad_order
        self.precision = precision
        self.f32 = precision == 'mixed'
        d = asm._device_state()
        self.d = d
        N, n, nnz = d['num_el'], asm.num_dofs, d['nnz']
        jt = torch.float32 if self.f32 else torch.float64
        dev = device()
        self.loc6 = torch.empty(6 * N, dtype=torch.float64, device=dev)
        self.R = torch.empty(n, dtype=torch.float64, device=dev)
        self.b = torch.empty(n, dtype=torch.float64, device=dev)
        self.dx = torch.empty(n, dtype=torch.float64, device=dev)
        # float32 CG on the element operator: every step with jacobian
        # "element", the tight steps of "csr" (built on first use there)
        self.cg_el = None
        self.Mr = None
        if self.el_all:
            self.locJ = self.J = self.cg = None
            self.Mr = torch.empty(6 * N, dtype=torch.float32, device=dev)
            self.cg_el = _CG(asm, 32, lanes_per_row, chunk, element_operator=True)
            self._tuned_grids(self.cg_el)
        else:
            self.locJ = torch.empty(21 * N, dtype=jt, device=dev)
            self.J = torch.empty(nnz, dtype=jt, device=dev)
            self.cg = _CG(asm, 32 if self.f32 else 64, lanes_per_row, chunk)
        # float64 Jacobian and CG of the mixed solver's fallback (built on first use)
        self.cg_safe = None
        self._J64 = None
        self.eta_fp64 = 0.0   # steps with eta <= this go straight to float64
        self._r0 = None
        self.nu_el = torch.empty(N, dtype=torch.float32, device=dev) if precond == 'ams' else None
        self.ams = None
        self.ams_options = {}   # omega, weight, theta, coarse_max of AmsPreconditioner
        if precond == 'ams' and self.f32:
            # with the auxiliary-space preconditioner the float32 CG on the
            # rounded CSR matrix reaches a few 1e-3 before the rounding of J
            # stalls it; tighter forcing terms go to the tight solver from
            # the start
            self.eta_fp64 = 4.9e-3

    def _tuned_grids(self, cg=None):
        """Launch shapes tuned for a mesh of this size (autotune), if any:
        the element-operator workspace's grids, the preconditioner's lanes."""
        from .autotune import cached_element_grids
        best = cached_element_grids(self.d['num_el'], self.asm.num_dofs)
        if best is None:
            return
        if cg is not None:
            nat.check(nat.load().sn_cg_set_grids(cg.h, best[0], best[1], nat.stream_handle()))
        elif self.ams is not None and self.el_all:
            nat.check(nat.load().sn_ams_set_lanes(self.ams.handle, best[2]))

    def _refresh_preconditioner(self):
        """Auxiliary-space preconditioner for the Jacobian of the last
        residual() call: built on first use, values refreshed after."""
        if self.precond != 'ams':
            return
        if self.ams is None:
            from .ams import AmsPreconditioner
            free = self.mask.cpu().numpy().astype(bool)
            self.ams = AmsPreconditioner(self.asm, free, self.nu_el, **self.ams_options)
            self._tuned_grids()
            for cg in (self.cg, self.cg_safe, self.cg_el):
                if cg is not None:
                    nat.check(nat.load().sn_cg_set_ams(cg.h, self.ams.handle))
        else:
            self.ams.update(self.nu_el)

    def retarget(self, src_dev):
        """Use new source fields / materials / node coordinates (drops the
        cached ||R(0)|| when any of them changed)."""
        key = (id(src_dev), tuple(id(m) for m in self.asm.material_list))
        if key != getattr(self, '_key', None):
            self.src = src_dev
            self._r0 = None
            self._key = key

    def residual_masked(self, x_dev):
        """Zero x's Dirichlet DOFs in place, then residual(x); the result is
        remembered with x, so that solve(x0=x) while x is unchanged and no
        other residual was formed reuses it (and the Jacobian at x) instead
        of forming it again."""
        x_dev[self.mask == 0] = 0.0
        rn = self.residual(x_dev)
        self._last = (x_dev, x_dev._version, rn)
        return rn

    def residual(self, x_dev, norm2_dev=None):
        """R(x) (FP64, device) and the Jacobian at x; returns ||R_free||_2.
        With ``norm2_dev`` (a float64 device scalar) nothing waits: the
        kernels are queued, ||R_free||^2 lands in norm2_dev, and None is
        returned (the square root of the value read back is the same number
        the waiting form returns)."""
        self._last = None
        lib = nat.load()
        if self.el_all:
            self.asm.newton_local_device(x_dev, self.quad_order, self.src, 2, True, self.Mr,
                                         self.loc6, nu_el=self.nu_el)
        else:
            self.asm.newton_local_device(x_dev, self.quad_order, self.src, 1, self.f32,
                                         self.locJ, self.loc6, nu_el=self.nu_el)
            nat.check(lib.sn_csr_gather_one(self.d['pattern'], int(self.f32), nat.ptr(self.locJ),
                                            nat.ptr(self.J), nat.stream_handle()))
        nat.check(lib.sn_dof_gather_f64(self.d['pattern'], nat.ptr(self.loc6), nat.ptr(self.R),
                                        nat.stream_handle()))
        if norm2_dev is not None:
            nat.check(lib.sn_dot_f64_dev(int(self.R.numel()), nat.ptr(self.R), nat.ptr(self.R),
                                         nat.ptr(self.mask), nat.ptr(norm2_dev),
                                         nat.stream_handle()))
            return None
        out = ctypes.c_double()
        nat.check(lib.sn_dot_f64(int(self.R.numel()), nat.ptr(self.R), nat.ptr(self.R),
                                 nat.ptr(self.mask), ctypes.byref(out), nat.stream_handle()))
        return float(np.sqrt(out.value))

    def _fp64_step(self, x, eta, max_cg):
        """One Newton step of the mixed solver in float64: the float64
        Jacobian at x (element kernel in float64 output mode) and float64 CG."""
        import torch
        lib = nat.load()
        if self._J64 is None:
            N = self.d['num_el']
            self._locJ64 = torch.empty(21 * N, dtype=torch.float64, device=device())
            self._J64 = torch.empty(self.d['nnz'], dtype=torch.float64, device=device())
            self._loc6b = torch.empty_like(self.loc6)
            self.cg_safe = _CG(self.asm, 64, self.lanes_per_row, self.chunk)
            if self.ams is not None:
                nat.check(lib.sn_cg_set_ams(self.cg_safe.h, self.ams.handle))
        self.asm.newton_local_device(x, self.quad_order, self.src, 1, False, self._locJ64,
                                     self._loc6b)
        nat.check(lib.sn_csr_gather_one(self.d['pattern'], 0, nat.ptr(self._locJ64),
                                        nat.ptr(self._J64), nat.stream_handle()))
        self.dx.zero_()
        its, _, _ = self.cg_safe.solve(self._J64, self.mask, self.b, self.dx, rtol=eta, atol=0.0,
                                       maxiter=max_cg)
        return its, x + self.dx

    def _element_step(self, x, eta, maxiter):
        """One Newton step of the mixed solver on the element operator: the
        Jacobian at x as float32 element matrices (sn_element_newton mode 2;
        residual() has formed them already with jacobian "element") and the
        float32 CG multiplying with them.  Returns (iterations, status, x + dx)."""
        import torch
        if self.cg_el is None:
            N = self.d['num_el']
            self.Mr = torch.empty(6 * N, dtype=torch.float32, device=device())
            self.cg_el = _CG(self.asm, 32, self.lanes_per_row, self.chunk, element_operator=True)
            self._tuned_grids(self.cg_el)
            if self.ams is not None:
                nat.check(nat.load().sn_cg_set_ams(self.cg_el.h, self.ams.handle))
        if not self.el_all:
            if getattr(self, '_loc6c', None) is None:
                self._loc6c = torch.empty_like(self.loc6)
            self.asm.newton_local_device(x, self.quad_order, self.src, 2, True, self.Mr,
                                         self._loc6c)
        x_try = x.clone()
        its, _, rc = self.cg_el.solve(self.Mr, self.mask, self.b, x_try, rtol=eta, atol=0.0,
                                      maxiter=maxiter)
        return its, rc, x_try

    def r0_queued(self, norm2_dev):
        """||R(0)||^2 queued into the float64 device scalar norm2_dev.  At
        x = 0 the elements of a domain without source fields see B = 0 and
        add nothing (their local values are set to zero instead of being
        formed), so only the domains with sources run the element kernel.
        The element matrices and reluctivities of the other domains are left
        as they were: the next residual() forms them."""
        import torch
        n = self.asm.num_dofs
        z = torch.zeros(n, dtype=torch.float64, device=device())
        with_src = [i for i, b in enumerate(self.src) if b is not None]
        if not self.el_all or len(with_src) == len(self.src):
            self.residual(z, norm2_dev)
            return
        self._last = None
        d = self.asm._device_state()
        offs, N = d['offsets'], d['num_el']
        loc = self.loc6.view(6, N)
        for i in range(len(self.src)):
            if self.src[i] is None:
                loc[:, int(offs[i]):int(offs[i + 1])].zero_()
        lib = nat.load()
        self.asm.newton_local_device(z, self.quad_order, self.src, 2, True, self.Mr, self.loc6,
                                     domains=with_src, nu_el=self.nu_el)
        nat.check(lib.sn_dof_gather_f64(self.d['pattern'], nat.ptr(self.loc6), nat.ptr(self.R),
                                        nat.stream_handle()))
        nat.check(lib.sn_dot_f64_dev(int(self.R.numel()), nat.ptr(self.R), nat.ptr(self.R),
                                     nat.ptr(self.mask), nat.ptr(norm2_dev), nat.stream_handle()))

    def r0(self):
        """||R(0)|| (the norm of the source term), cached."""
        if self._r0 is None:
            import torch
            z = torch.zeros(self.asm.num_dofs, dtype=torch.float64, device=device())
            self._r0 = self.residual(z)
        return self._r0

    def solve(self, x0=None, rtol=1e-8, max_newton=40, eta_max=0.1, eta_min=None, max_cg=-1,
              line_search=True, stats=None, log=None, stall_steps=3, eta_first=None,
              eta_margin=0.9, close_start=1e-2, plan_close=True):
        """Newton steps until ||R|| <= rtol ||R(0)||, ``max_newton`` steps, or
        ``stall_steps`` consecutive stalled steps: full steps that reduce ||R||
        by less than 10%, or steps that no length of the line search makes
        reduce it (the floating-point floor of the residual).  ``max_cg`` < 0 caps each
        linear solve at min(10 n, 50000) iterations; a float32 solve is cut at
        4x the previous step's count (at least 1500, 300 with the auxiliary-
        space preconditioner) and, if it stalls there, the step is redone by
        the tight solver (float32 on the element operator, then float64 if
        that stalls too), as are later steps with a forcing term at least as
        small.

        Forcing terms: Eisenstat-Walker choice 2 (0.9 (|R_k|/|R_k-1|)^2 within
        [eta_min, eta_max]), never below ``eta_margin`` times the distance
        still to go, so the step that can reach the target asks for just
        that.  The first step takes ``eta_first``: by default eta_max, except
        from a start x0 whose residual is already at most ``close_start``
        |R(0)|, where Newton is in its linear regime; that step takes the
        forcing term of the fewest equal steps to the target
        (``planned_forcing``).  With ``plan_close`` every later step whose
        residual is at most ``close_start`` |R(0)| is planned the same way
        (measured on one MI355X: 1M-tet dipole from zero 31.6 -> 30.7 ms, the
        6M-tet chain 559 -> 484 ms, the 64 variants of 20k-335k tets 9 % more
        CG iterations in the same time)."""
        import torch
        stats = stats if stats is not None else {}
        if eta_min is None:
            # float32 CG cannot be asked for much more than 3 digits per solve
            eta_min = 1e-3 if self.f32 else 1e-6
        if max_cg < 0:
            max_cg = min(10 * self.asm.num_dofs, 50000)
        n = self.asm.num_dofs
        r0 = self.r0()
        last = getattr(self, '_last', None)
        reuse = last is not None and x0 is not None and last[0] is x0 and \
            x0._version == last[1]
        x = torch.zeros(n, dtype=torch.float64, device=device()) if x0 is None else \
            as_dev_f64(x0).clone()
        x[self.mask == 0] = 0.0
        t0 = time.perf_counter()
        # R, the Jacobian and the preconditioner's reluctivities at x0 are
        # still those of residual_masked(x0)
        rn = last[2] if reuse else self.residual(x)
        hist, cg_its, etas, modes = [rn / r0], [], [], []
        eta = eta_max if eta_first is None else eta_first
        if eta_first is None and x0 is not None and 0.0 < rn <= close_start * r0:
            eta = planned_forcing(rn / r0, rtol, eta_min, eta_max)
        rn_prev = None
        k = 0
        stalls = 0
        converged = rn <= rtol * r0
        while not converged and k < max_newton and stalls < stall_steps:
            # forcing term (Eisenstat-Walker choice 2 with safeguards), never
            # asking for more than the remaining distance to the target
            eta = next_forcing(rn, rn_prev, r0, rtol, eta, eta_min, eta_max, eta_margin,
                               close_start if plan_close else None)
            etas.append(eta)
            torch.neg(self.R, out=self.b)
            self._refresh_preconditioner()
            mode = 'fp64'
            if self.f32:
                floor = 300 if self.precond == 'ams' else 1500
                cap32 = min(max(floor, 4 * max(cg_its[-1] if cg_its else 0, 1)), max_cg)
                its, tight, mode = 0, True, ''
                if self.el_all:
                    its, rc, x_try = self._element_step(x, eta, cap32)
                    mode = 'fp32 element'
                    tight = rc == nat.SN_ERR_NOT_CONVERGED
                    if tight:
                        mode += ' stalled -> '
                elif eta > self.eta_fp64:
                    x_try = x.clone()
                    its, _, rc = self.cg.solve(self.J, self.mask, self.b, x_try, rtol=eta,
                                               atol=0.0, maxiter=cap32)
                    tight = rc == nat.SN_ERR_NOT_CONVERGED
                    mode = 'fp32 stalled -> ' if tight else 'fp32'
                    if tight:
                        self.eta_fp64 = max(self.eta_fp64, eta)
                if tight and self.tight == 'element' and not self.el_all:
                    more, rc, x_try = self._element_step(x, eta, cap32)
                    its += more
                    mode += 'fp32 element'
                    tight = rc == nat.SN_ERR_NOT_CONVERGED
                    if tight:
                        mode += ' stalled -> '
                if tight:
                    more, x_try = self._fp64_step(x, eta, max_cg)
                    its += more
                    mode += 'fp64'
            else:
                self.dx.zero_()
                its, _, _ = self.cg.solve(self.J, self.mask, self.b, self.dx, rtol=eta, atol=0.0,
                                          maxiter=max_cg)
                x_try = x + self.dx
            modes.append(mode)
            cg_its.append(its)
            rn_try = self.residual(x_try)
            step = 1.0
            if line_search:
                dx = x_try - x
                tries = 0
                while not (rn_try < rn) and tries < 6:
                    step *= 0.5
                    x_try = x + step * dx
                    rn_try = self.residual(x_try)
                    tries += 1
            # stalled: no step length lowers ||R||, or the full step gains
            # less than 10%; damped steps that make progress (the first
            # steps into saturation) are not stalls
            stalled = rn_try >= rn or (step == 1.0 and rn_try > 0.9 * rn)
            stalls = stalls + 1 if stalled else 0
            rn_prev, rn, x = rn, rn_try, x_try
            hist.append(rn / r0)
            k += 1
            converged = rn <= rtol * r0
            if log is not None:
                log('  newton %d: |R|/|R0| = %.3e, cg %d, eta %.1e, step %.3g, %.3f s'
                    % (k, rn / r0, its, eta, step, time.perf_counter() - t0))
        torch.cuda.current_stream().synchronize()
        stats.update({'residual_history': hist, 'cg_iterations': cg_its, 'etas': etas,
                      'cg_modes': modes,
                      'newton_steps': k, 'converged': converged, 'r0': r0,
                      't_newton': time.perf_counter() - t0, 'precision': self.precision,
                      'precond': self.precond})
        if self.ams is not None:
            stats['ams_levels'] = self.ams.plan.sizes()
            stats['ams_setup_s'] = self.ams.plan.t_setup
        return x


class RedMVPSolverDiluted(RedMVPSolver):
    """Solver with diluted steel (snoopy/red_mvp_solver_diluted.py:13-229):
    the constructor takes one dilution per iron domain (air is never
    diluted); ``solve`` and ``solve_newton`` are those of RedMVPSolver on the
    diluted assembler."""

    def __init__(self, gmsh_model, coil_list, iron_tags, air_tags, reluctances, dilutions,
                 max_newton_iterations=25, element_order=1, quad_order=6, verbose=True):
        super().__init__(gmsh_model, coil_list, iron_tags, air_tags, reluctances,
                         max_newton_iterations, element_order, quad_order, verbose)
        from .curl_curl_assembler import CurlCurlAssemblerDiluted
        self.dilutions = list(dilutions) + [None] * len(self.air_tags)
        self.curl_curl_factory = CurlCurlAssemblerDiluted(
            gmsh_model.mesh, self.domain_tags, self.curl_curl_factory.material_list,
            self.dilutions, element_order)
