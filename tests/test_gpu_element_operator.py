"""GPU: the element operator (Jacobian as a sum of element products in
reference-curl form, sn_element_newton mode 2 + sn_cg_set_element_operator).

Tolerances: products with float64 element matrices within 1e-12 of the
oracle's assembled J = K + dK (only the summation order differs); float32
element matrices within 1e-6; CG solutions compared through what they
determine (the float64 residual and B = curl A).
"""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


@pytest.fixture(scope='module')
def dipole():
    from snoopy_amd.magnets import Dipole
    print('case: dipole h=0.12 (element operator)', flush=True)
    return Dipole(h=0.12)


@pytest.fixture(scope='module')
def solver(dipole):
    return dipole.solver(quad_order=2, verbose=False)


@pytest.fixture(scope='module')
def oprob(solver):
    src = [s.cpu().numpy().reshape(-1, 3) if s is not None else None
           for s in solver.source_fields_device()]
    return orc.OracleProblem(solver, source_fields=src)


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


def _mode2(solver, x, f32):
    torch = _torch()
    asm = solver.curl_curl_factory
    N = asm._device_state()['num_el']
    Mr = torch.empty(6 * N, dtype=torch.float32 if f32 else torch.float64, device='cuda')
    loc6 = torch.empty(6 * N, dtype=torch.float64, device='cuda')
    asm.newton_local_device(x, 2, solver.source_fields_device(), 2, f32, Mr, loc6)
    return Mr, loc6


def _apply(solver, Mr64, v, mask=None):
    torch = _torch()
    from snoopy_amd import _native as nat
    asm = solver.curl_curl_factory
    d = asm._device_state()
    work = torch.empty(6 * d['num_el'], dtype=torch.float64, device='cuda')
    y = torch.empty(asm.num_dofs, dtype=torch.float64, device='cuda')
    nat.check(nat.load().sn_element_apply_f64(
        d['pattern'], nat.ptr(d['gids']), nat.ptr(d['orient']), nat.ptr(Mr64),
        nat.ptr(mask) if mask is not None else None, nat.ptr(v), nat.ptr(work), nat.ptr(y),
        nat.stream_handle()))
    return y


def test_mode2_products_match_oracle_jacobian(solver, oprob):
    torch = _torch()
    from snoopy_amd import _native as nat
    asm = solver.curl_curl_factory
    d = asm._device_state()
    rng = np.random.default_rng(31)
    x = rng.standard_normal(asm.num_dofs) * 5e-3
    xd = torch.from_numpy(x).cuda()
    K, dK, rhs = oprob.assemble(x)
    Jo = (dK + K).tocsr()
    v = rng.standard_normal(asm.num_dofs)
    want = Jo @ v
    for f32, tol in ((False, 1e-12), (True, 1e-6)):
        Mr, loc6 = _mode2(solver, xd, f32)
        y = _apply(solver, Mr.double(), torch.from_numpy(v).cuda()).cpu().numpy()
        assert _rel(y, want) < tol, (f32, _rel(y, want))
        # the residual output is mode 1's
        R = torch.empty(asm.num_dofs, dtype=torch.float64, device='cuda')
        nat.check(nat.load().sn_dof_gather_f64(d['pattern'], nat.ptr(loc6), nat.ptr(R),
                                               nat.stream_handle()))
        assert _rel(R.cpu().numpy(), K @ x - rhs) < 1e-11
    # masked product: rows and columns of the Dirichlet DOFs removed
    mask = solver.mask_device()
    m = mask.cpu().numpy().astype(bool)
    Mr, _ = _mode2(solver, xd, False)
    y = _apply(solver, Mr, torch.from_numpy(v).cuda(), mask).cpu().numpy()
    wm = np.where(m, Jo @ np.where(m, v, 0.0), 0.0)
    assert _rel(y, wm) < 1e-12


def test_element_operator_keeps_the_gradient_null_space(solver):
    """Rounded to float32, the element matrices still annihilate gradients to
    float64 rounding; the rounded CSR values do not."""
    torch = _torch()
    from snoopy_amd import _native as nat
    asm = solver.curl_curl_factory
    d = asm._device_state()
    tm = asm.tet_mesh
    mask = solver.mask_device()
    m = mask.cpu().numpy().astype(bool)
    edges = np.asarray(tm.edges).reshape(-1, 2)
    rng = np.random.default_rng(5)
    phi = rng.standard_normal(len(tm.nodes))
    # phi = 0 at the end points of Dirichlet edges: the gradient then has no
    # Dirichlet component and stays a gradient on the free DOFs
    phi[edges[~m].ravel()] = 0.0
    g = phi[edges[:, 1]] - phi[edges[:, 0]]
    gd = torch.from_numpy(g).cuda()
    x = torch.from_numpy(rng.standard_normal(asm.num_dofs) * 5e-3).cuda()
    Mr, _ = _mode2(solver, x, True)
    v = torch.from_numpy(rng.standard_normal(asm.num_dofs)).cuda()
    scale = float(_apply(solver, Mr.double(), v, mask).norm() / v[mask == 1].norm())
    el = float(_apply(solver, Mr.double(), gd, mask).norm() / gd.norm()) / scale
    # the same Jacobian as float32 CSR values, rounded after assembly
    locJ = torch.empty(21 * d['num_el'], dtype=torch.float32, device='cuda')
    asm.newton_local_device(x, 2, solver.source_fields_device(), 1, True, locJ,
                            torch.empty(6 * d['num_el'], dtype=torch.float64, device='cuda'))
    J32 = torch.empty(d['nnz'], dtype=torch.float32, device='cuda')
    nat.check(nat.load().sn_csr_gather_one(d['pattern'], 1, nat.ptr(locJ), nat.ptr(J32),
                                           nat.stream_handle()))
    y = torch.empty_like(gd)
    nat.check(nat.load().sn_spmv_f64(d['pattern'], nat.ptr(J32.double()), nat.ptr(mask),
                                     nat.ptr(gd), nat.ptr(y), nat.stream_handle()))
    csr = float(y.norm() / gd.norm()) / scale
    print('gradient defect: element %.2e, rounded CSR %.2e' % (el, csr))
    assert el < 1e-13
    assert csr > 100 * el


@pytest.mark.parametrize('precond', ['jacobi', 'ams'])
def test_element_cg_matches_fp64_csr_cg(solver, precond):
    torch = _torch()
    from snoopy_amd import _native as nat
    from snoopy_amd.red_mvp_solver import _CG
    asm = solver.curl_curl_factory
    d = asm._device_state()
    mask = solver.mask_device()
    drv = solver.newton_driver('mixed', 8, 32, precond)
    drv.retarget(solver.source_fields_device())
    x0 = torch.zeros(asm.num_dofs, dtype=torch.float64, device='cuda')
    drv.residual(x0)
    drv._refresh_preconditioner()
    b = -drv.R
    Mr, _ = _mode2(solver, x0, True)
    cg_el = _CG(asm, 32, 8, 32, element_operator=True)
    assert nat.load().sn_cg_element_operator(cg_el.h) == 1
    if precond == 'ams':
        nat.check(nat.load().sn_cg_set_ams(cg_el.h, drv.ams.handle))
    xe = torch.zeros_like(x0)
    it_e, _, rc = cg_el.solve(Mr, mask, b, xe, rtol=1e-6, atol=0.0, maxiter=3000)
    assert rc == 0
    # float64 CSR CG on the float64 Jacobian of the same state
    locJ = torch.empty(21 * d['num_el'], dtype=torch.float64, device='cuda')
    asm.newton_local_device(x0, 2, solver.source_fields_device(), 1, False, locJ,
                            torch.empty(6 * d['num_el'], dtype=torch.float64, device='cuda'))
    J64 = torch.empty(d['nnz'], dtype=torch.float64, device='cuda')
    nat.check(nat.load().sn_csr_gather_one(d['pattern'], 0, nat.ptr(locJ), nat.ptr(J64),
                                           nat.stream_handle()))
    cg64 = _CG(asm, 64, 8, 32)
    if precond == 'ams':
        nat.check(nat.load().sn_cg_set_ams(cg64.h, drv.ams.handle))
    x64 = torch.zeros_like(x0)
    it_64, _, rc64 = cg64.solve(J64, mask, b, x64, rtol=1e-6, atol=0.0, maxiter=3000)
    assert rc64 == 0
    print('iterations: element f32 %d, csr f64 %d' % (it_e, it_64))
    assert abs(it_e - it_64) <= max(3, it_64 // 10)
    y = torch.empty_like(x0)
    nat.check(nat.load().sn_spmv_f64(d['pattern'], nat.ptr(J64), nat.ptr(mask), nat.ptr(xe),
                                     nat.ptr(y), nat.stream_handle()))
    res = float(((b - y)[mask == 1]).norm() / b[mask == 1].norm())
    Be = asm.element_B(xe.cpu().numpy())
    B64 = asm.element_B(x64.cpu().numpy())
    print('float64 residual of the float32 solution %.2e, B difference %.2e'
          % (res, _rel(Be, B64)))
    # float32 CG: the recursive residual (1e-6 here) and the true one part by
    # the float32 rounding of the updates
    assert res < 3e-5, res
    assert _rel(Be, B64) < 5e-5
    # enable = 0 returns the workspace to CSR values
    nat.check(nat.load().sn_cg_set_element_operator(cg_el.h, None, None, 0, None))
    assert nat.load().sn_cg_element_operator(cg_el.h) == 0


def test_element_cg_launch_shapes_give_the_same_solution(solver):
    """sn_cg_set_grids and sn_ams_set_lanes (the autotuner's knobs) change
    only the launch shapes: the element-operator CG with the auxiliary-space
    preconditioner gives the same iteration count (within 2) and solution
    (float32 summation order) for every candidate shape, and the autotuner
    keeps one of them."""
    import ctypes
    torch = _torch()
    from snoopy_amd import _native as nat
    from snoopy_amd.autotune import element_grid_candidates, tune_element_grids
    from snoopy_amd.red_mvp_solver import _CG
    asm = solver.curl_curl_factory
    mask = solver.mask_device()
    drv = solver.newton_driver('mixed', 8, 32, 'ams')
    drv.retarget(solver.source_fields_device())
    x0 = torch.zeros(asm.num_dofs, dtype=torch.float64, device='cuda')
    drv.residual(x0)
    drv._refresh_preconditioner()
    b = -drv.R
    cg = _CG(asm, 32, 8, 32, element_operator=True)
    nat.check(nat.load().sn_cg_set_ams(cg.h, drv.ams.handle))
    ref = None
    cands = element_grid_candidates(asm._device_state()['num_el'])
    assert len(cands) >= 2
    for g, v, lanes in cands:
        nat.check(nat.load().sn_ams_set_lanes(drv.ams.handle, lanes))
        nat.check(nat.load().sn_cg_set_grids(cg.h, g, v, None))
        ge, gv = ctypes.c_int(), ctypes.c_int()
        nat.check(nat.load().sn_cg_grids(cg.h, ctypes.byref(ge), ctypes.byref(gv)))
        assert ge.value % 8 == 0 and gv.value >= v
        x = torch.zeros_like(x0)
        it, _, rc = cg.solve(drv.Mr, mask, b, x, rtol=1e-6, atol=0.0, maxiter=3000)
        assert rc == 0
        print('shape', (g, v, lanes), '->', (ge.value, gv.value), 'iterations', it, flush=True)
        if ref is None:
            ref = (it, x.clone())
        else:
            assert abs(it - ref[0]) <= 2
            assert float((x - ref[1]).norm() / ref[1].norm()) < 1e-4
    from snoopy_amd import autotune
    try:
        best, times = tune_element_grids(drv, iters=16, reps=1)
        assert best in times and best in cands
    finally:
        # later tests start from the default shapes
        autotune._GRIDS.pop((asm._device_state()['num_el'], asm.num_dofs), None)
    with pytest.raises(RuntimeError):
        nat.check(nat.load().sn_cg_set_grids(cg.h, 0, 0, None))
    with pytest.raises(RuntimeError):
        nat.check(nat.load().sn_ams_set_lanes(drv.ams.handle, 5))
    nat.check(nat.load().sn_ams_set_lanes(drv.ams.handle, 8))


def test_element_operator_needs_a_float32_workspace(solver):
    from snoopy_amd import _native as nat
    from snoopy_amd.red_mvp_solver import _CG
    asm = solver.curl_curl_factory
    d = asm._device_state()
    cg = _CG(asm, 64)
    rc = nat.load().sn_cg_set_element_operator(cg.h, nat.ptr(d['gids']), nat.ptr(d['orient']), 1,
                                               nat.stream_handle())
    assert rc == nat.SN_ERR_ARG
    assert 'float32' in nat.load().sn_last_error().decode()
    assert nat.load().sn_cg_element_operator(cg.h) == 0


@pytest.mark.parametrize('precond', ['jacobi', 'ams'])
def test_mixed_newton_tight_steps_on_the_element_operator(solver, precond):
    x64 = solver.solve_newton('fp64', rtol=1e-10, precond=precond)
    asm = solver.curl_curl_factory
    B64 = asm.element_B(x64.cpu().numpy())
    out = {}
    for tight in ('element', 'fp64'):
        drv = solver.newton_driver('mixed', 8, 32, precond, tight, 'csr')
        # every step with a forcing term below 1e-2 goes to the tight solver
        drv.eta_fp64 = 1e-2
        xm = solver.solve_newton('mixed', rtol=1e-10, precond=precond, tight=tight,
                                 jacobian='csr')
        st = dict(solver.stats)
        assert st['converged'] and st['residual_history'][-1] <= 1e-10
        tight_modes = [m for m, e in zip(st['cg_modes'], st['etas']) if e <= 1e-2]
        assert tight_modes
        want = 'fp32 element' if tight == 'element' else 'fp64'
        assert all(m == want for m in tight_modes), st['cg_modes']
        out[tight] = asm.element_B(xm.cpu().numpy())
        assert _rel(out[tight], B64) < 1e-7


@pytest.mark.parametrize('precond', ['jacobi', 'ams'])
def test_mixed_newton_default_is_the_element_operator(solver, precond):
    """Mixed precision multiplies with the element operator in every step and
    keeps no CSR copy of the Jacobian."""
    x64 = solver.solve_newton('fp64', rtol=1e-10, precond=precond)
    asm = solver.curl_curl_factory
    xm = solver.solve_newton('mixed', rtol=1e-10, precond=precond)
    st = dict(solver.stats)
    assert st['converged'] and st['residual_history'][-1] <= 1e-10
    assert all(m == 'fp32 element' for m in st['cg_modes']), st['cg_modes']
    drv = solver.newton_driver('mixed', 8, 32, precond)
    assert drv.el_all and drv.J is None and drv.cg is None
    assert _rel(asm.element_B(xm.cpu().numpy()), asm.element_B(x64.cpu().numpy())) < 1e-7
    with pytest.raises(ValueError):
        solver.newton_driver('fp64', 8, 32, precond, jacobian='element')
