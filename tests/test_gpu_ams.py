"""GPU: the auxiliary-space preconditioner against its float64 oracle (level
matrices, one application), inside CG (same solution as Jacobi-PCG, fewer
iterations) and inside Newton (same field)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def case():
    import importlib
    import torch
    mg = importlib.import_module('snoopy_amd.magnets')
    d = mg.Dipole(h=0.08)
    solver = d.solver()
    print('case: AMS dipole h=0.08 tets', d.mesh.num_elements, flush=True)
    drv = solver.newton_driver('fp64', 8, 32, 'ams')
    drv.retarget(solver.source_fields_device())
    # the Jacobian (and element reluctivities) at A = 0
    x = torch.zeros(solver.curl_curl_factory.num_dofs, dtype=torch.float64, device='cuda')
    drv.residual(x)
    drv._refresh_preconditioner()
    return d, solver, drv


def _oracle(d, solver, drv):
    from oracle import ams_oracle as ao
    asm = solver.curl_curl_factory
    tm = d.mesh
    order = asm._device_state()['order']
    plan = drv.ams.plan
    l0 = np.full(tm.num_nodes, plan.n_free)
    l0[plan.free_ids] = plan.perm0
    nu = drv.nu_el.cpu().numpy().astype(np.float64)
    return ao.OracleAms(tm.nodes, tm.edges, tm.cells[order], drv.mask.cpu().numpy().astype(bool),
                        nu, l0, [lv['agg_ptr'] for lv in plan.levels[1:]],
                        omega=drv.ams.omega, weight=drv.ams.weight)


def test_level_matrices_match_oracle(case):
    import scipy.sparse as sp
    d, solver, drv = case
    o = _oracle(d, solver, drv)
    for l, lv in enumerate(drv.ams.plan.levels):
        vals = drv.ams.level_values(l).cpu().numpy().astype(np.float64)
        M = sp.csr_array((vals, lv['indices'], lv['indptr']), shape=(lv['n'], lv['n']))
        diff = (M - o.A[l]).tocsr()
        err = np.abs(diff.data).max() / np.abs(o.A[l].data).max() if diff.nnz else 0.0
        print('level', l, 'rows', lv['n'], 'max rel diff %.2e' % err, flush=True)
        assert err < 2e-6


@pytest.mark.parametrize('combined', [1, 0])
def test_apply_matches_oracle(case, combined):
    """One application against the oracle, with the down steps through the
    coarse levels' combined restrictions P^T (I - w A D^-1) and through the
    fine residual of the pre-smoothed iterate (the default)."""
    import torch
    from snoopy_amd import _native as nat
    d, solver, drv = case
    nat.check(nat.load().sn_ams_set_restriction(drv.ams.handle, combined))
    o = _oracle(d, solver, drv)
    rng = np.random.default_rng(0)
    n = solver.curl_curl_factory.num_dofs
    free = drv.mask.cpu().numpy().astype(bool)
    r = np.where(free, rng.standard_normal(n), 0.0)
    dinv = rng.uniform(0.5, 2.0, n)
    rd = torch.from_numpy(r).cuda()
    # the auxiliary part alone (zero diagonal part), which runs in float32
    z_ref = o.apply(r, np.zeros(n))
    z = drv.ams.apply(drv.mask, rd, torch.zeros(n, dtype=torch.float64, device='cuda'))
    z = z.cpu().numpy()
    err = np.abs(z - z_ref).max() / np.abs(z_ref).max()
    print('auxiliary part max rel diff %.2e' % err, flush=True)
    assert err < 1e-4
    # the whole preconditioner
    z_ref = o.apply(r, dinv)
    z = drv.ams.apply(drv.mask, rd, torch.from_numpy(dinv).cuda()).cpu().numpy()
    err = np.abs(z - z_ref).max() / np.abs(z_ref).max()
    print('apply max rel diff %.2e' % err, flush=True)
    assert err < 1e-6
    assert np.all(z[~free] == 0.0)
    nat.check(nat.load().sn_ams_set_restriction(drv.ams.handle, 0))


def _spmv(solver, vals, mask, x):
    import torch
    from snoopy_amd import _native as nat
    y = torch.empty_like(x)
    d = solver.curl_curl_factory._device_state()
    nat.check(nat.load().sn_spmv_f64(d['pattern'], nat.ptr(vals), nat.ptr(mask), nat.ptr(x),
                                     nat.ptr(y), nat.stream_handle()))
    return y


def test_cg_with_ams_solves_the_jacobi_system(case):
    import torch
    from snoopy_amd import _native as nat
    from snoopy_amd.red_mvp_solver import _CG
    d, solver, drv = case
    asm = solver.curl_curl_factory
    n = asm.num_dofs
    rng = np.random.default_rng(1)
    v = torch.from_numpy(rng.standard_normal(n)).cuda()
    b = _spmv(solver, drv.J, drv.mask, v)            # consistent right-hand side
    res = {}
    for name in ('jacobi', 'ams'):
        cg = _CG(asm, 64)
        if name == 'ams':
            nat.check(nat.load().sn_cg_set_ams(cg.h, drv.ams.handle))
        x = torch.zeros(n, dtype=torch.float64, device='cuda')
        its, rn, rc = cg.solve(drv.J, drv.mask, b, x, rtol=1e-10, atol=0.0, maxiter=20000)
        assert rc == nat.SN_OK
        r = b - _spmv(solver, drv.J, drv.mask, x)
        rel = float(torch.linalg.norm(r[drv.mask.bool()]) / torch.linalg.norm(b[drv.mask.bool()]))
        res[name] = (its, rel, x)
        print(name, 'iterations', its, 'true residual %.2e' % rel, flush=True)
        assert rel < 1e-9
    assert res['ams'][0] < res['jacobi'][0] / 2


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_newton_with_ams_matches_jacobi(case, precision):
    d, solver, drv = case
    asm = solver.curl_curl_factory
    xa = solver.solve_newton(precision, rtol=1e-8, precond='jacobi')
    sa = dict(solver.stats)
    xb = solver.solve_newton(precision, rtol=1e-8, precond='ams')
    sb = dict(solver.stats)
    assert sa['converged'] and sb['converged']
    Ba = asm.element_B(xa.cpu().numpy())
    Bb = asm.element_B(xb.cpu().numpy())
    err = np.abs(Ba - Bb).max() / np.abs(Ba).max()
    print(precision, 'cg jacobi', sa['cg_iterations'], 'ams', sb['cg_iterations'],
          'B diff %.1e' % err, flush=True)
    assert err < 1e-6
    assert sum(sb['cg_iterations']) < sum(sa['cg_iterations'])
