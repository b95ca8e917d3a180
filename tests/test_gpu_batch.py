"""GPU: batched solves (block-diagonal join of uneven meshes, segment-aware
CG) against the same variants solved one at a time."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def variants():
    from snoopy_amd.magnets import Dipole, sample_variant
    out = []
    for seed, h in ((0, 0.12), (1, 0.15), (2, 0.10)):
        out.append(Dipole(h=h, **sample_variant(seed)))
    print('case: batch of', [d.num_tets for d in out], 'tets', flush=True)
    return out


def test_segmented_cg_matches_single_solves(variants):
    import torch
    from snoopy_amd import _native as nat
    from snoopy_amd.batch import BatchedSolver
    from snoopy_amd.red_mvp_solver import _CG
    bs = BatchedSolver(variants, precision='fp64')
    bs.source_fields()
    bs.residual(torch.zeros(bs.n, dtype=torch.float64, device='cuda'))
    b = -bs.R.clone()
    x = torch.zeros_like(b)
    rt = (ctypes.c_double * bs.nvar)(*([1e-6] * bs.nvar))
    its = (ctypes.c_int * bs.nvar)()
    nat.check(nat.load().sn_cg_solve_batched(bs.cg, nat.ptr(bs.J), nat.ptr(bs.mask), nat.ptr(b),
                                             nat.ptr(x), rt, 0.0, -1, its, None, None,
                                             nat.stream_handle()))
    for v, d in enumerate(variants):
        s = d.solver(quad_order=2, verbose=False)
        asm = s.curl_curl_factory
        drv = s.newton_driver('fp64')
        drv.retarget(s.source_fields_device())
        drv.residual(torch.zeros(asm.num_dofs, dtype=torch.float64, device='cuda'))
        bv = -drv.R
        xv = torch.zeros_like(bv)
        it, _, _ = _CG(asm, 64).solve(drv.J, s.mask_device(), bv, xv, rtol=1e-6, atol=0.0)
        lo, hi = int(bs.row0[v]), int(bs.row0[v + 1])
        assert abs(its[v] - it) <= 2, (v, its[v], it)
        # the joined system reproduces the single system's right-hand side
        assert float((b[lo:hi] - bv).abs().max()) <= 1e-12 * float(bv.abs().max())
        Bb = asm.element_B(x[lo:hi].cpu().numpy())
        Bs = asm.element_B(xv.cpu().numpy())
        assert np.abs(Bb - Bs).max() <= 1e-4 * np.abs(Bs).max()


@pytest.mark.parametrize('precision,precond', [('mixed', 'jacobi'), ('fp64', 'jacobi'),
                                               ('mixed', 'ams'), ('fp64', 'ams')])
def test_batched_newton_matches_single_newton(variants, precision, precond):
    from snoopy_amd.batch import BatchedSolver
    bs = BatchedSolver(variants, precision=precision, precond=precond)
    x = bs.solve(rtol=1e-8)
    assert bs.stats['converged'].all()
    print(precision, precond, 'batched cg', bs.stats['cg_iterations'], flush=True)
    parts = bs.split(x)
    for v, d in enumerate(variants):
        s = d.solver(quad_order=2, verbose=False)
        xv = s.solve_newton(precision, rtol=1e-8, precond=precond)
        asm = s.curl_curl_factory
        Bb = asm.element_B(parts[v])
        Bs = asm.element_B(xv.cpu().numpy())
        assert np.abs(Bb - Bs).max() <= 1e-6 * np.abs(Bs).max()


@pytest.mark.parametrize('workers', [2, 3])
def test_concurrent_streams_match_single_solves(variants, workers):
    """Variants solved on several streams at once give the solutions of the
    same solves run one at a time (every solve is deterministic).  With the
    element operator's CG, concurrent streams delay some workgroups of a
    kernel behind others: a kernel that read a counter its first workgroup
    advances broke down there (regression)."""
    from snoopy_amd.batch import ConcurrentSolver
    cs = ConcurrentSolver(variants, workers=workers, precision='mixed', precond='ams')
    res = cs.solve(rtol=1e-8)
    for v, d in enumerate(variants):
        assert res[v][1]['converged']
        s = d.solver(quad_order=2, verbose=False)
        xv = s.solve_newton('mixed', rtol=1e-8, precond='ams')
        assert res[v][1]['cg_iterations'] == s.stats['cg_iterations']
        asm = s.curl_curl_factory
        Bc = asm.element_B(res[v][0].cpu().numpy())
        Bs = asm.element_B(xv.cpu().numpy())
        assert np.abs(Bc - Bs).max() <= 1e-12 * np.abs(Bs).max()


def test_anchor_warm_start_matches_cold_solves():
    """Variants of one dipole that differ in the coil current density: with the
    anchor warm start (middle variant from zero, the others from its solution
    scaled by the current ratio) every variant converges to ||R|| <= 1e-8
    ||R(0)|| of its own and gives the field of its solve from zero, with fewer
    CG iterations than from zero."""
    from snoopy_amd.batch import ConcurrentSolver
    from snoopy_amd.magnets import Dipole
    fam = [Dipole(h=0.1, J=J) for J in (1.2e6, 1.5e6, 1.8e6)]
    print('case: anchor warm start, 3 variants of', fam[0].num_tets, 'tets', flush=True)
    warm = ConcurrentSolver(fam, workers=2, precision='mixed', precond='ams',
                            warm_start='anchor').solve(rtol=1e-8)
    cold = ConcurrentSolver(fam, workers=2, precision='mixed', precond='ams').solve(rtol=1e-8)
    asm = fam[0].solver(quad_order=2, verbose=False).curl_curl_factory
    for v in range(3):
        assert warm[v][1]['converged'] and cold[v][1]['converged']
        assert warm[v][1]['residual_history'][-1] <= 1e-8
        Bw = asm.element_B(warm[v][0].cpu().numpy())
        Bc = asm.element_B(cold[v][0].cpu().numpy())
        err = np.abs(Bw - Bc).max() / np.abs(Bc).max()
        iw, ic = sum(warm[v][1]['cg_iterations']), sum(cold[v][1]['cg_iterations'])
        print('variant', v, 'B diff %.2e' % err, 'cg warm', iw, 'cold', ic, flush=True)
        assert err <= 1e-6
        if v != 1:
            assert iw < ic
    # the anchor itself is the solve from zero
    assert warm[1][1]['cg_iterations'] == cold[1][1]['cg_iterations']


def test_anchor_warm_start_needs_one_mesh():
    from snoopy_amd.batch import ConcurrentSolver
    from snoopy_amd.magnets import Dipole
    with pytest.raises(ValueError):
        ConcurrentSolver([Dipole(h=0.12), Dipole(h=0.15)], workers=2, warm_start='anchor')
