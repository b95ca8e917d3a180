"""GPU: the variant-minor batched CG (sn_vcg) and the lock-step variant
batch built on it.

* V identical systems give V bit-identical corrections and iteration counts
  (every reduction is per variant and fixed-order), and match the single
  element-operator CG with the same preconditioner (summation orders differ:
  within 2 iterations, corrections within the CG tolerance).
* Different right-hand sides and element matrices (variants of the coil
  current density): each system matches its own single CG.
* The lock-step Newton of 4 and of 3 variants (padded to 4 systems)
  converges every variant to 1e-8 of its own |R(0)|, with element B within
  1e-6 of each variant's own single mixed solve."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = 0.1


def _single_state(dp):
    """Solver, driver and the Jacobian / residual at x = 0 of one variant."""
    import torch
    s = dp.solver(quad_order=2, verbose=False)
    drv = s.newton_driver('mixed', 8, 32, 'ams')
    drv.retarget(s.source_fields_device())
    drv.residual(torch.zeros(s.curl_curl_factory.num_dofs, dtype=torch.float64, device='cuda'))
    drv._refresh_preconditioner()
    return s, drv


def _vcg(s, nv, ams):
    from snoopy_amd import _native as nat
    d = s.curl_curl_factory._device_state()
    h = ctypes.c_void_p()
    nat.check(nat.load().sn_vcg_create(ctypes.byref(h), d['pattern'], nat.ptr(d['gids']),
                                       nat.ptr(d['orient']), nv, 32, nat.stream_handle()))
    nat.check(nat.load().sn_vcg_set_ams(h, ams.handle, nat.stream_handle()))
    return h


def _solve(h, Mr, mask, b, rtol, nv):
    import torch
    from snoopy_amd import _native as nat
    x = torch.zeros_like(b)
    rt = (ctypes.c_double * nv)(*rtol)
    its = (ctypes.c_int * nv)()
    st = (ctypes.c_int * nv)()
    nat.check(nat.load().sn_vcg_solve(h, nat.ptr(Mr), nat.ptr(mask), nat.ptr(b), nat.ptr(x), rt,
                                      0.0, -1, its, None, st, nat.stream_handle()))
    torch.cuda.synchronize()
    return x, list(its), list(st)


def test_identical_systems_are_bit_identical_and_match_the_single_cg():
    import torch
    from snoopy_amd import _native as nat
    from snoopy_amd.magnets import Dipole
    dp = Dipole(h=H)
    print('case: 4 identical systems, %d tets' % dp.num_tets, flush=True)
    s, drv = _single_state(dp)
    nv = 4
    Mr = drv.Mr.reshape(-1, 1).repeat(1, nv).contiguous()
    b = (-drv.R).reshape(-1, 1).repeat(1, nv).contiguous()
    h = _vcg(s, nv, drv.ams)
    try:
        x, its, st = _solve(h, Mr, drv.mask, b, [1e-3] * nv, nv)
    finally:
        nat.load().sn_vcg_destroy(h)
    assert st == [0] * nv and len(set(its)) == 1, (its, st)
    for v in range(1, nv):
        assert torch.equal(x[:, v], x[:, 0]), v
    # the single element-operator CG with the same preconditioner
    xs = torch.zeros_like(drv.R)
    it1, _, _ = drv.cg_el.solve(drv.Mr, drv.mask, -drv.R, xs, rtol=1e-3, atol=0.0)
    assert abs(it1 - its[0]) <= 2, (it1, its[0])
    err = float((x[:, 0] - xs).norm() / xs.norm())
    assert err < 2e-3, err


def test_variants_match_their_own_single_cg():
    import torch
    from snoopy_amd import _native as nat
    from snoopy_amd.magnets import Dipole
    Js = [1.0e6, 1.4e6, 1.8e6, 2.6e6]
    print('case: 4 current densities, h=%g' % H, flush=True)
    states = [_single_state(Dipole(h=H, J=J)) for J in Js]
    s0, d0 = states[1]
    nv = 4
    Mr = torch.stack([d.Mr for _, d in states], dim=1).contiguous()
    b = torch.stack([-d.R for _, d in states], dim=1).contiguous()
    h = _vcg(s0, nv, d0.ams)
    try:
        x, its, st = _solve(h, Mr, d0.mask, b, [1e-4, 1e-3, 1e-2, 1e-3], nv)
    finally:
        nat.load().sn_vcg_destroy(h)
    assert st == [0] * nv, st
    for v, ((s, d), rt) in enumerate(zip(states, [1e-4, 1e-3, 1e-2, 1e-3])):
        # the same system, this variant's CG and the shared preconditioner
        nat.check(nat.load().sn_cg_set_ams(d.cg_el.h, d0.ams.handle))
        xs = torch.zeros_like(d.R)
        it1, _, _ = d.cg_el.solve(d.Mr, d.mask, -d.R, xs, rtol=rt, atol=0.0)
        nat.check(nat.load().sn_cg_set_ams(d.cg_el.h, d.ams.handle))
        assert abs(it1 - its[v]) <= 2, (v, it1, its[v])
        err = float((x[:, v] - xs).norm() / xs.norm())
        assert err < 2 * rt, (v, err)


def test_product_layouts_and_schedules_agree():
    """Element products in DOF-list order or by element slot: the same values
    summed in the same order, so bit-identical iterates; round-robin blocks and
    other grid sizes change only the reduction order, the preconditioner's
    down steps without the combined restrictions only float rounding."""
    import torch
    from snoopy_amd import _native as nat
    from snoopy_amd.magnets import Dipole
    Js = [1.0e6, 1.4e6, 1.8e6, 2.6e6]
    print('case: product layouts, 4 variants, h=%g' % H, flush=True)
    states = [_single_state(Dipole(h=H, J=J)) for J in Js]
    s0, d0 = states[1]
    Mr = torch.stack([d.Mr for _, d in states], dim=1).contiguous()
    b = torch.stack([-d.R for _, d in states], dim=1).contiguous()
    h = _vcg(s0, 4, d0.ams)
    out = {}
    try:
        for flags in (1, 3, 2):
            nat.check(nat.load().sn_vcg_set_schedule(h, flags))
            out[flags] = _solve(h, Mr, d0.mask, b, [1e-3] * 4, 4)
        # down steps through the fine residual of the pre-smoothed iterate
        # instead of the combined restrictions: the same preconditioner up to
        # float rounding
        nat.check(nat.load().sn_vcg_set_schedule(h, 7))
        out['no rt'] = _solve(h, Mr, d0.mask, b, [1e-3] * 4, 4)
        # other launch shapes: only the reductions' partial sums change
        nat.check(nat.load().sn_vcg_set_schedule(h, 3))
        nat.check(nat.load().sn_vcg_set_grids(h, 24, 2, nat.stream_handle()))
        out['grids'] = _solve(h, Mr, d0.mask, b, [1e-3] * 4, 4)
    finally:
        nat.load().sn_vcg_destroy(h)
    assert torch.equal(out[1][0], out[3][0]) and out[1][1] == out[3][1]
    for k in (2, 'grids', 'no rt'):
        assert all(abs(a - c) <= 2 for a, c in zip(out[1][1], out[k][1])), (k, out[1][1], out[k][1])
        err = float((out[1][0] - out[k][0]).norm() / out[1][0].norm())
        assert err < 2e-3, (k, err)


@pytest.mark.parametrize('nvar', [4, 3])
def test_lock_step_newton_matches_single_solves(nvar):
    from snoopy_amd.batch import VariantBatchSolver
    from snoopy_amd.magnets import Dipole
    Js = list(np.linspace(1.2e6, 1.8e6, nvar))
    dps = [Dipole(h=H, J=float(J)) for J in Js]
    print('case: lock-step Newton, %d variants, %d tets' % (nvar, dps[0].num_tets), flush=True)
    vb = VariantBatchSolver(dps)
    out = vb.solve(rtol=1e-8)
    assert vb.stats['systems'] == 4
    for v, (dp, (x, st)) in enumerate(zip(dps, out)):
        assert st['converged'] and st['residual_history'][-1] <= 1e-8, (v, st)
        assert not st['fallback'], (v, st)
        s = dp.solver(quad_order=2, verbose=False)
        xr = s.solve_newton('mixed', rtol=1e-8, precond='ams')
        asm = s.curl_curl_factory
        Bb = asm.element_B(x.cpu().numpy())
        Br = asm.element_B(xr.cpu().numpy())
        err = np.abs(Bb - Br).max() / np.abs(Br).max()
        assert err < 1e-6, (v, err)
    # a second solve on the same objects gives the same result
    out2 = vb.solve(rtol=1e-8)
    for (x, st), (x2, st2) in zip(out, out2):
        assert st['cg_iterations'] == st2['cg_iterations']
        assert bool((x == x2).all())


@pytest.mark.parametrize('n', [1, 1000, 3_000_001])
def test_device_dot_has_the_bits_of_the_host_dot(n):
    """sn_dot_f64_dev leaves on the device the same number sn_dot_f64 returns
    (the same partials, summed in the same order), with and without a mask;
    several queued on one stream land in their own slots."""
    import torch
    from snoopy_amd import _native as nat
    lib = nat.load()
    g = torch.Generator(device='cuda').manual_seed(5)
    a = torch.randn(n, dtype=torch.float64, device='cuda', generator=g) * 1e3
    b = torch.randn(n, dtype=torch.float64, device='cuda', generator=g)
    mask = (torch.rand(n, device='cuda', generator=g) > 0.3).to(torch.uint8)
    out = torch.full((4,), float('nan'), dtype=torch.float64, device='cuda')
    cases = [(a, b, None), (a, a, None), (a, b, mask), (b, b, mask)]
    for k, (x, y, m) in enumerate(cases):
        nat.check(lib.sn_dot_f64_dev(n, nat.ptr(x), nat.ptr(y), nat.ptr(m) if m is not None else None,
                                     nat.ptr(out[k]), nat.stream_handle()))
    got = out.cpu().tolist()
    for k, (x, y, m) in enumerate(cases):
        ref = ctypes.c_double()
        nat.check(lib.sn_dot_f64(n, nat.ptr(x), nat.ptr(y), nat.ptr(m) if m is not None else None,
                                 ctypes.byref(ref), nat.stream_handle()))
        assert got[k] == ref.value, (k, got[k], ref.value)
        # and the value itself, against a float64 sum on the host
        xs, ys = x.cpu().numpy(), y.cpu().numpy()
        w = np.ones(n, dtype=bool) if m is None else m.cpu().numpy().astype(bool)
        exact = float(np.sum(xs[w] * ys[w]))
        assert abs(got[k] - exact) <= 1e-12 * max(1.0, float(np.sum(np.abs(xs[w] * ys[w])))), k


def test_pack_matches_the_strided_copies():
    """sn_vcg_pack gives the variant-minor operands the strided copies give:
    padding systems repeat the last element operators, null residuals give
    b = 0."""
    import torch
    from snoopy_amd import _native as nat
    from snoopy_amd.magnets import Dipole
    dp = Dipole(h=H)
    s, drv = _single_state(dp)
    lib = nat.load()
    N6, n = drv.Mr.numel(), drv.R.numel()
    g = torch.Generator(device='cuda').manual_seed(3)
    mrs = [torch.randn(N6, dtype=torch.float32, device='cuda', generator=g) for _ in range(3)]
    rs = [torch.randn(n, dtype=torch.float64, device='cuda', generator=g) for _ in range(3)]
    live = [True, False, True]
    for nv in (4, 8):
        h = _vcg(s, nv, drv.ams)
        try:
            Mr = torch.full((N6, nv), float('nan'), dtype=torch.float32, device='cuda')
            b = torch.full((n, nv), float('nan'), dtype=torch.float64, device='cuda')
            mr_src = (ctypes.c_void_p * nv)(*[mrs[min(v, 2)].data_ptr() for v in range(nv)])
            r_src = (ctypes.c_void_p * nv)(*[rs[v].data_ptr() if v < 3 and live[v] else None
                                             for v in range(nv)])
            nat.check(lib.sn_vcg_pack(h, mr_src, r_src, nat.ptr(Mr), nat.ptr(b),
                                      nat.stream_handle()))
            torch.cuda.synchronize()
        finally:
            lib.sn_vcg_destroy(h)
        for v in range(nv):
            assert torch.equal(Mr[:, v], mrs[min(v, 2)]), (nv, v)
            ref = -rs[v] if v < 3 and live[v] else torch.zeros_like(rs[0])
            assert torch.equal(b[:, v], ref), (nv, v)


def test_source_only_residual_at_zero_has_the_bits_of_the_full_one():
    """R(0) from the domains with sources only (the others' local values set
    to zero) equals the full residual at zero bit for bit, also when the
    local values hold another state before."""
    import torch
    from snoopy_amd.magnets import Dipole
    dp = Dipole(h=H)
    s, drv = _single_state(dp)
    n = s.curl_curl_factory.num_dofs
    full = drv.residual(torch.zeros(n, dtype=torch.float64, device='cuda'))
    R_full = drv.R.clone()
    x = torch.from_numpy(np.random.default_rng(2).standard_normal(n) * 1e-2).cuda()
    x[drv.mask == 0] = 0.0
    drv.residual(x)                           # other values in every element
    out = torch.full((1,), float('nan'), dtype=torch.float64, device='cuda')
    drv.r0_queued(out[0])
    assert float(np.sqrt(out.item())) == full
    assert torch.equal(drv.R, R_full)
