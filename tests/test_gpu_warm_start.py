"""GPU: the learned warm start.

* The reduced-Newton kernels (sn_pod_element_modes, sn_pod_reduced_jacobian,
  sn_pod_project) against the oracle's float64 Jacobian: H = U^T J U with J
  assembled by the oracle at the same state (tolerance 1e-5 relative: the
  element matrices are float32 in the kernel), g = U^T v exactly.
* End to end at a reduced family size: snapshots from the solver, energy
  POD, BF16 MLP, network + one reduced Newton step, residual judge, then
  Newton to 1e-8 -- converging to the field of the solve from A = 0."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('r', [32, 64, 128])
def test_reduced_jacobian_matches_oracle(r):
    import torch
    from snoopy_amd import _native as nat
    from snoopy_amd.magnets import Dipole
    print('case: reduced jacobian r=%d, dipole h=0.2' % r, flush=True)
    dp = Dipole(h=0.2)
    sol = dp.solver(quad_order=2, verbose=False)
    asm = sol.curl_curl_factory
    d = asm._device_state()
    N, n = d['num_el'], asm.num_dofs
    rng = np.random.default_rng(r)
    x = rng.standard_normal(n) * 3e-3
    x[~sol.make_boundary_mask()] = 0.0
    drv = sol.newton_driver('mixed', 8, 32, 'ams')
    drv.retarget(sol.source_fields_device())
    xd = torch.from_numpy(x).cuda()
    drv.residual(xd)                                   # Mr (float32) and R at x
    U = rng.standard_normal((r, n)).astype(np.float32)
    U[:, ~sol.make_boundary_mask()] = 0.0
    Ud = torch.from_numpy(U).cuda()
    W = torch.empty(N * 3 * r, dtype=torch.float32, device='cuda')
    lib = nat.load()
    nat.check(lib.sn_pod_element_modes(N, n, r, nat.ptr(Ud), nat.ptr(d['gids']),
                                       nat.ptr(d['orient']), nat.ptr(W), nat.stream_handle()))
    H = np.empty(r * r)
    nat.check(lib.sn_pod_reduced_jacobian(N, r, nat.ptr(W), nat.ptr(drv.Mr),
                                          H.ctypes.data_as(ctypes.c_void_p), nat.stream_handle()))
    src = [s.cpu().numpy() if s is not None else None for s in sol.source_fields_device()]
    _, J, _ = orc.assemble_like_reference(asm, x, 2, src)
    U64 = U.astype(np.float64)
    Href = U64 @ (J @ U64.T)
    err = np.abs(H.reshape(r, r) - Href).max() / np.abs(Href).max()
    assert err < 1e-5, err
    g = np.empty(r)
    v = torch.from_numpy(rng.standard_normal(n)).cuda()
    nat.check(lib.sn_pod_project(n, r, nat.ptr(Ud), nat.ptr(v), g.ctypes.data_as(ctypes.c_void_p),
                                 nat.stream_handle()))
    gref = U64 @ v.cpu().numpy()
    assert np.abs(g - gref).max() <= 1e-12 * np.abs(gref).max()


def test_reduced_jacobian_reuses_the_constant_domains():
    """WarmStart.reduced_jacobian sums the air (constant reluctivity) once per
    geometry and the iron on every call: at a second, saturating state it
    still equals the full sum of the kernel over all elements (which then
    differs from the first state's), and a new geometry version re-forms
    the cached part."""
    import torch
    from snoopy_amd import _native as nat
    from snoopy_amd.magnets import Dipole
    from snoopy_amd.warm_start import WarmStart
    r = 32
    print('case: cached constant part of the reduced jacobian, dipole h=0.2', flush=True)
    dp = Dipole(h=0.2)
    sol = dp.solver(quad_order=2, verbose=False)
    asm = sol.curl_curl_factory
    d = asm._device_state()
    N, n = d['num_el'], asm.num_dofs
    rng = np.random.default_rng(5)
    free = sol.make_boundary_mask()
    U = rng.standard_normal((r, n)).astype(np.float32)
    U[:, ~free] = 0.0
    Ud = torch.from_numpy(U).cuda()
    W = torch.empty(N * 3 * r, dtype=torch.float32, device='cuda')
    lib = nat.load()
    nat.check(lib.sn_pod_element_modes(N, n, r, nat.ptr(Ud), nat.ptr(d['gids']),
                                       nat.ptr(d['orient']), nat.ptr(W), nat.stream_handle()))
    ws = WarmStart.__new__(WarmStart)
    ws._W, ws.r_galerkin = W, r
    drv = sol.newton_driver('mixed', 8, 32, 'ams')
    drv.retarget(sol.source_fields_device())
    var, const = ws._domain_ranges(asm)
    assert len(var) == 1 and len(const) == 1
    Hs = []
    for scale in (1e-4, 3e-2):
        x = rng.standard_normal(n) * scale
        x[~free] = 0.0
        drv.residual(torch.from_numpy(x).cuda())
        H = ws.reduced_jacobian(drv)
        full = np.empty(r * r)
        nat.check(lib.sn_pod_reduced_jacobian(N, r, nat.ptr(W), nat.ptr(drv.Mr),
                                              full.ctypes.data_as(ctypes.c_void_p),
                                              nat.stream_handle()))
        full = full.reshape(r, r)
        assert np.abs(H - full).max() <= 1e-5 * np.abs(full).max()
        Hs.append(full)
    assert np.abs(Hs[1] - Hs[0]).max() > 1e-3 * np.abs(Hs[0]).max()
    key0 = asm._galerkin_const[ws][0]
    asm.geometry_version = getattr(asm, 'geometry_version', 0) + 1
    ws.reduced_jacobian(drv)
    assert asm._galerkin_const[ws][0] != key0


@pytest.mark.parametrize('r', [32, 128])
def test_device_forms_of_the_reduced_step_match_the_host_forms(r):
    """The reduced Jacobian (with its cached constant part) and the batched
    projection left on the device equal their host forms bit for bit; the
    device Cholesky solve matches numpy's solve of the symmetrised systems,
    and gives no correction for a system that is not positive definite."""
    import torch
    from snoopy_amd import _native as nat
    from snoopy_amd.magnets import Dipole
    from snoopy_amd.warm_start import WarmStart
    print('case: device reduced step r=%d, dipole h=0.2' % r, flush=True)
    dp = Dipole(h=0.2)
    sol = dp.solver(quad_order=2, verbose=False)
    asm = sol.curl_curl_factory
    d = asm._device_state()
    N, n = d['num_el'], asm.num_dofs
    rng = np.random.default_rng(9)
    free = sol.make_boundary_mask()
    U = rng.standard_normal((r, n)).astype(np.float32)
    U[:, ~free] = 0.0
    Ud = torch.from_numpy(U).cuda()
    W = torch.empty(N * 3 * r, dtype=torch.float32, device='cuda')
    lib = nat.load()
    nat.check(lib.sn_pod_element_modes(N, n, r, nat.ptr(Ud), nat.ptr(d['gids']),
                                       nat.ptr(d['orient']), nat.ptr(W), nat.stream_handle()))
    ws = WarmStart.__new__(WarmStart)
    ws._W, ws.r_galerkin = W, r
    drv = sol.newton_driver('mixed', 8, 32, 'ams')
    drv.retarget(sol.source_fields_device())
    x = rng.standard_normal(n) * 1e-2
    x[~free] = 0.0
    drv.residual(torch.from_numpy(x).cuda())
    Hh = ws.reduced_jacobian(drv)
    Hd = torch.full((r, r), float('nan'), dtype=torch.float64, device='cuda')
    ws.reduced_jacobian_device(drv, Hd)
    assert np.array_equal(Hd.cpu().numpy(), Hh)
    # projections of 4 vectors
    V4 = torch.from_numpy(rng.standard_normal((4, n))).cuda()
    gh = np.empty((4, r))
    nat.check(lib.sn_pod_project_many(n, r, 4, nat.ptr(Ud), nat.ptr(V4),
                                      gh.ctypes.data_as(ctypes.c_void_p), nat.stream_handle()))
    gd = torch.full((4, r), float('nan'), dtype=torch.float64, device='cuda')
    nat.check(lib.sn_pod_project_many_dev(n, r, 4, nat.ptr(Ud), nat.ptr(V4), nat.ptr(gd),
                                          nat.stream_handle()))
    assert np.array_equal(gd.cpu().numpy(), gh)
    # each row has the bits of the single-vector projection
    for w in range(4):
        g1 = np.empty(r)
        nat.check(lib.sn_pod_project(n, r, nat.ptr(Ud), nat.ptr(V4[w]),
                                     g1.ctypes.data_as(ctypes.c_void_p), nat.stream_handle()))
        assert np.array_equal(g1, gh[w]), w
    # solves: the reduced Jacobian, a random SPD matrix, a singular one
    M = rng.standard_normal((r, r))
    S = np.zeros((r, r))
    S[:r // 2, :r // 2] = np.eye(r // 2)
    Hs = np.stack([Hh, M @ M.T + r * np.eye(r), S])
    g = rng.standard_normal((3, r))
    D = torch.full((3, r), float('nan'), dtype=torch.float32, device='cuda')
    nat.check(lib.sn_pod_reduced_solve_many(3, r, nat.ptr(torch.from_numpy(Hs).cuda()),
                                            nat.ptr(torch.from_numpy(g).cuda()), nat.ptr(D),
                                            nat.stream_handle()))
    D = D.cpu().numpy()
    for k in range(2):
        ref = np.linalg.solve(0.5 * (Hs[k] + Hs[k].T), -g[k])
        assert np.abs(D[k] - ref).max() <= 1e-5 * np.abs(ref).max(), k
    assert np.all(D[2] == 0.0)


@pytest.fixture(scope='module')
def trained():
    from snoopy_amd.warm_start import MorphFamily, WarmStart
    fam = MorphFamily(counts=(4, 3, 4, 3, 3, 5))
    print('case: warm start family tets', fam.mesh.num_elements, flush=True)
    ws = WarmStart.train(fam, n_train=128, r=64, epochs=800, rtol=1e-8, workers=2,
                         hidden=(128, 128, 128, 128))
    return fam, ws


def test_training_fits_coefficients(trained):
    fam, ws = trained
    st = ws.train_stats
    assert st['snapshots_converged'] == st['n_train']
    assert st['loss_last'] < 0.05 * st['loss_first'], st
    assert st['pod_energy'] > 0.999


def test_learned_start_converges_to_the_same_field(trained):
    """Held-out variants: Newton from the judged start (network + one reduced
    Newton step) reaches the same tolerance and the same field as Newton from
    A = 0.  The CG work of both is printed."""
    from snoopy_amd.warm_start import sample_params, solve_with_warm_start
    fam, ws = trained
    fs = ws.fsolver
    work = [0, 0]
    for seed in (5000, 5001, 5002, 5003):
        p = sample_params(seed)
        fs.set_variant(p)
        xa = fs.solver.solve_newton('mixed', rtol=1e-8, precond='ams')
        s0 = dict(fs.solver.stats)
        xb = solve_with_warm_start(fs.solver, ws, p, rtol=1e-8)
        s1 = dict(fs.solver.stats)
        assert s0['converged'] and s1['converged']
        assert s1['residual_history'][-1] <= 1e-8
        asm = fs.solver.curl_curl_factory
        Ba = asm.element_B(xa.cpu().numpy())
        Bb = asm.element_B(xb.cpu().numpy())
        err = np.abs(Ba - Bb).max() / np.abs(Ba).max()
        work[0] += sum(s0['cg_iterations'])
        work[1] += sum(s1['cg_iterations'])
        print('seed', seed, 'start residual %.3e' % s1['warm_start']['r_start_rel'], 'kept',
              s1['warm_start']['used'], 'cg', sum(s0['cg_iterations']), '->',
              sum(s1['cg_iterations']), 'B diff %.1e' % err, flush=True)
        assert err < 1e-6
    print('total CG from zero %d, from the learned start %d' % tuple(work), flush=True)


def _family_variant(fam, p):
    """A Dipole of parameters p on its own copy of the family mesh."""
    m = fam.private_model()
    m.tet_mesh.nodes[:] = fam.nodes(dict(fam.p0, **p))
    return fam.variant(p, m)


def test_concurrent_learned_batch_matches_the_batch_from_zero(trained):
    """ConcurrentSolver with warm_start='learned' on 2 streams: every variant
    converges to 1e-8 and to the field of the same batch solved from A = 0."""
    from snoopy_amd.batch import ConcurrentSolver
    from snoopy_amd.warm_start import sample_params
    fam, ws = trained
    dps = [_family_variant(fam, sample_params(s)) for s in (6000, 6001, 6002, 6003)]
    print('case: learned batch of 4 on 2 streams', flush=True)
    cs = ConcurrentSolver(dps, workers=2, precision='mixed', precond='ams',
                          warm_start='learned', model=ws)
    warm = cs.solve(rtol=1e-8)
    xs_warm = [x.cpu().numpy() for x, _ in warm]
    st_warm = [st for _, st in warm]
    cs.set_warm_start(None)
    cold = cs.solve(rtol=1e-8)
    for k, ((xc, sc), xw, sw, s) in enumerate(zip(cold, xs_warm, st_warm, cs.solvers)):
        assert sc['converged'] and sw['converged'], k
        assert sw['residual_history'][-1] <= 1e-8
        assert 'warm_start' in sw and sw['warm_start']['r_start_rel'] > 0
        asm = s.curl_curl_factory
        Bc, Bw = asm.element_B(xc.cpu().numpy()), asm.element_B(xw)
        err = np.abs(Bc - Bw).max() / np.abs(Bc).max()
        print('variant', k, 'cg', sum(sc['cg_iterations']), '->', sum(sw['cg_iterations']),
              'start %.2e' % sw['warm_start']['r_start_rel'], 'B diff %.1e' % err, flush=True)
        assert err < 1e-6


def test_batched_start_matches_the_single_starts(trained):
    """WarmStart.start_many (expansions and projections of all variants in
    one pass over the modes) against WarmStart.start per variant: the same
    residual judge, starts equal up to the order of the correction's sum."""
    import torch
    from snoopy_amd.warm_start import PARAM_NAMES, sample_params
    fam, ws = trained
    dps = [_family_variant(fam, sample_params(s)) for s in (6100, 6101, 6102)]
    print('case: batched learned start of 3 (padded to 4)', flush=True)
    sols = [d.solver(quad_order=2, verbose=False) for d in dps]
    for s in sols:
        ws.precompute_constant(s)
        drv = s.newton_driver('mixed', 8, 32, 'ams')
        drv.retarget(s.source_fields_device())
        drv.r0()
    params = [{k: d.p[k] for k in PARAM_NAMES} for d in dps]
    many = ws.start_many(sols, params)
    for s, p, (xm, im) in zip(sols, params, many):
        xs, info = ws.start(s, p)
        assert im['used'] == info['used']
        assert abs(im['r_start_rel'] - info['r_start_rel']) <= 1e-6 * info['r_start_rel']
        if xs is not None:
            err = float((xm - xs).abs().max() / xs.abs().max())
            assert err < 1e-9, err
    # the pass over the modes: bit-identical to the single expansion
    a = ws.coefficients(params)
    X = ws.expand_many(np.vstack([a, a[:1]]))
    for v in range(3):
        assert torch.equal(X[v], ws.expand(a[v]))


@pytest.mark.parametrize('batched_start', [False, True])
def test_variant_batch_with_learned_start_matches_the_batch_from_zero(trained, batched_start):
    """VariantBatchSolver (lock-step Newton, one variant-minor CG) from the
    learned start, each variant's start on its own stream or all starts in
    one pass over the modes: every variant converges to 1e-8 and to the field
    of the same batch from A = 0."""
    from snoopy_amd.batch import VariantBatchSolver
    from snoopy_amd.warm_start import sample_params
    fam, ws = trained
    dps = [_family_variant(fam, sample_params(s)) for s in (6200, 6201, 6202, 6203)]
    print('case: learned variant-minor batch of 4, batched start %s' % batched_start, flush=True)
    vb = VariantBatchSolver(dps, warm_start='learned', model=ws, batched_start=batched_start)
    warm = vb.solve(rtol=1e-8)
    xs_warm = [x.cpu().numpy() for x, _ in warm]
    st_warm = [st for _, st in warm]
    vb.set_warm_start(None)
    cold = vb.solve(rtol=1e-8)
    for k, ((xc, sc), xw, sw, s) in enumerate(zip(cold, xs_warm, st_warm, vb.solvers)):
        assert sc['converged'] and sw['converged'], k
        assert sw['residual_history'][-1] <= 1e-8
        assert sw['warm_start']['r_start_rel'] > 0
        asm = s.curl_curl_factory
        Bc, Bw = asm.element_B(xc.cpu().numpy()), asm.element_B(xw)
        err = np.abs(Bc - Bw).max() / np.abs(Bc).max()
        print('variant', k, 'cg', sum(sc['cg_iterations']), '->', sum(sw['cg_iterations']),
              'start %.2e' % sw['warm_start']['r_start_rel'], 'B diff %.1e' % err, flush=True)
        assert err < 1e-6


@pytest.mark.parametrize('steps', [1, 2])
def test_batch_starts_match_the_single_starts(trained, steps):
    """The variant batch's starts (sources, R(0), reduced steps queued for
    all variants with one host wait) against WarmStart.start per variant,
    with one and two reduced Newton steps: the same judge, starts equal up to
    the order of the corrections' sums."""
    import torch
    from snoopy_amd.batch import VariantBatchSolver
    from snoopy_amd.warm_start import PARAM_NAMES, sample_params
    fam, ws = trained
    dps = [_family_variant(fam, sample_params(s)) for s in (6300, 6301, 6302)]
    print('case: batch starts of 3, %d reduced steps' % steps, flush=True)
    old = ws.galerkin_steps
    ws.galerkin_steps = steps
    try:
        vb = VariantBatchSolver(dps, warm_start='learned', model=ws)
        drvs = [vb._driver(v) for v in range(3)]
        params = [{k: d.p[k] for k in PARAM_NAMES} for d in dps]
        coeffs = ws.coefficients(params)
        r0, x0s, infos = [None] * 3, [None] * 3, [None] * 3
        vb._batched_starts(drvs, coeffs, r0, x0s, infos)
        for v, (s, p) in enumerate(zip(vb.solvers, params)):
            xs, info = ws.start(s, p, coeffs=coeffs[v])
            assert infos[v]['used'] == info['used'], v
            assert abs(infos[v]['r_start_rel'] - info['r_start_rel']) <= 1e-6 * info['r_start_rel']
            assert r0[v] == s.newton_driver('mixed', 8, 32, 'ams').r0()
            if xs is not None:
                err = float((x0s[v] - xs).abs().max() / xs.abs().max())
                assert err < 1e-9, (v, err)
    finally:
        ws.galerkin_steps = old


def test_saved_model_predicts_the_same_coefficients(trained, tmp_path):
    from snoopy_amd.warm_start import MorphFamily, WarmStart, sample_params
    fam, ws = trained
    path = str(tmp_path / 'ws.npz')
    ws.save(path)
    ps = [sample_params(s) for s in (7000, 7001, 7002)]
    a = ws.coefficients(ps)
    for fam2 in (fam, None):
        ws2 = WarmStart.load(path, family=fam2)
        np.testing.assert_array_equal(ws2.coefficients(ps), a)
    with pytest.raises(ValueError):
        WarmStart.load(path, family=MorphFamily(counts=(3, 3, 4, 3, 3, 5)))
