"""Benchmark of the nonlinear reduced-vector-potential solve on one MI355X.

Workload (the metric's configuration): the window-frame dipole (yoke
1.0 x 0.6 x 0.5 m, aperture 0.3 x 0.2 m, two racetrack coils at 1.5e6 A/m^2,
0.5 m air pad) meshed with ~1M Kuhn tetrahedra (h = 0.02 m), lowest-order
Nedelec elements, 4-point quadrature, iron nu = cubic B-spline in B^2 fitted
to 60 samples of B(H) = mu0 H + 1.9 tanh(mu0 4000 H / 1.9).

A solve is complete: source field at the iron quadrature points
(Biot-Savart) plus Newton from the start vector to ||R|| <= 1e-8 ||R(0)||.
Mesh, CSR pattern and element-to-nonzero map are set up before the timed
region (inputs resident in HBM).  One step is one batch: --batch variants of
the dipole (coil current density 1.2e6..1.8e6 A/m^2) solved in the chosen
form (mixed precision with the AMS preconditioner, the learned warm start,
Newton in lock step with one variant-minor CG for all variants; --batch-warm,
--batch-solver), each to ||R|| <= 1e-8 ||R(0)|| of its own; value = solves per
second of the batch.  Beside it: the same batch from zero
(batch_cold_solves_per_s), and on concurrent HIP streams with the learned,
zero and anchor starts.  With --batch 0 a step is
one solve.  The single solve of the chosen form and the plain FP64 GPU solve
are timed beside it, and the relative B error at 10k aperture probes is
reported against FP64 solves at 1e-11 (for the single solve and for every
variant of the batch).

Prints one JSON line (rank 0, N = 1).
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/s at 1M tets, mixed+learned warm start vs plain FP64, B err<1e-6"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--tets', type=float, default=1.0e6)
    ap.add_argument('--rtol', type=float, default=1e-8)
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
    ap.add_argument('--no-extra', action='store_true',
                    help='skip the 64-variant batch and the 6-yoke chain in the default line')
    ap.add_argument('--cpu-cg-iters', type=int, default=300)
    ap.add_argument('--cpu-threads', type=int, default=16)
    ap.add_argument('--lanes', type=int, default=0, help='CG lanes per row (0 = autotune)')
    ap.add_argument('--precond', default='ams', choices=['jacobi', 'ams'],
                    help='CG preconditioner of the timed solves')
    ap.add_argument('--ams-weight', type=float, default=None)
    ap.add_argument('--ams-omega', type=float, default=None)
    ap.add_argument('--ams-theta', type=float, default=None)
    ap.add_argument('--ams-passes', type=int, default=None)
    ap.add_argument('--ams-coarse', type=int, default=None)
    ap.add_argument('--workload', default='dipole1m',
                    choices=['dipole1m', 'batch64', 'chain6', 'warmstart', 'gemm'],
                    help='dipole1m: the metric; batch64: 64 seeded variants 20k-400k tets, '
                         'batched vs one at a time; chain6: 6-yoke chain')
    ap.add_argument('--variants', type=int, default=64)
    ap.add_argument('--batch', type=int, default=8,
                    help='dipole1m: variants solved concurrently in one step (0: one solve '
                         'per step)')
    ap.add_argument('--batch-warm', default='learned', choices=['none', 'anchor', 'learned'],
                    help='dipole1m: start of the timed solves -- learned (network + reduced '
                         'Newton step, trained before the timed region), none (A = 0), or '
                         'anchor (batch of J-only variants: median-J variant from zero, the '
                         'others from its scaled solution); the other two batch starts are '
                         'timed beside it')
    ap.add_argument('--tune-grids', type=int, default=1,
                    help='dipole1m: autotune the element-operator CG launch shapes (1) or not (0)')
    ap.add_argument('--batch-solver', default='variant', choices=['variant', 'concurrent'],
                    help='dipole1m: the chosen batch form -- variant (lock-step Newton, one '
                         'variant-minor CG for all variants, at most 8) or concurrent (one '
                         'solver per variant on --streams HIP streams); the other is timed '
                         'beside it')
    ap.add_argument('--batch-only', action='store_true',
                    help='dipole1m: time the chosen batch start only (for traces)')
    ap.add_argument('--streams', type=int, default=None,
                    help='concurrent HIP streams (default 8 for dipole1m, 4 for batch64)')
    ap.add_argument('--ws-train', type=int, default=None,
                    help='warm start: training snapshots (default 512 for dipole1m, 2048 for '
                         'the warmstart workload)')
    ap.add_argument('--ws-test', type=int, default=256, help='warm start: held-out variants')
    ap.add_argument('--ws-modes', type=int, default=256, help='warm start: POD modes')
    ap.add_argument('--ws-epochs', type=int, default=2000, help='warm start: training epochs')
    ap.add_argument('--ws-galerkin', type=int, default=128,
                    help='warm start: modes of the reduced Newton step (32, 64 or 128)')
    ap.add_argument('--ws-workers', type=int, default=4,
                    help='warm start: concurrent snapshot solvers (one stream each)')
    ap.add_argument('--ws-galerkin-steps', type=int, default=1,
                    help='warm start: reduced Newton steps before the full Newton')
    ap.add_argument('--copy-calib', action='store_true',
                    help='only run a plain 256 MiB device copy 5 times (HBM counter calibration)')
    ap.add_argument('--quick', action='store_true',
                    help='timed solves only (no B-error reference, roofline or CPU leg)')
    return ap.parse_args()


def probe_B(dp, solver, x_dev, probes):
    """B at the probe points as the reference evaluates it: coil field plus
    element-averaged curl A at the nodes of each domain, P1-interpolated
    (evaluate_fem_solution_node_interpolation)."""
    from snoopy_amd.evaluators import evaluate_fem_solution_node_interpolation
    return evaluate_fem_solution_node_interpolation(dp.model, probes, solver, dp.coils,
                                                    x_dev.cpu().numpy(), [1, 2])


def b_error(B, Bref):
    d = np.linalg.norm(B - Bref, axis=1)
    return float(d.max() / np.linalg.norm(Bref, axis=1).max()), \
        float(np.linalg.norm(B - Bref) / np.linalg.norm(Bref))


def cpu_baseline(dp, solver, gpu_stats, cg_iters, threads):
    """The reference path on the host cores, stage by stage as the reference
    runs it: Biot-Savart (oracle C restatement, OpenMP, ``threads``), one
    assembly (restatement element loops, 1 thread; scipy COO->CSR and sparse
    sums, the reference's own calls) and ``cg_iters`` iterations of scipy's cg
    with the Jacobi preconditioner (the reference's own solver) on the full
    1M-tet system; the solve time is extrapolated with the FP64 Jacobi GPU
    solve's Newton and CG counts (the same algorithm, so the same counts).
    The restatement's element loop is faster than the reference's compiled
    loop and every other stage is the reference's own code, so the figure does
    not flatter the GPU.  Beside it, one complete, unextrapolated CPU solve of
    the ~50k-tet dipole (the CPU-runnable configuration): Newton from A = 0 to
    1e-8 with scipy's Jacobi-cg."""
    from oracle import oracle as orc
    from snoopy_amd.magnets import Dipole, mesh_size_for
    t0 = time.perf_counter()
    prob = orc.OracleProblem(solver, threads=threads, csr='scipy')
    t_bs = prob.t_source
    n = prob.n
    t1 = time.perf_counter()
    K, dK, rhs = prob.assemble(np.zeros(n))
    J = (dK + K).tocsr()
    t_asm = time.perf_counter() - t1
    t2 = time.perf_counter()
    _, it, _ = orc.scipy_pcg(J, rhs, np.zeros(n), prob.mask, rtol=1e-30, atol=0.0,
                             maxiter=cg_iters)
    t_cg = (time.perf_counter() - t2) / max(it, 1)
    n_asm = len(gpu_stats['residual_history']) + 1      # residual evaluations incl. R(0)
    n_cg = int(sum(gpu_stats['cg_iterations']))
    t_solve = t_bs + n_asm * t_asm + n_cg * t_cg
    # config 1: a complete CPU solve, not extrapolated
    h1, _ = mesh_size_for(5.0e4)
    d1 = Dipole(h=h1)
    s1 = d1.solver(quad_order=2, verbose=False)
    t3 = time.perf_counter()
    p1 = orc.OracleProblem(s1, threads=threads, csr='scipy')
    _, st1 = orc.residual_newton(p1, rtol=1e-8, linear=orc.scipy_pcg)
    t_full = time.perf_counter() - t3
    sample_s = time.perf_counter() - t0
    return {'value': 1.0 / t_solve, 'unit': 'solves/s', 'cores': threads, 'kind': 'port',
            'sample': ('reference path on the host: Biot-Savart at %d iron quadrature points '
                       '(oracle C restatement, %d OpenMP threads, %.2f s), one assembly of %d '
                       'tets (restatement element loops + scipy COO->CSR and sums, 1 thread, '
                       '%.2f s), %d iterations of scipy cg with the Jacobi preconditioner '
                       '(1 thread, %.1f ms each); solve time extrapolated with the FP64 GPU '
                       'solve counts (%d assemblies, %d CG iterations) = %.1f s; sample wall '
                       'time %.1f s; cpu_count=%s, %s'
                       % (prob.src[0].shape[0], threads, t_bs, dp.num_tets, t_asm, it,
                          1e3 * t_cg, n_asm, n_cg, t_solve, sample_s, os.cpu_count(),
                          platform.processor() or platform.machine())),
            'calibration': 'profiles/cpu_calibration_dipole1M.md',
            'config1_full_solve': {'tets': d1.num_tets, 'seconds': t_full,
                                   'solves_per_s': 1.0 / t_full,
                                   'newton_steps': len(st1['cg_iterations']),
                                   'cg_iterations': st1['cg_iterations'],
                                   'final_residual': st1['residual_history'][-1]}}


def main():
    args = parse()
    if args.ws_train is None:
        args.ws_train = 2048 if args.workload == 'warmstart' else 512
    if args.streams is None:
        # measured best: 8 streams for the 1M-tet variants (26.4 vs 24.9 solves/s at 4),
        # 4 for the 64 small variants
        args.streams = 8 if args.workload == 'dipole1m' else 4
    if args.gpus != 1:
        print('bench.py runs on one GPU only (this build has no multi-GPU path); use --gpus 1',
              file=sys.stderr)
        sys.exit(2)
    import torch
    if args.copy_calib:
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({'copy_calib_bytes_read': src.numel() * 4,
                          'copy_calib_bytes_written': dst.numel() * 4}))
        return
    import __graft_entry__
    __graft_entry__.load_package()
    if args.workload == 'batch64':
        return bench_batch(args)
    if args.workload == 'chain6':
        return bench_chain(args)
    if args.workload == 'warmstart':
        return bench_warmstart(args)
    if args.workload == 'gemm':
        return bench_gemm(args)
    from snoopy_amd import _native as nat
    from snoopy_amd.magnets import Dipole, mesh_size_for
    from snoopy_amd.red_mvp_solver import _CG

    h, _ = mesh_size_for(args.tets)
    t0 = time.perf_counter()
    dp = Dipole(h=h)
    solver = dp.solver(quad_order=2, verbose=False)
    asm = solver.curl_curl_factory
    asm._device_state()
    log('mesh: h=%.4f tets=%d edges=%d nnz=%d iron=%.3f setup %.1fs'
        % (h, dp.num_tets, asm.num_dofs, asm.nnz, dp.iron_fraction(), time.perf_counter() - t0))

    if args.lanes == 0:
        args.lanes = solver.tuned_lanes('mixed')
        log('autotuned lanes per row: %s' % solver.stats['autotune_lanes'])
    if args.tune_grids:
        solver.tuned_element_grids(args.lanes)
        log('autotuned element-operator CG launch shapes: %s' % solver.stats['autotune_grids'])
    # learned warm start: a family of this dipole's mesh (same topology, grid
    # lines moved with the 8 parameters), snapshots from the solver, energy
    # POD and the BF16 network -- all set-up, outside the timed regions
    ws, ws_stats, p0 = None, None, None
    if args.batch_warm == 'learned':
        from snoopy_amd.warm_start import PARAM_NAMES, MorphFamily, WarmStart
        t = time.perf_counter()
        fam = MorphFamily(base=dp)
        ws = WarmStart.train(fam, n_train=args.ws_train, r=args.ws_modes, epochs=args.ws_epochs,
                             rtol=args.rtol, workers=args.ws_workers, log=log)
        ws.fsolver = ws.fsolvers = None          # the training solvers are not needed again
        ws.to_device(asm, r_galerkin=args.ws_galerkin)
        ws.galerkin_steps = args.ws_galerkin_steps
        torch.cuda.synchronize()
        t_c = time.perf_counter()
        ws.precompute_constant(solver, lanes_per_row=args.lanes)
        torch.cuda.synchronize()
        t_const = time.perf_counter() - t_c
        ws_stats = {k: (float(v) if isinstance(v, (int, float, np.floating)) else v)
                    for k, v in ws.train_stats.items()}
        ws_stats['t_setup_total'] = time.perf_counter() - t
        ws_stats['galerkin_modes'] = ws.r_galerkin
        ws_stats['t_constant_H_per_mesh'] = t_const
        p0 = {k: dp.p[k] for k in PARAM_NAMES}
        log('learned warm start trained: %s' % ws_stats)

    ams_opts = {k: v for k, v in (('weight', args.ams_weight), ('omega', args.ams_omega),
                                   ('theta', args.ams_theta), ('passes', args.ams_passes),
                                   ('coarse_max', args.ams_coarse)) if v is not None}

    def one_solve(precision, rtol, use_warm, precond=None):
        precond = precond or args.precond
        solver._src = None                      # Biot-Savart is part of every solve
        if precond == 'ams':
            solver.newton_driver(precision, args.lanes, 32, 'ams').ams_options = ams_opts
        if use_warm and ws is not None:
            from snoopy_amd.warm_start import solve_with_warm_start
            return solve_with_warm_start(solver, ws, p0, rtol=rtol, lanes_per_row=args.lanes)
        return solver.solve_newton(precision, rtol=rtol, lanes_per_row=args.lanes, log=log,
                                   precond=precond)

    def timed_once(precision, precond):
        one_solve(precision, args.rtol, False, precond)
        torch.cuda.synchronize()
        t = time.perf_counter()
        x = one_solve(precision, args.rtol, False, precond)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        st = dict(solver.stats)
        log('%s/%s: %.3f s, newton %d, cg %s' % (precision, precond, dt, st['newton_steps'],
                                                 st['cg_iterations']))
        return dt, st, x

    # FP64 GPU solves: with the reference's Jacobi preconditioner (the
    # drop-in method) and with the chosen preconditioner
    t_jac, st_jac, _ = timed_once('fp64', 'jacobi')
    if args.precond == 'jacobi':
        t_fp64, st64, x64 = t_jac, st_jac, _
    else:
        t_fp64, st64, x64 = timed_once('fp64', args.precond)

    def timed_steps(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / args.steps, out

    # single solves in mixed precision: from A = 0, and from the learned start
    t_cold, xc = timed_steps(lambda: one_solve('mixed', args.rtol, False))
    st_cold = dict(solver.stats)
    log('mixed from zero: %.3f s/solve, newton %d, cg %s' % (t_cold, st_cold['newton_steps'],
                                                           st_cold['cg_iterations']))
    t_mixed, xm, stm = t_cold, xc, st_cold
    if ws is not None:
        t_mixed, xm = timed_steps(lambda: one_solve('mixed', args.rtol, True))
        stm = dict(solver.stats)
        log('mixed, learned start: %.3f s/solve, newton %d, cg %s, start %s'
            % (t_mixed, stm['newton_steps'], stm['cg_iterations'], stm.get('warm_start')))

    # batched: variants of the dipole (coil current density spread over
    # 1.2e6..1.8e6 A/m^2, the metric's 1.5e6 in the middle), each a complete
    # solve.  The chosen form (--batch-solver variant) runs them in lock step
    # through the variant-minor CG: one set of kernel launches per CG
    # iteration for all variants.  The chosen start and the start from zero
    # are timed on it; beside it the same variants on --streams concurrent
    # HIP streams (one solver each), with the learned, zero and anchor starts.
    t_batch, bout, bdps, bsolver = None, None, None, None
    batch_runs = {}
    if args.batch > 0:
        from snoopy_amd.batch import ConcurrentSolver, VariantBatchSolver
        bdps = [Dipole(h=h, J=float(J)) for J in np.linspace(1.2e6, 1.8e6, args.batch)]

        def record(key, mode, tb, res, form):
            its = [int(sum(o[1]['cg_iterations'])) for o in res]
            batch_runs[key] = {'s': tb, 'solves_per_s': args.batch / tb, 'cg': its,
                               'form': form,
                               'converged': all(bool(o[1]['converged']) for o in res)}
            if mode == 'learned':
                batch_runs[key]['start_residual'] = [
                    o[1].get('warm_start', {}).get('r_start_rel') for o in res]
                batch_runs[key]['cg_per_step_first'] = res[0][1]['cg_iterations']
                batch_runs[key]['residual_history_first'] = res[0][1]['residual_history']
            log('batch %s (%s): %d variants, %.3f s per batch (%.1f solves/s), cg %s'
                % (form, mode, args.batch, tb, args.batch / tb, its))

        runs = []
        if args.batch_solver == 'variant' and args.batch <= 8:
            vb = VariantBatchSolver(bdps, lanes_per_row=args.lanes, workers=args.streams)
            vmodes = [m for m in [args.batch_warm] + ([] if args.batch_only else ['none'])
                      if m != 'anchor']
            runs += [('variant', m, vb) for m in dict.fromkeys(vmodes)]
        if args.batch_solver == 'concurrent' or not args.batch_only:
            cs = ConcurrentSolver(bdps, workers=args.streams, precision='mixed',
                                  precond=args.precond, lanes_per_row=args.lanes,
                                  ams_options=ams_opts)
            cmodes = [args.batch_warm] + ([] if args.batch_only else
                                          [m for m in ('none', 'anchor') if m != args.batch_warm])
            runs += [('concurrent', m, cs) for m in cmodes]
        chosen = 'variant' if any(r[0] == 'variant' for r in runs) else 'concurrent'
        for form, mode, bs in runs:
            bs.set_warm_start(None if mode == 'none' else mode, ws if mode == 'learned' else None)

            def one_batch():
                for s_ in bs.solvers:
                    s_._src = None           # every solve computes its own sources
                return bs.solve(args.rtol)
            tb, res = timed_steps(one_batch)
            key = mode if form == chosen else 'concurrent_' + mode
            record(key, mode, tb, res, form)
            if form == 'variant':
                # the last timed batch: sources, starts and first residuals; the
                # batched CG; the residuals after each CG
                batch_runs[key]['phases_s'] = {k: bs.stats[k] for k in
                                               ('t_sources', 't_prepare', 't_cg',
                                                't_advance')}
            if form == chosen and mode == args.batch_warm:
                t_batch, bout, bsolver = tb, res, bs
    if args.quick:
        print(json.dumps({'quick': True, 'fp64_s': t_fp64, 'mixed_s': t_mixed,
                          'mixed_cold_s': t_cold, 'fp64_jacobi_s': t_jac, 'batch': batch_runs,
                          'learned_warm_start': ws_stats,
                          'fp64_stats': {k: v for k, v in st64.items()},
                          'mixed_stats': {k: v for k, v in stm.items()}}, default=str),
              flush=True)
        return
    # FP64 reference for the B error: Newton down to the FP64 residual floor
    log('fp64 reference solve (rtol 1e-11 or the residual floor)')
    xref = solver.solve_newton('fp64', rtol=1e-11, lanes_per_row=args.lanes, log=log)
    stref = dict(solver.stats)
    probes = dp.probes(10000, seed=2)
    Bref = probe_B(dp, solver, xref, probes)
    Bm = probe_B(dp, solver, xm, probes)
    err_cold, _ = b_error(probe_B(dp, solver, xc, probes), Bref)
    B64 = probe_B(dp, solver, x64, probes)
    err_m, err_m_l2 = b_error(Bm, Bref)
    err_64, _ = b_error(B64, Bref)
    # every batched variant against its own FP64 reference
    batch_err = None
    if bout is not None:
        batch_err = 0.0
        for bd, bs, (xb, _) in zip(bdps, bsolver.solvers, bout):
            xr = bs.solve_newton('fp64', rtol=1e-11, lanes_per_row=args.lanes, precond='ams')
            batch_err = max(batch_err, b_error(probe_B(bd, bs, xb, probes),
                                               probe_B(bd, bs, xr, probes))[0])
        log('batch B error (max over variants): %.2e' % batch_err)

    # roofline of the dominant operation: the CG's product with the element
    # operator and the CG update (cg_elem_kernel, then cg_edge_kernel), 200
    # back-to-back launch pairs on the solver's stream between two HIP events
    # (the kernel trace summary in profiles/ gives the in-solve durations too)
    drv = solver.newton_driver('mixed', args.lanes, 32, args.precond)
    lib = nat.load()
    ms = ctypes.c_double()
    nat.check(lib.sn_cg_time_spmv(drv.cg_el.h, nat.ptr(drv.Mr), nat.ptr(drv.mask), 200,
                                  ctypes.byref(ms), nat.stream_handle()))
    n, nnz = asm.num_dofs, asm.nnz
    N = dp.num_tets
    # per element: 6 signed DOF ids, 6 matrix values, 6 products written
    # (cg_elem_kernel) and read back with their 6 list entries (cg_edge_kernel);
    # per DOF: the {z, p_old} pair in each kernel, DOF range, mask, p written,
    # r and x read and written
    op_bytes = 4 * 6 * N * 5 + (8 + 8 + 4 + 1 + 4 + 8 + 8) * n + 4
    op_gbs = op_bytes / (ms.value * 1e-3) / 1e9
    # the CSR SpMV of the same Jacobian rounded after assembly (jacobian
    # "csr"), for comparison
    drv_csr = solver.newton_driver('mixed', args.lanes, 32, args.precond, jacobian='csr')
    drv_csr.retarget(solver.source_fields_device())
    drv_csr.residual(xm)
    ms_csr = ctypes.c_double()
    nat.check(lib.sn_cg_time_spmv(drv_csr.cg.h, nat.ptr(drv_csr.J), nat.ptr(drv_csr.mask), 200,
                                  ctypes.byref(ms_csr), nat.stream_handle()))
    ib = lib.sn_cg_index_bytes(drv_csr.cg.h)   # 2: 16-bit column offsets, 4: int32
    csr_bytes = nnz * (4 + ib) + 4 * (n + 1) + 4 * 4 * n + n
    # element kernel (residual + element matrices, mode 2) timed with HIP
    # events on its stream
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    reps = 10
    for _ in range(reps):
        asm.newton_local_device(xm, 2, solver.source_fields_device(), 2, True, drv.Mr, drv.loc6)
    ev1.record()
    torch.cuda.synchronize()
    asm_ms = ev0.elapsed_time(ev1) / reps
    n_iron = int(np.sum(dp.mesh.tags == 1))
    asm_bytes = 16 * N + 24 * N + 4 * N + 24 * asm.tet_mesh.num_nodes + 8 * n + \
        4 * 3 * 8 * n_iron + 6 * 4 * N + 6 * 8 * N

    log('roofline: element operator product %.4f ms (%.0f GB/s), csr spmv %.4f ms, '
        'element kernel %.3f ms' % (ms.value, op_gbs, ms_csr.value, asm_ms))
    def pmc_traffic(name, **match):
        # HBM bytes per launch pair from the PMC runs summarised in profiles/
        tfile = os.path.join(ROOT, 'profiles', name)
        if os.path.exists(tfile):
            tr = json.load(open(tfile))
            if all(tr.get(k) == v for k, v in match.items()):
                return tr['hbm_bytes_per_product']
        return None
    traffic = pmc_traffic('cg_element_product_dipole1M_traffic.json', tets=N)
    roof_single = {'bound': 'hbm', 'achieved': op_gbs, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                   'frac': op_gbs / HBM_PEAK_GBS, 'traffic': traffic,
                   'kernel': 'cg_elem_kernel<float> + cg_edge_kernel<float,float,false> '
                             '(element operator product with the CG update, one variant)',
                   'avg_ms': ms.value, 'launches': 200, 'algorithmic_bytes': op_bytes}
    roof = roof_single
    if bsolver is not None and getattr(bsolver, 'h', None) is not None:
        # the batch's own product: the variant-minor element and DOF passes,
        # 200 back-to-back pairs for all variants at once
        V = bsolver.nv
        Mr_v = torch.stack([bsolver._driver(min(v, bsolver.nvar - 1)).Mr for v in range(V)],
                           dim=1).contiguous()
        ms_v = ctypes.c_double()
        nat.check(lib.sn_vcg_time_pair(bsolver.h, nat.ptr(Mr_v), nat.ptr(drv.mask), 200,
                                       ctypes.byref(ms_v), nat.stream_handle()))
        del Mr_v
        # per element: 6 DOF ids and 6 list entries (shared), per variant 6
        # matrix values, 6 products written and read back; per DOF: range and
        # mask (shared), per variant the {z, p} pair in each pass, p written,
        # r and x read and written
        v_bytes = N * (24 + 24 + 72 * V) + n * (5 + 36 * V) + 4
        v_gbs = v_bytes / (ms_v.value * 1e-3) / 1e9
        log('roofline: variant-minor element pair (%d variants) %.4f ms (%.0f GB/s)'
            % (V, ms_v.value, v_gbs))
        roof = {'bound': 'hbm', 'achieved': v_gbs, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                'frac': v_gbs / HBM_PEAK_GBS,
                'traffic': pmc_traffic('vcg_element_pair_dipole1M_traffic.json', tets=N,
                                       variants=V),
                'kernel': 'vcg_elem_kernel<%d> + vcg_edge_kernel<%d,false> (the batch\'s '
                          'element operator product with the CG update, %d variants at once)'
                          % (V, V, V),
                'avg_ms': ms_v.value, 'launches': 200, 'algorithmic_bytes': v_bytes}
    cpu = None
    if not args.no_cpu:
        log('cpu baseline leg')
        try:
            cpu = cpu_baseline(dp, solver, st_jac, args.cpu_cg_iters, args.cpu_threads)
        except Exception as e:
            log('cpu baseline failed:', repr(e))

    # the chosen form: batched solves (concurrent variants) when --batch > 0
    start_desc = {
        'learned': ('learned warm start: the BF16 network (8 -> 4x512 GELU -> %d POD '
                    'coefficients) trained before the timed region on %d solver-generated '
                    'solutions of the morph family of this mesh, one reduced Newton step in the '
                    'leading %d modes, kept when its residual is below |R(0)|'
                    % (args.ws_modes, args.ws_train, ws.r_galerkin if ws else 0)),
        'anchor': ('anchor start (J-only variants: the median-J variant from zero, the others '
                   'from its solution scaled by the current-density ratio)'),
        'none': 'every variant from A = 0'}
    if t_batch is not None:
        value, step_ms = args.batch / t_batch, 1e3 * t_batch
        form = ('in lock step, one variant-minor CG for all variants (their own element '
                'operators, forcing terms and stop tests; one auxiliary-space preconditioner)'
                if batch_runs.get(args.batch_warm, {}).get('form') == 'variant'
                else 'on %d concurrent HIP streams' % args.streams)
        workload = ('%d variants (coil current density 1.2e6-1.8e6 A/m^2) of the window-frame '
                    'dipole, %d tets, %d edge DOFs, nnz %d each, Q=4, B-spline-in-B^2 iron, '
                    'Newton to 1e-8 of each variant\'s own |R(0)|, %s, %s; one step = the batch'
                    % (args.batch, N, n, nnz, form, start_desc[args.batch_warm]))
    else:
        value, step_ms = 1.0 / t_mixed, 1e3 * t_mixed
        workload = ('window-frame dipole, %d tets, %d edge DOFs, nnz %d, Q=4, '
                    'B-spline-in-B^2 iron, Newton to 1e-8' % (N, n, nnz))
    out = {
        'metric': METRIC,
        'value': value,
        'unit': 'solves/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': step_ms,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32+f64',
        'data': 'synthetic',
        'config': {'workload': workload,
                   'tets': N, 'edges': n, 'nnz': nnz, 'iron_fraction': dp.iron_fraction(),
                   'batch': args.batch, 'streams': args.streams,
                   'precision': 'mixed (FP32 element-operator Jacobian and CG, FP64 residual '
                                'and dots)',
                   'preconditioner': args.precond,
                   'warm_start': ws is not None,
                   'batch_warm_start': start_desc[args.batch_warm],
                   'parallelism': 'single GPU'},
        'autotune': {k: solver.stats[k] for k in ('autotune_lanes', 'autotune_grids')
                     if k in solver.stats},
        'single_solve_per_s': 1.0 / t_mixed,
        'single_cold_solves_per_s': 1.0 / t_cold,
        'single_cold_cg_iterations': st_cold['cg_iterations'],
        'learned_warm_start': ws_stats,
        'learned_start_residual': (stm.get('warm_start') or {}).get('r_start_rel'),
        'batch_B_rel_err_max': batch_err,
        'batch_runs': batch_runs,
        'batch_cold_solves_per_s': batch_runs.get('none', {}).get('solves_per_s'),
        'batch_concurrent_solves_per_s': batch_runs.get('concurrent_' + args.batch_warm,
                                                        {}).get('solves_per_s'),
        'batch_anchor_solves_per_s_J_only': (batch_runs.get('anchor') or
                                             batch_runs.get('concurrent_anchor') or
                                             {}).get('solves_per_s'),
        'batch_learned_solves_per_s': batch_runs.get('learned', {}).get('solves_per_s'),
        'batch_cg_iterations': ([int(sum(o[1]['cg_iterations'])) for o in bout]
                                if bout is not None else None),
        'fp64_solves_per_s': 1.0 / t_fp64,
        'fp64_jacobi_solves_per_s': 1.0 / t_jac,
        'speedup_vs_fp64': value * t_fp64,
        'speedup_vs_fp64_jacobi': value * t_jac,
        'speedup_single_vs_fp64_jacobi': t_jac / t_mixed,
        'cg_iterations_fp64_jacobi': st_jac['cg_iterations'],
        'ams_levels': stm.get('ams_levels'),
        'ams_setup_s': stm.get('ams_setup_s'),
        'B_rel_err_max': err_m,
        'B_rel_err_max_cold': err_cold,
        'B_rel_err_l2': err_m_l2,
        'B_rel_err_fp64': err_64,
        'B_reference_residual': stref['residual_history'][-1],
        'final_residual': stm['residual_history'][-1],
        'newton_steps': stm['newton_steps'],
        'cg_iterations': stm['cg_iterations'],
        'newton_steps_fp64': st64['newton_steps'],
        'cg_iterations_fp64': st64['cg_iterations'],
        'roofline': roof,
        'roofline_single_variant': roof_single,
        'csr_spmv': {'kernel': 'cg_spmv_kernel<float,float,%d,%s>'
                               % (args.lanes, 'short' if ib == 2 else 'int'),
                     'avg_ms': ms_csr.value, 'algorithmic_bytes': csr_bytes,
                     'GB/s': csr_bytes / (ms_csr.value * 1e-3) / 1e9},
        'assembly_kernel': {'avg_ms': asm_ms, 'algorithmic_bytes': asm_bytes,
                            'GB/s': asm_bytes / (asm_ms * 1e-3) / 1e9},
        'cpu_baseline': cpu,
    }
    if cpu:
        out['speedup_vs_cpu'] = value / cpu['value']
    if not args.no_extra:
        # the batched and multi-magnet workloads on the same box, same run
        bsolver = bout = bdps = runs = vb = cs = None      # noqa: F841
        import gc
        gc.collect()
        torch.cuda.empty_cache()
        if ws is not None:
            out['family8'] = measure_family8(args, ws, steps=2)
        out['batch64'] = measure_batch64(args, steps=2)
        out['chain6'] = measure_chain6(args)
    print(json.dumps(out), flush=True)


def measure_family8(args, ws, steps=2):
    """8 seeded variants of the learned model's own family at the bench's
    size: all 8 parameters drawn (yoke and aperture sizes, length, J, B_sat,
    mu_r), each variant on its own moved copy of the bench mesh, solved
    concurrently from the learned start and from zero on the same solver
    objects, each to 1e-8 of its own |R(0)|.  Set-up (the moved meshes,
    preconditioner plans, the learned start's per-mesh terms) is timed
    apart."""
    import torch
    from snoopy_amd.batch import ConcurrentSolver
    from snoopy_amd.warm_start import sample_params
    fam = ws.family
    t0 = time.perf_counter()
    dps = []
    for k in range(8):
        p = sample_params(9100 + k)
        m = fam.private_model()
        m.tet_mesh.nodes[:] = fam.nodes(dict(fam.p0, **p))
        dps.append(fam.variant(p, m))
    cs = ConcurrentSolver(dps, workers=args.streams, precision='mixed', precond='ams',
                          lanes_per_row=args.lanes, warm_start='learned', model=ws)
    cs.solve(rtol=args.rtol)                  # first solve: device set-up and AMS plans
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    runs = {}
    for mode in ('learned', None):
        cs.set_warm_start(mode, ws if mode else None)
        cs.solve(rtol=args.rtol)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            res = cs.solve(rtol=args.rtol)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / steps
        runs[mode or 'none'] = {
            's': dt, 'solves_per_s': len(dps) / dt,
            'cg': [int(sum(o[1]['cg_iterations'])) for o in res],
            'converged': all(bool(o[1]['converged']) for o in res),
            'worst_final_residual': max(float(o[1]['residual_history'][-1]) for o in res)}
        if mode:
            runs['learned']['start_residual'] = [o[1].get('warm_start', {}).get('r_start_rel')
                                                 for o in res]
    del cs
    # the same variants in lock step through the variant-minor CG (one shared
    # preconditioner: the anchor variant's geometry and reluctivity)
    from snoopy_amd.batch import VariantBatchSolver
    vb = VariantBatchSolver(dps, lanes_per_row=args.lanes, workers=args.streams,
                            warm_start='learned', model=ws)
    vb.solve(rtol=args.rtol)                  # first solve: the shared preconditioner's plan
    for mode in ('learned', None):
        vb.set_warm_start(mode, ws if mode else None)
        vb.solve(rtol=args.rtol)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            for s_ in vb.solvers:
                s_._src = None
            res = vb.solve(rtol=args.rtol)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / steps
        runs['variant_' + (mode or 'none')] = {
            's': dt, 'solves_per_s': len(dps) / dt,
            'cg': [int(sum(o[1]['cg_iterations'])) for o in res],
            'converged': all(bool(o[1]['converged']) for o in res),
            'fallback': [bool(o[1]['fallback']) for o in res],
            'worst_final_residual': max(float(o[1]['residual_history'][-1]) for o in res)}
    out = {'variants': len(dps), 'tets': dps[0].num_tets, 'streams': args.streams,
           'params_seeds': [9100, 9107], 'setup_s': t_setup, 'runs': runs,
           'learned_solves_per_s': runs['learned']['solves_per_s'],
           'cold_solves_per_s': runs['none']['solves_per_s'],
           'variant_learned_solves_per_s': runs['variant_learned']['solves_per_s'],
           'variant_cold_solves_per_s': runs['variant_none']['solves_per_s']}
    log('family8: %s' % out)
    del vb
    import gc
    gc.collect()
    torch.cuda.empty_cache()
    return out


def measure_batch64(args, steps=None):
    """The batched workload (magnets.batch_variants: 64 seeded dipoles of
    20k-400k tets) solved concurrently on 8 HIP streams from one queue, mixed precision with
    the auxiliary-space preconditioner, each to 1e-8 of its own |R(0)|: set-up
    time (meshes, patterns, preconditioner plans), timed batch, and every
    variant's B at 1000 aperture probes against its own FP64 + AMS solve."""
    import torch
    from snoopy_amd.batch import ConcurrentSolver
    from snoopy_amd.magnets import batch_variants
    steps = steps or args.steps
    t0 = time.perf_counter()
    dps = batch_variants(64)
    t_mesh = time.perf_counter() - t0
    cs = ConcurrentSolver(dps, workers=8, precision='mixed', precond='ams')
    cs.solve(rtol=args.rtol)                  # first solve: device set-up and AMS plans
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    t = time.perf_counter()
    for _ in range(steps):
        res = cs.solve(rtol=args.rtol)
    torch.cuda.synchronize()
    t_solve = (time.perf_counter() - t) / steps
    err = 0.0
    for k, (d, s, (x, _)) in enumerate(zip(dps, cs.solvers, res)):
        xr = s.solve_newton('fp64', rtol=1e-10, precond='ams')
        pts = d.probes(1000, seed=k)
        err = max(err, b_error(probe_B(d, s, x, pts), probe_B(d, s, xr, pts))[0])
    tets = [d.num_tets for d in dps]
    ams_setup = float(sum(r[1].get('ams_setup_s') or 0.0 for r in res))
    out = {'variants': len(dps), 'tets_min': min(tets), 'tets_max': max(tets),
           'tets_total': int(sum(tets)), 'streams': 8, 'solve_s': t_solve,
           'solves_per_s': len(dps) / t_solve,
           'setup_s': t_setup - t_solve, 'mesh_s': t_mesh, 'ams_plan_s_total': ams_setup,
           'all_converged': all(bool(r[1]['converged']) for r in res),
           'worst_final_residual': float(max(r[1]['residual_history'][-1] for r in res)),
           'cg_total': int(sum(sum(r[1]['cg_iterations']) for r in res)),
           'B_rel_err_max_vs_fp64_ams': err}
    log('batch64: %s' % out)
    del cs, res
    torch.cuda.empty_cache()
    return out


def measure_chain6(args, jacobi=False):
    """The 6-yoke chain (magnets.chain_dipole, about 6M tets): the mixed solve
    (AMS) timed, the FP64 + AMS solve timed and used as the B reference at
    5000 beam-axis points, optionally the FP64 + Jacobi solve."""
    import torch
    from snoopy_amd.magnets import chain_dipole
    t0 = time.perf_counter()
    dp = chain_dipole(6.0e6)
    s = dp.solver(quad_order=2, verbose=False)
    s.curl_curl_factory._device_state()
    t_mesh = time.perf_counter() - t0
    out = {'tets': dp.num_tets, 'edges': s.curl_curl_factory.num_dofs,
           'nnz': s.curl_curl_factory.nnz, 'mesh_s': t_mesh}
    xs = {}
    forms = [('mixed', 'ams'), ('fp64', 'ams')] + ([('fp64', 'jacobi')] if jacobi else [])
    for prec, pc in forms:
        s._src = None
        s.solve_newton(prec, rtol=args.rtol, precond=pc)      # untimed: set-up
        torch.cuda.synchronize()
        s._src = None
        t = time.perf_counter()
        xs[(prec, pc)] = s.solve_newton(prec, rtol=args.rtol, precond=pc)
        torch.cuda.synchronize()
        st = dict(s.stats)
        out['%s_%s_s' % (prec, pc)] = time.perf_counter() - t
        out['%s_%s_cg' % (prec, pc)] = st['cg_iterations']
        if prec == 'mixed':
            out['mixed_converged'] = bool(st['converged'])
            out['mixed_final_residual'] = st['residual_history'][-1]
            out['ams_plan_s'] = st.get('ams_setup_s')
    xr = s.solve_newton('fp64', rtol=1e-10, precond='ams')
    axis = dp.beam_axis(5000)
    Bref = probe_B(dp, s, xr, axis)
    out['B_axis_rel_err_mixed_vs_fp64_ams'] = b_error(probe_B(dp, s, xs[('mixed', 'ams')], axis),
                                                      Bref)[0]
    out['solves_per_s'] = 1.0 / out['mixed_ams_s']
    out['speedup_vs_fp64_ams'] = out['fp64_ams_s'] / out['mixed_ams_s']
    log('chain6: %s' % out)
    del s, xs, xr
    torch.cuda.empty_cache()
    return out


def bench_batch(args):
    """Seeded dipole variants (batched-workload ranges), meshes of 20k-400k
    tets, Newton-solved together (one joined block-diagonal system, one
    segment-aware CG graph) and one at a time; mixed precision, 1e-8."""
    import torch
    from snoopy_amd.batch import BatchedSolver
    from snoopy_amd.magnets import Dipole, mesh_size_for, sample_variant
    rng = np.random.default_rng(12345)
    dps = []
    t0 = time.perf_counter()
    for seed in range(args.variants):
        v = sample_variant(seed)
        target = float(np.exp(rng.uniform(np.log(2e4), np.log(4e5))))
        h, _ = mesh_size_for(target, **v)
        dps.append(Dipole(h=h, **v))
    tets = [d.num_tets for d in dps]
    log('batch: %d variants, %d..%d tets, total %d (mesh %.1f s)'
        % (len(dps), min(tets), max(tets), sum(tets), time.perf_counter() - t0))
    bs = BatchedSolver(dps, precision='mixed', lanes_per_row=8, precond=args.precond)
    log('joined: %d edges, nnz %d, setup %.1f s' % (bs.n, bs.asm.nnz, bs.t_setup))
    bs.solve(rtol=args.rtol, log=log)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        bs.solve(rtol=args.rtol)
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t) / args.steps
    conv = int(bs.stats['converged'].sum())
    log('batched: %.3f s for %d variants (%d converged), newton %d'
        % (t_batch, len(dps), conv, bs.stats['newton_steps']))
    # the same variants one at a time
    t_seq = 0.0
    for d in dps:
        s = d.solver(quad_order=2, verbose=False)
        s.solve_newton('mixed', rtol=args.rtol, lanes_per_row=8, precond=args.precond)
        torch.cuda.synchronize()
        t = time.perf_counter()
        s.solve_newton('mixed', rtol=args.rtol, lanes_per_row=8, precond=args.precond)
        torch.cuda.synchronize()
        t_seq += time.perf_counter() - t
    log('one at a time: %.3f s' % t_seq)
    # the same variants on several streams at once
    from snoopy_amd.batch import ConcurrentSolver
    cs = ConcurrentSolver(dps, workers=args.streams, precision='mixed', precond=args.precond)
    cs.solve(rtol=args.rtol)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        res = cs.solve(rtol=args.rtol)
    torch.cuda.synchronize()
    t_conc = (time.perf_counter() - t) / args.steps
    conc_ok = all(r[1]['converged'] for r in res)
    log('%d streams: %.3f s (%d/%d converged)' % (args.streams, t_conc,
                                                  sum(r[1]['converged'] for r in res), len(res)))
    print(json.dumps({
        'metric': 'variant solves/s, %d seeded dipoles of 20k-400k tets, Newton to %g'
                  % (len(dps), args.rtol),
        'value': len(dps) / min(t_batch, t_conc), 'unit': 'solves/s', 'n_gpus': 1,
        'steps': args.steps,
        'higher_is_better': True, 'dtype': 'f32+f64', 'data': 'synthetic',
        'config': {'workload': 'batch of %d seeded dipole variants' % len(dps),
                   'tets_total': sum(tets), 'tets_min': min(tets), 'tets_max': max(tets),
                   'edges_total': bs.n, 'nnz_total': bs.asm.nnz, 'precond': args.precond},
        'batched_s': t_batch, 'one_at_a_time_s': t_seq, 'speedup': t_seq / t_batch,
        'streams': args.streams, 'concurrent_s': t_conc, 'concurrent_converged': conc_ok,
        'concurrent_solves_per_s': len(dps) / t_conc,
        'concurrent_speedup_vs_one_at_a_time': t_seq / t_conc,
        'joined_solves_per_s': len(dps) / t_batch,
        'converged': conv, 'newton_steps': bs.stats['newton_steps']}), flush=True)


def bench_warmstart(args):
    """Learned warm start on the morph family of the 205k-tet reference mesh
    (fixed topology, grid lines following the 8 parameters): --ws-train
    solver-generated snapshots, energy POD with --ws-modes modes, the BF16
    network 8 -> 4x512 GELU -> modes trained by this project's kernels; then
    --ws-test held-out seeds solved from A = 0 and from the learned start
    (network, one reduced Newton step, residual judge), mixed precision,
    Newton to --rtol.  Both timings include the source field; the learned
    one includes the network, the expansion and the reduced step."""
    import torch
    from snoopy_amd.warm_start import (MorphFamily, WarmStart, sample_params,
                                       solve_with_warm_start)
    fam = MorphFamily()
    log('warm start family: %d tets, %d edges' % (fam.mesh.num_elements, fam.mesh.num_edges))
    t0 = time.perf_counter()
    ws = WarmStart.train(fam, n_train=args.ws_train, r=args.ws_modes, epochs=args.ws_epochs,
                         rtol=args.rtol, workers=args.ws_workers, log=log)
    ws.to_device(ws.fsolver.asm, r_galerkin=args.ws_galerkin)
    ws.galerkin_steps = args.ws_galerkin_steps
    t_setup = time.perf_counter() - t0
    log('trained in %.1f s: %s' % (t_setup, ws.train_stats))
    fs = ws.fsolver
    sol = fs.solver
    rows = []
    for seed in range(5000, 5000 + args.ws_test):
        p = sample_params(seed)
        fs.set_variant(p)
        sol._src = None
        sol.solve_newton('mixed', rtol=args.rtol, precond='ams')     # untimed: first use of p
        torch.cuda.synchronize()
        sol._src = None
        t = time.perf_counter()
        xa = sol.solve_newton('mixed', rtol=args.rtol, precond='ams')
        torch.cuda.synchronize()
        t_zero, s0 = time.perf_counter() - t, dict(sol.stats)
        sol._src = None
        t = time.perf_counter()
        xb = solve_with_warm_start(sol, ws, p, rtol=args.rtol)
        torch.cuda.synchronize()
        t_warm, s1 = time.perf_counter() - t, dict(sol.stats)
        asm = sol.curl_curl_factory
        Ba, Bb = asm.element_B(xa.cpu().numpy()), asm.element_B(xb.cpu().numpy())
        dB = float(np.abs(Ba - Bb).max() / np.abs(Ba).max())
        info = s1['warm_start']
        rows.append((info['r_start_rel'], info['used'], s0['newton_steps'], s1['newton_steps'],
                     sum(s0['cg_iterations']), sum(s1['cg_iterations']), t_zero, t_warm, dB,
                     bool(s0['converged']), bool(s1['converged'])))
        if seed < 5008 or seed % 32 == 0:
            log('seed %d: start |R|/|R0| %.3e kept %s newton %d -> %d cg %d -> %d  %.4f -> %.4f s'
                '  B diff %.1e' % ((seed,) + rows[-1][:9]))
    a = np.array([r[:9] for r in rows], dtype=float)
    gemm = gemm_rates([(args.ws_train, 512, 512), (8192, 512, 512), (8192, 2048, 2048)])
    print(json.dumps({
        'metric': 'learned warm start on %d held-out variants (%d-tet morph family), Newton to '
                  '%g: time from A = 0 / time from the learned start'
                  % (len(rows), fam.mesh.num_elements, args.rtol),
        'value': float(a[:, 6].sum() / a[:, 7].sum()), 'unit': 'speedup', 'n_gpus': 1,
        'higher_is_better': True, 'dtype': 'f32+f64', 'data': 'synthetic',
        'config': {'workload': 'learned warm start', 'tets': fam.mesh.num_elements,
                   'train_snapshots': args.ws_train, 'test_seeds': len(rows),
                   'pod_modes': ws.pod.U.shape[1], 'galerkin_modes': ws.r_galerkin,
                   'hidden': [512] * 4, 'inner_product': 'K0 (unit-reluctivity curl-curl)'},
        'setup_s': t_setup,
        'train_stats': {k: (float(v) if isinstance(v, (int, float, np.floating)) else v)
                        for k, v in ws.train_stats.items()},
        'start_residual_median': float(np.median(a[:, 0])), 'kept_fraction': float(a[:, 1].mean()),
        'newton_from_zero_mean': float(a[:, 2].mean()), 'newton_warm_mean': float(a[:, 3].mean()),
        'cg_from_zero_mean': float(a[:, 4].mean()), 'cg_warm_mean': float(a[:, 5].mean()),
        'cg_from_zero_median': float(np.median(a[:, 4])),
        'cg_warm_median': float(np.median(a[:, 5])),
        's_from_zero_mean': float(a[:, 6].mean()), 's_warm_mean': float(a[:, 7].mean()),
        'fewer_cg_fraction': float(np.mean(a[:, 5] < a[:, 4])),
        'all_converged': bool(all(r[9] and r[10] for r in rows)),
        'B_diff_max': float(a[:, 8].max()),
        'gemm_bf16': gemm}), flush=True)


BF16_PEAK_TFLOPS = 2500.0   # MI355X dense BF16 matrix peak


def gemm_rates(shapes, reps=50):
    """TFLOP/s of the BF16 MFMA GEMM (C = A B, GELU epilogue as in the MLP's
    hidden layers) at the given (M, N, K), timed with events on its stream."""
    import torch
    from snoopy_amd import _native as nat
    out = []
    for M, N, K in shapes:
        A = torch.randn(M, K, device='cuda').to(torch.bfloat16)
        B = torch.randn(K, N, device='cuda').to(torch.bfloat16)
        bias = torch.zeros(N, device='cuda')
        Z = torch.empty(M, N, device='cuda')
        Cb = torch.empty(M, N, device='cuda', dtype=torch.bfloat16)
        call = lambda: nat.check(nat.load().sn_gemm_bf16(  # noqa: E731
            0, 0, 1, M, N, K, nat.ptr(A), K, nat.ptr(B), N, nat.ptr(bias), None, nat.ptr(Z),
            None, nat.ptr(Cb), N, nat.stream_handle()))
        for _ in range(3):
            call()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            call()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        tf = 2.0 * M * N * K / (ms * 1e-3) / 1e12
        out.append({'M': M, 'N': N, 'K': K, 'avg_us': ms * 1e3, 'TFLOP/s': tf,
                    'frac_of_peak': tf / BF16_PEAK_TFLOPS})
    return out


def bench_gemm(args):
    """The learned warm start's matrix-core work alone: the BF16 GEMM at the
    network's shapes and larger ones, and training epochs of the network of
    the learned warm start (8 -> 4x512 GELU -> 256, batch 2048, forward +
    backward + Adam) on random data."""
    import torch
    from snoopy_amd.mlp import MLP
    shapes = [(2048, 512, 512), (2048, 256, 512), (8192, 512, 512), (8192, 2048, 2048),
              (16384, 4096, 4096)]
    gemm = gemm_rates(shapes)
    for g in gemm:
        log('gemm %dx%dx%d: %.1f us, %.1f TFLOP/s' % (g['M'], g['N'], g['K'], g['avg_us'],
                                                     g['TFLOP/s']))
    B = 2048
    X = torch.rand(B, 8, device='cuda') * 2 - 1
    T = torch.randn(B, 256, device='cuda')
    mlp = MLP((8, 512, 512, 512, 512, 256), seed=0)
    mlp.train(X, T, epochs=5)
    torch.cuda.synchronize()
    epochs = 200
    t = time.perf_counter()
    mlp.train(X, T, epochs=epochs)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / epochs
    flop = 3 * 2 * B * (8 * 512 + 3 * 512 * 512 + 512 * 256)   # forward + 2x backward
    print(json.dumps({'metric': 'BF16 MFMA GEMM and network training on one MI355X',
                      'gemm_bf16': gemm, 'mlp_epoch_ms': dt * 1e3,
                      'mlp_epoch_TFLOP/s': flop / dt / 1e12, 'mlp_batch': B,
                      'mlp_sizes': list(mlp.sizes), 'peak_TFLOP/s': BF16_PEAK_TFLOPS}),
          flush=True)


def bench_chain(args):
    """Six yokes with alternating polarity along z, 0.1 m apart, seeded
    dimensions; mixed vs FP64 GPU solve."""
    import torch
    from snoopy_amd.magnets import Dipole, chain_yokes, chain_mesh_size_for
    yokes = chain_yokes(6, seed=7)
    h, est = chain_mesh_size_for(args.tets if args.tets != 1.0e6 else 6.0e6, yokes)
    t0 = time.perf_counter()
    dp = Dipole(h=h, yokes=yokes)
    solver = dp.solver(quad_order=2, verbose=False)
    asm = solver.curl_curl_factory
    asm._device_state()
    log('chain: %d tets, %d edges, nnz %d, setup %.1f s'
        % (dp.num_tets, asm.num_dofs, asm.nnz, time.perf_counter() - t0))
    out = {}
    forms = [('fp64', 'jacobi'), ('mixed', args.precond)]
    if args.precond != 'jacobi':
        forms.insert(1, ('fp64', args.precond))
    for prec, pc in forms:
        solver._src = None
        solver.solve_newton(prec, rtol=args.rtol, lanes_per_row=8, log=log, precond=pc)
        torch.cuda.synchronize()
        t = time.perf_counter()
        solver._src = None
        x = solver.solve_newton(prec, rtol=args.rtol, lanes_per_row=8, precond=pc)
        torch.cuda.synchronize()
        out[(prec, pc)] = (time.perf_counter() - t, dict(solver.stats), x)
        log('%s/%s: %.3f s, cg %s' % (prec, pc, out[(prec, pc)][0],
                                      out[(prec, pc)][1]['cg_iterations']))
    axis = dp.beam_axis(5000)
    B = {p: probe_B(dp, solver, out[p][2], axis) for p in out}
    ref = B[('fp64', 'jacobi')]
    chosen = ('mixed', args.precond)
    err = float(np.linalg.norm(B[chosen] - ref, axis=1).max() / np.linalg.norm(ref, axis=1).max())
    print(json.dumps({
        'metric': 'solves/s, 6-yoke chain, %d tets, Newton to %g' % (dp.num_tets, args.rtol),
        'value': 1.0 / out[chosen][0], 'unit': 'solves/s', 'n_gpus': 1,
        'higher_is_better': True, 'dtype': 'f32+f64', 'data': 'synthetic',
        'config': {'workload': '6 alternating yokes along z', 'tets': dp.num_tets,
                   'edges': asm.num_dofs, 'nnz': asm.nnz, 'precond': args.precond},
        'seconds': {'%s/%s' % k: v[0] for k, v in out.items()},
        'cg_iterations': {'%s/%s' % k: v[1]['cg_iterations'] for k, v in out.items()},
        'B_axis_rel_err_vs_fp64_jacobi': err}), flush=True)


if __name__ == '__main__':
    main()
