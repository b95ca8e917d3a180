"""GPU: the learned warm start end to end at a small size -- snapshots from
the solver, POD, BF16 MLP training, prediction -- and what it gains over
Newton from A = 0 on held-out variants."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def trained():
    from snoopy_amd.warm_start import MorphFamily, WarmStart
    fam = MorphFamily(counts=(4, 3, 4, 3, 3, 5))
    print('case: warm start family tets', fam.mesh.num_elements, flush=True)
    ws = WarmStart.train(fam, n_train=96, r=24, epochs=800, rtol=1e-7, hidden=(128, 128, 128, 128))
    return fam, ws


def test_training_fits_coefficients(trained):
    fam, ws = trained
    st = ws.train_stats
    assert st['loss_last'] < 0.2 * st['loss_first'], st
    assert st['pod_energy'] > 0.99


def test_warm_started_newton_on_held_out_variants(trained):
    """Held-out variants: the residual judge keeps the prediction only when
    its residual is below the residual at A = 0, and Newton from the kept
    start converges to the same tolerance and the same field as Newton from
    A = 0.  The CG work of both is printed (the gain is reported, not
    asserted: on this 8-parameter family the prediction lands at a few
    percent of the initial residual at best)."""
    import torch
    from snoopy_amd.warm_start import sample_params
    fam, ws = trained
    fs = ws.fsolver
    for seed in (5000, 5001, 5002, 5003):
        p = sample_params(seed)
        fs.set_variant(p)
        x0 = ws.predict_family(p)
        assert torch.isfinite(x0).all()
        drv = fs.solver.newton_driver('mixed')
        drv.retarget(fs.solver.source_fields_device())
        r_pred = drv.residual(x0) / drv.r0()
        xa = fs.solver.solve_newton('mixed', rtol=1e-8)
        s0 = dict(fs.solver.stats)
        use = r_pred < 1.0
        xb = fs.solver.solve_newton('mixed', rtol=1e-8, x0=x0 if use else None)
        s1 = dict(fs.solver.stats)
        assert s0['converged'] and s1['converged']
        assert s1['residual_history'][-1] <= 1e-8
        asm = fs.solver.curl_curl_factory
        Ba = asm.element_B(xa.cpu().numpy())
        Bb = asm.element_B(xb.cpu().numpy())
        err = np.abs(Ba - Bb).max() / np.abs(Ba).max()
        print('seed', seed, 'pred residual %.3e' % r_pred, 'kept', bool(use), 'cg',
              sum(s0['cg_iterations']), '->', sum(s1['cg_iterations']), 'B diff %.1e' % err,
              flush=True)
        assert err < 1e-5
