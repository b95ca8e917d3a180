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


def test_prediction_beats_zero_start_on_held_out_variants(trained):
    from snoopy_amd.warm_start import sample_params
    fam, ws = trained
    fs = ws.fsolver
    its0, its1, better = [], [], 0
    for seed in (5000, 5001, 5002, 5003):
        p = sample_params(seed)
        fs.set_variant(p)
        x0 = ws.predict_family(p)
        drv = fs.solver.newton_driver('mixed')
        drv.retarget(fs.solver.source_fields_device())
        r_pred = drv.residual(x0) / drv.r0()
        fs.solver.solve_newton('mixed', rtol=1e-8)
        s0 = dict(fs.solver.stats)
        fs.solver.solve_newton('mixed', rtol=1e-8, x0=x0)
        s1 = dict(fs.solver.stats)
        assert s0['converged'] and s1['converged']
        its0.append(sum(s0['cg_iterations']))
        its1.append(sum(s1['cg_iterations']))
        better += r_pred < 1.0
        print('seed', seed, 'pred residual', r_pred, 'cg', its0[-1], '->', its1[-1], flush=True)
    assert better == 4
    assert sum(its1) < sum(its0)
