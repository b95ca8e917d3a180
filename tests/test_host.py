"""CPU tests of host-side pieces: mesh-to-mesh transfer of edge DOFs."""
import numpy as np


def test_transfer_reproduces_a_uniform_field():
    """Whitney interpolation between two different meshes is exact for the
    fields the element space contains (uniform B: linear A)."""
    from snoopy_amd.magnets import Dipole
    from snoopy_amd.warm_start import transfer_operator
    a = Dipole(h=0.2)
    b = Dipole(h=0.13)
    B0 = np.array([0.2, 1.1, -0.4])
    A = lambda p: 0.5 * np.cross(B0[None, :], p)  # noqa: E731
    xa = a.mesh.edge_dofs_of_field(A)
    idx, w = transfer_operator(a.mesh, b.mesh)
    xb = np.where(idx >= 0, w * xa[np.maximum(idx, 0)], 0.0).sum(axis=1)
    np.testing.assert_allclose(xb, b.mesh.edge_dofs_of_field(A), atol=1e-12)


def test_anchor_is_the_median_current_and_falls_back_to_zero():
    import torch
    from snoopy_amd.batch import anchor_index, anchor_usable
    assert anchor_index([1.2e6, 1.8e6, 1.5e6]) == 2
    assert anchor_index([4.0, 1.0, 3.0, 2.0]) == 3          # lower median of 1, 2, 3, 4
    x = torch.ones(5, dtype=torch.float64)
    assert anchor_usable({'converged': True}, x)
    assert not anchor_usable({'converged': False}, x)
    x[2] = float('nan')
    assert not anchor_usable({'converged': True}, x)
    x[2] = float('inf')
    assert not anchor_usable({'converged': True}, x)


def test_anchor_rejects_variants_that_differ_beyond_the_current():
    import pytest
    from snoopy_amd.batch import ConcurrentSolver
    from snoopy_amd.magnets import Dipole
    a = Dipole(h=0.3, J=1.0e6)
    for other in (Dipole(h=0.3, J=2.0e6, B_sat=2.0), Dipole(h=0.3, J=2.0e6, mu_r=3000.0),
                  Dipole(h=0.3, J=2.0e6, coil_gap=0.03)):
        with pytest.raises(ValueError, match='anchor'):
            ConcurrentSolver([a, other], warm_start='anchor')
    with pytest.raises(ValueError, match='share the mesh'):
        ConcurrentSolver([a, Dipole(h=0.25, J=2.0e6)], warm_start='anchor')


def test_morph_family_contains_its_base_mesh_and_stays_valid():
    from snoopy_amd.magnets import Dipole
    from snoopy_amd.warm_start import MorphFamily, sample_params
    dp = Dipole(h=0.2)
    fam = MorphFamily(base=dp)
    np.testing.assert_array_equal(fam.nodes(fam.p0), dp.mesh.nodes)
    for seed in range(20):
        p = sample_params(seed)
        X = fam.nodes(p)[dp.mesh.cells]
        E = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=2)
        assert np.linalg.det(E).min() > 0
        # the yoke's faces are grid lines of the variant: iron elements fill
        # exactly the yoke box minus the aperture
        vol = np.abs(np.linalg.det(E)) / 6.0
        iron = vol[dp.mesh.tags == 1].sum()
        expect = (p['W'] * p['H'] - p['ax'] * p['ay']) * p['L']
        assert abs(iron - expect) < 1e-9 * expect


def test_energy_pod_is_blind_to_the_null_space():
    """POD in the inner product of a positive semi-definite K: adding null
    vectors of K (the solver's gauge) to the snapshots changes neither the
    coefficients nor the captured energy, and the modes are K-orthonormal."""
    from snoopy_amd.warm_start import POD
    rng = np.random.default_rng(3)
    n, m, S = 60, 10, 40
    B = rng.standard_normal((n - m, n))
    K = B.T @ B                                    # null space of dimension m
    N = np.linalg.svd(B)[2][n - m:].T              # basis of the null space
    Z = rng.standard_normal((S, 3)) @ rng.standard_normal((3, n))   # rank-3 family
    Z = Z + 0.01 * rng.standard_normal((S, n))
    junk = (rng.standard_normal((S, m)) @ N.T) * 5.0
    p1 = POD(Z, 8, Z @ K)
    p2 = POD(Z + junk, 8, (Z + junk) @ K)
    # the three family modes exactly; the noise modes to the float32 Gram's
    # rounding (the null-space part is 500x the noise in amplitude)
    np.testing.assert_allclose(p1.lam[:3], p2.lam[:3], rtol=1e-4)
    np.testing.assert_allclose(p1.lam, p2.lam, rtol=1e-2)
    np.testing.assert_allclose(np.abs(p1.coeffs[:, :3]), np.abs(p2.coeffs[:, :3]), rtol=1e-3,
                               atol=1e-4)
    G = p2.U.T.astype(np.float64) @ K @ p2.U.astype(np.float64)
    np.testing.assert_allclose(G, np.eye(G.shape[0]), atol=1e-3)
    np.testing.assert_allclose(p2.project((Z + junk)[5])[:3], p2.coeffs[5, :3], rtol=1e-3, atol=1e-4)


def test_planned_forcing_splits_the_remaining_digits_into_equal_steps():
    """A start at 1.6e-3 |R(0)| with rtol 1e-8 and at most 3 digits per
    float32 step needs 2 steps of 10^-2.6 each; starts that 3 digits or
    fewer separate from the target take one step; a start already at the
    target takes eta_max."""
    import math
    from snoopy_amd.red_mvp_solver import planned_forcing
    eta = planned_forcing(1.6e-3, 1e-8, 1e-3, 0.1)
    assert math.isclose(eta, math.sqrt(1e-8 / 1.6e-3), rel_tol=1e-12)
    assert math.isclose(planned_forcing(5e-6, 1e-8, 1e-3, 0.1), 2e-3, rel_tol=1e-12)
    assert math.isclose(planned_forcing(1e-5, 1e-8, 1e-3, 0.1), 1e-3, rel_tol=1e-9)
    # 6 digits exactly: two steps of 1e-3
    assert math.isclose(planned_forcing(1e-2, 1e-8, 1e-3, 0.1), 1e-3, rel_tol=1e-9)
    # 9 digits: three steps of 1e-3, never below eta_min
    assert math.isclose(planned_forcing(10.0, 1e-8, 1e-3, 0.1), 1e-3, rel_tol=1e-9)
    assert planned_forcing(5e-9, 1e-8, 1e-3, 0.1) == 0.1
    assert planned_forcing(2e-8, 1e-8, 1e-3, 0.1) == 0.1     # 0.5 capped at eta_max
