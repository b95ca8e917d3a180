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
