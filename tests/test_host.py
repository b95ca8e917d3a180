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
