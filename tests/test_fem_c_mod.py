"""CPU tests of the fem_c_mod drop-in's host checks: arguments the
reference's FFI would read out of bounds (or print about and continue,
fem_c/interface_c.pyx:233-234,951-959) raise ValueError before any device
work."""
import numpy as np
import pytest


def _args():
    from snoopy_amd import finite_element as fe
    from snoopy_amd.magnets import Dipole
    from snoopy_amd.materials import NU0, ConstantReluctance
    m = Dipole(h=0.3).mesh
    w, q = fe.quadrature_rule(2)
    Q = len(w)
    tab = np.zeros((24, Q, 6, 3))
    orient = np.zeros(m.num_elements, dtype=np.int32)
    return dict(nodes=m.nodes, cells=m.cells.reshape(-1), glob_ids=m.glob_ids, curls=tab,
                d_phi=fe.lagrange_basis_derivative(q), orientations=orient,
                x=np.zeros(m.num_edges), B_s=np.zeros((0, 3)),
                reluctance=ConstantReluctance(NU0), q=q, w=w, num_dofs=m.num_edges,
                num_el_dofs=6)


@pytest.mark.parametrize('change,match', [
    (lambda a: a['orientations'].__setitem__(3, 24), r'orientation numbers must lie in \[0, 24\)'),
    (lambda a: a['orientations'].__setitem__(0, -1), 'orientation numbers'),
    (lambda a: a.__setitem__('curls', np.zeros(10)), 'not a multiple'),
    (lambda a: a.__setitem__('glob_ids', a['glob_ids'][:-1]), 'glob_ids has'),
    (lambda a: a.__setitem__('B_s', np.zeros((7, 3))), 'B_s has'),
    (lambda a: a.__setitem__('num_el_dofs', 12), 'lowest-order'),
    (lambda a: a.__setitem__('x', np.zeros(5)), 'x has'),
    (lambda a: a.__setitem__('d_phi', np.zeros((4, 3))), 'd_phi must be'),
])
def test_stiffness_ffi_rejects_bad_arguments(change, match):
    from snoopy_amd import fem_c_mod
    a = _args()
    a['orientations'] = a['orientations'].copy()
    change(a)
    with pytest.raises(ValueError, match=match):
        fem_c_mod.compute_stiffness_and_jacobi_matrix_curl_curl(**a)


def test_B_hcurl_ffi_rejects_bad_phi():
    from snoopy_amd import fem_c_mod
    a = _args()
    with pytest.raises(ValueError, match='phi must be'):
        fem_c_mod.compute_B_Hcurl(a['nodes'], a['cells'], a['glob_ids'], a['curls'],
                                  np.zeros((3, 4)), a['d_phi'], a['orientations'], a['x'], a['q'],
                                  a['w'], a['num_dofs'], 6)


def test_biot_savart_ffi_rejects_bad_shapes():
    from snoopy_amd import fem_c_mod
    with pytest.raises(ValueError, match='N x 6'):
        fem_c_mod.compute_B_line_segs_cpp(np.zeros((3, 5)), np.zeros((2, 3)), 1.0, 0.01)
    with pytest.raises(ValueError, match='M x 3'):
        fem_c_mod.compute_B_line_segs_cpp(np.zeros((3, 6)), np.zeros((2, 2)), 1.0, 0.01)
