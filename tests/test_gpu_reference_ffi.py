"""GPU parity of the fem_c_mod drop-in (snoopy_amd.fem_c_mod) called as the
reference calls its FFI (snoopy/curl_curl_assembler.py:915-927): 0-based
cells and DOF ids, a per-element orientation number and a curl table
[n_orient, Q, 6, 3] with n_orient = 24 (one orientation per ordering of the
element's four global vertices, as a mesher enumerates them), and tables that
vary per quadrature point.  The oracle (oracle/snoo_oracle.c, bit-identical
to the compiled fem_c/ctools.c) gets the same arrays.  Tolerance: 1e-12
relative to the largest entry (float64, only the summation order differs);
the CSR pattern bit-exact."""
import ctypes
import itertools
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

PERMS = list(itertools.permutations(range(4)))     # 24 vertex orderings


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def _gmsh_like(cells, Q):
    """Orientation number of each element (index of the ranking of its global
    vertex ids among the 24 orderings) and the 24-entry curl table: edge
    (a, b) of ordering r points from its lower- to its higher-ranked vertex."""
    from snoopy_amd.finite_element import LOCAL_EDGES, reference_curls
    rank = np.argsort(np.argsort(cells, axis=1), axis=1)
    index = {p: k for k, p in enumerate(PERMS)}
    orient = np.array([index[tuple(r)] for r in rank], dtype=np.int32)
    c = reference_curls()
    tab = np.empty((len(PERMS), Q, 6, 3))
    for k, r in enumerate(PERMS):
        s = np.array([1.0 if r[a] < r[b] else -1.0 for a, b in LOCAL_EDGES])
        tab[k] = (s[:, None] * c)[None]
    return orient, tab


@pytest.fixture(scope='module')
def case():
    from snoopy_amd.magnets import Dipole
    from snoopy_amd import finite_element as fe
    print('case: reference-shaped FFI, dipole h=0.15', flush=True)
    dp = Dipole(h=0.15)
    m = dp.mesh
    w, q = fe.quadrature_rule(2)
    Q = len(w)
    rng = np.random.default_rng(21)
    x = rng.standard_normal(m.num_edges) * 2e-3
    iron = np.nonzero(m.tags == 1)[0]
    air = np.nonzero(m.tags != 1)[0]
    Bs = rng.uniform(-1.2, 1.2, size=(len(iron) * Q, 3))
    return dp, m, w, q, Q, x, iron, air, Bs


def _bh_rel():
    from snoopy_amd.materials import Reluctance
    d = json.load(open(os.path.join(GOLDEN, 'aisi1010_bh.json')))
    return Reluctance(B_data=d['B(T)'], H_data=d['H(A/m)'])


def _oracle_domain(m, sel, curls, d_phi, w, model, x, Bs, orient):
    r, c, vK, vdK, rhs = orc.element_triplets(m.nodes, m.cells[sel], m.glob_ids[sel], orient,
                                              curls, d_phi, w, model, x, Bs)
    n = m.num_edges
    return orc.coo_to_csr(n, r, c, vK), orc.coo_to_csr(n, r, c, vdK), \
        (rhs if rhs is not None else np.zeros(n))


def _compare(K, dK, rhs, Ko, dKo, rhso):
    for a, b in ((K, Ko), (dK, dKo)):
        np.testing.assert_array_equal(a.indptr, b.indptr)
        np.testing.assert_array_equal(a.indices, b.indices)
        if np.abs(b.data).max() > 0:
            assert _rel(a.data, b.data) < 1e-12
        else:                                   # dK of a constant reluctivity
            assert not np.any(a.data)
    if np.abs(rhso).max() > 0:
        assert _rel(rhs, rhso) < 1e-12
    else:
        assert not np.any(rhs)


@pytest.mark.parametrize('domain', ['iron', 'air'])
def test_stiffness_ffi_with_24_orientations_matches_oracle(case, domain):
    from snoopy_amd import fem_c_mod
    from snoopy_amd import finite_element as fe
    from snoopy_amd.materials import NU0, ConstantReluctance, model_tables
    dp, m, w, q, Q, x, iron, air, Bs = case
    sel = iron if domain == 'iron' else air
    rel = _bh_rel() if domain == 'iron' else ConstantReluctance(NU0)
    B_s = Bs if domain == 'iron' else np.zeros((0, 3))
    orient, tab = _gmsh_like(m.cells[sel], Q)
    d_phi = fe.lagrange_basis_derivative(q)
    K, dK, rhs = fem_c_mod.compute_stiffness_and_jacobi_matrix_curl_curl(
        m.nodes, m.cells[sel].reshape(-1), m.glob_ids[sel], tab, d_phi, orient, x, B_s, rel, q, w,
        m.num_edges, 6)
    Ko, dKo, rhso = _oracle_domain(m, sel, tab.reshape(-1), d_phi, w, model_tables(rel), x,
                                   Bs if domain == 'iron' else None, orient)
    _compare(K, dK, rhs, Ko, dKo, rhso)
    # the 24-orientation description is the mesh's own orientation: the same
    # matrices as the 64-entry mask table of the product assembler
    mask = fe.orientations_for_cells(m.cells[sel])
    Km, dKm, rhsm = _oracle_domain(m, sel, fe.curl_table(Q), d_phi, w, model_tables(rel), x,
                                   Bs if domain == 'iron' else None, mask)
    _compare(K, dK, rhs, Km, dKm, rhsm)


def test_stiffness_ffi_reads_the_tables_per_quadrature_point(case):
    """Tables that differ at every quadrature point (a non-affine element map
    and arbitrary curls, 5 orientations): the FFI evaluates them as given."""
    from snoopy_amd import fem_c_mod
    from snoopy_amd import finite_element as fe
    from snoopy_amd.materials import model_tables, tanh_reluctance
    dp, m, w, q, Q, x, iron, air, Bs = case
    sel = iron[:3000]
    rng = np.random.default_rng(5)
    n_or = 5
    tab = rng.uniform(-1.0, 1.0, size=(n_or, Q, 6, 3))
    orient = rng.integers(0, n_or, size=len(sel)).astype(np.int32)
    d_phi = fe.lagrange_basis_derivative(q) + rng.uniform(-0.05, 0.05, size=(Q, 4, 3))
    B_s = Bs[:len(sel) * Q]
    for rel in (_bh_rel(), tanh_reluctance(1.9, 4000.0)):
        K, dK, rhs = fem_c_mod.compute_stiffness_and_jacobi_matrix_curl_curl(
            m.nodes, m.cells[sel].reshape(-1), m.glob_ids[sel], tab, d_phi, orient, x, B_s, rel,
            q, w, m.num_edges, 6)
        Ko, dKo, rhso = _oracle_domain(m, sel, tab.reshape(-1), d_phi.reshape(-1), w,
                                       model_tables(rel), x, B_s, orient)
        _compare(K, dK, rhs, Ko, dKo, rhso)


def test_B_hcurl_ffi_with_24_orientations_matches_oracle(case):
    from snoopy_amd import fem_c_mod
    from snoopy_amd import finite_element as fe
    dp, m, w, q, Q, x, iron, air, Bs = case
    orient, tab = _gmsh_like(m.cells, Q)
    phi = fe.lagrange_basis(q)
    d_phi = fe.lagrange_basis_derivative(q)
    pts, B = fem_c_mod.compute_B_Hcurl(m.nodes, m.cells.reshape(-1), m.glob_ids, tab, phi, d_phi,
                                       orient, x, q, w, m.num_edges, 6)
    po, Bo = orc.B_hcurl(m.nodes, m.cells, m.glob_ids, orient, tab.reshape(-1), phi, d_phi, x)
    assert _rel(pts, po) < 1e-14
    assert _rel(B, Bo) < 1e-12


def test_biot_savart_ffi_matches_oracle(case):
    from snoopy_amd import fem_c_mod
    dp = case[0]
    c = dp.coils[0]
    rng = np.random.default_rng(2)
    tar = rng.uniform(-0.5, 0.5, size=(2000, 3))
    B = fem_c_mod.compute_B_line_segs_cpp(c.get_segments(), tar, c.current, c.rad)
    assert _rel(B, orc.biot_savart(c.get_segments(), tar, c.current, c.rad)) < 1e-12


def test_table_kernel_rejects_an_orientation_beyond_the_table(case):
    """The device-side check of the C entry point itself: an orientation
    number >= n_orient is reported, never read."""
    import torch
    from snoopy_amd import _native as nat
    from snoopy_amd import finite_element as fe
    from snoopy_amd.device import as_dev_f64, as_dev_i32
    dp, m, w, q, Q, x, iron, air, Bs = case
    sel = air[:100]
    orient, tab = _gmsh_like(m.cells[sel], Q)
    orient = orient.copy()
    orient[17] = 24
    lib = nat.load()
    out = [torch.empty(21 * len(sel), dtype=torch.float64, device='cuda') for _ in range(2)]
    loc6 = torch.empty(6 * len(sel), dtype=torch.float64, device='cuda')
    h = [as_dev_f64(m.nodes.reshape(-1)), as_dev_i32(m.cells[sel]), as_dev_i32(m.glob_ids[sel]),
         as_dev_i32(orient), as_dev_f64(tab.reshape(-1)),
         as_dev_f64(fe.lagrange_basis_derivative(q).reshape(-1)), as_dev_f64(w),
         as_dev_f64(np.array([1.0])), as_dev_f64(x)]
    rc = lib.sn_element_table_f64(
        len(sel), nat.ptr(h[0]), nat.ptr(h[1]), nat.ptr(h[2]), nat.ptr(h[3]), 24, nat.ptr(h[4]),
        nat.ptr(h[5]), nat.ptr(h[6]), Q, 0, -1, nat.ptr(h[7]), None, nat.ptr(h[8]), None,
        nat.ptr(out[0]), nat.ptr(out[1]), nat.ptr(loc6), nat.stream_handle())
    assert rc == nat.SN_ERR_ARG
    assert b'orientation' in lib.sn_last_error()
