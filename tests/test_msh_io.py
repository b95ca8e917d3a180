"""CPU: gmsh .msh input and output (formats 4.1 and 2.2)."""
import numpy as np


def test_msh41_round_trip(tmp_path):
    import importlib
    mg = importlib.import_module('snoopy_amd.magnets')
    mio = importlib.import_module('snoopy_amd.msh_io')
    d = mg.Dipole(h=0.15)
    tm = d.mesh
    p = tmp_path / 'dipole.msh'
    mio.write_msh(str(p), tm)
    m2 = mio.read_msh(str(p)).tet_mesh
    np.testing.assert_array_equal(m2.nodes, tm.nodes)
    # the file groups elements by volume; the set of (cell, tag) is kept
    key = lambda m: sorted(zip(map(tuple, m.cells.tolist()), m.tags.tolist()))  # noqa: E731
    assert key(m2) == key(tm)
    np.testing.assert_array_equal(m2.edges, tm.edges)
    np.testing.assert_array_equal(np.sort(m2.boundary_edges), np.sort(tm.boundary_edges))


MSH22 = """$MeshFormat
2.2 0 8
$EndMeshFormat
$PhysicalNames
3
2 7 "Dirichlet Boundary"
3 1 "iron"
3 2 "air"
$EndPhysicalNames
$Nodes
5
10 0 0 0
11 1 0 0
12 0 1 0
13 0 0 1
20 1 1 1
$EndNodes
$Elements
4
1 4 2 1 1 10 11 12 13
2 4 2 2 2 11 12 20 13
3 2 2 7 5 10 11 12
4 2 2 9 6 11 12 20
$EndElements
"""


def test_msh22_tags_orientation_and_dirichlet_group(tmp_path):
    import importlib
    mio = importlib.import_module('snoopy_amd.msh_io')
    p = tmp_path / 'two.msh'
    p.write_text(MSH22)
    tm = mio.read_msh(str(p)).tet_mesh
    assert tm.num_nodes == 5 and tm.num_elements == 2
    np.testing.assert_array_equal(tm.tags, [1, 2])
    X = tm.nodes[tm.cells]
    det = np.einsum('ij,ij->i', X[:, 1] - X[:, 0], np.cross(X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]))
    assert np.all(det > 0)
    # only the triangle of the "Dirichlet Boundary" group (nodes 10, 11, 12)
    be = {tuple(e) for e in tm.edges[tm.boundary_edges]}
    assert be == {(0, 1), (0, 2), (1, 2)}


def test_binned_point_location_matches_structured():
    import importlib
    mg = importlib.import_module('snoopy_amd.magnets')
    mesh = importlib.import_module('snoopy_amd.mesh')
    d = mg.Dipole(h=0.12)
    tm = d.mesh
    rng = np.random.default_rng(9)
    lo, hi = tm.nodes.min(axis=0), tm.nodes.max(axis=0)
    pts = rng.uniform(lo - 0.05, hi + 0.05, (3000, 3))
    a = mesh.locate_points(tm, pts)
    b = mesh.locate_points_binned(tm, pts)
    np.testing.assert_array_equal(a < 0, b < 0)
    # points strictly inside an element are found in that element
    ok = a >= 0
    np.testing.assert_array_equal(a[ok], b[ok])


def test_gmsh_edge_numbering_of_the_two_tetrahedra_file(tmp_path):
    """createEdges order worked out by hand for MSH22 (0-based node ids 0..4
    for tags 10, 11, 12, 13, 20): entities by (dim, tag) -- triangle entity
    5 (0 1 2), triangle entity 6 (1 2 4), tetrahedron entity 1 (0 1 2 3),
    tetrahedron entity 2 (1 2 4 3, the file's order although it is
    reoriented) -- each element's edges in gmsh's local order, an edge
    numbered when first met."""
    import importlib
    mio = importlib.import_module('snoopy_amd.msh_io')
    p = tmp_path / 'two.msh'
    p.write_text(MSH22)
    g = mio.read_msh(str(p), edge_numbering='gmsh').tet_mesh
    s = mio.read_msh(str(p)).tet_mesh
    want = [(0, 1), (1, 2), (0, 2), (2, 4), (1, 4), (0, 3), (2, 3), (1, 3), (3, 4)]
    assert [tuple(e) for e in g.edges.tolist()] == want
    # the same elements, orientations and boundary edges, under the new ids
    np.testing.assert_array_equal(g.cells, s.cells)
    np.testing.assert_array_equal(g.orient, s.orient)
    np.testing.assert_array_equal(g.edges[g.glob_ids], s.edges[s.glob_ids])
    assert {tuple(e) for e in g.edges[g.boundary_edges]} == {(0, 1), (0, 2), (1, 2)}


def test_gmsh_edge_numbering_is_a_relabelling(tmp_path):
    """On a dipole mesh written to 4.1 and read back, the gmsh order is a
    permutation of the sorted edge ids: every element's local edges, the
    Dirichlet edges and the orientations describe the same node pairs."""
    import importlib
    mg = importlib.import_module('snoopy_amd.magnets')
    mio = importlib.import_module('snoopy_amd.msh_io')
    d = mg.Dipole(h=0.15)
    p = tmp_path / 'dipole.msh'
    mio.write_msh(str(p), d.mesh)
    s = mio.read_msh(str(p)).tet_mesh
    g = mio.read_msh(str(p), edge_numbering='gmsh').tet_mesh
    assert g.num_edges == s.num_edges
    assert not np.array_equal(g.edges, s.edges)
    np.testing.assert_array_equal(np.sort(g.edges, axis=0), np.sort(s.edges, axis=0))
    np.testing.assert_array_equal(g.edges[g.glob_ids], s.edges[s.glob_ids])
    np.testing.assert_array_equal(g.orient, s.orient)
    key = lambda m: {tuple(e) for e in m.edges[m.boundary_edges].tolist()}   # noqa: E731
    assert key(g) == key(s)
    # the first edges are those of the Dirichlet triangles, the file's first block
    first = {tuple(e) for e in g.edges[:3].tolist()}
    assert first <= key(g)


def test_edge_numbering_option_is_checked(tmp_path):
    import importlib
    import pytest
    mio = importlib.import_module('snoopy_amd.msh_io')
    p = tmp_path / 'two.msh'
    p.write_text(MSH22)
    with pytest.raises(ValueError):
        mio.read_msh(str(p), edge_numbering='lexicographic')
