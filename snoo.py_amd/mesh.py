"""Tetrahedral meshes, edge DOF numbering and a gmsh-like model facade.

The reference takes mesh, DOF numbering and orientations from gmsh
(snoopy/mesh_tools.py:7-61,106-203,350-427; snoopy/curl_curl_assembler.py:19-137).
gmsh is not available, so this module provides:

* a structured mesher: a tensor grid (uniform spacing in the magnet, geometric
  grading in the air pad) whose hexahedra are split into 6 Kuhn tetrahedra,
  each stored positively oriented (the reference's kernel divides by the
  signed det J, fem_c/ctools.c:1319, so orientation matters);
* edge numbering: the unique (lo, hi) node pairs sorted lexicographically;
  DOF id = edge id; the global direction of an edge is lo -> hi;
* Dirichlet edges: every edge lying in the outer boundary of the grid;
* ``TetModel``: the subset of the gmsh model API that ``RedMVPSolver`` and
  ``CurlCurlAssembler`` read (nodes, elements per volume tag, edges and the
  "Dirichlet Boundary" physical group), with 1-based tags as gmsh has them.
"""
import itertools
import math

import numpy as np

from .finite_element import LOCAL_EDGES, orientations_for_cells

IRON_TAG = 1
AIR_TAG = 2


# ---------------------------------------------------------------------------
# axis construction
# ---------------------------------------------------------------------------

def _uniform(a, b, h):
    n = max(1, int(math.ceil((b - a) / h - 1e-9)))
    return np.linspace(a, b, n + 1)


def _graded(a, b, h0, ratio, small_at_a=True):
    """Cells growing geometrically from h0 at one end, rescaled to fit [a,b]."""
    L = b - a
    sizes = []
    s, tot = h0, 0.0
    while tot + 1e-12 < L:
        sizes.append(s)
        tot += s
        s *= ratio
    sizes = np.array(sizes) * (L / tot)
    if not small_at_a:
        sizes = sizes[::-1]
    return np.concatenate([[a], a + np.cumsum(sizes)])


def axis_lines(core_breaks, h, pad, ratio=1.2):
    """Grid lines: uniform spacing h between core_breaks[0] and core_breaks[-1]
    (each interval between breaks meshed on its own so the breaks are grid
    lines), and graded cells over ``pad`` on both sides."""
    parts = []
    lo, hi = core_breaks[0], core_breaks[-1]
    parts.append(_graded(lo - pad, lo, h, ratio, small_at_a=False)[:-1])
    for a, b in zip(core_breaks[:-1], core_breaks[1:]):
        parts.append(_uniform(a, b, h)[:-1])
    parts.append(_graded(hi, hi + pad, h, ratio, small_at_a=True))
    return np.concatenate(parts)


def refine_axis(lines):
    """Midpoint bisection of every cell (keeps the Kuhn hierarchy nested)."""
    mid = 0.5 * (lines[:-1] + lines[1:])
    out = np.empty(2 * len(lines) - 1)
    out[0::2] = lines
    out[1::2] = mid
    return out


# ---------------------------------------------------------------------------
# mesh container
# ---------------------------------------------------------------------------

class TetMesh:
    """Linear tetrahedral mesh with edge numbering.

    nodes    float64 [N_n,3]
    cells    int32   [N_el,4]   0-based node ids, positively oriented
    tags     int32   [N_el]     volume tag of each element
    edges    int32   [N_e,2]    (lo, hi) node ids, sorted lexicographically
    glob_ids int32   [N_el,6]   0-based edge id of local edge j
    orient   int32   [N_el]     6-bit orientation mask
    boundary_edges int32        0-based ids of Dirichlet edges
    """

    def __init__(self, nodes, cells, tags, boundary_node_mask=None,
                 boundary_faces=None, grid=None):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self.cells = np.ascontiguousarray(cells, dtype=np.int32)
        self.tags = np.ascontiguousarray(tags, dtype=np.int32)
        self.grid = grid
        self._build_edges()
        self.orient = orientations_for_cells(self.cells)
        if boundary_faces is not None:
            self.boundary_edges = self._edges_of_faces(boundary_faces)
        else:
            self.boundary_edges = np.zeros(0, dtype=np.int32)

    @property
    def num_nodes(self):
        return self.nodes.shape[0]

    @property
    def num_elements(self):
        return self.cells.shape[0]

    @property
    def num_edges(self):
        return self.edges.shape[0]

    def _build_edges(self):
        n = self.num_nodes
        a = self.cells[:, LOCAL_EDGES[:, 0]].astype(np.int64)
        b = self.cells[:, LOCAL_EDGES[:, 1]].astype(np.int64)
        key = np.minimum(a, b) * n + np.maximum(a, b)
        uniq, inv = np.unique(key.ravel(), return_inverse=True)
        self.edges = np.stack([uniq // n, uniq % n], axis=1).astype(np.int32)
        self.glob_ids = inv.reshape(-1, 6).astype(np.int32)

    def _edges_of_faces(self, faces):
        n = self.num_nodes
        f = np.asarray(faces, dtype=np.int64)
        pairs = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [0, 2]]])
        key = pairs.min(axis=1) * n + pairs.max(axis=1)
        ekey = self.edges[:, 0].astype(np.int64) * n + self.edges[:, 1]
        ids = np.searchsorted(ekey, np.unique(key))
        return ids.astype(np.int32)

    def element_volumes(self):
        X = self.nodes[self.cells]
        d = np.linalg.det(np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=2))
        return d / 6.0

    def discrete_gradient(self):
        """Edge-node incidence G (scipy CSR, N_e x N_n): (G phi)_e = phi_hi - phi_lo."""
        from scipy.sparse import csr_array
        ne = self.num_edges
        rows = np.repeat(np.arange(ne), 2)
        cols = self.edges.reshape(-1)
        vals = np.tile([-1.0, 1.0], ne)
        return csr_array((vals, (rows, cols)), shape=(ne, self.num_nodes))

    def edge_dofs_of_field(self, A_fn):
        """Edge DOFs (line integrals lo->hi) of a field that is linear along
        each edge: A(midpoint) . (x_hi - x_lo)."""
        p0 = self.nodes[self.edges[:, 0]]
        p1 = self.nodes[self.edges[:, 1]]
        return np.einsum('ij,ij->i', A_fn(0.5 * (p0 + p1)), p1 - p0)


# ---------------------------------------------------------------------------
# structured Kuhn mesher
# ---------------------------------------------------------------------------

# the 6 Kuhn tetrahedra of the unit cube: v0=(0,0,0), v1=e_p0, v2=e_p0+e_p1, v3=(1,1,1)
def _kuhn_local():
    tets = []
    for perm in itertools.permutations(range(3)):
        v = [np.zeros(3, dtype=int)]
        cur = np.zeros(3, dtype=int)
        for ax in perm:
            cur = cur.copy()
            cur[ax] = 1
            v.append(cur)
        M = np.stack([v[1] - v[0], v[2] - v[0], v[3] - v[0]], axis=1)
        if np.linalg.det(M) < 0:
            v[1], v[2] = v[2], v[1]
        tets.append([tuple(x) for x in v])
    return tets


_KUHN = _kuhn_local()


class StructuredGrid:
    """Tensor-product grid lines and index helpers."""

    def __init__(self, xs, ys, zs):
        self.xs, self.ys, self.zs = (np.asarray(v, dtype=np.float64) for v in (xs, ys, zs))
        self.nx, self.ny, self.nz = len(xs) - 1, len(ys) - 1, len(zs) - 1

    def node_id(self, i, j, k):
        return i + (self.nx + 1) * (j + (self.ny + 1) * k)


def kuhn_mesh(xs, ys, zs, tag_fn):
    """Split every hexahedron of the tensor grid into 6 positive tetrahedra.

    tag_fn(xc, yc, zc) -> int32 tags at the cell centres.  Elements are ordered
    hexahedron by hexahedron (i fastest), 6 per hexahedron.
    """
    g = StructuredGrid(xs, ys, zs)
    X, Y, Z = np.meshgrid(g.xs, g.ys, g.zs, indexing='ij')
    nodes = np.stack([X.transpose(2, 1, 0).ravel(), Y.transpose(2, 1, 0).ravel(),
                      Z.transpose(2, 1, 0).ravel()], axis=1)
    I, J, K, cells = _kuhn_cells(g)
    xc = 0.5 * (g.xs[I] + g.xs[I + 1])
    yc = 0.5 * (g.ys[J] + g.ys[J + 1])
    zc = 0.5 * (g.zs[K] + g.zs[K + 1])
    tags = np.repeat(np.asarray(tag_fn(xc, yc, zc), dtype=np.int32), 6)
    faces = _boundary_faces(g)
    return TetMesh(nodes, cells, tags, boundary_faces=faces, grid=g)


def _kuhn_cells(g):
    """Hexahedron indices (I, J, K) and the 6 Kuhn tetrahedra of each, as
    node ids of the grid; hexahedra ordered i fastest."""
    nx, ny, nz = g.nx, g.ny, g.nz
    I, J, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing='ij')
    I, J, K = (v.transpose(2, 1, 0).ravel() for v in (I, J, K))
    cells = np.empty((I.size, 6, 4), dtype=np.int64)
    for t, verts in enumerate(_KUHN):
        for a, (di, dj, dk) in enumerate(verts):
            cells[:, t, a] = g.node_id(I + di, J + dj, K + dk)
    return I, J, K, cells.reshape(-1, 4)


ALL_SIDES = ('x0', 'x1', 'y0', 'y1', 'z0', 'z1')


def mapped_kuhn_mesh(nodes, shape, hex_tags, sides=ALL_SIDES):
    """Kuhn split of a structured hexahedral grid whose nodes are placed
    freely: node (i, j, k) of the (nx+1) x (ny+1) x (nz+1) lattice is
    ``nodes[i + (nx+1) (j + (ny+1) k)]``.  Used for grids bent to follow
    tapered faces.  ``hex_tags`` [nx*ny*nz] (i fastest) are the volume tags;
    ``sides`` names the outer faces that carry the Dirichlet condition
    ('x0' is i = 0, 'x1' is i = nx, and so on).  The mapping must keep every
    tetrahedron positively oriented; a ValueError says otherwise.  The mesh
    has no tensor grid for point location (element bins are used)."""
    nx, ny, nz = shape
    g = StructuredGrid(np.arange(nx + 1.0), np.arange(ny + 1.0), np.arange(nz + 1.0))
    nodes = np.ascontiguousarray(nodes, dtype=np.float64).reshape(-1, 3)
    if nodes.shape[0] != (nx + 1) * (ny + 1) * (nz + 1):
        raise ValueError('mapped_kuhn_mesh: %d nodes for a %dx%dx%d grid'
                         % (nodes.shape[0], nx, ny, nz))
    _, _, _, cells = _kuhn_cells(g)
    X = nodes[cells]
    det = np.einsum('ij,ij->i', X[:, 1] - X[:, 0],
                    np.cross(X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]))
    if not np.all(det > 0):
        raise ValueError('mapped_kuhn_mesh: %d tetrahedra are not positively oriented '
                         '(the mapping folds the grid)' % int(np.sum(det <= 0)))
    tags = np.repeat(np.asarray(hex_tags, dtype=np.int32).ravel(), 6)
    return TetMesh(nodes, cells, tags, boundary_faces=_boundary_faces(g, sides), grid=None)


def _boundary_faces(g, sides=ALL_SIDES):
    """Triangles of the outer surface (each grid face quad split in two along
    the diagonal the Kuhn split uses; every edge of the face plane is
    covered either way), on the named sides only."""
    tris = []
    nx, ny, nz = g.nx, g.ny, g.nz

    def quads(fix_axis, fix_val):
        dims = [nx, ny, nz]
        ax = [a for a in range(3) if a != fix_axis]
        A, B = np.meshgrid(np.arange(dims[ax[0]]), np.arange(dims[ax[1]]), indexing='ij')
        A, B = A.ravel(), B.ravel()
        idx = [None, None, None]

        def nid(da, db):
            idx[fix_axis] = np.full(A.shape, fix_val)
            idx[ax[0]] = A + da
            idx[ax[1]] = B + db
            return g.node_id(idx[0], idx[1], idx[2])
        n00, n10, n01, n11 = nid(0, 0), nid(1, 0), nid(0, 1), nid(1, 1)
        tris.append(np.stack([n00, n10, n11], axis=1))
        tris.append(np.stack([n00, n01, n11], axis=1))

    for name, axis, val in (('x0', 0, 0), ('x1', 0, nx), ('y0', 1, 0), ('y1', 1, ny),
                            ('z0', 2, 0), ('z1', 2, nz)):
        if name in sides:
            quads(axis, val)
    return np.concatenate(tris) if tris else np.zeros((0, 3), dtype=np.int64)


def locate_points_binned(mesh, pts, bins_per_axis=None):
    """Element containing each point of any tetrahedral mesh (-1 outside):
    elements are hashed into a uniform grid of bins by their bounding boxes,
    and each point tests the candidates of its bin by barycentric
    coordinates, keeping the element whose smallest coordinate is largest
    (the containing one; ties on shared faces go to the first best)."""
    pts = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
    X = mesh.nodes[mesh.cells]                                    # [m,4,3]
    lo_e, hi_e = X.min(axis=1), X.max(axis=1)
    lo, hi = mesh.nodes.min(axis=0), mesh.nodes.max(axis=0)
    if bins_per_axis is None:
        bins_per_axis = max(1, int(round((mesh.num_elements / 4.0) ** (1.0 / 3.0))))
    nb = np.full(3, bins_per_axis)
    size = (hi - lo) / nb * (1 + 1e-12) + 1e-300
    b0 = np.clip(((lo_e - lo) / size).astype(np.int64), 0, nb - 1)
    b1 = np.clip(((hi_e - lo) / size).astype(np.int64), 0, nb - 1)
    span = b1 - b0 + 1
    cnt = span.prod(axis=1)
    el = np.repeat(np.arange(mesh.num_elements), cnt)
    off = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    sp = span[el]
    i = b0[el, 0] + off % sp[:, 0]
    j = b0[el, 1] + (off // sp[:, 0]) % sp[:, 1]
    k = b0[el, 2] + off // (sp[:, 0] * sp[:, 1])
    key = (k * nb[1] + j) * nb[0] + i
    order = np.argsort(key, kind='stable')
    key, el = key[order], el[order]
    starts = np.searchsorted(key, np.arange(nb.prod() + 1))
    out = np.full(len(pts), -1, dtype=np.int64)
    best = np.full(len(pts), -1e-12)
    inside = np.all((pts >= lo - 1e-12) & (pts <= hi + 1e-12), axis=1)
    pb = np.clip(((pts - lo) / size).astype(np.int64), 0, nb - 1)
    pkey = (pb[:, 2] * nb[1] + pb[:, 1]) * nb[0] + pb[:, 0]
    ncand = np.where(inside, starts[pkey + 1] - starts[pkey], 0)
    pid = np.repeat(np.arange(len(pts)), ncand)
    cidx = np.arange(ncand.sum()) - np.repeat(np.cumsum(ncand) - ncand, ncand)
    cand = el[starts[pkey[pid]] + cidx]
    Xc = X[cand]
    M = np.stack([Xc[:, 1] - Xc[:, 0], Xc[:, 2] - Xc[:, 0], Xc[:, 3] - Xc[:, 0]], axis=2)
    lam = np.linalg.solve(M, (pts[pid] - Xc[:, 0])[:, :, None])[:, :, 0]
    score = np.minimum(1.0 - lam.sum(axis=1), lam.min(axis=1))
    # best candidate of each point: first of its group after sorting by -score
    order = np.lexsort((-score, pid))
    first = np.ones(len(order), dtype=bool)
    first[1:] = pid[order][1:] != pid[order][:-1]
    q = order[first]
    keep = score[q] > best[pid[q]]
    out[pid[q][keep]] = cand[q][keep]
    return out


def locate_points(mesh, pts):
    """Element containing each point (-1 outside): by the structured grid of
    a Kuhn mesh, or by element bins for any other mesh."""
    if getattr(mesh, 'grid', None) is None:
        return locate_points_binned(mesh, pts)
    g = mesh.grid
    pts = np.asarray(pts, dtype=np.float64)
    out = np.full(pts.shape[0], -1, dtype=np.int64)
    i = np.clip(np.searchsorted(g.xs, pts[:, 0], side='right') - 1, 0, g.nx - 1)
    j = np.clip(np.searchsorted(g.ys, pts[:, 1], side='right') - 1, 0, g.ny - 1)
    k = np.clip(np.searchsorted(g.zs, pts[:, 2], side='right') - 1, 0, g.nz - 1)
    inside = ((pts[:, 0] >= g.xs[0]) & (pts[:, 0] <= g.xs[-1]) &
              (pts[:, 1] >= g.ys[0]) & (pts[:, 1] <= g.ys[-1]) &
              (pts[:, 2] >= g.zs[0]) & (pts[:, 2] <= g.zs[-1]))
    hexid = i + g.nx * (j + g.ny * k)
    best = np.full(pts.shape[0], -np.inf)
    for t in range(6):
        e = hexid * 6 + t
        X = mesh.nodes[mesh.cells[e]]                       # [M,4,3]
        M = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=2)
        lam = np.linalg.solve(M, (pts - X[:, 0])[:, :, None])[:, :, 0]
        bary = np.concatenate([1.0 - lam.sum(axis=1, keepdims=True), lam], axis=1)
        score = bary.min(axis=1)
        better = score > best
        out[better] = e[better]
        best[better] = score[better]
    out[~inside] = -1
    return out


# ---------------------------------------------------------------------------
# gmsh-like facade
# ---------------------------------------------------------------------------

class _MeshFacade:
    """The subset of ``gmsh.model.mesh`` read by the reference's assembler."""

    def __init__(self, tm):
        self._tm = tm

    def getNodes(self):
        tm = self._tm
        return (np.arange(1, tm.num_nodes + 1, dtype=np.uint64),
                tm.nodes.reshape(-1).copy(), np.zeros(0))

    def getNode(self, tag):
        return self._tm.nodes[int(tag) - 1].copy(), np.zeros(0), 3, -1

    def getElements(self, dim=3, tag=-1):
        tm = self._tm
        if dim != 3:
            return [], [], []
        sel = np.arange(tm.num_elements) if tag == -1 else np.nonzero(tm.tags == tag)[0]
        return ([4], [sel.astype(np.uint64) + 1],
                [(tm.cells[sel].astype(np.uint64) + 1).reshape(-1)])

    def createEdges(self):
        return None

    def createFaces(self):
        return None

    def getAllEdges(self):
        tm = self._tm
        return (np.arange(1, tm.num_edges + 1, dtype=np.uint64),
                (tm.edges.astype(np.uint64) + 1).reshape(-1))

    def getKeysForElement(self, elementTag, functionSpaceType='HcurlLegendre0', returnCoord=False):
        e = int(elementTag) - 1
        keys = self._tm.glob_ids[e].astype(np.int64) + 1
        return np.ones(6, dtype=np.int32), keys, np.zeros(0)

    def getBasisFunctionsOrientationForElement(self, elementTag, functionSpaceType=''):
        return int(self._tm.orient[int(elementTag) - 1])


class TetModel:
    """A mesh wrapped to look like the gmsh model the reference solver takes.

    Volume tags are the element tags of ``TetMesh.tags``; the physical surface
    group "Dirichlet Boundary" holds the outer boundary edges.
    """

    def __init__(self, tet_mesh, name='model'):
        self.tet_mesh = tet_mesh
        self.mesh = _MeshFacade(tet_mesh)
        self.name = name

    def getPhysicalGroups(self, dim=-1):
        return [(2, 1)] if dim in (-1, 2) else []

    def getEntitiesForPhysicalName(self, name):
        return [(2, 1)] if name == 'Dirichlet Boundary' else []

    def dirichlet_dofs(self):
        """1-based Dirichlet DOF ids, as get_edge_boundary_dofs returns them
        (snoopy/mesh_tools.py:350-427)."""
        return np.unique(self.tet_mesh.boundary_edges.astype(np.int64) + 1)
