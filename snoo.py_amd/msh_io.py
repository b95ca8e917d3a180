"""gmsh ``.msh`` files (ASCII formats 4.1 and 2.2) to and from the build's mesh.

The reference meshes with gmsh and reads the mesh through gmsh's API
(snoopy/mesh_tools.py, snoopy/curl_curl_assembler.py:56-129): volumes carry
physical tags (``iron_tags`` / ``air_tags`` of the solver) and the physical
surface group named "Dirichlet Boundary" marks the faces whose edges get
homogeneous Dirichlet conditions (snoopy/mesh_tools.py:350-427).  ``read_msh``
turns such a file into a ``TetModel``: linear tetrahedra (element type 4)
with the physical tag of their volume, and the triangles (type 2) of the
"Dirichlet Boundary" group -- or, when the file has no such group, the outer
boundary faces.  Node ids are renumbered 0..N-1 in ascending gmsh tag order
and negatively oriented tetrahedra are reoriented.

Edge DOFs are numbered by this build's rule (lexicographic node pairs) by
default.  ``edge_numbering='gmsh'`` numbers them as gmsh's ``createEdges``
tags minus one, the reference's DOF ids (snoopy/mesh_tools.py:32-36,201-203,
snoopy/curl_curl_assembler.py:125-129): gmsh numbers an edge when it first
meets it, walking the model's entities by dimension and then tag (curves,
surfaces, volumes), each entity's elements in file order, and each
element's edges in gmsh's local order -- line (0,1); triangle (0,1), (1,2),
(2,0); tetrahedron (0,1), (1,2), (2,0), (3,0), (3,2), (3,1) -- with the
element's nodes as the file lists them.  This restates gmsh's public
createEdges; gmsh is not in this image, so the numbering is not checked
against gmsh itself.
"""
import numpy as np

from .mesh import TetMesh, TetModel

DIRICHLET = 'Dirichlet Boundary'


def _sections(text):
    out = {}
    lines = text.splitlines()
    i = 0
    while i < len(lines):
        ln = lines[i].strip()
        if ln.startswith('$') and not ln.startswith('$End'):
            name = ln[1:]
            j = i + 1
            while j < len(lines) and lines[j].strip() != '$End' + name:
                j += 1
            out[name] = lines[i + 1:j]
            i = j + 1
        else:
            i += 1
    return out


def _ints(line):
    return [int(t) for t in line.split()]


def outer_faces(cells):
    """Faces that belong to exactly one tetrahedron (sorted node triples)."""
    f = np.concatenate([cells[:, [0, 1, 2]], cells[:, [0, 1, 3]], cells[:, [0, 2, 3]],
                        cells[:, [1, 2, 3]]])
    f = np.sort(f, axis=1)
    u, cnt = np.unique(f, axis=0, return_counts=True)
    return u[cnt == 1]


def _orient(nodes, cells):
    X = nodes[cells]
    det = np.einsum('ij,ij->i', X[:, 1] - X[:, 0],
                    np.cross(X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]))
    cells = cells.copy()
    neg = det < 0
    cells[neg, 2], cells[neg, 3] = cells[neg, 3].copy(), cells[neg, 2].copy()
    return cells


# gmsh's local edge order of lines, triangles and tetrahedra
_GMSH_EDGES = {1: [(0, 1)], 2: [(0, 1), (1, 2), (2, 0)],
               4: [(0, 1), (1, 2), (2, 0), (3, 0), (3, 2), (3, 1)]}
_GMSH_NODES = {1: 2, 2: 3, 4: 4}


def gmsh_edge_order(blocks, num_nodes):
    """Node-pair keys (lo * num_nodes + hi, 0-based node ids) in the order
    gmsh's createEdges numbers them: ``blocks`` = [((dim, entity tag),
    element type, node array [m, k]) ...] in any order; entities are walked
    by (dim, tag), elements in the given order."""
    seq = []
    for _, etype, conn in sorted(blocks, key=lambda b: b[0]):
        if etype not in _GMSH_EDGES or len(conn) == 0:
            continue
        c = np.asarray(conn, dtype=np.int64)
        pairs = np.stack([np.stack([c[:, a], c[:, b]], axis=1)
                          for a, b in _GMSH_EDGES[etype]], axis=1).reshape(-1, 2)
        seq.append(pairs.min(axis=1) * num_nodes + pairs.max(axis=1))
    keys = np.concatenate(seq) if seq else np.zeros(0, dtype=np.int64)
    uk, first = np.unique(keys, return_index=True)
    return uk[np.argsort(first, kind='stable')]


def renumber_edges(mesh, key_order):
    """Give the edges of ``mesh`` the ids of their node-pair keys' positions
    in ``key_order`` (in place): edges, glob_ids and boundary_edges."""
    n = mesh.num_nodes
    key_order = np.asarray(key_order, dtype=np.int64)
    ekey = mesh.edges[:, 0].astype(np.int64) * n + mesh.edges[:, 1]
    sorter = np.argsort(key_order, kind='stable')
    new_id = sorter[np.clip(np.searchsorted(key_order, ekey, sorter=sorter), 0,
                            max(len(key_order) - 1, 0))] if len(key_order) else ekey * 0
    if len(key_order) != mesh.num_edges or not np.array_equal(key_order[new_id], ekey):
        raise ValueError('edge numbering: the numbered edges are not the edges of the '
                         'tetrahedra (%d numbered, %d in the mesh)'
                         % (len(key_order), mesh.num_edges))
    edges = np.empty_like(mesh.edges)
    edges[new_id] = mesh.edges
    mesh.edges = edges
    mesh.glob_ids = new_id[mesh.glob_ids].astype(np.int32)
    mesh.boundary_edges = new_id[mesh.boundary_edges].astype(np.int32)
    return mesh


def read_msh(path, edge_numbering='sorted'):
    """TetModel of a gmsh ASCII mesh file (format 4.1 or 2.2); edge DOFs
    numbered 'sorted' (this build's rule) or 'gmsh' (createEdges order)."""
    if edge_numbering not in ('sorted', 'gmsh'):
        raise ValueError("edge_numbering must be 'sorted' or 'gmsh'")
    with open(path) as fh:
        sec = _sections(fh.read())
    if 'MeshFormat' not in sec:
        raise ValueError('%s: no $MeshFormat section' % path)
    fmt = sec['MeshFormat'][0].split()
    version, ftype = float(fmt[0]), int(fmt[1])
    if ftype != 0:
        raise ValueError('%s: binary .msh files are not supported' % path)
    names = {}
    for ln in sec.get('PhysicalNames', [])[1:]:
        t = ln.split(maxsplit=2)
        names[(int(t[0]), int(t[1]))] = t[2].strip().strip('"')
    dir_phys = {tag for (dim, tag), nm in names.items() if dim == 2 and nm == DIRICHLET}
    tets, tet_tag, tris = [], [], []
    blocks = []          # ((dim, entity), element type, nodes) for gmsh's edge order
    if version >= 4.0:
        ent_phys = {}
        if 'Entities' in sec:
            e = sec['Entities']
            npts, ncur, nsurf, nvol = _ints(e[0])
            k = 1 + npts + ncur
            for dim, count in ((2, nsurf), (3, nvol)):
                for ln in e[k:k + count]:
                    t = ln.split()
                    tag = int(t[0])
                    nphys = int(t[7])
                    ent_phys[(dim, tag)] = [int(v) for v in t[8:8 + nphys]]
                k += count
        nd = sec['Nodes']
        nblocks = _ints(nd[0])[0]
        k = 1
        ntags, coords = [], []
        for _ in range(nblocks):
            _, _, _, nin = _ints(nd[k])
            k += 1
            ntags += [int(nd[k + i]) for i in range(nin)]
            coords += [[float(v) for v in nd[k + nin + i].split()[:3]] for i in range(nin)]
            k += 2 * nin
        el = sec['Elements']
        nblocks = _ints(el[0])[0]
        k = 1
        for _ in range(nblocks):
            dim, etag, etype, nin = _ints(el[k])
            k += 1
            rows = [_ints(el[k + i]) for i in range(nin)]
            k += nin
            if etype in _GMSH_NODES and nin:
                blocks.append(((dim, etag), etype, [r[1:1 + _GMSH_NODES[etype]] for r in rows]))
            phys = ent_phys.get((dim, etag), [])
            if etype == 4:
                tets += [r[1:5] for r in rows]
                tet_tag += [phys[0] if phys else etag] * nin
            elif etype == 2 and (not dir_phys or set(phys) & dir_phys):
                tris += [r[1:4] for r in rows]
    else:
        nd = sec['Nodes']
        n = int(nd[0])
        ntags = [int(nd[1 + i].split()[0]) for i in range(n)]
        coords = [[float(v) for v in nd[1 + i].split()[1:4]] for i in range(n)]
        el = sec['Elements']
        for ln in el[1:1 + int(el[0])]:
            r = _ints(ln)
            etype, ntag = r[1], r[2]
            phys = r[3] if ntag > 0 else 0
            conn = r[3 + ntag:]
            if etype in _GMSH_NODES:
                dim = {1: 1, 2: 2, 4: 3}[etype]
                ent = r[4] if ntag > 1 else phys
                blocks.append(((dim, ent), etype, [conn[:_GMSH_NODES[etype]]]))
            if etype == 4:
                tets.append(conn[:4])
                tet_tag.append(phys)
            elif etype == 2 and (not dir_phys or phys in dir_phys):
                tris.append(conn[:3])
    order = np.argsort(ntags)
    tag_sorted = np.asarray(ntags, dtype=np.int64)[order]
    nodes = np.asarray(coords, dtype=np.float64)[order]
    remap = lambda a: np.searchsorted(tag_sorted, np.asarray(a, dtype=np.int64))  # noqa: E731
    cells = _orient(nodes, remap(tets).astype(np.int32))
    if dir_phys:
        faces = remap(tris) if tris else np.zeros((0, 3), dtype=np.int64)
    else:
        faces = outer_faces(cells)
    mesh = TetMesh(nodes, cells, np.asarray(tet_tag, dtype=np.int32), boundary_faces=faces)
    if edge_numbering == 'gmsh':
        # the element nodes as the file lists them (before any reorientation)
        merged = {}
        for ent, etype, conn in blocks:
            merged.setdefault((ent, etype), []).extend(conn)
        renumber_edges(mesh, gmsh_edge_order(
            [(ent, etype, remap(conn)) for (ent, etype), conn in merged.items()], len(nodes)))
    return TetModel(mesh)


def write_msh(path, tet_mesh, boundary_faces=None):
    """Write a TetMesh as a gmsh 4.1 ASCII file: one volume entity per tag
    (physical group = tag), one surface entity with the physical group
    "Dirichlet Boundary" (the given faces, or the outer boundary faces)."""
    tm = tet_mesh
    faces = outer_faces(tm.cells) if boundary_faces is None else np.asarray(boundary_faces)
    vtags = sorted(int(t) for t in np.unique(tm.tags))
    dtag = max(vtags) + 1
    L = ['$MeshFormat', '4.1 0 8', '$EndMeshFormat', '$PhysicalNames', str(len(vtags) + 1)]
    L.append('2 %d "%s"' % (dtag, DIRICHLET))
    L += ['3 %d "volume %d"' % (t, t) for t in vtags]
    L.append('$EndPhysicalNames')
    lo, hi = tm.nodes.min(axis=0), tm.nodes.max(axis=0)
    box = ' '.join('%.17g' % v for v in list(lo) + list(hi))
    L += ['$Entities', '0 0 1 %d' % len(vtags), '1 %s 1 %d 0' % (box, dtag)]
    L += ['%d %s 1 %d 0' % (t, box, t) for t in vtags]
    L.append('$EndEntities')
    nn = tm.num_nodes
    L += ['$Nodes', '1 %d 1 %d' % (nn, nn), '3 %d 0 %d' % (vtags[0], nn)]
    L += [str(i + 1) for i in range(nn)]
    L += ['%.17g %.17g %.17g' % tuple(p) for p in tm.nodes]
    L.append('$EndNodes')
    blocks = [(2, 1, 2, faces)] + [(3, t, 4, tm.cells[tm.tags == t]) for t in vtags]
    ntot = sum(len(b[3]) for b in blocks)
    L += ['$Elements', '%d %d 1 %d' % (len(blocks), ntot, ntot)]
    eid = 1
    for dim, etag, etype, conn in blocks:
        L.append('%d %d %d %d' % (dim, etag, etype, len(conn)))
        for c in conn:
            L.append(' '.join([str(eid)] + [str(int(v) + 1) for v in c]))
            eid += 1
    L.append('$EndElements')
    with open(path, 'w') as fh:
        fh.write('\n'.join(L) + '\n')
