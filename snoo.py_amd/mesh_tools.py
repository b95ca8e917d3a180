"""Mesh helpers of the reference's mesh_tools that the solver uses.

``get_edge_boundary_dofs`` (snoopy/mesh_tools.py:350-427) and
``get_cotree_dofs`` (snoopy/mesh_tools.py:468-566) restated for ``TetModel``.
"""
import numpy as np


def get_edge_boundary_dofs(model, boundary_names, element_order=1):
    """1-based DOF ids of the edges on the named boundaries (unique, sorted)."""
    ids = []
    for name in boundary_names:
        if model.getEntitiesForPhysicalName(name):
            ids.append(model.dirichlet_dofs())
    if not ids:
        return np.zeros(0, dtype=np.int64)
    return np.unique(np.concatenate(ids))


def get_cotree_dofs(tet_mesh):
    """(tree, cotree) 0-based edge ids of a spanning tree of the edge graph.

    The reference uses networkx' minimum spanning tree with unit weights
    (mesh_tools.py:498-516); any spanning tree gauges the gradient kernel, and
    scipy's csgraph tree is used here.
    """
    from scipy.sparse import csr_array
    from scipy.sparse.csgraph import connected_components, minimum_spanning_tree
    e = tet_mesh.edges.astype(np.int64)
    nn = tet_mesh.num_nodes
    ids = np.arange(len(e), dtype=np.float64) + 1.0   # weight = edge id + 1 (unique, > 0)
    G = csr_array((ids, (e[:, 0], e[:, 1])), shape=(nn, nn))
    ncomp, _ = connected_components(G, directed=False)
    if ncomp != 1:
        raise ValueError('The input graph must be connected.')
    T = minimum_spanning_tree(G).tocoo()
    tree = np.sort(T.data.astype(np.int64) - 1)
    cotree = np.setdiff1d(np.arange(len(e), dtype=np.int64), tree)
    return tree, cotree
