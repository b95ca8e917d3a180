"""Field evaluation at points (snoopy/evaluators.py), on the device.

``evaluate_fem_solution_node_interpolation`` follows the reference
(snoopy/evaluators.py:413-453): the coil field is evaluated at the mesh nodes,
the element-constant curl A is averaged at the nodes separately for each
listed domain (compute_nodal_fields, snoopy/curl_curl_assembler.py:2089-2197)
and the coil field added, and each point takes the P1 interpolant of the
nodal field of the first listed domain that contains it
(evaluate_interpolation, snoopy/evaluators.py:297-411); points in no listed
domain get 0.  Point location uses the structured grid of the build's own
mesher instead of gmsh's element search.
"""
import numpy as np

from . import _native as nat
from .device import as_dev_f64, as_dev_i32, device
from .mesh import locate_points


def _node_lists(asm, domain_index):
    """(node_ptr, node_el) of one domain: the device-order elements of that
    domain at every mesh node, ascending (cached on the assembler)."""
    cache = asm.__dict__.setdefault('_node_lists', {})
    if domain_index not in cache:
        d = asm._device_state()
        b, e = int(d['offsets'][domain_index]), int(d['offsets'][domain_index + 1])
        cells = asm.tet_mesh.cells[d['order'][b:e]]
        nn = asm.tet_mesh.num_nodes
        nodes = cells.reshape(-1)
        els = np.repeat(np.arange(b, e, dtype=np.int64), 4)
        order = np.lexsort((els, nodes))
        ptr = np.zeros(nn + 1, dtype=np.int32)
        np.cumsum(np.bincount(nodes, minlength=nn), out=ptr[1:])
        cache[domain_index] = (as_dev_i32(ptr), as_dev_i32(els[order]))
    return cache[domain_index]


def nodal_fields_device(asm, x_dev, mat_ids):
    """Element-averaged curl A at the mesh nodes over the elements of the
    listed material tags (device [n_nodes, 3])."""
    import torch
    d = asm._device_state()
    _, Be = asm.compute_B_device(x_dev, quad_order=1)          # device order
    nn = asm.tet_mesh.num_nodes
    if len(mat_ids) == 1:
        ptr, els = _node_lists(asm, asm.volume_tags.index(mat_ids[0]))
    else:
        # union of the domains' lists, merged per node in ascending element order
        parts = [(_node_lists(asm, asm.volume_tags.index(m))) for m in mat_ids]
        ptr_h = [p.cpu().numpy() for p, _ in parts]
        el_h = [e.cpu().numpy() for _, e in parts]
        node_of = np.concatenate([np.repeat(np.arange(nn), np.diff(p)) for p in ptr_h])
        el_all = np.concatenate(el_h)
        order = np.lexsort((el_all, node_of))
        ptr = np.zeros(nn + 1, dtype=np.int32)
        np.cumsum(np.bincount(node_of, minlength=nn), out=ptr[1:])
        ptr, els = as_dev_i32(ptr), as_dev_i32(el_all[order])
    out = torch.empty((nn, 3), dtype=torch.float64, device=device())
    nat.check(nat.load().sn_nodal_average_f64(nn, nat.ptr(ptr), nat.ptr(els), nat.ptr(Be),
                                              nat.ptr(out), nat.stream_handle()))
    return out


def evaluate_fem_solution_node_interpolation(gmsh_model, points, solver, coil_list, x,
                                             domain_list):
    """B at ``points`` [M, 3] by nodal interpolation (host array [M, 3])."""
    import torch
    from .sources import source_field_device
    asm = solver.curl_curl_factory
    d = asm._device_state()
    tm = asm.tet_mesh
    x_dev = as_dev_f64(x)
    nodes_dev = d['nodes'].reshape(-1, 3)
    b_coil = source_field_device(coil_list, nodes_dev.contiguous())
    fields = torch.stack([nodal_fields_device(asm, x_dev, [m]) + b_coil for m in domain_list])
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
    el_mesh = locate_points(tm, pts)
    inv = np.empty(len(d['order']), dtype=np.int64)
    inv[d['order']] = np.arange(len(d['order']))
    ok = el_mesh >= 0
    el_dev = np.where(ok, inv[np.where(ok, el_mesh, 0)], -1)
    tags = np.where(ok, tm.tags[np.where(ok, el_mesh, 0)], -1)
    field = np.full(len(pts), -1, dtype=np.int64)
    for j, m in enumerate(domain_list):
        field[(field < 0) & (tags == m)] = j
    el_dev[field < 0] = -1
    out = torch.empty((len(pts), 3), dtype=torch.float64, device=device())
    # device copies held by name until the kernel is enqueued
    pts_d, el_d, field_d = as_dev_f64(pts), as_dev_i32(el_dev), as_dev_i32(field)
    fields = fields.contiguous()
    nat.check(nat.load().sn_p1_interp_f64(len(pts), nat.ptr(pts_d), nat.ptr(el_d),
                                          nat.ptr(field_d), tm.num_nodes, nat.ptr(d['nodes']),
                                          nat.ptr(d['cells']), nat.ptr(fields), nat.ptr(out),
                                          nat.stream_handle()))
    return out.cpu().numpy()
