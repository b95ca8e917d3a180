"""Drop-in for the reference's compiled extension ``snoopy.fem_c_mod``.

The reference builds ``fem_c_mod`` from Cython (fem_c/interface_c.pyx, built
by build.py:25-45) and calls it with host numpy arrays.  This module keeps
the three entry points of the nonlinear solve's path with the same names,
argument order, argument meaning and return values, and runs them on the GPU:

* ``compute_stiffness_and_jacobi_matrix_curl_curl`` (interface_c.pyx:162-374),
  called per domain by snoopy/curl_curl_assembler.py:915-927;
* ``compute_B_Hcurl`` (interface_c.pyx:555-696);
* ``compute_B_line_segs_cpp`` (interface_c.pyx:928-991).

The arguments are the reference's own: 0-based ``cells`` and ``glob_ids``,
the per-element ``orientations`` number, and the curl table of every
orientation its mesher enumerates, ``curls`` [n_orient, Q, 6, 3] (n_orient is
read from the table's size).  The element kernels evaluate those tables as
given (sn_element_table_f64, sn_B_Hcurl_table_f64).  Where the reference
would read out of bounds or print and continue, these functions raise
``ValueError``: an orientation outside [0, n_orient), a table of the wrong
size, a source field of the wrong length, an unknown reluctivity type.

A maintainer switches the reference to the GPU with one import in
snoopy/curl_curl_assembler.py and snoopy/racetrack_coil.py::

    from snoopy_amd import fem_c_mod as fem_c
"""
import ctypes
import hashlib

import numpy as np
from scipy.sparse import csr_array

from . import _native as nat
from .device import as_dev_f64, as_dev_i32, device
from .materials import model_tables

_PATTERNS = {}
_MAX_PATTERNS = 8


def _pattern(num_dofs, gids_dev, gids_host):
    """sn_pattern of one domain's DOF ids, cached by content (the reference
    calls the FFI once per domain and Newton step with the same ids)."""
    key = (int(num_dofs), gids_host.shape, hashlib.blake2b(gids_host.tobytes(),
                                                            digest_size=16).hexdigest())
    if key not in _PATTERNS:
        if len(_PATTERNS) >= _MAX_PATTERNS:
            _PATTERNS.pop(next(iter(_PATTERNS))).close()
        _PATTERNS[key] = _Pattern(num_dofs, gids_dev, gids_host.shape[0])
    return _PATTERNS[key]


class _Pattern:
    def __init__(self, num_dofs, gids_dev, num_el):
        lib = nat.load()
        self.h = ctypes.c_void_p()
        nat.check(lib.sn_pattern_create(ctypes.byref(self.h), int(num_dofs), int(num_el),
                                        nat.ptr(gids_dev), nat.stream_handle()))
        nnz = ctypes.c_longlong()
        ip, ix, dg = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        nat.check(lib.sn_pattern_info(self.h, ctypes.byref(nnz), ctypes.byref(ip),
                                      ctypes.byref(ix), ctypes.byref(dg)))
        self.nnz = int(nnz.value)
        import torch
        self.indptr = torch.empty(num_dofs + 1, dtype=torch.int32, device=device())
        self.indices = torch.empty(self.nnz, dtype=torch.int32, device=device())
        nat.check(lib.sn_pattern_copy(self.h, nat.ptr(self.indptr), nat.ptr(self.indices), None,
                                      nat.stream_handle()))
        self.indptr_host = self.indptr.cpu().numpy()
        self.indices_host = self.indices.cpu().numpy()

    def close(self):
        if self.h and nat._lib is not None:
            nat._lib.sn_pattern_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _as_host(a, dtype):
    return np.ascontiguousarray(np.asarray(a).reshape(-1), dtype=dtype)


def _tables(cells, glob_ids, curls, d_phi, orientations, num_el_dofs, el_nodes=4):
    """Checked flat host copies and the counts they imply
    (interface_c.pyx:213-222: el_nodes = d_phi.shape[1], num_el =
    len(cells) / el_nodes)."""
    d_phi = np.asarray(d_phi, dtype=np.float64)
    if d_phi.ndim != 3 or d_phi.shape[1:] != (el_nodes, 3):
        raise ValueError('d_phi must be [Q, 4, 3] (got {})'.format(d_phi.shape))
    if int(num_el_dofs) != 6:
        raise ValueError('only lowest-order tetrahedra (6 edge DOFs) are supported '
                         '(num_el_dofs={})'.format(num_el_dofs))
    Q = d_phi.shape[0]
    cells = _as_host(cells, np.int32)
    if cells.size % el_nodes:
        raise ValueError('len(cells) = {} is not a multiple of {}'.format(cells.size, el_nodes))
    num_el = cells.size // el_nodes
    gids = _as_host(glob_ids, np.int32)
    if gids.size != 6 * num_el:
        raise ValueError('glob_ids has {} entries, expected {} (= 6 x {} elements)'
                         .format(gids.size, 6 * num_el, num_el))
    curls = _as_host(curls, np.float64)
    if curls.size == 0 or curls.size % (Q * 18):
        raise ValueError('curls has {} values, not a multiple of Q x 6 x 3 = {}'
                         .format(curls.size, Q * 18))
    n_orient = curls.size // (Q * 18)
    orient = _as_host(orientations, np.int32)
    if orient.size != num_el:
        raise ValueError('orientations has {} entries for {} elements'.format(orient.size,
                                                                             num_el))
    if num_el and (orient.min() < 0 or orient.max() >= n_orient):
        raise ValueError('orientation numbers must lie in [0, {}) (the curl table holds {} '
                         'orientations; got {}..{})'.format(n_orient, n_orient, orient.min(),
                                                            orient.max()))
    return cells, gids.reshape(num_el, 6), curls, d_phi.reshape(-1), orient, Q, num_el, n_orient


def _model(reluctance):
    ts = reluctance.get_type_spec()
    if ts not in (0, 1, 2):
        raise ValueError('reluctance type specifier {} unknown'.format(ts))
    return model_tables(reluctance)


def compute_stiffness_and_jacobi_matrix_curl_curl(nodes, cells, glob_ids, curls, d_phi,
                                                  orientations, x, B_s, reluctance, q, w,
                                                  num_dofs, num_el_dofs):
    """(K, dK, rhs) of one domain: K and dK as ``scipy.sparse.csr_array``
    [num_dofs, num_dofs] over the union of the elements' 6x6 blocks (the COO
    sums of interface_c.pyx:367-368, duplicates summed in element order),
    rhs float64 [num_dofs] (zero when ``B_s`` is empty: the source-free
    kernel, fem_c/ctools.c:960-1193)."""
    import torch
    lib = nat.load()
    num_dofs = int(num_dofs)
    cells, gids, curls, d_phi, orient, Q, num_el, n_orient = _tables(
        cells, glob_ids, curls, d_phi, orientations, num_el_dofs)
    w = _as_host(w, np.float64)
    if w.size != Q:
        raise ValueError('w has {} weights for {} quadrature points'.format(w.size, Q))
    if len(np.asarray(q).reshape(-1)) != 3 * Q:
        raise ValueError('q must hold 3 x {} coordinates'.format(Q))
    x = _as_host(x, np.float64)
    if x.size != num_dofs:
        raise ValueError('x has {} entries for {} DOFs'.format(x.size, num_dofs))
    if num_el and (gids.min() < 0 or gids.max() >= num_dofs):
        raise ValueError('glob_ids must lie in [0, {})'.format(num_dofs))
    B_s = np.asarray(B_s, dtype=np.float64)
    with_src = B_s.size > 0
    if with_src and B_s.size != num_el * Q * 3:
        raise ValueError('B_s has {} values, expected {} (= {} elements x {} points x 3)'
                         .format(B_s.size, num_el * Q * 3, num_el, Q))
    ts, n_int, c, xk = _model(reluctance)
    if num_el == 0:
        z = csr_array((num_dofs, num_dofs))
        return z, z.copy(), np.zeros(num_dofs)
    gids_dev = as_dev_i32(gids.reshape(-1))
    pat = _pattern(num_dofs, gids_dev, gids)
    dev = device()
    locK = torch.empty(21 * num_el, dtype=torch.float64, device=dev)
    locdK = torch.empty_like(locK)
    loc6 = torch.empty(6 * num_el, dtype=torch.float64, device=dev)
    # device copies held in names until the kernel has run (a temporary
    # freed early would hand its memory to the next copy)
    Bs_dev = as_dev_f64(B_s.reshape(-1)) if with_src else None
    nodes_d = as_dev_f64(np.asarray(nodes, dtype=np.float64).reshape(-1))
    cells_d, orient_d = as_dev_i32(cells), as_dev_i32(orient)
    curls_d, dphi_d, w_d = as_dev_f64(curls), as_dev_f64(d_phi), as_dev_f64(w)
    c_d, xk_d, x_d = as_dev_f64(c), as_dev_f64(xk), as_dev_f64(x)
    nat.check(lib.sn_element_table_f64(
        num_el, nat.ptr(nodes_d), nat.ptr(cells_d), nat.ptr(gids_dev), nat.ptr(orient_d),
        n_orient, nat.ptr(curls_d), nat.ptr(dphi_d), nat.ptr(w_d), Q, ts, n_int, nat.ptr(c_d),
        nat.ptr(xk_d), nat.ptr(x_d), nat.ptr(Bs_dev), nat.ptr(locK), nat.ptr(locdK),
        nat.ptr(loc6), nat.stream_handle()))
    K = torch.empty(pat.nnz, dtype=torch.float64, device=dev)
    dK = torch.empty_like(K)
    rhs = torch.empty(num_dofs, dtype=torch.float64, device=dev)
    nat.check(lib.sn_csr_gather_f64(pat.h, nat.ptr(locK), nat.ptr(locdK), nat.ptr(K),
                                    nat.ptr(dK), None, nat.stream_handle()))
    nat.check(lib.sn_dof_gather_f64(pat.h, nat.ptr(loc6), nat.ptr(rhs), nat.stream_handle()))
    shape = (num_dofs, num_dofs)
    Kh = csr_array((K.cpu().numpy(), pat.indices_host.copy(), pat.indptr_host.copy()),
                   shape=shape)
    dKh = csr_array((dK.cpu().numpy(), pat.indices_host.copy(), pat.indptr_host.copy()),
                    shape=shape)
    return Kh, dKh, rhs.cpu().numpy()


def compute_B_Hcurl(nodes, cells, glob_ids, curls, phi, d_phi, orientations, x, q, w, num_dofs,
                    num_el_dofs):
    """(points, B), each float64 [num_el * Q, 3]: the quadrature points of
    every element and curl A there (fem_c/ctools.c:1500-1563)."""
    import torch
    lib = nat.load()
    cells, gids, curls, d_phi, orient, Q, num_el, n_orient = _tables(
        cells, glob_ids, curls, d_phi, orientations, num_el_dofs)
    phi = _as_host(phi, np.float64)
    if phi.size != 4 * Q:
        raise ValueError('phi must be [Q, 4] (got {} values for Q = {})'.format(phi.size, Q))
    x = _as_host(x, np.float64)
    if num_el and (gids.min() < 0 or gids.max() >= x.size):
        raise ValueError('glob_ids must lie in [0, {})'.format(x.size))
    dev = device()
    pts = torch.empty(num_el * Q * 3, dtype=torch.float64, device=dev)
    B = torch.empty_like(pts)
    if num_el:
        hold = [as_dev_f64(np.asarray(nodes, dtype=np.float64).reshape(-1)), as_dev_i32(cells),
                as_dev_i32(gids.reshape(-1)), as_dev_i32(orient), as_dev_f64(curls),
                as_dev_f64(phi), as_dev_f64(d_phi), as_dev_f64(x)]
        nat.check(lib.sn_B_Hcurl_table_f64(
            num_el, nat.ptr(hold[0]), nat.ptr(hold[1]), nat.ptr(hold[2]), nat.ptr(hold[3]),
            n_orient, nat.ptr(hold[4]), nat.ptr(hold[5]), nat.ptr(hold[6]), Q, nat.ptr(hold[7]),
            nat.ptr(pts), nat.ptr(B), nat.stream_handle()))
    return pts.cpu().numpy().reshape(-1, 3), B.cpu().numpy().reshape(-1, 3)


def compute_B_line_segs_cpp(src, tar, current, radius):
    """Biot-Savart field [M, 3] of straight segments src [N, 6] at targets
    tar [M, 3] with core radius ``radius`` (fem_c/ctools.c:1761-1923).  The
    reference prints and returns -1 on a wrong shape (interface_c.pyx:951-959);
    this raises ValueError."""
    from .sources import biot_savart_host
    return biot_savart_host(src, tar, current, radius)
