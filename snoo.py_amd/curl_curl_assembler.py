"""Curl-curl assembler on the device (snoopy/curl_curl_assembler.py:17-2197).

Keeps the reference's constructor and methods:

* ``CurlCurlAssembler(mesh, volume_tags, material_list, element_order=1)``
  with attributes ``nodes, cells, cell_types, cell_tags, global_ids,
  num_dofs, material_list, num_materials, volume_tags``
  (snoopy/curl_curl_assembler.py:19-137);
* ``compute_stiffness_and_jacobi_matrix_c(x, quad_order=2, source_fields=[])
  -> (K, J, rhs)`` with K, J as ``scipy.sparse.csr_array``
  (snoopy/curl_curl_assembler.py:815-932);
* ``get_quadrature_points`` (1521-1593), ``compute_B`` (1722-1802),
  ``compute_nodal_fields`` (2089-2197).

Behind them: the element kernel, the deterministic CSR scatter and the field
kernels of libsnoo_hip.so.  Mesh arrays are uploaded once; the element-to-
nonzero map is built once per mesh (``sn_pattern_create``).  Device-resident
variants (``*_device``) serve the Newton solvers without host copies.
"""
import ctypes

import numpy as np
from scipy.sparse import csr_array

from . import _native as nat
from .device import as_dev_f64, as_dev_i32, device
from .finite_element import (NUM_ORIENTATIONS, curl_table, lagrange_basis,
                             lagrange_basis_derivative, quadrature_rule)
from .materials import model_tables


class _Tables:
    """Device copies of the reference-element tables for one quadrature order."""

    def __init__(self, order):
        self.order = order
        w, q = quadrature_rule(order)
        self.w_host, self.q_host = w, q
        self.Q = len(w)
        self.phi_host = lagrange_basis(q)
        self.dphi_host = lagrange_basis_derivative(q)
        self.curls_host = curl_table(self.Q)
        self.w = as_dev_f64(w)
        self.phi = as_dev_f64(self.phi_host.reshape(-1))
        self.dphi = as_dev_f64(self.dphi_host.reshape(-1))
        self.curls = as_dev_f64(self.curls_host)


class _Model:
    """Device copy of one reluctivity model's table."""

    def __init__(self, rel):
        self.type_spec, self.n_int, c, x = model_tables(rel)
        self.c = as_dev_f64(c)
        self.x = as_dev_f64(x)


class CurlCurlAssembler:

    def __init__(self, mesh, volume_tags, material_list, element_order=1):
        if element_order != 1:
            raise NotImplementedError('only lowest-order (Whitney) edge elements are implemented')
        tm = mesh._tm if hasattr(mesh, '_tm') else getattr(mesh, 'tet_mesh', mesh)
        self.mesh = mesh
        self.tet_mesh = tm
        self.element_order = element_order
        self.material_list = list(material_list)
        self.num_dofs = tm.num_edges
        self.nodes = tm.nodes
        self.volume_tags = list(volume_tags)
        if len(self.volume_tags) == 0:
            sel_list = [np.arange(tm.num_elements)]
            self.num_materials = 1
        else:
            sel_list = [np.nonzero(tm.tags == tag)[0] for tag in self.volume_tags]
            self.num_materials = len(self.material_list)
        # reference layout: per domain, lists over cell types (one type here)
        self.cell_types, self.cells, self.cell_tags, self.global_ids = [], [], [], []
        self._sel = sel_list
        for sel in sel_list:
            self.cell_types.append([4])
            self.cells.append([(tm.cells[sel].astype(np.int64) + 1).reshape(-1)])
            self.cell_tags.append(sel.astype(np.int64) + 1)
            self.global_ids.append(tm.glob_ids[sel] + 1)
        self._dev = None
        self._tables = {}
        self._models = {}
        self.dilutions = [None] * len(self.material_list)

    def _dilution(self, n):
        """(c_num, m_num, c_den, m_den) of domain n as a C double[4], or None."""
        dil = self.dilutions[n] if n < len(self.dilutions) else None
        if dil is None:
            return None
        return (ctypes.c_double * 4)(dil.c_num, dil.m_num, dil.c_den, dil.m_den)

    # ------------------------------------------------------------ device state
    def _device_state(self):
        if self._dev is not None:
            return self._dev
        import torch
        tm = self.tet_mesh
        order = np.concatenate(self._sel) if self._sel else np.zeros(0, dtype=np.int64)
        offs = np.cumsum([0] + [len(s) for s in self._sel])
        d = {
            'order': order,
            'offsets': offs,
            'nodes': as_dev_f64(tm.nodes.reshape(-1)),
            'cells': as_dev_i32(tm.cells[order].reshape(-1)),
            'gids': as_dev_i32(tm.glob_ids[order].reshape(-1)),
            'orient': as_dev_i32(tm.orient[order]),
            'num_el': int(len(order)),
        }
        lib = nat.load()
        handle = ctypes.c_void_p()
        nat.check(lib.sn_pattern_create(ctypes.byref(handle), int(self.num_dofs), d['num_el'],
                                        nat.ptr(d['gids']), nat.stream_handle()))
        d['pattern'] = handle
        nnz = ctypes.c_longlong()
        ip, ix, dg = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        nat.check(lib.sn_pattern_info(handle, ctypes.byref(nnz), ctypes.byref(ip),
                                      ctypes.byref(ix), ctypes.byref(dg)))
        d['nnz'] = int(nnz.value)
        d['indptr_ptr'], d['indices_ptr'], d['diag_ptr'] = ip.value, ix.value, dg.value
        self._dev = d
        self._torch = torch
        return d

    def __del__(self):
        try:
            if self._dev is not None and nat._lib is not None:
                nat._lib.sn_pattern_destroy(self._dev['pattern'])
                self._dev = None
        except Exception:
            pass

    def tables(self, order):
        if order not in self._tables:
            self._tables[order] = _Tables(order)
        return self._tables[order]

    def model(self, n):
        rel = self.material_list[n]
        key = id(rel)
        if key not in self._models:
            self._models[key] = (_Model(rel), rel)
        return self._models[key][0]

    @property
    def pattern(self):
        return self._device_state()['pattern']

    @property
    def nnz(self):
        return self._device_state()['nnz']

    def csr_pattern_host(self):
        """(indptr, indices) of the device CSR pattern as numpy int32."""
        import torch
        d = self._device_state()
        n, nnz = self.num_dofs, d['nnz']
        ip = torch.empty(n + 1, dtype=torch.int32, device=device())
        ix = torch.empty(nnz, dtype=torch.int32, device=device())
        nat.check(nat.load().sn_pattern_copy(d['pattern'], nat.ptr(ip), nat.ptr(ix), None,
                                             nat.stream_handle()))
        return ip.cpu().numpy(), ix.cpu().numpy()

    def pattern_device(self):
        """(indptr, indices, diag) as device int32 tensors (copies)."""
        import torch
        d = self._device_state()
        n, nnz = self.num_dofs, d['nnz']
        ip = torch.empty(n + 1, dtype=torch.int32, device=device())
        ix = torch.empty(nnz, dtype=torch.int32, device=device())
        dg = torch.empty(n, dtype=torch.int32, device=device())
        nat.check(nat.load().sn_pattern_copy(d['pattern'], nat.ptr(ip), nat.ptr(ix), nat.ptr(dg),
                                             nat.stream_handle()))
        return ip, ix, dg

    # ------------------------------------------------------------ kernels
    def _source_dev(self, source_fields, n, Q):
        if len(source_fields) != self.num_materials:
            return None
        bs = source_fields[n]
        if bs is None or len(bs) == 0:
            return None
        t = as_dev_f64(bs).reshape(-1)
        ne = len(self._sel[n])
        if t.numel() != ne * Q * 3:
            raise ValueError('source field of domain {} has {} values, expected {} (= {} elements x '
                             '{} points x 3)'.format(n, t.numel(), ne * Q * 3, ne, Q))
        return t

    def assemble_local_device(self, x_dev, quad_order, source_fields):
        """Local K, dK (21 x N_el) and rhs (6 x N_el) in the SoA layout."""
        import torch
        d = self._device_state()
        T = self.tables(quad_order)
        N = d['num_el']
        locK = torch.empty(21 * N, dtype=torch.float64, device=device())
        locdK = torch.empty(21 * N, dtype=torch.float64, device=device())
        loc6 = torch.empty(6 * N, dtype=torch.float64, device=device())
        lib = nat.load()
        keep = []
        for n in range(self.num_materials):
            b, e = int(d['offsets'][n]), int(d['offsets'][n + 1])
            m = self.model(n)
            bs = self._source_dev(source_fields, n, T.Q)
            keep.append(bs)
            nat.check(lib.sn_element_local_f64(
                N, b, e - b, nat.ptr(d['nodes']), nat.ptr(d['cells']), nat.ptr(d['gids']),
                nat.ptr(d['orient']), nat.ptr(T.curls), nat.ptr(T.dphi), nat.ptr(T.w), T.Q,
                m.type_spec, m.n_int, nat.ptr(m.c), nat.ptr(m.x), nat.ptr(x_dev), nat.ptr(bs),
                nat.ptr(locK), nat.ptr(locdK), nat.ptr(loc6), nat.ptr(T.phi), self._dilution(n),
                nat.stream_handle()))
        return locK, locdK, loc6

    def assemble_device(self, x, quad_order=2, source_fields=()):
        """Device CSR values of K, dK, J = dK + K and rhs (float64)."""
        import torch
        d = self._device_state()
        x_dev = as_dev_f64(x)
        locK, locdK, loc6 = self.assemble_local_device(x_dev, quad_order, source_fields)
        nnz = d['nnz']
        K = torch.empty(nnz, dtype=torch.float64, device=device())
        dK = torch.empty_like(K)
        J = torch.empty_like(K)
        rhs = torch.empty(self.num_dofs, dtype=torch.float64, device=device())
        lib = nat.load()
        nat.check(lib.sn_csr_gather_f64(d['pattern'], nat.ptr(locK), nat.ptr(locdK), nat.ptr(K),
                                        nat.ptr(dK), nat.ptr(J), nat.stream_handle()))
        nat.check(lib.sn_dof_gather_f64(d['pattern'], nat.ptr(loc6), nat.ptr(rhs),
                                        nat.stream_handle()))
        return K, dK, J, rhs

    def newton_local_device(self, x_dev, quad_order, source_fields_dev, mode, f32, locJ, loc6,
                            domains=None, nu_el=None):
        """Fused element kernel of the Newton solvers over all domains (or the
        listed ones; the local values of the others are left as they are).
        ``nu_el`` (float32 device tensor, one entry per element in device
        order, optional) receives each element's mean reluctivity."""
        d = self._device_state()
        T = self.tables(quad_order)
        N = d['num_el']
        lib = nat.load()
        for n in (range(self.num_materials) if domains is None else domains):
            b, e = int(d['offsets'][n]), int(d['offsets'][n + 1])
            m = self.model(n)
            bs = source_fields_dev[n] if source_fields_dev is not None else None
            nat.check(lib.sn_element_newton(
                int(bool(f32)), int(mode), N, b, e - b, nat.ptr(d['nodes']), nat.ptr(d['cells']),
                nat.ptr(d['gids']), nat.ptr(d['orient']), nat.ptr(T.curls), nat.ptr(T.dphi),
                nat.ptr(T.w), T.Q, m.type_spec, m.n_int, nat.ptr(m.c), nat.ptr(m.x),
                nat.ptr(x_dev), nat.ptr(bs), nat.ptr(locJ), nat.ptr(loc6),
                nat.ptr(nu_el) if nu_el is not None else None, nat.ptr(T.phi), self._dilution(n),
                nat.stream_handle()))

    def compute_stiffness_and_jacobi_matrix_c(self, x, quad_order=2, source_fields=[]):
        """(K, J, rhs) as the reference returns them (curl_curl_assembler.py:815-932).

        K and J are scipy csr_arrays without explicitly stored zeros, which is
        what the reference's sparse additions ``K += this_K`` and ``dK + K``
        leave (scipy's binary ops drop zero results)."""
        K, dK, J, rhs = self.assemble_device(x, quad_order, source_fields)
        ip, ix = self.csr_pattern_host()
        n = self.num_dofs
        Kh = csr_array((K.cpu().numpy(), ix.copy(), ip.copy()), shape=(n, n))
        Jh = csr_array((J.cpu().numpy(), ix.copy(), ip.copy()), shape=(n, n))
        Kh.eliminate_zeros()
        Jh.eliminate_zeros()
        return Kh, Jh, rhs.cpu().numpy()

    # ------------------------------------------------------------ fields
    def get_quadrature_points_device(self, quad_order=3, domain_ids=()):
        import torch
        d = self._device_state()
        T = self.tables(quad_order)
        ids = list(domain_ids) if len(domain_ids) > 0 else list(range(self.num_materials))
        lib = nat.load()
        out = []
        for n in ids:
            b, e = int(d['offsets'][n]), int(d['offsets'][n + 1])
            pts = torch.empty(((e - b) * T.Q, 3), dtype=torch.float64, device=device())
            cells = d['cells'][4 * b:4 * e]
            nat.check(lib.sn_quad_points_f64(e - b, nat.ptr(d['nodes']), nat.ptr(cells),
                                             nat.ptr(T.phi), T.Q, nat.ptr(pts),
                                             nat.stream_handle()))
            out.append(pts)
        return out

    def get_quadrature_points(self, quad_order=3, domain_ids=[]):
        """Per-domain (M x 3) quadrature points (curl_curl_assembler.py:1521-1593)."""
        return [p.cpu().numpy() for p in self.get_quadrature_points_device(quad_order, domain_ids)]

    def compute_B_device(self, x_dev, quad_order=8):
        import torch
        d = self._device_state()
        T = self.tables(quad_order)
        N = d['num_el']
        pts = torch.empty((N * T.Q, 3), dtype=torch.float64, device=device())
        B = torch.empty_like(pts)
        lib = nat.load()
        nat.check(lib.sn_B_Hcurl_f64(N, nat.ptr(d['nodes']), nat.ptr(d['cells']), nat.ptr(d['gids']),
                                     nat.ptr(d['orient']), nat.ptr(T.curls), nat.ptr(T.phi), T.Q,
                                     nat.ptr(x_dev), nat.ptr(pts), nat.ptr(B),
                                     nat.stream_handle()))
        return pts, B

    def compute_B(self, x, quad_order=8):
        """(points, curl A) at the quadrature points of every domain
        (curl_curl_assembler.py:1722-1802); the caller adds the coil field."""
        pts, B = self.compute_B_device(as_dev_f64(x), quad_order)
        return pts.cpu().numpy(), B.cpu().numpy()

    def element_B(self, x):
        """Element-constant curl A per element of the mesh (mesh element order)."""
        d = self._device_state()
        _, B = self.compute_B_device(as_dev_f64(x), quad_order=1)
        out = np.empty((self.tet_mesh.num_elements, 3))
        out[d['order']] = B.cpu().numpy()
        return out

    def compute_nodal_fields(self, x, mat_ids):
        """Element-averaged B at the mesh nodes (curl_curl_assembler.py:2089-2197):
        the element-constant curl A of the elements of the listed material tags,
        averaged at each node over those elements (0 at nodes none touches)."""
        from .evaluators import nodal_fields_device
        return nodal_fields_device(self, as_dev_f64(x), list(mat_ids)).cpu().numpy()


class CurlCurlAssemblerDiluted(CurlCurlAssembler):
    """Curl-curl assembler with diluted steel (snoopy/curl_curl_assembler_diluted.py:17-300):
    one dilution object (or None) per domain; domains with a dilution and a
    source field use the diluted reluctivity s nu(s |B|) with s from the
    quadrature point's z (fem_c/ctools.c:1565-1759)."""

    def __init__(self, mesh, volume_tags, material_list, dilutions_list, element_order=1):
        super().__init__(mesh, volume_tags, material_list, element_order)
        self.dilutions = list(dilutions_list) + [None] * (len(self.material_list) -
                                                          len(dilutions_list))
