"""Auxiliary-space (Hiptmair-Xu) preconditioner for the Newton CG.

No counterpart in the reference, which preconditions scipy's ``cg`` with the
Jacobi diagonal only (snoopy/red_mvp_solver.py:198-205).  The kernels, the
numeric refresh and the apply live in ``csrc/sn_ams.hip``; this module builds
the host-side plan once per mesh:

- Dirichlet nodes are the end points of Dirichlet edges; the auxiliary space
  is the P1 vector fields on the other ("free") nodes.
- Level 0 is the P1 Laplacian with element reluctivity on the free nodes.
  Coarser levels come from greedy aggregation of the strength graph
  (``sn_amg_aggregate``) and Galerkin sums A_c = P^T A P with piecewise
  constant P, until at most ``coarse_max`` rows remain.
- Every level is renumbered so that each aggregate's rows are contiguous
  (coarsest level first, then each finer level sorted by the new index of its
  aggregate); restriction is then a contiguous range sum on the device.
- Gather plans give each nonzero of every level the ordered list of values
  that sum to it (element pair values for level 0, finer-level nonzeros
  otherwise), so the per-Newton-step refresh is a fixed-order gather.

The aggregation uses the reluctivities of the first Jacobian; later updates
keep the aggregates and refresh all values.
"""
import ctypes
import time

import numpy as np

from . import _native as nat
from .device import device as nu_device

# local node pairs (a <= b) of the 4x4 element matrix, in slot order
PAIR_SLOTS = [(a, b) for a in range(4) for b in range(a, 4)]


def element_laplacians(nodes, cells, nu_e):
    """[n_el, 10] values nu_e vol_e grad(l_a).grad(l_b) for the slot pairs."""
    X = nodes[cells]                                            # [m,4,3]
    E = X[:, 1:, :] - X[:, :1, :]                               # rows e1, e2, e3
    g = np.empty((len(cells), 4, 3))
    for c in range(3):
        g[:, c + 1] = np.cross(E[:, (c + 1) % 3], E[:, (c + 2) % 3])
    det = np.einsum('mk,mk->m', E[:, 0], g[:, 1])
    g[:, 1:] /= det[:, None, None]
    g[:, 0] = -g[:, 1:].sum(axis=1)
    s = np.asarray(nu_e, dtype=np.float64) * np.abs(det) / 6.0
    return np.stack([s * np.einsum('mk,mk->m', g[:, a], g[:, b]) for a, b in PAIR_SLOTS], axis=1)


def _coo_to_csr(n, rows, cols, vals):
    """Sorted CSR (scipy csr_array) summing duplicates and keeping every
    structurally present entry, also those whose sum is zero."""
    from scipy.sparse import csr_array
    keys = np.asarray(rows, dtype=np.int64) * n + np.asarray(cols, dtype=np.int64)
    uk, inv = np.unique(keys, return_inverse=True)
    data = np.bincount(inv.ravel(), weights=np.asarray(vals, dtype=np.float64), minlength=len(uk))
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(uk // n, minlength=n), out=indptr[1:])
    return csr_array((data, (uk % n).astype(np.int64), indptr), shape=(n, n))


def _csr_positions(indptr, indices, n_cols, rows, cols):
    """Positions of the entries (rows, cols) in a CSR pattern with sorted rows."""
    rowid = np.repeat(np.arange(len(indptr) - 1, dtype=np.int64), np.diff(indptr))
    keys = rowid * n_cols + indices.astype(np.int64)
    q = rows.astype(np.int64) * n_cols + cols.astype(np.int64)
    pos = np.searchsorted(keys, q)
    assert np.all(keys[pos] == q), 'entry outside the pattern'
    return pos


def _gather_plan(pos, values, nnz):
    """(ptr [nnz+1], list): the values landing on each position, in input order."""
    order = np.argsort(pos, kind='stable')
    ptr = np.zeros(nnz + 1, dtype=np.int32)
    np.cumsum(np.bincount(pos, minlength=nnz), out=ptr[1:])
    return ptr, np.ascontiguousarray(values[order], dtype=np.int32)


def aggregate(A, theta):
    """Aggregate id of every row of the scipy CSR matrix A (C++ helper)."""
    ip = np.ascontiguousarray(A.indptr, dtype=np.int32)
    ix = np.ascontiguousarray(A.indices, dtype=np.int32)
    v = np.ascontiguousarray(A.data, dtype=np.float64)
    agg = np.empty(A.shape[0], dtype=np.int32)
    na = ctypes.c_int()
    nat.check(nat.load().sn_amg_aggregate(A.shape[0], nat.ptr_np(ip), nat.ptr_np(ix),
                                          nat.ptr_np(v), float(theta), nat.ptr_np(agg),
                                          ctypes.byref(na)))
    return agg, int(na.value)


# strength threshold of the aggregation.  Measured on one MI355X (1M-tet
# dipole: levels 156 555 / 24 671 / 3 847 / 556 / 75 at 0.08, 156 555 /
# 6 120 / 240 / 12 at 0.25):
#   0.08: 8 learned-start variants 97.4 solves/s, CG 77-79 each; from zero
#         46.5; 64 variants of 20k-335k tets 226; 6M-tet chain 0.583 s
#   0.12, 0.18: 94.7 and 84.1; 44.5 and 38.6; 204 and 218; 0.584 and 0.648 s
#   0.25: 81.4, CG 146 for half the variants and 85 for the others (its
#         coarser levels leave a plateau in the CG near a relative residual
#         of 4e-3); 47.9; 227; 0.557 s
# 0.08 is as fast elsewhere within 5 % and keeps the learned start's CG
# counts even.
AMS_THETA = 0.08


class AmsPlan:
    """Host arrays of the preconditioner for one mesh and Dirichlet mask."""

    def __init__(self, nodes, edges, cells, free_edge, nu_e, theta=AMS_THETA, coarse_max=128,
                 max_levels=12, passes=1):
        t0 = time.perf_counter()
        nodes = np.asarray(nodes, dtype=np.float64).reshape(-1, 3)
        edges = np.asarray(edges).reshape(-1, 2)
        cells = np.asarray(cells).reshape(-1, 4)
        free_edge = np.asarray(free_edge).astype(bool)
        nn, n_el = len(nodes), len(cells)
        fixed = np.zeros(nn, dtype=bool)
        fixed[edges[~free_edge].ravel()] = True
        free_ids = np.nonzero(~fixed)[0]
        nf = len(free_ids)
        comp = np.full(nn, -1, dtype=np.int64)
        comp[free_ids] = np.arange(nf)

        # level 0 in the compact numbering of the free nodes
        loc = element_laplacians(nodes, cells, nu_e)                     # [n_el, 10]
        ca = comp[cells]                                                 # [n_el, 4]
        pairs = [(a, b, s) for s, (a, b) in enumerate(PAIR_SLOTS)]
        pairs += [(b, a, s) for s, (a, b) in enumerate(PAIR_SLOTS) if a != b]
        R = np.stack([ca[:, a] for a, b, s in pairs], axis=1)            # [n_el, 16]
        C = np.stack([ca[:, b] for a, b, s in pairs], axis=1)
        S = np.array([s for a, b, s in pairs])
        V = loc[:, S]
        ok = (R >= 0) & (C >= 0)
        A = _coo_to_csr(nf, R[ok], C[ok], V[ok])
        levels, aggs = [A], []
        while levels[-1].shape[0] > coarse_max and len(levels) < max_levels:
            Al = levels[-1]
            agg, na = aggregate(Al, theta)
            if na > 0.5 * Al.shape[0]:
                # too few strong connections (anisotropic or weakly coupled
                # rows): aggregate over every connection instead
                agg, na = aggregate(Al, 0.0)
            for _ in range(passes - 1):
                # aggregates of aggregates: fewer, larger levels
                fr = np.repeat(np.arange(Al.shape[0]), np.diff(Al.indptr))
                At = _coo_to_csr(na, agg[fr], agg[Al.indices], Al.data)
                agg2, na2 = aggregate(At, theta)
                if na2 > 0.5 * na:
                    agg2, na2 = aggregate(At, 0.0)
                agg, na = agg2[agg], na2
            if na >= Al.shape[0]:
                break
            fr = np.repeat(np.arange(Al.shape[0]), np.diff(Al.indptr))
            levels.append(_coo_to_csr(na, agg[fr], agg[Al.indices], Al.data))
            aggs.append(agg)
        L = len(levels)
        if levels[-1].shape[0] > 1024:
            raise RuntimeError('auxiliary-space hierarchy: coarsest level has %d rows'
                               % levels[-1].shape[0])

        # renumber: aggregates contiguous on every level
        perm = [None] * L
        perm[L - 1] = np.arange(levels[L - 1].shape[0])
        for l in range(L - 2, -1, -1):
            key = perm[l + 1][aggs[l]]
            order = np.argsort(key, kind='stable')
            perm[l] = np.empty(len(order), dtype=np.int64)
            perm[l][order] = np.arange(len(order))
        self.levels = []
        for l in range(L):
            Al = levels[l]
            fr = np.repeat(np.arange(Al.shape[0]), np.diff(Al.indptr))
            B = _coo_to_csr(Al.shape[0], perm[l][fr], perm[l][Al.indices], Al.data)
            lev = {'n': B.shape[0], 'nnz': B.nnz, 'indptr': B.indptr.astype(np.int32),
                   'indices': B.indices.astype(np.int32), 'init_vals': B.data.copy()}
            if l > 0:
                coarse_new = perm[l][aggs[l - 1]][np.argsort(perm[l - 1])]  # of new fine rows
                cnt = np.bincount(coarse_new, minlength=B.shape[0])
                agg_ptr = np.zeros(B.shape[0] + 1, dtype=np.int32)
                np.cumsum(cnt, out=agg_ptr[1:])
                assert np.all(np.diff(coarse_new) >= 0), 'aggregates not contiguous'
                lev['agg_ptr'] = agg_ptr
                lev['coarse_of_fine'] = coarse_new
            self.levels.append(lev)

        # gather plans
        p0 = perm[0]
        lv0 = self.levels[0]
        rows = np.where(ok, p0[np.where(R >= 0, R, 0)], -1)
        cols = np.where(ok, p0[np.where(C >= 0, C, 0)], -1)
        src = S[None, :] * n_el + np.arange(n_el)[:, None]               # slot*n_el + e
        pos = _csr_positions(lv0['indptr'], lv0['indices'], nf, rows[ok], cols[ok])
        lv0['gptr'], lv0['glist'] = _gather_plan(pos, src[ok], lv0['nnz'])
        for l in range(1, L):
            f, c = self.levels[l - 1], self.levels[l]
            fr = np.repeat(np.arange(f['n']), np.diff(f['indptr']))
            cof = c['coarse_of_fine']
            pos = _csr_positions(c['indptr'], c['indices'], c['n'], cof[fr], cof[f['indices']])
            c['gptr'], c['glist'] = _gather_plan(pos, np.arange(f['nnz']), c['nnz'])

        # edge side
        node_new = np.full(nn, nf, dtype=np.int64)
        node_new[free_ids] = p0
        self.edge_nodes = np.ascontiguousarray(node_new[edges].astype(np.int32))   # [n_e, 2]
        tan = np.zeros((len(edges), 4), dtype=np.float32)
        tan[:, :3] = 0.5 * (nodes[edges[:, 1]] - nodes[edges[:, 0]])
        self.edge_tan = tan
        self.edge_mesh = np.ascontiguousarray(edges.astype(np.int32))
        fe = np.nonzero(free_edge)[0]
        inc_node = np.concatenate([node_new[edges[fe, 0]], node_new[edges[fe, 1]]])
        inc_edge = np.concatenate([fe, fe])
        keep = inc_node < nf
        inc_node, inc_edge = inc_node[keep], inc_edge[keep]
        order = np.lexsort((inc_edge, inc_node))
        self.inc_edge = np.ascontiguousarray(inc_edge[order].astype(np.int32))
        self.inc_ptr = np.zeros(nf + 1, dtype=np.int32)
        np.cumsum(np.bincount(inc_node, minlength=nf), out=self.inc_ptr[1:])
        self.n_free, self.n_el, self.n_edges = nf, n_el, len(edges)
        self.free_ids, self.perm0 = free_ids, p0
        self.t_setup = time.perf_counter() - t0

    def sizes(self):
        return [lv['n'] for lv in self.levels]

    def operator_complexity(self):
        return sum(lv['nnz'] for lv in self.levels) / self.levels[0]['nnz']


class AmsPreconditioner:
    """Device handle of the preconditioner for one assembler and mask."""

    def __init__(self, asm, free_edge, nu_el_dev, omega=0.7, weight=2.0, theta=AMS_THETA,
                 coarse_max=128, passes=1):
        d = asm._device_state()
        tm = asm.tet_mesh
        cells = tm.cells[d['order']]
        nu_host = nu_el_dev.cpu().numpy().astype(np.float64)
        self.plan = AmsPlan(tm.nodes, tm.edges, cells, free_edge, nu_host, theta=theta,
                            coarse_max=coarse_max, passes=passes)
        p = self.plan
        lib = nat.load()
        self.handle = ctypes.c_void_p()
        nat.check(lib.sn_ams_create(ctypes.byref(self.handle), p.n_edges, p.n_free, p.n_el,
                                    nat.ptr_np(p.edge_nodes), nat.ptr_np(p.edge_mesh),
                                    nat.ptr_np(p.inc_ptr), nat.ptr_np(p.inc_edge),
                                    float(omega), float(weight), nat.stream_handle()))
        for l, lv in enumerate(p.levels):
            nat.check(lib.sn_ams_add_level(self.handle, lv['n'], lv['nnz'],
                                           nat.ptr_np(lv['indptr']), nat.ptr_np(lv['indices']),
                                           nat.ptr_np(lv['gptr']), nat.ptr_np(lv['glist']),
                                           nat.ptr_np(lv['agg_ptr']) if l > 0 else None,
                                           nat.stream_handle()))
        self._asm = asm
        self.omega, self.weight = float(omega), float(weight)
        self.update(nu_el_dev)

    def update(self, nu_el_dev):
        d = self._asm._device_state()
        nat.check(nat.load().sn_ams_update(self.handle, nat.ptr(d['nodes']), nat.ptr(d['cells']),
                                           nat.ptr(nu_el_dev), nat.stream_handle()))

    def apply(self, mask_dev, r_dev, dinv_dev):
        import torch
        z = torch.empty_like(r_dev)
        nat.check(nat.load().sn_ams_apply(self.handle, nat.ptr(mask_dev), nat.ptr(r_dev),
                                          nat.ptr(dinv_dev), nat.ptr(z), nat.stream_handle()))
        return z

    def level_values(self, level):
        """Current values of one level (float32 device tensor)."""
        import torch
        n, nnz = ctypes.c_int(), ctypes.c_int()
        lib = nat.load()
        nat.check(lib.sn_ams_level_info(self.handle, level, ctypes.byref(n), ctypes.byref(nnz)))
        out = torch.empty(nnz.value, dtype=torch.float32, device=nu_device())
        nat.check(lib.sn_ams_level_values(self.handle, level, nat.ptr(out), nat.stream_handle()))
        return out

    def __del__(self):
        try:
            if self.handle and nat._lib is not None:
                nat._lib.sn_ams_destroy(self.handle)
                self.handle = None
        except Exception:
            pass
