"""CPU oracle of the auxiliary-space preconditioner -- TEST INFRASTRUCTURE.

The preconditioner has no counterpart in the reference (its CG uses the
Jacobi diagonal, snoopy/red_mvp_solver.py:198-199), so this restatement
follows the textbook construction it implements, in float64 with scipy:

- Hiptmair & Xu, "Nodal auxiliary space preconditioning in H(curl) and
  H(div) spaces", SIAM J. Numer. Anal. 45 (2007): M = S + Pi B_L Pi^T, with
  Pi the Whitney interpolation of P1 vector fields (edge DOF = line integral
  of the linear interpolant, (x_hi - x_lo)/2 . (y_lo + y_hi)) and B_L a
  preconditioner of the nodal vector Laplacian.
- B_L: V(1,1)-cycle with damped Jacobi smoothing, piecewise-constant
  prolongation P from the aggregates and Galerkin coarse matrices P^T A P,
  exact solve on the coarsest level.

It builds every operator from the mesh itself (element stiffness by inverting
the element Jacobian, P from the aggregate ranges, products with scipy) and
takes from the product only the level-0 numbering and the aggregate ranges,
so the product's gather plans, kernels and coarse inverse are all checked.
Pins (tests/test_ams_oracle.py): Pi reproduces the edge DOFs of linear
fields; the Laplacian has zero row sums, is exact on linear functions and its
energy of a linear function is sum nu_e vol_e |grad|^2; the recursive V-cycle
equals the closed two-level operator S + (I - S A) S + (I - S A) P C^-1 P^T
(I - A S); the preconditioner is symmetric.
"""
import numpy as np
import scipy.sparse as sp


def element_stiffness(nodes, cells, nu_e):
    """[n_el, 4, 4] nu_e vol_e grad(l_a).grad(l_b) via the inverse Jacobian."""
    X = nodes[cells]
    Jm = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=2)  # cols
    Jinv = np.linalg.inv(Jm)
    ref = np.array([[-1.0, -1.0, -1.0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    G = np.einsum('ak,mkj->maj', ref, Jinv)           # grad l_a = J^-T grad_ref
    vol = np.abs(np.linalg.det(Jm)) / 6.0
    return np.einsum('maj,mbj->mab', G, G) * (np.asarray(nu_e) * vol)[:, None, None]


def laplacian_full(nodes, cells, nu_e):
    nn = len(nodes)
    Ke = element_stiffness(nodes, cells, nu_e)
    r = np.repeat(cells, 4, axis=1).ravel()
    c = np.tile(cells, (1, 4)).ravel()
    return sp.csr_array((Ke.ravel(), (r, c)), shape=(nn, nn))


def interpolation_full(nodes, edges):
    """Pi [n_edges, 3 n_nodes] for edges (lo, hi)."""
    ne = len(edges)
    t = 0.5 * (nodes[edges[:, 1]] - nodes[edges[:, 0]])
    rows = np.repeat(np.arange(ne), 6)
    cols = np.concatenate([3 * edges[:, :1] + np.arange(3), 3 * edges[:, 1:] + np.arange(3)],
                          axis=1).ravel()
    vals = np.concatenate([t, t], axis=1).ravel()
    return sp.csr_array((vals, (rows, cols)), shape=(ne, 3 * len(nodes)))


def prolongation(agg_ptr):
    agg_ptr = np.asarray(agg_ptr)
    nc = len(agg_ptr) - 1
    nf = int(agg_ptr[-1])
    rows = np.arange(nf)
    cols = np.repeat(np.arange(nc), np.diff(agg_ptr))
    return sp.csr_array((np.ones(nf), (rows, cols)), shape=(nf, nc))


class OracleAms:
    """Float64 preconditioner for one mesh; ``level0_of_node[i]`` is the
    level-0 index of mesh node i (>= n_free: Dirichlet node), ``agg_ptrs``
    the aggregate ranges of levels 1.. (finest first)."""

    def __init__(self, nodes, edges, cells, free_edge, nu_e, level0_of_node, agg_ptrs,
                 omega=0.6, weight=1.0):
        nodes = np.asarray(nodes, dtype=np.float64).reshape(-1, 3)
        edges = np.asarray(edges).reshape(-1, 2)
        self.free_edge = np.asarray(free_edge).astype(bool)
        l0 = np.asarray(level0_of_node)
        nf = int(np.count_nonzero(~self._fixed_nodes(edges, self.free_edge, len(nodes))))
        self.n_free = nf
        free_nodes = np.nonzero(l0 < nf)[0]
        # selection/permutation: level-0 index -> mesh node
        node_of_l0 = np.empty(nf, dtype=np.int64)
        node_of_l0[l0[free_nodes]] = free_nodes
        Q = sp.csr_array((np.ones(nf), (np.arange(nf), node_of_l0)), shape=(nf, len(nodes)))
        Lf = laplacian_full(nodes, cells, nu_e)
        A0 = (Q @ Lf @ Q.T).tocsr()
        self.A = [A0]
        self.P = []
        for ap in agg_ptrs:
            P = prolongation(ap)
            self.P.append(P)
            self.A.append((P.T @ self.A[-1] @ P).tocsr())
        Q3 = sp.kron(Q, sp.identity(3), format='csr')
        self.Pi = (interpolation_full(nodes, edges) @ Q3.T).tocsr()   # free-node columns
        self.omega, self.weight = omega, weight

    @staticmethod
    def _fixed_nodes(edges, free_edge, nn):
        fixed = np.zeros(nn, dtype=bool)
        fixed[edges[~free_edge].ravel()] = True
        return fixed

    def vcycle(self, b, l=0):
        A = self.A[l]
        if l == len(self.A) - 1:
            return np.linalg.solve(A.toarray(), b)
        w = self.omega / A.diagonal()
        x = w * b
        P = self.P[l]
        x = x + P @ self.vcycle(P.T @ (b - A @ x), l + 1)
        return x + w * (b - A @ x)

    def aux(self, b3):
        """B_L on three components: b3 [n_free, 3]."""
        return np.stack([self.vcycle(b3[:, k].copy()) for k in range(3)], axis=1)

    def apply(self, r, dinv):
        r = np.where(self.free_edge, r, 0.0)
        b3 = (self.Pi.T @ r).reshape(-1, 3)
        y = self.aux(b3).reshape(-1)
        z = dinv * r + self.weight * (self.Pi @ y)
        return np.where(self.free_edge, z, 0.0)
