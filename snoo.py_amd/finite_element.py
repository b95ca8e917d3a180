"""Reference-element tables for lowest-order Nedelec (Whitney) tetrahedra.

The reference obtains these tables from gmsh (snoopy/finite_element.py:153-220:
``getIntegrationPoints("Gauss<order>")``, ``getBasisFunctions(..., "GradLagrange")``
and ``"CurlHcurlLegendre0"`` for every orientation).  There is no gmsh here, so
they are derived in closed form:

* reference tetrahedron (0,0,0),(1,0,0),(0,1,0),(0,0,1) with barycentrics
  lam0 = 1-u-v-w, lam1 = u, lam2 = v, lam3 = w;
* Whitney edge function of local edge (a,b): w = lam_a grad lam_b - lam_b grad lam_a,
  whose reference curl is the constant 2 grad lam_a x grad lam_b; its edge
  integral along a->b is 1, so an edge DOF is the line integral of A;
* an element orientation is a 6-bit mask; bit j set means that the global
  direction of local edge j (lower global node id -> higher) runs b->a, which
  flips the sign of that basis function.  The curl table is laid out as the
  reference's ``curls`` buffer, ``curls[((o*Q + q)*6 + j)*3 + c]`` with o in
  [0, 64) (fem_c/ctools.c:1280,1315 index it the same way).

Quadrature rules (weights sum to the reference volume 1/6, as the reference's
kernel requires since it multiplies w_q by det J, fem_c/ctools.c:1319):
order 0/1 centroid, order 2 the Hammer-Stroud 4-point rule, order 3 the
5-point Keast rule, order >= 4 the Stroud conical-product rule with
ceil((order+1)/2) Gauss-Jacobi points per direction (exact to degree
2*ceil((order+1)/2)-1 >= order).
"""
import math

import numpy as np

# local edges (a, b), a < b
LOCAL_EDGES = np.array([(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)], dtype=np.int32)
NUM_ORIENTATIONS = 64

# gradients of the barycentrics on the reference tetrahedron
_GRAD_LAM = np.array([[-1.0, -1.0, -1.0],
                      [1.0, 0.0, 0.0],
                      [0.0, 1.0, 0.0],
                      [0.0, 0.0, 1.0]])


def _gauss_jacobi_01(n, alpha):
    """n-point Gauss rule on [0,1] for the weight (1-x)^alpha (Golub-Welsch)."""
    a, b = float(alpha), 0.0
    # recurrence of the Jacobi polynomials P^(a,b) on [-1,1]
    diag = np.empty(n)
    for i in range(n):
        den = (2 * i + a + b) * (2 * i + a + b + 2)
        diag[i] = (b * b - a * a) / den if den != 0 else (b - a) / (a + b + 2)
    off = np.empty(max(n - 1, 0))
    for i in range(1, n):
        num = 4 * i * (i + a) * (i + b) * (i + a + b)
        den = (2 * i + a + b) ** 2 * (2 * i + a + b + 1) * (2 * i + a + b - 1)
        off[i - 1] = math.sqrt(num / den)
    T = np.diag(diag) + np.diag(off, 1) + np.diag(off, -1)
    x, V = np.linalg.eigh(T)
    mu0 = 2.0 ** (a + b + 1) * math.gamma(a + 1) * math.gamma(b + 1) / math.gamma(a + b + 2)
    w = mu0 * V[0, :] ** 2
    # map [-1,1] -> [0,1]: (1-x)^a becomes 2^a (1-t)^a, dx = 2 dt
    t = 0.5 * (x + 1.0)
    w = w / 2.0 ** (a + 1)
    return t, w


def quadrature_rule(order):
    """Return (w[Q], q[3Q]) for the reference tetrahedron, gmsh-style flat q."""
    order = int(order)
    if order <= 1:
        pts = np.array([[0.25, 0.25, 0.25]])
        w = np.array([1.0 / 6.0])
    elif order == 2:
        a = (5.0 + 3.0 * math.sqrt(5.0)) / 20.0
        b = (5.0 - math.sqrt(5.0)) / 20.0
        pts = np.array([[b, b, b], [a, b, b], [b, a, b], [b, b, a]])
        w = np.full(4, 1.0 / 24.0)
    elif order == 3:
        pts = np.array([[0.25, 0.25, 0.25],
                        [1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0],
                        [0.5, 1.0 / 6.0, 1.0 / 6.0],
                        [1.0 / 6.0, 0.5, 1.0 / 6.0],
                        [1.0 / 6.0, 1.0 / 6.0, 0.5]])
        w = np.array([-2.0 / 15.0, 3.0 / 40.0, 3.0 / 40.0, 3.0 / 40.0, 3.0 / 40.0])
    else:
        n = (order + 2) // 2
        t1, w1 = _gauss_jacobi_01(n, 2)
        t2, w2 = _gauss_jacobi_01(n, 1)
        t3, w3 = _gauss_jacobi_01(n, 0)
        P, W = [], []
        for i in range(n):
            for j in range(n):
                for k in range(n):
                    u = t1[i]
                    v = t2[j] * (1.0 - t1[i])
                    ww = t3[k] * (1.0 - t1[i]) * (1.0 - t2[j])
                    P.append((u, v, ww))
                    W.append(w1[i] * w2[j] * w3[k])
        pts = np.array(P)
        w = np.array(W)
    return w.astype(np.float64), pts.reshape(-1).astype(np.float64)


def lagrange_basis(q):
    """P1 basis values phi[Q,4] at flat reference points q[3Q]."""
    p = np.asarray(q, dtype=np.float64).reshape(-1, 3)
    return np.stack([1.0 - p.sum(axis=1), p[:, 0], p[:, 1], p[:, 2]], axis=1)


def lagrange_basis_derivative(q):
    """P1 basis gradients d_phi[Q,4,3] (reference layout of GradLagrange)."""
    Q = len(q) // 3
    return np.broadcast_to(_GRAD_LAM, (Q, 4, 3)).copy()


def reference_curls():
    """Curl of the 6 reference Whitney functions, [6,3] (positive direction)."""
    c = np.empty((6, 3))
    for j, (a, b) in enumerate(LOCAL_EDGES):
        c[j] = 2.0 * np.cross(_GRAD_LAM[a], _GRAD_LAM[b])
    return c


def curl_table(Q):
    """Curl table for all 64 orientations: flat [64*Q*6*3]."""
    c = reference_curls()
    tab = np.empty((NUM_ORIENTATIONS, Q, 6, 3))
    for o in range(NUM_ORIENTATIONS):
        s = np.array([-1.0 if (o >> j) & 1 else 1.0 for j in range(6)])
        tab[o] = (s[:, None] * c)[None, :, :]
    return tab.reshape(-1)


def orientations_for_cells(cells):
    """6-bit orientation mask per element from the global node ids [N,4]."""
    cells = np.asarray(cells)
    o = np.zeros(cells.shape[0], dtype=np.int32)
    for j, (a, b) in enumerate(LOCAL_EDGES):
        o |= (cells[:, a] > cells[:, b]).astype(np.int32) << j
    return o


class FiniteElement:
    """Tables of the linear tetrahedron, mirroring snoopy/finite_element.py:4-220."""

    def __init__(self, gmsh_code=4):
        self.gmsh_code = gmsh_code   # 4 = 4-node tetrahedron in gmsh numbering
        self.num_nodes = 4

    def get_number_of_nodes(self):
        return 4

    def get_quadrature_rule(self, order):
        return quadrature_rule(order)

    def evaluate_basis(self, q):
        return lagrange_basis(q)

    def evaluate_basis_derivative(self, q):
        return lagrange_basis_derivative(q)

    def get_curl_edge_basis_functions(self, mesh, loc, element_type=4, element_order=1):
        Q = len(loc) // 3
        return curl_table(Q), NUM_ORIENTATIONS
