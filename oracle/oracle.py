"""CPU FP64 oracle of the reduced-vector-potential solve -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker and as the CPU baseline.  The product (snoo.py_amd)
never imports it.

Contents and how each piece is pinned (tests/test_oracle.py):

* ``liboracle.so`` (oracle/snoo_oracle.c): element kernels (ctools.c:960-1388),
  PCHIP evaluation (ctools.c:130-204), B at quadrature points
  (ctools.c:1500-1563), Biot-Savart (ctools.c:1761-1923), COO->CSR,
  Jacobi-PCG (scipy 1.15.3 cg).  Pinned against the reference's own C compiled
  from /root/reference/fem_c/ctools.c (oracle/_ref, where it builds) on
  identical inputs, against committed golden vectors made from it
  (tests/golden/), and against closed forms: uniform-field reproduction,
  K.G = 0, finite-difference Jacobian, straight-wire and square-loop
  Biot-Savart, scipy PchipInterpolator and scipy.sparse.linalg.cg.
* ``assemble_like_reference``: the Python glue of
  compute_stiffness_and_jacobi_matrix_c (curl_curl_assembler.py:815-932):
  per-domain COO->CSR, K += K_d, dK += dK_d, J = dK + K with scipy's sparse
  addition (which drops zero results).
* ``reference_newton``: RedMVPSolver.solve (red_mvp_solver.py:121-224).
* ``residual_newton``: the configuration's inexact Newton on R(x) = K(x)x -
  rhs(x), same forcing-term rule as the product; used for timing.
"""
import ctypes
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_vp = ctypes.c_void_p
_ci = ctypes.c_int
_cd = ctypes.c_double
_cll = ctypes.c_longlong
NU0 = 1.0 / 4.0 / np.pi * 1e7


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, 'liboracle.so')
        if not os.path.exists(path):
            raise RuntimeError('oracle not built: make -C oracle')
        L = ctypes.CDLL(path)
        L.or_pchip_eval.restype = _cd
        L.or_pchip_eval.argtypes = [_ci, _vp, _vp, _cd]
        L.or_pchip_deriv.restype = _cd
        L.or_pchip_deriv.argtypes = [_ci, _vp, _vp, _cd]
        L.or_nu_eval.restype = None
        L.or_nu_eval.argtypes = [_ci, _ci, _vp, _vp, _cd, ctypes.POINTER(_cd), ctypes.POINTER(_cd)]
        L.or_K_dK_red.restype = None
        L.or_K_dK_red.argtypes = [_ci, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _ci, _ci, _ci, _vp, _vp,
                                  _vp, _vp, _vp, _vp, _vp, _vp, _vp]
        L.or_K_dK_red_dil.restype = None
        L.or_K_dK_red_dil.argtypes = [_ci, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _ci, _ci, _ci,
                                      _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
        L.or_K_dK_air.restype = None
        L.or_K_dK_air.argtypes = [_ci, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _ci, _ci, _ci, _vp, _vp,
                                  _vp, _vp, _vp, _vp, _vp]
        L.or_B_Hcurl.restype = None
        L.or_B_Hcurl.argtypes = [_ci, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _ci, _vp, _vp, _vp]
        L.or_biot_savart.restype = None
        L.or_biot_savart.argtypes = [_vp, _ci, _vp, _ci, _cd, _cd, _vp, _ci]
        L.or_coo_to_csr.restype = _cll
        L.or_coo_to_csr.argtypes = [_ci, _cll, _vp, _vp, _vp, _vp, _vp, _vp]
        L.or_spmv.restype = None
        L.or_spmv.argtypes = [_ci, _vp, _vp, _vp, _vp, _vp]
        L.or_pcg.restype = _ci
        L.or_pcg.argtypes = [_ci, _vp, _vp, _vp, _vp, _vp, _vp, _cd, _cd, _ci, ctypes.POINTER(_ci)]
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's own ctools.c compiled into oracle/_ref (None if absent)."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, '_ref', 'libctools_ref.so')
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        R.make_triplet_list.restype = _vp
        R.make_triplet_list.argtypes = [_ci, _vp, _vp, _vp]
        R.make_mesh.restype = _vp
        R.make_mesh.argtypes = [_ci, _ci, _ci, _vp, _vp]
        R.make_quadrature_rule.restype = _vp
        R.make_quadrature_rule.argtypes = [_ci, _vp, _vp]
        R.make_reluctance.restype = _vp
        R.make_reluctance.argtypes = [_ci, _ci, _vp, _vp]
        R.compute_K_dK_Hcurl.restype = None
        R.compute_K_dK_Hcurl.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
        R.compute_K_dK_Hcurl_red.restype = None
        R.compute_K_dK_Hcurl_red.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                             _vp, _vp]
        R.make_dilution.restype = _vp
        R.make_dilution.argtypes = [_cd, _cd, _cd, _cd]
        R.compute_K_dK_Hcurl_red_dil.restype = None
        R.compute_K_dK_Hcurl_red_dil.argtypes = [_vp] * 14
        R.compute_B_Hcurl.restype = None
        R.compute_B_Hcurl.argtypes = [_vp, _vp, _ci, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
        R.compute_B_line_segs.restype = None
        R.compute_B_line_segs.argtypes = [_vp, _vp, _vp, _cd, _cd, _ci, _ci]
        R.eval_pchip_interpolator.restype = _cd
        R.eval_pchip_interpolator.argtypes = [_ci, _vp, _vp, _cd]
        R.eval_pchip_interpolator_derivative.restype = _cd
        R.eval_pchip_interpolator_derivative.argtypes = [_ci, _vp, _vp, _cd]
        _REF = R
    return _REF


def _p(a):
    return a.ctypes.data_as(_vp) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ----------------------------------------------------------------- reluctivity

def nu_eval(type_spec, c, x, B):
    """nu, dnu/dB at the points B (arrays), type 0/1/2."""
    c = _f64(c).reshape(-1)
    x = _f64(x)
    n_int = -1 if type_spec == 0 else len(x) - 1
    nu = np.empty(len(B))
    dnu = np.empty(len(B))
    a, b = _cd(), _cd()
    for i, v in enumerate(np.asarray(B, dtype=np.float64)):
        lib().or_nu_eval(type_spec, n_int, _p(c), _p(x), float(v), ctypes.byref(a), ctypes.byref(b))
        nu[i], dnu[i] = a.value, b.value
    return nu, dnu


# ----------------------------------------------------------------- elements

def element_triplets(nodes, cells, gids, orient, curls, d_phi, w, model, x, Bs=None,
                     phi=None, dilution=None):
    """36 triplets per element and rhs, reference order (ctools.c:960-1388;
    with ``dilution`` = (c_num, m_num, c_den, m_den) and ``phi`` [Q, 4] the
    diluted kernel, ctools.c:1565-1759).

    model = (type_spec, n_int, c_flat, knots)."""
    ts, n_int, c, kn = model
    ne = len(cells) // 4 if np.ndim(cells) == 1 else len(cells)
    Q = len(w)
    nt = 36 * ne
    rows = np.empty(nt, dtype=np.int32)
    cols = np.empty(nt, dtype=np.int32)
    vK = np.empty(nt)
    vdK = np.empty(nt)
    rhs = np.zeros(len(x))
    hold = [_f64(nodes).reshape(-1), _i32(cells).reshape(-1), _i32(gids).reshape(-1),
            _i32(orient), _f64(curls), _f64(d_phi).reshape(-1), _f64(w), _f64(c), _f64(kn),
            _f64(x)]
    args = [ne, _p(hold[0]), _p(hold[1]), _p(hold[2]), _p(hold[3]), _p(hold[4]), _p(hold[5]),
            _p(hold[6]), Q, ts, n_int, _p(hold[7]), _p(hold[8]), _p(hold[9])]
    if Bs is not None and dilution is not None:
        Bsf = _f64(Bs).reshape(-1)
        ph = _f64(phi).reshape(-1)
        dl = _f64(dilution)
        a = args[:6] + [_p(ph)] + args[6:]
        lib().or_K_dK_red_dil(*a, _p(Bsf), _p(dl), _p(rows), _p(cols), _p(vK), _p(vdK), _p(rhs))
    elif Bs is not None:
        Bsf = _f64(Bs).reshape(-1)
        lib().or_K_dK_red(*args, _p(Bsf), _p(rows), _p(cols), _p(vK), _p(vdK), _p(rhs))
    else:
        lib().or_K_dK_air(*args, _p(rows), _p(cols), _p(vK), _p(vdK))
    return rows, cols, vK, vdK, rhs


def coo_to_csr(n, rows, cols, vals):
    """scipy csr_array((v,(i,j))) semantics (duplicates summed, sorted)."""
    from scipy.sparse import csr_array
    nt = len(rows)
    indptr = np.empty(n + 1, dtype=np.int32)
    indices = np.empty(max(nt, 1), dtype=np.int32)
    data = np.empty(max(nt, 1))
    h = [_i32(rows), _i32(cols), _f64(vals)]
    nnz = lib().or_coo_to_csr(n, nt, _p(h[0]), _p(h[1]), _p(h[2]), _p(indptr), _p(indices),
                              _p(data))
    return csr_array((data[:nnz].copy(), indices[:nnz].copy(), indptr), shape=(n, n))


def _model_of(rel):
    import importlib
    mats = importlib.import_module(type(rel).__module__)
    ts, n_int, c, x = mats.model_tables(rel)
    return ts, n_int, c, x


def domain_inputs(asm, n, quad_order):
    """Host arrays of domain n of a product assembler, in the reference layout."""
    from importlib import import_module
    fe = import_module(type(asm).__module__.rsplit('.', 1)[0] + '.finite_element')
    w, q = fe.quadrature_rule(quad_order)
    Q = len(w)
    tm = asm.tet_mesh
    sel = asm._sel[n]
    return dict(nodes=tm.nodes, cells=tm.cells[sel], gids=tm.glob_ids[sel], orient=tm.orient[sel],
                curls=fe.curl_table(Q), d_phi=fe.lagrange_basis_derivative(q), w=w,
                phi=fe.lagrange_basis(q), Q=Q)


def assemble_like_reference(asm, x, quad_order, source_fields, dilutions=None):
    """(K, J, rhs) following curl_curl_assembler.py:815-932 with the oracle's
    element kernels (source_fields[n] None/empty -> air kernel); with
    ``dilutions`` (one (c_num, m_num, c_den, m_den) or None per domain) the
    diluted assembler, curl_curl_assembler_diluted.py:165-304."""
    from scipy.sparse import csr_array
    n = asm.num_dofs
    K = csr_array((n, n))
    dK = csr_array((n, n))
    rhs = np.zeros(n)
    use_src = len(source_fields) == asm.num_materials
    for d in range(asm.num_materials):
        inp = domain_inputs(asm, d, quad_order)
        model = _model_of(asm.material_list[d])
        bs = source_fields[d] if use_src else None
        if bs is not None and len(bs) == 0:
            bs = None
        dil = dilutions[d] if dilutions is not None else None
        r, c, vK, vdK, rh = element_triplets(inp['nodes'], inp['cells'], inp['gids'],
                                             inp['orient'], inp['curls'], inp['d_phi'], inp['w'],
                                             model, x, bs, phi=inp['phi'], dilution=dil)
        K += coo_to_csr(n, r, c, vK)
        dK += coo_to_csr(n, r, c, vdK)
        rhs += rh
    return K, dK + K, rhs


def B_hcurl(nodes, cells, gids, orient, curls, phi, d_phi, x):
    ne = len(cells)
    Q = len(phi)
    pts = np.empty(ne * Q * 3)
    B = np.empty(ne * Q * 3)
    h = [_f64(nodes).reshape(-1), _i32(cells).reshape(-1), _i32(gids).reshape(-1), _i32(orient),
         _f64(curls), _f64(phi).reshape(-1), _f64(d_phi).reshape(-1), _f64(x)]
    lib().or_B_Hcurl(ne, _p(h[0]), _p(h[1]), _p(h[2]), _p(h[3]), _p(h[4]), _p(h[5]), _p(h[6]), Q,
                     _p(h[7]), _p(pts), _p(B))
    return pts.reshape(-1, 3), B.reshape(-1, 3)


def quad_points(nodes, cells, phi):
    X = np.asarray(nodes)[np.asarray(cells)]            # [ne,4,3]
    return np.einsum('qa,eak->eqk', np.asarray(phi), X).reshape(-1, 3)


def biot_savart(src, tar, current, rad, threads=1):
    src = _f64(src).reshape(-1, 6)
    tar = _f64(tar).reshape(-1, 3)
    B = np.empty_like(tar)
    lib().or_biot_savart(_p(src), len(src), _p(tar), len(tar), float(current), float(rad), _p(B),
                         int(threads))
    return B


def coil_field(coils, pts, threads=1, disc_arc=5):
    """Sum of RacetrackCoil fields (segments from the coil's get_segments)."""
    B = np.zeros((len(pts), 3))
    for c in coils:
        B += biot_savart(c.get_segments(disc_arc=disc_arc), pts, c.current, c.rad, threads)
    return B


# ----------------------------------------------------------------- solvers

def pcg(A, b, x0, mask, rtol=1e-5, atol=0.0, maxiter=-1):
    """Masked Jacobi-PCG; returns (x, iterations, info)."""
    A = A.tocsr()
    A.sort_indices()
    n = A.shape[0]
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    if maxiter < 0:
        maxiter = 10 * int(m.sum())
    x = _f64(x0).copy()
    info = _ci()
    h = [_i32(A.indptr), _i32(A.indices), _f64(A.data), _f64(b)]
    it = lib().or_pcg(n, _p(h[0]), _p(h[1]), _p(h[2]), _p(m), _p(h[3]), _p(x), float(rtol),
                      float(atol), int(maxiter), ctypes.byref(info))
    return x, it, info.value


def scipy_pcg(A, b, x0, mask, rtol=1e-5, atol=0.0, maxiter=-1):
    """The reference's own linear solve: scipy's cg with the Jacobi
    preconditioner on J[mask][:, mask] (snoopy/red_mvp_solver.py:195-205);
    returns (x, iterations, info) like pcg."""
    from scipy.sparse import diags
    from scipy.sparse.linalg import cg
    A = A.tocsr()
    m = np.asarray(mask, dtype=bool)
    Ac = A[m][:, m]
    M = diags(1.0 / Ac.diagonal(), shape=Ac.shape)
    count = [0]

    def cb(_):
        count[0] += 1
    kw = {} if maxiter < 0 else {'maxiter': int(maxiter)}
    xc, info = cg(Ac, np.asarray(b, dtype=np.float64)[m], x0=_f64(x0)[m], rtol=rtol, atol=atol,
                  M=M, callback=cb, **kw)
    x = _f64(x0).copy()
    x[m] = xc
    return x, count[0], info


class OracleProblem:
    """Everything the CPU solvers need, extracted from a product solver object
    (mesh arrays, tables, models, source fields, Dirichlet mask).  ``csr``:
    "oracle" converts the element triplets with or_coo_to_csr, "scipy" with
    scipy's csr_array((v, (i, j))) as fem_c/interface_c.pyx:367-368 does."""

    def __init__(self, solver, source_fields=None, threads=1, csr='oracle'):
        self.csr = csr
        asm = solver.curl_curl_factory
        self.asm = asm
        self.n = asm.num_dofs
        self.quad_order = solver.quad_order
        self.mask = solver.make_boundary_mask()
        self.domains = [domain_inputs(asm, d, self.quad_order) for d in range(asm.num_materials)]
        self.models = [_model_of(m) for m in asm.material_list]
        n_iron = len(solver.iron_tags)
        if source_fields is None:
            t0 = time.perf_counter()
            source_fields = []
            for d in range(asm.num_materials):
                if d < n_iron:
                    inp = self.domains[d]
                    pts = quad_points(inp['nodes'], inp['cells'], inp['phi'])
                    source_fields.append(coil_field(solver.coil_list, pts, threads))
                else:
                    source_fields.append(None)
            self.t_source = time.perf_counter() - t0
        self.src = source_fields

    def assemble(self, x):
        """(K, dK, rhs) as scipy CSR (per-domain COO->CSR then summed)."""
        from scipy.sparse import csr_array
        n = self.n
        K = csr_array((n, n))
        dK = csr_array((n, n))
        rhs = np.zeros(n)
        for d, inp in enumerate(self.domains):
            bs = self.src[d]
            r, c, vK, vdK, rh = element_triplets(inp['nodes'], inp['cells'], inp['gids'],
                                                 inp['orient'], inp['curls'], inp['d_phi'],
                                                 inp['w'], self.models[d], x, bs)
            if self.csr == 'scipy':
                K = K + csr_array((vK, (r, c)), shape=(n, n))
                dK = dK + csr_array((vdK, (r, c)), shape=(n, n))
            else:
                K = K + coo_to_csr(n, r, c, vK)
                dK = dK + coo_to_csr(n, r, c, vdK)
            rhs += rh
        return K, dK, rhs


def reference_newton(prob, tolerance=1e-4, x_0=None, maxiter=-1, max_newton=25):
    """RedMVPSolver.solve (red_mvp_solver.py:121-224) on the CPU."""
    n = prob.n
    x_n = np.zeros(n) if x_0 is None else _f64(x_0).copy()
    mask = prob.mask
    hist = []
    x_np1 = x_n
    for i in range(max_newton):
        K, dK, rhs = prob.assemble(x_n)
        J = dK + K
        b = J @ x_n - K @ x_n + rhs
        if np.max(np.abs(b)) < 1e-14:
            x_np1 = 0.0 * x_n
            break
        xx, it, info = pcg(J, b, x_n, mask, rtol=1e-5, atol=tolerance, maxiter=maxiter)
        x_np1 = np.where(mask, xx, 0.0)
        h = x_np1 - x_n
        hist.append((it, np.max(np.abs(h)) / np.max(np.abs(x_np1))))
        if np.max(np.abs(h)) < tolerance * np.max(np.abs(x_np1)):
            break
        x_n = x_np1.copy()
    return x_np1, hist


def residual_norm(prob, x):
    K, dK, rhs = prob.assemble(x)
    R = K @ x - rhs
    return float(np.linalg.norm(R[prob.mask])), K, dK, R


def residual_newton(prob, rtol=1e-8, x0=None, max_newton=40, eta_max=0.1, eta_min=1e-6,
                    line_search=True, linear=None):
    """Inexact Newton on R(x) = K(x)x - rhs(x) with the product's forcing rule;
    ``linear`` is the Jacobi-PCG used (pcg, the restatement, by default)."""
    linear = linear or pcg
    n = prob.n
    r0, _, _, _ = residual_norm(prob, np.zeros(n))
    x = np.zeros(n) if x0 is None else np.where(prob.mask, _f64(x0), 0.0)
    rn, K, dK, R = residual_norm(prob, x)
    hist, its = [rn / r0], []
    rn_prev = None
    eta = eta_max
    k = 0
    while rn > rtol * r0 and k < max_newton:
        if rn_prev is not None:
            eta = min(eta_max, max(0.9 * (rn / rn_prev) ** 2, eta_min))
        eta = max(eta, min(eta_max, 0.5 * rtol * r0 / rn))
        J = dK + K
        dx, it, info = linear(J, -R, np.zeros(n), prob.mask, rtol=eta, atol=0.0)
        its.append(it)
        x_try = x + np.where(prob.mask, dx, 0.0)
        rn_try, K2, dK2, R2 = residual_norm(prob, x_try)
        step = 1.0
        tries = 0
        while line_search and not (rn_try < rn) and tries < 6:
            step *= 0.5
            x_try = x + step * np.where(prob.mask, dx, 0.0)
            rn_try, K2, dK2, R2 = residual_norm(prob, x_try)
            tries += 1
        rn_prev, rn, x, K, dK, R = rn, rn_try, x_try, K2, dK2, R2
        hist.append(rn / r0)
        k += 1
    return x, {'residual_history': hist, 'cg_iterations': its, 'r0': r0}


def _quad_eval(p1, p2, p3, p4, u, v):
    """QuadGeometry.eval (brick_field.py:21-37)."""
    return (p1 * 0.25 * (u + 1) * (v + 1) + p2 * 0.25 * (1 - u) * (v + 1)
            + p3 * 0.25 * (1 - u) * (1 - v) + p4 * 0.25 * (u + 1) * (1 - v))


def _quad_normal(p1, p2, p3, p4, u, v):
    """QuadGeometry.eval_normal (brick_field.py:39-70)."""
    du = p1 * 0.25 * (v + 1) - p2 * 0.25 * (v + 1) - p3 * 0.25 * (1 - v) + p4 * 0.25 * (1 - v)
    dv = p1 * 0.25 * (u + 1) + p2 * 0.25 * (1 - u) - p3 * 0.25 * (1 - u) - p4 * 0.25 * (u + 1)
    n = np.array([du[1] * dv[2] - du[2] * dv[1], du[2] * dv[0] - du[0] * dv[2],
                  du[0] * dv[1] - du[1] * dv[0]])
    return n / np.linalg.norm(n)


def brick_field(brick, points):
    """BrickField.compute_B restated point by point (brick_field.py:400-433 with
    BrickGeometry sides, brick_field.py:137-158, and is_inside, :165-194): a
    point is inside when no face has (p - centre).normal > 0; the first brick
    that contains it sets the field."""
    order = [(3, 2, 1, 0), (4, 5, 6, 7), (5, 4, 0, 1), (6, 5, 1, 2), (2, 3, 7, 6), (3, 0, 4, 7)]
    vals = [np.array([0.0, brick.Phi / brick.area_1, 0.0]),
            np.array([brick.Phi / brick.area_2, 0.0, 0.0]),
            np.array([0.0, -1.0 * brick.Phi / brick.area_3, 0.0])]
    sides = []
    for P in brick.bricks:
        sides.append([(_quad_eval(*[P[i] for i in o], 0.0, 0.0),
                       _quad_normal(*[P[i] for i in o], 0.0, 0.0)) for o in order])
    out = np.zeros((len(points), 3))
    for m, p in enumerate(np.asarray(points, dtype=np.float64)):
        for d in range(3):
            if all(np.sum((p - r) * nrm) <= 0.0 for r, nrm in sides[d]):
                out[m] = vals[d]
                break
    return out


# ------------------------------------------------------------ probe fields

def nodal_fields(num_nodes, cells_list, B_el_list):
    """compute_nodal_fields (curl_curl_assembler.py:2089-2197): element vectors
    of the chosen domains summed into their nodes, divided by the number of
    those elements at each node (nodes with none stay 0).  cells_list /
    B_el_list: one [n_el, 4] / [n_el, 3] array per chosen domain."""
    B = np.zeros((num_nodes, 3))
    mult = np.zeros(num_nodes, dtype=np.int64)
    for cells, Be in zip(cells_list, B_el_list):
        for j in range(len(cells)):
            for a in cells[j]:
                B[a] += Be[j]
                mult[a] += 1
    m = mult != 0
    B[m] /= mult[m][:, None]
    return B


def p1_interpolate(points, nodes, cells, el, field_of_point, nodal_list):
    """evaluate_interpolation (evaluators.py:297-411) for located points:
    local coordinates (u, v, w) of the point in its element, P1 basis
    (1-u-v-w, u, v, w), the nodal field of the point's domain; points with
    no element or domain get 0."""
    out = np.zeros((len(points), 3))
    for p in range(len(points)):
        e, f = el[p], field_of_point[p]
        if e < 0 or f < 0:
            continue
        X = nodes[cells[e]]
        M = np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]], axis=1)
        u, v, w = np.linalg.solve(M, points[p] - X[0])
        phi = np.array([1.0 - u - v - w, u, v, w])
        out[p] = phi @ nodal_list[f][cells[e]]
    return out
