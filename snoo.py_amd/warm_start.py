"""Learned warm start: parameters -> POD coefficients -> initial edge DOFs.

Pipeline:

1. ``MorphFamily``: one tensor-grid topology (fixed cell counts per segment)
   whose grid lines follow the dipole parameters (yoke W, H, L, aperture ax,
   ay).  Every variant shares cells, edges, DOF numbering and Dirichlet DOFs,
   so converged solutions of different variants are vectors of one space.
2. ``generate``: the solver itself produces the converged A of many seeded
   variants on that topology (mixed-precision Newton).
3. ``POD``: mean and leading modes of the snapshots (method of snapshots on
   the host, float64 eigen-decomposition of the Gram matrix).
4. ``MLP``: 8 normalised parameters -> 4 x 512 GELU -> r normalised POD
   coefficients, trained with the BF16 matrix-core kernels (mlp.py).
5. ``WarmStart.predict``: MLP forward, POD expansion (``sn_pod_expand``) and
   Whitney interpolation onto the target mesh (``sn_edge_transfer``).

Plain Newton from A = 0 stays the fallback: ``solve_with_warm_start`` keeps
the prediction only if its residual is below the residual at A = 0.
"""
import os
import time

import numpy as np

from . import _native as nat
from .device import as_dev_f64, as_dev_i32, device
from .finite_element import LOCAL_EDGES
from .magnets import DEFAULT, Dipole
from .materials import tanh_reluctance
from .mesh import AIR_TAG, IRON_TAG, TetModel, kuhn_mesh, locate_points

# parameter ranges of the warm-start family (uniform sampling)
PARAM_NAMES = ('W', 'H', 'ax', 'ay', 'J', 'B_sat', 'mu_r', 'L')
PARAM_LO = np.array([0.6, 0.4, 0.15, 0.1, 0.5e6, 1.7, 2000.0, 0.3])
PARAM_HI = np.array([1.4, 0.9, 0.4, 0.3, 3.0e6, 2.1, 6000.0, 0.8])

MODEL_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'models')


def sample_params(seed):
    rng = np.random.default_rng(seed)
    p = rng.uniform(PARAM_LO, PARAM_HI)
    v = dict(zip(PARAM_NAMES, p))
    v['ax'] = min(v['ax'], v['W'] - 0.2)
    v['ay'] = min(v['ay'], v['H'] - 0.2)
    return v


def normalise(params_list):
    P = np.array([[float(p[k]) for k in PARAM_NAMES] for p in params_list])
    return (2.0 * (P - PARAM_LO) / (PARAM_HI - PARAM_LO) - 1.0).astype(np.float32)


def _fixed_axis(seg_bounds, counts, pad, n_pad, ratio):
    """Grid lines with fixed counts: graded pads, uniform inner segments."""
    g = ratio ** np.arange(n_pad)
    sizes = g / g.sum() * pad
    lo, hi = seg_bounds[0], seg_bounds[-1]
    left = lo - np.cumsum(sizes)[::-1]
    parts = [left]
    for a, b, n in zip(seg_bounds[:-1], seg_bounds[1:], counts):
        parts.append(np.linspace(a, b, n + 1)[:-1])
    parts.append(hi + np.concatenate([[0.0], np.cumsum(sizes)]))
    return np.concatenate(parts)


def _breaks(p, pad):
    """Geometry break points of each axis of a single-yoke dipole, outer
    boundary (yoke bounds +- pad) included."""
    bx = [-p['W'] / 2 - pad, -p['W'] / 2, -p['ax'] / 2, p['ax'] / 2, p['W'] / 2, p['W'] / 2 + pad]
    by = [-p['H'] / 2 - pad, -p['H'] / 2, -p['ay'] / 2, p['ay'] / 2, p['H'] / 2, p['H'] / 2 + pad]
    bz = [-p['L'] / 2 - pad, -p['L'] / 2, p['L'] / 2, p['L'] / 2 + pad]
    return bx, by, bz


class MorphFamily:
    """Fixed-topology dipole meshes over the warm-start parameter ranges.

    Two constructions share one rule: every variant's grid lines are those of
    a reference grid moved so that the geometry's break points (pad
    boundary, yoke faces, aperture faces) land on the variant's.
    * ``counts = (n_pad, n_wall_x, n_ap_x, n_wall_y, n_ap_y, n_len)``: a grid
      with these cell counts per segment (graded pads, uniform inner
      segments); the default gives 205 632 tetrahedra.
    * ``base`` (a single-yoke Dipole): that dipole's own mesh is the
      reference, moved piecewise-linearly between the break points, so the
      family contains the dipole's mesh exactly (its parameters give its
      nodes)."""

    def __init__(self, counts=(7, 6, 10, 6, 8, 14), pad=0.5, ratio=1.3, base=None):
        self.counts = tuple(int(c) for c in counts)
        self.pad, self.ratio = pad, ratio
        self.base = base
        if base is not None:
            if len(base.yokes) != 1 or base.p.get('z0', 0.0) != 0.0:
                raise ValueError('MorphFamily: the base dipole must have one yoke at z0 = 0')
            self.pad = base.p['pad']
            self.p0 = {k: base.p[k] for k in PARAM_NAMES}
            self._ref_axes = base.axes
            self._ref_breaks = _breaks(base.p, self.pad)
            self.mesh = base.mesh
            self.model = base.model
            return
        self.p0 = {k: DEFAULT[k] for k in PARAM_NAMES}
        base_axes = self.axes(dict(DEFAULT))
        p0 = dict(DEFAULT)

        def tag_fn(xc, yc, zc):
            in_box = (np.abs(xc) < p0['W'] / 2) & (np.abs(yc) < p0['H'] / 2) & \
                     (np.abs(zc) < p0['L'] / 2)
            in_ap = (np.abs(xc) < p0['ax'] / 2) & (np.abs(yc) < p0['ay'] / 2)
            return np.where(in_box & ~in_ap, IRON_TAG, AIR_TAG)
        self.mesh = kuhn_mesh(*base_axes, tag_fn)
        self.model = TetModel(self.mesh)

    def axes(self, p):
        p = dict(self.base.p if self.base is not None else DEFAULT, **p)
        if self.base is not None:
            new = _breaks(p, self.pad)
            return tuple(np.interp(a, r, b) for a, r, b in zip(self._ref_axes, self._ref_breaks,
                                                               new))
        n_pad, nwx, nax, nwy, nay, nl = self.counts
        xs = _fixed_axis([-p['W'] / 2, -p['ax'] / 2, p['ax'] / 2, p['W'] / 2], [nwx, nax, nwx],
                         self.pad, n_pad, self.ratio)
        ys = _fixed_axis([-p['H'] / 2, -p['ay'] / 2, p['ay'] / 2, p['H'] / 2], [nwy, nay, nwy],
                         self.pad, n_pad, self.ratio)
        zs = _fixed_axis([-p['L'] / 2, p['L'] / 2], [nl], self.pad, n_pad, self.ratio)
        return xs, ys, zs

    def nodes(self, p):
        xs, ys, zs = self.axes(p)
        X, Y, Z = np.meshgrid(xs, ys, zs, indexing='ij')
        return np.stack([X.transpose(2, 1, 0).ravel(), Y.transpose(2, 1, 0).ravel(),
                         Z.transpose(2, 1, 0).ravel()], axis=1)

    def variant(self, p, model=None):
        """Dipole object of parameters p on this topology (coils and reluctivity);
        ``model`` (a TetModel of a private copy of the family mesh) replaces
        the shared one."""
        full = dict(self.base.p if self.base is not None else DEFAULT)
        full.update(p)
        d = Dipole.__new__(Dipole)
        d.p = full
        d.h = None
        d.yokes = [dict(W=full['W'], H=full['H'], L=full['L'], ax=full['ax'], ay=full['ay'],
                        z0=0.0, sign=1.0)]
        d.coils = d._coils_for(d.yokes[0])
        d.reluctance = tanh_reluctance(full['B_sat'], full['mu_r'])
        d.model = model if model is not None else self.model
        d.mesh = d.model.tet_mesh
        return d

    def private_model(self):
        """A TetModel over a copy of the family mesh with its own node array
        (solvers on several streams hold different variants at once)."""
        import copy
        m = copy.copy(self.mesh)
        m.nodes = self.mesh.nodes.copy()
        return TetModel(m)


class FamilySolver:
    """One RedMVPSolver on the family topology, re-targeted per variant by
    replacing node coordinates, coils and reluctivity on the device."""

    def __init__(self, family, quad_order=2, private=True):
        self.family = family
        self.model = family.private_model() if private else family.model
        d0 = family.variant(dict(family.p0), self.model)
        self.solver = d0.solver(quad_order=quad_order, verbose=False)
        self.asm = self.solver.curl_curl_factory
        self.asm._device_state()

    def set_variant(self, p):
        import torch
        d = self.family.variant(p, self.model)
        nodes = self.family.nodes(d.p)
        self.asm.tet_mesh.nodes[:] = nodes
        self.asm._dev['nodes'].copy_(torch.from_numpy(nodes.reshape(-1)).to(device()))
        self.asm.geometry_version = getattr(self.asm, 'geometry_version', 0) + 1
        self.solver.coil_list = d.coils
        self.solver.curl_curl_factory.material_list[0] = d.reluctance
        self.solver._src = None
        return d

    def solve(self, p, rtol=1e-6, precision='mixed', x0=None, precond='ams'):
        self.set_variant(p)
        return self.solver.solve_newton(precision, rtol=rtol, x0=x0, precond=precond)


class EnergyInner:
    """The inner product (x, y)_K0 = x^T K0 y with K0 the curl-curl matrix of
    unit reluctivity on the family's reference geometry (the default
    parameters).  K0 annihilates discrete gradients (the curl of a gradient
    is zero whatever the geometry), so this inner product sees only the field
    B = curl A: two solutions that differ by a gradient, which the Newton
    solver's iterates pick up from the preconditioner, are the same vector in
    it.  K0 is held as float64 element matrices in reference-curl form
    (sn_element_newton mode 2 at A = 0, no source) and applied with
    sn_element_apply_f64."""

    def __init__(self, fsolver):
        import torch
        lib = nat.load()
        fsolver.set_variant(dict(fsolver.family.p0))
        asm = fsolver.asm
        d = asm._device_state()
        self.d, self.n = d, asm.num_dofs
        N = d['num_el']
        T = asm.tables(fsolver.solver.quad_order)
        dev = device()
        self.Mr = torch.empty(6 * N, dtype=torch.float64, device=dev)
        loc6 = torch.empty(6 * N, dtype=torch.float64, device=dev)
        x0 = torch.zeros(self.n, dtype=torch.float64, device=dev)
        one = as_dev_f64(np.array([1.0]))
        knots = as_dev_f64(np.zeros(1))
        nat.check(lib.sn_element_newton(
            0, 2, N, 0, N, nat.ptr(d['nodes']), nat.ptr(d['cells']), nat.ptr(d['gids']),
            nat.ptr(d['orient']), nat.ptr(T.curls), nat.ptr(T.dphi), nat.ptr(T.w), T.Q, 0, -1,
            nat.ptr(one), nat.ptr(knots), nat.ptr(x0), None, nat.ptr(self.Mr), nat.ptr(loc6), None,
            nat.ptr(T.phi), None, nat.stream_handle()))
        self.mask = fsolver.solver.mask_device()
        torch.cuda.current_stream().synchronize()

    def apply(self, x, out=None, work=None):
        """K0 x (float64 device vectors; Dirichlet rows and columns removed)."""
        import torch
        if out is None:
            out = torch.empty(self.n, dtype=torch.float64, device=device())
        if work is None:
            work = torch.empty(6 * self.d['num_el'], dtype=torch.float64, device=device())
        nat.check(nat.load().sn_element_apply_f64(
            self.d['pattern'], nat.ptr(self.d['gids']), nat.ptr(self.d['orient']),
            nat.ptr(self.Mr), nat.ptr(self.mask), nat.ptr(x), nat.ptr(work), nat.ptr(out),
            nat.stream_handle()))
        return out


def generate(family_solvers, seeds, rtol=1e-8, energy=None, log=None):
    """Converged edge DOFs of the variants with the given seeds, float32
    [len(seeds), n] on the host, and (with ``energy``) K0 x of each.  With
    several family solvers the variants are shared out over as many threads,
    each on its own HIP stream (default priority)."""
    import threading
    import torch
    fss = family_solvers if isinstance(family_solvers, (list, tuple)) else [family_solvers]
    seeds = list(seeds)
    n = fss[0].asm.num_dofs
    X = np.empty((len(seeds), n), dtype=np.float32)
    KX = np.empty((len(seeds), n), dtype=np.float32) if energy is not None else None
    P = [sample_params(s) for s in seeds]
    stats = np.zeros((len(seeds), 3))             # newton steps, CG iterations, converged
    t0 = time.perf_counter()
    errors = []

    def work(w):
        fs = fss[w]
        stream = torch.cuda.Stream()
        try:
            with torch.cuda.stream(stream):
                buf = torch.empty(6 * fs.asm._device_state()['num_el'], dtype=torch.float64,
                                  device=device()) if energy is not None else None
                kx = torch.empty(n, dtype=torch.float64, device=device()) \
                    if energy is not None else None
                for k in range(w, len(seeds), len(fss)):
                    x = fs.solve(P[k], rtol=rtol)
                    st = fs.solver.stats
                    stats[k] = (st['newton_steps'], sum(st['cg_iterations']), st['converged'])
                    X[k] = x.cpu().numpy()
                    if energy is not None:
                        energy.apply(x, kx, buf)
                        KX[k] = kx.cpu().numpy()
                    if log is not None and w == 0 and k % max(len(seeds) // 8, 1) < len(fss):
                        log('  snapshot %d/%d newton %d (%.1f s)'
                            % (k + 1, len(seeds), st['newton_steps'], time.perf_counter() - t0))
                stream.synchronize()
        except Exception as e:     # re-raised in the calling thread
            errors.append(e)

    threads = [threading.Thread(target=work, args=(w,)) for w in range(len(fss))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        raise errors[0]
    return X, P, KX, stats


class POD:
    """x ~ mean + U a with U [n, r] (method of snapshots).

    With ``KX`` (K0 x of every snapshot, EnergyInner) the modes are
    orthonormal in the K0 inner product and a = U^T K0 (x - mean): the
    decomposition and its coefficients see only B, not the gradient part of
    x, so no mode is spent on the solver's arbitrary gauge.  Without it, the
    Euclidean inner product."""

    def __init__(self, X, r, KX=None):
        X = np.asarray(X, dtype=np.float32)
        self.mean = X.astype(np.float64).mean(axis=0)
        Xc = X - self.mean.astype(np.float32)
        if KX is not None:
            kmean = np.asarray(KX, dtype=np.float64).mean(axis=0)
            Kc = np.asarray(KX, dtype=np.float32) - kmean.astype(np.float32)
            C = (Xc @ Kc.T).astype(np.float64)
            C = 0.5 * (C + C.T)
        else:
            Kc = None
            C = (Xc @ Xc.T).astype(np.float64)
        lam_all, V = np.linalg.eigh(C)
        order = np.argsort(lam_all)[::-1][:r]
        lam, V = lam_all[order], V[:, order]
        keep = lam > lam[0] * 1e-14
        lam, V = lam[keep], V[:, keep]
        s = np.sqrt(lam)
        self.U = (Xc.T @ V.astype(np.float32)) / s.astype(np.float32)      # [n, r]
        self.KU = (Kc.T @ V.astype(np.float32)) / s.astype(np.float32) if Kc is not None else None
        # deterministic signs: largest-magnitude entry of each mode positive
        sign = np.sign(self.U[np.abs(self.U).argmax(axis=0), np.arange(self.U.shape[1])])
        self.U *= sign
        if self.KU is not None:
            self.KU *= sign
        self.coeffs = (V * s * sign).astype(np.float64)                    # [n_snap, r]
        self.energy = lam / np.sum(np.clip(lam_all, 0, None))
        self.lam = lam

    def project(self, x):
        """Coefficients of a host vector x (in the POD's inner product)."""
        xc = np.asarray(x, dtype=np.float64) - self.mean
        B = self.KU if self.KU is not None else self.U
        return B.T.astype(np.float64) @ xc


def transfer_operator(src_mesh, dst_mesh):
    """(idx [n_dst_edges, 6], w [n_dst_edges, 6]): edge DOFs of dst from the
    Whitney field of src, DOF = A(midpoint) . (x_hi - x_lo)."""
    p0 = dst_mesh.nodes[dst_mesh.edges[:, 0]]
    p1 = dst_mesh.nodes[dst_mesh.edges[:, 1]]
    mid = 0.5 * (p0 + p1)
    tvec = p1 - p0
    el = locate_points(src_mesh, mid)
    ok = el >= 0
    el = np.where(ok, el, 0)
    X = src_mesh.nodes[src_mesh.cells[el]]                              # [m,4,3]
    Jm = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=2)
    Jinv = np.linalg.inv(Jm)
    lam = np.einsum('mij,mj->mi', Jinv, mid - X[:, 0])
    lam = np.concatenate([1.0 - lam.sum(axis=1, keepdims=True), lam], axis=1)   # [m,4]
    grad_hat = np.array([[-1.0, -1.0, -1.0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    G = np.einsum('ak,mkj->maj', grad_hat, Jinv)                       # grad lam_a = J^-T grad_hat
    o = src_mesh.orient[el]
    idx = np.empty((len(mid), 6), dtype=np.int32)
    w = np.empty((len(mid), 6))
    for j, (a, b) in enumerate(LOCAL_EDGES):
        wj = lam[:, a:a + 1] * G[:, b] - lam[:, b:b + 1] * G[:, a]
        sgn = np.where((o >> j) & 1, -1.0, 1.0)
        w[:, j] = sgn * np.einsum('mk,mk->m', wj, tvec)
        idx[:, j] = src_mesh.glob_ids[el, j]
    idx[~ok] = -1
    w[~ok] = 0.0
    return idx, w


class WarmStart:
    """Trained network + POD basis for one morph family."""

    def __init__(self, family, pod, mlp, coeff_std):
        self.family, self.pod, self.mlp, self.coeff_std = family, pod, mlp, coeff_std
        self._U = None
        self._W = None
        self.r_galerkin = 0
        self._transfer = {}

    @classmethod
    def train(cls, family, n_train=2048, r=256, seeds0=1000, epochs=3000, rtol=1e-8,
              hidden=(512, 512, 512, 512), log=None, fsolver=None, workers=4, energy=True,
              lr=2e-3):
        """Snapshots of ``n_train`` seeded variants (Newton to ``rtol``, on
        ``workers`` streams), the POD (in the K0 inner product with
        ``energy``), and the MLP trained on the normalised coefficients."""
        import torch
        from .mlp import MLP
        fs = fsolver or FamilySolver(family)
        fss = [fs] + [FamilySolver(family) for _ in range(max(workers, 1) - 1)]
        t0 = time.perf_counter()
        K0 = EnergyInner(fs) if energy else None
        X, P, KX, gst = generate(fss, range(seeds0, seeds0 + n_train), rtol=rtol, energy=K0,
                                 log=log)
        t_data = time.perf_counter() - t0
        r = min(r, n_train)
        pod = POD(X, r, KX)
        del X, KX
        t_pod = time.perf_counter() - t0 - t_data
        r = pod.U.shape[1]
        std = pod.coeffs.std(axis=0) + 1e-30
        T = np.ascontiguousarray(pod.coeffs / std, dtype=np.float32)
        Xin = normalise(P)
        mlp = MLP((8,) + tuple(hidden) + (r,), seed=0)
        hist = mlp.train(torch.from_numpy(Xin).to(device()), torch.from_numpy(T).to(device()),
                         epochs=epochs, lr=lr, log=log)
        torch.cuda.synchronize()
        ws = cls(family, pod, mlp, std)
        ws.energy_inner = K0
        ws.train_stats = {'t_data': t_data, 't_pod': t_pod,
                          't_train': time.perf_counter() - t0 - t_data - t_pod,
                          'loss_first': hist[0], 'loss_last': hist[-1], 'n_train': n_train,
                          'modes': r, 'pod_energy': float(pod.energy[:r].sum()),
                          'snapshot_newton_mean': float(gst[:, 0].mean()),
                          'snapshot_cg_mean': float(gst[:, 1].mean()),
                          'snapshots_converged': int(gst[:, 2].sum()), 'workers': len(fss),
                          'inner_product': 'K0' if energy else 'euclidean'}
        ws.fsolver = fs
        ws.fsolvers = fss
        return ws

    def coefficients(self, params_list):
        """POD coefficients [len(params_list), r] (host float64) predicted by
        the network in one batched forward pass."""
        import torch
        Xin = torch.from_numpy(normalise(params_list)).to(device())
        a = self.mlp.forward(Xin).cpu().numpy().astype(np.float64)
        return a * self.coeff_std

    def to_device(self, asm=None, r_galerkin=64):
        """Upload the modes and the mean (once), and, given an assembler of the
        family topology, the face circulations W of the first ``r_galerkin``
        modes for the reduced Newton step (sn_pod_element_modes)."""
        import torch
        if self._U is None:
            self._U = torch.from_numpy(np.ascontiguousarray(self.pod.U.T)).to(device())
            self._mean = as_dev_f64(self.pod.mean)
        sizes = [k for k in (32, 64, 128) if k <= min(self.pod.U.shape[1], r_galerkin)]
        rg = max(sizes) if sizes else 0
        if asm is not None and rg and self._W is None:
            d = asm._device_state()
            N, n = d['num_el'], asm.num_dofs
            self._W = torch.empty(N * 3 * rg, dtype=torch.float32, device=device())
            nat.check(nat.load().sn_pod_element_modes(N, n, rg, nat.ptr(self._U),
                                                      nat.ptr(d['gids']), nat.ptr(d['orient']),
                                                      nat.ptr(self._W), nat.stream_handle()))
            self.r_galerkin = rg
        return self

    def expand(self, a, out=None):
        """mean + U a on the family topology (device, float64)."""
        import torch
        self.to_device()
        c = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(device())
        n = self.pod.U.shape[0]
        if out is None:
            out = torch.empty(n, dtype=torch.float64, device=device())
        nat.check(nat.load().sn_pod_expand(n, int(c.numel()), nat.ptr(self._U), nat.ptr(c),
                                           nat.ptr(self._mean), nat.ptr(out),
                                           nat.stream_handle()))
        return out

    def expand_many(self, A, mean=True):
        """mean + U a for every row a of A ([V, r], V in 1, 2, 4, 8) in one pass
        over the modes: device float64 [V, n], row v bit-identical to
        expand(A[v]) (without the mean when ``mean`` is False).  A may be a
        host array or a float32 device tensor (then nothing waits)."""
        import torch
        self.to_device()
        if isinstance(A, torch.Tensor):
            c = A.to(device=device(), dtype=torch.float32).contiguous()
        else:
            A = np.ascontiguousarray(A, dtype=np.float32)
            c = torch.from_numpy(A).to(device())
        n = self.pod.U.shape[0]
        out = torch.empty((A.shape[0], n), dtype=torch.float64, device=device())
        nat.check(nat.load().sn_pod_expand_many(n, int(A.shape[1]), int(A.shape[0]),
                                                nat.ptr(self._U), nat.ptr(c),
                                                nat.ptr(self._mean) if mean else None,
                                                nat.ptr(out), nat.stream_handle()))
        return out

    def start_many(self, solvers, params, coeffs=None, each=None, lanes_per_row=8):
        """WarmStart.start for several solvers of the family topology at once:
        the expansions and the projections of all of them in one pass over
        the modes each (sn_pod_expand_many, sn_pod_project_many_dev), their
        reduced systems in one launch (sn_pod_reduced_solve_many); the
        residuals and reduced Jacobians per solver through ``each(fn,
        indices)``, which may run them on several streams (default: in turn).
        Each solver's driver must already hold its sources and |R(0)|.
        Returns [(x0 or None, info)]."""
        import torch
        V = len(solvers)
        NV = 1
        while NV < V:
            NV *= 2
        if NV > 8:
            raise ValueError('start_many: at most 8 solvers')
        each = each or (lambda fn, idx: [fn(v) for v in idx])
        a = self.coefficients(params) if coeffs is None else np.asarray(coeffs)
        A = np.zeros((NV, a.shape[1]))
        A[:V] = a
        drvs = [s.newton_driver('mixed', lanes_per_row, 32, 'ams') for s in solvers]
        X = self.expand_many(A)
        rg = self.r_galerkin
        lib = nat.load()
        n = X.shape[1]
        for _ in range(self.galerkin_steps if rg else 0):
            H = torch.zeros((NV, rg, rg), dtype=torch.float64, device=device())
            scratch = torch.empty(V, dtype=torch.float64, device=device())

            def jacobians(v):
                drvs[v].residual(X[v], scratch[v])
                self.reduced_jacobian_device(drvs[v], H[v])
            each(jacobians, range(V))
            R = torch.zeros((NV, n), dtype=torch.float64, device=device())
            for v in range(V):
                R[v].copy_(drvs[v].R)
            G = torch.zeros((NV, rg), dtype=torch.float64, device=device())
            nat.check(lib.sn_pod_project_many_dev(n, rg, NV, nat.ptr(self._U), nat.ptr(R),
                                                  nat.ptr(G), nat.stream_handle()))
            D = torch.zeros((NV, rg), dtype=torch.float32, device=device())
            nat.check(lib.sn_pod_reduced_solve_many(V, rg, nat.ptr(H), nat.ptr(G), nat.ptr(D),
                                                    nat.stream_handle()))
            X += self.expand_many(D, mean=False)
        # one tensor per start (rows of X would share one version counter,
        # which the judge's in-place masking bumps for all of them)
        xs = [X[v].clone() for v in range(V)]
        out = [None] * V

        def judge(v):
            r = drvs[v].residual_masked(xs[v]) / drvs[v].r0()
            use = bool(np.isfinite(r) and r < 1.0)
            out[v] = ((xs[v] if use else None), {'r_start_rel': float(r), 'used': use})
        each(judge, range(V))
        return out

    def predict_family(self, p):
        """Predicted edge DOFs on the family topology (device, float64)."""
        return self.expand(self.coefficients([p])[0])

    @staticmethod
    def _domain_ranges(asm):
        """Element ranges [e0, e1) of the device element order, split into
        those whose element matrices depend on the state (nonlinear or
        diluted reluctivity) and those that do not (constant reluctivity):
        ([(e0, e1), ...], [(n, e0, e1), ...])."""
        from .materials import ConstantReluctance
        offs = asm._device_state()['offsets']
        var, const = [], []
        for n in range(min(len(asm.material_list), len(offs) - 1)):
            e0, e1 = int(offs[n]), int(offs[n + 1])
            if e1 <= e0:
                continue
            if isinstance(asm.material_list[n], ConstantReluctance) and asm._dilution(n) is None:
                const.append((n, e0, e1))
            else:
                var.append((e0, e1))
        return var, const

    def _range_H(self, drv, e0, e1):
        H = np.empty(self.r_galerkin ** 2)
        nat.check(nat.load().sn_pod_reduced_jacobian_range(
            drv.d['num_el'], e0, e1, self.r_galerkin, nat.ptr(self._W), nat.ptr(drv.Mr),
            H.ctypes.data_as(nat._vp), nat.stream_handle()))
        return H

    def reduced_jacobian(self, drv):
        """H = U^T J U over the first r_galerkin modes, J the element
        Jacobian of drv's last residual() call.  The domains of constant
        reluctivity contribute a state-independent term, summed once per
        geometry of the assembler (asm.geometry_version, bumped where node
        coordinates are changed in place) and kept with the assembler; the
        other domains are summed on every call."""
        import weakref
        asm = drv.asm
        var, const = self._domain_ranges(asm)
        H = np.zeros(self.r_galerkin ** 2)
        for e0, e1 in var:
            H += self._range_H(drv, e0, e1)
        if const:
            key = (getattr(asm, 'geometry_version', 0), self.r_galerkin,
                   tuple((e0, e1, id(asm.material_list[n])) for n, e0, e1 in const))
            cache = asm.__dict__.setdefault('_galerkin_const', weakref.WeakKeyDictionary())
            hit = cache.get(self)
            if hit is None or hit[0] != key:
                Hc = np.zeros(self.r_galerkin ** 2)
                for _, e0, e1 in const:
                    Hc += self._range_H(drv, e0, e1)
                hit = (key, Hc)
                cache[self] = hit
            H += hit[1]
        return H.reshape(self.r_galerkin, self.r_galerkin)

    def reduced_jacobian_device(self, drv, out):
        """reduced_jacobian(drv) queued on the current stream into the
        float64 device tensor ``out`` [r_galerkin, r_galerkin]: the same
        sums, in the same order, without a host wait (the constant domains'
        term is the host cache's, uploaded once per geometry; forming that
        cache, the first time, waits)."""
        import torch
        import weakref
        asm = drv.asm
        var, const = self._domain_ranges(asm)
        lib = nat.load()
        out.zero_()
        if var:
            tmp = torch.empty_like(out)
            for e0, e1 in var:
                nat.check(lib.sn_pod_reduced_jacobian_range_dev(
                    drv.d['num_el'], e0, e1, self.r_galerkin, nat.ptr(self._W), nat.ptr(drv.Mr),
                    nat.ptr(tmp), nat.stream_handle()))
                out += tmp
        if const:
            dev = asm.__dict__.setdefault('_galerkin_const_dev', weakref.WeakKeyDictionary())
            cache = asm.__dict__.get('_galerkin_const')
            hit = cache.get(self) if cache is not None else None
            key = (getattr(asm, 'geometry_version', 0), self.r_galerkin,
                   tuple((e0, e1, id(asm.material_list[n])) for n, e0, e1 in const))
            if hit is None or hit[0] != key:
                # the host form computes (and caches) the constant term; the
                # state-dependent part it sums as well is dropped here
                self.reduced_jacobian(drv)
                hit = asm.__dict__['_galerkin_const'][self]
            d = dev.get(self)
            if d is None or d[0] != key:
                d = (key, torch.from_numpy(hit[1].reshape(out.shape)).to(out.device))
                dev[self] = d
            out += d[1]
        return out

    def precompute_constant(self, solver, lanes_per_row=8):
        """Form the state-independent part of the reduced Jacobian for
        ``solver``'s geometry now (set-up), so that no start pays for it."""
        import torch
        self.to_device(solver.curl_curl_factory)
        if not self.r_galerkin:
            return self
        drv = solver.newton_driver('mixed', lanes_per_row, 32, 'ams')
        drv.retarget(solver.source_fields_device())
        drv.residual(torch.zeros(solver.curl_curl_factory.num_dofs, dtype=torch.float64,
                                 device=device()))
        self.reduced_jacobian(drv)
        return self

    def galerkin(self, drv, x, steps=1):
        """Reduced Newton steps in the span of the first r_galerkin modes:
        solve (U^T J U) da = -U^T R(x) and set x += U da, with J and R at x
        from the mixed driver ``drv`` (element-operator Jacobian).  x is
        updated in place and returned.  Everything stays on the device and
        nothing waits (the reduced system is solved by Cholesky,
        sn_pod_reduced_solve_many; one that is not numerically positive
        definite gets no correction)."""
        import torch
        lib = nat.load()
        rg = self.r_galerkin
        n = drv.asm.num_dofs
        dev = x.device
        H = torch.empty((rg, rg), dtype=torch.float64, device=dev)
        g = torch.empty(rg, dtype=torch.float64, device=dev)
        da = torch.empty(rg, dtype=torch.float32, device=dev)
        norm2 = torch.empty(1, dtype=torch.float64, device=dev)
        for _ in range(steps):
            drv.residual(x, norm2[0])
            self.reduced_jacobian_device(drv, H)
            nat.check(lib.sn_pod_project_many_dev(n, rg, 1, nat.ptr(self._U), nat.ptr(drv.R),
                                                  nat.ptr(g), nat.stream_handle()))
            nat.check(lib.sn_pod_reduced_solve_many(1, rg, nat.ptr(H), nat.ptr(g), nat.ptr(da),
                                                    nat.stream_handle()))
            nat.check(lib.sn_pod_expand(n, rg, nat.ptr(self._U), nat.ptr(da), nat.ptr(x),
                                        nat.ptr(x), nat.stream_handle()))
        return x

    galerkin_steps = 1

    def start(self, solver, p, coeffs=None, galerkin_steps=None, lanes_per_row=8):
        """Start vector for ``solver`` (a solver on the family topology whose
        current variant has parameters p): the network's prediction, one
        reduced Newton step, and the residual judge -- the start is kept only
        if ||R(x0)|| < ||R(0)||.  Returns (x0 or None, info)."""
        a = self.coefficients([p])[0] if coeffs is None else coeffs
        drv = solver.newton_driver('mixed', lanes_per_row, 32, 'ams')
        drv.retarget(solver.source_fields_device())
        self.to_device(solver.curl_curl_factory)
        x = self.expand(a)
        if galerkin_steps is None:
            galerkin_steps = self.galerkin_steps
        if galerkin_steps and self.r_galerkin:
            self.galerkin(drv, x, galerkin_steps)
        r0 = drv.r0()              # before the judged residual, which Newton then reuses
        r = drv.residual_masked(x) / r0
        use = bool(np.isfinite(r) and r < 1.0)
        return (x if use else None), {'r_start_rel': float(r), 'used': use}

    def prepare(self, dp):
        """Host set-up of the transfer onto dipole ``dp``'s mesh (once per mesh)."""
        import copy
        key = id(dp.mesh)
        if key not in self._transfer:
            p = {k: dp.p[k] for k in PARAM_NAMES}
            src = copy.copy(self.family.mesh)
            src.nodes = self.family.nodes(p)
            from .mesh import StructuredGrid
            src.grid = StructuredGrid(*self.family.axes(p))
            idx, w = transfer_operator(src, dp.mesh)
            self._transfer[key] = (as_dev_i32(idx.reshape(-1)), as_dev_f64(w.reshape(-1)), p)
        return self._transfer[key]

    def predict(self, dp):
        """Initial edge DOFs (device float64) for dipole ``dp`` on its own mesh."""
        import torch
        idx, w, p = self.prepare(dp)
        xc = self.predict_family(p)
        nf = dp.mesh.num_edges
        out = torch.empty(nf, dtype=torch.float64, device=device())
        nat.check(nat.load().sn_edge_transfer(nf, nat.ptr(idx), nat.ptr(w), nat.ptr(xc),
                                              nat.ptr(out), nat.stream_handle()))
        return out

    # ------------------------------------------------------------ persistence
    def save(self, path):
        st = self.mlp.state()
        np.savez(path, counts=np.array(self.family.counts), mean=self.pod.mean, U=self.pod.U,
                 coeff_std=self.coeff_std, based=np.array(self.family.base is not None),
                 **{'mlp_' + k: v for k, v in st.items()})

    @classmethod
    def load(cls, path, family=None):
        """A saved model.  A model of a family built on a dipole's own mesh
        (MorphFamily(base=...)) needs that family passed back in; the mode
        count must match the family's edge count either way."""
        from .mlp import MLP
        z = np.load(path)
        if family is None and 'based' in z.files and bool(z['based']):
            raise ValueError('WarmStart.load: this model was trained on MorphFamily(base=...); '
                             'pass family=MorphFamily(base=<the same dipole>)')
        fam = family or MorphFamily(tuple(int(c) for c in z['counts']))
        if z['U'].shape[0] != fam.mesh.num_edges:
            raise ValueError('WarmStart.load: the model has %d edge DOFs, the family mesh %d'
                             % (z['U'].shape[0], fam.mesh.num_edges))
        pod = POD.__new__(POD)
        pod.mean, pod.U = z['mean'], z['U']
        mlp = MLP.from_state({k[4:]: z[k] for k in z.files if k.startswith('mlp_')})
        return cls(fam, pod, mlp, z['coeff_std'])


def default_path():
    return os.path.join(MODEL_DIR, 'warm_start_dipole.npz')


def load_default(dp=None):
    """The cached model of the default family, or an error if none is cached."""
    path = default_path()
    if not os.path.exists(path):
        raise FileNotFoundError('no trained warm-start model at ' + path)
    return WarmStart.load(path)


def solve_with_warm_start(solver, ws, p, rtol=1e-8, log=None, galerkin_steps=None, coeffs=None,
                          lanes_per_row=8):
    """Newton (mixed precision, auxiliary-space preconditioner) from the
    learned start of ``ws`` for the variant with parameters p on the family
    topology, or from A = 0 when the start's residual is not below the
    residual at A = 0 (plain Newton is the fallback)."""
    x0, info = ws.start(solver, p, coeffs=coeffs, galerkin_steps=galerkin_steps,
                        lanes_per_row=lanes_per_row)
    x = solver.solve_newton('mixed', rtol=rtol, x0=x0, log=log, precond='ams',
                            lanes_per_row=lanes_per_row)
    solver.stats['warm_start'] = info
    return x
