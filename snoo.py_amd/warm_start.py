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


class MorphFamily:
    """Fixed-topology dipole meshes over the warm-start parameter ranges.

    counts = (n_pad, n_wall_x, n_ap_x, n_wall_y, n_ap_y, n_len).  The default
    gives 205 632 tetrahedra."""

    def __init__(self, counts=(7, 6, 10, 6, 8, 14), pad=0.5, ratio=1.3):
        self.counts = tuple(int(c) for c in counts)
        self.pad, self.ratio = pad, ratio
        base = self.axes(dict(DEFAULT))
        p0 = dict(DEFAULT)

        def tag_fn(xc, yc, zc):
            in_box = (np.abs(xc) < p0['W'] / 2) & (np.abs(yc) < p0['H'] / 2) & \
                     (np.abs(zc) < p0['L'] / 2)
            in_ap = (np.abs(xc) < p0['ax'] / 2) & (np.abs(yc) < p0['ay'] / 2)
            return np.where(in_box & ~in_ap, IRON_TAG, AIR_TAG)
        self.mesh = kuhn_mesh(*base, tag_fn)
        self.model = TetModel(self.mesh)

    def axes(self, p):
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

    def variant(self, p):
        """Dipole object of parameters p on this topology (coils and reluctivity)."""
        full = dict(DEFAULT)
        full.update(p)
        d = Dipole.__new__(Dipole)
        d.p = full
        d.h = None
        d.yokes = [dict(W=full['W'], H=full['H'], L=full['L'], ax=full['ax'], ay=full['ay'],
                        z0=0.0, sign=1.0)]
        d.coils = d._coils_for(d.yokes[0])
        d.reluctance = tanh_reluctance(full['B_sat'], full['mu_r'])
        d.mesh = self.mesh
        d.model = self.model
        return d


class FamilySolver:
    """One RedMVPSolver on the family topology, re-targeted per variant by
    replacing node coordinates, coils and reluctivity on the device."""

    def __init__(self, family, quad_order=2):
        self.family = family
        d0 = family.variant(dict(DEFAULT))
        self.solver = d0.solver(quad_order=quad_order, verbose=False)
        self.asm = self.solver.curl_curl_factory
        self.asm._device_state()

    def set_variant(self, p):
        import torch
        d = self.family.variant(p)
        nodes = self.family.nodes(d.p)
        self.asm.tet_mesh.nodes[:] = nodes
        self.asm._dev['nodes'].copy_(torch.from_numpy(nodes.reshape(-1)).to(device()))
        self.solver.coil_list = d.coils
        self.solver.curl_curl_factory.material_list[0] = d.reluctance
        self.solver._src = None
        return d

    def solve(self, p, rtol=1e-6, precision='mixed', x0=None, precond='ams'):
        self.set_variant(p)
        return self.solver.solve_newton(precision, rtol=rtol, x0=x0, precond=precond)


def generate(family_solver, seeds, rtol=1e-6, log=None):
    """Converged edge DOFs (float64, host) of the variants with the given seeds."""
    X, P = [], []
    t0 = time.perf_counter()
    for k, s in enumerate(seeds):
        p = sample_params(s)
        x = family_solver.solve(p, rtol=rtol)
        X.append(x.cpu().numpy())
        P.append(p)
        if log is not None and (k % max(len(seeds) // 8, 1) == 0):
            log('  snapshot %d/%d seed %d newton %d (%.1f s)' % (
                k + 1, len(seeds), s, family_solver.solver.stats['newton_steps'],
                time.perf_counter() - t0))
    return np.stack(X), P


class POD:
    """x ~ mean + U a, U orthonormal columns (method of snapshots)."""

    def __init__(self, X, r):
        X = np.asarray(X, dtype=np.float64)
        self.mean = X.mean(axis=0)
        Xc = (X - self.mean).astype(np.float32)
        C = (Xc @ Xc.T).astype(np.float64)
        lam_all, V = np.linalg.eigh(C)
        order = np.argsort(lam_all)[::-1][:r]
        lam, V = np.clip(lam_all[order], 1e-300, None), V[:, order]
        s = np.sqrt(lam)
        self.U = (Xc.T @ V.astype(np.float32)) / s.astype(np.float32)      # [n, r]
        # deterministic signs: largest-magnitude entry of each mode positive
        sign = np.sign(self.U[np.abs(self.U).argmax(axis=0), np.arange(self.U.shape[1])])
        self.U *= sign
        self.coeffs = (V * s * sign).astype(np.float64)                    # [n_snap, r]
        self.energy = lam / np.sum(np.clip(lam_all, 0, None))


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
        self._transfer = {}

    @classmethod
    def train(cls, family, n_train=2048, r=256, seeds0=1000, epochs=3000, rtol=1e-6,
              hidden=(512, 512, 512, 512), log=None, fsolver=None):
        import torch
        from .mlp import MLP
        fs = fsolver or FamilySolver(family)
        t0 = time.perf_counter()
        X, P = generate(fs, range(seeds0, seeds0 + n_train), rtol=rtol, log=log)
        t_data = time.perf_counter() - t0
        r = min(r, n_train)
        pod = POD(X, r)
        t_pod = time.perf_counter() - t0 - t_data
        std = pod.coeffs.std(axis=0) + 1e-30
        T = (pod.coeffs / std).astype(np.float32)
        Xin = normalise(P)
        mlp = MLP((8,) + tuple(hidden) + (r,), seed=0)
        hist = mlp.train(torch.from_numpy(Xin).to(device()), torch.from_numpy(T).to(device()),
                         epochs=epochs, lr=2e-3, log=log)
        torch.cuda.synchronize()
        ws = cls(family, pod, mlp, std)
        ws.train_stats = {'t_data': t_data, 't_pod': t_pod,
                          't_train': time.perf_counter() - t0 - t_data - t_pod,
                          'loss_first': hist[0], 'loss_last': hist[-1], 'n_train': n_train,
                          'modes': r, 'pod_energy': float(pod.energy[:r].sum())}
        ws.fsolver = fs
        return ws

    def coefficients(self, params_list):
        import torch
        Xin = torch.from_numpy(normalise(params_list)).to(device())
        a = self.mlp.forward(Xin).cpu().numpy().astype(np.float64)
        return a * self.coeff_std

    def predict_family(self, p):
        """Predicted edge DOFs on the family topology (device, float64)."""
        import torch
        if self._U is None:
            self._U = torch.from_numpy(np.ascontiguousarray(self.pod.U.T)).to(device())
            self._mean = as_dev_f64(self.pod.mean)
        c = torch.from_numpy(self.coefficients([p])[0].astype(np.float32)).to(device())
        n = self.pod.U.shape[0]
        out = torch.empty(n, dtype=torch.float64, device=device())
        nat.check(nat.load().sn_pod_expand(n, int(c.numel()), nat.ptr(self._U), nat.ptr(c),
                                           nat.ptr(self._mean), nat.ptr(out),
                                           nat.stream_handle()))
        return out

    def prepare(self, dp):
        """Host set-up of the transfer onto dipole ``dp``'s mesh (once per mesh)."""
        import copy
        key = id(dp.mesh)
        if key not in self._transfer:
            p = {k: dp.p[k] for k in PARAM_NAMES}
            src = copy.copy(self.family.mesh)
            src.nodes = self.family.nodes(dict(DEFAULT, **p))
            from .mesh import StructuredGrid
            src.grid = StructuredGrid(*self.family.axes(dict(DEFAULT, **p)))
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
                 coeff_std=self.coeff_std, **{'mlp_' + k: v for k, v in st.items()})

    @classmethod
    def load(cls, path, family=None):
        from .mlp import MLP
        z = np.load(path)
        fam = family or MorphFamily(tuple(int(c) for c in z['counts']))
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


def solve_with_warm_start(solver, ws, dp, precision='mixed', rtol=1e-8, log=None):
    """Newton from the network's prediction, or from A = 0 when the
    prediction's residual is not below the residual at A = 0 (fallback)."""
    x0 = ws.predict(dp)
    drv = solver.newton_driver(precision)
    drv.retarget(solver.source_fields_device())
    r_pred = drv.residual(x0)
    use = r_pred < drv.r0()
    solver.stats_warm = {'r_pred_rel': r_pred / drv.r0(), 'used': bool(use)}
    return solver.solve_newton(precision, rtol=rtol, x0=x0 if use else None, log=log)
