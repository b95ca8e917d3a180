"""Synthetic window-frame dipoles: geometry, mesh, coils, probes, variants.

A box yoke W x H x L (x, y, z) centred at the origin with a rectangular
aperture ax x ay through z; one racetrack coil next to each pole face inside
the aperture (square cross-section ``coil_side``, corner radius
``corner_radius``), both carrying the same current so the aperture field is
along +y; an air box padded ``pad`` beyond the yoke, with the Dirichlet
condition on its outer surface.  Each coil is ``levels`` filaments stacked in
y (snoopy/racetrack_coil.py stacks one turn per y level), each carrying
J * coil_side^2 / levels, with core radius ``filament_radius``.

The mesh is a tensor grid: spacing h inside the yoke's bounding box (the yoke
and aperture faces are grid lines) and cells graded by ``ratio`` through the
air pad; every hexahedron is split into 6 Kuhn tetrahedra.
"""
import math

import numpy as np

from .materials import tanh_reluctance
from .mesh import AIR_TAG, IRON_TAG, TetModel, axis_lines, kuhn_mesh, refine_axis
from .racetrack_coil import RacetrackCoil

DEFAULT = dict(W=1.0, H=0.6, L=0.5, ax=0.3, ay=0.2, pad=0.5, J=1.5e6, coil_side=0.05,
               corner_radius=0.05, coil_gap=0.025, levels=5, filament_radius=0.005,
               B_sat=1.9, mu_r=4000.0, z0=0.0)


class Dipole:
    """Parameters, mesh and sources of one dipole (or a chain of yokes)."""

    def __init__(self, h=0.05, ratio=1.2, refine=0, **params):
        p = dict(DEFAULT)
        p.update(params)
        self.p = p
        self.h = h
        self.yokes = p.get('yokes') or [dict(W=p['W'], H=p['H'], L=p['L'], ax=p['ax'],
                                             ay=p['ay'], z0=p['z0'], sign=1.0)]
        self._build_mesh(h, ratio, refine)
        self.coils = []
        for yk in self.yokes:
            self.coils += self._coils_for(yk)
        self.reluctance = tanh_reluctance(p['B_sat'], p['mu_r'])

    # ------------------------------------------------------------------ mesh
    def _build_mesh(self, h, ratio, refine):
        ys = self.yokes
        xb = sorted({v for yk in ys for v in (-yk['W'] / 2, -yk['ax'] / 2, yk['ax'] / 2, yk['W'] / 2)})
        yb = sorted({v for yk in ys for v in (-yk['H'] / 2, -yk['ay'] / 2, yk['ay'] / 2, yk['H'] / 2)})
        zb = sorted({v for yk in ys for v in (yk['z0'] - yk['L'] / 2, yk['z0'] + yk['L'] / 2)})
        pad = self.p['pad']
        xs = axis_lines(xb, h, pad, ratio)
        yy = axis_lines(yb, h, pad, ratio)
        zs = axis_lines(zb, h, pad, ratio)
        for _ in range(refine):
            xs, yy, zs = refine_axis(xs), refine_axis(yy), refine_axis(zs)
        self.axes = (xs, yy, zs)

        def tag_fn(xc, yc, zc):
            iron = np.zeros(xc.shape, dtype=bool)
            for yk in ys:
                in_box = ((np.abs(xc) < yk['W'] / 2) & (np.abs(yc) < yk['H'] / 2) &
                          (np.abs(zc - yk['z0']) < yk['L'] / 2))
                in_ap = (np.abs(xc) < yk['ax'] / 2) & (np.abs(yc) < yk['ay'] / 2)
                iron |= in_box & ~in_ap
            return np.where(iron, IRON_TAG, AIR_TAG)

        self.mesh = kuhn_mesh(xs, yy, zs, tag_fn)
        self.model = TetModel(self.mesh)

    def _coils_for(self, yk):
        p = self.p
        c = p['coil_side']
        a = yk['ax'] / 2 - c / 2                       # x of the long sides
        b = yk['L'] / 2 + p['coil_gap'] + c / 2       # z of the end runs
        r = min(p['corner_radius'], 0.8 * a, 0.8 * b)
        z0 = yk['z0']
        kp = np.array([[-a, z0 - b + r], [-a, z0 + b - r], [-a + r, z0 + b], [a - r, z0 + b],
                       [a, z0 + b - r], [a, z0 - b + r], [a - r, z0 - b], [-a + r, z0 - b]])
        # reorder so the first straight run kp7 -> kp0 goes along +z at x = -a:
        # the loop then runs clockwise in the (x, z) plane and B_y > 0 inside
        kp = np.roll(kp, -1, axis=0)
        n = p['levels']
        I_fil = yk.get('sign', 1.0) * p['J'] * c * c / n
        coils = []
        for yc in (yk['ay'] / 2 - c / 2, -yk['ay'] / 2 + c / 2):
            y = yc + (np.arange(n) - (n - 1) / 2) * (c / n)
            coils.append(RacetrackCoil(kp, y, p['filament_radius'], I_fil))
        return coils

    # ------------------------------------------------------------------ helpers
    @property
    def num_tets(self):
        return self.mesh.num_elements

    def iron_fraction(self):
        return float(np.mean(self.mesh.tags == IRON_TAG))

    def solver(self, quad_order=2, reluctance=None, **kw):
        from .red_mvp_solver import RedMVPSolver
        rel = reluctance if reluctance is not None else self.reluctance
        return RedMVPSolver(self.model, self.coils, [IRON_TAG], [AIR_TAG], [rel],
                            quad_order=quad_order, **kw)

    def probes(self, n=10000, seed=2):
        """Uniform random points in the aperture box of the first yoke."""
        yk = self.yokes[0]
        rng = np.random.default_rng(seed)
        u = rng.uniform(-1.0, 1.0, size=(n, 3))
        return u * np.array([yk['ax'] / 2, yk['ay'] / 2, yk['L'] / 2]) + np.array([0, 0, yk['z0']])

    def beam_axis(self, n=5000):
        z = [yk['z0'] for yk in self.yokes]
        L = [yk['L'] for yk in self.yokes]
        zz = np.linspace(min(z) - max(L) / 2, max(z) + max(L) / 2, n)
        return np.stack([np.zeros(n), np.zeros(n), zz], axis=1)


def mesh_size_for(target_tets, ratio=1.2, **params):
    """Grid spacing h whose dipole mesh has about ``target_tets`` tetrahedra."""
    from .mesh import axis_lines as _al
    p = dict(DEFAULT)
    p.update(params)

    def count(h):
        xb = [-p['W'] / 2, -p['ax'] / 2, p['ax'] / 2, p['W'] / 2]
        yb = [-p['H'] / 2, -p['ay'] / 2, p['ay'] / 2, p['H'] / 2]
        zb = [p['z0'] - p['L'] / 2, p['z0'] + p['L'] / 2]
        n = [(len(_al(b, h, p['pad'], ratio)) - 1) for b in (xb, yb, zb)]
        return 6 * n[0] * n[1] * n[2]
    lo, hi = 1e-3, 1.0
    for _ in range(60):
        mid = math.sqrt(lo * hi)
        if count(mid) > target_tets:
            lo = mid
        else:
            hi = mid
    return hi, count(hi)


def sample_variant(seed, with_warm_start_inputs=False):
    """Seed-drawn dipole parameters (uniform ranges of the batched workload)."""
    rng = np.random.default_rng(seed)
    v = dict(W=rng.uniform(0.6, 1.4), H=rng.uniform(0.4, 0.9), ax=rng.uniform(0.15, 0.4),
             ay=rng.uniform(0.1, 0.3), J=rng.uniform(0.5e6, 3e6), B_sat=rng.uniform(1.7, 2.1))
    if with_warm_start_inputs:
        v['mu_r'] = rng.uniform(2000.0, 6000.0)
        v['L'] = rng.uniform(0.3, 0.8)
    # keep a yoke wall of at least 0.1 m around the aperture
    v['ax'] = min(v['ax'], v['W'] - 0.2)
    v['ay'] = min(v['ay'], v['H'] - 0.2)
    return v


def chain_yokes(n=6, seed=7, gap=0.1, L=0.5):
    """n yokes along z with alternating polarity, ``gap`` apart, dimensions
    drawn from the batched-workload ranges (W, H, ax, ay)."""
    out = []
    z = 0.0
    for k in range(n):
        v = sample_variant(seed * 1000 + k)
        out.append(dict(W=v['W'], H=v['H'], L=L, ax=v['ax'], ay=v['ay'], z0=z,
                        sign=1.0 if k % 2 == 0 else -1.0))
        z += L + gap
    zc = 0.5 * (out[0]['z0'] + out[-1]['z0'])
    for y in out:
        y['z0'] -= zc
    return out


def chain_mesh_size_for(target_tets, yokes, pad=0.5, ratio=1.2):
    """Grid spacing h whose chain mesh has about ``target_tets`` tetrahedra."""
    xb = sorted({v for yk in yokes for v in (-yk['W'] / 2, -yk['ax'] / 2, yk['ax'] / 2, yk['W'] / 2)})
    yb = sorted({v for yk in yokes for v in (-yk['H'] / 2, -yk['ay'] / 2, yk['ay'] / 2, yk['H'] / 2)})
    zb = sorted({v for yk in yokes for v in (yk['z0'] - yk['L'] / 2, yk['z0'] + yk['L'] / 2)})

    def count(h):
        n = [len(axis_lines(b, h, pad, ratio)) - 1 for b in (xb, yb, zb)]
        return 6 * n[0] * n[1] * n[2]
    lo, hi = 1e-3, 1.0
    for _ in range(50):
        mid = math.sqrt(lo * hi)
        if count(mid) > target_tets:
            lo = mid
        else:
            hi = mid
    return hi, count(hi)


def batch_variants(n=64, seed=12345, tets_lo=2e4, tets_hi=4e5):
    """The batched workload: dipoles with parameters ``sample_variant(k)``,
    k = 0..n-1, each meshed to a tetrahedron count drawn log-uniformly from
    [tets_lo, tets_hi] with generator ``seed``."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        v = sample_variant(k)
        target = float(np.exp(rng.uniform(np.log(tets_lo), np.log(tets_hi))))
        h, _ = mesh_size_for(target, **v)
        out.append(Dipole(h=h, **v))
    return out


def chain_dipole(target_tets=6.0e6, n=6, seed=7):
    """The multi-magnet chain: ``chain_yokes(n, seed)`` in one air box,
    meshed to about ``target_tets`` tetrahedra."""
    yokes = chain_yokes(n, seed=seed)
    h, _ = chain_mesh_size_for(target_tets, yokes)
    return Dipole(h=h, yokes=yokes)
