"""SHiP magnet templates meshed by this build's structured mesher.

The reference builds the SHiP yokes as gmsh OCC volumes
(snoopy/geometry_tools.py:133-1119) and meshes them with gmsh.  Every one of
those volumes is a ruled solid: its cross-section in the (x, y) plane at the
entrance z = Z_pos and at the exit z = Z_pos + Z_len is a union of
axis-aligned rectangles whose corners are the template's break coordinates
(X_core, Y_yoke, ...), and the solid joins the two sections by straight lines
in z, so every break moves linearly in z.  Only the quarter x >= 0, y >= 0 is
meshed, inside the air box [0, lim_x] x [0, lim_y] x [z_min, z_max].

``ruled_quarter_mesh`` meshes such a template with a bent tensor grid: the
grid lines of x (and y) run through the breaks at every z, each interval
between two breaks is split into a fixed number of cells, and the cells of a
hexahedron are split into 6 Kuhn tetrahedra.  Faces of the solid are then
faces of tetrahedra, so volumes are exact and the iron surface is resolved
without staircasing.  The Dirichlet condition sits on x = 0, x = lim_x,
y = lim_y, z = z_min and z = z_max; y = 0 keeps the natural condition (the
reference's boundary selection, snoopy/SHiP_functions.py:536-566).

Each template function returns the mesh model and the volume tags in the
order in which the reference's driver names its gmsh domains.
"""
import math

import numpy as np

from .mesh import TetModel, _graded, mapped_kuhn_mesh

SHIP_SIDES = ('x0', 'x1', 'y1', 'z0', 'z1')


def compute_area_polygon(kp, isclosed=True):
    """Shoelace area of a polygon [M, 2]; ``isclosed`` means the first and last
    key points coincide (snoopy/geometry_tools.py:1121-1145)."""
    kp = np.asarray(kp, dtype=np.float64)
    area = 0.0
    for i in range(kp.shape[0] - 1):
        area += kp[i, 0] * kp[i + 1, 1] - kp[i + 1, 0] * kp[i, 1]
    if not isclosed:
        area += kp[-1, 0] * kp[0, 1] - kp[0, 0] * kp[-1, 1]
    return 0.5 * area


def _uniform_fracs(width, h):
    n = max(1, int(math.ceil(width / h - 1e-9)))
    return np.arange(n) / n


def _graded_fracs(width, h, ratio):
    """Cell boundaries (as fractions, without the last) growing from h."""
    if width <= 0:
        return np.zeros(1)
    g = _graded(0.0, width, min(h, width), ratio, small_at_a=True)
    return g[:-1] / width


def _sorted_breaks(breaks, tol):
    """Sort named breaks {name: values at the z-knots}; merge breaks that
    coincide at every knot; reject breaks that cross or that coincide at some
    knots only."""
    names = sorted(breaks, key=lambda k: tuple(breaks[k]))
    vals, index = [], {}
    for nm in names:
        v = np.asarray(breaks[nm], dtype=np.float64)
        if vals:
            d = v - vals[-1]
            if np.all(np.abs(d) <= tol):
                index[nm] = len(vals) - 1
                continue
            if np.any(d < -tol):
                raise ValueError('break %r crosses its neighbour (values %s)' % (nm, v.tolist()))
            if np.any(np.abs(d) <= tol):
                raise ValueError('break %r meets its neighbour at one end only; the structured '
                                 'mesher needs breaks that are distinct everywhere or equal '
                                 'everywhere' % (nm,))
        index[nm] = len(vals)
        vals.append(v)
    return np.asarray(vals, dtype=np.float64), index


def _axis(breaks, h, ratio):
    """Cell boundaries of one axis as (interval, fraction) pairs: uniform
    cells of size <= h in every interval but the last, graded cells growing
    from h in the last (the air pad up to the box)."""
    vals, _ = breaks
    n_int = len(vals) - 1
    iv, fr = [], []
    for m in range(n_int):
        w = float(np.max(vals[m + 1] - vals[m]))
        f = _graded_fracs(w, h, ratio) if m == n_int - 1 else _uniform_fracs(w, h)
        iv.append(np.full(len(f), m))
        fr.append(f)
    iv.append([n_int - 1])
    fr.append([1.0])
    return np.concatenate(iv).astype(np.int64), np.concatenate(fr)


def _line_positions(vals, iv, fr, zk, z):
    """Positions of the grid lines at height z: every break piecewise linear
    between the knots zk and constant beyond them."""
    if len(zk) == 1:
        b = vals[:, 0]
    else:
        j = int(np.clip(np.searchsorted(zk, z, side='right') - 1, 0, len(zk) - 2))
        t = min(max((z - zk[j]) / (zk[j + 1] - zk[j]), 0.0), 1.0)
        b = (1.0 - t) * vals[:, j] + t * vals[:, j + 1]
    return b[iv] + fr * (b[np.minimum(iv + 1, len(b) - 1)] - b[iv])


def ruled_knot_mesh(xbreaks, ybreaks, zk, lim_x, lim_y, pad_lo, pad_hi, regions, air_tag, h,
                    hz=None, ratio=1.3):
    """Mesh ruled solids stacked along z in the quarter box.

    zk                z-knots (ascending): the ends of every solid
    xbreaks, ybreaks  {name: values at the knots}; the names '0' and 'lim'
                      (0 and lim_x / lim_y) are added
    pad_lo, pad_hi    air below the first knot and above the last
    regions           [(tag, [(x_lo, x_hi, y_lo, y_hi), ...], (z_lo, z_hi)), ...]:
                      a cell between z_lo and z_hi inside one of the rectangles
                      (named breaks) takes the tag; the first region that holds
                      it wins; every other cell is air_tag
    h, hz             cell size across and along z (hz defaults to h)
    ratio             growth of the graded cells in the air pads
    """
    hz = h if hz is None else hz
    zk = [float(z) for z in zk]
    nk = len(zk)
    tol = 1e-12 * max(lim_x, lim_y, 1.0)
    xb = dict(xbreaks, **{'0': [0.0] * nk, 'lim': [lim_x] * nk})
    yb = dict(ybreaks, **{'0': [0.0] * nk, 'lim': [lim_y] * nk})
    xs, ys = _sorted_breaks(xb, tol), _sorted_breaks(yb, tol)
    xiv, xfr = _axis(xs, h, ratio)
    yiv, yfr = _axis(ys, h, ratio)
    parts = [_graded(zk[0] - pad_lo, zk[0], hz, ratio, small_at_a=False)[:-1]]
    for a, b in zip(zk[:-1], zk[1:]):
        if b - a > tol:
            parts.append(np.linspace(a, b, max(1, int(math.ceil((b - a) / hz - 1e-9))) + 1)[:-1])
    parts.append(_graded(zk[-1], zk[-1] + pad_hi, hz, ratio, small_at_a=True))
    zs = np.concatenate(parts)
    nx, ny, nz = len(xfr) - 1, len(yfr) - 1, len(zs) - 1
    X = np.stack([_line_positions(xs[0], xiv, xfr, zk, z) for z in zs])     # [nz+1, nx+1]
    Y = np.stack([_line_positions(ys[0], yiv, yfr, zk, z) for z in zs])     # [nz+1, ny+1]
    nodes = np.empty((nz + 1, ny + 1, nx + 1, 3))
    nodes[..., 0] = X[:, None, :]
    nodes[..., 1] = Y[:, :, None]
    nodes[..., 2] = zs[:, None, None]
    # interval of each cell: the interval of its lower line
    cx, cy = xiv[:-1], yiv[:-1]
    zc = 0.5 * (zs[:-1] + zs[1:])
    tags = np.full((nz, ny, nx), air_tag, dtype=np.int32)
    done = np.zeros((nz, ny, nx), dtype=bool)
    for tag, rects, (z_lo, z_hi) in regions:
        hit = np.zeros((ny, nx), dtype=bool)
        for x_lo, x_hi, y_lo, y_hi in rects:
            ix = (cx >= xs[1][x_lo]) & (cx < xs[1][x_hi])
            iy = (cy >= ys[1][y_lo]) & (cy < ys[1][y_hi])
            hit |= iy[:, None] & ix[None, :]
        inside = (zc > z_lo) & (zc < z_hi)
        sel = inside[:, None, None] & hit[None] & ~done
        tags[sel] = tag
        done |= sel
    tm = mapped_kuhn_mesh(nodes.reshape(-1, 3), (nx, ny, nz), tags.ravel(), SHIP_SIDES)
    return TetModel(tm)


def ruled_quarter_mesh(xbreaks, ybreaks, z_pos, z_len, lim_x, lim_y, delta_z, regions,
                       air_tag, h, hz=None, ratio=1.3):
    """Mesh one ruled template in the quarter box: breaks {name: (value at
    z_pos, value at z_pos + z_len)}, regions [(tag, rects)] between z_pos and
    z_pos + z_len, delta_z of air on both ends (see ruled_knot_mesh)."""
    zr = (z_pos, z_pos + z_len)
    return ruled_knot_mesh(xbreaks, ybreaks, zr, lim_x, lim_y, delta_z, delta_z,
                           [(tag, rects, zr) for tag, rects in regions], air_tag, h, hz, ratio)


# ---------------------------------------------------------------------------
# the templates (tags in the order of the reference driver's domains)
# ---------------------------------------------------------------------------

def _pair(p, key):
    return (float(p[key + '_1']), float(p[key + '_2']))


def _box(p, delta_x, delta_y, x_outer='X_yoke'):
    lim_x = max(p[x_outer + '_1'], p[x_outer + '_2']) + delta_x
    lim_y = max(p['Y_yoke_1'], p['Y_yoke_2']) + delta_y
    return lim_x, lim_y


def mag_1_3_quarter(p, delta_x, delta_y, delta_z, h, hz=None, diluted_wing=False, ratio=1.3):
    """Yoke types 1 and 3 (snoopy/geometry_tools.py:216-304): core
    [X_mgap, X_core] x [0, Y_core], leg [X_void, X_yoke] x [0, Y_core] and
    top plate [X_mgap, X_yoke] x [Y_core, Y_yoke].  With ``diluted_wing``
    the leg is its own volume [X_void, X_yoke] x [0, Y_void] and the main
    volume is core plus top plate (snoopy/geometry_tools.py:601-805).
    Tags: (iron, air) or (iron, wing, air) = (1, 2) or (1, 2, 3)."""
    lim_x, lim_y = _box(p, delta_x, delta_y)
    xb = {n: _pair(p, 'X_' + n) for n in ('mgap', 'core', 'void', 'yoke')}
    yb = {'core': _pair(p, 'Y_core'), 'yoke': _pair(p, 'Y_yoke')}
    if diluted_wing:
        yb['void'] = _pair(p, 'Y_void')
        regions = [(1, [('mgap', 'core', '0', 'core'), ('mgap', 'yoke', 'core', 'yoke')]),
                   (2, [('void', 'yoke', '0', 'void')])]
        air = 3
    else:
        regions = [(1, [('mgap', 'core', '0', 'core'), ('void', 'yoke', '0', 'core'),
                        ('mgap', 'yoke', 'core', 'yoke')])]
        air = 2
    return ruled_quarter_mesh(xb, yb, p['Z_pos'], p['Z_len'], lim_x, lim_y, delta_z,
                              regions, air, h, hz, ratio), (lim_x, lim_y)


def mag_2_quarter(p, delta_x, delta_y, delta_z, h, hz=None, ratio=1.3):
    """The superconducting template: core [0, X_core] x [0, Y_core]
    (snoopy/geometry_tools.py:392-500) and yoke type 2, [0, X_yoke] x
    [0, Y_yoke] without [0, X_void] x [0, Y_void]
    (snoopy/geometry_tools.py:305-389).  Tags (core, yoke, air) = (1, 2, 3)."""
    lim_x, lim_y = _box(p, delta_x, delta_y)
    xb = {n: _pair(p, 'X_' + n) for n in ('core', 'void', 'yoke')}
    yb = {n: _pair(p, 'Y_' + n) for n in ('core', 'void', 'yoke')}
    regions = [(1, [('0', 'core', '0', 'core')]),
               (2, [('void', 'yoke', '0', 'void'), ('0', 'yoke', 'void', 'yoke')])]
    return ruled_quarter_mesh(xb, yb, p['Z_pos'], p['Z_len'], lim_x, lim_y, delta_z,
                              regions, 3, h, hz, ratio), (lim_x, lim_y)


def mag_4_quarter(p, delta_x, delta_y, delta_z, h, hz=None, ratio=1.3):
    """The two-part template (snoopy/geometry_tools.py:983-1119): core
    [X_C, X_D] x [0, Y_core]; yoke [X_A, X_F] x [0, Y_yoke] without
    [X_B, X_E] x [0, Y_core].  Tags (core, yoke, air) = (1, 2, 3)."""
    lim_x = max(p['X_F_1'], p['X_F_2']) + delta_x
    lim_y = max(p['Y_yoke_1'], p['Y_yoke_2']) + delta_y
    xb = {n: _pair(p, 'X_' + n) for n in 'ABCDEF'}
    yb = {'core': _pair(p, 'Y_core'), 'yoke': _pair(p, 'Y_yoke')}
    regions = [(1, [('C', 'D', '0', 'core')]),
               (2, [('A', 'B', '0', 'core'), ('E', 'F', '0', 'core'), ('A', 'F', 'core', 'yoke')])]
    return ruled_quarter_mesh(xb, yb, p['Z_pos'], p['Z_len'], lim_x, lim_y, delta_z,
                              regions, 3, h, hz, ratio), (lim_x, lim_y)


def _stack_two(a, b, lim):
    """Breaks of two solids stacked along z, A first, at the knots (A
    entrance, A exit, B entrance, B exit): {name: 4 values}, with '0' and
    'lim'.  Each solid's breaks follow it; between the solids every break is
    constant.  Inside a solid the other solid's breaks are inactive: each
    keeps the fraction of the way between its nearest active neighbours that
    it has between the solids, so no two grid lines cross or meet where
    they are apart between the solids (lines that meet there stay merged)."""
    gap = {n: v[1] for n, v in a.items()}
    gap.update({n: v[0] for n, v in b.items()})
    gap['0'], gap['lim'] = 0.0, float(lim)
    owner = {n: 'a' for n in a}
    owner.update({n: 'b' for n in b})
    owner['0'] = owner['lim'] = 'box'
    order = sorted(gap, key=lambda n: (gap[n], n))
    out = {n: [0.0, gap[n], gap[n], 0.0] for n in order}

    def place(knot, solid, own_value):
        active = [n for n in order if owner[n] in (solid, 'box')]
        for n in active:
            out[n][knot] = own_value(n)
        for i, n in enumerate(order):
            if owner[n] in (solid, 'box'):
                continue
            lo = next(m for m in reversed(order[:i]) if owner[m] in (solid, 'box'))
            hi = next(m for m in order[i + 1:] if owner[m] in (solid, 'box'))
            f = (gap[n] - gap[lo]) / (gap[hi] - gap[lo]) if gap[hi] > gap[lo] else 0.0
            out[n][knot] = out[lo][knot] + f * (out[hi][knot] - out[lo][knot])

    place(0, 'a', lambda n: a[n][0] if owner[n] == 'a' else gap[n])
    place(3, 'b', lambda n: b[n][1] if owner[n] == 'b' else gap[n])
    return {n: v for n, v in out.items() if n not in ('0', 'lim')}


def ncsc_quarter(p1, p2, lim_x, lim_y, pad_lo, pad_hi, h, hz=None, ratio=1.3):
    """A normal-conducting magnet (yoke type 1, geometry p1) followed along z
    by a superconducting one (core and yoke type 2, geometry p2) in one air
    box (snoopy/SHiP_functions.py:2298-2352).  Every break of one magnet
    keeps its end value along the other magnet, so the breaks of the two
    must not cross.  Tags (normal-conducting yoke, SC core, SC yoke, air) =
    (1, 2, 3, 4)."""
    z1a, z1b = p1['Z_pos'], p1['Z_pos'] + p1['Z_len']
    z2a, z2b = p2['Z_pos'], p2['Z_pos'] + p2['Z_len']
    if z2a < z1b:
        raise ValueError('the superconducting magnet starts (z = %g) inside the normal-'
                         'conducting one (ends at z = %g)' % (z2a, z1b))
    xb = _stack_two({'n_' + n: _pair(p1, 'X_' + n) for n in ('mgap', 'core', 'void', 'yoke')},
                    {'s_' + n: _pair(p2, 'X_' + n) for n in ('core', 'void', 'yoke')}, lim_x)
    yb = _stack_two({'n_' + n: _pair(p1, 'Y_' + n) for n in ('core', 'yoke')},
                    {'s_' + n: _pair(p2, 'Y_' + n) for n in ('core', 'void', 'yoke')}, lim_y)
    regions = [(1, [('n_mgap', 'n_core', '0', 'n_core'), ('n_void', 'n_yoke', '0', 'n_core'),
                    ('n_mgap', 'n_yoke', 'n_core', 'n_yoke')], (z1a, z1b)),
               (2, [('0', 's_core', '0', 's_core')], (z2a, z2b)),
               (3, [('s_void', 's_yoke', '0', 's_void'), ('0', 's_yoke', 's_void', 's_yoke')],
                (z2a, z2b))]
    return ruled_knot_mesh(xb, yb, (z1a, z1b, z2a, z2b), lim_x, lim_y, pad_lo, pad_hi, regions,
                           4, h, hz, ratio)


def ruled_volume(section_area, z_len):
    """Exact volume of a ruled solid whose section area is quadratic in z:
    Simpson's rule on the entrance, middle and exit sections."""
    return z_len / 6.0 * (section_area(0.0) + 4.0 * section_area(0.5) + section_area(1.0))
