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
    """Sort named breaks {name: (entrance, exit)}; merge breaks that coincide
    at both ends; reject breaks that cross or touch at one end only."""
    names = sorted(breaks, key=lambda k: (breaks[k][0], breaks[k][1]))
    vals, index = [], {}
    for nm in names:
        a, b = breaks[nm]
        if vals and abs(a - vals[-1][0]) <= tol and abs(b - vals[-1][1]) <= tol:
            index[nm] = len(vals) - 1
            continue
        if vals and (a < vals[-1][0] - tol or b < vals[-1][1] - tol):
            raise ValueError('break %r crosses its neighbour (entrance %g, exit %g)' % (nm, a, b))
        if vals and (abs(a - vals[-1][0]) <= tol or abs(b - vals[-1][1]) <= tol):
            raise ValueError('break %r meets its neighbour at one end only; the structured '
                             'mesher needs breaks that are distinct at both ends or equal '
                             'at both' % (nm,))
        index[nm] = len(vals)
        vals.append((a, b))
    return np.asarray(vals, dtype=np.float64), index


def _axis(breaks, h, ratio):
    """Cell boundaries of one axis as (interval, fraction) pairs: uniform
    cells of size <= h in every interval but the last, graded cells growing
    from h in the last (the air pad up to the box)."""
    vals, _ = breaks
    n_int = len(vals) - 1
    iv, fr = [], []
    for m in range(n_int):
        w = max(vals[m + 1, 0] - vals[m, 0], vals[m + 1, 1] - vals[m, 1])
        f = _graded_fracs(w, h, ratio) if m == n_int - 1 else _uniform_fracs(w, h)
        iv.append(np.full(len(f), m))
        fr.append(f)
    iv.append([n_int - 1])
    fr.append([1.0])
    return np.concatenate(iv).astype(np.int64), np.concatenate(fr)


def _line_positions(vals, iv, fr, t):
    """Positions of the grid lines at the relative z-position t in [0, 1]."""
    b = (1.0 - t) * vals[:, 0] + t * vals[:, 1]
    return b[iv] + fr * (b[np.minimum(iv + 1, len(b) - 1)] - b[iv])


def ruled_quarter_mesh(xbreaks, ybreaks, z_pos, z_len, lim_x, lim_y, delta_z, regions,
                       air_tag, h, hz=None, ratio=1.3):
    """Mesh a ruled template in the quarter box.

    xbreaks, ybreaks  {name: (value at z_pos, value at z_pos + z_len)}; the
                      names '0' and 'lim' (0 and lim_x / lim_y) are added
    regions           [(tag, [(x_lo, x_hi, y_lo, y_hi), ...]), ...]: a cell
                      between z_pos and z_pos + z_len inside one of the
                      rectangles (named breaks) takes the tag; the first
                      region that holds it wins; every other cell is air_tag
    h, hz             cell size across and along z (hz defaults to h)
    ratio             growth of the graded cells in the air pads
    """
    hz = h if hz is None else hz
    tol = 1e-12 * max(lim_x, lim_y, 1.0)
    xb = dict(xbreaks, **{'0': (0.0, 0.0), 'lim': (lim_x, lim_x)})
    yb = dict(ybreaks, **{'0': (0.0, 0.0), 'lim': (lim_y, lim_y)})
    xs, ys = _sorted_breaks(xb, tol), _sorted_breaks(yb, tol)
    xiv, xfr = _axis(xs, h, ratio)
    yiv, yfr = _axis(ys, h, ratio)
    z1, z2 = z_pos, z_pos + z_len
    zlo = _graded(z1 - delta_z, z1, hz, ratio, small_at_a=False)[:-1]
    zin = np.linspace(z1, z2, max(1, int(math.ceil(z_len / hz - 1e-9))) + 1)[:-1]
    zhi = _graded(z2, z2 + delta_z, hz, ratio, small_at_a=True)
    zs = np.concatenate([zlo, zin, zhi])
    nx, ny, nz = len(xfr) - 1, len(yfr) - 1, len(zs) - 1
    t = np.clip((zs - z1) / z_len, 0.0, 1.0)
    X = np.stack([_line_positions(xs[0], xiv, xfr, tk) for tk in t])     # [nz+1, nx+1]
    Y = np.stack([_line_positions(ys[0], yiv, yfr, tk) for tk in t])     # [nz+1, ny+1]
    nodes = np.empty((nz + 1, ny + 1, nx + 1, 3))
    nodes[..., 0] = X[:, None, :]
    nodes[..., 1] = Y[:, :, None]
    nodes[..., 2] = zs[:, None, None]
    # interval of each cell: the interval of its lower line
    cx, cy = xiv[:-1], yiv[:-1]
    zc = 0.5 * (zs[:-1] + zs[1:])
    inside = (zc > z1) & (zc < z2)
    tags = np.full((nz, ny, nx), air_tag, dtype=np.int32)
    done = np.zeros((nz, ny, nx), dtype=bool)
    for tag, rects in regions:
        hit = np.zeros((ny, nx), dtype=bool)
        for x_lo, x_hi, y_lo, y_hi in rects:
            ix = (cx >= xs[1][x_lo]) & (cx < xs[1][x_hi])
            iy = (cy >= ys[1][y_lo]) & (cy < ys[1][y_hi])
            hit |= iy[:, None] & ix[None, :]
        sel = inside[:, None, None] & hit[None] & ~done
        tags[sel] = tag
        done |= sel
    tm = mapped_kuhn_mesh(nodes.reshape(-1, 3), (nx, ny, nz), tags.ravel(), SHIP_SIDES)
    return TetModel(tm)


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


def ruled_volume(section_area, z_len):
    """Exact volume of a ruled solid whose section area is quadratic in z:
    Simpson's rule on the entrance, middle and exit sections."""
    return z_len / 6.0 * (section_area(0.0) + 4.0 * section_area(0.5) + section_area(1.0))
