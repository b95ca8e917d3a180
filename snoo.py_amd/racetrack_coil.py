"""Racetrack coil source plugin (snoopy/racetrack_coil.py:14-270).

``compute_B`` evaluates Biot-Savart on the GPU (HIP kernel behind the C ABI);
``get_segments`` and ``get_length`` are host-side geometry, restated from the
reference: four straight runs between the 8 key points in the x-z plane and
four corners, each a quadratic B-spline with knots [0,0,0,1,1,1] (a quadratic
Bezier curve) whose middle control point is the intersection of the two
adjacent straight lines, cut into ``disc_arc`` chords; the level is repeated
at every y of ``self.y``.
"""
import numpy as np


def compute_2D_line_intersection(intercept_1, slope_1, intercept_2, slope_2):
    """Intersection of two lines in 2D (snoopy/racetrack_coil.py:14-40)."""
    M = np.array([[slope_1[0], -slope_2[0]],
                  [slope_1[1], -slope_2[1]]])
    b = np.array([intercept_2[0] - intercept_1[0],
                  intercept_2[1] - intercept_1[1]])
    s = np.linalg.solve(M, b)
    return slope_1 * s[0] + intercept_1


def _bezier2(p0, p1, p2, t):
    """Quadratic B-spline on knots [0,0,0,1,1,1] = Bernstein form."""
    t = t[:, None]
    return (1.0 - t) ** 2 * p0 + 2.0 * t * (1.0 - t) * p1 + t ** 2 * p2


class RacetrackCoil:
    """Racetrack coil made of straight filaments (snoopy/racetrack_coil.py:43-70).

    :param kp: the 8 key points in the x,z plane, shape (8, 2).
    :param y: vertical coordinates; one turn per entry.
    :param rad: filament core radius of the Biot-Savart kernel.
    :param current: current of each filament (A).
    """

    def __init__(self, kp, y, rad, current=1.0):
        self.kp = np.asarray(kp, dtype=np.float64)
        self.y = np.asarray(y, dtype=np.float64)
        self.rad = rad
        self.current = current

    def compute_B(self, positions, disc_arc=5):
        """B at ``positions`` [M,3] (snoopy/racetrack_coil.py:72-88).

        Host arrays in and out; the evaluation runs on the GPU.  For device
        tensors use :func:`snoopy_amd.sources.biot_savart_device`.
        """
        from . import sources
        return sources.biot_savart_host(self.get_segments(disc_arc=disc_arc),
                                        positions, self.current, self.rad)

    def get_segments(self, disc_arc=5, enable_plot=False):
        """(M x 6) segment array [x1,y1,z1,x2,y2,z2] (snoopy/racetrack_coil.py:90-242)."""
        kp = self.kp
        t_vec = np.linspace(0.0, 1.0, disc_arc + 1)
        per_lvl = 4 * disc_arc + 4
        base = np.zeros((per_lvl, 6))

        def straight(row, a, b):
            base[row, 0], base[row, 2] = kp[a, 0], kp[a, 1]
            base[row, 3], base[row, 5] = kp[b, 0], kp[b, 1]

        def corner(row0, a, b, la, lb):
            # control point: intersection of the straight lines through a and b
            ctrl = compute_2D_line_intersection(kp[la[0], :], kp[la[1], :] - kp[la[0], :],
                                                kp[lb[0], :], kp[lb[1], :] - kp[lb[0], :])
            pts = _bezier2(kp[a], ctrl, kp[b], t_vec)
            for i in range(disc_arc):
                base[row0 + i, 0], base[row0 + i, 2] = pts[i, 0], pts[i, 1]
                base[row0 + i, 3], base[row0 + i, 5] = pts[i + 1, 0], pts[i + 1, 1]

        straight(0, 7, 0)
        corner(1, 0, 1, (7, 0), (1, 2))
        straight(disc_arc + 1, 1, 2)
        corner(2 + disc_arc, 2, 3, (1, 2), (3, 4))
        straight(2 + 2 * disc_arc, 3, 4)
        corner(3 + 2 * disc_arc, 4, 5, (3, 4), (5, 6))
        straight(3 + 3 * disc_arc, 5, 6)
        corner(4 + 3 * disc_arc, 6, 7, (5, 6), (7, 0))

        out = np.empty((len(self.y) * per_lvl, 6))
        for i, yy in enumerate(self.y):
            base[:, 1] = yy
            base[:, 4] = yy
            out[i * per_lvl:(i + 1) * per_lvl] = base
        return out

    def get_length(self, one_layer=True):
        """Total filament length, per layer by default (snoopy/racetrack_coil.py:244-270)."""
        segs = self.get_segments()
        lens = np.sqrt((segs[:, 3] - segs[:, 0]) ** 2 + (segs[:, 4] - segs[:, 1]) ** 2
                       + (segs[:, 5] - segs[:, 2]) ** 2)
        length = np.sum(lens)
        if one_layer:
            length /= len(self.y)
        return length
