"""Brick-field source plugin (snoopy/brick_field.py:6-433).

A piecewise-constant field in three hexahedral iron paths of a SHiP-style
magnet: B = Phi/area_1 along +y in brick 1, Phi/area_2 along +x in brick 2,
-Phi/area_3 along y in brick 3, zero elsewhere (snoopy/brick_field.py:400-433).
A point is in a brick when it lies on the inner side (or on) each of its six
faces: (p - r_f) . n_f <= 0 with r_f, n_f the centre and unit normal of the
bilinear face at (u,v) = (0,0) (snoopy/brick_field.py:42-70,165-194).

The face planes are computed here on the host; ``compute_B`` evaluates the
point tests on the GPU.
"""
import numpy as np


def _quad_centre_normal(p1, p2, p3, p4):
    """Centre and unit normal of QuadGeometry at (0,0) (brick_field.py:21-70)."""
    r = 0.25 * (p1 + p2 + p3 + p4)
    du = 0.25 * p1 - 0.25 * p2 - 0.25 * p3 + 0.25 * p4
    dv = 0.25 * p1 + 0.25 * p2 - 0.25 * p3 - 0.25 * p4
    n = np.cross(du, dv)
    return r, n / np.linalg.norm(n)


def _brick_faces(points):
    """The six faces of BrickGeometry in the reference's order (brick_field.py:137-158)."""
    P = points
    quads = [(P[3], P[2], P[1], P[0]), (P[4], P[5], P[6], P[7]),
             (P[5], P[4], P[0], P[1]), (P[6], P[5], P[1], P[2]),
             (P[2], P[3], P[7], P[6]), (P[3], P[0], P[4], P[7])]
    return [_quad_centre_normal(*q) for q in quads]


class BrickField:
    """Brick-field magnet (snoopy/brick_field.py:277-398)."""

    def __init__(self, X_mgap_1, X_core_1, X_void_1, X_yoke_1,
                 X_mgap_2, X_core_2, X_void_2, X_yoke_2,
                 Y_void_1, Y_yoke_1, Y_void_2, Y_yoke_2,
                 Z_len, Z_pos=0.0, B_goal=1.0, type=1):
        p1 = np.array([[X_mgap_1, 0., Z_pos], [X_core_1, 0., Z_pos],
                       [X_core_1, Y_void_1, Z_pos], [X_mgap_1, Y_yoke_1, Z_pos],
                       [X_mgap_2, 0., Z_pos + Z_len], [X_core_2, 0., Z_pos + Z_len],
                       [X_core_2, Y_void_2, Z_pos + Z_len], [X_mgap_2, Y_yoke_2, Z_pos + Z_len]])
        p2 = np.array([[X_mgap_1, Y_yoke_1, Z_pos], [X_core_1, Y_void_1, Z_pos],
                       [X_void_1, Y_void_1, Z_pos], [X_yoke_1, Y_yoke_1, Z_pos],
                       [X_mgap_2, Y_yoke_2, Z_pos + Z_len], [X_core_2, Y_void_2, Z_pos + Z_len],
                       [X_void_2, Y_void_2, Z_pos + Z_len], [X_yoke_2, Y_yoke_2, Z_pos + Z_len]])
        p3 = np.array([[X_void_1, 0.0, Z_pos], [X_yoke_1, 0.0, Z_pos],
                       [X_yoke_1, Y_yoke_1, Z_pos], [X_void_1, Y_void_1, Z_pos],
                       [X_void_2, 0.0, Z_pos + Z_len], [X_yoke_2, 0.0, Z_pos + Z_len],
                       [X_yoke_2, Y_yoke_2, Z_pos + Z_len], [X_void_2, Y_void_2, Z_pos + Z_len]])
        self.bricks = [p1, p2, p3]
        self.B_goal = B_goal
        self.area_1 = 0.5 * ((X_core_1 - X_mgap_1) * Z_len + (X_core_2 - X_mgap_2) * Z_len)
        self.area_2 = 0.5 * ((Y_yoke_1 - Y_void_1) * Z_len + (Y_yoke_2 - Y_void_2) * Z_len)
        self.area_3 = 0.5 * ((X_yoke_1 - X_void_1) * Z_len + (X_yoke_2 - X_void_2) * Z_len)
        if type == 1:
            self.Phi = self.B_goal * self.area_1
        elif type == 2:
            self.Phi = self.B_goal * self.area_3
        else:
            raise ValueError('BrickField type {} unknown'.format(type))

    def face_table(self):
        """float64 [3 bricks * 6 faces * 6] = (r, n) per face."""
        out = []
        for P in self.bricks:
            for r, n in _brick_faces(P):
                out.extend(list(r) + list(n))
        return np.array(out, dtype=np.float64)

    def field_values(self):
        """Field vector of each brick, float64 [3*3]."""
        return np.array([0.0, self.Phi / self.area_1, 0.0,
                         self.Phi / self.area_2, 0.0, 0.0,
                         0.0, -1.0 * self.Phi / self.area_3, 0.0])

    def compute_B(self, points):
        """B at ``points`` [M,3] (snoopy/brick_field.py:400-433); runs on the GPU."""
        from . import sources
        return sources.brick_field_host(self, points)
