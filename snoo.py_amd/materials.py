"""Reluctivity models (snoopy/materials.py:271-536) plus a B-spline-in-B^2 model.

Every model hands the device kernels one table format: a piecewise cubic in
power basis, coefficients ``c`` of shape (4, n) and knots ``x`` of length n+1,
evaluated as sum_j c[j,i] (t - x_i)^(3-j) on x_i <= t < x_{i+1} and with the
last cubic extrapolated at and beyond x_n (fem_c/ctools.c:130-204).

* ``ConstantReluctance``: type_spec 0, the value in c[0] (n = -1).
* ``Reluctance``: type_spec 1, scipy PCHIP of nu = H/B over t = |B|, with
  nu[0] = nu[1] (snoopy/materials.py:298-311).
* ``BSplineReluctance``: type_spec 2, a cubic B-spline of nu over t = |B|^2,
  converted to its piecewise-polynomial form; nu(B) = P(B^2) and
  dnu/dB = 2 B P'(B^2).
"""
import json

import numpy as np
from scipy.interpolate import PchipInterpolator, PPoly, make_interp_spline

MU0 = 4e-7 * np.pi
NU0 = 0.25 / np.pi * 1e7          # vacuum reluctivity used by the reference (red_mvp_solver.py:67)

TYPE_CONST = 0
TYPE_PCHIP = 1
TYPE_BSQ = 2


def _eval_ppoly_ref(c, x, t, deriv=False):
    """Host evaluation with the kernel's interval rule (ctools.c:130-204)."""
    t = np.asarray(t, dtype=np.float64)
    n = c.shape[1]
    i = np.clip(np.searchsorted(x, t, side='right') - 1, 0, n - 1)
    below = t < x[0]
    dt = t - x[i]
    if deriv:
        f = (3 * c[0, i] * dt + 2 * c[1, i]) * dt + c[2, i]
    else:
        f = ((c[0, i] * dt + c[1, i]) * dt + c[2, i]) * dt + c[3, i]
    return np.where(below, 0.0, f)


class Reluctance:
    """PCHIP reluctivity from a BH json (snoopy/materials.py:271-437)."""

    def __init__(self, input_filename=None, B_data=None, H_data=None):
        if input_filename is not None:
            with open(input_filename) as f:
                mat_data = json.load(f)
            B_data = mat_data['BH_data']['B(T)']
            H_data = mat_data['BH_data']['H(A/m)']
        self.B_data = np.array(B_data, dtype=np.float64)
        self.H_data = np.array(H_data, dtype=np.float64)
        self.initialize_interpolation()

    def initialize_interpolation(self):
        # nu = H/B, nu[0] = nu[1] (snoopy/materials.py:305-308)
        self.nu = 0.0 * self.H_data
        self.nu[1:] = self.H_data[1:] / self.B_data[1:]
        self.nu[0] = self.nu[1]
        self.spl = PchipInterpolator(self.B_data, self.nu, extrapolate=True)

    def evaluate_nu(self, B_mag):
        return self.spl(B_mag)

    def evaluate_nu_derivative(self, B_mag):
        return self.spl.derivative()(B_mag)

    def get_type_spec(self):
        return TYPE_PCHIP

    def get_spline_information(self):
        """(c[4,n], x[n+1]) of the PCHIP (snoopy/materials.py:416-424)."""
        return self.spl.c, self.spl.x


class ConstantReluctance:
    """Constant reluctivity (snoopy/materials.py:486-536)."""

    def __init__(self, value):
        self.value = value

    def evaluate_nu(self, B_mag):
        return self.value + 0.0 * np.asarray(B_mag, dtype=np.float64)

    def evaluate_nu_derivative(self, B_mag):
        return 0.0 * np.asarray(B_mag, dtype=np.float64)

    def get_type_spec(self):
        return TYPE_CONST


class BSplineReluctance:
    """Cubic B-spline of nu over s = |B|^2, interpolating (B_i^2, H_i/B_i).

    The spline is held as scipy ``BSpline`` (``self.spl``) and handed to the
    kernels as its piecewise-polynomial form on the distinct knots.
    """

    def __init__(self, B_data, H_data, nu_at_zero=None):
        B = np.asarray(B_data, dtype=np.float64)
        H = np.asarray(H_data, dtype=np.float64)
        nu = np.empty_like(B)
        nz = B > 0
        nu[nz] = H[nz] / B[nz]
        if (~nz).any():
            nu[~nz] = nu_at_zero if nu_at_zero is not None else nu[nz][0]
        order = np.argsort(B)
        self.B_data, self.H_data, self.nu = B[order], H[order], nu[order]
        s = self.B_data ** 2
        self.spl = make_interp_spline(s, self.nu, k=3)
        pp = PPoly.from_spline(self.spl)
        # keep only the intervals between distinct breakpoints inside the data
        x = pp.x
        keep = np.nonzero(np.diff(x) > 0)[0]
        self._c = np.ascontiguousarray(pp.c[:, keep])
        self._x = np.ascontiguousarray(np.concatenate([x[keep], [x[keep[-1] + 1]]]))

    def evaluate_nu(self, B_mag):
        return _eval_ppoly_ref(self._c, self._x, np.asarray(B_mag) ** 2)

    def evaluate_nu_derivative(self, B_mag):
        B_mag = np.asarray(B_mag, dtype=np.float64)
        return 2.0 * B_mag * _eval_ppoly_ref(self._c, self._x, B_mag ** 2, deriv=True)

    def get_type_spec(self):
        return TYPE_BSQ

    def get_spline_information(self):
        return self._c, self._x


def tanh_bh_samples(B_sat=1.9, mu_r=4000.0, n=60, H_max=1e6):
    """B(H) = mu0 H + B_sat tanh(mu0 mu_r H / B_sat) at H=0 and n-1
    log-spaced H in [1, H_max]."""
    H = np.concatenate([[0.0], np.logspace(0.0, np.log10(H_max), n - 1)])
    B = MU0 * H + B_sat * np.tanh(MU0 * mu_r * H / B_sat)
    return B, H


def tanh_reluctance(B_sat=1.9, mu_r=4000.0, n=60):
    """The configuration's iron: cubic B-spline in B^2 fitted to ``n`` samples
    of B(H) = mu0 H + B_sat tanh(mu0 mu_r H / B_sat); nu(0) is the limit
    1/(mu0 (1 + mu_r))."""
    B, H = tanh_bh_samples(B_sat, mu_r, n)
    return BSplineReluctance(B, H, nu_at_zero=1.0 / (MU0 * (1.0 + mu_r)))


def model_tables(rel):
    """(type_spec, n_int, c_flat[4n] or [1], x[n+1] or [0]) as the kernels take it,
    following fem_c/interface_c.pyx:224-234,304-307."""
    ts = rel.get_type_spec()
    if ts == TYPE_CONST:
        return ts, -1, np.array([float(rel.value)]), np.zeros(1)
    c, x = rel.get_spline_information()
    c = np.ascontiguousarray(np.asarray(c, dtype=np.float64))
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    return ts, len(x) - 1, c.reshape(-1), x


class Dilution:
    """Dilution of the iron along z (snoopy/materials.py:538-599): the ratio of
    two linear functions of z built from the entrance (1) and exit (2) x
    coordinates of middle gap, core, void and yoke; magnet type 1 uses
    (yoke - void) / (core - middle gap), type 3 the inverse."""

    def __init__(self, x_mgap_1, x_core_1, x_void_1, x_yoke_1,
                 x_mgap_2, x_core_2, x_void_2, x_yoke_2, z_pos, z_len, mag_type=1):
        slope = lambda a, b: (b - a) / z_len
        icpt = lambda a, b: (a * (z_len + z_pos) - b * z_pos) / z_len
        m_yoke, c_yoke = slope(x_yoke_1, x_yoke_2), icpt(x_yoke_1, x_yoke_2)
        m_core, c_core = slope(x_core_1, x_core_2), icpt(x_core_1, x_core_2)
        m_void, c_void = slope(x_void_1, x_void_2), icpt(x_void_1, x_void_2)
        m_mgap, c_mgap = slope(x_mgap_1, x_mgap_2), icpt(x_mgap_1, x_mgap_2)
        if mag_type == 1:
            self.m_num, self.m_den = m_yoke - m_void, m_core - m_mgap
            self.c_num, self.c_den = c_yoke - c_void, c_core - c_mgap
        elif mag_type == 3:
            self.m_num, self.m_den = m_core - m_mgap, m_yoke - m_void
            self.c_num, self.c_den = c_core - c_mgap, c_yoke - c_void
        else:
            raise ValueError('magnet type %r unknown (1 or 3)' % (mag_type,))

    @classmethod
    def from_coefficients(cls, c_num, m_num, c_den, m_den):
        d = cls.__new__(cls)
        d.c_num, d.m_num, d.c_den, d.m_den = float(c_num), float(m_num), float(c_den), float(m_den)
        return d

    def evaluate(self, z):
        """s(z) = (c_num + z m_num) / (c_den + z m_den) (fem_c/ctools.c:124-128)."""
        z = np.asarray(z, dtype=np.float64)
        return (self.c_num + z * self.m_num) / (self.c_den + z * self.m_den)
