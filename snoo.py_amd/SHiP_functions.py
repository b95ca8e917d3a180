"""SHiP magnet templates and their cost model (snoopy/SHiP_functions.py:40-2620).

Same names, arguments and return values as the reference:

* ``get_NI``, ``compute_prices``, ``compute_J_opt``: the cost model
  (snoopy/SHiP_functions.py:40-279), host arithmetic in the reference's
  order of operations so the figures agree to the last bit;
* ``get_vector_field_mag_1`` / ``_mag_2`` / ``_mag_3`` / ``_mag_4``: read a
  template row of the parameter table, build the quarter-magnet mesh, the
  racetrack coils and the solver, solve, and return
  ``[points, B, M_iron, M_coil, Q, current_density]`` with B at the
  quadrature points of every element (iron + coil field), optionally
  writing x, the parameter row and B at ``eval_pos`` to ``result_directory``.

The geometry comes from this build's structured mesher
(``geometry_tools``) instead of gmsh OCC: ``lc`` sets the cell size
(``h = lc / 4`` across, ``hz = lc / 2`` along z unless given).  The iron mass
is the exact volume of the ruled solids, as gmsh's getMass gives it.  The
solve is the reference's Newton iteration (``RedMVPSolver.solve``) unless
``method='newton'`` asks for the inexact-Newton solver with the AMS
preconditioner.  Plotting (pyvista) and the gmsh GUI are not provided.
"""
import json
import os
import time

import numpy as np

from .evaluators import evaluate_fem_solution_node_interpolation
from .geometry_tools import (compute_area_polygon, mag_1_3_quarter, mag_2_quarter,
                             mag_4_quarter, ncsc_quarter, ruled_volume)
from .materials import Dilution, Reluctance
from .racetrack_coil import RacetrackCoil
from .red_mvp_solver import RedMVPSolver, RedMVPSolverDiluted

_COLS = {'mgap': 'Xmgap', 'core': 'Xcore', 'void': 'Xvoid', 'yoke': 'Xyoke'}

# solver statistics, mesh size and wall time of the last driver call
last_run = {}


def _load(materials_directory, name):
    with open(os.path.join(materials_directory, name)) as f:
        return json.load(f)


def _row(parameters, i):
    """The geometry of template row i as {'X_core_1': ..., 'Z_pos': ...}."""
    p = {}
    for n, col in _COLS.items():
        for e in (1, 2):
            p['X_%s_%d' % (n, e)] = float(parameters['%s%d(m)' % (col, e)][i])
    for n in ('core', 'void', 'yoke'):
        for e in (1, 2):
            p['Y_%s_%d' % (n, e)] = float(parameters['Y%s%d(m)' % (n, e)][i])
    p['Z_len'] = float(parameters['Z_len(m)'][i])
    p['Z_pos'] = float(parameters['Z_pos(m)'][i])
    return p


def _drive(parameters, i):
    """Coil and box settings of template row i (snoopy/SHiP_functions.py:402-411)."""
    return dict(delta_x=float(parameters['delta_x(m)'][i]),
                delta_y=float(parameters['delta_y(m)'][i]),
                delta_z=float(parameters['delta_z(m)'][i]),
                yoke_spacer=parameters['yoke_spacer(mm)'][i] * 1e-3,
                ins=parameters['insulation(mm)'][i] * 1e-3,
                current=parameters['NI(A)'][i],
                coil_radius=0.5 * parameters['coil_diam(mm)'][i] * 1e-3,
                field_density=0.5 * parameters['field_density'][i],
                max_turns=np.int64(parameters['max_turns'][i]))


# ---------------------------------------------------------------------------
# cost model
# ---------------------------------------------------------------------------

def compute_J_opt(conductor_material_data, kappa_elec):
    """Current density in A/mm^2 that balances conductor and electricity cost
    (snoopy/SHiP_functions.py:260-279): sqrt(kappa_cu dens 1e-3 / kappa_elec /
    rho) 1e-6 with kappa_cu = material + manufacturing cost per kg."""
    kappa_cu = conductor_material_data['material_cost(CHF/kg)']
    kappa_cu += conductor_material_data['manufacturing_cost(CHF/kg)']
    dens = conductor_material_data['density(g/m3)']
    rho = conductor_material_data['resistivity(Ohm.m)']
    return np.sqrt(kappa_cu * dens * 1e-3 / kappa_elec / rho) * 1e-6


def compute_prices(magnet_parameters, df_index, M_iron, M_coil, Q,
                   electricity_costs=5.0 / 72000.0, runtime=72000,
                   materials_directory='files/materials', price_peanuts=0.5):
    """Iron, coil and electricity cost (snoopy/SHiP_functions.py:204-258);
    masses in g, Q in W."""
    iron = _load(materials_directory, magnet_parameters['material'][df_index])
    coil = _load(materials_directory, magnet_parameters['coil_material'][df_index])
    C_iron = 1e-3 * M_iron * (iron['material_cost(CHF/kg)']
                              + iron['manufacturing_cost(CHF/kg)']) * price_peanuts
    C_coil = 1e-3 * M_coil * (coil['material_cost(CHF/kg)']
                              + coil['manufacturing_cost(CHF/kg)']) * price_peanuts
    C_edf = Q * electricity_costs * runtime * price_peanuts
    return C_iron, C_coil, C_edf


def _cut_area(a1, b1, a2, b2, z_len):
    """Area of the cut between the lines a and b (entrance 1, exit 2) over
    the magnet length, through the reference's key-point polygon (its first
    point is the origin, snoopy/SHiP_functions.py:93-144)."""
    kp = np.zeros((5, 2))
    kp[1] = (b1, 0.0)
    kp[2] = (b2, z_len)
    kp[3] = (a2, z_len)
    kp[4] = (a1, 0.0)
    return compute_area_polygon(kp)


def get_NI(B_goal, parameters_df, df_index, materials_directory='files/materials'):
    """Ampere-turns that drive B_goal through the template by the magnetic
    circuit estimate, and whether the geometry suits an air-cooled magnet
    (snoopy/SHiP_functions.py:40-202)."""
    rel = Reluctance(os.path.join(materials_directory, parameters_df['material'][df_index]))
    p = _row(parameters_df, df_index)
    L = p['Z_len']
    A1 = _cut_area(p['X_mgap_1'], p['X_core_1'], p['X_mgap_2'], p['X_core_2'], L)
    A2 = _cut_area(p['Y_void_1'], p['Y_yoke_1'], p['Y_void_2'], p['Y_yoke_2'], L)
    A3 = _cut_area(p['X_void_1'], p['X_yoke_1'], p['X_void_2'], p['X_yoke_2'], L)
    yoke_type = parameters_df['yoke_type'][df_index]
    if yoke_type not in ('Mag1', 'Mag2', 'Mag3'):
        raise ValueError('get_NI: yoke type %r has no circuit estimate (Mag1, Mag2, Mag3)'
                         % (yoke_type,))
    # path lengths through core (1), top plate (2) and leg (3)
    l1 = 0.25 * (p['Y_core_1'] + p['Y_core_2'] + p['Y_yoke_1'] + p['Y_yoke_2'])
    l2 = (0.5 * (p['X_yoke_1'] + p['X_yoke_2']) + 0.5 * (p['X_void_1'] + p['X_void_2'])
          - 0.5 * (p['X_core_1'] + p['X_core_2']) - 0.5 * (p['X_mgap_1'] + p['X_mgap_2']))
    l3 = l1
    if yoke_type == 'Mag3':
        # the coil sits on the leg: the flux of the leg at B_goal
        Phi = abs(B_goal) * A3
        H1 = rel.evaluate_nu(Phi / A1) * Phi / A1
    else:
        Phi = abs(B_goal) * A1
        H1 = rel.evaluate_nu(B_goal) * B_goal
    H2 = rel.evaluate_nu(Phi / A2) * Phi / A2
    H3 = rel.evaluate_nu(Phi / A3) * Phi / A3
    NI = H1 * l1 + H2 * l2 + H3 * l3
    if B_goal < 0.0:
        NI *= -1.0
    valid_geometry = (A1 > A3) if yoke_type == 'Mag3' else (A1 < A3)
    return NI, valid_geometry


# ---------------------------------------------------------------------------
# coils and their figures
# ---------------------------------------------------------------------------

def _levels(Y_core_1, Y_core_2, coil_radius, ins, max_turns):
    """Conductor stack of a coil in y (snoopy/SHiP_functions.py:579-591)."""
    slot_size = 2 * min(Y_core_1, Y_core_2)
    num_cond = np.int32(slot_size / 2 / (coil_radius + ins))
    if num_cond > max_turns:
        num_cond = max_turns
    y = np.linspace(-0.5 * slot_size + coil_radius + ins,
                    0.5 * slot_size - coil_radius - ins, num_cond)
    return slot_size, num_cond, y


def _around(x_in_1, x_out_1, x_in_2, x_out_2, dv, z_pos, z_len, lift=0.0):
    """Key points (x, z) of a racetrack around [x_in, x_out] at the clearance
    yoke spacer + insulation + conductor radius, as the templates lay them
    out (snoopy/SHiP_functions.py:617-645,1523-1531).  The clearance is
    added term by term, left to right, so the key points (and the coil
    lengths and masses after them) agree with the reference's to the bit."""
    s, ins, r = dv['yoke_spacer'], dv['ins'], dv['coil_radius']
    out = lambda v: v + s + ins + r  # noqa: E731
    inn = lambda v: v - s - ins - r  # noqa: E731
    top = z_pos + z_len
    return np.array([[inn(x_in_2), top + lift],
                     [x_in_2, out(top)],
                     [x_out_2, out(top)],
                     [out(x_out_2), top],
                     [out(x_out_1), z_pos],
                     [x_out_1, inn(z_pos)],
                     [x_in_1, inn(z_pos)],
                     [inn(x_in_1), z_pos]])


def _figures(current, J_tar, conductor, turn_perimeter, A_geo):
    """Copper and coil cross-sections, current density, coil mass and power
    (snoopy/SHiP_functions.py:596-660)."""
    A_cu = abs(current) / J_tar
    A_coil = abs(current) / J_tar / conductor['filling_factor']
    if current == 0.0:
        current_density = 0.0
    else:
        current_density = abs(current) / min([A_geo, A_coil]) / conductor['filling_factor']
    M_coil = A_cu * turn_perimeter * conductor['density(g/m3)']
    if current == 0:
        Q = 0.0
    else:
        Q = abs(current * current) * turn_perimeter * conductor['resistivity(Ohm.m)'] / A_cu
    return A_coil, current_density, M_coil, Q


def _target_density(parameters, i, conductor, kappa_elec):
    if parameters['J_tar(A/mm2)'][i] < 0:
        return compute_J_opt(conductor, kappa_elec) * 1e6
    return parameters['J_tar(A/mm2)'][i] * 1e6


class Template:
    """Everything the drivers build before the solve: mesh model, domain
    tags, reluctances, dilutions, coils and the cost figures."""

    def __init__(self, model, iron_tags, air_tags, reluctances, dilutions, coils, M_iron,
                 M_coil, Q, current_density, A_coil, turn_perimeter, field_density, current):
        self.model = model
        self.iron_tags = iron_tags
        self.air_tags = air_tags
        self.reluctances = reluctances
        self.dilutions = dilutions
        self.coils = coils
        self.M_iron = M_iron
        self.M_coil = M_coil
        self.Q = Q
        self.current_density = current_density
        self.A_coil = A_coil
        self.turn_perimeter = turn_perimeter
        self.field_density = field_density
        self.current = current


def _sizes(lc, h, hz):
    return (0.25 * lc if h is None else h), (0.5 * lc if hz is None else hz)


def build_template(kind, parameters, df_index=0, lc=0.2, materials_directory='files/materials',
                   use_diluted_steel=False, kappa_elec=5.0, h=None, hz=None):
    """Mesh, coils and cost figures of template ``kind`` (1, 2, 3 or 4) for
    row ``df_index`` (template 4 also reads row df_index + 1)."""
    i = df_index
    rel = Reluctance(os.path.join(materials_directory, parameters['material'][i]))
    iron = _load(materials_directory, parameters['material'][i])
    conductor = _load(materials_directory, parameters['coil_material'][i])
    J_tar = _target_density(parameters, i, conductor, kappa_elec)
    dv = _drive(parameters, i)
    h, hz = _sizes(lc, h, hz)
    dens = iron['density(g/m3)']
    if kind == 4:
        p = _mag4_row(parameters, i)
    else:
        p = _row(parameters, i)
    L, Zp = p['Z_len'], p['Z_pos']
    lin = lambda k, s: (1.0 - s) * p[k + '_1'] + s * p[k + '_2']  # noqa: E731
    slot_size, num_cond, y = _levels(p['Y_core_1'], p['Y_core_2'], dv['coil_radius'],
                                     dv['ins'], dv['max_turns'])
    I_lvl = dv['current'] / num_cond
    dilutions = None
    if kind in (1, 3):
        model, _ = mag_1_3_quarter(p, dv['delta_x'], dv['delta_y'], dv['delta_z'], h, hz,
                                   diluted_wing=use_diluted_steel)
        if use_diluted_steel:
            main = lambda s: ((lin('X_core', s) - lin('X_mgap', s)) * lin('Y_core', s)  # noqa: E731
                              + (lin('X_yoke', s) - lin('X_mgap', s))
                              * (lin('Y_yoke', s) - lin('Y_core', s)))
            wing = lambda s: (lin('X_yoke', s) - lin('X_void', s)) * lin('Y_void', s)  # noqa: E731
            M_iron = 4 * ruled_volume(main, L) * dens
            M_iron += 4 * ruled_volume(wing, L) * dens
            iron_tags, air_tags = [1, 2], [3]
            dil = Dilution(p['X_mgap_1'], p['X_core_1'], p['X_void_1'], p['X_yoke_1'],
                           p['X_mgap_2'], p['X_core_2'], p['X_void_2'], p['X_yoke_2'],
                           Zp, L, mag_type=1)
            dilutions = [None, dil]
            reluctances = [rel, rel]
        else:
            sec = lambda s: ((lin('X_yoke', s) - lin('X_mgap', s)) * lin('Y_yoke', s)  # noqa: E731
                             - (lin('X_void', s) - lin('X_core', s)) * lin('Y_core', s))
            M_iron = 4 * ruled_volume(sec, L) * dens
            iron_tags, air_tags, reluctances = [1], [2], [rel]
        if kind == 1 and (p['X_mgap_1'] == 0.0 or p['X_mgap_2'] == 0.0):
            # one coil around the whole core (snoopy/SHiP_functions.py:614-628)
            kp = _around(-p['X_core_1'], p['X_core_1'], -p['X_core_2'], p['X_core_2'],
                         dv, Zp, L)
            coils = [RacetrackCoil(kp, y, dv['coil_radius'], I_lvl)]
            turn_perimeter = coils[-1].get_length()
        else:
            if kind == 1:
                kp_1 = _around(p['X_mgap_1'], p['X_core_1'], p['X_mgap_2'], p['X_core_2'],
                               dv, Zp, L, lift=dv['coil_radius'])
            else:
                kp_1 = _around(p['X_void_1'], p['X_yoke_1'], p['X_void_2'], p['X_yoke_2'],
                               dv, Zp, L)
            kp_2 = kp_1.copy()
            kp_2[:, 0] *= -1.0
            coils = [RacetrackCoil(kp_1, y, dv['coil_radius'], I_lvl),
                     RacetrackCoil(kp_2, y, dv['coil_radius'], I_lvl)]
            turn_perimeter = coils[0].get_length() + coils[1].get_length()
        slot_size_horz = min(p['X_void_1'] - p['X_core_1'], p['X_void_2'] - p['X_core_2'])
    elif kind == 2:
        model, _ = mag_2_quarter(p, dv['delta_x'], dv['delta_y'], dv['delta_z'], h, hz)
        core = lambda s: lin('X_core', s) * lin('Y_core', s)  # noqa: E731
        yoke = lambda s: (lin('X_yoke', s) * lin('Y_yoke', s)  # noqa: E731
                          - lin('X_void', s) * lin('Y_void', s))
        M_iron = 4 * (ruled_volume(core, L) + ruled_volume(yoke, L)) * dens
        iron_tags, air_tags, reluctances = [1, 2], [3], [rel, rel]
        kp = _around(-p['X_core_1'], p['X_core_1'], -p['X_core_2'], p['X_core_2'], dv, Zp, L)
        coils = [RacetrackCoil(kp, y, dv['coil_radius'], I_lvl)]
        turn_perimeter = coils[0].get_length()
        slot_size_horz = min(p['X_void_1'] - p['X_core_1'], p['X_void_2'] - p['X_core_2'])
    elif kind == 4:
        model, _ = mag_4_quarter(p, dv['delta_x'], dv['delta_y'], dv['delta_z'], h, hz)
        core = lambda s: (lin('X_D', s) - lin('X_C', s)) * lin('Y_core', s)  # noqa: E731
        yoke = lambda s: ((lin('X_F', s) - lin('X_A', s)) * lin('Y_yoke', s)  # noqa: E731
                          - (lin('X_E', s) - lin('X_B', s)) * lin('Y_core', s))
        M_iron = 4 * ruled_volume(core, L) * dens
        M_iron += 4 * ruled_volume(yoke, L) * dens
        iron_tags, air_tags, reluctances = [1, 2], [3], [rel, rel]
        if use_diluted_steel:
            dil = Dilution(p['X_A_1'] + p['X_E_1'], p['X_B_1'] + p['X_F_1'], p['X_C_1'],
                           p['X_D_1'], p['X_A_2'] + p['X_E_2'], p['X_B_2'] + p['X_F_2'],
                           p['X_C_2'], p['X_D_2'], Zp, L)
            dilutions = [dil, None]
        kp_1 = _around(p['X_C_1'], p['X_D_1'], p['X_C_2'], p['X_D_2'], dv, Zp, L)
        kp_2 = kp_1.copy()
        kp_2[:, 0] *= -1.0
        coils = [RacetrackCoil(kp_1, y, dv['coil_radius'], I_lvl),
                 RacetrackCoil(kp_2, y, dv['coil_radius'], I_lvl)]
        turn_perimeter = coils[0].get_length() + coils[1].get_length()
        slot_size_horz = min([p['X_C_1'] - p['X_B_1'], p['X_C_2'] - p['X_B_2'],
                              p['X_E_1'] - p['X_D_1'], p['X_E_2'] - p['X_D_2']])
    else:
        raise ValueError('template %r unknown (1, 2, 3 or 4)' % (kind,))
    A_geo = slot_size_horz * slot_size
    A_coil, current_density, M_coil, Q = _figures(dv['current'], J_tar, conductor,
                                                  turn_perimeter, A_geo)
    return Template(model, iron_tags, air_tags, reluctances, dilutions, coils, M_iron, M_coil,
                    Q, current_density, A_coil, turn_perimeter, dv['field_density'],
                    dv['current'])


def build_ncsc(parameters, df_index=0, lc=0.2, materials_directory='files/materials', h=None,
               hz=None):
    """Mesh, coils and figures of the normal-conducting / superconducting
    pair of rows df_index and df_index + 1 (snoopy/SHiP_functions.py:2139-2502):
    the material, conductor and target current density come from the second
    row; iron mass, coil mass and power count the superconducting magnet
    only; the coils keep yoke spacer + insulation (no conductor radius)
    from the iron."""
    i, j = df_index, df_index + 1
    rel = Reluctance(os.path.join(materials_directory, parameters['material'][j]))
    iron = _load(materials_directory, parameters['material'][j])
    conductor = _load(materials_directory, parameters['coil_material'][j])
    J_tar = parameters['J_tar(A/mm2)'][j] * 1e6
    p1, p2 = _row(parameters, i), _row(parameters, j)
    d1, d2 = _drive(parameters, i), _drive(parameters, j)
    lim_x = max([max([p1['X_yoke_1'], p1['X_yoke_2']]) + d1['delta_x'],
                 max([p2['X_yoke_1'], p2['X_yoke_2']]) + d2['delta_x']])
    lim_y = max([max([p1['Y_yoke_1'], p1['Y_yoke_2']]) + d1['delta_y'],
                 max([p2['Y_yoke_1'], p2['Y_yoke_2']]) + d2['delta_y']])
    h, hz = _sizes(lc, h, hz)
    model = ncsc_quarter(p1, p2, lim_x, lim_y, d1['delta_z'], d2['delta_z'], h, hz)
    dens = iron['density(g/m3)']
    lin = lambda k, s: (1.0 - s) * p2[k + '_1'] + s * p2[k + '_2']  # noqa: E731
    core = lambda s: lin('X_core', s) * lin('Y_core', s)  # noqa: E731
    yoke = lambda s: lin('X_yoke', s) * lin('Y_yoke', s) - lin('X_void', s) * lin('Y_void', s)  # noqa: E731
    M_iron = 4 * ruled_volume(core, p2['Z_len']) * dens
    M_iron += 4 * ruled_volume(yoke, p2['Z_len']) * dens
    coils = []
    for p, dv in ((p1, d1), (p2, d2)):
        _, num_cond, y = _levels(p['Y_core_1'], p['Y_core_2'], dv['coil_radius'], dv['ins'],
                                 dv['max_turns'])
        kp = _around(-p['X_core_1'], p['X_core_1'], -p['X_core_2'], p['X_core_2'],
                     dict(dv, coil_radius=0.0), p['Z_pos'], p['Z_len'])
        coils.append(RacetrackCoil(kp, y, dv['coil_radius'], dv['current'] / num_cond))
    turn_perimeter = coils[1].get_length()
    current = d2['current']
    A_cu = abs(current) / J_tar
    M_coil = A_cu * turn_perimeter * conductor['density(g/m3)']
    Q = abs(current * current) * turn_perimeter * conductor['resistivity(Ohm.m)'] / A_cu
    return Template(model, [1, 2, 3], [4], [rel, rel, rel], None, coils, M_iron, M_coil, Q,
                    J_tar, A_cu / conductor['filling_factor'], turn_perimeter,
                    max([d1['field_density'], d2['field_density']]), current)


def get_vector_field_ncsc(parameters, df_index=0, lc=0.2, geo_th=1e-5, run_gmsh=False,
                          plot_geo=False, plot_result=False, result_directory='none',
                          result_spec='', eval_pos=np.zeros((0, 3)),
                          materials_directory='files/materials', quad_order=6, h=None, hz=None,
                          method='reference', verbose=True):
    """Cross-talk of a normal-conducting magnet (row df_index) and a
    superconducting one (row df_index + 1) in one air box
    (snoopy/SHiP_functions.py:2139-2620); returns
    [points, B, M_iron, M_coil, Q, J_tar]."""
    tp = build_ncsc(parameters, df_index, lc, materials_directory, h, hz)
    return _solve_template(tp, parameters, df_index, result_directory, result_spec, eval_pos,
                           quad_order, 50, method, verbose, time.time(),
                           eval_tags=[1, 2, 4, 3])


def _mag4_row(parameters, i):
    """Template 4 takes its six x key points from rows i and i + 1
    (snoopy/SHiP_functions.py:1795-1808)."""
    p = {}
    src = {'A': ('Xmgap', i), 'B': ('Xcore', i), 'C': ('Xvoid', i),
           'D': ('Xcore', i + 1), 'E': ('Xvoid', i + 1), 'F': ('Xyoke', i + 1)}
    for k, (col, r) in src.items():
        for e in (1, 2):
            p['X_%s_%d' % (k, e)] = float(parameters['%s%d(m)' % (col, e)][r])
    for n in ('core', 'void', 'yoke'):
        for e in (1, 2):
            p['Y_%s_%d' % (n, e)] = float(parameters['Y%s%d(m)' % (n, e)][i])
    p['Z_len'] = float(parameters['Z_len(m)'][i])
    p['Z_pos'] = float(parameters['Z_pos(m)'][i])
    return p


# ---------------------------------------------------------------------------
# drivers
# ---------------------------------------------------------------------------

def _run(kind, parameters, df_index, lc, result_directory, result_spec, eval_pos,
         materials_directory, quad_order, use_diluted_steel, kappa_elec, max_newton, h, hz,
         method, verbose):
    t_s = time.time()
    tp = build_template(kind, parameters, df_index, lc, materials_directory,
                        use_diluted_steel, kappa_elec, h, hz)
    return _solve_template(tp, parameters, df_index, result_directory, result_spec, eval_pos,
                           quad_order, max_newton, method, verbose, t_s)


def _solve_template(tp, parameters, df_index, result_directory, result_spec, eval_pos,
                    quad_order, max_newton, method, verbose, t_s, eval_tags=None):
    """Solve a built template, form the field map and the probe fields, write
    the reference's result files; eval_tags orders the domains for the probe
    interpolation (default: iron then air)."""
    if tp.dilutions is not None:
        solver = RedMVPSolverDiluted(tp.model, tp.coils, tp.iron_tags, tp.air_tags,
                                     tp.reluctances, tp.dilutions, quad_order=quad_order,
                                     max_newton_iterations=max_newton, verbose=verbose)
    else:
        solver = RedMVPSolver(tp.model, tp.coils, tp.iron_tags, tp.air_tags, tp.reluctances,
                              quad_order=quad_order, max_newton_iterations=max_newton,
                              verbose=verbose)
    if tp.current == 0.0:
        x = np.zeros((solver.num_dofs_mvp,))
    elif method == 'reference':
        x = solver.solve()
    elif method == 'newton':
        x = solver.solve_newton(precision='mixed', precond='ams').cpu().numpy()
    else:
        raise ValueError("method must be 'reference' or 'newton' (got %r)" % (method,))
    points, B_i = solver.curl_curl_factory.compute_B(x, quad_order=tp.field_density)
    B_coil = 0.0 * B_i
    for coil in tp.coils:
        B_coil += coil.compute_B(points)
    if verbose:
        print('elapsed time = {:.2f} sec'.format(time.time() - t_s))
    if result_directory != 'none':
        np.save(os.path.join(result_directory, 'x' + result_spec + '.npy'), x)
        parameters.iloc[df_index].to_frame().T.to_csv(
            os.path.join(result_directory, 'parameters' + result_spec + '.csv'), index=False)
    ret_vals = [points, B_i + B_coil, tp.M_iron, tp.M_coil, tp.Q, tp.current_density]
    if eval_pos.shape[0] > 0:
        B_tot = evaluate_fem_solution_node_interpolation(
            tp.model, eval_pos, solver, tp.coils, x,
            eval_tags if eval_tags is not None else tp.iron_tags + tp.air_tags)
        if result_directory != 'none':
            fn = os.path.join(result_directory, 'B' + result_spec + '.csv')
        else:
            fn = 'B' + result_spec + '.csv'
        import pandas as pd
        pd.DataFrame(data=np.append(eval_pos, B_tot, axis=1),
                     columns=['x(m)', 'y(m)', 'z(m)', 'Bx(T)', 'By(T)', 'Bz(T)']).to_csv(
                         fn, index=False)
    if verbose:
        print('The iron mass is = {:.2f} t'.format(tp.M_iron * 1e-6))
        print('The coil mass is {:.2f} kg'.format(tp.M_coil * 1e-3))
        print('The coil surface is = {:.2f} m2'.format(tp.A_coil))
        print('The coil perimeter is = {:.2f} m'.format(tp.turn_perimeter))
        print('The current density is = {:.2f} A/mm2'.format(tp.current_density * 1e-6))
        print('The power consumption = {:.2f} W'.format(tp.Q))
    last_run.clear()
    last_run.update(solver.stats, num_elements=tp.model.tet_mesh.num_elements,
                    num_dofs=solver.num_dofs_mvp, seconds=time.time() - t_s)
    return ret_vals


def get_vector_field_mag_1(parameters, df_index=0, lc=0.2, geo_th=1e-5, run_gmsh=False,
                           plot_geo=False, plot_result=False, result_directory='none',
                           result_spec='', eval_pos=np.zeros((0, 3)),
                           materials_directory='files/materials', quad_order=8,
                           max_coil_size=0.05, use_diluted_steel=False, kappa_elec=5.0,
                           h=None, hz=None, method='reference', verbose=True):
    """Normal-conducting template 1, coil around the core
    (snoopy/SHiP_functions.py:282-808)."""
    return _run(1, parameters, df_index, lc, result_directory, result_spec, eval_pos,
                materials_directory, quad_order, use_diluted_steel, kappa_elec, 25, h, hz,
                method, verbose)


def get_vector_field_mag_2(parameters, df_index=0, lc=0.4, geo_th=1e-5, run_gmsh=False,
                           plot_geo=False, plot_result=False, result_directory='none',
                           result_spec='', eval_pos=np.zeros((0, 3)),
                           materials_directory='files/materials', quad_order=8,
                           max_coil_size=0.05, kappa_elec=5.0, h=None, hz=None,
                           method='reference', verbose=True):
    """Superconducting template 2: core and yoke of type 2, one coil around
    the core (snoopy/SHiP_functions.py:812-1216)."""
    return _run(2, parameters, df_index, lc, result_directory, result_spec, eval_pos,
                materials_directory, quad_order, False, kappa_elec, 25, h, hz, method,
                verbose)


def get_vector_field_mag_3(parameters, df_index=0, lc=0.2, geo_th=1e-5, run_gmsh=False,
                           plot_geo=False, plot_result=False, result_directory='none',
                           result_spec='', eval_pos=np.zeros((0, 3)),
                           materials_directory='files/materials', quad_order=8,
                           max_coil_size=0.05, use_diluted_steel=False, kappa_elec=5.0,
                           h=None, hz=None, method='reference', verbose=True):
    """Normal-conducting template 3, coils around the legs
    (snoopy/SHiP_functions.py:1218-1710)."""
    return _run(3, parameters, df_index, lc, result_directory, result_spec, eval_pos,
                materials_directory, quad_order, use_diluted_steel, kappa_elec, 25, h, hz,
                method, verbose)


def get_vector_field_mag_4(parameters, df_index=0, lc=0.2, geo_th=1e-5, run_gmsh=False,
                           plot_geo=False, plot_result=False, result_directory='none',
                           result_spec='', eval_pos=np.zeros((0, 3)),
                           materials_directory='files/materials', quad_order=8,
                           max_coil_size=0.05, use_diluted_steel=False, kappa_elec=5.0,
                           h=None, hz=None, method='reference', verbose=True):
    """Two-part template 4 from rows df_index and df_index + 1, coils around
    the core (snoopy/SHiP_functions.py:1712-2137)."""
    return _run(4, parameters, df_index, lc, result_directory, result_spec, eval_pos,
                materials_directory, quad_order, use_diluted_steel, kappa_elec, 150, h, hz,
                method, verbose)
