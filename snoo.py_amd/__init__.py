"""snoopy_amd: the nonlinear reduced-vector-potential solve of snoo.py on
AMD Instinct MI355X (gfx950).

Drop-in classes with the reference's names and signatures:
RedMVPSolver, CurlCurlAssembler, Reluctance, ConstantReluctance,
RacetrackCoil, BrickField.  Behind them the hot path runs as hand-written HIP
kernels in libsnoo_hip.so (C ABI: include/snoo_hip.h).
"""
from .brick_field import BrickField
from .curl_curl_assembler import CurlCurlAssembler, CurlCurlAssemblerDiluted
from .evaluators import evaluate_fem_solution_node_interpolation
from .finite_element import FiniteElement
from .materials import (BSplineReluctance, ConstantReluctance, Dilution, Reluctance,
                        tanh_reluctance)
from .mesh import TetMesh, TetModel, kuhn_mesh
from .mesh_tools import get_cotree_dofs, get_edge_boundary_dofs
from .racetrack_coil import RacetrackCoil
from .red_mvp_solver import NewtonDriver, RedMVPSolver, RedMVPSolverDiluted
from .geometry_tools import compute_area_polygon
from .SHiP_functions import (compute_J_opt, compute_prices, get_NI, get_vector_field_mag_1,
                             get_vector_field_mag_2, get_vector_field_mag_3,
                             get_vector_field_mag_4, get_vector_field_ncsc)

__all__ = ['BrickField', 'CurlCurlAssembler', 'FiniteElement', 'BSplineReluctance',
           'ConstantReluctance', 'Reluctance', 'tanh_reluctance', 'TetMesh', 'TetModel',
           'kuhn_mesh', 'get_cotree_dofs', 'get_edge_boundary_dofs', 'RacetrackCoil',
           'NewtonDriver', 'RedMVPSolver', 'RedMVPSolverDiluted', 'CurlCurlAssemblerDiluted',
           'Dilution', 'evaluate_fem_solution_node_interpolation', 'compute_area_polygon',
           'compute_J_opt', 'compute_prices', 'get_NI', 'get_vector_field_mag_1',
           'get_vector_field_mag_2', 'get_vector_field_mag_3', 'get_vector_field_mag_4',
           'get_vector_field_ncsc']
