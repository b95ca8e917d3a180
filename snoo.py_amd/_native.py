"""ctypes binding of libsnoo_hip.so (declarations in include/snoo_hip.h).

The library is loaded after torch so that it binds to the HIP runtime torch
has already loaded.  There is no fallback: if the library is missing or fails
to load, every GPU entry point raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libsnoo_hip.so')

_c_int, _c_dbl, _c_ll, _vp = ctypes.c_int, ctypes.c_double, ctypes.c_longlong, ctypes.c_void_p
_P_int = ctypes.POINTER(ctypes.c_int)
_P_dbl = ctypes.POINTER(ctypes.c_double)
_P_ll = ctypes.POINTER(ctypes.c_longlong)

# name -> (restype, argtypes)
SIGNATURES = {
    'sn_last_error': (ctypes.c_char_p, []),
    'sn_version': (_c_int, []),
    'sn_biot_savart_f64': (_c_int, [_vp, _c_int, _vp, _c_int, _c_dbl, _c_dbl, _vp, _c_int, _vp]),
    'sn_brick_field_f64': (_c_int, [_vp, _vp, _vp, _c_int, _vp, _c_int, _vp]),
    'sn_quad_points_f64': (_c_int, [_c_int, _vp, _vp, _vp, _c_int, _vp, _vp]),
    'sn_B_Hcurl_f64': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _vp]),
    'sn_element_table_f64': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _c_int,
                                      _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    'sn_B_Hcurl_table_f64': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _c_int,
                                      _vp, _vp, _vp, _vp]),
    'sn_pattern_create': (_c_int, [ctypes.POINTER(_vp), _c_int, _c_int, _vp, _vp]),
    'sn_pattern_destroy': (None, [_vp]),
    'sn_pattern_info': (_c_int, [_vp, _P_ll, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                 ctypes.POINTER(_vp)]),
    'sn_pattern_copy': (_c_int, [_vp, _vp, _vp, _vp, _vp]),
    'sn_element_local_f64': (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                      _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                      _vp, _vp, _vp]),
    'sn_csr_gather_f64': (_c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    'sn_dof_gather_f64': (_c_int, [_vp, _vp, _vp, _vp]),
    'sn_element_newton': (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp,
                                   _vp, _vp, _vp, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp,
                                   _vp, _vp, _vp, _vp, _vp, _vp]),
    'sn_csr_gather_one': (_c_int, [_vp, _c_int, _vp, _vp, _vp]),
    'sn_spmv_f64': (_c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    'sn_dot_f64': (_c_int, [_c_int, _vp, _vp, _vp, _P_dbl, _vp]),
    'sn_dot_f64_dev': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp]),
    'sn_seg_dot_f64': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    'sn_seg_dot_f64_noalloc': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    'sn_absmax_f64': (_c_int, [_c_int, _vp, _vp, _P_dbl, _vp]),
    'sn_cg_create': (_c_int, [ctypes.POINTER(_vp), _vp, _c_int, _c_int, _c_int, _vp]),
    'sn_cg_destroy': (None, [_vp]),
    'sn_cg_create_batched': (_c_int, [ctypes.POINTER(_vp), _vp, _c_int, _c_int, _c_int, _c_int,
                                      _vp, _vp]),
    'sn_cg_solve_batched': (_c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _c_dbl, _c_int, _vp, _vp, _vp,
                                     _vp]),
    'sn_cg_time_spmv': (_c_int, [_vp, _vp, _vp, _c_int, _P_dbl, _vp]),
    'sn_cg_solve': (_c_int, [_vp, _vp, _vp, _vp, _vp, _c_dbl, _c_dbl, _c_int, _P_int, _P_dbl,
                             _vp]),
    'sn_gemm_bf16': (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _c_int, _vp,
                              _c_int, _vp, _vp, _vp, _vp, _vp, _c_int, _vp]),
    'sn_mse_grad': (_c_int, [_c_int, _vp, _vp, ctypes.c_float, _vp, _vp, _P_dbl, _vp]),
    'sn_colsum_f32': (_c_int, [_c_int, _c_int, _vp, _vp, _vp]),
    'sn_adam_f32': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, ctypes.c_float, ctypes.c_float,
                             ctypes.c_float, ctypes.c_float, _c_int, ctypes.c_float, _vp]),
    'sn_f32_to_bf16': (_c_int, [_c_int, _vp, _vp, _vp]),
    'sn_pod_expand': (_c_int, [_c_int, _c_int, _vp, _vp, _vp, _vp, _vp]),
    'sn_pod_expand_many': (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp]),
    'sn_pod_project_many': (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp]),
    'sn_pod_reduced_jacobian_range_dev': (_c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp]),
    'sn_pod_project_many_dev': (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp]),
    'sn_pod_reduced_solve_many': (_c_int, [_c_int, _c_int, _vp, _vp, _vp, _vp]),
    'sn_pod_element_modes': (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp]),
    'sn_pod_reduced_jacobian': (_c_int, [_c_int, _c_int, _vp, _vp, _vp, _vp]),
    'sn_pod_reduced_jacobian_range': (_c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp,
                                               _vp]),
    'sn_pod_project': (_c_int, [_c_int, _c_int, _vp, _vp, _vp, _vp]),
    'sn_edge_transfer': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp]),
    'sn_ams_create': (_c_int, [ctypes.POINTER(_vp), _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp,
                               ctypes.c_float, ctypes.c_float, _vp]),
    'sn_ams_add_level': (_c_int, [_vp, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    'sn_ams_update': (_c_int, [_vp, _vp, _vp, _vp, _vp]),
    'sn_ams_apply': (_c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    'sn_ams_num_levels': (_c_int, [_vp]),
    'sn_amg_aggregate': (_c_int, [_c_int, _vp, _vp, _vp, ctypes.c_double, _vp,
                                  ctypes.POINTER(_c_int)]),
    'sn_ams_level_info': (_c_int, [_vp, _c_int, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int)]),
    'sn_ams_level_values': (_c_int, [_vp, _c_int, _vp, _vp]),
    'sn_ams_destroy': (None, [_vp]),
    'sn_cg_set_ams': (_c_int, [_vp, _vp]),
    'sn_nodal_average_f64': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp]),
    'sn_p1_interp_f64': (_c_int, [_c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _vp, _vp]),
    'sn_cg_set_schedule': (_c_int, [_vp, _c_int]),
    'sn_cg_set_col_offsets': (_c_int, [_vp, _c_int]),
    'sn_cg_index_bytes': (_c_int, [_vp]),
    'sn_cg_set_element_operator': (_c_int, [_vp, _vp, _vp, _c_int, _vp]),
    'sn_cg_element_operator': (_c_int, [_vp]),
    'sn_cg_set_grids': (_c_int, [_vp, _c_int, _c_int, _vp]),
    'sn_ams_set_lanes': (_c_int, [_vp, _c_int]),
    'sn_ams_set_restriction': (_c_int, [_vp, _c_int]),
    'sn_cg_grids': (_c_int, [_vp, _vp, _vp]),
    'sn_vcg_create': (_c_int, [ctypes.POINTER(_vp), _vp, _vp, _vp, _c_int, _c_int, _vp]),
    'sn_vcg_set_ams': (_c_int, [_vp, _vp, _vp]),
    'sn_vcg_set_schedule': (_c_int, [_vp, _c_int]),
    'sn_vcg_solve': (_c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _c_dbl, _c_int, _vp, _vp, _vp, _vp]),
    'sn_vcg_info': (_c_int, [_vp, _P_int, _P_int, _P_int]),
    'sn_vcg_time_pair': (_c_int, [_vp, _vp, _vp, _c_int, _P_dbl, _vp]),
    'sn_vcg_set_grids': (_c_int, [_vp, _c_int, _c_int, _vp]),
    'sn_vcg_pack': (_c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    'sn_vcg_destroy': (None, [_vp]),
    'sn_element_apply_f64': (_c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
}

SN_OK, SN_ERR_ARG, SN_ERR_HIP, SN_ERR_NOT_CONVERGED, SN_ERR_BREAKDOWN = 0, 1, 2, 3, 4


class SnooError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__('snoo_hip error {}: {}'.format(code, msg))
        self.code = code


class NotConverged(SnooError):
    pass


class Breakdown(SnooError):
    pass


_lib = None


def load(path=LIB_PATH):
    """Load (once) and return the ctypes library object."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError('libsnoo_hip.so not built ({}); run __graft_entry__.build()'.format(path))
    import torch  # noqa: F401  (bind to torch's HIP runtime)
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc == SN_OK:
        return
    msg = load().sn_last_error().decode(errors='replace')
    if rc == SN_ERR_NOT_CONVERGED:
        raise NotConverged(rc, msg)
    if rc == SN_ERR_BREAKDOWN:
        raise Breakdown(rc, msg)
    raise SnooError(rc, msg)


def ptr(t):
    """Device (or host) address of a torch tensor, None for None.  The kernels
    read dense row-major arrays: a strided view (a transpose, a column slice,
    or a tensor made from a Fortran-ordered numpy array) raises instead of
    handing its storage over in the wrong order."""
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError('a strided tensor (shape {}, strides {}) cannot be passed to the C ABI; '
                         'call .contiguous() first'.format(tuple(t.shape), t.stride()))
    return ctypes.c_void_p(t.data_ptr())


def ptr_np(a):
    """Host address of a C-contiguous numpy array, None for None."""
    if a is None:
        return None
    assert a.flags['C_CONTIGUOUS']
    return ctypes.c_void_p(a.ctypes.data)


def stream_handle(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)
