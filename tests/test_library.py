"""CPU tests of the C-ABI library: it loads, and exports every symbol that
include/snoo_hip.h declares (no compute calls without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT


def _declared():
    txt = open(os.path.join(ROOT, 'include', 'snoo_hip.h')).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(sn_[A-Za-z0-9_]+)\s*\(', txt)))


def test_header_declares_the_expected_entry_points():
    names = _declared()
    for n in ('sn_biot_savart_f64', 'sn_pattern_create', 'sn_element_local_f64',
              'sn_csr_gather_f64', 'sn_element_newton', 'sn_cg_solve', 'sn_last_error'):
        assert n in names


def test_library_exports_every_declared_symbol():
    from snoopy_amd import _native as nat
    lib = ctypes.CDLL(nat.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    from snoopy_amd import _native as nat
    assert sorted(nat.SIGNATURES) == _declared()
    lib = nat.load()
    assert lib.sn_version() >= 1
    assert isinstance(lib.sn_last_error(), bytes)


def test_strided_tensors_are_refused_at_the_abi():
    """The kernels read dense row-major arrays: a transposed view, a column
    slice or a tensor built from a Fortran-ordered numpy array must not reach
    them (its storage order differs from its index order)."""
    import numpy as np
    import pytest
    import torch
    from snoopy_amd import _native as nat
    a = torch.zeros(4, 3)
    assert nat.ptr(a).value == a.data_ptr()
    for bad in (a.t(), a[:, :2], torch.from_numpy(np.asfortranarray(np.zeros((4, 3))))):
        with pytest.raises(ValueError, match='strided'):
            nat.ptr(bad)
