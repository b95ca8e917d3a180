"""GPU test of the allocation-free segmented dot, sn_seg_dot_f64_noalloc,
against sn_seg_dot_f64 through the C ABI: bit-identical sums (same kernel,
same order) on two calls in a row, the second with fewer and smaller
segments, so the scratch array's leftover contents from the first call must
not show; and the size limit rejected before any device work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_vp, _ci = ctypes.c_void_p, ctypes.c_int


def _fn(lib):
    f = lib.sn_seg_dot_f64_noalloc
    f.restype = _ci
    f.argtypes = [_ci, _vp, _vp, _vp, _vp, _vp, _vp]
    return f


def _run(nat, fn, row0, a, b, mask):
    import torch
    nseg = len(row0) - 1
    r0 = torch.from_numpy(np.asarray(row0, dtype=np.int32)).cuda()
    out = np.full(nseg, np.nan)
    ref = np.full(nseg, np.nan)
    args = (nat.ptr(r0), nat.ptr(a), nat.ptr(b), nat.ptr(mask))
    nat.check(fn(nseg, *args, out.ctypes.data_as(_vp), nat.stream_handle()))
    nat.check(nat.load().sn_seg_dot_f64(nseg, *args, ref.ctypes.data_as(_vp),
                                        nat.stream_handle()))
    return out, ref


def test_noalloc_seg_dot_matches_the_allocating_entry_point():
    import torch
    from snoopy_amd import _native as nat
    lib = nat.load()
    fn = _fn(lib)
    rng = np.random.default_rng(4)
    n = 3_000_000
    a = torch.from_numpy(rng.standard_normal(n)).cuda()
    b = torch.from_numpy(rng.standard_normal(n)).cuda()
    mask = torch.from_numpy((rng.random(n) > 0.1).astype(np.uint8)).cuda()
    print('case: noalloc seg dot, 4096 segments then 37', flush=True)
    row0 = np.sort(rng.integers(0, n, 4095))
    row0 = np.concatenate([[0], row0, [n]])
    out, ref = _run(nat, fn, row0, a, b, mask)
    np.testing.assert_array_equal(out, ref)
    # host float64 check of a few segments (order differs: 1e-12 relative)
    ah, bh, mh = a.cpu().numpy(), b.cpu().numpy(), mask.cpu().numpy().astype(bool)
    for v in (0, 1000, 4095):
        sl = slice(row0[v], row0[v + 1])
        exact = np.sum((ah * bh)[sl][mh[sl]])
        assert abs(out[v] - exact) <= 1e-12 * max(1.0, np.abs(ah[sl] * bh[sl]).sum())
    # second call: fewer, smaller segments, no mask
    row1 = np.linspace(0, 50_000, 38).astype(np.int64)
    out1, ref1 = _run(nat, fn, row1, a, b, None)
    np.testing.assert_array_equal(out1, ref1)
    assert np.all(np.isfinite(out1))


def test_noalloc_seg_dot_rejects_more_segments_than_its_scratch():
    from snoopy_amd import _native as nat
    fn = _fn(nat.load())
    dummy = ctypes.c_void_p(1)
    out = np.zeros(4097)
    rc = fn(4097, dummy, dummy, dummy, None, out.ctypes.data_as(_vp), None)
    assert rc == nat.SN_ERR_ARG
    assert b'scratch maximum of 4096' in nat.load().sn_last_error()
