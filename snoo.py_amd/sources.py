"""Source-field evaluation on the device (Biot-Savart and brick fields).

The reference evaluates coil fields with OpenMP C (fem_c/ctools.c:1761-1923)
behind compute_B_line_segs_cpp (fem_c/interface_c.pyx:928-991) and brick
fields with a Python loop (snoopy/brick_field.py:400-433).  Here both are HIP
kernels; ``source_field_device`` sums any list of coil objects into one device
buffer, keeping the reference's duck-typed contract (any object with
``compute_B(points) -> (M,3)``) through a host round trip for unknown types.
"""
import numpy as np

from . import _native as nat
from .device import device, as_dev_f64


def biot_savart_device(segs_dev, tar_dev, current, rad, out_dev, accumulate=False, stream=None):
    lib = nat.load()
    nat.check(lib.sn_biot_savart_f64(nat.ptr(segs_dev), int(segs_dev.shape[0]), nat.ptr(tar_dev),
                                     int(tar_dev.shape[0]), float(current), float(rad),
                                     nat.ptr(out_dev), int(bool(accumulate)),
                                     nat.stream_handle(stream)))
    return out_dev


def biot_savart_host(segs, positions, current, rad):
    """(M,3) numpy B of line segments at numpy positions; errors raise."""
    segs = np.asarray(segs, dtype=np.float64)
    positions = np.asarray(positions, dtype=np.float64)
    if segs.ndim != 2 or segs.shape[1] != 6:
        raise ValueError('the source array must be N x 6 (got {})'.format(segs.shape))
    if positions.ndim != 2 or positions.shape[1] != 3:
        raise ValueError('the target array must be M x 3 (got {})'.format(positions.shape))
    import torch
    s = as_dev_f64(segs)
    t = as_dev_f64(positions)
    out = torch.empty_like(t)
    biot_savart_device(s, t, current, rad, out)
    return out.cpu().numpy()


def brick_field_device(brick, tar_dev, out_dev, accumulate=False, stream=None):
    lib = nat.load()
    faces = np.ascontiguousarray(brick.face_table())
    fields = np.ascontiguousarray(brick.field_values())
    nat.check(lib.sn_brick_field_f64(faces.ctypes.data_as(nat._vp), fields.ctypes.data_as(nat._vp),
                                     nat.ptr(tar_dev), int(tar_dev.shape[0]), nat.ptr(out_dev),
                                     int(bool(accumulate)), nat.stream_handle(stream)))
    return out_dev


def brick_field_host(brick, points):
    import torch
    t = as_dev_f64(np.asarray(points, dtype=np.float64).reshape(-1, 3))
    out = torch.empty_like(t)
    brick_field_device(brick, t, out)
    return out.cpu().numpy()


def segments_device(coil, disc_arc=5):
    """The coil's segments (device float64 [M,6]), kept with the coil and
    formed again when its key points, levels or ``disc_arc`` change.  Their
    upload from pageable memory would otherwise wait for the work already
    queued on the stream (the previous coil's Biot-Savart)."""
    import torch
    key = (np.asarray(coil.kp, dtype=np.float64).tobytes(),
           np.asarray(coil.y, dtype=np.float64).tobytes(), int(disc_arc))
    hit = getattr(coil, '_segments_dev', None)
    if hit is None or hit[0] != key:
        if hit is not None:
            torch.cuda.synchronize()       # no queued kernel may still read the old array
        hit = (key, as_dev_f64(coil.get_segments(disc_arc=disc_arc)))
        coil._segments_dev = hit
    return hit[1]


def source_field_device(coil_list, points_dev, disc_arc=5):
    """Sum of the coil fields at device points [M,3] (float64)."""
    import torch
    from .brick_field import BrickField
    from .racetrack_coil import RacetrackCoil
    out = torch.zeros_like(points_dev)
    for coil in coil_list:
        if isinstance(coil, RacetrackCoil):
            segs = segments_device(coil, disc_arc)
            biot_savart_device(segs, points_dev, coil.current, coil.rad, out, accumulate=True)
        elif isinstance(coil, BrickField):
            brick_field_device(coil, points_dev, out, accumulate=True)
        else:
            # duck-typed plugin: evaluate on the host, add on the device
            B = np.asarray(coil.compute_B(points_dev.cpu().numpy()), dtype=np.float64)
            out += torch.from_numpy(B.reshape(-1, 3)).to(device())
    return out
