"""Helpers shared by the GPU tests (device buffers through the C ABI)."""
import numpy as np

import del3d_amd as D


def require_gpu():
    n = D.device_count()
    assert n >= 1, "no GPU visible to the HIP runtime"
    D.lib()


def predicates_dev(pts, op, idx, stats=False):
    """Runs the batched C ABI on device copies; returns (signs, stats or None)."""
    xyz = D.DeviceArray.from_host(np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3))
    dop = D.DeviceArray.from_host(np.ascontiguousarray(op, dtype=np.uint8))
    didx = D.DeviceArray.from_host(np.ascontiguousarray(idx, dtype=np.int32).reshape(-1, 5))
    out = D.DeviceArray((len(op),), np.int8)
    st = None
    if stats:
        st = D.predicates_stats(xyz, dop, didx, out)
    else:
        D.predicates(xyz, dop, didx, out)
        D.synchronize()
    return out.to_host(), st


def say(*a):
    print(*a, flush=True)
