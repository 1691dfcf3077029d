"""The predicate trace format: the predicate calls of one run of a Delaunay
driver, as (op, result, point indices) records over its point array, for
replay through the mixed batch (del3d_predicates_mixed).

File layout (little-endian):
  "D3TR"                 4 bytes
  version                u32 (2)
  n_points, n_records    i64, i64
  points                 3 * n_points f64 (x, y, z per point)
  op                     n_records u8: the mixed-batch codes of
                         include/del3d_amd.h (0 orient3d, 1 in_sphere,
                         2 in_sphere_unperturbed, 16 + DEL3D_OP2_*)
  result                 n_records i8: the value the call returned
  idx                    5 * n_records i32: the call's points in argument
                         order, unused entries -1
Version 1 files (orient3d and in_sphere only) read the same way.

A record's points are indices, not coordinates: the reference's
perturbations compare pointer identity (src/predicates.cpp:220-227,
src/predicates2.cpp:249-255), which indices preserve.
"""
from __future__ import annotations

import struct
from pathlib import Path

import numpy as np

MAGIC = b"D3TR"
VERSION = 2


def read(path):
    """(points (n,3) f64, op u8, result i8, idx (m,5) i32) of a trace file."""
    raw = Path(path).read_bytes()
    if raw[:4] != MAGIC:
        raise ValueError(f"{path}: not a predicate trace")
    (version,) = struct.unpack_from("<I", raw, 4)
    if version not in (1, 2):
        raise ValueError(f"{path}: trace version {version} not supported")
    n, m = struct.unpack_from("<qq", raw, 8)
    off = 24
    pts = np.frombuffer(raw, np.float64, 3 * n, off).reshape(n, 3)
    off += 24 * n
    op = np.frombuffer(raw, np.uint8, m, off)
    off += m
    res = np.frombuffer(raw, np.int8, m, off)
    off += m
    idx = np.frombuffer(raw, np.int32, 5 * m, off).reshape(m, 5)
    if off + 20 * m != len(raw):
        raise ValueError(f"{path}: truncated or oversized trace")
    return pts, op, res, idx


def write(path, pts, op, res, idx) -> None:
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    op = np.ascontiguousarray(op, np.uint8)
    res = np.ascontiguousarray(res, np.int8)
    idx = np.ascontiguousarray(idx, np.int32).reshape(-1, 5)
    if not (len(op) == len(res) == len(idx)):
        raise ValueError("op, result and idx lengths differ")
    with open(path, "wb") as f:
        f.write(MAGIC + struct.pack("<Iqq", VERSION, len(pts), len(op)))
        for a in (pts, op, res, idx):
            f.write(a.tobytes())


def replay(pts, op, idx, stream=None) -> np.ndarray:
    """Results of every record, recomputed on the GPU in one mixed batch."""
    from . import DeviceArray, predicates_mixed, synchronize
    xyz = DeviceArray.from_host(np.ascontiguousarray(pts, np.float64).reshape(-1, 3))
    dop = DeviceArray.from_host(np.ascontiguousarray(op, np.uint8))
    didx = DeviceArray.from_host(np.ascontiguousarray(idx, np.int32).reshape(-1, 5))
    out = DeviceArray((len(op),), np.int8)
    predicates_mixed(xyz, dop, didx, out, stream)
    synchronize()
    return out.to_host()
