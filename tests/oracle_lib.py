"""ctypes access to the checker libraries (test infrastructure only).

* ``oracle/liboracle.so`` - the plain-C restatement of del3d's hot path.
* ``oracle/_ref/libdel3d_ref.so`` - the reference itself, compiled from its own
  sources by ``oracle/Makefile`` where those sources exist. Optional: on a
  machine without it, ``ref()`` returns None and dependent checks skip.

Nothing in the product imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libdel3d_ref.so"
REF_DRIVER_SO = ROOT / "oracle" / "_ref" / "libdel3d_ref_driver.so"
TRACE_RECORDER = ROOT / "oracle" / "_ref" / "trace_recorder"

_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_u32p = C.POINTER(C.c_uint32)
_u8p = C.POINTER(C.c_uint8)
_i8p = C.POINTER(C.c_int8)

_oracle = None
_ref = None
_ref_driver = None


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def oracle() -> C.CDLL:
    global _oracle
    if _oracle is None:
        if not ORACLE_SO.exists():
            raise RuntimeError(f"{ORACLE_SO} is not built (run __graft_entry__.build())")
        lib = C.CDLL(str(ORACLE_SO))
        for name in ("orc_orient3d", "orc_int_orient3d"):
            getattr(lib, name).argtypes = [_dp] * 4
            getattr(lib, name).restype = C.c_int
        for name in ("orc_in_sphere", "orc_in_sphere_unperturbed",
                     "orc_int_in_sphere", "orc_int_in_sphere_raw"):
            getattr(lib, name).argtypes = [_dp] * 5
            getattr(lib, name).restype = C.c_int
        for name in ("orc_pred_batch", "orc_int_pred_batch", "orc_pred2_batch",
                     "orc_int_pred2_batch"):
            getattr(lib, name).argtypes = [_dp, C.c_int64, _u8p, _i32p, C.c_int64, _i8p]
            getattr(lib, name).restype = None
        lib.orc_last_stage_counts.argtypes = [_i64p]
        lib.orc_spatial_sort.argtypes = [_dp, C.c_int64, _i32p]
        lib.orc_shuffle.argtypes = [C.c_int64, _i32p]
        lib.orc_shuffle_rejections.argtypes = [C.c_int64]
        lib.orc_shuffle_rejections.restype = C.c_int64
        lib.orc_rand48_draws.argtypes = [C.c_uint32, C.c_int64, _u32p]
        lib.orc_spatial_sort_visits.argtypes = [_dp, C.c_int64]
        lib.orc_spatial_sort_visits.restype = C.c_int64
        lib.orc_hash_order.argtypes = [_i32p, C.c_int64]
        lib.orc_hash_order.restype = C.c_uint64
        for name in ("orc_pts_cylinder", "orc_pts_sphere", "orc_pts_torus"):
            getattr(lib, name).argtypes = [C.c_int64, C.c_int, C.c_uint32, _dp]
            getattr(lib, name).restype = C.c_int64
        lib.orc_pts_box_shell.argtypes = [C.c_int, _dp]
        lib.orc_pts_box_shell.restype = C.c_int64
        lib.orc_pts_lattice.argtypes = [C.c_int, C.c_int, C.c_int, _dp]
        lib.orc_pts_lattice.restype = C.c_int64
        lib.orc_pts_random_cube.argtypes = [C.c_int64, C.c_uint32, _dp]
        lib.orc_pts_random_cube.restype = C.c_int64
        lib.orc_circumcenters.argtypes = [_dp, _i32p, C.c_int64, C.c_int, _dp]
        lib.orc_mt19937_words.argtypes = [C.c_uint32, C.c_int64, _u32p]
        lib.orc_rng_below.argtypes = [C.c_uint32, C.c_uint64, C.c_int64, _i64p]
        _oracle = lib
    return _oracle


def ref():
    """The compiled reference, or None where it was not built."""
    global _ref
    if _ref is None:
        if not REF_SO.exists():
            return None
        lib = C.CDLL(str(REF_SO))
        lib.ref_orient3d.argtypes = [_dp] * 4
        lib.ref_orient3d.restype = C.c_int
        lib.ref_in_sphere.argtypes = [_dp] * 5
        lib.ref_in_sphere.restype = C.c_int
        lib.ref_in_sphere_unperturbed.argtypes = [_dp] * 5
        lib.ref_in_sphere_unperturbed.restype = C.c_int
        lib.ref_pred_batch.argtypes = [_dp, C.c_int64, _u8p, _i32p, C.c_int64, _i8p]
        lib.ref_pred2_batch.argtypes = [_dp, C.c_int64, _u8p, _i32p, C.c_int64, _i8p]
        lib.ref_circumcenters.argtypes = [_dp, _i32p, C.c_int64, C.c_int, _dp]
        lib.ref_pred_batch_mt.argtypes = [_dp, C.c_int64, _u8p, _i32p, C.c_int64, _i8p, C.c_int]
        lib.ref_spatial_sort.argtypes = [_dp, C.c_int64, _i32p]
        lib.ref_delaunay_cells.argtypes = [_dp, C.c_int64, _i32p, C.c_int64]
        lib.ref_delaunay_cells.restype = C.c_int64
        _ref = lib
    return _ref


# ---- convenience wrappers ---------------------------------------------------

def _xyz(pts) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(pts, dtype=np.float64).reshape(-1, 3))
    return a


def pred_batch(pts, op, idx5, which: str = "oracle", threads: int = 1) -> np.ndarray:
    """Signs of a query batch from the restatement ("oracle"), the big-integer
    oracle ("int") or the compiled reference ("ref"; `threads` > 1 splits the
    batch over that many std::threads)."""
    xyz = _xyz(pts)
    op = np.ascontiguousarray(op, dtype=np.uint8)
    idx5 = np.ascontiguousarray(np.asarray(idx5, dtype=np.int32).reshape(-1, 5))
    out = np.zeros(len(op), dtype=np.int8)
    if which == "oracle":
        oracle().orc_pred_batch(_ptr(xyz, _dp), len(xyz), _ptr(op, _u8p), _ptr(idx5, _i32p),
                                len(op), _ptr(out, _i8p))
    elif which == "int":
        oracle().orc_int_pred_batch(_ptr(xyz, _dp), len(xyz), _ptr(op, _u8p), _ptr(idx5, _i32p),
                                    len(op), _ptr(out, _i8p))
    elif which == "ref":
        r = ref()
        if r is None:
            raise RuntimeError("reference library not built")
        if threads > 1:
            r.ref_pred_batch_mt(_ptr(xyz, _dp), len(xyz), _ptr(op, _u8p), _ptr(idx5, _i32p),
                                len(op), _ptr(out, _i8p), int(threads))
        else:
            r.ref_pred_batch(_ptr(xyz, _dp), len(xyz), _ptr(op, _u8p), _ptr(idx5, _i32p),
                             len(op), _ptr(out, _i8p))
    else:
        raise ValueError(which)
    return out


def pred2_batch(pts, op, idx5, which: str = "oracle") -> np.ndarray:
    """Results of a coplanar batch (op codes ORC_OP2_*) from the restatement
    ("oracle"), the big-integer apex oracle ("int", exactly coplanar queries
    only) or the compiled reference ("ref")."""
    xyz = _xyz(pts)
    op = np.ascontiguousarray(op, dtype=np.uint8)
    idx5 = np.ascontiguousarray(np.asarray(idx5, dtype=np.int32).reshape(-1, 5))
    out = np.zeros(len(op), dtype=np.int8)
    args = (_ptr(xyz, _dp), len(xyz), _ptr(op, _u8p), _ptr(idx5, _i32p), len(op), _ptr(out, _i8p))
    if which == "oracle":
        oracle().orc_pred2_batch(*args)
    elif which == "int":
        oracle().orc_int_pred2_batch(*args)
    elif which == "ref":
        r = ref()
        if r is None:
            raise RuntimeError("reference library not built")
        r.ref_pred2_batch(*args)
        inexact = np.nonzero(op == 8)[0]
        if len(inexact):
            sub = np.ascontiguousarray(idx5[inexact])
            res = np.zeros(len(inexact), np.int8)
            ref_driver().ref_inexact_orientation_batch(_ptr(xyz, _dp), _ptr(sub, _i32p),
                                                       len(inexact), _ptr(res, _i8p))
            out[inexact] = res
    else:
        raise ValueError(which)
    return out


def ref_driver():
    """The reference driver's private plain-double orientation (oracle/
    ref_driver.cpp), or an error where it was not built."""
    global _ref_driver
    if _ref_driver is None:
        if not REF_DRIVER_SO.exists():
            raise RuntimeError("reference driver library not built")
        lib = C.CDLL(str(REF_DRIVER_SO))
        lib.ref_inexact_orientation_batch.argtypes = [_dp, _i32p, C.c_int64, _i8p]
        _ref_driver = lib
    return _ref_driver


def circumcenters(pts, cells, lazy: int, which: str = "oracle") -> np.ndarray:
    """(n, 3) circumcentres of cells (4 point indices each) from the
    restatement ("oracle") or the compiled reference ("ref")."""
    xyz = _xyz(pts)
    cells = np.ascontiguousarray(np.asarray(cells, dtype=np.int32).reshape(-1, 4))
    out = np.zeros((len(cells), 3))
    args = (_ptr(xyz, _dp), _ptr(cells, _i32p), len(cells), int(lazy), _ptr(out, _dp))
    if which == "oracle":
        oracle().orc_circumcenters(*args)
    elif which == "ref":
        r = ref()
        if r is None:
            raise RuntimeError("reference library not built")
        r.ref_circumcenters(*args)
    else:
        raise ValueError(which)
    return out


def delaunay_cells(pts) -> np.ndarray:
    """The finite cells of the reference's Delaunay triangulation, in its order."""
    r = ref()
    if r is None:
        raise RuntimeError("reference library not built")
    xyz = _xyz(pts)
    n = r.ref_delaunay_cells(_ptr(xyz, _dp), len(xyz), None, 0)
    cells = np.zeros((n, 4), np.int32)
    r.ref_delaunay_cells(_ptr(xyz, _dp), len(xyz), _ptr(cells, _i32p), n)
    return cells


def stage_counts() -> tuple[int, int, int]:
    out = np.zeros(3, dtype=np.int64)
    oracle().orc_last_stage_counts(_ptr(out, _i64p))
    return int(out[0]), int(out[1]), int(out[2])


def spatial_sort(pts, order=None, which: str = "oracle") -> np.ndarray:
    """order (default 0..n-1) holds indices into pts; it may be shorter."""
    xyz = _xyz(pts)
    o = np.arange(len(xyz), dtype=np.int32) if order is None else np.array(order, dtype=np.int32)
    n = len(o)
    if which == "oracle":
        oracle().orc_spatial_sort(_ptr(xyz, _dp), n, _ptr(o, _i32p))
    elif which == "ref":
        r = ref()
        if r is None:
            raise RuntimeError("reference library not built")
        r.ref_spatial_sort(_ptr(xyz, _dp), n, _ptr(o, _i32p))
    else:
        raise ValueError(which)
    return o


def shuffle(n: int) -> np.ndarray:
    o = np.arange(n, dtype=np.int32)
    oracle().orc_shuffle(n, _ptr(o, _i32p))
    return o


def hash_order(order) -> int:
    o = np.ascontiguousarray(order, dtype=np.int32)
    return int(oracle().orc_hash_order(_ptr(o, _i32p), len(o)))


def rand48_draws(seed: int, count: int) -> np.ndarray:
    out = np.zeros(count, dtype=np.uint32)
    oracle().orc_rand48_draws(seed, count, _ptr(out, _u32p))
    return out


def mt19937_words(seed: int, count: int) -> np.ndarray:
    out = np.zeros(count, dtype=np.uint32)
    oracle().orc_mt19937_words(seed, count, _ptr(out, _u32p))
    return out


def rng_below(seed: int, n: int, count: int) -> np.ndarray:
    out = np.zeros(count, dtype=np.int64)
    oracle().orc_rng_below(seed, n, count, _ptr(out, _i64p))
    return out


# ---- the reference tests' point sets (tests/pointsets.h) --------------------

def cylinder(n: int, q: int = 32, seed: int = 1) -> np.ndarray:
    out = np.zeros((n, 3))
    oracle().orc_pts_cylinder(n, q, seed, _ptr(out, _dp))
    return out


def sphere(n: int, q: int = 32, seed: int = 2) -> np.ndarray:
    out = np.zeros((n, 3))
    oracle().orc_pts_sphere(n, q, seed, _ptr(out, _dp))
    return out


def torus(n: int, q: int = 32, seed: int = 3) -> np.ndarray:
    out = np.zeros((n, 3))
    oracle().orc_pts_torus(n, q, seed, _ptr(out, _dp))
    return out


def box_shell(m: int) -> np.ndarray:
    k = oracle().orc_pts_box_shell(m, C.cast(None, _dp))
    out = np.zeros((k, 3))
    oracle().orc_pts_box_shell(m, _ptr(out, _dp))
    return out


def lattice(nx: int, ny: int, nz: int) -> np.ndarray:
    out = np.zeros((nx * ny * nz, 3))
    oracle().orc_pts_lattice(nx, ny, nz, _ptr(out, _dp))
    return out


def random_cube(n: int, seed: int = 4) -> np.ndarray:
    out = np.zeros((n, 3))
    oracle().orc_pts_random_cube(n, seed, _ptr(out, _dp))
    return out


def have_ref() -> bool:
    return REF_SO.exists()


def env_flag(name: str) -> bool:
    return os.environ.get(name, "") not in ("", "0")
