"""del3d_amd - del3d's exact predicates and spatial sort on MI355X (gfx950).

Python access to libdel3d_amd.so (include/del3d_amd.h). The names mirror the
reference's interface:

  orient3d(a, b, c, d)                 include/del3d/predicates.h:26
  in_sphere(a, b, c, d, e)             include/del3d/predicates.h:49-50
  in_sphere_unperturbed(a, b, c, d, e) include/del3d/predicates.h:54-55
  spatial_sort(points, order)          include/del3d/spatial_sort.h:37
  orient2d, coplanar_orientation, collinear, equal, collinear_position,
  coplanar_in_circle(_unperturbed)     include/del3d/predicates2.h:26-69
  circumcenter, circumcenter_lazy      include/del3d/delaunay.h:29-49

with the same sign convention (NEGATIVE=-1, ZERO=0, POSITIVE=1) and the same
results, plus batched forms over device buffers (``predicates``,
``spatial_sort_device``, ``spatial_sort_sets``). Every call runs on the GPU;
there is no CPU fallback, and importing the package fails if the library has
not been built.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import _hip
from ._hip import DeviceArray, Event, Stream, device_count, synchronize

NEGATIVE, ZERO, POSITIVE = -1, 0, 1
ORIENT3D, IN_SPHERE, IN_SPHERE_UNPERTURBED = 0, 1, 2
# coplanar batch op codes (DEL3D_OP2_*)
(COPLANAR_ORIENTATION, COPLANAR_ORIENTATION4, COLLINEAR, EQUAL, COLLINEAR_POSITION,
 IN_CIRCLE, IN_CIRCLE_UNPERTURBED, ORIENT2D_XY) = range(8)
CP_BEFORE, CP_SOURCE, CP_MIDDLE, CP_TARGET, CP_AFTER = range(5)

LIB_PATH = Path(__file__).resolve().parent / "libdel3d_amd.so"

_lib = None


def lib() -> C.CDLL:
    """The loaded C library. Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        _hip._lib()  # the HIP runtime the library links, loaded first and shared
        L = C.CDLL(str(LIB_PATH))
        vp, i64 = C.c_void_p, C.c_int64
        dp = C.POINTER(C.c_double)
        L.del3d_last_error.restype = C.c_char_p
        L.del3d_version.restype = C.c_char_p
        L.del3d_predicates.argtypes = [vp, i64, vp, vp, i64, vp, vp]
        L.del3d_predicates_stats.argtypes = [vp, i64, vp, vp, i64, vp, vp, C.POINTER(C.c_int64)]
        L.del3d_predicates_host.argtypes = [vp, i64, vp, vp, i64, vp]
        L.del3d_predicates_noalloc.argtypes = [vp, i64, vp, vp, i64, vp, vp]
        L.del3d_predicates_timed.argtypes = [vp, i64, vp, vp, i64, vp, vp, C.POINTER(C.c_float)]
        L.del3d_predicates_ev.argtypes = [vp, i64, vp, vp, i64, vp, vp, vp, vp]
        i64p = C.POINTER(C.c_int64)
        L.del3d_predicates_sets.argtypes = [vp, i64p, i64, vp, vp, i64p, vp, vp]
        L.del3d_pred_plan_create.argtypes = [i64p, i64p, i64, C.POINTER(C.c_void_p)]
        L.del3d_pred_plan_run.argtypes = [vp, vp, vp, vp, vp, vp]
        L.del3d_pred_plan_run_stats.argtypes = [vp, vp, vp, vp, vp, vp, i64p]
        L.del3d_pred_plan_run_graph.argtypes = [vp, vp, vp, vp, vp, vp]
        L.del3d_pred_plan_autotune.argtypes = [vp, vp, vp, vp, vp, vp, C.POINTER(C.c_float)]
        L.del3d_pred_plan_set_knobs.argtypes = [vp, C.POINTER(C.c_int32), C.c_int32]
        L.del3d_pred_plan_get_knobs.argtypes = [vp, C.POINTER(C.c_int32)]
        L.del3d_pred_plan_destroy.argtypes = [vp]
        L.del3d_server_create.argtypes = [i64, C.c_int32, i64, C.POINTER(C.c_void_p)]
        L.del3d_server_put_points.argtypes = [vp, vp, i64, i64p]
        L.del3d_server_drop_points.argtypes = [vp, i64]
        L.del3d_server_submit.argtypes = [vp, i64, vp, vp, i64, vp, i64p]
        L.del3d_server_wait.argtypes = [vp, i64, C.POINTER(C.c_double)]
        L.del3d_server_submit_sort.argtypes = [vp, i64, vp, i64p]
        L.del3d_release_cache.argtypes = []
        L.del3d_server_info.argtypes = [vp, i64p]
        L.del3d_server_destroy.argtypes = [vp]
        L.del3d_spatial_sort.argtypes = [vp, i64, vp, i64, vp]
        L.del3d_spatial_sort_sets.argtypes = [vp, C.POINTER(C.c_int64), i64, vp, vp]
        L.del3d_spatial_sort_host.argtypes = [vp, i64, vp, i64]
        L.del3d_orient3d.argtypes = [dp] * 4
        L.del3d_predicates2.argtypes = [vp, i64, vp, vp, i64, vp, vp]
        L.del3d_predicates2_host.argtypes = [vp, i64, vp, vp, i64, vp]
        L.del3d_predicates_mixed.argtypes = [vp, i64, vp, vp, i64, vp, vp]
        L.del3d_circumcenters.argtypes = [vp, i64, vp, i64, C.c_int, vp, vp]
        L.del3d_circumcenters_host.argtypes = [vp, i64, vp, i64, C.c_int, vp]
        L.del3d_circumcenter.argtypes = [dp, dp, dp, dp, C.c_int, dp]
        L.del3d_circumcenter.restype = None
        L.del3d_orient2d.argtypes = [C.c_double] * 6
        for name, n in (("del3d_coplanar_orientation", 3), ("del3d_coplanar_orientation4", 4),
                        ("del3d_collinear", 3), ("del3d_equal", 2), ("del3d_collinear_position", 3),
                        ("del3d_coplanar_in_circle", 4), ("del3d_coplanar_in_circle_unperturbed", 4)):
            getattr(L, name).argtypes = [dp] * n
        L.del3d_in_sphere.argtypes = [dp] * 5
        L.del3d_in_sphere_unperturbed.argtypes = [dp] * 5
        _lib = L
    return _lib


class Del3dError(RuntimeError):
    pass


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise Del3dError(f"{what}: {lib().del3d_last_error().decode()} (code {rc})")


def _stream_handle(stream):
    return stream.handle if stream is not None else None


# ---- batched, device-resident ------------------------------------------------

def predicates(xyz: DeviceArray, op: DeviceArray, idx: DeviceArray, out: DeviceArray,
               stream: Stream | None = None) -> None:
    """Signs of a mixed batch: op[q] in {ORIENT3D, IN_SPHERE, IN_SPHERE_UNPERTURBED},
    idx[q] = 5 point indices (orient3d ignores the fifth). Enqueued on `stream`."""
    n_q = int(op.shape[0])
    _check(lib().del3d_predicates(xyz.ptr, int(xyz.shape[0]), op.ptr, idx.ptr, n_q, out.ptr,
                                  _stream_handle(stream)), "del3d_predicates")


def predicates2(xyz: DeviceArray, op: DeviceArray, idx: DeviceArray, out: DeviceArray,
                stream: Stream | None = None) -> None:
    """Results of a coplanar batch: op[q] one of COPLANAR_ORIENTATION ..
    ORIENT2D_XY, idx[q] = the op's point indices in argument order (5 int32,
    the unused ones ignored). Enqueued on `stream`."""
    n_q = int(op.shape[0])
    _check(lib().del3d_predicates2(xyz.ptr, int(xyz.shape[0]), op.ptr, idx.ptr, n_q, out.ptr,
                                   _stream_handle(stream)), "del3d_predicates2")


MIXED_OP2_BASE = 16
INEXACT_ORIENTATION = 8  # DEL3D_OP2_INEXACT_ORIENTATION


def predicates_mixed(xyz: DeviceArray, op: DeviceArray, idx: DeviceArray, out: DeviceArray,
                     stream: Stream | None = None) -> None:
    """A mixed batch: op bytes ORIENT3D / IN_SPHERE / IN_SPHERE_UNPERTURBED or
    MIXED_OP2_BASE + a coplanar op code (the predicate trace format's codes)."""
    n_q = int(op.shape[0])
    _check(lib().del3d_predicates_mixed(xyz.ptr, int(xyz.shape[0]), op.ptr, idx.ptr, n_q, out.ptr,
                                        _stream_handle(stream)), "del3d_predicates_mixed")


def circumcenters(xyz: DeviceArray, cells: DeviceArray, out: DeviceArray, lazy: bool = False,
                  stream: Stream | None = None) -> None:
    """Circumcentres of cells (n, 4) int32 into out (n, 3) f64: the exact
    circumcenter, or circumcenter_lazy with lazy=True. Enqueued on `stream`."""
    _check(lib().del3d_circumcenters(xyz.ptr, int(xyz.shape[0]), cells.ptr, int(cells.shape[0]),
                                     1 if lazy else 0, out.ptr, _stream_handle(stream)),
           "del3d_circumcenters")


STAT_NAMES = ("filter", "certified", "exact_regs", "exact_wide", "perturbed", "bad_index")


def predicates_stats(xyz: DeviceArray, op: DeviceArray, idx: DeviceArray, out: DeviceArray,
                     stream: Stream | None = None) -> dict:
    st = (C.c_int64 * 6)()
    _check(lib().del3d_predicates_stats(xyz.ptr, int(xyz.shape[0]), op.ptr, idx.ptr,
                                        int(op.shape[0]), out.ptr, _stream_handle(stream), st),
           "del3d_predicates_stats")
    return dict(zip(STAT_NAMES, st))


def _i64p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def predicates_sets(xyz: DeviceArray, pt_off, op: DeviceArray, idx: DeviceArray, q_off,
                    out: DeviceArray, stream: Stream | None = None) -> None:
    """Many point sets in one call: set s owns points pt_off[s]:pt_off[s+1] and
    queries q_off[s]:q_off[s+1], with set-local indices. Synchronizes."""
    po = np.ascontiguousarray(pt_off, dtype=np.int64)
    qo = np.ascontiguousarray(q_off, dtype=np.int64)
    _check(lib().del3d_predicates_sets(xyz.ptr, _i64p(po), len(po) - 1, op.ptr, idx.ptr, _i64p(qo),
                                       out.ptr, _stream_handle(stream)), "del3d_predicates_sets")


class PredPlan:
    """A plan over a fixed set layout (include/del3d_amd.h, del3d_pred_plan_*)."""

    KNOBS = 5

    def __init__(self, pt_off, q_off):
        self._po = np.ascontiguousarray(pt_off, dtype=np.int64)
        self._qo = np.ascontiguousarray(q_off, dtype=np.int64)
        h = C.c_void_p()
        _check(lib().del3d_pred_plan_create(_i64p(self._po), _i64p(self._qo), len(self._po) - 1,
                                            C.byref(h)), "del3d_pred_plan_create")
        self.handle = h

    def run(self, xyz, op, idx, out, stream=None, graph=False):
        f = lib().del3d_pred_plan_run_graph if graph else lib().del3d_pred_plan_run
        _check(f(self.handle, xyz.ptr, op.ptr, idx.ptr, out.ptr, _stream_handle(stream)),
               "del3d_pred_plan_run")

    def run_stats(self, xyz, op, idx, out, stream=None) -> dict:
        st = (C.c_int64 * 6)()
        _check(lib().del3d_pred_plan_run_stats(self.handle, xyz.ptr, op.ptr, idx.ptr, out.ptr,
                                               _stream_handle(stream), st), "del3d_pred_plan_run_stats")
        return dict(zip(STAT_NAMES, st))

    def autotune(self, xyz, op, idx, out, stream=None) -> float:
        ms = C.c_float()
        _check(lib().del3d_pred_plan_autotune(self.handle, xyz.ptr, op.ptr, idx.ptr, out.ptr,
                                              _stream_handle(stream), C.byref(ms)),
               "del3d_pred_plan_autotune")
        return float(ms.value)

    @property
    def knobs(self) -> list:
        k = (C.c_int32 * self.KNOBS)()
        _check(lib().del3d_pred_plan_get_knobs(self.handle, k), "del3d_pred_plan_get_knobs")
        return list(k)

    @knobs.setter
    def knobs(self, values) -> None:
        k = (C.c_int32 * len(values))(*values)
        _check(lib().del3d_pred_plan_set_knobs(self.handle, k, len(values)), "del3d_pred_plan_set_knobs")

    def close(self) -> None:
        if self.handle:
            lib().del3d_pred_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def predicates_timed(xyz: DeviceArray, op: DeviceArray, idx: DeviceArray, out: DeviceArray,
                     stream: Stream | None = None) -> tuple[float, float, float]:
    """del3d_predicates plus HIP-event durations (ms) of its three kernels."""
    ms = (C.c_float * 3)()
    _check(lib().del3d_predicates_timed(xyz.ptr, int(xyz.shape[0]), op.ptr, idx.ptr,
                                        int(op.shape[0]), out.ptr, _stream_handle(stream), ms),
           "del3d_predicates_timed")
    return float(ms[0]), float(ms[1]), float(ms[2])


def predicates_ev(xyz: DeviceArray, op: DeviceArray, idx: DeviceArray, out: DeviceArray,
                  stream: Stream, ev_begin: Event, ev_end: Event) -> None:
    """del3d_predicates, with the two events recorded around the fast kernel
    (no synchronization)."""
    _check(lib().del3d_predicates_ev(xyz.ptr, int(xyz.shape[0]), op.ptr, idx.ptr, int(op.shape[0]),
                                     out.ptr, _stream_handle(stream), ev_begin.handle, ev_end.handle),
           "del3d_predicates_ev")


def spatial_sort_device(xyz: DeviceArray, order: DeviceArray, stream: Stream | None = None) -> None:
    _check(lib().del3d_spatial_sort(xyz.ptr, int(xyz.shape[0]), order.ptr, int(order.shape[0]),
                                    _stream_handle(stream)), "del3d_spatial_sort")


def spatial_sort_sets(xyz: DeviceArray, offsets: np.ndarray, order: DeviceArray,
                      stream: Stream | None = None) -> None:
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    _check(lib().del3d_spatial_sort_sets(xyz.ptr, off.ctypes.data_as(C.POINTER(C.c_int64)),
                                         len(off) - 1, order.ptr, _stream_handle(stream)),
           "del3d_spatial_sort_sets")


class Server:
    """The serving loop (include/del3d_amd.h, del3d_server_*): resident point
    sets and continuously batched predicate jobs. Host arrays handed to it
    are kept referenced until they may be released."""

    INFO = ("used_bytes", "resident_bytes", "resident_sets", "evictions", "uploads", "batches",
            "jobs_done", "budget")

    def __init__(self, budget_bytes: int, slots: int = 3, max_batch_queries: int = 1 << 24):
        h = C.c_void_p()
        _check(lib().del3d_server_create(int(budget_bytes), int(slots), int(max_batch_queries),
                                         C.byref(h)), "del3d_server_create")
        self.handle = h
        self._sets = {}
        self._jobs = {}

    def put_points(self, points) -> int:
        xyz = points if (isinstance(points, np.ndarray) and points.dtype == np.float64 and points.ndim == 2
                         and points.shape[1] == 3 and points.flags.c_contiguous) \
            else np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
        sid = C.c_int64()
        _check(lib().del3d_server_put_points(self.handle, xyz.ctypes.data, len(xyz), C.byref(sid)),
               "del3d_server_put_points")
        self._sets[sid.value] = xyz
        return sid.value

    def drop_points(self, set_id: int) -> None:
        _check(lib().del3d_server_drop_points(self.handle, int(set_id)), "del3d_server_drop_points")
        self._sets.pop(set_id, None)

    def submit(self, set_id: int, op, idx, out=None) -> int:
        """A predicate job; `out` (int8, e.g. pinned) receives the signs."""
        op = np.ascontiguousarray(op, dtype=np.uint8)
        idx = np.ascontiguousarray(np.asarray(idx, dtype=np.int32).reshape(-1, 5))
        if out is None:
            out = np.empty(len(op), dtype=np.int8)
        assert out.dtype == np.int8 and len(out) == len(op) and out.flags.c_contiguous
        jid = C.c_int64()
        _check(lib().del3d_server_submit(self.handle, int(set_id), op.ctypes.data, idx.ctypes.data,
                                         len(op), out.ctypes.data, C.byref(jid)), "del3d_server_submit")
        self._jobs[jid.value] = (op, idx, out)
        return jid.value

    def submit_sort(self, set_id: int, out=None) -> int:
        """A sort job: the spatial sort of the whole resident set; `out` (an
        int32 array of the set's size, e.g. pinned) receives the order."""
        n = len(self._sets[set_id])
        if out is None:
            out = np.empty(n, dtype=np.int32)
        assert out.dtype == np.int32 and len(out) == n and out.flags.c_contiguous
        jid = C.c_int64()
        _check(lib().del3d_server_submit_sort(self.handle, int(set_id), out.ctypes.data, C.byref(jid)),
               "del3d_server_submit_sort")
        self._jobs[jid.value] = (out,)
        return jid.value

    def wait(self, job_id: int):
        """(signs or sort order, latency in ms) of a finished job."""
        ms = C.c_double()
        rc = lib().del3d_server_wait(self.handle, int(job_id), C.byref(ms))
        bufs = self._jobs.pop(job_id)
        _check(rc, "del3d_server_wait")
        return bufs[-1], ms.value

    def info(self) -> dict:
        v = (C.c_int64 * 8)()
        _check(lib().del3d_server_info(self.handle, v), "del3d_server_info")
        return dict(zip(self.INFO, v))

    def close(self) -> None:
        if self.handle:
            lib().del3d_server_destroy(self.handle)
            self.handle = None
            self._jobs.clear()
            self._sets.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- host-memory conveniences ------------------------------------------------

def predicates_np(points, op, idx) -> np.ndarray:
    xyz = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
    op = np.ascontiguousarray(op, dtype=np.uint8)
    idx = np.ascontiguousarray(np.asarray(idx, dtype=np.int32).reshape(-1, 5))
    out = np.zeros(len(op), dtype=np.int8)
    _check(lib().del3d_predicates_host(xyz.ctypes.data, len(xyz), op.ctypes.data, idx.ctypes.data,
                                       len(op), out.ctypes.data), "del3d_predicates_host")
    return out


def predicates2_np(points, op, idx) -> np.ndarray:
    """The coplanar batch on host arrays (copied in and out)."""
    xyz = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
    op = np.ascontiguousarray(op, dtype=np.uint8)
    idx = np.ascontiguousarray(np.asarray(idx, dtype=np.int32).reshape(-1, 5))
    out = np.zeros(len(op), dtype=np.int8)
    _check(lib().del3d_predicates2_host(xyz.ctypes.data, len(xyz), op.ctypes.data,
                                        idx.ctypes.data, len(op), out.ctypes.data),
           "del3d_predicates2_host")
    return out


def circumcenters_np(points, cells, lazy: bool = False) -> np.ndarray:
    """Circumcentres of cells on host arrays; returns (n, 3)."""
    xyz = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
    cells = np.ascontiguousarray(np.asarray(cells, dtype=np.int32).reshape(-1, 4))
    out = np.zeros((len(cells), 3))
    _check(lib().del3d_circumcenters_host(xyz.ctypes.data, len(xyz), cells.ctypes.data, len(cells),
                                          1 if lazy else 0, out.ctypes.data),
           "del3d_circumcenters_host")
    return out


def spatial_sort(points, order=None) -> np.ndarray:
    """The reference's spatial_sort(points, order) on host arrays; returns the
    permuted order (a new array)."""
    xyz = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
    o = (np.arange(len(xyz), dtype=np.int32) if order is None
         else np.array(order, dtype=np.int32, copy=True))
    _check(lib().del3d_spatial_sort_host(xyz.ctypes.data, len(xyz), o.ctypes.data, len(o)),
           "del3d_spatial_sort_host")
    return o


def _pts(*args):
    """One double[3] per argument; the same Python object passed twice becomes
    the same pointer, which is what the reference's perturbation compares."""
    seen = {}
    keep, ptrs = [], []
    for p in args:
        if id(p) not in seen:
            a = np.ascontiguousarray(p, dtype=np.float64).reshape(3)
            keep.append(a)
            seen[id(p)] = a.ctypes.data_as(C.POINTER(C.c_double))
        ptrs.append(seen[id(p)])
    return keep, ptrs


def orient3d(a, b, c, d) -> int:
    """Sign of |b-a, c-a, d-a| (one GPU query)."""
    keep, p = _pts(a, b, c, d)
    return int(lib().del3d_orient3d(*p))


def in_sphere(a, b, c, d, e) -> int:
    """POSITIVE iff e is inside the sphere through a, b, c, d (positively
    oriented); never ZERO (symbolic perturbation). One GPU query."""
    keep, p = _pts(a, b, c, d, e)
    return int(lib().del3d_in_sphere(*p))


def in_sphere_unperturbed(a, b, c, d, e) -> int:
    keep, p = _pts(a, b, c, d, e)
    return int(lib().del3d_in_sphere_unperturbed(*p))


def orient2d(px, py, qx, qy, rx, ry) -> int:
    """Sign of |q-p, r-p| in the plane (one GPU query)."""
    return int(lib().del3d_orient2d(px, py, qx, qy, rx, ry))


def coplanar_orientation(p, q, r, s=None) -> int:
    """Orientation of (p,q,r) in the first non-degenerate of the xy, yz, xz
    projections; with s, the side of s against line (p,q), signed by (p,q,r)."""
    if s is None:
        keep, ptr = _pts(p, q, r)
        return int(lib().del3d_coplanar_orientation(*ptr))
    keep, ptr = _pts(p, q, r, s)
    return int(lib().del3d_coplanar_orientation4(*ptr))


def collinear(p, q, r) -> bool:
    keep, ptr = _pts(p, q, r)
    return bool(lib().del3d_collinear(*ptr))


def equal(p, q) -> bool:
    keep, ptr = _pts(p, q)
    return bool(lib().del3d_equal(*ptr))


def collinear_position(p, q, r) -> int:
    """CP_BEFORE, CP_SOURCE, CP_MIDDLE, CP_TARGET or CP_AFTER."""
    keep, ptr = _pts(p, q, r)
    return int(lib().del3d_collinear_position(*ptr))


def coplanar_in_circle(p0, p1, p2, p) -> int:
    """POSITIVE iff p is inside the circle through the coplanar p0, p1, p2;
    never ZERO (symbolic perturbation)."""
    keep, ptr = _pts(p0, p1, p2, p)
    return int(lib().del3d_coplanar_in_circle(*ptr))


def coplanar_in_circle_unperturbed(p0, p1, p2, p) -> int:
    keep, ptr = _pts(p0, p1, p2, p)
    return int(lib().del3d_coplanar_in_circle_unperturbed(*ptr))


def circumcenter(p0, p1, p2, p3) -> tuple:
    """The exact circumcentre (one GPU evaluation); p0 if coplanar."""
    keep, p = _pts(p0, p1, p2, p3)
    out = (C.c_double * 3)()
    lib().del3d_circumcenter(*p, 0, out)
    return tuple(out)


def circumcenter_lazy(p, q, r, s) -> tuple:
    """The lazy circumcentre, bit-compatible with the reference's."""
    keep, ptr = _pts(p, q, r, s)
    out = (C.c_double * 3)()
    lib().del3d_circumcenter(*ptr, 1, out)
    return tuple(out)


__all__ = [
    "NEGATIVE", "ZERO", "POSITIVE", "ORIENT3D", "IN_SPHERE", "IN_SPHERE_UNPERTURBED",
    "DeviceArray", "Stream", "Event", "device_count", "synchronize", "lib", "Del3dError",
    "predicates", "predicates_stats", "predicates_timed", "predicates_sets", "PredPlan", "Server", "spatial_sort_device", "spatial_sort_sets",
    "predicates_np", "spatial_sort", "orient3d", "in_sphere", "in_sphere_unperturbed",
    "predicates2", "predicates2_np", "orient2d", "coplanar_orientation", "collinear", "equal",
    "collinear_position", "coplanar_in_circle", "coplanar_in_circle_unperturbed",
    "COPLANAR_ORIENTATION", "COPLANAR_ORIENTATION4", "COLLINEAR", "EQUAL", "COLLINEAR_POSITION",
    "IN_CIRCLE", "IN_CIRCLE_UNPERTURBED", "ORIENT2D_XY", "INEXACT_ORIENTATION", "MIXED_OP2_BASE",
    "predicates_mixed", "circumcenters", "circumcenters_np", "circumcenter", "circumcenter_lazy",
]
