"""CPU checks of the C ABI: the library loads, exports every declared entry
point (C and the C++ drop-in), and rejects bad arguments before touching a
device."""
import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "del3d_amd" / "libdel3d_amd.so"

DROPIN_SYMBOLS = [
    "_ZN5del3d8orient3dEPKdS1_S1_S1_",
    "_ZN5del3d9in_sphereEPKdS1_S1_S1_S1_",
    "_ZN5del3d21in_sphere_unperturbedEPKdS1_S1_S1_S1_",
    "_ZN5del3d12spatial_sortERKSt6vectorINS_6Point3ESaIS1_EERS0_IiSaIiEE",
    # include/del3d/predicates2.h:26-69
    "_ZN5del3d8orient2dEdddddd",
    "_ZN5del3d20coplanar_orientationEPKdS1_S1_",
    "_ZN5del3d20coplanar_orientationEPKdS1_S1_S1_",
    "_ZN5del3d9collinearEPKdS1_S1_",
    "_ZN5del3d5equalEPKdS1_",
    "_ZN5del3d18collinear_positionEPKdS1_S1_",
    "_ZN5del3d18coplanar_in_circleEPKdS1_S1_S1_",
    "_ZN5del3d30coplanar_in_circle_unperturbedEPKdS1_S1_S1_",
    # include/del3d/delaunay.h:29-49
    "_ZN5del3d12circumcenterERKNS_6Point3ES2_S2_S2_",
    "_ZN5del3d17circumcenter_lazyERKNS_6Point3ES2_S2_S2_",
]


def declared_functions():
    text = (ROOT / "include" / "del3d_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(del3d_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not LIB.exists():
        pytest.fail(f"{LIB} not built")
    return ctypes.CDLL(str(LIB))


def test_exports_every_declared_function(lib):
    names = declared_functions()
    assert "del3d_predicates" in names and "del3d_spatial_sort" in names
    for n in names:
        assert hasattr(lib, n), n


def test_exports_reference_cpp_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True,
                         text=True, check=True).stdout
    for sym in DROPIN_SYMBOLS:
        assert re.search(rf"\bT {re.escape(sym)}$", out, flags=re.M), sym


def test_argument_errors_without_device(lib):
    lib.del3d_last_error.restype = ctypes.c_char_p
    vp = ctypes.c_void_p
    lib.del3d_predicates.argtypes = [vp, ctypes.c_int64, vp, vp, ctypes.c_int64, vp, vp]
    assert lib.del3d_predicates(None, 0, None, None, 0, None, None) == 0   # nothing to do
    assert lib.del3d_predicates(None, 0, None, None, -1, None, None) == 1
    assert b"negative" in lib.del3d_last_error()
    assert lib.del3d_predicates(None, 0, None, None, 5, None, None) == 1
    lib.del3d_spatial_sort.argtypes = [vp, ctypes.c_int64, vp, ctypes.c_int64, vp]
    assert lib.del3d_spatial_sort(None, 0, None, 1, None) == 0           # n < 2: no-op
    assert lib.del3d_spatial_sort(None, 0, None, 5, None) == 1
    off = (ctypes.c_int64 * 3)(0, 5, 3)
    lib.del3d_spatial_sort_sets.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ctypes.c_int64, vp, vp]
    assert lib.del3d_spatial_sort_sets(None, off, 2, None, None) == 1
    assert b"decrease" in lib.del3d_last_error()
    lib.del3d_predicates2.argtypes = [vp, ctypes.c_int64, vp, vp, ctypes.c_int64, vp, vp]
    assert lib.del3d_predicates2(None, 0, None, None, 0, None, None) == 0
    assert lib.del3d_predicates2(None, 0, None, None, -3, None, None) == 1
    assert lib.del3d_predicates2(None, 0, None, None, 4, None, None) == 1
    assert b"null" in lib.del3d_last_error()


def test_noalloc_refuses_sets_above_its_limit(lib):
    """The allocation-free form checks the point count against its fixed
    scratch before touching a device."""
    lib.del3d_last_error.restype = ctypes.c_char_p
    vp = ctypes.c_void_p
    lib.del3d_predicates_noalloc.argtypes = [vp, ctypes.c_int64, vp, vp, ctypes.c_int64, vp, vp]
    dummy = ctypes.create_string_buffer(64)
    p = ctypes.cast(dummy, vp)
    assert lib.del3d_predicates_noalloc(p, (1 << 20) + 1, p, p, 1, p, None) == 1
    assert b"at most 1048576 points" in lib.del3d_last_error()
    assert lib.del3d_predicates_noalloc(None, 0, None, None, 0, None, None) == 0  # nothing to do
    assert lib.del3d_predicates_noalloc(None, 0, None, None, -1, None, None) == 1


def test_version(lib):
    lib.del3d_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.del3d_version()


class TdsView(ctypes.Structure):
    _fields_ = [("cell_v", ctypes.c_void_p), ("cell_n", ctypes.c_void_p), ("cell_stride", ctypes.c_int64),
                ("n_cells", ctypes.c_int64), ("vertex_point", ctypes.c_void_p),
                ("vertex_stride", ctypes.c_int64), ("n_vertices", ctypes.c_int64),
                ("point_base", ctypes.c_int64), ("infinite", ctypes.c_int32), ("dimension", ctypes.c_int32)]


def test_bridge_argument_errors_without_device(lib):
    """The bridge's host-only form needs no device; bad views and short
    output buffers are errors, not reads outside the caller's tables."""
    lib.del3d_last_error.restype = ctypes.c_char_p
    vp, i64 = ctypes.c_void_p, ctypes.c_int64
    lib.del3d_bridge_create.argtypes = [vp, i64, i64, ctypes.POINTER(vp)]
    lib.del3d_bridge_conflicts.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(TdsView), vp, vp, vp, i64, vp, vp,
                                           i64, vp]
    lib.del3d_bridge_info.argtypes = [vp, vp]
    lib.del3d_bridge_destroy.argtypes = [vp]
    br = vp()
    assert lib.del3d_bridge_create(None, 5, 0, ctypes.byref(br)) == 1
    pts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0.3, 0.3, 0.3], [0, 0, -5]], np.float64)
    assert lib.del3d_bridge_create(pts.ctypes.data, 6, -1, ctypes.byref(br)) == 1
    assert lib.del3d_bridge_create(pts.ctypes.data, 6, 2**63 - 1, ctypes.byref(br)) == 0
    # two cells, each the other's neighbour across all four facets: cell 0
    # over points 0..3 (the start), cell 1 over 0, 1, 2, 5, positively
    # oriented, whose circumsphere (centre (0.5, 0.5, -2.5)) excludes point
    # 4. The region of point 4 is cell 0 alone, with four facets, all
    # towards cell 1.
    lib.del3d_orient3d.argtypes = [vp] * 4
    c1 = [0, 1, 2, 5]
    if lib.del3d_orient3d(*[pts[k].ctypes.data for k in c1]) < 0:
        c1 = [1, 0, 2, 5]
    cv = np.array([0, 1, 2, 3] + c1, np.int32)
    cn = np.array([1, 1, 1, 1, 0, 0, 0, 0], np.int32)
    vpt = np.arange(6, dtype=np.int32)
    view = TdsView(cv.ctypes.data, cn.ctypes.data, 4, 2, vpt.ctypes.data, 1, 6, 0, 99, 3)
    starts = np.array([0], np.int32)
    point = np.array([4], np.int32)
    cells = np.zeros(8, np.int32)
    facets = np.zeros(16, np.int32)
    coff = np.zeros(2, np.int64)
    foff = np.zeros(2, np.int64)
    args = lambda v, ccap=8, fcap=8: (br, 1, ctypes.byref(v), starts.ctypes.data, point.ctypes.data,
                                      cells.ctypes.data, ccap, coff.ctypes.data, facets.ctypes.data, fcap,
                                      foff.ctypes.data)
    assert lib.del3d_bridge_conflicts(*args(view)) == 0, lib.del3d_last_error()
    assert list(coff) == [0, 1] and list(foff) == [0, 4] and cells[0] == 0
    assert facets[:8].tolist() == [0, 0, 0, 1, 0, 2, 0, 3]
    assert lib.del3d_bridge_conflicts(*args(view, fcap=2)) == 1      # short facet buffer
    assert b"capacity" in lib.del3d_last_error() and foff[1] == 4
    cn[0] = 7                                                          # neighbour outside the view
    assert lib.del3d_bridge_conflicts(*args(view)) == 1
    assert b"bad view" in lib.del3d_last_error()
    cn[0] = 1
    vpt[5] = 6                                                         # point outside the bridge's
    assert lib.del3d_bridge_conflicts(*args(view)) == 1
    assert b"bad view" in lib.del3d_last_error()
    vpt[5] = 5
    bad_dim = TdsView(cv.ctypes.data, cn.ctypes.data, 4, 2, vpt.ctypes.data, 1, 6, 0, 99, 4)
    assert lib.del3d_bridge_conflicts(*args(bad_dim)) == 1
    info = np.zeros(6, np.int64)
    assert lib.del3d_bridge_info(br, info.ctypes.data) == 0
    assert info[0] >= 1 and info[4] == info[2] and info[3] == 0       # all on the host
    assert lib.del3d_bridge_destroy(br) == 0
