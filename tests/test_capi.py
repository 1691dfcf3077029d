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
