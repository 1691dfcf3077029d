"""Circumcentres (src/triangulation.cpp:805-930) on the CPU.

The oracle's restatement (expansions; intervals under FE_UPWARD) against the
compiled reference, bit for bit, on random cells, slivers (where the interval
pass refuses and the exact form decides), degenerate lattice cells and the
cells of the reference's own Delaunay triangulation; against exact rational
arithmetic (Python fractions) within a few ulps for the exact form and within
the 1e-5 relative acceptance for the lazy form. Then the GPU's per-cell core
(del3d_amd/csrc/circ_core.h: upward rounding emulated from round-to-nearest,
expansions on a strided workspace), host build, against the oracle bit for
bit.
"""
import ctypes as C
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
from del3d_amd import workloads as W

HOST = Path(__file__).resolve().parents[1] / "del3d_amd" / "libpredcore_host.so"


def host_core(pts, cells, lazy):
    lib = C.CDLL(str(HOST))
    pts = np.ascontiguousarray(pts, np.float64)
    cells = np.ascontiguousarray(cells, np.int32)
    out = np.zeros((len(cells), 3))
    lib.circcore_host_batch(pts.ctypes.data_as(C.c_void_p), cells.ctypes.data_as(C.c_void_p),
                            C.c_int64(len(cells)), C.c_int(lazy), out.ctypes.data_as(C.c_void_p))
    return out


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def cell_sets():
    rng = np.random.default_rng(17)
    sets = {}
    pts = rng.random((400, 3))
    sets["random"] = (pts, rng.integers(0, 400, size=(3000, 4)).astype(np.int32))
    # slivers: four points near a common plane (offsets of a few ulps to 1e-9)
    base = rng.random((3000, 3, 2))
    sl = []
    for i in range(3000):
        uv = np.vstack([base[i], rng.random((1, 2))])
        z = 0.2 * uv[:, 0] + 0.3 * uv[:, 1] + 0.1
        z[3] += rng.choice([1e-9, 1e-12, 1e-15, 4e-16, 0.0]) * rng.choice([-1, 1])
        sl.append(np.column_stack([uv, z]))
    sets["slivers"] = (np.vstack(sl), np.arange(12000, dtype=np.int32).reshape(3000, 4))
    lat = W.lattice(4, 4, 4)
    sets["lattice"] = (lat, rng.integers(0, 64, size=(3000, 4)).astype(np.int32))
    wide = rng.normal(size=(300, 3)) * np.exp2(rng.integers(-60, 60, size=(300, 1)))
    sets["scales"] = (wide, rng.integers(0, 300, size=(3000, 4)).astype(np.int32))
    return sets


@pytest.mark.parametrize("lazy", [0, 1])
@pytest.mark.parametrize("name", ["random", "slivers", "lattice", "scales"])
def test_oracle_and_gpu_core_match_reference(name, lazy):
    pts, cells = cell_sets()[name]
    want = O.circumcenters(pts, cells, lazy, "oracle")
    if O.ref() is not None:
        assert np.array_equal(bits(O.circumcenters(pts, cells, lazy, "ref")), bits(want))
    if not HOST.exists():
        pytest.fail(f"{HOST} not built")
    lib = C.CDLL(str(HOST))
    lib.circcore_overflows.restype = C.c_int64
    before = lib.circcore_overflows()
    got = host_core(pts, cells, lazy)
    bad = np.nonzero(np.any(bits(got) != bits(want), axis=1))[0]
    assert len(bad) == 0, (name, cells[bad[:3]], got[bad[:3]], want[bad[:3]])
    if name == "scales" and not lazy:
        # expansions that outgrow the small workspace: the worst-case pass decides
        assert lib.circcore_overflows() - before > 300


def test_reference_delaunay_cells():
    """The cells of the reference's triangulation of 3,000 rand48 points."""
    if O.ref() is None:
        pytest.skip("reference library not built")
    pts = W.c1_points(3000, 1)
    cells = O.delaunay_cells(pts)
    assert len(cells) > 15000
    for lazy in (0, 1):
        want = O.circumcenters(pts, cells, lazy, "ref")
        assert np.array_equal(bits(O.circumcenters(pts, cells, lazy, "oracle")), bits(want))
        if HOST.exists():
            assert np.array_equal(bits(host_core(pts, cells, lazy)), bits(want))


def _exact_center(p):
    """The circumcentre in rational arithmetic (None if coplanar)."""
    P = [[Fraction(float(v)) for v in row] for row in p]
    a, b, c = ([P[i][k] - P[0][k] for k in range(3)] for i in (1, 2, 3))

    def cross(u, v):
        return [u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]]

    def dot(u, v):
        return sum(x * y for x, y in zip(u, v))
    den = 2 * dot(a, cross(b, c))
    if den == 0:
        return None
    bc, ca, ab = cross(b, c), cross(c, a), cross(a, b)
    return [P[0][k] + (dot(a, a) * bc[k] + dot(b, b) * ca[k] + dot(c, c) * ab[k]) / den
            for k in range(3)]


@pytest.mark.parametrize("name", ["random", "slivers"])
def test_against_rational_arithmetic(name):
    pts, cells = cell_sets()[name]
    cells = cells[:400]
    ex = O.circumcenters(pts, cells, 0, "oracle")
    lz = O.circumcenters(pts, cells, 1, "oracle")
    checked = 0
    for i, c in enumerate(cells):
        e = _exact_center(pts[c])
        if e is None:
            continue
        checked += 1
        for k in range(3):
            t = float(e[k])
            # exact form: two roundings (quotient, then the sum with p0)
            tol = 4 * np.spacing(abs(t)) + 4 * np.spacing(abs(float(pts[c[0], k])))
            assert abs(ex[i, k] - t) <= tol, (i, k, ex[i, k], t)
            # lazy form: a midpoint accepted at 1e-5 relative radius, or exact
            assert abs(lz[i, k] - t) <= 1.01e-5 * abs(t) + tol, (i, k, lz[i, k], t)
    assert checked > 300
