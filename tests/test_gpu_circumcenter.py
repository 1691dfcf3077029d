"""GPU parity of the circumcentres (del3d_circumcenters, the scalar entry
point and the C++ drop-in symbols) with the oracle's restatement of
src/triangulation.cpp:805-930, bit for bit, in both forms: random cells,
slivers (the lazy pass refuses, the exact kernel decides), degenerate lattice
cells, mixed scales, the cells of the reference's own triangulation, and a
bad index (NaN)."""
import numpy as np
import pytest

import del3d_amd as D
import oracle_lib as O
import test_circumcenter as TC
from del3d_amd import workloads as W
from gpu_util import require_gpu, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


def run(pts, cells, lazy):
    xyz = D.DeviceArray.from_host(np.ascontiguousarray(pts, np.float64).reshape(-1, 3))
    dc = D.DeviceArray.from_host(np.ascontiguousarray(cells, np.int32).reshape(-1, 4))
    out = D.DeviceArray((len(cells), 3), np.float64)
    D.circumcenters(xyz, dc, out, lazy=bool(lazy))
    D.synchronize()
    return out.to_host()


@pytest.mark.parametrize("lazy", [0, 1])
@pytest.mark.parametrize("name", ["random", "slivers", "lattice", "scales"])
def test_cells_match_oracle(name, lazy):
    pts, cells = TC.cell_sets()[name]
    say("circumcenters", name, "lazy" if lazy else "exact", len(cells))
    got = run(pts, cells, lazy)
    want = O.circumcenters(pts, cells, lazy, "oracle")
    bad = np.nonzero(np.any(TC.bits(got) != TC.bits(want), axis=1))[0]
    assert len(bad) == 0, (name, cells[bad[:3]], got[bad[:3]], want[bad[:3]])


def test_reference_triangulation_cells():
    if O.ref() is None:
        pytest.skip("reference library not built")
    pts = W.c1_points(10000, 1)
    cells = O.delaunay_cells(pts)
    say("circumcenters of", len(cells), "Delaunay cells")
    for lazy in (0, 1):
        got = run(pts, cells, lazy)
        assert np.array_equal(TC.bits(got), TC.bits(O.circumcenters(pts, cells, lazy, "ref")))


def test_large_batch_and_host_form():
    rng = np.random.default_rng(3)
    pts = rng.random((100000, 3))
    cells = rng.integers(0, len(pts), size=(1 << 20, 4)).astype(np.int32)
    say("large circumcenter batch", len(cells))
    for lazy in (1, 0):
        got = run(pts, cells, lazy)
        s = rng.choice(len(cells), 20000, replace=False)
        assert np.array_equal(TC.bits(got[s]), TC.bits(O.circumcenters(pts, cells[s], lazy, "oracle")))
    small = cells[:3000]
    assert np.array_equal(TC.bits(D.circumcenters_np(pts, small, lazy=True)),
                          TC.bits(O.circumcenters(pts, small, 1, "oracle")))


def test_scalar_and_bad_index():
    a, b, c, d = [0.0, 0, 0], [1.0, 0, 0], [0.0, 1, 0], [0.0, 0, 1]
    assert D.circumcenter(a, b, c, d) == (0.5, 0.5, 0.5)
    assert D.circumcenter_lazy(a, b, c, d) == (0.5, 0.5, 0.5)
    assert D.circumcenter(a, b, c, [1.0, 1, 0]) == (0.0, 0.0, 0.0)   # coplanar: p0
    pts = np.array([a, b, c, d])
    got = run(pts, np.array([[0, 1, 2, 3], [0, 1, 2, 4], [-1, 1, 2, 3]], np.int32), 1)
    assert got[0].tolist() == [0.5, 0.5, 0.5]
    assert np.isnan(got[1:]).all()
