"""The driver bridge with its waves on the GPU: the reference's unchanged
Delaunay driver, its flood fill's conflict tests answered one mixed
predicate batch per breadth-first wave, must build the same cells in the same
order as the driver's own fill (oracle/bridge_delaunay.cpp)."""
import pytest

from bridge_util import check_cases, run
from gpu_util import require_gpu, say

pytestmark = pytest.mark.gpu


def test_bridge_reference_cases_gpu_waves():
    """The reference test's inputs, every wave on the GPU; the rest of the
    driver runs on the reference's own predicates and sort."""
    require_gpu()
    rows = run("ref_bridge_delaunay", "cases", 0)
    say([(r["case"], r["bridge"]["gpu_batches"]) for r in rows])
    check_cases(rows)
    assert sum(r["bridge"]["gpu_tests"] for r in rows) == sum(r["bridge"]["tests"] for r in rows) > 100000


def test_bridge_rand48_build_gpu_waves():
    """The 10^4-point rand48 build with all 53,517 waves on the GPU: the same
    66,624 cells in the same order, finite cells hashing as the reference's
    own build."""
    require_gpu()
    (row,) = run("ref_bridge_delaunay", "rand48", 10000, 0)
    say(row)
    assert row["same_as_driver"] is True and row["cells"] == 66624
    (ref,) = run("ref_delaunay_timing", 10000, 1)
    assert row["timing_hash"] == ref["hash"]
    assert row["bridge"]["gpu_tests"] == row["bridge"]["tests"]


def test_bridge_through_dropin_library():
    """The whole build on this library: the sort on the GPU, the driver's
    other predicate calls on the host path, the conflict waves of at least 64
    tests on the GPU and the smaller ones on the host."""
    require_gpu()
    (row,) = run("dropin_bridge_delaunay", "rand48", 10000, 64)
    say(row)
    assert row["same_as_driver"] is True and row["cells"] == 66624


def test_bridge_lockstep_builds_gpu_waves():
    """Sixteen builds advanced together, every step's waves of all pending
    regions in one GPU batch: each build equal to its own driver build."""
    require_gpu()
    (row,) = run("ref_bridge_delaunay", "lockstep", 16, 2000, 0)
    say(row)
    assert row["same_as_driver"] == 16
    assert row["bridge"]["gpu_batches"] * 40 < row["bridge"]["tests"]
