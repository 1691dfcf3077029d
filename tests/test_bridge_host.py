"""The driver bridge (del3d_bridge_conflicts) on the CPU: the reference's
unchanged Delaunay driver, with its flood fill's conflict tests answered in
breadth-first waves by the bridge (every wave on the host here), must build
the same cells in the same order as the driver's own fill. The programs link
the reference's own predicates and sort, so only the bridge is this
library's (oracle/bridge_delaunay.cpp)."""
import pytest

from bridge_util import HOST_ONLY, check_cases, run


def test_bridge_reference_cases_host():
    check_cases(run("ref_bridge_delaunay", "cases", HOST_ONLY))


def test_bridge_rand48_build_host():
    """The 10^4-point rand48 build: every one of its 66,624 cells in the
    driver's order, and the finite cells' hash equal to the reference's own
    build (oracle/_ref/ref_delaunay_timing)."""
    (row,) = run("ref_bridge_delaunay", "rand48", 10000, HOST_ONLY)
    assert row["same_as_driver"] is True and row["cells"] == 66624 and row["finite"] == 66368
    (ref,) = run("ref_delaunay_timing", 10000, 1)
    assert row["timing_hash"] == ref["hash"]
    assert row["bridge"]["tests"] == row["bridge"]["host_tests"] > 400000


def test_bridge_lockstep_builds_host():
    """Eight builds advanced together, one bridge call per step for all
    pending regions: each equal to its own driver build."""
    (row,) = run("ref_bridge_delaunay", "lockstep", 8, 1500, HOST_ONLY)
    assert row["same_as_driver"] == 8
