"""The coplanar predicates (src/predicates2.cpp:111-259) on the CPU.

The oracle's restatement is pinned against the known answers of
tests/test_predicates2.cpp:64-90, against the big-integer apex method of
tests/exact.h:294-378 (which shares no formula with the library's 4x4
determinant) on exactly coplanar queries, including the reference's full
enumeration of a cocircular ring, and against the compiled reference on every
input, coplanar or not. The GPU's per-query core (del3d_amd/csrc/pred2_core.h,
host build) is then checked against the oracle with every tier, the exact
tiers also forced on queries the filters decide.
"""
import ctypes as C
import itertools
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
from del3d_amd import workloads as W

HOST = Path(__file__).resolve().parents[1] / "del3d_amd" / "libpredcore_host.so"
PENDING = 9  # the fast path's "left to the exact kernel" (pred2_core.h kPending2)
OPS = {"orient3": 0, "orient4": 1, "collinear": 2, "equal": 3, "position": 4, "in_circle": 5,
       "in_circle_raw": 6, "orient2d_xy": 7, "inexact": 8}


def host_core(pts, op, idx, mode):
    lib = C.CDLL(str(HOST))
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    op = np.ascontiguousarray(op, np.uint8)
    idx = np.ascontiguousarray(idx, np.int32)
    out = np.zeros(len(op), np.int8)
    lib.predcore2_host_batch(pts.ctypes.data_as(C.c_void_p), C.c_int64(len(pts)),
                             op.ctypes.data_as(C.c_void_p), idx.ctypes.data_as(C.c_void_p),
                             C.c_int64(len(op)), out.ctypes.data_as(C.c_void_p), C.c_int(mode))
    return out


def query(pts_list, ops_and_slots):
    pts = np.array(pts_list, dtype=np.float64)
    op = np.array([o for o, _ in ops_and_slots], np.uint8)
    idx = np.zeros((len(op), 5), np.int32)
    for i, (_, sl) in enumerate(ops_and_slots):
        idx[i, : len(sl)] = sl
    return pts, op, idx


def test_known_answers():
    # tests/test_predicates2.cpp:64-90, answers read off the geometry
    pts = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0.5, 0.5, 0), (9, 9, 0), (2, 0, 0),
           (0, 0, 0), (1, 0, 1), (0, 1, 1), (0.5, 0.5, 1.0)]
    a, b, c, centre, far, online, ta, tb, tc, tin = range(10)
    cases = [
        (OPS["collinear"], (a, b, online), 1),
        (OPS["collinear"], (a, b, c), 0),
        (OPS["orient3"], (a, b, c), 1),
        (OPS["orient3"], (a, c, b), -1),
        (OPS["in_circle_raw"], (a, b, c, centre), 1),
        (OPS["in_circle_raw"], (a, b, c, far), -1),
        (OPS["in_circle_raw"], (a, c, b, centre), 1),
        (OPS["in_circle_raw"], (ta, tb, tc, tin), 1),
        (OPS["in_circle"], (a, b, c, centre), 1),
        (OPS["in_circle"], (a, b, c, far), -1),
        # collinear_position on the line y = z = 0: before, source, middle,
        # target, after (include/del3d/predicates2.h:50-51)
        (OPS["position"], (b, a, online), 0),
        (OPS["position"], (a, a, online), 1),
        (OPS["position"], (a, b, online), 2),
        (OPS["position"], (a, online, online), 3),
        (OPS["position"], (a, online, b), 4),
        (OPS["equal"], (a, ta), 1),
        (OPS["equal"], (a, b), 0),
    ]
    P, op, idx = query(pts, [(o, s) for o, s, _ in cases])
    want = np.array([w for _, _, w in cases], np.int8)
    for which in ("oracle", "int"):
        assert np.array_equal(O.pred2_batch(P, op, idx, which), want), which
    if HOST.exists():
        assert np.array_equal(host_core(P, op, idx, 0), want)
    if O.ref() is not None:
        assert np.array_equal(O.pred2_batch(P, op, idx, "ref"), want)


def _precondition_mask(pts, op, idx):
    """Queries whose answer the apex oracle defines: a non-collinear reference
    triple for the orientation and in-circle ops, a collinear triple with
    p != r for collinear_position, every orient2d / collinear / equal."""
    n = len(op)
    col = O.pred2_batch(pts, np.full(n, OPS["collinear"], np.uint8), idx, "int")
    pr_eq = np.all(pts[idx[:, 0]] == pts[idx[:, 2]], axis=1)
    ok = np.ones(n, bool)
    tri = np.isin(op, [OPS["orient3"], OPS["orient4"], OPS["in_circle"], OPS["in_circle_raw"]])
    ok[tri] = col[tri] == 0
    pos = op == OPS["position"]
    ok[pos] = (col[pos] == 1) & ~pr_eq[pos]
    return ok


COPLANAR = [k for k in W.COPLANAR_KINDS if k != "random"]


@pytest.mark.parametrize("kind", COPLANAR)
def test_oracle_matches_apex_oracle(kind):
    rng = np.random.default_rng(abs(hash(kind)) % 2 ** 32)
    pts = W.coplanar_points(kind, rng)
    op, idx = W.coplanar_queries(len(pts), 6000, rng)
    ok = _precondition_mask(pts, op, idx)
    op, idx = op[ok], idx[ok]
    assert len(op) > 1000
    got = O.pred2_batch(pts, op, idx, "oracle")
    want = O.pred2_batch(pts, op, idx, "int")
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0, (kind, op[bad[:5]], idx[bad[:5]], got[bad[:5]], want[bad[:5]])


def test_ring_enumeration_matches_apex_oracle():
    # every (a, b, d, e) with a, b, d distinct over the ring, its centre and an
    # outside point, as tests/test_predicates2.cpp:122-139 enumerates them
    pts = W.coplanar_points("ring", None)
    n = len(pts)
    quads = np.array([(a, b, d, e) for a, b, d in itertools.permutations(range(n), 3)
                      for e in range(n)], np.int32)
    idx = np.zeros((len(quads), 5), np.int32)
    idx[:, :4] = quads
    for name in ("orient4", "in_circle", "in_circle_raw"):
        op = np.full(len(quads), OPS[name], np.uint8)
        ok = _precondition_mask(pts, op, idx)
        got = O.pred2_batch(pts, op[ok], idx[ok], "oracle")
        want = O.pred2_batch(pts, op[ok], idx[ok], "int")
        assert np.array_equal(got, want), name
        if name == "in_circle":
            assert O.stage_counts()[2] > 5000  # the perturbation decides many


@pytest.mark.parametrize("kind", W.COPLANAR_KINDS)
def test_oracle_matches_reference(kind):
    if O.ref() is None:
        pytest.skip("reference library not built")
    rng = np.random.default_rng(7 + len(kind))
    pts = W.coplanar_points(kind, rng)
    op, idx = W.coplanar_queries(len(pts), 8000, rng)
    got = O.pred2_batch(pts, op, idx, "oracle")
    want = O.pred2_batch(pts, op, idx, "ref")
    assert np.array_equal(got, want)


@pytest.mark.parametrize("kind", W.COPLANAR_KINDS)
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_gpu_core_on_host_matches_oracle(kind, mode):
    if not HOST.exists():
        pytest.fail(f"{HOST} not built")
    rng = np.random.default_rng(11 + 3 * len(kind))
    pts = W.coplanar_points(kind, rng)
    op, idx = W.coplanar_queries(len(pts), 6000, rng)
    want = O.pred2_batch(pts, op, idx, "oracle")
    got = host_core(pts, op, idx, mode)
    if mode == 1:  # the fast path's result: where it decides, it is final
        dec = got != PENDING
        assert dec.mean() > 0.2
        assert np.array_equal(got[dec], want[dec])
    else:
        assert np.array_equal(got, want)


def test_gpu_core_extreme_magnitudes():
    """Exactly coplanar queries spanning the double range, against the
    big-integer apex oracle: the exact tiers at their widest workspace (limb
    counts near the maximum). Products of such coordinates leave FP64's range,
    where the reference's expansion stage is no longer exact, so the check is
    against the true sign."""
    if not HOST.exists():
        pytest.fail(f"{HOST} not built")
    rng = np.random.default_rng(5)
    base = W.coplanar_points("scaled", rng, 80)[:40] * 2.0 ** 60  # on z = 3x/4 - 5y/8
    pts = np.vstack([base * 2.0 ** -1030, base * 2.0 ** 900, base, np.zeros((1, 3))])
    op, idx = W.coplanar_queries(len(pts), 3000, rng)
    ok = _precondition_mask(pts, op, idx)
    op, idx = op[ok], idx[ok]
    want = O.pred2_batch(pts, op, idx, "int")
    for mode in (0, 2):
        assert np.array_equal(host_core(pts, op, idx, mode), want), mode


def inexact_points(rng, n=300):
    """Points for the plain-double orientation: random doubles, near-coplanar
    quadruples (a random plane plus perturbations of a few ulps) and a lattice,
    so that many determinants round to tiny values of either sign or to 0."""
    a = rng.random((n, 3))
    base = rng.random((n, 2))
    plane = np.column_stack([base, 0.3 * base[:, 0] + 0.7 * base[:, 1] + 0.1])
    plane[:, 2] += rng.integers(-3, 4, size=n) * np.spacing(plane[:, 2])
    return np.vstack([a, plane, W.lattice(4, 4, 4)])


def test_inexact_orientation_matches_compiled_driver():
    """The driver's plain-double orientation (src/triangulation.cpp:296-308)
    against the driver's own code, compiled from its source (bit-exact, also
    where the rounded determinant's sign is wrong), and against the exact sign
    where the determinant is far from zero."""
    rng = np.random.default_rng(8)
    pts = inexact_points(rng)
    n = 20000
    idx = rng.integers(0, len(pts), size=(n, 5)).astype(np.int32)
    near = rng.integers(300, 600, size=(n // 2, 4))
    idx[: n // 2, :4] = near  # quadruples from the near-coplanar group
    op = np.full(n, OPS["inexact"], np.uint8)
    got = O.pred2_batch(pts, op, idx, "oracle")
    ex = O.pred2_batch(pts, op, idx, "int")
    assert (got != ex).sum() > 20          # the near-coplanar group does flip signs
    far = np.abs(np.linalg.det(pts[idx[:, 1:4]] - pts[idx[:, :1]])) > 1e-6
    assert np.array_equal(got[far], ex[far]) and far.sum() > n // 3
    if HOST.exists():
        assert np.array_equal(host_core(pts, op, idx, 0), got)
        assert np.array_equal(host_core(pts, op, idx, 1), got)
    if O.REF_DRIVER_SO.exists():
        assert np.array_equal(O.pred2_batch(pts, op, idx, "ref"), got)


@pytest.mark.parametrize("scale_exp", [-100, 0, 90])
def test_certificate_decides_small_grids_and_agrees(scale_exp):
    """The certified small-grid tier of the fast path (filter-only mode): on
    integer points whose differences stay below 2^6 (in-circle) every query is
    decided, including exactly zero determinants, and agrees with the oracle;
    with one coordinate planted half a unit off the grid, or differences of
    2^7 and more, the certificate must refuse, and whatever the fast path does
    decide still agrees with the oracle. Points are not coplanar: the bound
    holds for any points."""
    if not HOST.exists():
        pytest.fail(f"{HOST} not built")
    rng = np.random.default_rng(400 + scale_exp)
    s = 2.0 ** scale_exp
    small = rng.integers(0, 31, size=(200, 3)).astype(np.float64)   # differences < 2^5
    coplanar = np.column_stack([rng.integers(0, 31, size=(200, 2)), np.zeros(200)])
    coplanar[:, 2] = coplanar[:, 0] + 2 * coplanar[:, 1] - 40       # z = x + 2y - 40
    wide = rng.integers(0, 1 << 9, size=(200, 3)).astype(np.float64)
    planted = small.copy()
    planted[::7, 1] += 0.5
    n = 8000
    for name, pts, must_decide in (("small", small, True), ("coplanar", coplanar, False),
                                   ("wide", wide, False), ("planted", planted, False)):
        pts = pts * s
        op = rng.choice(np.array([OPS["in_circle_raw"], OPS["in_circle"], OPS["orient3"],
                                  OPS["collinear"], OPS["orient4"]], np.uint8), size=n)
        idx = rng.integers(0, len(pts), size=(n, 5)).astype(np.int32)
        idx[: n // 10, 3] = idx[: n // 10, 0]
        want = O.pred2_batch(pts, op, idx, "oracle")
        fast = host_core(pts, op, idx, 1)
        dec = fast != PENDING
        assert np.array_equal(fast[dec], want[dec]), name
        if name == "small":
            raw = op == OPS["in_circle_raw"]
            assert dec[raw].all(), name      # every in-circle determinant certified
        if name == "coplanar":
            assert dec.mean() > 0.95, (name, dec.mean())
