"""GPU parity of the coplanar predicates (del3d_predicates2 and the scalar
entry points) with the oracle's restatement of src/predicates2.cpp, on the
reference test's point sets (tests/test_predicates2.cpp:92-139: integer grids
on the z = 0 and a tilted plane, an exactly cocircular ring) and on harder
ones: dyadic planes, two clusters 2^120 apart, collinear input, points off
any common plane (no precondition), identical points and aliased slots.
Tolerance: none (every result bit-exact)."""
import itertools

import numpy as np
import pytest

import del3d_amd as D
import oracle_lib as O
from del3d_amd import workloads as W
from gpu_util import require_gpu, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


def run2(pts, op, idx):
    xyz = D.DeviceArray.from_host(np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3))
    dop = D.DeviceArray.from_host(np.ascontiguousarray(op, dtype=np.uint8))
    didx = D.DeviceArray.from_host(np.ascontiguousarray(idx, dtype=np.int32).reshape(-1, 5))
    out = D.DeviceArray((len(op),), np.int8)
    D.predicates2(xyz, dop, didx, out)
    D.synchronize()
    return out.to_host()


@pytest.mark.parametrize("kind", W.COPLANAR_KINDS)
def test_batch_matches_oracle(kind):
    rng = np.random.default_rng(100 + len(kind))
    pts = W.coplanar_points(kind, rng)
    op, idx = W.coplanar_queries(len(pts), 40000, rng)
    say("coplanar batch", kind, len(op))
    got = run2(pts, op, idx)
    want = O.pred2_batch(pts, op, idx, "oracle")
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0, (kind, op[bad[:5]], idx[bad[:5]], got[bad[:5]], want[bad[:5]])
    if O.ref() is not None:
        assert np.array_equal(got, O.pred2_batch(pts, op, idx, "ref"))


def test_ring_enumeration():
    """All (a, b, d, e) with a, b, d distinct over the ring of
    tests/test_predicates2.cpp:122-139: mostly exactly cocircular, so nearly
    every in-circle query goes through the exact tier and the perturbation."""
    pts = W.coplanar_points("ring_tilted", None)
    n = len(pts)
    quads = np.array([(a, b, d, e) for a, b, d in itertools.permutations(range(n), 3)
                      for e in range(n)], np.int32)
    idx = np.zeros((3 * len(quads), 5), np.int32)
    idx[:, :4] = np.vstack([quads] * 3)
    op = np.repeat(np.array([D.COPLANAR_ORIENTATION4, D.IN_CIRCLE, D.IN_CIRCLE_UNPERTURBED],
                            np.uint8), len(quads))
    say("ring enumeration", len(op))
    got = run2(pts, op, idx)
    assert np.array_equal(got, O.pred2_batch(pts, op, idx, "oracle"))
    assert O.stage_counts()[2] > 5000


def test_extreme_magnitudes_true_sign():
    """Coplanar points 2^-1030 .. 2^960: the widest exact workspace. Checked
    against the big-integer apex oracle (products leave FP64's range there, so
    the true sign is the reference's only where its arithmetic is exact)."""
    rng = np.random.default_rng(9)
    base = W.coplanar_points("scaled", rng, 80)[:40] * 2.0 ** 60
    pts = np.vstack([base * 2.0 ** -1030, base * 2.0 ** 900, base, np.zeros((1, 3))])
    op, idx = W.coplanar_queries(len(pts), 3000, rng)
    col = O.pred2_batch(pts, np.full(len(op), D.COLLINEAR, np.uint8), idx, "int")
    tri = np.isin(op, [0, 1, 5, 6])
    pos = op == D.COLLINEAR_POSITION
    ok = (~tri | (col == 0)) & ~pos
    op, idx = op[ok], idx[ok]
    got = run2(pts, op, idx)
    assert np.array_equal(got, O.pred2_batch(pts, op, idx, "int"))


def test_bad_index_and_unknown_op_give_zero():
    pts = W.coplanar_points("grid_z0", None)
    op = np.array([D.IN_CIRCLE, D.COPLANAR_ORIENTATION, 9, 255, D.EQUAL, D.EQUAL], np.uint8)
    idx = np.array([[0, 1, 2, len(pts), 0], [-1, 1, 2, 0, 0], [0, 1, 2, 3, 0], [0, 1, 2, 3, 0],
                    [5, 5, 10**6, -7, 3], [5, 6, 0, 0, 0]], np.int32)
    got = run2(pts, op, idx)
    assert got.tolist() == [0, 0, 0, 0, 1, 0]


def test_empty_batch_and_host_form():
    pts = W.coplanar_points("tilted", None)
    assert len(run2(pts, np.zeros(0, np.uint8), np.zeros((0, 5), np.int32))) == 0
    rng = np.random.default_rng(4)
    op, idx = W.coplanar_queries(len(pts), 5000, rng)
    assert np.array_equal(D.predicates2_np(pts, op, idx), O.pred2_batch(pts, op, idx, "oracle"))


def test_large_batch_matches_host_core_and_oracle_sample():
    """2^22 queries in one call on a dyadic plane and the tilted ring: the
    whole batch against the host build of the same per-query core, and a
    sample of 100,000 against the oracle."""
    import test_coplanar as TC
    rng = np.random.default_rng(12)
    pts = np.vstack([W.coplanar_points("dyadic_plane", rng, 3000),
                     W.coplanar_points("ring_tilted", None)])
    op, idx = W.coplanar_queries(len(pts), 1 << 22, rng)
    say("large coplanar batch", len(op))
    got = run2(pts, op, idx)
    assert np.array_equal(got, TC.host_core(pts, op, idx, 0))
    s = rng.choice(len(op), 100000, replace=False)
    assert np.array_equal(got[s], O.pred2_batch(pts, op[s], idx[s], "oracle"))


def test_scalar_entry_points_known_answers():
    """tests/test_predicates2.cpp:64-90 through the scalar C entry points."""
    a, b, c = [0.0, 0, 0], [1.0, 0, 0], [0.0, 1, 0]
    centre, far, online = [0.5, 0.5, 0.0], [9.0, 9, 0], [2.0, 0, 0]
    assert D.collinear(a, b, online) and not D.collinear(a, b, c)
    assert D.coplanar_orientation(a, b, c) == 1 and D.coplanar_orientation(a, c, b) == -1
    assert D.coplanar_orientation(a, b, c, centre) == 1
    assert D.coplanar_in_circle_unperturbed(a, b, c, centre) == 1
    assert D.coplanar_in_circle_unperturbed(a, b, c, far) == -1
    assert D.coplanar_in_circle_unperturbed(a, c, b, centre) == 1
    ta, tb, tc, tin = [0.0, 0, 0], [1.0, 0, 1], [0.0, 1, 1], [0.5, 0.5, 1.0]
    assert D.coplanar_in_circle_unperturbed(ta, tb, tc, tin) == 1
    assert D.coplanar_in_circle(a, b, c, centre) == 1 and D.coplanar_in_circle(a, b, c, far) == -1
    assert D.orient2d(0, 0, 1, 0, 0, 1) == 1 and D.orient2d(0, 0, 1, 1, 2, 2) == 0
    assert D.equal(a, [0.0, 0.0, 0.0]) and not D.equal(a, b)
    assert [D.collinear_position(b, a, online), D.collinear_position(a, a, online),
            D.collinear_position(a, b, online), D.collinear_position(a, online, online),
            D.collinear_position(a, online, b)] == [0, 1, 2, 3, 4]
    # a cocircular query decided by the perturbation, with the query point
    # aliasing a vertex (same object = same pointer)
    r = [[5.0, 0, 0], [0.0, 5, 0], [-5.0, 0, 0], [0.0, -5, 0]]
    pts = np.array(r)
    for quad in ([0, 1, 2, 3], [1, 0, 2, 3], [0, 1, 2, 0], [2, 1, 0, 1]):
        args = [r[i] for i in quad]
        idx = np.array([quad + [0]], np.int32)
        want = O.pred2_batch(pts, np.array([D.IN_CIRCLE], np.uint8), idx, "oracle")[0]
        assert D.coplanar_in_circle(*args) == want, quad


TRACES = {
    "rand48_10000": ["10000", "1"],        # the 471,224 calls of the C1 build
    "lattice_16^3": ["lattice", "16", "16", "16"],
    "lattice_40x30x1": ["lattice", "40", "30", "1"],
    "lattice_64x1x1": ["lattice", "64", "1", "1"],
}


@pytest.mark.skipif(not O.TRACE_RECORDER.exists(), reason="trace recorder not built")
@pytest.mark.parametrize("name", list(TRACES))
def test_driver_trace_replay(name, tmp_path):
    """Every exact predicate call (3D and coplanar) the reference's Delaunay
    build makes, recorded by the reference itself, replayed in one mixed
    batch on the GPU: identical results."""
    import subprocess
    from del3d_amd import trace as T
    out = tmp_path / "t.trace"
    say("recording", name)
    subprocess.run([str(O.TRACE_RECORDER)] + TRACES[name] + [str(out)], check=True, timeout=300,
                   capture_output=True)
    pts, op, res, idx = T.read(out)
    say("replaying", name, len(op), "records,", int((op >= 16).sum()), "coplanar")
    got = T.replay(pts, op, idx)
    bad = np.nonzero(got != res)[0]
    assert len(bad) == 0, (name, op[bad[:5]], idx[bad[:5]], got[bad[:5]], res[bad[:5]])


def test_mixed_batch_all_ops():
    """3D, coplanar and inexact-orientation queries interleaved in one mixed
    batch, against the per-family batches and the oracle."""
    rng = np.random.default_rng(31)
    pts = np.vstack([W.coplanar_points("ring_tilted", None), W.lattice(4, 4, 4),
                     rng.random((200, 3))])
    n = 60000
    op = rng.choice(np.array([0, 1, 2] + [16 + k for k in range(9)] + [3, 15, 25, 255], np.uint8),
                    size=n)
    idx = rng.integers(0, len(pts), size=(n, 5)).astype(np.int32)
    say("mixed batch", n)
    got = T_replay(pts, op, idx)
    want = np.zeros(n, np.int8)
    m3 = op <= 2
    want[m3] = O.pred_batch(pts, op[m3], idx[m3], "oracle")
    m2 = (op >= 16) & (op <= 24)
    want[m2] = O.pred2_batch(pts, op[m2] - 16, idx[m2], "oracle")
    assert np.array_equal(got, want)
    assert np.array_equal(got[m2], run2(pts, op[m2] - 16, idx[m2]))


def T_replay(pts, op, idx):
    from del3d_amd import trace as T
    return T.replay(pts, op, idx)


def test_inexact_orientation_matches_driver():
    """The driver's plain-double orientation on near-coplanar quadruples
    (signs of rounded determinants) against the oracle and, where built, the
    driver's own code."""
    import test_coplanar as TC
    rng = np.random.default_rng(8)
    pts = TC.inexact_points(rng)
    n = 20000
    idx = rng.integers(0, len(pts), size=(n, 5)).astype(np.int32)
    idx[: n // 2, :4] = rng.integers(300, 600, size=(n // 2, 4))
    op = np.full(n, D.INEXACT_ORIENTATION, np.uint8)
    got = run2(pts, op, idx)
    assert np.array_equal(got, O.pred2_batch(pts, op, idx, "oracle"))
    if O.REF_DRIVER_SO.exists():
        assert np.array_equal(got, O.pred2_batch(pts, op, idx, "ref"))
