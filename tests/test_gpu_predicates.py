"""GPU predicate parity: every sign equals the oracle's (bit-exact, int8),
on the reference's own test workloads, the measured configurations and the
hard cases (wide exponent spans, denormals, aliases, ties)."""
import json
from pathlib import Path

import numpy as np
import pytest

import del3d_amd as D
import oracle_lib as O
from del3d_amd import workloads as W
from gpu_util import predicates_dev, require_gpu, say

pytestmark = pytest.mark.gpu
PINS = json.loads((Path(__file__).parent / "golden" / "reference_pins.json").read_text())


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


def test_known_signs_scalar():
    k = PINS["known_signs"]
    o, x, y, z = (np.array(k[n], dtype=np.float64) for n in "oxyz")
    ins, out = np.array(k["inside"]), np.array(k["outside"])
    say("scalar known signs")
    assert D.orient3d(o, x, y, z) == k["orient3d(o,x,y,z)"]
    assert D.orient3d(o, y, x, z) == k["orient3d(o,y,x,z)"]
    assert D.in_sphere_unperturbed(o, x, y, z, ins) == k["in_sphere_raw(o,x,y,z,inside)"]
    assert D.in_sphere_unperturbed(o, x, y, z, out) == k["in_sphere_raw(o,x,y,z,outside)"]
    assert D.in_sphere(o, x, y, z, ins) == k["in_sphere(o,x,y,z,inside)"]
    assert D.in_sphere(o, x, y, z, out) == k["in_sphere(o,x,y,z,outside)"]
    # aliased e (same object = same pointer) goes through the perturbation
    assert D.in_sphere(o, x, y, z, o) == O.oracle().orc_in_sphere(
        *[a.ctypes.data_as(O._dp) for a in (o, x, y, z, o)])


@pytest.mark.parametrize("name", ["cylinder", "sphere", "torus", "box_shell"])
def test_reference_stage2_workload(name):
    """tests/test_predicates.cpp:107-129: Rng(12345) 5-tuples over quantised sets."""
    pts = {"cylinder": O.cylinder(400, 8), "sphere": O.sphere(400, 8),
           "torus": O.torus(400, 8), "box_shell": O.box_shell(6)}[name]
    draws = O.rng_below(12345, len(pts), 20000 * 5).reshape(20000, 5).astype(np.int32)
    for op in (0, 1, 2):
        ops = np.full(20000, op, np.uint8)
        say(name, "op", op)
        got, _ = predicates_dev(pts, ops, draws)
        want = O.pred_batch(pts, ops, draws, "oracle")
        assert np.array_equal(got, want), (name, op, int((got != want).sum()))
    if name == "box_shell":
        orient, _ = predicates_dev(pts, np.zeros(20000, np.uint8), draws)
        raw, _ = predicates_dev(pts, np.full(20000, 2, np.uint8), draws)
        assert int((orient == 0).sum()) == PINS["degenerate_counts"]["box_shell_6_exactly_coplanar"]
        assert int((raw[orient == 1] == 0).sum()) == PINS["degenerate_counts"]["box_shell_6_exactly_cospherical"]


def test_c2_lattice_sample_vs_oracle():
    pts, op, idx = W.c2(1 << 18, seed=2)
    say("c2 2^18")
    got, st = predicates_dev(pts, op, idx, stats=True)
    want = O.pred_batch(pts, op, idx, "oracle")
    assert np.array_equal(got, want), int((got != want).sum())
    assert st["perturbed"] == int((op == 1).sum())   # every C2 in_sphere query is cospherical
    assert st["filter"] + st["certified"] + st["exact_regs"] + st["exact_wide"] == len(op)


def test_c2_full_size_properties():
    """2^24 queries: a 2^16 sample against the oracle, and over the whole batch
    the exact identities orient3d(b,a,c,d) = -orient3d(a,b,c,d) and
    in_sphere != 0."""
    pts, op, idx = W.c2(1 << 24, seed=2)
    say("c2 2^24")
    got, st = predicates_dev(pts, op, idx, stats=True)
    rng = np.random.default_rng(0)
    sample = rng.choice(len(op), size=1 << 16, replace=False)
    want = O.pred_batch(pts, op[sample], idx[sample], "oracle")
    assert np.array_equal(got[sample], want)
    assert np.all(got[op == 1] != 0)
    swapped = idx.copy()
    swapped[:, [0, 1]] = swapped[:, [1, 0]]
    o = np.zeros(len(op), np.uint8)
    g1, _ = predicates_dev(pts, o, idx)
    g2, _ = predicates_dev(pts, o, swapped)
    assert np.array_equal(g1, -g2)
    assert np.array_equal(g1[op == 0], got[op == 0])
    say("stats", st)


def test_random_doubles_vs_oracle():
    pts = W.c3_points(1_000_000, seed=3)
    op, idx = W.random_queries(len(pts), 1 << 20, seed=5)
    say("c3-like 1e6 points 2^20 queries")
    got, st = predicates_dev(pts, op, idx, stats=True)
    want = O.pred_batch(pts, op, idx, "oracle")
    assert np.array_equal(got, want)


def hard_sets():
    rng = np.random.default_rng(17)
    v = rng.normal(size=(300, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return {
        "near_sphere_ulps": v * (1 + rng.integers(-3, 4, size=(300, 1)) * 2.0 ** -52),
        "wide_exponents": rng.normal(size=(200, 3)) * np.exp2(rng.integers(-80, 80, size=(200, 3))),
        "very_wide": rng.normal(size=(60, 3)) * np.exp2(rng.integers(-900, 900, size=(60, 3))),
        "tiny": rng.normal(size=(100, 3)) * 1e-100,
        "grid_2^-48": rng.integers(0, 2 ** 48, size=(300, 3)) * 2.0 ** -48,
        "quantised_2^-10_sphere": np.round(v * 1024) / 1024,
        "duplicates": np.vstack([O.lattice(3, 3, 3)] * 2),
        "identical": np.tile([0.5, -1.25, 3.0], (20, 1)),
        "denormals": np.array([[5e-324, 0, 0], [0, 5e-324, 0], [0, 0, 5e-324], [1e-310, 1e-310, 0],
                               [0, 0, 0], [2.5e-323, 1e-323, 0], [-0.0, 0.0, -5e-324]]),
    }


@pytest.mark.parametrize("name", list(hard_sets().keys()))
def test_hard_cases_vs_int_oracle(name):
    pts = hard_sets()[name]
    rng = np.random.default_rng(abs(hash(name)) % 2 ** 32)
    n = 20000
    idx = rng.integers(0, len(pts), size=(n, 5)).astype(np.int32)
    idx[:3000, 4] = idx[:3000, 0]
    idx[3000:5000, 3] = idx[3000:5000, 1]
    op = rng.integers(0, 3, size=n).astype(np.uint8)
    say("hard", name)
    got, st = predicates_dev(pts, op, idx, stats=True)
    want = O.pred_batch(pts, op, idx, "int")
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0, (name, bad[:5], op[bad[:5]], got[bad[:5]], want[bad[:5]], st)
    if name == "very_wide":
        assert st["exact_wide"] > 0          # the LDS tier ran


def test_tails_and_empty():
    """Ragged sizes through both forms a single call can take: the one-pass
    kernel (a plain call of at most 2^19 queries) and the pipeline (the
    statistics form, and a plain call above 2^19)."""
    pts = O.lattice(3, 3, 3)
    for n in (0, 1, 63, 255, 256, 257, 1000, (1 << 19) + 1):
        rng = np.random.default_rng(n)
        idx = rng.integers(0, len(pts), size=(n, 5)).astype(np.int32)
        op = rng.integers(0, 3, size=n).astype(np.uint8)
        say("tail", n)
        want = O.pred_batch(pts, op, idx, "oracle")
        got, _ = predicates_dev(pts, op, idx)
        assert np.array_equal(got, want)
        got, _ = predicates_dev(pts, op, idx, stats=True)
        assert np.array_equal(got, want)


def test_host_memory_form():
    pts, op, idx = W.c2(5000, seed=9)
    say("host form")
    assert np.array_equal(D.predicates_np(pts, op, idx), O.pred_batch(pts, op, idx, "oracle"))


def _plan_signs(pts_list, op_list, idx_list, knobs=None, graph=False, replays=1):
    """Runs a plan over several sets (set-local indices); returns the signs."""
    sizes = [len(p) for p in pts_list]
    nqs = [len(o) for o in op_list]
    po = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    qo = np.concatenate([[0], np.cumsum(nqs)]).astype(np.int64)
    xyz = D.DeviceArray.from_host(np.vstack(pts_list).astype(np.float64))
    dop = D.DeviceArray.from_host(np.concatenate(op_list).astype(np.uint8))
    didx = D.DeviceArray.from_host(np.vstack(idx_list).astype(np.int32))
    out = D.DeviceArray((int(qo[-1]),), np.int8)
    plan = D.PredPlan(po, qo)
    if knobs is not None:
        k = plan.knobs
        for i, v in knobs.items():
            k[i] = v
        plan.knobs = k
    s = D.Stream()
    results = []
    for _ in range(replays):
        out.zero()
        plan.run(xyz, dop, didx, out, s, graph=graph)
        s.synchronize()
        results.append(out.to_host())
    plan.close()
    return results


@pytest.mark.parametrize("name", ["lattice_5^3", "very_wide", "quantised_2^-10_sphere", "denormals"])
def test_lds_and_global_paths_agree_with_oracle(name):
    """The same queries through the LDS kernel (set held in LDS) and the
    global-gather kernel (routing threshold 0), both against the oracle."""
    pts = O.lattice(5, 5, 5) if name == "lattice_5^3" else hard_sets()[name]
    rng = np.random.default_rng(3)
    n = 30000
    idx = rng.integers(0, len(pts), size=(n, 5)).astype(np.int32)
    idx[:4000, 4] = idx[:4000, 2]
    op = rng.integers(0, 3, size=n).astype(np.uint8)
    want = O.pred_batch(pts, op, idx, "int")
    say("paths", name)
    lds = _plan_signs([pts], [op], [idx], knobs={0: 5632, 1: 4096})[0]
    glob = _plan_signs([pts], [op], [idx], knobs={0: 0, 2: 1024})[0]
    assert np.array_equal(lds, want)
    assert np.array_equal(glob, want)


def test_out_of_range_index_gives_zero():
    pts = O.lattice(3, 3, 3)
    op = np.array([0, 1, 2, 1, 3, 200], np.uint8)
    idx = np.array([[0, 1, 3, 27, 0], [0, 1, 3, 9, -1], [0, 1, 3, 9, 13], [0, 1, 3, 9, 1 << 30],
                    [0, 1, 3, 9, 13], [0, 1, 3, 9, 13]], np.int32)
    xyz = D.DeviceArray.from_host(pts)
    dop, didx = D.DeviceArray.from_host(op), D.DeviceArray.from_host(idx)
    out = D.DeviceArray((6,), np.int8)
    st = D.predicates_stats(xyz, dop, didx, out)
    got = out.to_host()
    assert st["bad_index"] == 5  # three bad indices, two unknown op codes
    assert got[0] == 0 and got[1] == 0 and got[3] == 0 and got[4] == 0 and got[5] == 0
    assert got[2] == O.pred_batch(pts, op[2:3], idx[2:3], "oracle")[0]


def _sphere_points(radius):
    """Every integer point on the sphere x^2 + y^2 + z^2 = radius^2."""
    x = np.arange(-radius, radius + 1)
    X, Y = np.meshgrid(x, x, indexing="ij")
    r = radius * radius - X * X - Y * Y
    ok = r >= 0
    Z = np.sqrt(np.where(ok, r, 0)).astype(np.int64)
    ok &= Z * Z == r
    neg = ok & (Z > 0)
    return np.concatenate([np.stack([X[ok], Y[ok], Z[ok]], 1),
                           np.stack([X[neg], Y[neg], -Z[neg]], 1)]).astype(np.float64)


@pytest.mark.parametrize("radius,kernel", [(625, "lds"), (975, "global")])
def test_exact_queue_buffers_with_bad_entries(radius, kernel):
    """Five points of an integer sphere are exactly cospherical and their
    differences exceed the certified grid, so every in_sphere goes through the
    exact queue: each wavefront's queue buffer fills and is appended many
    times, in the LDS kernel (3,750 points) and the global one (9,750).
    Orient3d queries, unknown op codes and out-of-range indices are mixed in,
    so the regrouped packs mix kinds; the bad ones give 0 and are counted."""
    pts = _sphere_points(radius)
    rng = np.random.default_rng(radius)
    n = 1 << 17
    idx = rng.integers(0, len(pts), size=(n, 5)).astype(np.int32)
    op = np.where(rng.random(n) < 0.8, 1, rng.integers(0, 3, size=n)).astype(np.uint8)
    bad_op = rng.random(n) < 0.01
    op[bad_op] = rng.integers(3, 256, size=int(bad_op.sum()))
    bad_idx = (rng.random(n) < 0.01) & ~bad_op
    rows = np.nonzero(bad_idx)[0]
    idx[rows, rng.integers(0, 4, size=len(rows))] = len(pts) + rng.integers(0, 1000, size=len(rows))
    say("exact queue", kernel, len(pts), "points")
    got, st = predicates_dev(pts, op, idx, stats=True)
    bad = bad_op | bad_idx
    assert st["bad_index"] == int(bad.sum())
    assert np.all(got[bad] == 0)
    ok = ~bad
    want = O.pred_batch(pts, op[ok], idx[ok], "oracle")
    assert np.array_equal(got[ok], want)
    assert st["exact_regs"] + st["exact_wide"] > n // 2


def _mixed_sets(n_sets, seed, max_pts=12000):
    """Sets of log-uniform size (uniform cube, quantised sphere, jittered
    lattice), 4 queries per point, set-local indices."""
    rng = np.random.default_rng(seed)
    pts_l, op_l, idx_l = [], [], []
    for s in range(n_sets):
        n = int(10 * (max_pts / 10) ** rng.random())
        kind = s % 3
        if kind == 0:
            p = rng.random((n, 3))
        elif kind == 1:
            v = rng.normal(size=(n, 3))
            p = np.round(v / np.linalg.norm(v, axis=1, keepdims=True) * 1024) / 1024
        else:
            m = max(2, int(round(n ** (1 / 3))))
            g = np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n]
            p = g + np.round(rng.random(g.shape) * 64) / 1024
            n = len(p)
        nq = 4 * n
        pts_l.append(p)
        op_l.append(rng.integers(0, 3, size=nq).astype(np.uint8))
        idx_l.append(rng.integers(0, n, size=(nq, 5)).astype(np.int32))
    return pts_l, op_l, idx_l


def test_sets_match_single_set_calls():
    pts_l, op_l, idx_l = _mixed_sets(60, seed=9)
    say("sets", sum(len(p) for p in pts_l), "points", sum(len(o) for o in op_l), "queries")
    got = _plan_signs(pts_l, op_l, idx_l)[0]
    off = 0
    for p, o, i in zip(pts_l, op_l, idx_l):
        alone, _ = predicates_dev(p, o, i)
        assert np.array_equal(got[off:off + len(o)], alone)
        off += len(o)
    want = np.concatenate([O.pred_batch(p, o, i, "oracle") for p, o, i in zip(pts_l[:12], op_l[:12], idx_l[:12])])
    assert np.array_equal(got[:len(want)], want)


def test_sets_call_and_graph_replays():
    """del3d_predicates_sets, a plan run, and five replays of the captured
    graph (with other GPU work between them) all give the same signs."""
    pts_l, op_l, idx_l = _mixed_sets(40, seed=10)
    eager = _plan_signs(pts_l, op_l, idx_l)[0]
    sizes = [len(p) for p in pts_l]
    po = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    qo = np.concatenate([[0], np.cumsum([len(o) for o in op_l])]).astype(np.int64)
    xyz = D.DeviceArray.from_host(np.vstack(pts_l))
    dop = D.DeviceArray.from_host(np.concatenate(op_l))
    didx = D.DeviceArray.from_host(np.vstack(idx_l))
    out = D.DeviceArray((int(qo[-1]),), np.int8)
    D.predicates_sets(xyz, po, dop, didx, qo, out)
    assert np.array_equal(out.to_host(), eager)
    plan = D.PredPlan(po, qo)
    s = D.Stream()
    for r in range(5):
        out.zero()
        plan.run(xyz, dop, didx, out, s, graph=True)
        s.synchronize()
        assert np.array_equal(out.to_host(), eager), r
        predicates_dev(pts_l[0], op_l[0], idx_l[0])  # eager work between replays
    ms = plan.autotune(xyz, dop, didx, out, s)
    say("autotuned", plan.knobs, ms)
    out.zero()
    plan.run(xyz, dop, didx, out, s, graph=True)
    s.synchronize()
    assert np.array_equal(out.to_host(), eager)
    plan.close()


def test_plan_xcd_bins_and_shared_list_agree():
    """Global tasks in per-XCD bins (with stealing once a bin is empty) and
    in one shared list, at several task sizes and occupancies, all routed to
    the global kernel: every sign equals the single-set calls'."""
    pts_l, op_l, idx_l = _mixed_sets(48, seed=12)
    sizes = [len(p) for p in pts_l]
    po = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    qo = np.concatenate([[0], np.cumsum([len(o) for o in op_l])]).astype(np.int64)
    want = np.concatenate([predicates_dev(p, o, i)[0] for p, o, i in zip(pts_l, op_l, idx_l)])
    xyz = D.DeviceArray.from_host(np.vstack(pts_l))
    dop = D.DeviceArray.from_host(np.concatenate(op_l))
    didx = D.DeviceArray.from_host(np.vstack(idx_l))
    out = D.DeviceArray((int(qo[-1]),), np.int8)
    plan = D.PredPlan(po, qo)
    s = D.Stream()
    for knobs in ([0, 32768, 256, 1, 1], [0, 32768, 256, 1, 0], [0, 32768, 4096, 8, 1],
                  [5632, 32768, 1024, 4, 1], [0, 32768, 1 << 20, 2, 1]):
        say("knobs", knobs)
        plan.knobs = knobs
        out.zero()
        plan.run(xyz, dop, didx, out, s, graph=True)
        s.synchronize()
        assert np.array_equal(out.to_host(), want), knobs
    plan.close()


def test_plan_prepass_across_set_boundaries():
    """The plan's prepasses (each set's grid summary and the float copy of
    the sets exact in float) read contiguous ranges that cross set
    boundaries, with empty sets, one-point sets and sets whose coordinates
    start at odd doubles: every sign equals the oracle's and the single-set
    calls'."""
    rng = np.random.default_rng(31)
    kinds = [(0, "cube"), (5, "quant"), (0, "cube"), (1, "cube"), (341, "quant"), (342, "cube"),
             (700, "lattice"), (0, "quant"), (6000, "quant"), (3, "lattice"), (9001, "cube"),
             (40000, "lattice"), (70001, "cube"), (2, "quant"), (20000, "quant")]
    pts_l, op_l, idx_l = [], [], []
    for n, kind in kinds:
        if kind == "cube":
            p = rng.random((n, 3))
        elif kind == "quant":  # sphere points on a 2^-10 grid: exact in float
            v = rng.normal(size=(n, 3))
            p = np.round(v / np.maximum(np.linalg.norm(v, axis=1, keepdims=True), 1e-300) * 1024) / 1024
        else:  # integer lattice, jittered on a 2^-10 grid: exact in float
            m = max(2, int(np.ceil(n ** (1 / 3))))
            g = np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n]
            p = g + np.round(rng.random(g.shape) * 8) / 1024
        nq = 3 * n
        pts_l.append(np.asarray(p, dtype=np.float64).reshape(n, 3))
        op_l.append(rng.integers(0, 3, size=nq).astype(np.uint8))
        idx_l.append(rng.integers(0, max(n, 1), size=(nq, 5)).astype(np.int32))
    say("prepass sets", [n for n, _ in kinds])
    got = _plan_signs(pts_l, op_l, idx_l, graph=True, replays=2)
    assert np.array_equal(got[0], got[1])
    off = 0
    for p, o, i in zip(pts_l, op_l, idx_l):
        nq = len(o)
        if nq:
            assert np.array_equal(got[0][off:off + nq], O.pred_batch(p, o, i, "oracle")), len(p)
            assert np.array_equal(got[0][off:off + nq], predicates_dev(p, o, i)[0]), len(p)
        off += nq


@pytest.mark.parametrize("kind", ["float_exact_lattice", "lattice_2^-120", "lattice_2^-135",
                                  "wide_integers", "random_doubles", "packed_lattice",
                                  "packed_lattice_2^-200"])
def test_global_kernel_point_formats(kind):
    """Single sets above the packed LDS kernel's capacity (34^3 = 39,304
    points), so the global kernel gathers them: a half-integer lattice moved
    by 1000 (too wide to pack) and the same scaled by 2^-120 (exact normal
    floats: gathered from the 16-byte float copy), scaled by 2^-135 (float
    subnormals: doubles), integers up to 2^30 (certifiable, not exact in
    float: doubles), random doubles (neither), and the unmoved lattice at 2^0
    and 2^-200 (packable: the global kernel's packed variant gathers 4-byte
    words and decides on the grid) — each against the oracle."""
    rng = np.random.default_rng(77)
    m = 34
    if "lattice" in kind:
        pts = W.lattice(m, m, m) * np.array([0.5, 0.25, 1.0]) - 3.0
        if not kind.startswith("packed"):
            pts = pts + 1000.0
        if "120" in kind or "135" in kind:
            pts = pts * 2.0 ** (-120 if "120" in kind else -135)
        if "200" in kind:
            pts = pts * 2.0 ** -200
    elif kind == "wide_integers":
        pts = rng.integers(-(1 << 30), 1 << 30, size=(40000, 3)).astype(np.float64)
        pts[::3] = np.round(pts[::3] / 1024.0)  # some close neighbours as well
    else:
        pts = rng.random((40000, 3))
    n = 1 << 17
    op = rng.integers(0, 3, size=n).astype(np.uint8)
    idx = rng.integers(0, len(pts), size=(n, 5)).astype(np.int32)
    if "lattice" in kind:  # local cells: many degenerate queries
        base = rng.integers(0, m - 2, size=(n, 3))
        off = rng.integers(0, 3, size=(n, 5, 3))
        ijk = base[:, None, :] + off
        idx = ((ijk[..., 0] * m + ijk[..., 1]) * m + ijk[..., 2]).astype(np.int32)
    say("global kernel", kind, len(pts), "points")
    got, st = predicates_dev(pts, op, idx, stats=True)
    assert np.array_equal(got, O.pred_batch(pts, op, idx, "oracle"))
    if "lattice" in kind:
        assert st["perturbed"] > 1000


def _packed_case(kind, rng):
    """(points, lattice side or 0) of a set routed to the packed LDS kernel or,
    when its values span too many bits, to the global kernel beside it."""
    if kind == "lattice32":  # the headline set
        return W.lattice(32, 32, 32), 32
    if kind == "lattice32_2^-200":  # near the smallest packable grid, 2^-206
        return W.lattice(32, 32, 32) * 2.0 ** -200, 32
    if kind == "lattice32_2^-1000":
        return W.lattice(32, 32, 32) * 2.0 ** -1000, 32
    if kind == "centered_signed_zeros":  # -0.0 where the lattice has 0
        pts = W.lattice(24, 24, 24) - 12.0
        return np.where(pts == 0.0, -0.0, pts), 24
    if kind == "span255":  # the widest packable span, random integers
        return rng.integers(-255, 256, size=(20000, 3)).astype(np.float64), 0
    if kind == "span256":  # one bit too wide: the global kernel runs
        pts = rng.integers(-255, 256, size=(20000, 3)).astype(np.float64)
        pts[7, 1] = 256.0
        return pts, 0
    if kind == "smallest":  # kLdsPts + 1 points
        return rng.integers(-8, 8, size=(5633, 3)).astype(np.float64) * 0.125, 0
    if kind == "largest":  # kPackPts points
        return rng.integers(-20, 20, size=(33792, 3)).astype(np.float64), 0
    return rng.random((20000, 3)), 0  # random doubles: the global kernel runs


@pytest.mark.parametrize("kind", ["lattice32", "lattice32_2^-200", "lattice32_2^-1000", "centered_signed_zeros", "span255",
                                  "span256", "smallest", "largest", "random_doubles"])
def test_packed_kernel_point_formats(kind):
    """Single sets of 5,633..33,792 points run the packed LDS kernel (one
    32-bit word a point, the whole set in LDS) when their values span at most
    8 bits on a grid 2^gmin with -206 <= gmin <= 171, else the global kernel:
    each against the oracle, with the perturbation exercised on the lattices.
    The lattice at 2^-1000 (global kernel) has products that underflow FP64,
    where the reference's expansions are not exact and the library returns
    the true sign: that case is checked against the big-integer oracle."""
    rng = np.random.default_rng(91)
    pts, m = _packed_case(kind, rng)
    n = 1 << 17
    op = rng.integers(0, 3, size=n).astype(np.uint8)
    idx = rng.integers(0, len(pts), size=(n, 5)).astype(np.int32)
    if m:  # local cells: many degenerate queries
        base = rng.integers(0, m - 2, size=(n, 3))
        off = rng.integers(0, 3, size=(n, 5, 3))
        ijk = base[:, None, :] + off
        idx = ((ijk[..., 0] * m + ijk[..., 1]) * m + ijk[..., 2]).astype(np.int32)
    say("packed route", kind, len(pts), "points")
    got, st = predicates_dev(pts, op, idx, stats=True)
    assert np.array_equal(got, O.pred_batch(pts, op, idx, "int" if "1000" in kind else "oracle"))
    if m:
        assert st["perturbed"] > 1000


def test_plan_more_queries_than_one_launch_sequence():
    """A plan of 2^28 + 2^20 queries, above the 2^28 of one of a plan's launch
    sequences: an LDS set first, then the global sets largest first, so the
    second global set is split between the two sequences. Each set's queries
    are copies of one 2^24 batch: every sign, in a run and a graph replay,
    equals that batch's single-set signs, which a sample checks against the
    oracle."""
    rng = np.random.default_rng(35)
    m = 34
    g = np.arange(m, dtype=np.float64)
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    n = 1 << 24
    base = rng.integers(0, m - 2, size=(n, 3))
    ijk = base[:, None, :] + rng.integers(0, 3, size=(n, 5, 3))  # local cells: mostly degenerate
    sets = [(rng.random((4000, 3)), None, 1), (lat, ((ijk[..., 0] * m + ijk[..., 1]) * m + ijk[..., 2]), 8),
            (rng.random((9000, 3)), None, 7)]
    pts_l, op_l, idx_l, want = [], [], [], []
    for pts, idx, copies in sets:
        op = rng.integers(0, 3, size=n).astype(np.uint8)
        if idx is None:
            idx = rng.integers(0, len(pts), size=(n, 5))
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        one, _ = predicates_dev(pts, op, idx)
        s = rng.choice(n, 1 << 12, replace=False)
        assert np.array_equal(one[s], O.pred_batch(pts, op[s], idx[s], "oracle"))
        tail = (1 << 20) if copies == 7 else 0
        pts_l.append(pts)
        op_l.append(np.concatenate([op] * copies + [op[:tail]]))
        idx_l.append(np.concatenate([idx] * copies + [idx[:tail]]))
        want.append(np.concatenate([one] * copies + [one[:tail]]))
    want = np.concatenate(want)
    assert len(want) == (1 << 28) + (1 << 20)
    say("plan of two launch sequences", len(want), "queries")
    got = _plan_signs(pts_l, op_l, idx_l, graph=True, replays=2)
    for r in got:
        assert np.array_equal(r, want)


@pytest.mark.parametrize("kind", ["lattice32", "lattice34", "random9000"])
def test_more_queries_than_one_launch_sequence(kind):
    """2^28 + 2^18 queries in one call, above the 2^28 of one launch sequence
    (so two sequences with their own queues and counters), on the 32^3
    lattice (packed LDS kernel), the 34^3 lattice (above the packed LDS
    kernel's capacity: the global kernel's packed variant) and on 9,000
    random points (the global kernel beside it): equal to the same queries in
    calls of 2^24."""
    if kind == "lattice32":
        pts, op, idx = W.c2(1 << 24, seed=2)
    elif kind == "lattice34":
        m = 34
        g = np.arange(m, dtype=np.float64)
        pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
        rng = np.random.default_rng(34)
        n = 1 << 24
        op = rng.integers(0, 3, size=n).astype(np.uint8)
        base = rng.integers(0, m - 2, size=(n, 3))
        ijk = base[:, None, :] + rng.integers(0, 3, size=(n, 5, 3))   # local cells: mostly degenerate
        idx = ((ijk[..., 0] * m + ijk[..., 1]) * m + ijk[..., 2]).astype(np.int32)
    else:
        rng = np.random.default_rng(33)
        pts = rng.random((9000, 3))
        op = rng.integers(0, 3, size=1 << 24).astype(np.uint8)
        idx = rng.integers(0, len(pts), size=(1 << 24, 5)).astype(np.int32)
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    one, _ = predicates_dev(pts, op, idx)
    tail = 1 << 18
    say("two launch sequences", kind, 16 * len(op) + tail, "queries")
    got, _ = predicates_dev(pts, np.concatenate([op] * 16 + [op[:tail]]),
                            np.concatenate([idx] * 16 + [idx[:tail]]))
    assert np.array_equal(got, np.concatenate([one] * 16 + [one[:tail]]))
    s = np.random.default_rng(1).choice(len(op), 1 << 14, replace=False)
    assert np.array_equal(one[s], O.pred_batch(pts, op[s], idx[s], "oracle"))
