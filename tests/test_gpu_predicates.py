"""GPU predicate parity: every sign equals the oracle's (bit-exact, int8),
on the reference's own test workloads, the measured configurations and the
hard cases (wide exponent spans, denormals, aliases, ties)."""
import json
import struct
import subprocess
import tempfile
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
    pts = O.lattice(3, 3, 3)
    for n in (0, 1, 63, 255, 256, 257, 1000):
        rng = np.random.default_rng(n)
        idx = rng.integers(0, len(pts), size=(n, 5)).astype(np.int32)
        op = rng.integers(0, 3, size=n).astype(np.uint8)
        say("tail", n)
        got, _ = predicates_dev(pts, op, idx)
        assert np.array_equal(got, O.pred_batch(pts, op, idx, "oracle"))


def test_host_memory_form():
    pts, op, idx = W.c2(5000, seed=9)
    say("host form")
    assert np.array_equal(D.predicates_np(pts, op, idx), O.pred_batch(pts, op, idx, "oracle"))


def _read_trace(path):
    raw = Path(path).read_bytes()
    assert raw[:4] == b"D3TR"
    n, nrec = struct.unpack_from("<qq", raw, 8)
    off = 24
    pts = np.frombuffer(raw, dtype=np.float64, count=3 * n, offset=off).reshape(n, 3)
    off += 24 * n
    op = np.frombuffer(raw, dtype=np.uint8, count=nrec, offset=off)
    off += nrec
    sign = np.frombuffer(raw, dtype=np.int8, count=nrec, offset=off)
    off += nrec
    idx = np.frombuffer(raw, dtype=np.int32, count=5 * nrec, offset=off).reshape(nrec, 5)
    return pts, op, sign, idx


@pytest.mark.skipif(not O.TRACE_RECORDER.exists(), reason="trace recorder not built (needs the reference sources at build time)")
def test_c1_trace_replay():
    """Every orient3d / in_sphere call of the reference's Delaunay build on
    10,000 rand48 points, replayed on the GPU: identical signs."""
    with tempfile.TemporaryDirectory() as td:
        out = Path(td) / "c1.trace"
        say("recording C1 trace")
        subprocess.run([str(O.TRACE_RECORDER), "10000", "1", str(out)], check=True, timeout=300)
        pts, op, sign, idx = _read_trace(out)
    assert np.array_equal(pts, W.c1_points(10000, 1))
    idx = idx.copy()
    idx[op == 0, 4] = idx[op == 0, 3]
    say("replaying", len(op))
    got, st = predicates_dev(pts, op, idx, stats=True)
    assert np.array_equal(got, sign), int((got != sign).sum())
    say("trace stats", st)
