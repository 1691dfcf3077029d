"""The predicate trace format and traces of the reference's own Delaunay
driver (oracle/trace_recorder, built from the reference's unchanged sources):
every recorded result recomputed by the oracle. The lattice traces are the
degenerate builds, where the driver's dimension bootstrap and exactly
coplanar hull facets call the coplanar predicates."""
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
from del3d_amd import trace as T

TRACES = {
    "rand48_2000": ["2000", "1"],
    "lattice_8x8x8": ["lattice", "8", "8", "8"],
    "lattice_9x7x1": ["lattice", "9", "7", "1"],    # planar: coplanar predicates only
    "lattice_15x1x1": ["lattice", "15", "1", "1"],  # collinear
}


def record(name, td):
    out = Path(td) / f"{name}.trace"
    subprocess.run([str(O.TRACE_RECORDER)] + TRACES[name] + [str(out)], check=True, timeout=300,
                   capture_output=True)
    return T.read(out)


def oracle_replay(pts, op, idx):
    got = np.zeros(len(op), np.int8)
    m3 = op < 16
    if m3.any():
        i3 = idx[m3].copy()
        i3[op[m3] == 0, 4] = i3[op[m3] == 0, 3]
        got[m3] = O.pred_batch(pts, op[m3], i3, "oracle")
    if (~m3).any():
        got[~m3] = O.pred2_batch(pts, op[~m3] - 16, np.maximum(idx[~m3], 0), "oracle")
    return got


@pytest.mark.skipif(not O.TRACE_RECORDER.exists(), reason="trace recorder not built")
@pytest.mark.parametrize("name", list(TRACES))
def test_recorded_driver_calls_match_oracle(name):
    with tempfile.TemporaryDirectory() as td:
        pts, op, res, idx = record(name, td)
    assert len(op) > 0
    if name.startswith("lattice"):
        assert (op >= 16).any()  # the coplanar predicates were called
    got = oracle_replay(pts, op, idx)
    bad = np.nonzero(got != res)[0]
    assert len(bad) == 0, (name, op[bad[:5]], idx[bad[:5]], got[bad[:5]], res[bad[:5]])


def test_format_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    pts = rng.random((50, 3))
    op = rng.choice(np.array([0, 1, 2, 16, 21, 24], np.uint8), size=300)
    res = rng.integers(-1, 2, size=300).astype(np.int8)
    idx = rng.integers(-1, 50, size=(300, 5)).astype(np.int32)
    T.write(tmp_path / "t.trace", pts, op, res, idx)
    p2, o2, r2, i2 = T.read(tmp_path / "t.trace")
    assert np.array_equal(p2, pts) and np.array_equal(o2, op)
    assert np.array_equal(r2, res) and np.array_equal(i2, idx)
    raw = (tmp_path / "t.trace").read_bytes()
    (tmp_path / "bad.trace").write_bytes(raw[:-3])
    with pytest.raises(ValueError):
        T.read(tmp_path / "bad.trace")
