"""Runs the bridged Delaunay driver programs of oracle/_ref (built by
oracle/Makefile's dropin target where the reference sources exist) and checks
their JSON lines."""
import json
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF_BIN = ROOT / "oracle" / "_ref"
HOST_ONLY = str(2**63 - 1)   # gpu_min above any wave: every test on the host

PINS = json.loads((ROOT / "tests" / "golden" / "reference_pins.json").read_text())["delaunay_cell_hash"]


def pinned_hash(case):
    """The pinned hash of a case name printed by bridge_delaunay, or None."""
    if case.startswith("random_cube_seed4 "):
        return int(PINS["random_cube_seed4"][case.split()[1]], 16)
    return int(PINS[case], 16) if case in PINS else None


def run(exe_name, *args, timeout=600):
    exe = REF_BIN / exe_name
    if not exe.exists():
        pytest.skip(f"{exe_name} not built (needs the reference sources at build time)")
    r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    return [json.loads(l) for l in r.stdout.splitlines() if l.startswith("{")]


def check_cases(rows):
    """Every case equal to the driver's own build; the pinned ones at their
    pinned hash (tests/test_delaunay.cpp:209-243)."""
    pinned = 0
    for row in rows:
        assert row["same_as_driver"] is True, row
        want = pinned_hash(row["case"])
        if want is not None:
            assert int(row["hash"], 16) == want, row
            pinned += 1
    assert pinned == 11
