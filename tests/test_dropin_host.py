"""The reference's own predicate test programs, linked against
libdel3d_amd.so in place of src/predicates.cpp and src/predicates2.cpp
(oracle/Makefile `dropin`). Their calls are scalar, which the library
evaluates on the host with the kernels' per-query code, so they run here
without a GPU: every check of tests/test_predicates.cpp and
tests/test_predicates2.cpp must pass, as against the reference."""
import subprocess
from pathlib import Path

import pytest

REF_BIN = Path(__file__).resolve().parents[1] / "oracle" / "_ref"


@pytest.mark.parametrize("name,expect", [
    ("predicates", "ALL PREDICATES MATCH THE REFERENCE"),
    ("predicates2", "ALL COPLANAR PREDICATES MATCH THE REFERENCE"),
])
def test_reference_predicate_tests_on_host_path(name, expect):
    exe = REF_BIN / f"dropin_test_{name}"
    if not exe.exists():
        pytest.skip("drop-in binaries not built (need the reference sources at build time)")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert expect in r.stdout
