"""The reference's own test programs, and its unchanged Delaunay driver,
linked against libdel3d_amd.so in place of the reference's predicates and
spatial sort (oracle/Makefile `dropin`): they must pass with every pinned
hash, exactly as against the reference."""
import subprocess
from pathlib import Path

import pytest

from gpu_util import require_gpu, say

pytestmark = pytest.mark.gpu
REF_BIN = Path(__file__).resolve().parents[1] / "oracle" / "_ref"


@pytest.mark.parametrize("name", [
    "predicates",
    "predicates2",
    "spatial_sort",
    "delaunay",
])
def test_reference_test_program_on_gpu_library(name):
    exe = REF_BIN / f"dropin_test_{name}"
    if not exe.exists():
        pytest.skip("drop-in binaries not built (need the reference sources at build time)")
    require_gpu()
    say("running", exe.name)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    say(r.stdout[-1500:])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
