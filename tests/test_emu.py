"""The library's kernels and host code (C ABI, predicate pipelines 3D and
coplanar, spatial sort), run on the CPU by the emulator in tests/emu under AddressSanitizer,
against the oracle.

EMU_DEFER=1 models the runtime's asynchrony at its least favourable:
kernels run only when the host synchronizes, while copies to or from
pageable host memory and stream-ordered frees take effect at once. A host
sequence that needs a wait it does not have (reading results before the
kernels producing them ran, or refilling a buffer a queued kernel still
reads) then fails here instead of on a GPU. EMU_BLOCK_ORDER reverses or
shuffles the order of blocks, which exposes dependencies between blocks of
one launch."""
import os
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
EXE = Path("/tmp") / f"del3d_emu_sort_{os.getuid()}"


@pytest.fixture(scope="module")
def emu_exe():
    csrc = HERE.parent / "del3d_amd" / "csrc"
    subprocess.run(["python", str(HERE / "emu" / "build_emu.py"), str(EXE),
                    str(csrc / "capi.hip"), str(csrc / "predicates.hip"), str(csrc / "predicates2.hip"),
                    str(csrc / "circumcenter.hip"), str(csrc / "batch.hip"), str(csrc / "server.hip"), str(csrc / "spatial_sort.hip"),
                    str(csrc / "host_scalar.cpp")],
                   check=True)
    return EXE


@pytest.mark.parametrize("defer,order,sizes,sort_only", [("1", "2", ["2", "65", "1100"], ""),
                                                         ("0", "1", ["2", "65", "1100"], "1")])
def test_emulated_sort_matches_oracle(emu_exe, defer, order, sizes, sort_only):
    """Threads per lane under AddressSanitizer: the predicate pipelines and
    small sorts (subtree and LDS passes); the second run sorts only."""
    env = dict(os.environ, EMU_DEFER=defer, EMU_BLOCK_ORDER=order, ASAN_OPTIONS="detect_leaks=0",
               EMU_SKIP_SETS="1")
    if sort_only:
        env["EMU_SORT_ONLY"] = "1"
    r = subprocess.run([str(emu_exe)] + sizes, capture_output=True, text=True, timeout=900, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]


@pytest.fixture(scope="module")
def fiber_exe():
    exe = Path("/tmp") / f"del3d_emu_fibers_{os.getuid()}"
    csrc = HERE.parent / "del3d_amd" / "csrc"
    subprocess.run(["python", str(HERE / "emu" / "build_emu.py"), str(exe),
                    str(csrc / "capi.hip"), str(csrc / "predicates.hip"), str(csrc / "predicates2.hip"),
                    str(csrc / "circumcenter.hip"), str(csrc / "batch.hip"), str(csrc / "server.hip"),
                    str(csrc / "spatial_sort.hip"), str(csrc / "host_scalar.cpp")],
                   check=True, env=dict(os.environ, EMU_SANITIZER="none", EMU_OPT="-O2"))
    return exe


def test_emulated_sort_device_rounds_fibers(fiber_exe):
    """The sort's device-wide rounds (selects above 4,096 elements), the
    LDS finish and subtree passes, at sizes up to 20,000 and on tie-heavy
    sets, a permuted subset order, a bad index and eight sets in one call,
    against the oracle: the kernels run as fibers (one OS thread, switched
    at every barrier and wavefront operation), which keeps these sizes to
    seconds; AddressSanitizer is off in this mode (its stack poisoning does
    not follow switched stacks), the guards in the kernels are not."""
    env = dict(os.environ, EMU_FIBERS="1", EMU_SORT_ONLY="1", EMU_SORT_CASES="1")
    r = subprocess.run([str(fiber_exe), "2", "9", "65", "1100", "4681", "20000"], capture_output=True,
                       text=True, timeout=900, env=env)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert r.stdout.count(" ok") >= 12 and "MISMATCH" not in r.stdout


@pytest.fixture(scope="module")
def tsan_exe():
    exe = Path("/tmp") / f"del3d_emu_tsan_{os.getuid()}"
    csrc = HERE.parent / "del3d_amd" / "csrc"
    subprocess.run(["python", str(HERE / "emu" / "build_emu.py"), str(exe),
                    str(csrc / "capi.hip"), str(csrc / "predicates.hip"), str(csrc / "predicates2.hip"),
                    str(csrc / "circumcenter.hip"), str(csrc / "batch.hip"), str(csrc / "server.hip"), str(csrc / "spatial_sort.hip"),
                    str(csrc / "host_scalar.cpp")],
                   check=True,
                   env=dict(os.environ, EMU_SANITIZER="thread"))
    return exe


def test_emulated_predicates_under_thread_sanitizer(tsan_exe):
    """The 3D predicate pipeline (LDS, packed and global kernels over several
    tiles per workgroup, plans; a third of the other runs' queries) under
    ThreadSanitizer: a missing barrier
    between a tile's LDS writes and another thread's reads is a data race
    between the emulated threads."""
    env = dict(os.environ, EMU_PRED_ONLY="1", EMU_QUERY_DIV="3", TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([str(tsan_exe)], capture_output=True, text=True, timeout=900, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "pred lattice20 (packed) ok" in r.stdout and "MISMATCH" not in r.stdout


def test_emulated_server_under_thread_sanitizer(tsan_exe):
    """The serving loop (scheduler thread, one worker thread per slot, two
    client threads, and the predicate pipeline it drives) under
    ThreadSanitizer."""
    exe = tsan_exe
    env = dict(os.environ, EMU_SERVER_ONLY="1", TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=900, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "server jobs" in r.stdout and "MISMATCH" not in r.stdout
