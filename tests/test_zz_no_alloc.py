"""The allocation-free predicate entry point (del3d_predicates_noalloc),
called through the C ABI on device buffers: its signs equal those of
del3d_predicates on the same buffers, call after call on the shared fixed
scratch, the second call on a smaller input than the first; every route of
the pipeline (the global kernel over two launch sequences, the packed LDS
kernel, the LDS kernel, the global kernel's packed variant) is exercised."""
import numpy as np
import pytest

import del3d_amd as D
from del3d_amd import workloads as W
from gpu_util import require_gpu, say

pytestmark = pytest.mark.gpu


def _device(a, dtype):
    return D.DeviceArray.from_host(np.ascontiguousarray(a, dtype=dtype))


def _both(pts, op, idx):
    """Signs from del3d_predicates and from del3d_predicates_noalloc on the
    same device buffers, each into its own output (filled with 9 first)."""
    lib = D.lib()
    pts = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
    xyz = _device(pts, np.float64)
    dop, didx = _device(op, np.uint8), _device(np.asarray(idx).reshape(-1, 5), np.int32)
    n_q = len(op)
    outs = []
    for fn in (lib.del3d_predicates, lib.del3d_predicates_noalloc):
        out = D.DeviceArray.from_host(np.full(n_q, 9, np.int8))
        rc = fn(xyz.ptr, len(pts), dop.ptr, didx.ptr, n_q, out.ptr, None)
        assert rc == 0, lib.del3d_last_error()
        D.synchronize()
        outs.append(out.to_host())
    return outs


def _random_set(rng, n_pts, n_q):
    pts = rng.random((n_pts, 3))
    pts[: n_pts // 4] = np.round(pts[: n_pts // 4] * 8) / 8  # ties: exact tiers and perturbation
    op = rng.integers(0, 3, size=n_q).astype(np.uint8)
    idx = rng.integers(0, n_pts, size=(n_q, 5)).astype(np.int32)
    idx[::7, 4] = idx[::7, 0]  # aliased slots
    return pts, op, idx


def test_noalloc_equals_predicates_on_consecutive_calls():
    require_gpu()
    rng = np.random.default_rng(21)
    cases = [
        # first call: 300,000 random doubles (global kernel), 2^21 + 777
        # queries, i.e. three launch sequences of the fixed 2^20-query queues
        ("global, three launch sequences", *_random_set(rng, 300_000, (1 << 21) + 777)),
        # second call, smaller: the 32^3 lattice (packed LDS kernel)
        ("packed lattice", *W.c2(1 << 18, seed=5)),
        # third and fourth: a small set (LDS kernel), and a 40^3 lattice
        # (the global kernel's packed variant) at the point limit's scale
        ("LDS", *_random_set(rng, 3000, 50_000)),
        ("global packed", W.lattice(40, 40, 40), *_random_set(rng, 64_000, 200_000)[1:]),
    ]
    for name, pts, op, idx in cases:
        say("noalloc", name, len(pts), len(op))
        want, got = _both(pts, op, idx)
        assert not np.any(want == 9), name
        assert np.array_equal(got, want), f"{name}: {int((got != want).sum())} signs differ"


def test_noalloc_refuses_more_points_than_its_scratch():
    require_gpu()
    lib = D.lib()
    n = (1 << 20) + 1
    xyz = D.DeviceArray((n, 3), np.float64)
    op = D.DeviceArray((4,), np.uint8)
    idx = D.DeviceArray((4, 5), np.int32)
    out = D.DeviceArray((4,), np.int8)
    assert lib.del3d_predicates_noalloc(xyz.ptr, n, op.ptr, idx.ptr, 4, out.ptr, None) == 1
    assert b"at most" in lib.del3d_last_error()
