"""GPU checks of the serving loop: predicate and sort jobs arriving as a
Poisson stream from two client threads, over resident point sets whose sizes
together exceed the HBM budget (so sets are evicted and uploaded again, and
batches wait for memory), each job's signs or sort order equal to the same
work run alone (tolerance 0), with pinned and pageable host arrays."""
import threading

import numpy as np
import pytest

import del3d_amd as D
import oracle_lib as O
from del3d_amd import workloads as W
from gpu_util import predicates_dev, require_gpu, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


def _sets(rng):
    sizes = [700, 3000, 5000, 20000, 90000, 200000, 32768, 12000]
    sets = []
    for i, n in enumerate(sizes):
        if i == 6:
            sets.append(W.lattice(32, 32, 32))
        else:
            sets.append(W.set_points(i % 3, n, rng))
    return sets


def _jobs(rng, sets, n_jobs):
    jobs = []
    for _ in range(n_jobs):
        s = int(rng.integers(0, len(sets)))
        n = len(sets[s])
        nq = int(1000 * 300 ** rng.random())
        op = rng.integers(0, 3, size=nq).astype(np.uint8)
        idx = rng.integers(0, n, size=(nq, 5)).astype(np.int32)
        idx[: nq // 8, 4] = idx[: nq // 8, 0]
        jobs.append((s, op, idx))
    return jobs


def test_served_jobs_equal_alone():
    rng = np.random.default_rng(21)
    sets = _sets(rng)
    jobs = _jobs(rng, sets, 120)
    alone = [predicates_dev(sets[s], op, idx)[0] for s, op, idx in jobs]
    # 16 MB: the sets (about 9 MB) and the buffers of a few in-flight batches
    # (up to 9 MB each) do not fit at once
    srv = D.Server(16 << 20, slots=3, max_batch_queries=1 << 19)
    ids = [srv.put_points(p) for p in sets]
    results = [None] * len(jobs)
    lat = [0.0] * len(jobs)

    def client(first):
        r = np.random.default_rng(100 + first)
        mine = list(range(first, len(jobs), 2))
        sub = {}
        for k in mine:
            s, op, idx = jobs[k]
            sub[k] = srv.submit(ids[s], op, idx)
            threading.Event().wait(r.exponential(0.0005))  # Poisson arrivals
        for k in mine:
            results[k], lat[k] = srv.wait(sub[k])

    say("serving", len(jobs), "jobs")
    th = [threading.Thread(target=client, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    info = srv.info()
    say("server", info, "p50 ms", np.percentile(lat, 50), "p99 ms", np.percentile(lat, 99))
    srv.close()
    for k, (got, want) in enumerate(zip(results, alone)):
        assert np.array_equal(got, want), k
    assert info["evictions"] > 0 and info["uploads"] > len(sets)
    assert info["jobs_done"] == len(jobs)
    # one job of each set against the oracle
    for s in range(len(sets)):
        k = next(k for k, j in enumerate(jobs) if j[0] == s)
        _, op, idx = jobs[k]
        m = min(len(op), 50000)
        assert np.array_equal(results[k][:m], O.pred_batch(sets[s], op[:m], idx[:m], "oracle"))


def test_oversized_set_is_refused_and_others_run():
    rng = np.random.default_rng(5)
    srv = D.Server(8 << 20, slots=2, max_batch_queries=1 << 16)
    small = rng.random((2000, 3))
    sid = srv.put_points(small)
    with pytest.raises(D.Del3dError):
        srv.put_points(rng.random((400000, 3)))  # 9.6 MB > 8 MB budget
    op = rng.integers(0, 3, size=50000).astype(np.uint8)
    idx = rng.integers(0, 2000, size=(50000, 5)).astype(np.int32)
    got, _ = srv.wait(srv.submit(sid, op, idx))
    srv.close()
    assert np.array_equal(got, predicates_dev(small, op, idx)[0])


def test_sort_and_predicate_jobs_mixed_equal_alone():
    rng = np.random.default_rng(33)
    sizes = [5000, 60000, 300000, 12000]
    sets = [W.set_points(i % 3, n, rng) for i, n in enumerate(sizes)]
    sorted_alone = [O.spatial_sort(p) for p in sets]
    jobs = []
    for k in range(40):
        s = int(rng.integers(0, len(sets)))
        if k % 3 == 0:
            jobs.append(("sort", s, None, None))
        else:
            nq = int(2000 * 100 ** rng.random())
            op = rng.integers(0, 3, size=nq).astype(np.uint8)
            idx = rng.integers(0, len(sets[s]), size=(nq, 5)).astype(np.int32)
            jobs.append(("pred", s, op, idx))
    # 160 MB: the sets (9 MB), a 300,000-point sort's scratch (93 MB as
    # budgeted) and the batch buffers (up to 31 MB per slot) do not all fit
    # at once, so work waits for memory
    srv = D.Server(160 << 20, slots=3, max_batch_queries=1 << 20)
    # every other set from a pinned host array, the others pageable
    ids = [srv.put_points(D._hip.pinned_copy(p) if i % 2 else p) for i, p in enumerate(sets)]
    sub = []
    for k, (kind, s, op, idx) in enumerate(jobs):
        if kind == "sort":
            out = D._hip.pinned_empty(len(sets[s]), np.int32) if k % 2 else None
            sub.append(srv.submit_sort(ids[s], out))
        else:
            out = D._hip.pinned_empty(len(op), np.int8) if k % 2 else None
            sub.append(srv.submit(ids[s], op, idx, out))
    say("serving", len(jobs), "mixed jobs")
    res = [srv.wait(j) for j in sub]
    info = srv.info()
    srv.close()
    say("server", info)
    for k, (kind, s, op, idx) in enumerate(jobs):
        got = res[k][0]
        if kind == "sort":
            assert np.array_equal(got, sorted_alone[s]), k
        else:
            assert np.array_equal(got, predicates_dev(sets[s], op, idx)[0]), k
