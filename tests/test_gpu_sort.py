"""GPU spatial-sort parity: the permutation equals the oracle's, and the
reference tests' 17 pinned hashes hold."""
import json
from pathlib import Path

import numpy as np
import pytest

import del3d_amd as D
import oracle_lib as O
from del3d_amd import workloads as W
from gpu_util import require_gpu, say

pytestmark = pytest.mark.gpu
PINS = json.loads((Path(__file__).parent / "golden" / "reference_pins.json").read_text())


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


def _pinned():
    pins = PINS["spatial_sort_fnv1a"]
    cases = [(f"random n={n}", lambda n=int(n): O.random_cube(n, 7), h)
             for n, h in pins["random_cube_seed7"].items()]
    cases += [("lattice 12x12x6", lambda: O.lattice(12, 12, 6), pins["lattice_12_12_6"]),
              ("box shell 8", lambda: O.box_shell(8), pins["box_shell_8"]),
              ("300 identical", lambda: np.tile([1.0, 2.0, 3.0], (300, 1)), pins["identical_300_at_1_2_3"]),
              ("cylinder", lambda: O.cylinder(2000), pins["cylinder_2000_q32"]),
              ("sphere", lambda: O.sphere(2000), pins["sphere_2000_q32"]),
              ("torus", lambda: O.torus(2000), pins["torus_2000_q32"])]
    return cases


@pytest.mark.parametrize("name,make,want", _pinned(), ids=[c[0] for c in _pinned()])
def test_pinned_hashes(name, make, want):
    pts = make()
    say("sort", name)
    order = D.spatial_sort(pts)
    assert O.hash_order(order) == int(want, 16)


def _device_sort(pts, order=None):
    xyz = D.DeviceArray.from_host(np.ascontiguousarray(pts, np.float64).reshape(-1, 3))
    o = np.arange(len(pts), dtype=np.int32) if order is None else np.asarray(order, np.int32)
    d = D.DeviceArray.from_host(o)
    D.spatial_sort_device(xyz, d)
    D.synchronize()
    return d.to_host()


@pytest.mark.parametrize("n", [2, 3, 9, 64, 65, 1023, 1024, 1025, 2049, 4681, 5000, 20000, 100000,
                               1000000])
def test_random_sizes_vs_oracle(n):
    pts = O.random_cube(n, 11)
    say("sort random", n)
    assert np.array_equal(_device_sort(pts), O.spatial_sort(pts))


@pytest.mark.parametrize("name", ["lattice32", "sphere_q1024", "identical4000", "grid_ties"])
def test_tie_heavy_vs_oracle(name):
    rng = np.random.default_rng(5)
    if name == "lattice32":
        pts = W.lattice(32, 32, 32)
    elif name == "sphere_q1024":
        v = rng.normal(size=(60000, 3))
        pts = np.round(v / np.linalg.norm(v, axis=1, keepdims=True) * 1024) / 1024
    elif name == "identical4000":
        pts = np.tile([0.25, 0.5, 0.75], (4000, 1))
    else:
        pts = rng.integers(0, 6, size=(30000, 3)).astype(np.float64)
    say("sort ties", name, len(pts))
    assert np.array_equal(_device_sort(pts), O.spatial_sort(pts))


def test_order_is_a_subset_and_permuted():
    pts = O.random_cube(5000, 3)
    rng = np.random.default_rng(2)
    order = rng.permutation(5000)[:3000].astype(np.int32)
    say("sort subset order")
    assert np.array_equal(D.spatial_sort(pts, order), O.spatial_sort(pts, order))


def test_sets_match_individual_sorts():
    rng = np.random.default_rng(8)
    sizes = [0, 1, 2, 7, 9, 64, 500, 1024, 1500, 7000, 30000]
    rng.shuffle(sizes)
    sets = [O.random_cube(n, 20 + i) if n else np.zeros((0, 3)) for i, n in enumerate(sizes)]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    pts = np.vstack(sets)
    order = np.concatenate([np.arange(n, dtype=np.int32) for n in sizes])
    xyz = D.DeviceArray.from_host(pts)
    d = D.DeviceArray.from_host(order)
    say("sort sets", sizes)
    D.spatial_sort_sets(xyz, off, d)
    D.synchronize()
    got = d.to_host()
    for i, n in enumerate(sizes):
        assert np.array_equal(got[off[i]:off[i + 1]], O.spatial_sort(sets[i])), n


def test_many_mid_size_sets_vs_oracle():
    """300 sets of 4,097 to 40,000 points in one call: every phase has at
    least one select per CU, none above 262,144 elements, so each select runs
    all its rounds in one workgroup (the workgroup tier); sets of tie-heavy
    grids among them. Each set against the oracle."""
    rng = np.random.default_rng(81)
    sizes = rng.integers(4097, 40001, size=300).tolist()
    sets = []
    for i, n in enumerate(sizes):
        p = O.random_cube(n, 100 + i)
        if i % 3 == 0:
            p = np.floor(p * 16) / 16  # ties on every axis
        sets.append(p)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    order = np.concatenate([np.arange(n, dtype=np.int32) for n in sizes])
    xyz = D.DeviceArray.from_host(np.vstack(sets))
    d = D.DeviceArray.from_host(order)
    say("sort 300 mid-size sets", sum(sizes), "points")
    D.spatial_sort_sets(xyz, off, d)
    D.synchronize()
    got = d.to_host()
    for i, n in enumerate(sizes):
        assert np.array_equal(got[off[i]:off[i + 1]], O.spatial_sort(sets[i])), (i, n)


def test_ten_million_points_vs_oracle():
    pts = W.c3_points(10_000_000, seed=3)
    say("sort 1e7")
    got = _device_sort(pts)
    want = O.spatial_sort(pts)
    assert np.array_equal(got, want)


def test_bad_order_entry_is_an_error():
    pts = O.random_cube(6000, 4)
    order = np.arange(6000, dtype=np.int32)
    order[123] = 6000
    say("sort bad index")
    with pytest.raises(RuntimeError, match="outside"):
        D.spatial_sort(pts, order)
    # the library still sorts afterwards
    assert np.array_equal(D.spatial_sort(pts), O.spatial_sort(pts))
