"""Pins the CPU oracle against what it does not derive from.

* glibc's srand48/lrand48 (same LCG and output) for the rand48 draw stream;
* the C++ standard's required mt19937 output for the point-set generator;
* the reference tests' 17 pinned spatial-sort hashes;
* the known signs of tests/test_predicates.cpp:64-76;
* the big-integer oracle (a different exact method) on degenerate inputs;
* closed forms: orient3d of a unit tetrahedron under permutations, the
  lifted determinant's brute-force sign on tiny integer inputs.
"""
import ctypes
import json
from itertools import permutations
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O

PINS = json.loads((Path(__file__).parent / "golden" / "reference_pins.json").read_text())


def test_rand48_matches_glibc_lrand48():
    libc = ctypes.CDLL("libc.so.6")
    libc.lrand48.restype = ctypes.c_long
    libc.srand48.argtypes = [ctypes.c_long]
    for seed in (1, 2, 3, 12345):
        libc.srand48(seed)
        want = np.array([libc.lrand48() for _ in range(2000)], dtype=np.uint32)
        got = O.rand48_draws(seed, 2000)
        assert np.array_equal(got, want)


def test_mt19937_standard_value():
    w = O.mt19937_words(PINS["mt19937"]["seed"], 10000)
    assert int(w[-1]) == PINS["mt19937"]["word_10000"]


def _pinned_inputs():
    pins = PINS["spatial_sort_fnv1a"]
    cases = []
    for n, h in pins["random_cube_seed7"].items():
        cases.append((f"random n={n}", lambda n=int(n): O.random_cube(n, 7), h))
    cases.append(("lattice 12x12x6", lambda: O.lattice(12, 12, 6), pins["lattice_12_12_6"]))
    cases.append(("box shell 8", lambda: O.box_shell(8), pins["box_shell_8"]))
    cases.append(("300 identical", lambda: np.tile([1.0, 2.0, 3.0], (300, 1)),
                  pins["identical_300_at_1_2_3"]))
    cases.append(("cylinder", lambda: O.cylinder(2000), pins["cylinder_2000_q32"]))
    cases.append(("sphere", lambda: O.sphere(2000), pins["sphere_2000_q32"]))
    cases.append(("torus", lambda: O.torus(2000), pins["torus_2000_q32"]))
    return cases


@pytest.mark.parametrize("name,make,want", _pinned_inputs(), ids=[c[0] for c in _pinned_inputs()])
def test_oracle_sort_reproduces_pinned_hashes(name, make, want):
    pts = make()
    order = O.spatial_sort(pts)
    assert sorted(order.tolist()) == list(range(len(pts)))
    assert O.hash_order(order) == int(want, 16), name


def test_known_signs():
    k = PINS["known_signs"]
    o, x, y, z = (np.array(k[n], dtype=np.float64) for n in "oxyz")
    inside = np.array(k["inside"], dtype=np.float64)
    outside = np.array(k["outside"], dtype=np.float64)
    P = ctypes.POINTER(ctypes.c_double)

    def p(a):
        return a.ctypes.data_as(P)

    lib = O.oracle()
    for fn in (lib.orc_orient3d, lib.orc_int_orient3d):
        assert fn(p(o), p(x), p(y), p(z)) == k["orient3d(o,x,y,z)"]
        assert fn(p(o), p(y), p(x), p(z)) == k["orient3d(o,y,x,z)"]
    for fn in (lib.orc_in_sphere_unperturbed, lib.orc_int_in_sphere_raw):
        assert fn(p(o), p(x), p(y), p(z), p(inside)) == k["in_sphere_raw(o,x,y,z,inside)"]
        assert fn(p(o), p(x), p(y), p(z), p(outside)) == k["in_sphere_raw(o,x,y,z,outside)"]
    for fn in (lib.orc_in_sphere, lib.orc_int_in_sphere):
        assert fn(p(o), p(x), p(y), p(z), p(inside)) == k["in_sphere(o,x,y,z,inside)"]
        assert fn(p(o), p(x), p(y), p(z), p(outside)) == k["in_sphere(o,x,y,z,outside)"]


def test_orient3d_is_alternating_on_unit_tetrahedron():
    """Closed form: the determinant changes sign with the parity of a permutation."""
    pts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    for perm in permutations(range(4)):
        inv = sum(1 for i in range(4) for j in range(i + 1, 4) if perm[i] > perm[j])
        want = 1 if inv % 2 == 0 else -1
        idx = np.array([list(perm) + [0]], dtype=np.int32)
        for which in ("oracle", "int"):
            assert O.pred_batch(pts, [0], idx, which)[0] == want


def _det(m):
    if len(m) == 1:
        return m[0][0]
    return sum((-1) ** j * m[0][j] * _det([r[:j] + r[j + 1:] for r in m[1:]])
               for j in range(len(m)))


def _sgn(v):
    return (v > 0) - (v < 0)


def _brute_orient(a, b, c, d):
    """sign |b-a, c-a, d-a| by cofactor expansion in Python integers."""
    return _sgn(_det([[b[i] - a[i] for i in range(3)], [c[i] - a[i] for i in range(3)],
                      [d[i] - a[i] for i in range(3)]]))


def _brute_lifted5(a, b, c, d, e):
    """sign of the 5x5 determinant with rows (x, y, z, x^2+y^2+z^2, 1)."""
    return _sgn(_det([[p[0], p[1], p[2], p[0] ** 2 + p[1] ** 2 + p[2] ** 2, 1]
                      for p in (a, b, c, d, e)]))


def test_insphere_brute_force_small_integers():
    """The textbook 5x5 lifted determinant, with its global sign fixed once by
    the known inside answer, must agree with the raw in_sphere everywhere."""
    conv = 1 * _brute_lifted5((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (0.25, 0.25, 0.25))
    assert conv in (1, -1)
    rng = np.random.default_rng(11)
    pts = rng.integers(-3, 4, size=(60, 3)).astype(np.float64)
    idx = rng.integers(0, 60, size=(2000, 5)).astype(np.int32)
    ip = [tuple(int(v) for v in pts[i]) for i in range(60)]
    for which in ("oracle", "int"):
        raw = O.pred_batch(pts, np.full(2000, 2, np.uint8), idx, which)
        orient = O.pred_batch(pts, np.zeros(2000, np.uint8), idx, which)
        for q in range(2000):
            a, b, c, d, e = (ip[i] for i in idx[q])
            assert orient[q] == _brute_orient(a, b, c, d)
            assert raw[q] == conv * _brute_lifted5(a, b, c, d, e)


def test_restatement_matches_int_oracle_on_reference_stage2_sets():
    """The reference's stage-2 workload (tests/test_predicates.cpp:107-129):
    Rng(12345) index draws over quantised surfaces and box_shell(6)."""
    sets = {"cylinder": O.cylinder(400, 8), "sphere": O.sphere(400, 8),
            "torus": O.torus(400, 8), "box_shell": O.box_shell(6)}
    for name, pts in sets.items():
        draws = O.rng_below(12345, len(pts), 20000 * 5).reshape(20000, 5).astype(np.int32)
        for op in (0, 1, 2):
            ops = np.full(20000, op, np.uint8)
            a = O.pred_batch(pts, ops, draws, "oracle")
            b = O.pred_batch(pts, ops, draws, "int")
            assert np.array_equal(a, b), (name, op)
        if name == "box_shell":
            orient = O.pred_batch(pts, np.zeros(20000, np.uint8), draws, "int")
            raw = O.pred_batch(pts, np.full(20000, 2, np.uint8), draws, "int")
            assert int((orient == 0).sum()) == PINS["degenerate_counts"]["box_shell_6_exactly_coplanar"]
            pos = orient == 1
            assert int((raw[pos] == 0).sum()) == PINS["degenerate_counts"]["box_shell_6_exactly_cospherical"]


def test_restatement_matches_int_oracle_random_doubles():
    rng = np.random.default_rng(5)
    # near-degenerate: points on a sphere perturbed by a few ulps, plus wide exponent spans
    n = 400
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    v = v * (1 + rng.integers(-3, 4, size=(n, 1)) * 2.0 ** -52)
    wide = rng.normal(size=(n, 3)) * np.exp2(rng.integers(-60, 60, size=(n, 3)))
    for pts in (v, wide, np.vstack([v, wide])):
        idx = rng.integers(0, len(pts), size=(20000, 5)).astype(np.int32)
        for op in (0, 1, 2):
            ops = np.full(20000, op, np.uint8)
            assert np.array_equal(O.pred_batch(pts, ops, idx, "oracle"),
                                  O.pred_batch(pts, ops, idx, "int")), op


def test_perturbation_with_aliases_matches_int_oracle():
    """Aliased indices (pointer identity in the reference) and equal-coordinate
    ties, both restated independently."""
    rng = np.random.default_rng(7)
    pts = np.vstack([O.lattice(3, 3, 3), O.lattice(3, 3, 3)])   # duplicates of every point
    idx = rng.integers(0, len(pts), size=(30000, 5)).astype(np.int32)
    idx[:5000, 4] = idx[:5000, 0]          # e aliases a
    idx[5000:10000, 3] = idx[5000:10000, 1]
    ops = np.full(30000, 1, np.uint8)
    assert np.array_equal(O.pred_batch(pts, ops, idx, "oracle"), O.pred_batch(pts, ops, idx, "int"))


@pytest.mark.skipif(not O.have_ref(), reason="reference not built here")
def test_restatement_matches_compiled_reference():
    rng = np.random.default_rng(9)
    sets = [O.random_cube(300, 4), O.sphere(300), O.lattice(4, 4, 4),
            np.vstack([O.lattice(3, 3, 3)] * 2), O.box_shell(5)]
    for pts in sets:
        idx = rng.integers(0, len(pts), size=(20000, 5)).astype(np.int32)
        ops = rng.integers(0, 3, size=20000).astype(np.uint8)
        assert np.array_equal(O.pred_batch(pts, ops, idx, "oracle"), O.pred_batch(pts, ops, idx, "ref"))
    for n in (0, 1, 2, 3, 9, 64, 100, 1000, 5000):
        pts = O.random_cube(n, 7) if n else np.zeros((0, 3))
        assert np.array_equal(O.spatial_sort(pts), O.spatial_sort(pts, which="ref"))
    lat = O.lattice(10, 10, 10)
    assert np.array_equal(O.spatial_sort(lat), O.spatial_sort(lat, which="ref"))


@pytest.mark.skipif(not O.have_ref(), reason="reference not built here")
def test_reference_delaunay_pins_hold():
    """Confirms the compiled reference is the pinned one (tests/test_delaunay.cpp:209-243)."""
    pins = PINS["delaunay_cell_hash"]

    def cell_hash(cells):
        h = 1469598103934665603
        def mix(v):
            nonlocal h
            for b in range(8):
                h ^= (v >> (8 * b)) & 0xFF
                h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
        mix(len(cells))
        for t in cells:
            for v in t:
                mix(int(v) & 0xFFFFFFFFFFFFFFFF)
        return h

    def cells_of(pts):
        xyz = np.ascontiguousarray(pts, dtype=np.float64)
        cap = 20 * max(len(xyz), 1) + 64
        out = np.zeros((cap, 4), dtype=np.int32)
        P = ctypes.POINTER(ctypes.c_double)
        I = ctypes.POINTER(ctypes.c_int32)
        m = O.ref().ref_delaunay_cells(xyz.ctypes.data_as(P), len(xyz), out.ctypes.data_as(I), cap)
        return out[:m]

    for n, h in pins["random_cube_seed4"].items():
        assert cell_hash(cells_of(O.random_cube(int(n), 4))) == int(h, 16)
    assert cell_hash(cells_of(O.lattice(5, 5, 5))) == int(pins["lattice_5_5_5"], 16)
    assert cell_hash(cells_of(O.cylinder(300))) == int(pins["cylinder_300_q32"], 16)
