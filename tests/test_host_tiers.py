"""The GPU's per-query arithmetic (del3d_amd/csrc/pred_core.h), built for the
host, against the big-integer oracle: every tier, including tiers forced on
queries the kernels would decide earlier."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O

LIB = Path(__file__).resolve().parents[1] / "del3d_amd" / "libpredcore_host.so"


def run(pts, op, idx, mode):
    lib = C.CDLL(str(LIB))
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    op = np.ascontiguousarray(op, np.uint8)
    idx = np.ascontiguousarray(idx, np.int32)
    out = np.zeros(len(op), np.int8)
    tier = np.zeros(len(op), np.int8)
    lib.predcore_host_batch(pts.ctypes.data_as(C.c_void_p), C.c_int64(len(pts)),
                            op.ctypes.data_as(C.c_void_p), idx.ctypes.data_as(C.c_void_p),
                            C.c_int64(len(op)), out.ctypes.data_as(C.c_void_p), C.c_int(mode),
                            tier.ctypes.data_as(C.c_void_p))
    return out, tier


def point_sets():
    rng = np.random.default_rng(3)
    return {
        "lattice": O.lattice(4, 4, 4),
        "duplicates": np.vstack([O.lattice(3, 3, 3)] * 2),
        "sphere_q8": O.sphere(400, 8),
        "box_shell": O.box_shell(6),
        "cube": O.random_cube(300, 4),
        "wide_exponents": rng.normal(size=(200, 3)) * np.exp2(rng.integers(-80, 80, size=(200, 3))),
        "tiny": rng.normal(size=(100, 3)) * 1e-100,
        "mixed_scales": np.vstack([rng.normal(size=(100, 3)), rng.normal(size=(100, 3)) * 2.0 ** -40]),
        "grid_2^-48": rng.integers(0, 2 ** 48, size=(300, 3)) * 2.0 ** -48,
        "signed_zero": np.array([[0.0, -0.0, 0.0], [-0.0, 0.0, 1.0], [1.0, 0.0, -0.0],
                                 [0.0, 1.0, 0.0], [0.0, 0.0, -0.0], [1.0, 1.0, 0.0]]),
        "denormals": np.array([[5e-324, 0, 0], [0, 5e-324, 0], [0, 0, 5e-324], [1e-310, 1e-310, 0],
                               [0, 0, 0], [2.5e-323, 1e-323, 0]]),
    }


@pytest.mark.parametrize("name", list(point_sets().keys()))
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_tiers_match_int_oracle(name, mode):
    if not LIB.exists():
        pytest.fail(f"{LIB} not built")
    pts = point_sets()[name]
    rng = np.random.default_rng(hash(name) % 2 ** 32)
    n = 3000
    idx = rng.integers(0, len(pts), size=(n, 5)).astype(np.int32)
    idx[:400, 4] = idx[:400, 0]      # e aliases a
    idx[400:600, 2] = idx[400:600, 1]
    op = rng.integers(0, 3, size=n).astype(np.uint8)
    want = O.pred_batch(pts, op, idx, "int")
    got, tier = run(pts, op, idx, mode)
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0, (name, mode, bad[:5], op[bad[:5]], got[bad[:5]], want[bad[:5]])
    if mode == 1:
        assert set(np.unique(tier)) <= {2, 3}
    if mode == 2:
        assert set(np.unique(tier)) == {3}


def _set_grid(pts):
    """(gmin, mexp) of a point set from the doubles' bits: the lowest set
    bit's exponent and floor(log2 |c|) over nonzero coordinates."""
    c = np.asarray(pts, dtype=np.float64).ravel()
    c = c[c != 0]
    m, e = np.frexp(np.abs(c))  # |c| = m * 2^e, m in [0.5, 1)
    mexp = int((e - 1).max())
    lsb = []
    for v in c:
        n, ex = float(abs(v)).as_integer_ratio()  # v = n / 2^k exactly
        t = (ex & -ex).bit_length() - 1  # ex is a power of two
        lsb.append((n & -n).bit_length() - 1 - t)
    return int(min(lsb)), mexp


@pytest.mark.parametrize("scale", [0, -10, -48, 20, -300])
def test_set_level_certificate_implies_per_coordinate(scale):
    """Whenever the set summary certifies a query, the per-coordinate test
    does too (so both tiers accept exactly the exact cases); on coarse grids
    the set test accepts most degenerate queries."""
    lib = C.CDLL(str(LIB))
    rng = np.random.default_rng(1000 + scale)
    for trial in range(40):
        span = int(rng.integers(1, 40))
        pts = rng.integers(-2 ** span, 2 ** span, size=(60, 3)).astype(np.float64) * 2.0 ** scale
        if trial % 4 == 1:
            pts[0, 0] += 2.0 ** (scale - 3)  # one finer coordinate lowers gmin
        gmin, mexp = _set_grid(pts)
        both = np.zeros(2, np.intc)
        accepted = 0
        for _ in range(60):
            np_ = int(rng.choice([4, 5]))
            q = np.ascontiguousarray(pts[rng.integers(0, len(pts), np_)])
            lib.predcore_host_t1(q.ctypes.data_as(C.c_void_p), C.c_int(np_), C.c_int32(gmin),
                                 C.c_int32(mexp), both.ctypes.data_as(C.c_void_p))
            assert not (both[1] and not both[0]), (scale, span, q)
            accepted += int(both[1])
        if span <= 5 and trial % 4 != 1:
            assert accepted > 0



def _pack(pts):
    lib = C.CDLL(str(LIB))
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    gmin, mexp = _set_grid(pts) if np.any(pts != 0) else (1 << 30, -(1 << 30))
    words = np.zeros(len(pts), np.uint32)
    back = np.zeros_like(pts)
    ok = lib.predcore_pack_points(pts.ctypes.data_as(C.c_void_p), C.c_int64(len(pts)), C.c_int32(gmin),
                                  C.c_int32(mexp), words.ctypes.data_as(C.c_void_p),
                                  back.ctypes.data_as(C.c_void_p))
    return bool(ok), words, back


def _packed_sets():
    rng = np.random.default_rng(7)
    lat = O.lattice(32, 32, 32)
    return rng, {
        "lattice32": lat,
        "lattice32_2^-40": lat * 2.0 ** -40,
        "lattice32_2^-200": lat * 2.0 ** -200,
        "lattice32_2^140": lat * 2.0 ** 140,
        "centered_span255": rng.integers(-255, 256, size=(3000, 3)).astype(np.float64),
        "span256": np.vstack([rng.integers(-255, 256, size=(3000, 3)), [[256, 0, 0]]]).astype(np.float64),
        "signed_zeros": np.where(rng.random((2000, 3)) < 0.3, -0.0,
                                 rng.integers(-200, 200, size=(2000, 3)).astype(np.float64)),
        "span512": np.vstack([lat, [[512.0, 0.0, 0.0]]]),
        "offset_lattice": lat + 1000.0,
        "lattice32_2^-300": lat * 2.0 ** -300,
        "lattice32_2^200": lat * 2.0 ** 200,
        "all_zero": np.zeros((100, 3)),
        "random": rng.normal(size=(500, 3)),
    }


@pytest.mark.parametrize("kind,packable", [
    ("lattice32", True), ("lattice32_2^-40", True), ("lattice32_2^-200", True), ("lattice32_2^140", True),
    ("centered_span255", True), ("signed_zeros", True), ("span256", False), ("span512", False),
    ("offset_lattice", False),
    ("lattice32_2^-300", False), ("lattice32_2^200", False), ("all_zero", False), ("random", False)])
def test_packed_point_format_round_trip(kind, packable):
    """The packed small-grid copy (one 32-bit word a point) that the packed
    LDS kernel reads: a set qualifies exactly when its values span at most 8
    significant bits on a grid 2^gmin with -206 <= gmin <= 171; then every
    coordinate comes back (-0.0 as 0.0) and comparing words is comparing
    points lexicographically."""
    _, sets = _packed_sets()
    pts = sets[kind]
    ok, words, back = _pack(pts)
    assert ok == packable
    if ok:
        assert np.array_equal(pts, back)  # == : -0.0 equals 0.0
        assert np.all(words >> 30 == 0)
        i = np.random.default_rng(1).integers(0, len(pts), size=(4000, 2))
        a, b = pts[i[:, 0]], pts[i[:, 1]]
        lex = (a[:, 0] < b[:, 0]) | ((a[:, 0] == b[:, 0]) & ((a[:, 1] < b[:, 1]) |
                                                              ((a[:, 1] == b[:, 1]) & (a[:, 2] < b[:, 2]))))
        assert np.array_equal(lex, words[i[:, 0]] < words[i[:, 1]])


@pytest.mark.parametrize("kind", ["lattice32", "lattice32_2^-40", "lattice32_2^-200", "lattice32_2^140",
                                  "centered_span255", "signed_zeros"])
def test_packed_decision_matches_oracle(kind):
    """The packed kernel's decision on the integer grid (the certified
    branch, taken by every query of a packable set), built for the host,
    against the oracle on local-cell queries (most of them degenerate) and
    random ones (the widest differences of the set among them)."""
    rng, sets = _packed_sets()
    pts = np.ascontiguousarray(sets[kind], dtype=np.float64)
    gmin, mexp = _set_grid(pts)
    n = 20000
    op = rng.integers(0, 3, size=n).astype(np.uint8)
    idx = rng.integers(0, len(pts), size=(n, 5)).astype(np.int32)
    if "lattice" in kind:
        base = rng.integers(0, 30, size=(n // 2, 3))
        off = rng.integers(0, 3, size=(n // 2, 5, 3))
        ijk = base[:, None, :] + off
        idx[: n // 2] = ((ijk[..., 0] * 32 + ijk[..., 1]) * 32 + ijk[..., 2]).astype(np.int32)
    out = np.zeros(n, np.int8)
    lib = C.CDLL(str(LIB))
    ok = lib.predcore_packed_batch(pts.ctypes.data_as(C.c_void_p), C.c_int64(len(pts)), C.c_int32(gmin),
                                   C.c_int32(mexp), op.ctypes.data_as(C.c_void_p),
                                   idx.ctypes.data_as(C.c_void_p), C.c_int64(n),
                                   out.ctypes.data_as(C.c_void_p))
    assert ok == 1
    assert np.array_equal(out, O.pred_batch(pts, op, idx, "oracle"))


@pytest.mark.parametrize("g,packable", [(-206, True), (-207, False), (171, True), (172, False)])
def test_packable_range_edges(g, packable):
    """Span-255 sets at the edges of the packable range (grid 2^-206 .. 2^171):
    random integer points in [-255, 255] with an 8^3 unit lattice block,
    scaled by 2^g. At -206 the in_sphere products of the lattice's unit
    differences are multiples of 2^-1030 (subnormal); at 171 the largest reach
    2^900. Inside the range the packed decision equals the oracle and the
    big-integer oracle on local-cell (mostly degenerate) and random queries;
    one step outside, the set is not packable."""
    rng = np.random.default_rng(1000 + g)
    lat = O.lattice(8, 8, 8) - 4.0
    ints = rng.integers(-255, 256, size=(1500, 3)).astype(np.float64)
    ints[0] = [255.0, -255.0, 1.0]  # the full span, and an odd value: the grid is exactly 2^g
    pts = np.ascontiguousarray(np.vstack([lat, ints]) * 2.0 ** g)
    ok, _, back = _pack(pts)
    assert ok == packable
    if not ok:
        return
    assert np.array_equal(pts, back)
    gmin, mexp = _set_grid(pts)
    assert gmin == g and mexp - gmin == 7
    n = 20000
    op = rng.integers(0, 3, size=n).astype(np.uint8)
    idx = rng.integers(0, len(pts), size=(n, 5)).astype(np.int32)
    base = rng.integers(0, 6, size=(n // 2, 3))
    ijk = base[:, None, :] + rng.integers(0, 3, size=(n // 2, 5, 3))
    idx[: n // 2] = ((ijk[..., 0] * 8 + ijk[..., 1]) * 8 + ijk[..., 2]).astype(np.int32)
    out = np.zeros(n, np.int8)
    lib = C.CDLL(str(LIB))
    assert lib.predcore_packed_batch(pts.ctypes.data_as(C.c_void_p), C.c_int64(len(pts)), C.c_int32(gmin),
                                     C.c_int32(mexp), op.ctypes.data_as(C.c_void_p),
                                     idx.ctypes.data_as(C.c_void_p), C.c_int64(n),
                                     out.ctypes.data_as(C.c_void_p)) == 1
    want = O.pred_batch(pts, op, idx, "oracle")
    assert np.array_equal(out, want)
    assert np.array_equal(out, O.pred_batch(pts, op, idx, "int"))
    # the local cells are degenerate: exact zeros of orient3d and of the
    # unperturbed in_sphere among them
    assert np.count_nonzero((out[: n // 2] == 0) & (op[: n // 2] != 1)) > 500
