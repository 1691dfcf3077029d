"""Synthetic inputs of the measured configurations (deterministic, seeded).

C1  10,000 points, coordinates = consecutive rand48 states (srand48(1)
    layout) times 2^-48, consumed x, y, z per point.
C2  the 32x32x32 integer lattice (index (i*32+j)*32+k, point (i,j,k)) and
    2^24 queries, half orient3d, half in_sphere, in random order:
      orient3d: 4 distinct points of one lattice plane through a 3x3x3 block
                (axis and diagonal planes), or 4 distinct corners of a unit cube;
      in_sphere: 5 distinct corners of a unit cube (always cospherical), or 5
                distinct lattice points on a sphere of radius^2 1, 2, 3 or 5
                around a lattice point (always cospherical); a and b are
                swapped when orient3d(a,b,c,d) < 0 and coplanar (a,b,c,d) are
                redrawn, so every in_sphere query meets its precondition.
C3  10^7 points: even indices N(0,1)^3, odd indices uniform [-1,1)^3, and
    random 5-point in_sphere / 4-point orient3d queries over them.
Everything is numpy, seeded with numpy's PCG64 except C1's rand48 stream.
"""
from __future__ import annotations

import numpy as np

MASK48 = (1 << 48) - 1


def rand48_states(seed: int, count: int) -> np.ndarray:
    """The rand48 states after 1..count steps from srand48(seed), as uint64."""
    a, c = 0x5DEECE66D, 0xB
    x = ((seed << 16) | 0x330E) & MASK48
    out = np.empty(count, dtype=np.uint64)
    for i in range(count):
        x = (a * x + c) & MASK48
        out[i] = x
    return out


def c1_points(n: int = 10_000, seed: int = 1) -> np.ndarray:
    st = rand48_states(seed, 3 * n)
    return (st.astype(np.float64) * (1.0 / 281474976710656.0)).reshape(n, 3)


def lattice(nx: int, ny: int, nz: int) -> np.ndarray:
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    return np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)


def _distinct_rows(rng, sizes: np.ndarray, k: int) -> np.ndarray:
    """Per row, k distinct integers in [0, sizes[row])."""
    n = len(sizes)
    out = np.empty((n, k), dtype=np.int64)
    todo = np.arange(n)
    while len(todo):
        draw = (rng.random((len(todo), k)) * sizes[todo, None]).astype(np.int64)
        s = np.sort(draw, axis=1)
        ok = np.all(s[:, 1:] != s[:, :-1], axis=1)
        out[todo[ok]] = draw[ok]
        todo = todo[~ok]
    return out


_CUBE = np.array([[i, j, k] for i in (0, 1) for j in (0, 1) for k in (0, 1)], dtype=np.int64)


def _sphere_shells():
    shells = {}
    for r2 in (1, 2, 3, 5):
        pts = [(x, y, z) for x in range(-2, 3) for y in range(-2, 3) for z in range(-2, 3)
               if x * x + y * y + z * z == r2]
        shells[r2] = np.array(pts, dtype=np.int64)
    return shells


def _plane_patches():
    """Lattice planes through the 27 points of a 3x3x3 block with >= 4 points."""
    local = np.array([[i, j, k] for i in range(3) for j in range(3) for k in range(3)])
    normals = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, -1, 0), (1, 0, 1), (1, 0, -1),
               (0, 1, 1), (0, 1, -1), (1, 1, 1), (1, 1, -1), (1, -1, 1), (-1, 1, 1)]
    patches = []
    for nrm in normals:
        dots = local @ np.array(nrm)
        for c in np.unique(dots):
            members = local[dots == c]
            if len(members) >= 4:
                patches.append(members)
    size = np.array([len(p) for p in patches])
    table = np.zeros((len(patches), 9, 3), dtype=np.int64)
    for i, p in enumerate(patches):
        table[i, :len(p)] = p
    return table, size


def _orient_sign_int(a, b, c, d) -> np.ndarray:
    """sign |b-a, c-a, d-a| for small integer coordinates (exact in int64)."""
    u, v, w = b - a, c - a, d - a
    det = (u[:, 0] * (v[:, 1] * w[:, 2] - v[:, 2] * w[:, 1])
           - u[:, 1] * (v[:, 0] * w[:, 2] - v[:, 2] * w[:, 0])
           + u[:, 2] * (v[:, 0] * w[:, 1] - v[:, 1] * w[:, 0]))
    return np.sign(det)


def c2(n_q: int = 1 << 24, seed: int = 2, side: int = 32):
    """(points, op, idx) of the degenerate-lattice configuration."""
    rng = np.random.default_rng(seed)
    pts = lattice(side, side, side)

    def lin(p):
        return (p[..., 0] * side + p[..., 1]) * side + p[..., 2]

    op = rng.integers(0, 2, size=n_q).astype(np.uint8)  # 0 orient3d, 1 in_sphere
    idx = np.zeros((n_q, 5), dtype=np.int64)

    # orient3d
    o_rows = np.nonzero(op == 0)[0]
    half = rng.random(len(o_rows)) < 0.5
    patch_tab, patch_size = _plane_patches()
    r = o_rows[half]
    pid = rng.integers(0, len(patch_tab), size=len(r))
    base = rng.integers(0, side - 2, size=(len(r), 3))
    pick = _distinct_rows(rng, patch_size[pid], 4)
    pp = patch_tab[pid[:, None], pick] + base[:, None, :]
    idx[r, :4] = lin(pp)
    r = o_rows[~half]
    base = rng.integers(0, side - 1, size=(len(r), 3))
    pick = _distinct_rows(rng, np.full(len(r), 8), 4)
    idx[r, :4] = lin(_CUBE[pick] + base[:, None, :])
    idx[o_rows, 4] = idx[o_rows, 3]

    # in_sphere, redrawn until (a,b,c,d) is not coplanar, then oriented
    s_rows = np.nonzero(op == 1)[0]
    shells = _sphere_shells()
    todo = s_rows
    while len(todo):
        kind = rng.random(len(todo)) < 0.5
        P = np.zeros((len(todo), 5, 3), dtype=np.int64)
        r = np.nonzero(kind)[0]
        base = rng.integers(0, side - 1, size=(len(r), 3))
        pick = _distinct_rows(rng, np.full(len(r), 8), 5)
        P[r] = _CUBE[pick] + base[:, None, :]
        r = np.nonzero(~kind)[0]
        r2 = rng.choice(np.array([1, 2, 3, 5]), size=len(r))
        center = rng.integers(2, side - 2, size=(len(r), 3))
        for val, shell in shells.items():
            m = r[r2 == val]
            if len(m) == 0:
                continue
            pick = _distinct_rows(rng, np.full(len(m), len(shell)), 5)
            P[m] = shell[pick] + center[r2 == val][:, None, :]
        s = _orient_sign_int(P[:, 0], P[:, 1], P[:, 2], P[:, 3])
        neg = s < 0
        P[neg, 0], P[neg, 1] = P[neg, 1].copy(), P[neg, 0].copy()
        ok = s != 0
        idx[todo[ok]] = lin(P[ok])
        todo = todo[~ok]
    return pts, op, idx.astype(np.int32)


def c3_points(n: int = 10_000_000, seed: int = 3) -> np.ndarray:
    rng = np.random.default_rng(seed)
    pts = np.empty((n, 3), dtype=np.float64)
    ne = (n + 1) // 2
    pts[0::2] = rng.standard_normal((ne, 3))
    pts[1::2] = 2.0 * rng.random((n - ne, 3)) - 1.0
    return pts


def random_queries(n_pts: int, n_q: int, seed: int):
    """Random mixed queries (half orient3d, half in_sphere) over n_pts points,
    indices drawn independently (no precondition)."""
    rng = np.random.default_rng(seed)
    op = rng.integers(0, 2, size=n_q).astype(np.uint8)
    idx = rng.integers(0, n_pts, size=(n_q, 5), dtype=np.int64).astype(np.int32)
    return op, idx


# ---- batched and served configurations ---------------------------------------

def c4_sizes(n_sets: int = 4096, seed: int = 4, lo: int = 100, hi: int = 200_000) -> np.ndarray:
    """Log-uniform set sizes n_s = floor(lo * (hi/lo)^U)."""
    rng = np.random.default_rng(seed)
    return np.floor(lo * (hi / lo) ** rng.random(n_sets)).astype(np.int64)


def set_points(kind: int, n: int, rng) -> np.ndarray:
    """kind 0: uniform in the unit cube; 1: unit-sphere surface snapped to the
    2^-10 grid (many ties and near-cospherical points); 2: an integer lattice
    plus a jitter snapped to 2^-10."""
    if kind == 0:
        return rng.random((n, 3))
    if kind == 1:
        v = rng.standard_normal((n, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        return np.round(v * 1024.0) / 1024.0
    m = max(2, int(np.ceil(n ** (1.0 / 3.0))))
    g = np.stack(np.unravel_index(np.arange(n), (m, m, m)), axis=1).astype(np.float64)
    return g + np.round(rng.random((n, 3)) * 64.0) / 1024.0


def c4_points(sizes: np.ndarray, seed: int = 4) -> np.ndarray:
    """All sets' points, concatenated (set s has kind s mod 3)."""
    rng = np.random.default_rng(seed + 1)
    out = np.empty((int(sizes.sum()), 3))
    off = 0
    for s, n in enumerate(sizes):
        out[off:off + n] = set_points(s % 3, int(n), rng)
        off += n
    return out


def _splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def hashed_queries(seed: int, n_set: int, q: np.ndarray):
    """Host regeneration of the device generator's queries (op, idx) with
    global indices q of a set of n_set points (libdel3d_workloads.so)."""
    q = np.asarray(q, dtype=np.uint64)
    base = np.uint64(seed) * np.uint64(1 << 40) + np.uint64(8) * q
    idx = np.empty((len(q), 5), dtype=np.int32)
    with np.errstate(over="ignore"):
        for k in range(5):
            h = _splitmix64(base + np.uint64(k)) >> np.uint64(32)
            idx[:, k] = ((h * np.uint64(n_set)) >> np.uint64(32)).astype(np.int32)
        op = (_splitmix64(base + np.uint64(5)) & np.uint64(1)).astype(np.uint8)
    return op, idx


# ---- coplanar predicates -----------------------------------------------------

_RING25 = [(5, 0), (4, 3), (3, 4), (0, 5), (-3, 4), (-4, 3), (-5, 0), (-4, -3), (-3, -4),
           (0, -5), (3, -4), (4, -3)]
# M M^T = 9 I: maps the z = 0 plane isometrically (up to the factor 3) onto a
# tilted plane with integer coordinates, so circles stay circles.
_TILT = np.array([[1, 2, 2], [2, 1, -2], [2, -2, 1]], dtype=np.float64)


def coplanar_points(kind: str, rng, n: int = 400) -> np.ndarray:
    """Point sets for the coplanar predicates.

    grid_z0      integer grid [-3,3]^2 in the plane z = 0
    tilted       the same grid on the plane z = x + y
    ring         the 12 lattice points of x^2 + y^2 = 25, its centre and (9,9)
    ring_tilted  the ring mapped by an integer similarity onto a tilted plane
    ring_offset  ring_tilted scaled by 1024 and translated by a 2^-10-grid vector
                 (still exactly cocircular, too wide for the certified FP64 tier)
    dyadic_plane random 2^-20-grid points on z = 3x/4 - 5y/8 + 1/2 (exact)
    line         points p + t d on one line, t small integers (collinear input)
    scaled       2^-20-grid points on z = 3x/4 - 5y/8, half scaled by 2^-60, half
                 by 2^+60 (two clusters on one plane)
    random       uniform points in the unit cube (no precondition holds)
    """
    if kind in ("grid_z0", "tilted"):
        g = np.array([(i, j, 0.0) for i in range(-3, 4) for j in range(-3, 4)])
        if kind == "tilted":
            g[:, 2] = g[:, 0] + g[:, 1]
        return g
    if kind in ("ring", "ring_tilted", "ring_offset"):
        r = np.array([(x, y, 0.0) for x, y in _RING25] + [(0, 0, 0.0), (9, 9, 0.0)])
        if kind == "ring":
            return r
        r = r @ _TILT.T
        if kind == "ring_offset":  # scaled and moved off the small grid: still cocircular
            r = r * 1024.0 + np.array([0.5, -0.25, 0.125]) + 3 * 2.0 ** -10
        return r
    if kind in ("dyadic_plane", "scaled"):
        xy = rng.integers(-(1 << 20), 1 << 20, size=(n, 2)) * 2.0 ** -20
        z = 0.75 * xy[:, 0] - 0.625 * xy[:, 1] + (0.5 if kind == "dyadic_plane" else 0.0)
        p = np.column_stack([xy, z])
        if kind == "scaled":  # the plane through the origin, so both clusters lie on it
            p = np.vstack([p[: n // 2] * 2.0 ** -60, p[n // 2:] * 2.0 ** 60])
        return p
    if kind == "line":
        t = rng.integers(-40, 40, size=n).astype(np.float64)
        p0 = np.array([0.5, -1.25, 3.0])
        d = np.array([0.125, 0.75, -0.5])
        return p0 + t[:, None] * d
    if kind == "random":
        return rng.random((n, 3))
    raise ValueError(kind)


COPLANAR_KINDS = ("grid_z0", "tilted", "ring", "ring_tilted", "ring_offset", "dyadic_plane", "line",
                  "scaled", "random")


def coplanar_queries(n_pts: int, n_q: int, rng, ops=tuple(range(8))):
    """Random coplanar-batch queries: op codes from ops, indices drawn
    independently, with about one query in eight repeating an earlier slot
    (equal index = the same point)."""
    op = rng.choice(np.asarray(ops, dtype=np.uint8), size=n_q).astype(np.uint8)
    idx = rng.integers(0, n_pts, size=(n_q, 5)).astype(np.int32)
    m = n_q // 8
    src = rng.integers(0, 3, size=m)
    idx[np.arange(m), 3] = idx[np.arange(m), src]
    idx[:, 4] = 0
    return op, idx


def coplanar_lattice(n_q: int = 1 << 24, seed: int = 6, side: int = 256):
    """(points, op, idx) of the coplanar benchmark: a side x side integer grid
    mapped by the similarity _TILT onto a tilted plane (so its cocircular and
    collinear subsets stay so in 3D), and n_q queries whose points come from a
    4x4 window at a random position (squares, rectangles and grid lines: many
    exactly cocircular or collinear). Op mix: 40% coplanar_in_circle, 10%
    unperturbed, 20% coplanar_orientation(p,q,r,s), 10% coplanar_orientation,
    10% collinear, 5% collinear_position, 5% orient2d."""
    rng = np.random.default_rng(seed)
    i, j = np.meshgrid(np.arange(side), np.arange(side), indexing="ij")
    g = np.stack([i.ravel(), j.ravel(), np.zeros(side * side)], axis=1).astype(np.float64)
    pts = g @ _TILT.T
    ops = np.array([5] * 8 + [6] * 2 + [1] * 4 + [0] * 2 + [2] * 2 + [4] + [7], np.uint8)
    op = ops[rng.integers(0, len(ops), size=n_q)]
    base = rng.integers(0, side - 3, size=(n_q, 2))
    off = rng.integers(0, 4, size=(n_q, 4, 2))
    ij = base[:, None, :] + off
    idx = np.zeros((n_q, 5), np.int32)
    idx[:, :4] = (ij[..., 0] * side + ij[..., 1]).astype(np.int32)
    return pts, op, idx
