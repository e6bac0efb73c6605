"""Synthetic NACA 0012 airfoil meshes and the reference's mesh file format.

The workloads are unstructured triangle meshes around a NACA 0012 airfoil
(thickness 0.12, chord 1): 128 cosine-clustered surface points, a far-field
circle of radius 20 chords with 48 points, and seeded interior points whose
spacing grows geometrically with distance from the airfoil.  The points,
rounded to multiples of 2^-40, are Delaunay-triangulated by this package's
own triangulator (csrc/meshgen_delaunay.cpp: Bowyer-Watson insertion with
exact integer predicates, built into libmeshgen.so), and the triangles inside
the airfoil are removed.  The result is put in a canonical order, so it does
not depend on the order in which the triangulator emits its triangles.

Interior points lie on offset curves of the (convex) airfoil polygon at
distances d_0 < d_1 < ... with spacing h(d) = h_wall + g*d along and across
the curves, so consecutive distances grow by the factor (1 + 0.866 g) far from
the body.  g is bisected to reach the requested triangle count.  Every point
inside the diametral circle of a wall or far-field edge is rejected, which
makes each boundary edge a Gabriel edge and therefore an edge of the Delaunay
triangulation; the generator checks that all 176 boundary edges are present.

The reference reads meshes in a plain-text format (gpu_init.cu:16-32):
"Nnode Ncell", then Nnode x coordinates, Nnode y coordinates, then Ncell
triangles as three 1-based node ids each.  Triangles are written
counter-clockwise (positive area, gpu_init.cu:95-112).
"""
from __future__ import annotations

import ctypes
import gzip
import math
import os
from dataclasses import dataclass

import numpy as np

N_SURFACE = 128
N_FAR = 48
R_FAR = 20.0
CENTER = (0.5, 0.0)


@dataclass
class Mesh:
    """A triangle mesh in the reference's conventions: x, y per node (0-based
    arrays) and tri (ncell, 3) of 1-based node ids, counter-clockwise."""
    x: np.ndarray
    y: np.ndarray
    tri: np.ndarray
    seed: int | None = None
    name: str = ""

    @property
    def nnode(self):
        return len(self.x)

    @property
    def ncell(self):
        return len(self.tri)


def naca0012_thickness(x, t=0.12):
    """Half thickness y_t(x) of the NACA 4-digit symmetric section."""
    return 5.0 * t * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x ** 2 + 0.2843 * x ** 3
                      - 0.1015 * x ** 4)


def naca0012_surface(n=N_SURFACE, t=0.12):
    """Counter-clockwise surface polygon: upper side from the trailing edge to
    the leading edge, then the mirrored lower side.  Cosine clustering at both
    ends; no point sits exactly at x = 0 or x = 1 (the trailing edge is blunt
    and closed by one short edge).  Upper and lower points are exact mirrors."""
    half = n // 2
    theta = math.pi * (np.arange(half) + 0.5) / half
    x = 0.5 * (1.0 + np.cos(theta))
    yt = naca0012_thickness(x, t)
    xs = np.concatenate([x, x[::-1]])
    ys = np.concatenate([yt, -yt[::-1]])
    return xs, ys


def far_circle(n=N_FAR, r=R_FAR):
    phi = 2.0 * math.pi * np.arange(n) / n
    return CENTER[0] + r * np.cos(phi), CENTER[1] + r * np.sin(phi)


def _inside_polygon(px, py, vx, vy):
    """Even-odd ray casting; vectorised over points."""
    inside = np.zeros(px.shape, dtype=bool)
    n = len(vx)
    for i in range(n):
        j = (i - 1) % n
        xi, yi, xj, yj = vx[i], vy[i], vx[j], vy[j]
        cond = (yi > py) != (yj > py)
        with np.errstate(divide="ignore", invalid="ignore"):
            xint = (xj - xi) * (py - yi) / (yj - yi) + xi
        inside ^= cond & (px < xint)
    return inside


class _OffsetCurves:
    """Offset curves of a convex polygon: in direction phi the curve at
    distance d is the polygon's support vertex plus d (cos phi, sin phi)."""

    def __init__(self, sx, sy, samples=4096):
        phi = 2.0 * math.pi * np.arange(samples) / samples
        self.c, self.s = np.cos(phi), np.sin(phi)
        support = np.argmax(np.outer(self.c, sx) + np.outer(self.s, sy), axis=1)
        self.bx, self.by = sx[support], sy[support]

    def at(self, d):
        return self.bx + d * self.c, self.by + d * self.s


def _ring_points(curves, d, h, rng, jitter=0.25):
    cx, cy = curves.at(d)
    seg = np.hypot(np.diff(np.append(cx, cx[0])), np.diff(np.append(cy, cy[0])))
    s = np.concatenate([[0.0], np.cumsum(seg)])
    perim = s[-1]
    n = max(6, int(round(perim / h)))
    u = (np.arange(n) + rng.uniform(0.0, 1.0) + rng.uniform(-jitter, jitter, n)) / n
    t = np.mod(u, 1.0) * perim
    xs = np.interp(t, s, np.append(cx, cx[0]))
    ys = np.interp(t, s, np.append(cy, cy[0]))
    return xs, ys


def _interior_points(g, curves, h_wall, rng):
    xs, ys = [], []
    d = 0.75 * h_wall
    while True:
        h = h_wall + g * d
        if d > R_FAR + 1.0:
            break
        dj = d + rng.uniform(-0.15, 0.15) * 0.866 * h
        px, py = _ring_points(curves, max(dj, 0.3 * h_wall), h, rng)
        xs.append(px)
        ys.append(py)
        d += 0.866 * h
    return np.concatenate(xs), np.concatenate(ys)


def _filter_points(px, py, sx, sy, fx, fy):
    keep = np.ones(px.shape, dtype=bool)
    # inside the far circle
    keep &= np.hypot(px - CENTER[0], py - CENTER[1]) < R_FAR * (1.0 - 1e-6)
    # outside the airfoil
    keep &= ~_inside_polygon(px, py, sx, sy)
    # outside every boundary edge's diametral circle (Gabriel protection)
    for bx, by in ((sx, sy), (fx, fy)):
        ax, ay = bx, by
        qx, qy = np.roll(bx, -1), np.roll(by, -1)
        mx, my = 0.5 * (ax + qx), 0.5 * (ay + qy)
        rad = 0.5 * np.hypot(qx - ax, qy - ay) * 1.02
        for i in range(len(bx)):
            near = np.abs(px - mx[i]) < rad[i]
            if near.any():
                keep[near] &= np.hypot(px[near] - mx[i], py[near] - my[i]) >= rad[i]
    # keep a clearance from surface vertices of 0.6 local surface spacing
    # (nearest vertex by brute force over the surface, in chunks of points)
    seg = np.hypot(np.roll(sx, -1) - sx, np.roll(sy, -1) - sy)
    local = 0.5 * (seg + np.roll(seg, 1))
    dist = np.empty(len(px))
    idx = np.empty(len(px), dtype=np.int64)
    for a in range(0, len(px), 8192):
        dx = px[a:a + 8192, None] - sx[None, :]
        dy = py[a:a + 8192, None] - sy[None, :]
        d2 = dx * dx + dy * dy
        k = np.argmin(d2, axis=1)
        idx[a:a + 8192] = k
        dist[a:a + 8192] = np.sqrt(d2[np.arange(len(k)), k])
    keep &= dist >= 0.6 * local[idx]
    return px[keep], py[keep]


GRID = 2.0 ** 40  # the triangulator's coordinate grid


def snap(v):
    """Rounds coordinates to the triangulator's grid (multiples of 2^-40)."""
    return np.round(np.asarray(v, dtype=np.float64) * GRID) / GRID


_LIB = None


def _meshgen_lib():
    """libmeshgen.so, built here first if it is missing (g++)."""
    global _LIB
    if _LIB is None:
        from . import build as _build
        if not os.path.exists(_build.MESHGEN):
            _build.build_meshgen()
        lib = ctypes.CDLL(_build.MESHGEN)
        lib.fvm_meshgen_delaunay.restype = ctypes.c_int
        lib.fvm_meshgen_delaunay.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_int)]
        _LIB = lib
    return _LIB


def delaunay(x, y):
    """Delaunay triangles (counter-clockwise, 0-based node ids) of points on
    the grid of `snap`, |x|, |y| < 32, whose convex hull edges have empty
    diametral circles (the generator's boundaries do)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.empty((2 * len(x) + 8, 3), dtype=np.int32)
    nt = ctypes.c_int(0)
    rc = _meshgen_lib().fvm_meshgen_delaunay(len(x), x.ctypes.data, y.ctypes.data, out.ctypes.data,
                                             len(out), ctypes.byref(nt))
    if rc != 0:
        raise RuntimeError({1: "points off the triangulator's grid or out of range",
                            2: "repeated point", 3: "triangle buffer too small"}.get(rc, f"error {rc}"))
    return out[:nt.value].astype(np.int64)


def _triangulate(x, y, sx, sy, n_surface, n_far):
    tri = delaunay(x, y)
    cxs = x[tri].mean(axis=1)
    cys = y[tri].mean(axis=1)
    tri = tri[~_inside_polygon(cxs, cys, sx, sy)]
    # counter-clockwise orientation
    x0, y0 = x[tri[:, 0]], y[tri[:, 0]]
    x1, y1 = x[tri[:, 1]], y[tri[:, 1]]
    x2, y2 = x[tri[:, 2]], y[tri[:, 2]]
    area2 = (y2 - y0) * (x1 - x0) - (y1 - y0) * (x2 - x0)
    flip = area2 < 0
    tri[flip, 1], tri[flip, 2] = tri[flip, 2].copy(), tri[flip, 1].copy()
    # boundary check: the boundary edges must be exactly the wall and far loops
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    e.sort(axis=1)
    keys, counts = np.unique(e[:, 0] * len(x) + e[:, 1], return_counts=True)
    if counts.max() > 2:
        raise RuntimeError("triangulation is not manifold")
    bnd = set(keys[counts == 1].tolist())
    want = set()
    for base, n in ((0, n_surface), (n_surface, n_far)):
        for i in range(n):
            a, b = base + i, base + (i + 1) % n
            want.add(min(a, b) * len(x) + max(a, b))
    if bnd != want:
        raise RuntimeError(f"boundary recovery failed: {len(want - bnd)} boundary edges missing, "
                           f"{len(bnd - want)} unexpected")
    return tri


def _hilbert_index(x, y, order=16):
    """Index along a Hilbert curve of points scaled into the unit square."""
    n = 1 << order
    xi = np.clip(((x - x.min()) / max(np.ptp(x), 1e-300) * (n - 1)).astype(np.int64), 0, n - 1)
    yi = np.clip(((y - y.min()) / max(np.ptp(y), 1e-300) * (n - 1)).astype(np.int64), 0, n - 1)
    d = np.zeros_like(xi)
    s = n >> 1
    while s > 0:
        rx = ((xi & s) > 0).astype(np.int64)
        ry = ((yi & s) > 0).astype(np.int64)
        d += s * s * ((3 * rx) ^ ry)
        # rotate
        swap = ry == 0
        flipm = swap & (rx == 1)
        xi = np.where(flipm, s - 1 - xi, xi)
        yi = np.where(flipm, s - 1 - yi, yi)
        xi, yi = np.where(swap, yi, xi), np.where(swap, xi, yi)
        s >>= 1
    return d


def naca0012_mesh(target_cells, seed, tol=0.02, max_iter=40):
    """Seeded NACA 0012 Delaunay mesh with about `target_cells` triangles.

    Node order: 128 surface points, 48 far-field points, interior points, all
    on the triangulator's grid.  Each triangle starts at its smallest node id
    (counter-clockwise), and the triangles are emitted in Hilbert order of
    their centroids (as a mesh generator with spatial locality would write
    them), ties broken by node ids."""
    sx, sy = (snap(v) for v in naca0012_surface())
    fx, fy = (snap(v) for v in far_circle())
    seg = np.hypot(np.roll(sx, -1) - sx, np.roll(sy, -1) - sy)
    h_wall = float(seg.mean())
    want_int = max(0, (int(target_cells) - (N_SURFACE + N_FAR)) // 2)

    curves = _OffsetCurves(sx, sy)

    def raw(g):
        return _interior_points(g, curves, h_wall, np.random.default_rng(seed))

    # Bisect g on the unfiltered count scaled by the filter's keep ratio, then
    # correct the ratio from one filtered attempt and repeat.
    keep = 0.95
    best = None
    for _ in range(4):
        lo, hi = 1e-4, 10.0
        for _ in range(max_iter):
            g = math.sqrt(lo * hi)
            n = len(raw(g)[0]) * keep
            if abs(n - want_int) <= 0.25 * tol * max(want_int, 1):
                break
            if n > want_int:
                lo = g
            else:
                hi = g
        rx, ry = raw(g)
        px, py = _filter_points(rx, ry, sx, sy, fx, fy)
        if best is None or abs(len(px) - want_int) < abs(len(best[0]) - want_int):
            best = (px, py)
        if abs(len(px) - want_int) <= tol * max(want_int, 1) or len(px) == 0:
            break
        keep = max(0.05, len(px) / max(len(rx), 1))
    px, py = (snap(v) for v in best)
    x = np.concatenate([sx, fx, px])
    y = np.concatenate([sy, fy, py])
    tri = _triangulate(x, y, sx, sy, N_SURFACE, N_FAR)
    # smallest node id first, the counter-clockwise order kept
    first = np.argmin(tri, axis=1)
    rows = np.arange(len(tri))
    tri = np.stack([tri[rows, (first + k) % 3] for k in range(3)], axis=1)
    cx = x[tri].mean(axis=1)
    cy = y[tri].mean(axis=1)
    # Hilbert key on a log-polar map keeps the dense near-wall region resolved
    r = np.hypot(cx - CENTER[0], cy - CENTER[1])
    key = _hilbert_index(np.log(r + 0.05), np.arctan2(cy - CENTER[1], cx - CENTER[0]))
    tri = tri[np.lexsort((tri[:, 2], tri[:, 1], tri[:, 0], key))]
    return Mesh(x=x, y=y, tri=(tri + 1).astype(np.int32), seed=seed,
                name=f"naca0012-{len(tri)}-s{seed}")


def write_dat(mesh: Mesh, path):
    """Writes the reference's text mesh format with round-trip precision."""
    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "wt") as f:
        f.write(f"{mesh.nnode} {mesh.ncell}\n")
        f.write("\n".join(repr(float(v)) for v in mesh.x))
        f.write("\n")
        f.write("\n".join(repr(float(v)) for v in mesh.y))
        f.write("\n")
        for a, b, c in mesh.tri:
            f.write(f"{a} {b} {c}\n")


def read_dat(path):
    """Reads the reference's text mesh format (gzip accepted by suffix)."""
    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "rt") as f:
        tok = f.read().split()
    nnode, ncell = int(tok[0]), int(tok[1])
    if len(tok) < 2 + 2 * nnode + 3 * ncell:
        raise ValueError(f"{path}: truncated mesh file")
    vals = tok[2:]
    x = np.array(vals[:nnode], dtype=np.float64)
    y = np.array(vals[nnode:2 * nnode], dtype=np.float64)
    tri = np.array(vals[2 * nnode:2 * nnode + 3 * ncell], dtype=np.int32).reshape(ncell, 3)
    return Mesh(x=x, y=y, tri=tri, name=str(path))
