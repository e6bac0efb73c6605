"""Small hand-built meshes for edge cases of the tiling and the boundaries.

The reference classifies a boundary edge as a wall when its midpoint lies in
|x| < 2 and |y| < 2 and as far field otherwise (gpu_init.cu:187-204)."""
import numpy as np

from gpueularsolver2d_amd.meshgen import Mesh


def _ccw(x, y, tri):
    """Orders each triangle counter-clockwise (1-based node ids)."""
    t = np.asarray(tri, dtype=np.int32)
    a, b, c = t[:, 0] - 1, t[:, 1] - 1, t[:, 2] - 1
    area = (x[b] - x[a]) * (y[c] - y[a]) - (x[c] - x[a]) * (y[b] - y[a])
    t[area < 0] = t[area < 0][:, [0, 2, 1]]
    return t


def annulus(nt, nr, r_wall=0.5, r_far=4.0):
    """nt points around, nr layers of cells between a wall circle and a far
    circle: 2 * nt * nr triangles, every far edge outside the wall box."""
    r = np.linspace(r_wall, r_far, nr + 1)
    th = 2 * np.pi * np.arange(nt) / nt
    x = np.concatenate([ri * np.cos(th) for ri in r])
    y = np.concatenate([ri * np.sin(th) for ri in r])
    tri = []
    for k in range(nr):
        for i in range(nt):
            a, b = k * nt + i, k * nt + (i + 1) % nt
            c, d = a + nt, b + nt
            tri += [(a + 1, c + 1, d + 1), (a + 1, d + 1, b + 1)]
    return Mesh(x, y, _ccw(x, y, tri), name=f"annulus{nt}x{nr}")


def square2(h=5.0):
    """Two triangles, every boundary edge far field."""
    x = np.array([-h, h, h, -h])
    y = np.array([-h, -h, h, h])
    return Mesh(x, y, _ccw(x, y, [(1, 2, 3), (1, 3, 4)]), name="square2")
