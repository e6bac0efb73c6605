"""The FP32 stage kernel forms a cell's reconstruction offsets (centroid to
the midpoint of local edge k, the side opposite vertex k) from the cell's own
side vectors instead of storing them per edge:

    r_k = (P[k+1] + P[k+2]) / 2 - (P0 + P1 + P2) / 3 = (d[k+2] - d[k+1]) / 6,
    d[i] = P[i+2] - P[i+1],

and the cell record holds g[i] = (x[N1] - x[N2], y[N2] - y[N1]) with
N1 = P[i+1], N2 = P[i+2] (the reference's gradient vectors,
gpu_solver.cu:261-265), i.e. g[i] = (-d[i].x, d[i].y), so

    r_k = (g[k+1].x - g[k+2].x, g[k+2].y - g[k+1].y) / 6.

Checked against midpoint - centroid as the reference forms them
(gpu_init.cu:100-228: centroid = vertex mean, mid = average of the edge's
nodes) on the reference's own meshes.  The swapped-index form, which a slip
in the kernel's (k + 1) % 3 / (k + 2) % 3 would give, is far off, so the check
tells the two apart."""
import os

import numpy as np
import pytest

from gpueularsolver2d_amd import read_dat

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def offsets(mesh):
    xy = np.stack([np.asarray(mesh.x, float), np.asarray(mesh.y, float)], axis=1)
    tri = np.asarray(mesh.tri, dtype=np.int64).reshape(-1, 3)
    if tri.min() == 1:  # the reference's 1-based node ids
        tri = tri - 1
    P = xy[tri]  # (cells, vertex, xy)
    cen = P.mean(axis=1)
    r = np.stack([(P[:, (k + 1) % 3] + P[:, (k + 2) % 3]) / 2 - cen for k in range(3)], axis=1)
    gx = np.stack([P[:, (i + 1) % 3, 0] - P[:, (i + 2) % 3, 0] for i in range(3)], axis=1)
    gy = np.stack([P[:, (i + 2) % 3, 1] - P[:, (i + 1) % 3, 1] for i in range(3)], axis=1)
    return r, gx, gy


def from_sides(gx, gy, a, b):
    # r_k from the side vectors k + a and k + b (mod 3)
    return np.stack([np.stack([(gx[:, (k + a) % 3] - gx[:, (k + b) % 3]) / 6,
                               (gy[:, (k + b) % 3] - gy[:, (k + a) % 3]) / 6], axis=1)
                     for k in range(3)], axis=1)


@pytest.mark.parametrize("name", ["2822-3k.dat.gz", "2822-7k.dat.gz"])
def test_offsets_from_side_vectors(name):
    r, gx, gy = offsets(read_dat(os.path.join(GOLDEN, name)))
    scale = np.linalg.norm(r, axis=2)
    assert scale.min() > 0
    err = np.linalg.norm(from_sides(gx, gy, 1, 2) - r, axis=2) / scale
    assert err.max() < 1e-12
    # the swapped form (sides k + 2 and k + 1) is the offset's negation
    bad = np.linalg.norm(from_sides(gx, gy, 2, 1) - r, axis=2) / scale
    assert bad.min() > 1.0
