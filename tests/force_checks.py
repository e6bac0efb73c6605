"""Independent checks of the lift and drag coefficients (test helpers).

The reference computes no force coefficients; the solver defines them as the
first-order Roe wall flux of the state, summed over the wall edges
(fvm_abi.h, fvm_get_forces).  These helpers provide what the mathematics
fixes independently of that formula:

* the wall polygon's geometry (edges oriented with the body on their left,
  outward normals of the fluid cells, the shoelace area), for the closed form
  of the force of a linear pressure field (Gauss: the surface integral of
  p n over the wall is the body's area times the pressure gradient);
* the force coefficients integrated from the wall pressure coefficient that
  the reference's post-processing writes (gpu_fvm.cu:182-185: node pressures
  averaged from the cells by Node_Computing, gpu_fvm.cu:71-131), a different
  discretisation of the same surface integral;
* a mesh mirrored in y, on which the flow at -alpha is the mirror image of the
  flow at alpha (Cl changes sign, Cd does not).
"""
import numpy as np

from gpueularsolver2d_amd.meshgen import Mesh


def wall_edges(topo, x, y, cx, cy):
    """Wall edges in wall-id order as (a, b, left) with 0-based node ids a, b
    ordered so that the body lies on the left of a -> b (the fluid cell's
    centroid on the right).  topo: arrays elog, wallid, node1, node2, left
    (1-based ids, as the reference numbers them, index 0 unused or not);
    cx, cy: centroids indexed by the same cell ids as topo['left']."""
    elog = np.asarray(topo["elog"])
    e = np.flatnonzero(elog == 2)
    e = e[np.argsort(np.asarray(topo["wallid"])[e])]
    a = np.asarray(topo["node1"])[e] - 1
    b = np.asarray(topo["node2"])[e] - 1
    left = np.asarray(topo["left"])[e]
    # cross((b - a), (c - a)) < 0: the fluid centroid is on the right
    cr = (x[b] - x[a]) * (cy[left] - y[a]) - (y[b] - y[a]) * (cx[left] - x[a])
    flip = cr > 0
    a2, b2 = np.where(flip, b, a), np.where(flip, a, b)
    return a2, b2, left


def polygon_area(x, y, a, b):
    """Area enclosed by the oriented edges a -> b (shoelace formula)."""
    return 0.5 * float(np.sum(x[a] * y[b] - x[b] * y[a]))


def body_normals(x, y, a, b):
    """Per edge n ds pointing into the body (out of the fluid cell)."""
    dx, dy = x[b] - x[a], y[b] - y[a]
    # the body is on the left of a -> b: its inward normal is (-dy, dx)
    return -dy, dx


def cp_force_coefficients(x, y, a, b, cp_node, alpha_deg):
    """Cl, Cd from the pressure coefficient at the wall nodes: the integral of
    Cp n ds over the wall with the trapezoidal rule on each edge (n into the
    body, chord 1), rotated by alpha into lift and drag."""
    nx, ny = body_normals(x, y, a, b)
    c = 0.5 * (cp_node[a] + cp_node[b])
    fx, fy = float(np.sum(c * nx)), float(np.sum(c * ny))
    al = np.radians(alpha_deg)
    return -fx * np.sin(al) + fy * np.cos(al), fx * np.cos(al) + fy * np.sin(al)


def mirrored(mesh):
    """The mesh reflected in the x axis, triangles kept counter-clockwise."""
    tri = np.asarray(mesh.tri)[:, [0, 2, 1]]
    return Mesh(np.array(mesh.x), -np.array(mesh.y), np.ascontiguousarray(tri), seed=mesh.seed,
                name=mesh.name + "-mirrored")


def thin_airfoil_cl(mach, alpha_deg):
    """Thin-airfoil lift 2 pi alpha with the Prandtl-Glauert factor."""
    return 2.0 * np.pi * np.radians(alpha_deg) / np.sqrt(1.0 - mach * mach)
