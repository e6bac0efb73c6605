"""The oracle's lift and drag coefficients against what the mathematics fixes
independently of their formula (orc_forces: first-order Roe wall flux
momentum, fvm_oracle.c).  The reference computes no forces; its only
force-related output is the wall -Cp of Output (gpu_fvm.cu:182-185).

* Linear pressure at rest: the wall force of p = p0 + k y is -k A_body in y
  (Gauss), A_body from the shoelace formula; likewise in x.  Exact to
  rounding when each wall cell carries the pressure of its wall edge's
  midpoint (the midpoint rule is exact for a linear p on straight edges).
* Mirror symmetry: on the mesh reflected in y, the flow at -alpha is the
  mirror image of the flow at alpha: Cl changes sign, Cd does not.
* Converged subsonic lift: its slope in alpha is positive and near the
  thin-airfoil value 2 pi / sqrt(1 - M^2), Cl is linear in alpha, and the
  slope scales with the Prandtl-Glauert factor between two Mach numbers.
  (The random interior points make the mesh slightly asymmetric, so Cl at
  alpha 0 is small but not zero: slopes are taken between alphas.)
* The integral of the reference's own wall pressure coefficient (node
  averages of Node_Computing) gives the same coefficients to discretisation
  level.
A force summed from the wrong flux components, or with the opposite sign,
fails each of the first, third and fourth checks, and the second catches the
wrong components at alpha != 0."""
import os

import numpy as np
import pytest

from force_checks import (cp_force_coefficients, mirrored, polygon_area, thin_airfoil_cl,
                          wall_edges)
from oracle_ctypes import LIB_OMP_PATH, OracleMesh, OracleSolver, read_dat
from test_host_driver import reference_postprocessing

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GAMMA = 1.4


def _oracle_mesh(mesh, library=None):
    kw = {} if library is None else dict(library=library)
    return OracleMesh(mesh.x, mesh.y, mesh.tri, **kw)


@pytest.fixture(scope="module")
def naca3k():
    from gpueularsolver2d_amd import naca0012_mesh
    return naca0012_mesh(3000, seed=1)


@pytest.fixture(scope="module")
def rae3k():
    from gpueularsolver2d_amd.meshgen import Mesh
    x, y, tri = read_dat(os.path.join(GOLDEN, "2822-3k.dat.gz"))
    return Mesh(x, y, tri, name="2822-3k")


def _wall(om, mesh):
    a, b, left = wall_edges(dict(elog=om.elog, wallid=om.wallid, node1=om.node1, node2=om.node2,
                                 left=om.left), np.asarray(mesh.x), np.asarray(mesh.y), om.cx, om.cy)
    return a, b, left


@pytest.mark.parametrize("which", ["naca3k", "rae3k"])
@pytest.mark.parametrize("axis", ["y", "x"])
def test_linear_pressure_force_is_gradient_times_body_area(which, axis, request):
    mesh = request.getfixturevalue(which)
    om = _oracle_mesh(mesh)
    x, y = np.asarray(mesh.x), np.asarray(mesh.y)
    a, b, left = _wall(om, mesh)
    assert len(np.unique(left)) == len(left)  # one wall edge per wall cell
    area = polygon_area(x, y, a, b)
    assert area > 0
    p0, k = 3.0, 0.25
    p = np.full(om.ncell, p0)
    coord = y if axis == "y" else x
    p[left - 1] = p0 + k * 0.5 * (coord[a] + coord[b])
    W = np.column_stack([np.ones(om.ncell), np.zeros(om.ncell), np.zeros(om.ncell),
                         p / (GAMMA - 1.0)])
    s = OracleSolver(om, 0.5, 0.0)   # alpha 0: Cl = Fy / q, Cd = Fx / q, q = 1/2
    s.set_state(W)
    cl, cd = s.forces()
    want = -k * area / 0.5
    if axis == "y":
        assert abs(cl - want) < 1e-11 and abs(cd) < 1e-11, (cl, cd, want)
    else:
        assert abs(cd - want) < 1e-11 and abs(cl) < 1e-11, (cl, cd, want)


@pytest.mark.parametrize("alpha", [2.31, -1.5])
def test_mirrored_mesh_flips_lift_keeps_drag(naca3k, alpha):
    ma = _oracle_mesh(naca3k)
    mb = _oracle_mesh(mirrored(naca3k))
    A = OracleSolver(ma, 0.729, alpha, cfl=1.0)
    B = OracleSolver(mb, 0.729, -alpha, cfl=1.0)
    A.run(30)
    B.run(30)
    (cla, cda), (clb, cdb) = A.forces(), B.forces()
    assert abs(cla) > 1e-2
    assert abs(clb + cla) < 1e-9 * abs(cla), (cla, clb)
    assert abs(cdb - cda) < 1e-9 * max(abs(cda), 1e-3), (cda, cdb)
    # the states are mirror images: rho, rho u, rho E equal, rho v opposite
    WA, WB = A.state(), B.state()
    np.testing.assert_allclose(WB[:, 2], -WA[:, 2], rtol=0, atol=1e-10)
    np.testing.assert_allclose(WB[:, [0, 1, 3]], WA[:, [0, 1, 3]], rtol=0, atol=1e-10)


@pytest.fixture(scope="module")
def converged_first_order(naca3k):
    """First order, no smoothing, CFL 1.5, run to Resd < 1e-7 at three
    (Mach, alpha) points (all-core oracle build; the same bits)."""
    om = _oracle_mesh(naca3k, LIB_OMP_PATH)
    out = {}
    for mach, alpha in ((0.5, 0.0), (0.5, 1.0), (0.5, 2.0), (0.3, 0.0), (0.3, 1.0)):
        s = OracleSolver(om, mach, alpha, cfl=1.5, order=1, smoothing=False)
        h, n = s.run(20000, epsl=1e-7)
        assert h[-1] < 1e-7, (mach, alpha, n, h[-1])
        out[(mach, alpha)] = (s.forces(), s.state())
    return om, out


def test_converged_lift_slope_linearity_and_compressibility(converged_first_order):
    _, r = converged_first_order
    cl = {k: v[0][0] for k, v in r.items()}
    slope05 = cl[(0.5, 1.0)] - cl[(0.5, 0.0)]   # per degree
    slope03 = cl[(0.3, 1.0)] - cl[(0.3, 0.0)]
    th = thin_airfoil_cl(0.5, 1.0)
    # positive and near thin-airfoil theory (first-order dissipation on this
    # coarse mesh, 128 wall edges, takes off a few per cent)
    assert 0.85 * th < slope05 < 1.10 * th, (slope05, th)
    # linear in alpha: the second difference is small against the slope
    curv = cl[(0.5, 2.0)] - 2.0 * cl[(0.5, 1.0)] + cl[(0.5, 0.0)]
    assert abs(curv) < 0.03 * slope05, (curv, slope05)
    # the alpha-0 lift of the slightly asymmetric mesh is small
    assert abs(cl[(0.5, 0.0)]) < 0.15 * slope05
    # Prandtl-Glauert: the slope scales as 1 / sqrt(1 - M^2)
    pg = np.sqrt(1 - 0.3 ** 2) / np.sqrt(1 - 0.5 ** 2)
    assert abs((slope05 / slope03) / pg - 1.0) < 0.02, (slope05 / slope03, pg)
    # drag is positive (numerical entropy production) and small
    for (c_l, c_d), _ in r.values():
        assert 0.0 < c_d < 0.05


def test_forces_match_integrated_wall_pressure(naca3k, converged_first_order):
    om, r = converged_first_order
    x, y = np.asarray(naca3k.x), np.asarray(naca3k.y)
    a, b, _ = _wall(om, naca3k)
    for (mach, alpha), ((cl, cd), W) in r.items():
        _, mcp = reference_postprocessing(naca3k, W, mach=mach)
        clp, cdp = cp_force_coefficients(x, y, a, b, -mcp, alpha)
        # two discretisations of one surface integral: node-averaged
        # pressures with the trapezoidal rule against the cells' wall fluxes
        assert abs(clp - cl) < 0.06 * max(abs(cl), 0.1), (mach, alpha, cl, clp)
        assert abs(cdp - cd) < 0.015, (mach, alpha, cd, cdp)
