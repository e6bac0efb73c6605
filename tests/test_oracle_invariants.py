"""Mesh-level invariants of the oracle: topology against the reference meshes,
and identities of the gradient, flux, gather, smoothing and forces that hold
whatever the mesh."""
import os

import numpy as np
import pytest

from oracle_ctypes import OracleMesh, OracleSolver, dptr, lib, read_dat

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# Counts of the reference meshes (Nnode, Ncell, Nedge, wall edges, far edges),
# measured independently of this code from the reference's .dat files.
REFERENCE_COUNTS = {"2822-3k": (1907, 3558, 5465, 128, 128),
                    "2822-7k": (7140, 13513, 20653, 511, 256)}


@pytest.fixture(scope="module")
def rae():
    x, y, tri = read_dat(os.path.join(GOLDEN, "2822-3k.dat.gz"))
    return OracleMesh(x, y, tri)


@pytest.mark.parametrize("name", sorted(REFERENCE_COUNTS))
def test_reference_mesh_counts_literal_and_sorted(name):
    x, y, tri = read_dat(os.path.join(GOLDEN, name + ".dat.gz"))
    lit = OracleMesh(x, y, tri, literal=True)
    fast = OracleMesh(x, y, tri, literal=False)
    assert (lit.nnode, lit.ncell, lit.nedge, lit.nwall, lit.nfar) == REFERENCE_COUNTS[name]
    for k in ("celledge", "neicell", "clog", "elog", "left", "right", "farid", "wallid", "node1",
              "node2", "vx", "vy", "vn", "midx", "midy"):
        assert np.array_equal(getattr(lit, k), getattr(fast, k)), k


def test_topology_structure(rae):
    m = rae
    E = m.nedge
    elog = m.elog[1:]
    # changeEdge order: interior, then far field, then wall
    assert np.all(np.diff(elog) >= 0)
    assert (elog == 1).sum() == m.nfar and (elog == 2).sum() == m.nwall
    # far / wall ids are a permutation of 1..n
    assert sorted(m.farid[1:][elog == 1]) == list(range(1, m.nfar + 1))
    assert sorted(m.wallid[1:][elog == 2]) == list(range(1, m.nwall + 1))
    # every interior edge appears in exactly its two cells, boundary edges in one
    cnt = np.bincount(m.celledge[1:].reshape(-1), minlength=E + 1)[1:]
    assert np.all(cnt == np.where(elog == 0, 2, 1))
    # left is the lower cell index of an interior edge
    interior = elog == 0
    assert np.all(m.left[1:][interior] < m.right[1:][interior])
    # CLog = number of interior neighbours
    assert np.array_equal(m.clog[1:], (m.neicell[1:] > 0).sum(axis=1))


def test_cell_closure(rae):
    # the three local edge vectors of a triangle sum to zero (n ds closes)
    m = rae
    s = OracleSolver(m, 0.5, 0.0)
    lefts = m.left[1:]
    acc = np.zeros((m.ncell + 1, 2))
    for e in range(1, m.nedge + 1):
        acc[m.left[e]] += (m.vy[e], m.vx[e])
        if m.right[e] > 0:
            acc[m.right[e]] -= (m.vy[e], m.vx[e])
    assert np.abs(acc[1:]).max() < 1e-15
    del s, lefts


def test_freestream_preservation_and_constant_gradient(rae):
    # a uniform freestream is a steady state away from the wall: the gradient
    # vanishes and the residual is zero to rounding in every cell that does
    # not touch the wall
    m = rae
    s = OracleSolver(m, 0.6, 2.0)
    r = s.eval_residual()
    wall_cells = set(m.left[1:][m.elog[1:] == 2].tolist())
    interior = np.array([c not in wall_cells for c in range(1, m.ncell + 1)])
    assert np.abs(r["res"][interior]).max() < 1e-13
    g = r["grad"][interior]
    assert np.abs(g).max() < 1e-9
    # the bounds of a uniform state collapse onto the state
    assert np.array_equal(r["umax"][interior], r["umin"][interior])


def test_conservation_telescopes(rae):
    # interior fluxes cancel pairwise: the sum of all cell residuals equals the
    # sum of the boundary-edge fluxes computed on their own
    m = rae
    s = OracleSolver(m, 0.729, 2.31, cfl=1.0)
    s.run(25)
    r = s.eval_residual()
    W = s.state()
    total = np.zeros(4)
    for e in range(1, m.nedge + 1):
        if m.elog[e] == 0:
            continue
        L = W[m.left[e] - 1].copy()
        ghost = np.zeros(4)
        if m.elog[e] == 1:
            lib().orc_far_ghost(s.h, dptr(L), m.vx[e], m.vy[e], dptr(ghost))
        else:
            lib().orc_wall_ghost(dptr(L), m.vx[e], m.vy[e], m.vn[e], dptr(ghost))
        F = np.zeros(4)
        lib().orc_roe(dptr(L), dptr(ghost), m.vy[e] / m.vn[e], m.vx[e] / m.vn[e], 1.4, dptr(F))
        total += F * m.vn[e]
    np.testing.assert_allclose(r["res"].sum(axis=0), total, rtol=1e-9, atol=1e-12)


def test_spectral_radius_sum_and_dt(rae):
    m = rae
    s = OracleSolver(m, 0.729, 2.31, cfl=1.0)
    s.run(5)
    r = s.eval_residual()
    np.testing.assert_array_equal(r["dt"], np.minimum(m.area[1:] * 1.0 / r["sr"], 0.1))


def test_smoothing_fixed_point(rae):
    # a constant residual field is a fixed point of both Jacobi sweeps
    m = rae
    s = OracleSolver(m, 0.5, 0.0)
    R = np.tile(np.array([1.0, -2.5, 0.125, 3.0]), (m.ncell, 1))
    np.testing.assert_allclose(s.smooth(R), R, rtol=1e-15)


def test_smoothing_hand_computed():
    # two triangles sharing one edge: CLog = 1 each
    x = np.array([0.0, 1.0, 1.0, 0.0])
    y = np.array([0.0, 0.0, 1.0, 1.0])
    m = OracleMesh(x, y, np.array([[1, 2, 3], [1, 3, 4]]))
    s = OracleSolver(m, 0.5, 0.0, eps=0.8)
    R = np.array([[1.0, 0, 0, 0], [3.0, 0, 0, 0]])
    out = s.smooth(R)
    r1 = np.array([(0.8 * 3.0 + 1.0) / 1.8, (0.8 * 1.0 + 3.0) / 1.8])
    r2 = np.array([(0.8 * r1[1] + 1.0) / 1.8, (0.8 * r1[0] + 3.0) / 1.8])
    expect = np.where(np.abs(R[:, 0]) > np.abs(r2), r2, R[:, 0])
    np.testing.assert_allclose(out[:, 0], expect, rtol=1e-15)


def test_gradient_edge_assembly(rae):
    # The oracle forms each cell's Green-Gauss gradient from its own three
    # node-ordered edges.  Assemble it independently by scattering over the
    # edge list with the edge orientation (n ds = (vy, vx) out of the left
    # cell), using face values 1/2 (U_L + U_R or ghost); orientation, index or
    # sign mistakes in either formulation break the agreement.
    m = rae
    s = OracleSolver(m, 0.729, 2.31, cfl=1.0)
    s.run(15)
    W = s.state()
    r = s.eval_residual()
    acc = np.zeros((m.ncell, 2, 4))
    for e in range(1, m.nedge + 1):
        L = m.left[e] - 1
        if m.elog[e] == 0:
            UR = W[m.right[e] - 1]
        else:
            UR = np.zeros(4)
            if m.elog[e] == 1:
                lib().orc_far_ghost(s.h, dptr(W[L].copy()), m.vx[e], m.vy[e], dptr(UR))
            else:
                lib().orc_wall_ghost(dptr(W[L].copy()), m.vx[e], m.vy[e], m.vn[e], dptr(UR))
        acc[L, 0] += 0.5 * (W[L] + UR) * m.vy[e]
        acc[L, 1] += 0.5 * (W[L] + UR) * m.vx[e]
        if m.elog[e] == 0:
            R = m.right[e] - 1
            acc[R, 0] -= 0.5 * (W[R] + W[L]) * m.vy[e]
            acc[R, 1] -= 0.5 * (W[R] + W[L]) * m.vx[e]
    acc /= m.area[1:, None, None]
    scale = np.abs(acc).max()
    assert np.abs(r["grad"] - acc).max() < 1e-11 * scale


def test_forces_vanish_for_uniform_pressure(rae):
    # at rest with uniform pressure the wall force is p * closed contour = 0
    m = rae
    s = OracleSolver(m, 0.5, 3.0)
    fs = s.freestream()
    W = np.tile(np.array([1.0, 0.0, 0.0, fs["PIS"] / 0.4]), (m.ncell, 1))
    s.set_state(W)
    cl, cd = s.forces()
    assert abs(cl) < 1e-12 and abs(cd) < 1e-12
    # and with a different uniform pressure, still zero
    s.set_state(np.tile(np.array([1.0, 0.0, 0.0, 2.0]), (m.ncell, 1)))
    cl, cd = s.forces()
    assert abs(cl) < 1e-12 and abs(cd) < 1e-12


def test_forces_rotate_with_alpha(rae):
    # the same state seen at angle alpha: (Cl, Cd) is the force rotated by alpha
    m = rae
    s0 = OracleSolver(m, 0.729, 0.0, cfl=1.0)
    s0.run(20)
    W = s0.state()
    s1 = OracleSolver(m, 0.729, 10.0, cfl=1.0)
    s1.set_state(W)
    cl0, cd0 = s0.forces()
    cl1, cd1 = s1.forces()
    # p_inf is the same (depends only on M); the force vector is (cd0, cl0)
    a = np.radians(10.0)
    np.testing.assert_allclose([cl1, cd1], [-cd0 * np.sin(a) + cl0 * np.cos(a),
                                            cd0 * np.cos(a) + cl0 * np.sin(a)], rtol=1e-10, atol=1e-12)


def test_openmp_build_gives_the_same_bits():
    # the CPU baseline's all-core build of the oracle runs the same arithmetic
    from oracle_ctypes import LIB_OMP_PATH, lib
    from gpueularsolver2d_amd import naca0012_mesh
    m = naca0012_mesh(3000, seed=7)
    a = OracleSolver(OracleMesh(m.x, m.y, m.tri), 0.8, 1.25)
    b = OracleSolver(OracleMesh(m.x, m.y, m.tri, library=LIB_OMP_PATH), 0.8, 1.25)
    ha, _ = a.run(30)
    hb, _ = b.run(30)
    assert lib(LIB_OMP_PATH).orc_threads() >= 1
    assert np.array_equal(a.state(), b.state())
    assert np.array_equal(ha, hb)
