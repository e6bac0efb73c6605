"""Known-answer tests of the oracle's point kernels against closed forms and
invariants that a plausible mistake (dropped term, wrong sign or index,
swapped operand) would break."""
import math

import numpy as np
import pytest

from oracle_ctypes import OracleMesh, OracleSolver, dptr, lib

G = 1.4


def cons(rho, u, v, p, g=G):
    return np.array([rho, rho * u, rho * v, p / (g - 1) + 0.5 * rho * (u * u + v * v)])


def phys_flux(U, nx, ny, g=G):
    rho, u, v = U[0], U[1] / U[0], U[2] / U[0]
    p = (g - 1) * (U[3] - 0.5 * rho * (u * u + v * v))
    vn = u * nx + v * ny
    return np.array([rho * vn, rho * u * vn + p * nx, rho * v * vn + p * ny, (U[3] + p) * vn])


def roe(UL, UR, nx, ny, g=G):
    out = np.zeros(4)
    lib().orc_roe(dptr(np.asarray(UL, float)), dptr(np.asarray(UR, float)), nx, ny, g, dptr(out))
    return out


def test_pta_freestream_closed_form():
    # non-dimensional freestream: p = 1/(gamma M^2), a = 1/M
    for M in (0.3, 0.5, 0.8, 2.0):
        rhoE = 1.0 / (G * (G - 1) * M * M) + 0.5
        out = np.zeros(3)
        lib().orc_pta(G, 287.06, dptr(np.array([1.0, 1.0, 0.0, rhoE])), dptr(out))
        assert out[0] == pytest.approx(1.0 / (G * M * M), rel=1e-14)
        assert out[2] == pytest.approx(1.0 / M, rel=1e-14)


@pytest.mark.parametrize("state", [(1.0, 0.3, -0.2, 0.9), (0.7, 1.5, 0.4, 0.4), (1.2, -0.1, 0.05, 2.0)])
@pytest.mark.parametrize("theta", [0.0, 0.7, 2.5, 4.0])
def test_roe_consistency(state, theta):
    # F(U, U, n) = F(U) . n  (dissipation vanishes)
    U = cons(*state)
    nx, ny = math.cos(theta), math.sin(theta)
    np.testing.assert_allclose(roe(U, U, nx, ny), phys_flux(U, nx, ny), rtol=1e-13, atol=1e-13)


def test_roe_supersonic_upwind():
    # all three eigenvalues positive and above the Harten threshold -> F = F(U_L)
    UL, UR = cons(1.0, 3.0, 0.1, 0.5), cons(0.9, 2.8, 0.0, 0.45)
    np.testing.assert_allclose(roe(UL, UR, 1.0, 0.0), phys_flux(UL, 1.0, 0.0), rtol=1e-12, atol=1e-12)
    # all negative -> F = F(U_R)
    UL, UR = cons(1.0, -3.0, 0.1, 0.5), cons(0.9, -2.8, 0.0, 0.45)
    np.testing.assert_allclose(roe(UL, UR, 1.0, 0.0), phys_flux(UR, 1.0, 0.0), rtol=1e-12, atol=1e-12)


def test_roe_antisymmetry_and_rotation():
    rng = np.random.default_rng(0)
    for _ in range(50):
        UL = cons(rng.uniform(0.5, 1.5), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.5, 2))
        UR = cons(rng.uniform(0.5, 1.5), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.5, 2))
        th = rng.uniform(0, 2 * math.pi)
        nx, ny = math.cos(th), math.sin(th)
        F = roe(UL, UR, nx, ny)
        np.testing.assert_allclose(roe(UR, UL, -nx, -ny), -F, rtol=1e-12, atol=1e-12)
        # rotate both states and the normal by phi: mass/energy invariant, momentum rotates
        ph = rng.uniform(0, 2 * math.pi)
        c, s = math.cos(ph), math.sin(ph)
        rot = lambda U: np.array([U[0], c * U[1] - s * U[2], s * U[1] + c * U[2], U[3]])
        Fr = roe(rot(UL), rot(UR), c * nx - s * ny, s * nx + c * ny)
        np.testing.assert_allclose(Fr, rot(F), rtol=1e-11, atol=1e-12)


def test_roe_harten_fix_branch():
    # a stagnation state: V = 0 so |lambda_2| uses the entropy fix (0 + d^2)/(2 d)
    U = cons(1.0, 0.0, 0.0, 1.0)
    F = roe(U, U, 1.0, 0.0)
    np.testing.assert_allclose(F, [0.0, 1.0, 0.0, 0.0], atol=1e-15)


@pytest.fixture(scope="module")
def solver():
    # a mesh is needed only to hold the freestream constants of the solver
    x = np.array([0.0, 1.0, 0.0]); y = np.array([0.0, 0.0, 1.0])
    m = OracleMesh(x, y, np.array([[1, 2, 3]]))
    return OracleSolver(m, 0.5, 1.25)


def far(s, U, vx, vy):
    out = np.zeros(4)
    lib().orc_far_ghost(s.h, dptr(np.asarray(U, float)), vx, vy, dptr(out))
    return out


def test_far_ghost_freestream_identity(solver):
    fs = solver.freestream()
    U = np.array([1.0, fs["UI"], fs["VI"], fs["rhoE"]])
    for th in np.linspace(0, 2 * math.pi, 13):
        g = far(solver, U, math.cos(th), math.sin(th))
        # both branches (inflow / outflow) return the freestream
        assert g[0] == 1.0 and g[1] == fs["UI"] and g[2] == fs["VI"]
        assert g[3] == pytest.approx(fs["rhoE"], rel=1e-15)


def test_far_ghost_outflow_pressure(solver):
    # subsonic outflow imposes p = p_inf and keeps the outgoing characteristic
    fs = solver.freestream()
    U = cons(1.1, 0.6, 0.1, 3.2)
    g = far(solver, U, 0.0, 1.0)  # n = (vy, vx)/|.| = (1, 0): outflow since u > 0
    p = (G - 1) * (g[3] - 0.5 * (g[1] ** 2 + g[2] ** 2) / g[0])
    assert p == pytest.approx(fs["PIS"], rel=1e-13)
    assert g[2] / g[0] == pytest.approx(0.1, rel=1e-13)  # tangential velocity kept


def test_far_ghost_supersonic_branches():
    # |u_inf . n| / AIN > 1 with the dimensional AIN = sqrt(gamma R TIN)
    # (gpu_solver.cu:79): TIN = 0.002 K gives AIN = 0.897, so an edge whose
    # normal is within ~26 degrees of the flow direction is supersonic.
    # Inflow (the cell's normal speed <= 0) takes the freestream state,
    # outflow the cell's own state, whatever the cell holds (:81-93).
    x = np.array([0.0, 1.0, 0.0]); y = np.array([0.0, 0.0, 1.0])
    s = OracleSolver(OracleMesh(x, y, np.array([[1, 2, 3]])), 0.5, 0.0, t_in=0.002)
    fs = s.freestream()
    assert fs["AIN"] == pytest.approx(math.sqrt(G * 287.06 * 0.002), rel=1e-15) and fs["AIN"] < 1
    rhoE_inf = 287.06 * 0.002 / fs["VIN"] ** 2 / (G - 1) + 0.5
    U_out = cons(1.3, 0.4, -0.2, 2.1)   # moves along +x
    U_in = cons(0.8, -0.3, 0.1, 1.7)    # moves along -x
    # n = (vy, vx) / |v| = (1, 0): u_inf . n = 1 > AIN
    g = far(s, U_out, 0.0, 2.0)
    assert np.array_equal(g, U_out)
    g = far(s, U_in, 0.0, 2.0)
    assert g[0] == 1.0 and g[1] == fs["UI"] and g[2] == fs["VI"] and g[3] == rhoE_inf
    # n = (-1, 0) (an inflow edge for the freestream): the cell's own motion decides
    assert np.array_equal(far(s, U_in, 0.0, -2.0), U_in)
    g = far(s, U_out, 0.0, -2.0)
    assert g[0] == 1.0 and g[1] == fs["UI"] and g[3] == rhoE_inf
    # a normal at 60 degrees to the flow is subsonic there (0.5 < AIN): the
    # characteristic branch, which does not return the cell state unchanged
    g = far(s, U_out, math.sin(math.radians(60)), math.cos(math.radians(60)))
    assert not np.array_equal(g, U_out)


def test_wall_ghost_mirror_and_zero_mass_flux():
    U = cons(1.0, 0.7, 0.2, 1.5)
    for th in (0.3, 1.9, 4.4):
        vx, vy = 2.0 * math.sin(th), 2.0 * math.cos(th)  # n = (vy, vx)/2 = (cos, sin)
        out = np.zeros(4)
        lib().orc_wall_ghost(dptr(U), vx, vy, 2.0, dptr(out))
        nx, ny = math.cos(th), math.sin(th)
        un, ung = (U[1] * nx + U[2] * ny), (out[1] * nx + out[2] * ny)
        assert ung == pytest.approx(-un, abs=1e-14)
        assert out[0] == U[0] and out[3] == U[3]
        F = roe(U, out, nx, ny)
        assert abs(F[0]) < 1e-14 and abs(F[3]) < 1e-14


def lim(u, umax, umin, d):
    d = np.array(d, float)
    lib().orc_limiter(dptr(np.array(u, float)), dptr(np.array(umax, float)),
                      dptr(np.array(umin, float)), dptr(d))
    return d


def test_limiter_cases():
    u = [1.0, 1.0, 1.0, 1.0]
    umax = [1.2, 1.0, 1.5, 1.1]
    umin = [0.9, 0.8, 1.0, 0.95]
    d = lim(u, umax, umin, [0.1, 0.3, -0.2, 0.0])
    assert d[0] == 0.1               # inside the bounds: unchanged
    assert d[1] == 0.0               # local maximum (Umax == U): limited to zero
    assert d[2] == 0.0               # local minimum (Umin == U): limited to zero
    assert d[3] == 0.0               # zero slope stays zero
    d = lim(u, umax, umin, [0.4, -0.4, 0.7, -0.1])
    np.testing.assert_allclose(d, [0.2, -0.2, 0.5, -0.05], rtol=1e-15)


def test_spectral_radius_uniform_state():
    U = cons(1.3, 0.4, -0.3, 1.1)
    vx, vy = 0.3, -0.4
    sr = lib().orc_spectral_radius(dptr(U), dptr(U), vx, vy, G)
    u, v = 0.4, -0.3
    a = math.sqrt(G * 1.1 / 1.3)
    assert sr == pytest.approx(abs(u * vy + v * vx) + a * math.hypot(vx, vy), rel=1e-14)
