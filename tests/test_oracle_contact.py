"""Pins of the oracle's contact stage (the collision step of SolidCloud::evolve,
src/solidcloud.cpp:477-519, with the geometry of src/libcollision/collision.cpp:7-32 and the
library's spring-dashpot law) against closed forms: bounding radii of the shape headers, the force
of a static overlap, antisymmetry, momentum conservation, the restitution coefficient of a linear
spring-dashpot and the static overlap of a resting sphere.  CPU only."""
import math

import numpy as np
import pytest

import oracle
from sdfibm_amd import solids as S
from test_oracle_dynamics import _solid


def _pair(c1, c2, r1=0.1, r2=0.2, v1=(0, 0, 0), v2=(0, 0, 0), w1=(0, 0, 0), w2=(0, 0, 0)):
    return np.concatenate([_solid("Sphere", center=c1, vel=v1, omega=w1, radius=r1),
                           _solid("Sphere", center=c2, vel=v2, omega=w2, radius=r2)])


def test_bounding_radii_of_the_shape_headers():
    assert oracle.bounding_radius(_solid("Sphere", radius=0.3)) == 0.3
    assert oracle.bounding_radius(_solid("Circle", radius=0.2)) == 0.2
    assert oracle.bounding_radius(_solid("Ellipsoid", radiusa=0.1, radiusb=0.4, radiusc=0.2)) == 0.4
    # the reference's box radius is the largest half-side, not the circumscribed radius (box.h:29)
    assert oracle.bounding_radius(_solid("Box", radiusa=0.1, radiusb=0.2, radiusc=0.3)) == 0.3
    assert oracle.bounding_radius(_solid("Rectangle", radiusa=0.1, radiusb=0.2)) == 0.2
    assert oracle.bounding_radius(_solid("Circle_Tail", radius=0.3, ratio=1.0, thickness=0.1)) == \
        pytest.approx(0.6, rel=1e-15)
    assert oracle.bounding_radius(_solid("Plane")) == 0.0


def test_static_overlap_force_and_antisymmetry():
    rng = np.random.default_rng(3)
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    c1 = np.array([0.3, -0.2, 0.1])
    sol = _pair(c1, c1 + 0.25 * u)  # overlap 0.1 + 0.2 - 0.25
    f = oracle.contact_forces(sol, 1e4)
    np.testing.assert_allclose(f[1], 1e4 * 0.05 * u, rtol=1e-12)
    np.testing.assert_array_equal(f[0], -f[1])  # bitwise
    # the spin of either sphere has no normal component at the contact point
    spun = _pair(c1, c1 + 0.25 * u, w1=(3, -1, 2), w2=(0.5, 4, -2))
    np.testing.assert_allclose(oracle.contact_forces(spun, 1e4), f, rtol=1e-12, atol=1e-12)
    # apart: no force
    assert not oracle.contact_forces(_pair(c1, c1 + 0.31 * u), 1e4).any()


def test_damping_term():
    """f = k delta - c v_n, v_n = (v2 - v1).n < 0 when approaching; no attraction."""
    c1 = np.zeros(3)
    n = np.array([1.0, 0, 0])
    sol = _pair(c1, c1 + 0.28 * n, v1=(0.5, 0, 0), v2=(-0.25, 0.3, 0))
    k, c = 1e3, 4.0
    f = oracle.contact_forces(sol, k, c)
    np.testing.assert_allclose(f[1], (k * 0.02 - c * (-0.75)) * n, rtol=1e-12)
    sep = _pair(c1, c1 + 0.28 * n, v1=(-5.0, 0, 0), v2=(5.0, 0, 0))  # fast separation: would attract
    assert not oracle.contact_forces(sep, k, c).any()


@pytest.mark.parametrize("deg", [0.0, 30.0, -90.0])
def test_plane_force_along_the_rotated_normal(deg):
    th = math.radians(deg)
    q = (math.cos(th / 2), 0.0, 0.0, math.sin(th / 2))  # rotation about z
    nrm = np.array([-math.sin(th), math.cos(th), 0.0])  # the local +y axis
    p0 = np.array([0.2, 0.1, 0.0])
    s = _solid("Sphere", center=p0 + 0.07 * nrm + 0.3 * np.array([nrm[1], -nrm[0], 0]), radius=0.1)
    sol = np.concatenate([_solid("Plane", center=p0, q=q), s])
    f = oracle.contact_forces(sol, 2e3)
    assert not f[0].any()  # nothing acts on the plane
    np.testing.assert_allclose(f[1], 2e3 * 0.03 * nrm, rtol=1e-12, atol=1e-13)


def test_zero_law_is_plain_evolve():
    rng = np.random.default_rng(5)
    sol = np.concatenate([_solid("Sphere", center=rng.uniform(0, 0.3, 3), vel=rng.normal(size=3),
                                 omega=rng.normal(size=3), radius=0.1) for _ in range(6)])
    dyn = oracle.make_dyn(6, rho=2.0)
    ft = rng.normal(size=(6, 6))
    g = np.array([0.0, -9.81, 0.0])
    a, da = oracle.evolve(sol, dyn, ft, 0.0, 1e-3, g, 1.0)
    b, db = oracle.evolve_contact(sol, dyn, ft, 0.0, 1e-3, g, 1.0, 0.0, 0.0)
    for key in ("center", "q", "vel", "omega"):
        np.testing.assert_array_equal(a[key], b[key])
    np.testing.assert_array_equal(da.view(np.uint8), db.view(np.uint8))


def _run(sol, dyn, steps, dt, k, c, g=(0.0, 0.0, 0.0)):
    ft = np.zeros((sol.size, 6))
    for _ in range(steps):
        sol, dyn = oracle.evolve_contact(sol, dyn, ft, 0.0, dt, np.array(g, float), 1.0, k, c)
    return sol, dyn


def test_head_on_collision_conserves_momentum_and_energy():
    """Equal spheres, no damping: the velocities are exchanged, to an error of first order in the
    time step (forces are taken at the start of each sub-step), and the momentum is conserved to
    rounding."""
    r = 0.1
    m = 1000.0 * 4 / 3 * math.pi * r ** 3
    k = 1e4  # contact time pi sqrt(m_eff / k) ~ 0.046
    err = []
    for dt, steps in ((1e-3, 400), (5e-4, 800)):
        sol = np.concatenate([_solid("Sphere", center=(-0.15, 0, 0), vel=(1.0, 0, 0), radius=r),
                              _solid("Sphere", center=(0.15, 0.0, 0), vel=(-0.5, 0, 0), radius=r)])
        sol, _ = _run(sol, oracle.make_dyn(2, rho=1000.0), steps, dt, k, 0.0)
        p = m * sol["vel"].sum(axis=0)
        np.testing.assert_allclose(p, [m * 0.5, 0, 0], rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(sol["vel"][:, 0], [-0.5, 1.0], rtol=1e-2)
        err.append(abs(sol["vel"][1, 0] - 1.0))
    assert err[1] < 0.6 * err[0]


def test_restitution_of_the_spring_dashpot():
    """Sphere hitting a plane at speed v0, no gravity: e = v_out / v0 for a linear spring-dashpot
    is exp(-pi zeta / sqrt(1 - zeta^2)), zeta = c / (2 sqrt(k m)).  The no-attraction clamp ends
    the contact when the force reaches zero, slightly before the overlap does, which raises e
    by O(zeta^2)."""
    r, rho = 0.05, 2000.0
    m = rho * 4 / 3 * math.pi * r ** 3
    k = 1e3
    for zeta in (0.0, 0.05, 0.2):
        c = 2 * zeta * math.sqrt(k * m)
        sol = np.concatenate([_solid("Plane", center=(0, 0, 0)),
                              _solid("Sphere", center=(0, r + 0.01, 0), vel=(0, -0.5, 0), radius=r)])
        dyn = oracle.make_dyn(2, rho=rho)
        dyn["mask"][0] = 0
        sol, _ = _run(sol, dyn, 300, 1e-3, k, c)
        assert sol["center"][1, 1] > r  # left the wall
        e = sol["vel"][1, 1] / 0.5
        e_lin = math.exp(-math.pi * zeta / math.sqrt(1 - zeta * zeta))
        assert e_lin - 2e-3 < e < e_lin + 2 * zeta * zeta + 2e-3, (zeta, e, e_lin)


def test_resting_sphere_static_overlap():
    """Under gravity with damping a sphere settles on a plane at overlap m g' / k."""
    r, rho = 0.05, 3000.0
    m = rho * 4 / 3 * math.pi * r ** 3
    k = 500.0
    c = 2 * 0.7 * math.sqrt(k * m)
    sol = np.concatenate([_solid("Plane", center=(0, 0, 0)), _solid("Sphere", center=(0, r, 0), radius=r)])
    dyn = oracle.make_dyn(2, rho=rho)
    dyn["mask"][0] = 0
    g = (0.0, -9.81, 0.0)
    sol, _ = _run(sol, dyn, 2000, 1e-3, k, c, g)
    gp = (rho - 1.0) / rho * 9.81
    assert r - sol["center"][1, 1] == pytest.approx(m * gp / k, rel=1e-6)
    assert abs(sol["vel"][1, 1]) < 1e-9
