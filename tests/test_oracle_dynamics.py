"""Pins of the oracle's rigid-body update (SolidCloud::evolve, src/solidcloud.cpp:521-562;
Solid::move, src/solid.h:182-202) against closed forms.  CPU only."""
import math

import numpy as np
import pytest

import oracle
from sdfibm_amd import solids as S


def _solid(type_name, center=(0.0, 0.0, 0.0), q=(1.0, 0.0, 0.0, 0.0), vel=(0, 0, 0), omega=(0, 0, 0), **prm):
    s = S.make_solids(1)
    shape, p = S.make_shape(type_name, **prm)
    s["shape"] = shape
    s["prm"][0] = p
    s["center"][0] = center
    s["q"][0] = q
    s["vel"][0] = vel
    s["omega"][0] = omega
    return s


def test_mass_properties_closed_forms():
    r, a, b, c = 0.3, 0.5, 0.2, 0.1
    v, moi = oracle.mass_properties(_solid("Sphere", radius=r)[0])
    assert v == pytest.approx(4 / 3 * math.pi * r ** 3, rel=1e-15)
    assert moi == pytest.approx([0.4 * v * r * r] * 3, rel=1e-15)
    v, moi = oracle.mass_properties(_solid("Ellipsoid", radiusa=a, radiusb=b, radiusc=c)[0])
    assert v == pytest.approx(4 / 3 * math.pi * a * b * c, rel=1e-15)
    assert moi == pytest.approx([0.2 * v * (b * b + c * c), 0.2 * v * (a * a + c * c), 0.2 * v * (a * a + b * b)],
                                rel=1e-15)
    v, moi = oracle.mass_properties(_solid("Box", radiusa=a, radiusb=b, radiusc=c)[0])
    assert v == pytest.approx(8 * a * b * c, rel=1e-15)
    assert moi[0] == pytest.approx(v * (b * b + c * c) / 3, rel=1e-15)
    v, moi = oracle.mass_properties(_solid("Circle", radius=r)[0])
    assert v == pytest.approx(math.pi * r * r, rel=1e-15) and moi[2] == pytest.approx(0.5 * v * r * r, rel=1e-15)
    v, moi = oracle.mass_properties(_solid("Plane")[0])
    assert v == 0.0 and list(moi) == [1.0, 1.0, 1.0]


@pytest.mark.parametrize("n", [1, 3])
def test_free_fall_is_exact(n):
    """No fluid force: v = g' t and x = x0 + g' t^2 / 2 with g' = (rho_s - rho_f)/rho_s g (the
    trapezoidal position update is exact for a linear velocity); one sub-step for a single solid,
    twenty otherwise."""
    sol = np.concatenate([_solid("Sphere", center=(0.1 * k, 0.5, 0.0), radius=0.05) for k in range(n)])
    dyn = oracle.make_dyn(n, rho=3.0)
    g = np.array([0.0, -9.81, 0.0])
    dt, steps = 1e-3, 25
    ft = np.zeros((n, 6))
    for _ in range(steps):
        sol, dyn = oracle.evolve(sol, dyn, ft, 0.0, dt, g, 1.0)
    t = dt * steps
    gp = (3.0 - 1.0) / 3.0 * g
    np.testing.assert_allclose(sol["vel"], np.tile(gp * t, (n, 1)), rtol=1e-12, atol=1e-15)
    expect = np.array([[0.1 * k, 0.5, 0.0] for k in range(n)]) + 0.5 * gp * t * t
    np.testing.assert_allclose(sol["center"], expect, rtol=1e-12, atol=1e-15)
    np.testing.assert_array_equal(sol["q"], np.tile([1.0, 0, 0, 0], (n, 1)))


def test_ab2_fluid_force_mix():
    """First evolve: old force := current, so the solid feels F; later steps feel 1.5 F - 0.5 F_old."""
    sol = _solid("Sphere", radius=0.1)
    vol, _ = oracle.mass_properties(sol[0])
    rho = 2.0
    dyn = oracle.make_dyn(1, rho=rho)
    dt = 1e-2
    F1 = np.array([[1.0, -2.0, 0.5, 0, 0, 0]])
    F2 = np.array([[-3.0, 1.0, 2.0, 0, 0, 0]])
    s1, d1 = oracle.evolve(sol, dyn, F1, 0.0, dt, np.zeros(3), rho)  # rho_f = rho_s: no gravity
    np.testing.assert_allclose(s1["vel"][0], F1[0, :3] / (vol * rho) * dt, rtol=1e-13)
    s2, d2 = oracle.evolve(s1, d1, F2, dt, dt, np.zeros(3), rho)
    dv = (1.5 * F2[0, :3] - 0.5 * F1[0, :3]) / (vol * rho) * dt
    np.testing.assert_allclose(s2["vel"][0] - s1["vel"][0], dv, rtol=1e-12, atol=1e-15)
    np.testing.assert_array_equal(d2["ff_old"][0], F2[0, :3])
    assert d2["first_fluid"][0] == 0


def test_spin_quaternion_closed_form():
    """Constant omega about z, no torque: each step maps the half angle h to h + atan(w dt / 2)
    (q += dt/2 (0, w) q, then normalise), so after n steps q = (cos(n atan(w dt/2)), 0, 0, sin(..))."""
    w, dt, n = 3.0, 1e-2, 40
    sol = _solid("Sphere", omega=(0.0, 0.0, w), radius=0.1)
    dyn = oracle.make_dyn(1, rho=1.0)
    for _ in range(n):
        sol, dyn = oracle.evolve(sol, dyn, np.zeros((1, 6)), 0.0, dt, np.zeros(3), 1.0)
    h = n * math.atan(w * dt / 2)
    np.testing.assert_allclose(sol["q"][0], [math.cos(h), 0, 0, math.sin(h)], rtol=0, atol=1e-13)
    np.testing.assert_allclose(sol["omega"][0], [0, 0, w], rtol=1e-15)


def test_torque_on_rotated_ellipsoid_principal_axis():
    """A torque along a body principal axis (world direction R e_k) accelerates the spin along that
    axis at T / I_k; pins R . inv(I) . R^T (a transposed or missing rotation fails)."""
    q = S.quaternion_xyz(np.radians([30.0, -50.0, 70.0]))
    sol = _solid("Ellipsoid", q=q, radiusa=0.3, radiusb=0.2, radiusc=0.1)
    rho = 1.7
    vol, moi = oracle.mass_properties(sol[0])
    w, x, y, z = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    dt = 1e-3
    for k in range(3):
        axis = R[:, k]
        ft = np.zeros((1, 6))
        ft[0, 3:] = 2.0 * axis
        s1, _ = oracle.evolve(sol, oracle.make_dyn(1, rho=rho), ft, 0.0, dt, np.zeros(3), rho)
        np.testing.assert_allclose(s1["omega"][0], 2.0 / (moi[k] * rho) * dt * axis, rtol=1e-12, atol=1e-14)


def test_motion_mask_and_plane():
    """Motion01Mask b110000: only vx, vy free; a masked wall (b000000) never moves."""
    sol = np.concatenate([_solid("Sphere", radius=0.1, omega=(1.0, 2.0, 3.0)), _solid("Plane")])
    dyn = oracle.make_dyn(2, rho=2.0)
    dyn["mask"] = [0b000011, 0]
    ft = np.array([[1.0, 1.0, 1.0, 1.0, 1.0, 1.0], [5.0, 5.0, 5.0, 5.0, 5.0, 5.0]])
    g = np.array([0.0, 0.0, -10.0])
    s1, _ = oracle.evolve(sol, dyn, ft, 0.0, 1e-2, g, 1.0)
    assert s1["vel"][0, 2] == 0.0 and list(s1["omega"][0]) == [0.0, 0.0, 0.0]
    assert s1["vel"][0, 0] > 0 and s1["vel"][0, 1] > 0
    np.testing.assert_array_equal(s1[1]["center"], sol[1]["center"])
    np.testing.assert_array_equal(s1[1]["vel"], [0, 0, 0])
    np.testing.assert_array_equal(s1[1]["q"], sol[1]["q"])
