"""Rigid-body update on the device (sdfibm_evolve: SolidCloud::evolve, src/solidcloud.cpp:521-562)
and the coupled step loop of the reference's solver (src/main.cpp:66-84 without the PISO solve:
interact -> U -= Fs dt -> evolve -> fixInternal), against the CPU oracle.

Tolerances: evolve from identical inputs runs the same FP64 operations on both sides (no FMA
contraction): state within 1e-14 relative.  The coupled loop feeds each side its own fluid force
(reduction order differs at 1e-16): state within 1e-9 relative, U within 1e-9 of max|U|."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from sdfibm_amd import scenes
from sdfibm_amd import solids as S
from sdfibm_amd.cloud import SolidCloud, motion_mask

pytestmark = pytest.mark.gpu


def _state(sol):
    return np.concatenate([sol["center"], sol["q"], sol["vel"], sol["omega"]], axis=1)


def test_evolve_matches_oracle_from_same_forces():
    sc = scenes.small_mixed_3d(20, seed=3)
    ns = sc.solids.size
    rng = np.random.default_rng(4)
    rho = rng.uniform(1.1, 3.0, ns)
    masks = np.full(ns, -1, dtype=np.int32)
    masks[1] = motion_mask("b110001")
    masks[2] = motion_mask("b000000")
    g = np.array([0.3, -9.81, 0.2])
    with SolidCloud(sc.mesh, rho_f=sc.rho_f, precision="fp64") as c:
        c.set_solids(sc.solids)
        c.set_velocity(sc.U)
        c.set_dynamics(rho, masks)
        dyn = oracle.make_dyn(ns, rho, masks)
        sol = sc.solids.copy()
        for step in range(3):
            ft = c.interact(step * sc.dt, sc.dt)
            c.evolve(step * sc.dt, sc.dt, g)
            sol, dyn = oracle.evolve(sol, dyn, ft, step * sc.dt, sc.dt, g, sc.rho_f)
            got = c.solids()
            np.testing.assert_allclose(_state(got), _state(sol), rtol=1e-14, atol=1e-15)
        # masked components stay zero (the fully masked solid still drifts by its initial
        # velocity in the first sub-step: the position update averages old and new velocity)
        assert got["vel"][1][2] == 0 and got["omega"][1][0] == 0 and got["omega"][1][1] == 0
        assert np.all(got["vel"][2] == 0) and np.all(got["omega"][2] == 0)


def test_single_solid_uses_one_substep():
    """One solid: one sub-step per evolve (solidcloud.cpp:524-526); free fall is then exact."""
    sc = scenes.small_mixed_3d(16, seed=5)
    one = sc.solids[:1].copy()
    g = np.array([0.0, 0.0, -10.0])
    with SolidCloud(sc.mesh, rho_f=0.0, precision="fp64") as c:
        c.set_solids(one)
        c.set_dynamics(2.0)
        v0 = one["vel"][0].copy()
        for _ in range(5):
            c.evolve(0.0, 1e-3, g)  # no interact: no fluid force (rho_f = 0 also zeroes it)
        got = c.solids()
    np.testing.assert_allclose(got["vel"][0], v0 + 5e-3 * g, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("prec", ["fp64"])
def test_coupled_loop_sedimentation(prec):
    """16 discs + 4 walls on a 100 x 200 mesh, three full coupling steps, GPU vs oracle."""
    sc = scenes.sedimentation_2d(100, 200, 16)
    ns = sc.solids.size
    rho = np.full(ns, 1.5)
    masks = np.full(ns, motion_mask("b110001"), dtype=np.int32)
    masks[16:] = motion_mask("b000000")  # walls
    g = np.array(sc.info["gravity"])
    dt = 2e-3
    rng = np.random.default_rng(9)
    U0 = np.zeros((sc.mesh.n_cells, 3))
    U0[:, :2] = rng.normal(0, 0.05, (sc.mesh.n_cells, 2))
    om = oracle.OracleMesh(sc.mesh.to_polymesh())
    sol = sc.solids.copy()
    dyn = oracle.make_dyn(ns, rho, masks)
    U = U0.copy()
    with SolidCloud(sc.mesh, two_d=True, rho_f=1.0, precision=prec) as c:
        c.set_solids(sc.solids)
        c.set_velocity(U0)
        c.set_dynamics(rho, masks)
        for step in range(3):
            t = step * dt
            ft = c.interact(t, dt)
            c.update_velocity(dt)
            c.evolve(t, dt, g)
            c.fix_internal()
            r = om.interact(sol, U, two_d=True, rho_f=1.0, dt=dt)
            np.testing.assert_array_equal(c.fields()["Ct"], r["Ct"])
            scale = np.maximum(np.abs(r["ft"]), 1e-4 * np.abs(r["ft"]).max())
            assert np.max(np.abs(ft - r["ft"]) / scale) <= 1e-9
            U = om.update_velocity(U, r["Fs"], dt)
            sol, dyn = oracle.evolve(sol, dyn, r["ft"], t, dt, g, 1.0)
            U = om.fix_internal(U, r["Ct"], sol)
            got = c.solids()
            np.testing.assert_allclose(_state(got), _state(sol), rtol=1e-9, atol=1e-12)
            Ug = c.velocity()
            assert np.max(np.abs(Ug - U)) <= 1e-9 * max(np.abs(U).max(), 1.0)
        # the discs sank, the walls stayed
        assert np.all(got["center"][:16, 1] < sc.solids["center"][:16, 1])
        np.testing.assert_array_equal(got["center"][16:], sc.solids["center"][16:])


def test_coupled_loop_3d_jittered_graph():
    """The 3-D coupled loop as replays of the captured graph (sdfibm_run_steps) on a jittered
    24^3 mesh (non-planar faces): spheres, ellipsoids and boxes between four fixed planes, with
    gravity, four steps against the oracle's interact -> U -= Fs dt -> evolve -> fixInternal.
    Tolerances of the 2-D loop above (each side feeds its own FP64 force sums forward)."""
    from test_gpu_graph import _walls
    sc = scenes.small_mixed_3d(24, seed=6, n_each=2, jitter=0.2)
    keep = np.isin(sc.solids["shape"], [S.SPHERE, S.ELLIPSOID, S.BOX])
    sol0 = np.concatenate([sc.solids[keep], _walls(0.06)])
    ns = sol0.size
    rho = np.linspace(1.4, 2.6, ns)
    masks = np.full(ns, -1, dtype=np.int32)
    masks[-4:] = motion_mask("b000000")
    sol0["vel"][-4:] = 0.0
    sol0["omega"][-4:] = 0.0
    g = np.array([0.0, -9.81, 0.0])
    dt = 1e-3
    n = 4
    om = oracle.OracleMesh(sc.mesh.to_polymesh())
    sol = sol0.copy()
    dyn = oracle.make_dyn(ns, rho, masks)
    U = sc.U.copy()
    for step in range(n):
        t = step * dt
        r = om.interact(sol, U, rho_f=sc.rho_f, dt=dt)
        U = om.update_velocity(U, r["Fs"], dt)
        sol, dyn = oracle.evolve(sol, dyn, r["ft"], t, dt, g, sc.rho_f)
        U = om.fix_internal(U, r["Ct"], sol)
    with SolidCloud(sc.mesh, rho_f=sc.rho_f, precision="fp64") as c:
        c.set_solids(sol0)
        c.set_velocity(sc.U)
        c.set_dynamics(rho, masks)
        ft = c.run_steps(0.0, dt, n, g)
        got = c.solids()
        Ug = c.velocity()
        Ct = c.fields()["Ct"]
    np.testing.assert_array_equal(Ct, r["Ct"])
    scale = np.maximum(np.abs(r["ft"]), 1e-4 * np.abs(r["ft"]).max())
    assert np.max(np.abs(ft - r["ft"]) / scale) <= 1e-9
    np.testing.assert_allclose(_state(got), _state(sol), rtol=1e-9, atol=1e-12)
    assert np.max(np.abs(Ug - U)) <= 1e-9 * max(np.abs(U).max(), 1.0)
    assert np.any(got["center"][:-4] != sol0["center"][:-4])
    np.testing.assert_array_equal(got["center"][-4:], sol0["center"][-4:])
