"""The general polyMesh path (sdfibm_create_polymesh) on the GPU (needs an MI355X).

* the reference's own multi-block case: examples/flow_past_cylinder/re200, its nine graded
  blocks regenerated from the blockMeshDict numbers, reproduces all 78,400 values of 0/As to
  1e-9 but the one documented cell, and the oracle on the same mesh to 1e-12;
* tool_vof/example's mesh (its topology pinned by the reference's polyMesh digests) through the
  polyMesh path reproduces alpha.water to 1e-11;
* a 3-D O-grid (five blocks with arc edges, graded, 8 layers) with spheres, ellipsoids, boxes, a
  CSG solid and a plane, both precisions, against the oracle on the same polyMesh;
* one jittered block through both paths (structured and polyMesh): the same candidates and fields,
  with and without plane walls (the polyMesh path against the oracle too);
* three coupled steps (interact, U -= Fs dt, evolve, fixInternal) against the oracle;
* the calls a polyMesh context refuses.
Tolerances of test_gpu_parity.py."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from _helpers import CYLINDER_ODD_CELL, cylinder_re200, split_candidates, toolvof_mesh, toolvof_solids
from sdfibm_amd import blockmesh, scenes
from sdfibm_amd import mesh as M
from sdfibm_amd import solids as S
from sdfibm_amd._lib import SdfibmError
from sdfibm_amd.cloud import SolidCloud, motion_mask
from test_gpu_parity import TOL

pytestmark = pytest.mark.gpu


def _compare(c, ref, ns, tol, ft=None, exact_lists=True):
    got = c.fields()
    parts = split_candidates(ref, ns)
    for s in range(ns):
        counts, cells, alpha = c.candidates(s)
        rc, rcells, ralpha = parts[s]
        assert np.array_equal(counts, rc), (s, counts, rc)
        if exact_lists:
            assert np.array_equal(cells, rcells), s
        assert np.max(np.abs(alpha - ralpha), initial=0.0) <= tol["a"], s
    assert np.array_equal(got["Ct"], ref["Ct"])
    assert np.max(np.abs(got["As"] - ref["As"])) <= tol["a"]
    assert np.max(np.abs(got["Fs"] - ref["Fs"])) <= tol["f"] * max(np.abs(ref["Fs"]).max(), 1e-300)
    if ft is not None:
        scale = np.maximum(np.abs(ref["ft"]), max(tol["floor"] * np.abs(ref["ft"]).max(), 1e-300))
        assert np.max(np.abs(ft - ref["ft"]) / scale) <= tol["ft"]
    return got


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_cylinder_nine_blocks(prec):
    pm, info, sol, As_file = cylinder_re200()
    U = np.random.default_rng(2).normal(size=(pm.n_cells, 3))
    U[:, 2] = 0.0
    ref = oracle.OracleMesh(pm).interact(sol, U, two_d=True, rho_f=1.0, dt=1.0, capture=True,
                                         capacity=4 * pm.n_cells)
    with SolidCloud(pm, two_d=True, rho_f=1.0, precision=prec) as c:
        c.set_solids(sol)
        c.set_velocity(U)
        ft = c.interact(0.0, 1.0)
        got = _compare(c, ref, 1, TOL[prec], ft)
    if prec == "fp64":
        bad = np.nonzero(np.abs(got["As"] - As_file) > 1e-9)[0]
        assert list(bad) == [CYLINDER_ODD_CELL]


def test_toolvof_alpha_water_through_polymesh():
    m, g = toolvof_mesh()
    pm = m.to_polymesh()
    sol = toolvof_solids()
    with SolidCloud(pm, two_d=True, rho_f=0.0, precision="fp64") as c:
        c.set_solids(sol)
        c.set_velocity(np.zeros((pm.n_cells, 3)))
        c.interact(0.0, 1.0)
        alpha = c.fields()["As"]
    assert np.max(np.abs(alpha - g["alpha"])) <= 1e-11


def _ogrid(n=16, nz=8, r_in=0.5, r_out=1.0, h=0.6):
    """Five blocks of an O-grid about the z axis (a square core, four ring blocks with arc
    edges), graded towards the core, extruded over nz layers."""
    a = r_in / np.sqrt(2.0)
    verts = []
    for z in (0.0, h):
        verts += [(a, -a, z), (-a, -a, z), (-a, a, z), (a, a, z)]  # core corners 0..3 (+8)
        b = r_out / np.sqrt(2.0)
        verts += [(b, -b, z), (-b, -b, z), (-b, b, z), (b, b, z)]  # outer corners 4..7 (+8)
    # ring blocks: local x from the core outwards (graded), local y along the core's edge
    blocks = [([1, 0, 3, 2, 9, 8, 11, 10], (n, n, nz), (1.0, 1.0, 1.0)),
              ([0, 4, 7, 3, 8, 12, 15, 11], (n // 2, n, nz), (2.0, 1.0, 1.0)),
              ([1, 5, 4, 0, 9, 13, 12, 8], (n // 2, n, nz), (2.0, 1.0, 1.0)),
              ([2, 6, 5, 1, 10, 14, 13, 9], (n // 2, n, nz), (2.0, 1.0, 1.0)),
              ([3, 7, 6, 2, 11, 15, 14, 10], (n // 2, n, nz), (2.0, 1.0, 1.0))]
    arcs = {}
    for zo, z in ((0, 0.0), (8, h)):
        for p, q, ang in ((4, 7, 0.0), (7, 6, 90.0), (6, 5, 180.0), (5, 4, 270.0)):
            t = np.radians(ang)
            arcs[(p + zo, q + zo)] = np.array([r_out * np.cos(t), r_out * np.sin(t), z])
    return blockmesh.block_mesh(blockmesh.BlockMesh(np.array(verts), blocks, arcs, []))[0]


def _ogrid_solids(rng, table):
    sol = S.make_solids(6)
    kinds = [("Sphere", dict(radius=0.12, com=(0.01, 0.0, 0.0))),
             ("Ellipsoid", dict(radiusa=0.18, radiusb=0.09, radiusc=0.1)),
             ("Box", dict(radiusa=0.08, radiusb=0.12, radiusc=0.07, com=(0.0, 0.01, 0.0))),
             ("Sphere", dict(radius=0.09))]
    for t, (name, prm) in enumerate(kinds):
        sid, p = S.make_shape(name, **prm)
        sol[t]["shape"] = sid
        sol[t]["prm"] = p
    sol["center"][:4] = [(0.0, 0.0, 0.3), (0.55, 0.35, 0.28), (-0.5, -0.45, 0.33), (0.05, -0.75, 0.2)]
    sol["shape"][4] = S.CSG
    sol["prm"][4, 0] = 0
    sol["center"][4] = (-0.62, 0.41, 0.31)
    sol["shape"][5] = S.PLANE
    sol["center"][5] = (0.0, 0.0, 0.07)
    sol["q"][5] = S.quaternion_xyz((np.radians(92.0), 0.0, 0.0))
    sol["q"][:5] = S.random_unit_quaternions(rng, 5)
    sol["vel"][:5] = rng.normal(0, 0.1, (5, 3))
    sol["omega"][:5] = rng.normal(0, 1.0, (5, 3))
    return sol


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_ogrid_mixed_solids(prec):
    pm = _ogrid()
    rng = np.random.default_rng(31)
    table = np.zeros(1, dtype=S.CSG_DTYPE)
    table[0] = scenes.csg_family(rng, 1, 0.3)
    sol = _ogrid_solids(rng, table)
    U = rng.normal(size=(pm.n_cells, 3))
    oracle.set_csg(table)
    ref = oracle.OracleMesh(pm).interact(sol, U, rho_f=1.2, dt=1e-3, capture=True, capacity=8 * pm.n_cells)
    assert all(int(c[:3].sum()) > 10 for c in ref["counts"])
    with SolidCloud(pm, rho_f=1.2, precision=prec) as c:
        c.set_csg_shapes(table)
        c.set_solids(sol)
        c.set_velocity(U)
        ft = c.interact(0.0, 1e-3)
        _compare(c, ref, sol.size, TOL[prec], ft)
        # a second step after moving the solids (the buffers are reused)
        sol2 = sol.copy()
        sol2["center"][:5] += 0.03
        ref2 = oracle.OracleMesh(pm).interact(sol2, U, rho_f=1.2, dt=1e-3, capture=True, capacity=8 * pm.n_cells)
        c.set_solids(sol2)
        ft2 = c.interact(0.0, 1e-3)
        _compare(c, ref2, sol.size, TOL[prec], ft2)


@pytest.mark.parametrize("walls", [False, True], ids=["bodies", "walls"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_jittered_block_both_paths(prec, walls):
    sc = scenes.small_mixed_3d(20, seed=8, jitter=0.2)
    sol = sc.solids
    if walls:  # six walls of the cube (floor and ceiling inset) and one tilted plane
        tilt = S.make_solids(1)
        tilt["shape"] = S.PLANE
        tilt["center"] = (0.45, 0.9, 0.55)
        tilt["q"] = S.quaternion_xyz((0.3, 0.2, 2.8))
        sol = np.concatenate([sol, scenes.walls_3d(0.13, 0.0), tilt])
    pm = sc.mesh.to_polymesh()
    with SolidCloud(sc.mesh, rho_f=sc.rho_f, precision=prec) as a, \
            SolidCloud(pm, rho_f=sc.rho_f, precision=prec) as b:
        for c in (a, b):
            c.set_solids(sol)
            c.set_velocity(sc.U)
        fa, fb = a.interact(0.0, sc.dt), b.interact(0.0, sc.dt)
        if walls and prec == "fp64":
            ref = oracle.OracleMesh(pm).interact(sol, sc.U, two_d=False, rho_f=sc.rho_f, dt=sc.dt, capture=True,
                                                 capacity=8 * pm.n_cells)
            _compare(b, ref, sol.size, TOL[prec], ft=fb)
        for s in range(sol.size):
            ca, cb = a.candidates(s), b.candidates(s)
            assert np.array_equal(ca[0], cb[0]) and np.array_equal(ca[1], cb[1]), s
            assert np.max(np.abs(ca[2] - cb[2]), initial=0.0) <= TOL[prec]["a"], s
        ga, gb = a.fields(), b.fields()
        assert np.array_equal(ga["Ct"], gb["Ct"])
        assert np.max(np.abs(ga["As"] - gb["As"])) <= TOL[prec]["a"]
        scale = np.maximum(np.abs(fa), 1e-4 * np.abs(fa).max())
        assert np.max(np.abs(fa - fb) / scale) <= TOL[prec]["ft"]


def test_coupled_steps_against_oracle():
    pm = _ogrid(n=12, nz=6)
    rng = np.random.default_rng(5)
    table = np.zeros(1, dtype=S.CSG_DTYPE)
    table[0] = scenes.csg_family(rng, 1, 0.3)
    sol0 = _ogrid_solids(rng, table)
    ns = sol0.size
    rho = np.linspace(1.5, 2.5, ns)
    masks = np.full(ns, -1, dtype=np.int32)
    masks[-1] = motion_mask("b000000")
    sol0["vel"][-1] = 0.0
    sol0["omega"][-1] = 0.0
    g = np.array([0.0, 0.0, -9.81])
    dt = 1e-3
    oracle.set_csg(table)
    om = oracle.OracleMesh(pm)
    sol, dyn, U = sol0.copy(), oracle.make_dyn(ns, rho, masks), rng.normal(0, 0.2, (pm.n_cells, 3))
    with SolidCloud(pm, rho_f=1.1, precision="fp64") as c:
        c.set_csg_shapes(table)
        c.set_solids(sol0)
        c.set_velocity(U)
        c.set_dynamics(rho, masks)
        for step in range(3):
            t = step * dt
            ft = c.interact(t, dt)
            c.update_velocity(dt)
            c.evolve(t, dt, g)
            c.fix_internal()
            r = om.interact(sol, U, rho_f=1.1, dt=dt)
            np.testing.assert_array_equal(c.fields()["Ct"], r["Ct"])
            scale = np.maximum(np.abs(r["ft"]), 1e-4 * np.abs(r["ft"]).max())
            assert np.max(np.abs(ft - r["ft"]) / scale) <= 1e-9
            U = om.update_velocity(U, r["Fs"], dt)
            sol, dyn = oracle.evolve(sol, dyn, r["ft"], t, dt, g, 1.1)
            U = om.fix_internal(U, r["Ct"], sol)
            got = c.solids()
            st = lambda x: np.concatenate([x["center"], x["q"], x["vel"], x["omega"]], axis=1)
            np.testing.assert_allclose(st(got), st(sol), rtol=1e-9, atol=1e-12)
            assert np.max(np.abs(c.velocity() - U)) <= 1e-9 * max(np.abs(U).max(), 1.0)


def test_refused_calls():
    pm = _ogrid(n=8, nz=4)
    sol = S.make_solids(1)
    sid, p = S.make_shape("Sphere", radius=0.2)
    sol["shape"], sol["prm"] = sid, p
    with SolidCloud(pm, precision="fp64") as c:
        c.set_solids(sol)
        c.set_dynamics(2.0)
        for call in (lambda: c.run_steps(0.0, 1e-3, 2), lambda: c.interact_steps(0.0, 1e-3, 2),
                     lambda: c.autotune(1e-3)):
            with pytest.raises(SdfibmError) as e:
                call()
            assert e.value.status == -6
        c.interact(0.0, 1e-3)  # still usable
        assert c.pair_counts()[0] + c.pair_counts()[1] > 0
