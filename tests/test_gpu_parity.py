"""HIP path vs the CPU oracle, through the C ABI (needs an MI355X).

Tolerances:
  candidate lists (sets, classes, order, seed)  bit-exact, both precisions
  FP64 mode   alpha / As / Ts  |d| <= 1e-12;  Fs, force, torque  rel 1e-9
  FP32 mode   alpha / As / Ts  |d| <= 1e-6;   Fs rel 1e-5;  force, torque rel 1e-4
(alpha is a fraction of the cell, so |d| <= 1e-6 is 1e-6 relative to a full cell; force and
torque errors are relative to max(|value|, floor), floor = 1e-4 (FP64) or 1e-3 (FP32) of the
largest solid force: a boundary wall whose inside vertices are rounding noise carries forces
~1e-4 of the others, built from pyramid heights ~1e-6 h that FP32 resolves to ~1e-3)
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from _helpers import cylinder_case, split_candidates, toolvof_mesh, toolvof_solids
from sdfibm_amd import _lib, scenes
from sdfibm_amd import mesh as M
from sdfibm_amd import solids as S
from sdfibm_amd.cloud import SolidCloud

pytestmark = pytest.mark.gpu

TOL = {"fp64": dict(a=1e-12, f=1e-9, ft=1e-9, floor=1e-4), "fp32": dict(a=1e-6, f=1e-5, ft=1e-4, floor=1e-3)}


def _compare(cloud, ref, ft, ns, prec, fscale):
    tol = TOL[prec]
    got = cloud.fields()
    parts = split_candidates(ref, ns)
    for s in range(ns):
        counts, cells, alpha = cloud.candidates(s)
        rc, rcells, ralpha = parts[s]
        assert np.array_equal(counts, rc), (s, counts, rc)
        assert np.array_equal(cells, rcells), s
        assert np.max(np.abs(alpha - ralpha), initial=0.0) <= tol["a"], (s, np.max(np.abs(alpha - ralpha)))
    assert np.array_equal(got["Ct"], ref["Ct"])
    assert np.max(np.abs(got["As"] - ref["As"])) <= tol["a"]
    assert np.max(np.abs(got["Ts"] - ref["Ts"])) <= tol["a"] * 4
    fs_scale = max(np.abs(ref["Fs"]).max(), 1e-30)
    assert np.max(np.abs(got["Fs"] - ref["Fs"])) <= tol["f"] * fs_scale
    ft_scale = np.maximum(np.abs(ref["ft"]), max(fscale, tol["floor"] * np.abs(ref["ft"]).max()))
    assert np.max(np.abs(ft - ref["ft"]) / ft_scale) <= tol["ft"]


def _run(mesh, sol, U, two_d, rho_f, dt, prec, structured=False):
    with SolidCloud(mesh, two_d=two_d, rho_f=rho_f, precision=prec, force_structured=structured) as c:
        c.set_solids(sol)
        c.set_velocity(U)
        ft = c.interact(0.0, dt)
        yield c, ft


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("structured", [False, True])
def test_mixed_shapes_3d(prec, structured):
    sc = scenes.small_mixed_3d(24, seed=7)
    om = oracle.OracleMesh(sc.mesh.to_polymesh())
    ref = om.interact(sc.solids, sc.U, rho_f=sc.rho_f, dt=sc.dt, capture=True)
    for c, ft in _run(sc.mesh, sc.solids, sc.U, False, sc.rho_f, sc.dt, prec, structured):
        fscale = 1e-4 * np.abs(ref["ft"]).max()
        _compare(c, ref, ft, sc.solids.size, prec, fscale)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_small_spheres_one_wave_per_solid(prec):
    """2-4-cell spheres: mean tiles under 512 cells select the one-wave cut-cell pass."""
    sc = scenes.cloud_3d(32, 150, frac_sphere=1.0, d_min=2.0, d_max=4.0, d_uniform=True)
    om = oracle.OracleMesh(sc.mesh.to_polymesh())
    ref = om.interact(sc.solids, sc.U, rho_f=sc.rho_f, dt=sc.dt, capture=True)
    for c, ft in _run(sc.mesh, sc.solids, sc.U, False, sc.rho_f, sc.dt, prec):
        _compare(c, ref, ft, sc.solids.size, prec, 1e-4 * np.abs(ref["ft"]).max())


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("structured", [False, True])
def test_oversized_solids(prec, structured):
    """Solids whose tile exceeds 63 cells along x or 4096 vertex rows take the unbounded path
    (mesh-wide cells, union-find component of the seed) among ordinary solids and a plane."""
    mesh = M.uniform_box((96, 40, 36), (0, 0, 0), (1.2, 0.5, 0.45))
    sol = S.make_solids(5)
    sh, p = S.make_shape("Box", radiusa=0.45, radiusb=0.08, radiusc=0.07)  # 72 cells along x
    sol["shape"][0], sol["prm"][0] = sh, p
    sol["center"][0] = (0.6, 0.25, 0.22)
    sol["q"][0] = S.quaternion_xyz((0.1, 0.2, 0.3))
    sh, p = S.make_shape("Sphere", radius=0.2)  # an ordinary large tile
    sol["shape"][1], sol["prm"][1] = sh, p
    sol["center"][1] = (0.3, 0.27, 0.21)
    sh, p = S.make_shape("Ellipsoid", radiusa=0.55, radiusb=0.1, radiusc=0.12)  # 90 cells along x
    sol["shape"][2], sol["prm"][2] = sh, p
    sol["center"][2] = (0.62, 0.2, 0.25)
    sol["q"][2] = S.quaternion_xyz((0.0, 0.0, 0.15))
    sh, p = S.make_shape("Sphere", radius=0.05)
    sol["shape"][3], sol["prm"][3] = sh, p
    sol["center"][3] = (0.9, 0.35, 0.3)
    sol["shape"][4] = S.PLANE
    sol["center"][4] = (0.0, 0.06, 0.0)
    sol["q"][4] = S.quaternion_xyz((0.0, 0.0, 0.05))
    rng = np.random.default_rng(3)
    sol["vel"] = rng.normal(0, 0.1, (5, 3))
    sol["omega"] = rng.normal(0, 1.0, (5, 3))
    U = rng.normal(0, 1.0, (mesh.n_cells, 3))
    om = oracle.OracleMesh(mesh.to_polymesh())
    ref = om.interact(sol, U, rho_f=1.2, dt=1e-3, capture=True)
    for c, ft in _run(mesh, sol, U, False, 1.2, 1e-3, prec, structured):
        _compare(c, ref, ft, sol.size, prec, 1e-4 * np.abs(ref["ft"]).max())


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_jittered_mesh(prec):
    sc = scenes.small_mixed_3d(20, seed=3, jitter=0.2)
    om = oracle.OracleMesh(sc.mesh.to_polymesh())
    ref = om.interact(sc.solids, sc.U, rho_f=sc.rho_f, dt=sc.dt, capture=True)
    for c, ft in _run(sc.mesh, sc.solids, sc.U, False, sc.rho_f, sc.dt, prec):
        _compare(c, ref, ft, sc.solids.size, prec, 1e-4 * np.abs(ref["ft"]).max())


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("structured", [False, True])
def test_ragged_stretched_mesh(prec, structured):
    """37 x 23 x 13 cells (no axis a multiple of the field pass's brick), solids crossing the
    mesh boundary, and a plane: every partial brick and every cell-per-thread slot past the mesh
    edge.  Rectilinear: stretched spacing along x and z; structured: interior vertices jittered
    by up to 0.2 cells (the structured kind takes meshes within 0.45 cells of a uniform grid)."""
    rng = np.random.default_rng(77)
    if structured:
        m = M.jittered_box((37, 23, 13), (0.0, 0.0, 0.0), (1.0, 23 / 37, 13 / 37), 0.2, 78)
    else:
        xs = np.cumsum(np.r_[0.0, rng.uniform(0.8, 1.25, 37)])
        xs /= xs[-1]
        ys = np.linspace(0.0, 23 / 37, 24)
        zs = np.cumsum(np.r_[0.0, rng.uniform(0.7, 1.3, 13)])
        zs *= (13 / 37) / zs[-1]
        m = M.rectilinear(xs, ys, zs)
    h = 1.0 / 37
    kinds = [("Sphere", dict(radius=3.1 * h)), ("Ellipsoid", dict(radiusa=4.2 * h, radiusb=2.3 * h, radiusc=1.7 * h)),
             ("Box", dict(radiusa=2.5 * h, radiusb=1.6 * h, radiusc=3.3 * h)), ("Sphere", dict(radius=2.2 * h)),
             ("Ellipsoid", dict(radiusa=3.0 * h, radiusb=3.0 * h, radiusc=2.0 * h)), ("Plane", {})]
    sol = S.make_solids(len(kinds))
    centres = [(0.02, 0.3, 0.1), (0.5, 0.31, 0.17), (0.97, 0.05, 0.33), (0.7, 0.6, 0.02), (0.3, 0.12, 0.3),
               (0.5, 0.1 * h + 0.04, 0.2)]
    for t, (name, prm) in enumerate(kinds):
        sid, pv = S.make_shape(name, **prm)
        sol[t]["shape"] = sid
        sol[t]["prm"] = pv
        sol[t]["center"] = centres[t]
        sol[t]["q"] = S.random_unit_quaternions(rng, 1)[0] if name != "Plane" else S.quaternion_xyz((0.2, 0.0, 0.1))
        if name != "Plane":
            sol[t]["vel"] = rng.normal(0, 0.1, 3)
            sol[t]["omega"] = rng.normal(0, 1.0, 3)
    U = rng.normal(size=(m.n_cells, 3))
    om = oracle.OracleMesh(m.to_polymesh())
    ref = om.interact(sol, U, rho_f=1.1, dt=1e-3, capture=True)
    assert (ref["counts"][:, :3].sum(1) > 0).all()
    for c, ft in _run(m, sol, U, False, 1.1, 1e-3, prec, structured):
        _compare(c, ref, ft, sol.size, prec, 1e-4 * np.abs(ref["ft"]).max())


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_toolvof_fixture_on_gpu(prec):
    """The reference's alpha.water through the structured-mesh path (mesh points from the file)."""
    m, g = toolvof_mesh()
    sol = toolvof_solids()
    U = np.zeros((m.n_cells, 3))
    for c, ft in _run(m, sol, U, True, 1.0, 1.0, prec):
        As = c.fields()["As"]
        assert np.max(np.abs(As - g["alpha"])) <= (1e-12 if prec == "fp64" else 1e-6)


def test_cylinder_fixture_on_gpu():
    m, sol, g = cylinder_case()
    for c, ft in _run(m, sol, np.zeros((m.n_cells, 3)), True, 1.0, 1.0, "fp64"):
        As = c.fields()["As"]
        bad = np.nonzero(np.abs(As - g["As"]) > 1e-9)[0]
        assert list(bad) == [57 + 120 * 50]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_sedimentation_with_walls(prec):
    """2-D discs + four walls on the boundary (walls rotated by 180 / +-90 degrees get rounding-
    level inside vertices on the boundary row; the classification must match bit for bit)."""
    sc = scenes.sedimentation_2d(100, 200, 16)
    rng = np.random.default_rng(5)
    U = rng.normal(size=(sc.mesh.n_cells, 3))
    U[:, 2] = 0
    sol = sc.solids.copy()
    sol["vel"][:16, :2] = rng.normal(0, 0.1, (16, 2))
    sol["omega"][:16, 2] = rng.normal(0, 1.0, 16)
    om = oracle.OracleMesh(sc.mesh.to_polymesh())
    ref = om.interact(sol, U, two_d=True, rho_f=1.0, dt=sc.dt, capture=True)
    for c, ft in _run(sc.mesh, sol, U, True, 1.0, sc.dt, prec):
        _compare(c, ref, ft, sol.size, prec, 1e-4 * np.abs(ref["ft"]).max())


@pytest.mark.parametrize("nz", [1, 3])
def test_disconnected_pieces_are_dropped(nz):
    """A disc with a thin tail at 60 poses: where the tail's cells do not connect to the disc's
    through face neighbours, only the seed's component is a candidate (the GPU's flood fill runs
    instead of its connected-set shortcut); elsewhere the shortcut takes the whole active set."""
    n = 24
    h = 1.0 / n
    m = M.uniform_box((n, n, nz), (0, 0, -0.5 * nz * h), (1, 1, 0.5 * nz * h))
    om = oracle.OracleMesh(m.to_polymesh())
    U = np.random.default_rng(3).normal(size=(m.n_cells, 3))
    P = m.points
    n_split = 0
    with SolidCloud(m, precision="fp64") as c:
        for k in range(60):
            s = S.make_solids(1)
            s["shape"] = S.CIRCLE_TAIL
            s["prm"][0, :3] = (2.2 * h, 4.0, 0.2 * h)
            s["center"] = (0.31 + 0.013 * k / n, 0.5 + 0.37 * h, 0.0)
            s["q"] = S.quaternion_xyz((0, 0, 0.3 + 0.05 * k))
            ref = om.interact(s, U, capture=True)
            c.set_solids(s)
            c.set_velocity(U)
            ft = c.interact(0.0, 1.0)
            _compare(c, ref, ft, 1, "fp64", 1e-4 * np.abs(ref["ft"]).max())
            ins = np.array([oracle.is_inside(s[0], p) for p in P]).reshape(nz + 1, n + 1, n + 1)
            act = np.zeros((nz, n, n), dtype=bool)
            for dz in (0, 1):
                for dy in (0, 1):
                    for dx in (0, 1):
                        act |= ins[dz:dz + nz, dy:dy + n, dx:dx + n]
            n_split += int(act.sum() > ref["counts"][0, :3].sum())
    assert n_split > 0, "no pose left an active piece outside the seed's component"


def test_many_solids_multi_pass_scan():
    """20,000 overlapping small spheres: more solids than the one-workgroup tile-offset scan
    takes, so the tile offsets come from the three-pass scan.  Fields, forces and a sample of
    the candidate lists against the oracle (FP64 tolerances above).  (128^3 cells keep every
    brick of the field pass under its 128-solid capacity.)"""
    n = 128
    h = 1.0 / n
    m = M.uniform_box((n, n, n), (0, 0, 0), (1, 1, 1))
    rng = np.random.default_rng(21)
    ns = 20000
    sol = S.make_solids(ns)
    sol["shape"] = S.SPHERE
    sol["prm"][:, 0] = rng.uniform(0.8, 2.0, ns) * h
    sol["center"] = rng.uniform(0.05, 0.95, (ns, 3))
    sol["vel"] = rng.normal(0, 0.1, (ns, 3))
    sol["omega"] = rng.normal(0, 1.0, (ns, 3))
    U = rng.normal(size=(m.n_cells, 3))
    om = oracle.OracleMesh(m.to_polymesh())
    ref = om.interact(sol, U, rho_f=1.0, dt=1e-3, capture=True, capacity=200 * ns)
    tol = TOL["fp64"]
    for c, ft in _run(m, sol, U, False, 1.0, 1e-3, "fp64"):
        got = c.fields()
        assert np.array_equal(got["Ct"], ref["Ct"])
        assert np.max(np.abs(got["As"] - ref["As"])) <= tol["a"]
        assert np.max(np.abs(got["Fs"] - ref["Fs"])) <= tol["f"] * np.abs(ref["Fs"]).max()
        ft_scale = np.maximum(np.abs(ref["ft"]), tol["floor"] * np.abs(ref["ft"]).max())
        assert np.max(np.abs(ft - ref["ft"]) / ft_scale) <= tol["ft"]
        parts = split_candidates(ref, ns)
        for s in range(0, ns, 97):
            counts, cells, _ = c.candidates(s)
            assert np.array_equal(counts, parts[s][0]) and np.array_equal(cells, parts[s][1]), s


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_code_buffer_grows_when_tiles_grow(prec):
    """Elongated ellipsoids set along x, then turned by the per-step pose update (same shapes: no
    re-estimate of the tile codes' size): their tiles outgrow the code buffer, which interact
    grows before running the step again.  The result equals a fresh context's bit for bit."""
    n = 24
    h = 1.0 / n
    m = M.uniform_box((n, n, n), (0, 0, 0), (1, 1, 1))
    rng = np.random.default_rng(31)
    ns = 20
    sol = S.make_solids(ns)
    sol["shape"] = S.ELLIPSOID
    sol["prm"][:, :3] = (6.0 * h, 1.0 * h, 1.0 * h)
    sol["center"] = rng.uniform(0.3, 0.7, (ns, 3))
    sol["q"] = (1.0, 0.0, 0.0, 0.0)
    sol["vel"] = rng.normal(0, 0.1, (ns, 3))
    U = rng.normal(size=(m.n_cells, 3))
    turned = sol.copy()
    # long axis from x to the diagonal (1, 1, 1)/sqrt(3): tiles of about 8^3 cells instead of 13 x 3 x 3
    th = np.arccos(1.0 / np.sqrt(3.0))
    turned["q"] = (np.cos(th / 2), 0.0, -np.sin(th / 2) / np.sqrt(2.0), np.sin(th / 2) / np.sqrt(2.0))
    with SolidCloud(m, precision=prec) as a, SolidCloud(m, precision=prec) as b:
        a.set_solids(sol)
        a.set_velocity(U)
        a.interact(0.0, 1e-3)
        cap0 = a.buffer_capacity()[0]
        a.set_solids(turned)  # same shapes: the fast path, no new size estimate
        fa = a.interact(0.0, 1e-3)
        assert a.buffer_capacity()[0] > cap0  # the growth path ran
        ga = a.fields()
        b.set_solids(turned)
        b.set_velocity(U)
        fb = b.interact(0.0, 1e-3)
        gb = b.fields()
        # the cut-cell pass's workgroup size follows the mean tile size at set_solids, which
        # differs between the two contexts: only the order of the FP64 force/torque sums
        np.testing.assert_allclose(fa, fb, rtol=1e-12, atol=1e-14 * np.abs(fb).max())
        for key in ("As", "Fs", "Ts", "Ct"):
            np.testing.assert_array_equal(ga[key], gb[key])
        for s_ in range(0, ns, 7):
            ca, xa, aa = a.candidates(s_)
            cb, xb, ab = b.candidates(s_)
            assert np.array_equal(ca, cb) and np.array_equal(xa, xb) and np.array_equal(aa, ab)


def test_seed_fallback_and_empty_solid():
    n = 16
    m = M.uniform_box((n, n, n), (0, 0, 0), (1, 1, 1))
    h = 1.0 / n
    sol = S.make_solids(3)
    sol["shape"] = S.SPHERE
    sol["prm"][0, :4] = (1.3 * h, -3.5 * h, 0.0, 0.0)       # body away from the solid position
    sol["center"][0] = (0.5, 0.5, 0.5)
    sol["prm"][1, 0] = 0.2 * h                              # smaller than a cell: no inside vertex
    sol["center"][1] = (0.5 + 0.5 * h, 0.5 + 0.5 * h, 0.5 + 0.5 * h)
    sol["prm"][2, 0] = 0.3
    sol["center"][2] = (1.2, 0.5, 0.5)                      # partly outside the mesh
    U = np.random.default_rng(1).normal(size=(m.n_cells, 3))
    om = oracle.OracleMesh(m.to_polymesh())
    ref = om.interact(sol, U, capture=True)
    assert ref["counts"][1, :3].sum() == 0
    for c, ft in _run(m, sol, U, False, 1.0, 1.0, "fp64"):
        _compare(c, ref, ft, 3, "fp64", 1e-4 * np.abs(ref["ft"]).max())
        assert np.all(ft[1] == 0)


def test_error_paths():
    m = M.uniform_box((8, 8, 8), (0, 0, 0), (1, 1, 1))
    with SolidCloud(m) as c:
        with pytest.raises(_lib.SdfibmError) as e:
            c.interact(0.0, 1.0)
        assert e.value.status == -6  # ERR_STATE: interact before set_solids
        bad = S.make_solids(1)
        bad["shape"] = 42
        with pytest.raises(_lib.SdfibmError) as e:
            c.set_solids(bad)
        assert e.value.status == -4
        nan = S.make_solids(1)
        nan["prm"][0, 0] = 0.1
        nan["center"] = (np.nan, 0, 0)
        with pytest.raises(_lib.SdfibmError) as e:
            c.set_solids(nan)
        assert e.value.status == -1
        ok = S.make_solids(1)
        ok["prm"][0, 0] = 0.2
        ok["center"] = (0.5, 0.5, 0.5)
        c.set_solids(ok)
        with pytest.raises(_lib.SdfibmError) as e:
            c.interact(0.0, 0.0)
        assert e.value.status == -1
        c.interact(0.0, 1.0)
        huge = S.make_solids(1)
        huge["shape"] = S.BOX
        huge["prm"][0, :3] = (0.45, 0.45, 0.45)
        huge["center"] = (0.5, 0.5, 0.5)
    big = M.uniform_box((80, 8, 8), (0, 0, 0), (1, 0.1, 0.1))
    with SolidCloud(big) as c:
        c.set_solids(np.concatenate([huge] * 33))  # tiles wider than 63 cells: the big-tile pass
        c.interact(0.0, 1.0)
        wall = S.make_solids(1)
        wall["shape"] = S.PLANE
        wall["center"] = (0.5, 0.05, 0.05)
        with pytest.raises(_lib.SdfibmError) as e:
            c.set_solids(np.concatenate([wall] * 33))  # more than 32 plane (or oversized) solids
        assert e.value.status == -1


def test_repeat_is_bitwise_deterministic():
    sc = scenes.small_mixed_3d(24, seed=8)
    with SolidCloud(sc.mesh, rho_f=sc.rho_f, precision="fp32") as c:
        c.set_solids(sc.solids)
        c.set_velocity(sc.U)
        a = c.interact(0.0, sc.dt)
        fa = c.fields()
        b = c.interact(0.0, sc.dt)
        fb = c.fields()
    assert np.array_equal(a, b)
    for k in fa:
        assert np.array_equal(fa[k], fb[k])


def test_cpp_host_demo():
    """The C++ mirror of SolidCloud (include/sdfibm_hip.hpp) driving interact from a g++ program:
    step 0 on a resting disc in uniform U = (1,0,0) gives F = rho_f (sum alpha V) U / dt."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sdfibm_amd", "host",
                       "coupling_demo")
    out = subprocess.run([exe, "2"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = [list(map(float, ln.split())) for ln in out.stdout.strip().splitlines()]
    assert len(rows) == 2
    step, vol, fx, fy = rows[0]
    assert abs(vol / (np.pi * 0.25) - 1) < 0.01
    assert abs(fx - vol * 1.0 / 0.01) < 1e-9 * abs(fx)
    assert abs(fy) < 1e-9 * abs(fx)
    assert abs(rows[1][2]) < abs(fx)  # U -= Fs dt removed the slip inside the disc


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_update_velocity_and_fix_internal(prec):
    """U -= Fs dt (main.cpp:70) then fixInternal with moved solids (solidcloud.cpp:288-301)."""
    sc = scenes.small_mixed_3d(20, seed=4)
    om = oracle.OracleMesh(sc.mesh.to_polymesh())
    ref = om.interact(sc.solids, sc.U, rho_f=sc.rho_f, dt=sc.dt)
    U1 = om.update_velocity(sc.U, ref["Fs"], sc.dt)
    moved = sc.solids.copy()
    moved["vel"] += 0.05
    moved["omega"] *= -1.0
    U2 = om.fix_internal(U1, ref["Ct"], moved)
    tol = 1e-12 if prec == "fp64" else 2e-5
    with SolidCloud(sc.mesh, rho_f=sc.rho_f, precision=prec) as c:
        c.set_solids(sc.solids)
        c.set_velocity(sc.U)
        c.interact(0.0, sc.dt)
        c.update_velocity(sc.dt)
        g1 = c.velocity()
        c.set_solids(moved)
        c.fix_internal()
        g2 = c.velocity()
    scale = np.abs(U1).max()
    assert np.max(np.abs(g1 - U1)) <= tol * scale
    assert np.max(np.abs(g2 - U2)) <= tol * scale
    assert np.any(ref["Ct"] >= 4)


def test_pose_update_matches_a_fresh_context():
    """set_solids with the same shapes and new poses (the per-step update of a host-side solver)
    gives what a fresh context gives for those poses, bit for bit."""
    sc = scenes.small_mixed_3d(24, seed=9)
    wall = S.make_solids(1)
    wall["shape"] = S.PLANE
    wall["center"] = (0.3, 0.2, 0.1)
    wall["q"] = S.quaternion_xyz((0.1, 0.0, 0.2))
    start = np.concatenate([sc.solids, wall])
    moved = start.copy()
    rng = np.random.default_rng(1)
    moved["center"] += rng.normal(0, 0.03, moved["center"].shape)  # the plane moves too
    q = moved["q"] + rng.normal(0, 0.1, moved["q"].shape)
    moved["q"] = q / np.linalg.norm(q, axis=1, keepdims=True)
    moved["vel"] += 0.1
    with SolidCloud(sc.mesh, rho_f=sc.rho_f, precision="fp32") as a, \
            SolidCloud(sc.mesh, rho_f=sc.rho_f, precision="fp32") as b:
        a.set_solids(start)
        a.set_velocity(sc.U)
        a.interact(0.0, sc.dt)
        a.interact(0.0, sc.dt)  # unmoved plane: its cell set is reused
        a.set_solids(moved)
        fa = a.interact(0.0, sc.dt)
        b.set_solids(moved)
        b.set_velocity(sc.U)
        fb = b.interact(0.0, sc.dt)
        np.testing.assert_array_equal(fa, fb)
        ga, gb = a.fields(), b.fields()
        for key in ("As", "Fs", "Ts", "Ct"):
            np.testing.assert_array_equal(ga[key], gb[key])
