"""Pins of the CPU oracle: reference-produced fixtures, closed forms, invariants and brute
force.  These run without a GPU."""
from __future__ import annotations

import hashlib
import itertools
import math

import numpy as np
import pytest
from scipy import ndimage

import oracle
from _helpers import (CYLINDER_ODD_CELL, cylinder_case, cylinder_re200, oracle_run, split_candidates,
                      toolvof_mesh, toolvof_solids)
from sdfibm_amd import mesh as M
from sdfibm_amd import solids as S


# ---------------------------------------------------------------------------------------------
# reference-produced fixtures
# ---------------------------------------------------------------------------------------------
def test_toolvof_topology_matches_blockmesh():
    """The structured mesh's polyMesh export reproduces blockMesh's faces/owner/neighbour
    (point numbering, internal-face order, hex-model vertex order, patch order)."""
    m, g = toolvof_mesh()
    pm = m.to_polymesh()

    def dig(a):
        return hashlib.sha256(np.ascontiguousarray(a.astype(np.int32)).tobytes()).hexdigest()

    nint = pm.neighbour.size
    faces = pm.face_points.reshape(-1, 4)
    assert dig(pm.owner[:nint]) == str(g["internal_owner_sha256"])
    assert dig(pm.neighbour) == str(g["neighbour_sha256"])
    assert dig(faces[:nint]) == str(g["internal_faces_sha256"])
    # boundary: patch sizes/starts and, per patch, the owner-sorted (owner, face) rows
    parts = []
    for st, nf in zip(g["patch_starts"], g["patch_sizes"]):
        sl = slice(int(st), int(st) + int(nf))
        order = np.argsort(pm.owner[sl], kind="stable")
        parts.append(np.concatenate([pm.owner[sl][order][:, None], faces[sl][order]], axis=1))
    assert pm.owner.size == int(g["patch_starts"][-1] + g["patch_sizes"][-1])
    assert dig(np.concatenate(parts)) == str(g["boundary_sha256"])


def test_toolvof_alpha_fixture():
    """tool_vof/example/0/alpha.water = clip(sum As) for 12 solids + 2 planes, written by the
    reference.  The oracle reproduces all 40,000 cells to 1e-12."""
    m, g = toolvof_mesh()
    sol = toolvof_solids()
    om, r = oracle_run(m, sol, np.zeros((m.n_cells, 3)), two_d=True)
    d = np.abs(r["As"] - g["alpha"])
    assert d.max() < 1e-12, d.max()
    assert int((g["alpha"] == 1.0).sum()) == 12192
    assert int(((g["alpha"] > 0) & (g["alpha"] < 1)).sum()) == 1670


def test_cylinder_fixture():
    """examples/flow_past_cylinder/re200/0/As (written by initialCorrect of an older revision).
    Block 5 regenerated as a uniform 120x120 mesh: 14,399 of 14,400 cells agree to 1e-9; the
    one exception, cell (i=57, j=50), is 1.0 (clipped) in the file and 0.6685 here, while its
    mirror image (62, 50) agrees (0.667591); no apex-order variant of the current algorithm
    produces the file's value there."""
    m, sol, g = cylinder_case()
    om, r = oracle_run(m, sol, np.zeros((m.n_cells, 3)), two_d=True)
    assert int(g["outside_nonzero"]) == 0
    bad = np.nonzero(np.abs(r["As"] - g["As"]) > 1e-9)[0]
    assert list(bad) == [57 + 120 * 50]
    assert abs(r["As"][62 + 120 * 50] - 0.667591) < 1e-6
    area = float((r["As"] * 0.01).sum())
    assert abs(area - (float((g["As"] * 0.01).sum()) - (1.0 - r["As"][bad[0]]) * 0.01)) < 1e-10


def test_cylinder_nine_blocks():
    """The same file, whole: the nine graded blocks regenerated from the blockMeshDict numbers
    (blockmesh.py: merged points, blockMesh's face order) and run as a general polyMesh.  The
    oracle reproduces all 78,400 values to 1e-9 except the one documented cell; the mesh has
    blockMesh's sizes (157,922 points, 156,240 internal faces, the six patches)."""
    pm, info, sol, As = cylinder_re200()
    assert pm.points.shape[0] == 157922 and pm.n_cells == 78400 and pm.neighbour.size == 156240
    assert info["patches"]["front"] == (156240, 78400) and info["patches"]["inlet"] == (313880, 280)
    om = oracle.OracleMesh(pm)
    r = om.interact(sol, np.zeros((pm.n_cells, 3)), two_d=True, rho_f=1.0, dt=1.0)
    bad = np.nonzero(np.abs(r["As"] - As) > 1e-9)[0]
    assert list(bad) == [CYLINDER_ODD_CELL]
    assert np.count_nonzero(As) == 352 and np.abs(r["As"]).sum() > 300


# ---------------------------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------------------------
def _solid(shape, prm=None, center=(0, 0, 0), q=(1, 0, 0, 0), vel=(0, 0, 0), omega=(0, 0, 0)):
    s = S.make_solids(1)
    s["shape"] = shape
    if prm is not None:
        s["prm"][0, :len(prm)] = prm
    s["center"] = center
    s["q"] = q
    s["vel"] = vel
    s["omega"] = omega
    return s[0]


def test_world_to_local_closed_forms():
    s = _solid(S.SPHERE, [1.0], center=(0.3, -0.2, 0.7))
    p = np.array([1.25, 2.5, -3.75])
    assert np.array_equal(oracle.world_to_local(s, p), p - np.array([0.3, -0.2, 0.7]))
    # +90 degrees about z: world +x is local -y
    s2 = _solid(S.SPHERE, [1.0], q=S.quaternion_xyz((0, 0, math.pi / 2)))
    loc = oracle.world_to_local(s2, np.array([1.0, 0.0, 0.0]))
    assert np.allclose(loc, [0.0, -1.0, 0.0], atol=1e-15)
    # a rotation preserves lengths
    rng = np.random.default_rng(3)
    q = S.random_unit_quaternions(rng, 1)[0]
    s3 = _solid(S.SPHERE, [1.0], q=q, center=(0.1, 0.2, 0.3))
    for _ in range(10):
        p = rng.normal(size=3)
        l = oracle.world_to_local(s3, p)
        assert abs(np.linalg.norm(l) - np.linalg.norm(p - [0.1, 0.2, 0.3])) < 1e-14


def test_sdf_closed_forms():
    sph = _solid(S.SPHERE, [0.5, 0.1, 0.0, 0.0])  # com = (0.1, 0, 0): body centred at -com
    assert abs(oracle.signed_distance(sph, [-0.1, 0.0, 0.0]) + 0.5) < 1e-15
    assert abs(oracle.signed_distance(sph, [0.9, 0.0, 0.0]) - 0.5) < 1e-15
    assert oracle.signed_distance(sph, [0.4, 0.0, 0.0]) == -1e-8  # |phi| < TOL -> -TOL
    assert not oracle.is_inside(sph, [0.4, 0.0, 0.0])  # on the surface: outside for the predicate
    box = _solid(S.BOX, [1.0, 2.0, 3.0])
    assert abs(oracle.signed_distance(box, [2.0, 3.0, 0.0]) - math.sqrt(2.0)) < 1e-15
    assert abs(oracle.signed_distance(box, [0.5, 0.0, 0.0]) + 0.5) < 1e-15
    rect = _solid(S.RECTANGLE, [1.0, 0.5])
    assert abs(oracle.signed_distance(rect, [0.0, 2.0, 7.0]) - 1.5) < 1e-15  # z ignored
    pl = _solid(S.PLANE, center=(0, 1, 0))
    assert oracle.signed_distance(pl, [3.0, 0.25, 1.0]) == -0.75
    # ellipsoid on an axis: 0.5 (x^2/a^2 - 1) / (|x| / a^2)
    ell = _solid(S.ELLIPSOID, [2.0, 1.0, 0.5])
    x = 3.0
    assert abs(oracle.signed_distance(ell, [x, 0, 0]) - 0.5 * (x * x / 4 - 1) / (x / 4)) < 1e-15
    # circle tail: the tail is the rectangle [0, 2 ra] x [-t, t], ra = (ratio+1) r / 2
    tail = _solid(S.CIRCLE_TAIL, [1.0, 1.0, 0.1])
    assert oracle.is_inside(tail, [1.9, 0.05, 0.0]) and not oracle.is_inside(tail, [1.9, 0.15, 0.0])
    assert abs(oracle.signed_distance(tail, [1.9, 0.3, 0.0]) - 0.2) < 1e-12
    # two-tail: tails at about +-30 degrees, half-thickness = thickness / 2
    tt = _solid(S.CIRCLE_TWOTAIL, [1.0, 1.0, 0.2])
    for sgn in (1, -1):
        p = 1.8 * np.array([math.cos(math.radians(30)), sgn * math.sin(math.radians(30)), 0.0])
        assert oracle.is_inside(tt, p)
    assert not oracle.is_inside(tt, [1.8, 0.0, 0.0])


def _cube_below_plane(n, c):
    """Exact volume of {x in [0,1]^3: n.x <= c} (all n_i != 0), by inclusion-exclusion."""
    n = np.asarray(n, dtype=np.float64)
    flip = n < 0
    c = c - np.sum(n[flip])
    n = np.abs(n)
    vol = 0.0
    for v in itertools.product((0, 1), repeat=3):
        t = c - float(np.dot(n, v))
        if t > 0:
            vol += (-1) ** sum(v) * t ** 3
    return vol / (6.0 * n[0] * n[1] * n[2])


def test_planar_fraction_exact_up_to_small():
    """A plane cut of a cube: the apex construction is exact for linear signed distances,
    up to the SMALL = 1e-6 shift of the apex (error O(SMALL / h))."""
    h = 0.1
    m = M.uniform_box((3, 3, 3), (0, 0, 0), (3 * h, 3 * h, 3 * h))
    om = oracle.OracleMesh(m.to_polymesh())
    rng = np.random.default_rng(11)
    centre_cell = 13  # (1,1,1)
    worst = 0.0
    for _ in range(200):
        q = S.random_unit_quaternions(rng, 1)[0]
        p0 = np.array([h, h, h]) + rng.uniform(0.05, 0.95, 3) * h
        s = _solid(S.PLANE, center=p0, q=q)
        # inside = local y < 0; local y = R^T(p - p0) . e_y = nrm . (p - p0), nrm = R e_y
        w, x, y, z = q
        nrm = np.array([2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)])
        if np.min(np.abs(nrm)) < 1e-3:
            continue
        # in cell-local unit coordinates u = (p - h)/h: inside <=> nrm . u <= nrm . (p0 - h)/h
        exact = _cube_below_plane(nrm, float(np.dot(nrm, (p0 - h) / h)))
        got = om.cell_fraction(s, centre_cell)
        worst = max(worst, abs(got - exact))
    assert worst < 20 * 1e-6 / h, worst


def test_sphere_volume_converges():
    """Sum of alpha V over a sphere converges to 4/3 pi r^3 at second order in h/r (the apex
    construction replaces the surface by planar facets per cell)."""
    errs = []
    for n, rc in ((16, 4), (32, 8)):
        m = M.uniform_box((n, n, n), (0, 0, 0), (1, 1, 1))
        r = rc / n
        sol = S.make_solids(1)
        sol["shape"] = S.SPHERE
        sol["prm"][0, 0] = r
        sol["center"] = (0.5 + 0.13 / n, 0.5 - 0.31 / n, 0.5 + 0.07 / n)
        om, res = oracle_run(m, sol, np.zeros((m.n_cells, 3)), two_d=False)
        errs.append(float(np.sum(res["As"])) / n ** 3 / (4 / 3 * math.pi * r ** 3) - 1)
        assert res["As"].max() <= 1.0 and res["As"].min() >= 0.0
    assert abs(errs[1]) < 0.012
    assert 3.0 < errs[0] / errs[1] < 5.0, errs


def test_rigid_motion_gives_zero_forcing():
    """uf = us at every cell (U is the solid's own rigid velocity field) -> Fs = 0, F = T = 0."""
    n = 16
    m = M.uniform_box((n, n, n), (0, 0, 0), (1, 1, 1))
    s = _solid(S.ELLIPSOID, [0.3, 0.2, 0.15], center=(0.52, 0.47, 0.5),
               q=S.random_unit_quaternions(np.random.default_rng(1), 1)[0],
               vel=(0.3, -0.2, 0.1), omega=(1.0, 2.0, -0.5))
    cc = m.cell_centres() if not m.rectilinear else None
    om = oracle.OracleMesh(m.to_polymesh())
    cc, V, _, _ = om.geometry()
    U = s["vel"] + np.cross(s["omega"], cc - s["center"])
    r = om.interact(np.array([s]), U, rho_f=1.7, dt=0.01)
    assert np.abs(r["Fs"]).max() < 1e-12
    assert np.abs(r["ft"]).max() < 1e-12


def test_uniform_flow_resting_solid_force():
    """U uniform, solid at rest: F = rho_f (sum alpha V) U / dt; T = rho_f sum alpha V (cc-c) x U / dt."""
    n = 20
    m = M.uniform_box((n, n, n), (0, 0, 0), (1, 1, 1))
    s = _solid(S.BOX, [0.2, 0.15, 0.1], center=(0.5, 0.51, 0.49),
               q=S.random_unit_quaternions(np.random.default_rng(2), 1)[0])
    om = oracle.OracleMesh(m.to_polymesh())
    cc, V, _, _ = om.geometry()
    u = np.array([0.7, -0.3, 0.2])
    rho, dt = 1.3, 0.05
    r = om.interact(np.array([s]), np.tile(u, (m.n_cells, 1)), rho_f=rho, dt=dt)
    a = r["Ts"]  # one solid: Ts = sum of that solid's alphas (not clipped)
    F = rho * np.sum(a * V) * u / dt
    T = rho * np.sum((a * V)[:, None] * np.cross(cc - s["center"], u), axis=0) / dt
    assert np.allclose(r["ft"][0, :3], F, rtol=1e-12, atol=0)
    assert np.allclose(r["ft"][0, 3:], T, rtol=1e-10, atol=1e-14)


def test_field_sum_equals_force_sum():
    """sum_cells rho_f Fs V = sum_solids F (Fs is not clipped; the identity is linear)."""
    from sdfibm_amd import scenes
    sc = scenes.small_mixed_3d(20, seed=5)
    om = oracle.OracleMesh(sc.mesh.to_polymesh())
    cc, V, _, _ = om.geometry()
    r = om.interact(sc.solids, sc.U, rho_f=sc.rho_f, dt=sc.dt)
    lhs = sc.rho_f * np.sum(r["Fs"] * V[:, None], axis=0)
    rhs = r["ft"][:, :3].sum(axis=0)
    assert np.allclose(lhs, rhs, rtol=1e-11, atol=1e-11)


def test_translation_by_whole_cells():
    """Shifting a solid by whole cells on a dyadic uniform mesh shifts its candidate set and
    leaves every fraction unchanged."""
    n = 16
    m = M.uniform_box((n, n, n), (0, 0, 0), (1, 1, 1))
    om = oracle.OracleMesh(m.to_polymesh())
    q = S.random_unit_quaternions(np.random.default_rng(4), 1)[0]
    a = _solid(S.ELLIPSOID, [0.19, 0.13, 0.09], center=(0.40625 + 0.01, 0.4375 + 0.003, 0.5), q=q)
    b = a.copy()
    b["center"] = a["center"] + np.array([3, -2, 1]) / n
    ra = om.interact(np.array([a]), np.zeros((m.n_cells, 3)), capture=True)
    rb = om.interact(np.array([b]), np.zeros((m.n_cells, 3)), capture=True)
    shift = 3 - 2 * n + n * n
    assert np.array_equal(ra["counts"][0, :3], rb["counts"][0, :3])
    assert np.array_equal(ra["cells"] + shift, rb["cells"])
    assert np.allclose(ra["alpha"], rb["alpha"], atol=1e-12)


# ---------------------------------------------------------------------------------------------
# enumeration semantics by brute force
# ---------------------------------------------------------------------------------------------
def _brute_candidates(m, om, s):
    """Connected component (face adjacency) of cells with >= 1 inside vertex, grown from the
    nearest cell (or the first such cell), found by exhaustive evaluation + labelling."""
    P = m.points
    ins = np.array([oracle.is_inside(s, p) for p in P]).reshape(m.nz + 1, m.ny + 1, m.nx + 1)
    act = np.zeros((m.nz, m.ny, m.nx), dtype=bool)
    cnt = np.zeros((m.nz, m.ny, m.nx), dtype=int)
    for dz, dy, dx in itertools.product((0, 1), repeat=3):
        cnt += ins[dz:dz + m.nz, dy:dy + m.ny, dx:dx + m.nx]
    act = cnt > 0
    lab, _ = ndimage.label(act)
    seed = om.nearest_cell(s["center"])
    flat = act.reshape(-1)
    if not flat[seed]:
        nz = np.nonzero(flat)[0]
        if nz.size == 0:
            return np.array([], dtype=np.int64), cnt.reshape(-1)
        seed = nz[0]
    comp = lab.reshape(-1) == lab.reshape(-1)[seed]
    return np.nonzero(comp)[0], cnt.reshape(-1)


@pytest.mark.parametrize("case", range(6))
def test_enumeration_matches_brute_force(case):
    rng = np.random.default_rng(100 + case)
    n = 12
    m = M.uniform_box((n, n, n), (0, 0, 0), (1, 1, 1))
    om = oracle.OracleMesh(m.to_polymesh())
    h = 1.0 / n
    kinds = [(S.ELLIPSOID, [4 * h, 1.2 * h, 0.9 * h]), (S.SPHERE, [2.6 * h, 0.0, 0.0, 0.0]),
             (S.CIRCLE_TWOTAIL, [1.6 * h, 2.5, 0.5 * h]), (S.BOX, [3 * h, 0.7 * h, 2 * h]),
             (S.SPHERE, [1.4 * h, -3 * h, 0.5 * h, 0.0]),  # body off the solid's position
             (S.CIRCLE_TAIL, [1.2 * h, 3.0, 0.45 * h])]
    shape, prm = kinds[case]
    q = S.random_unit_quaternions(rng, 1)[0] if shape not in S.TWO_D_SHAPES else \
        S.quaternion_xyz((0, 0, rng.uniform(0, 6.28)))
    s = _solid(shape, prm, center=rng.uniform(0.35, 0.65, 3), q=q)
    r = om.interact(np.array([s]), np.zeros((m.n_cells, 3)), capture=True)
    want, cnt = _brute_candidates(m, om, s)
    got = np.sort(r["cells"])
    assert np.array_equal(got, want)
    c = r["counts"][0]
    inside = r["cells"][:c[0]]
    assert np.all(cnt[inside] == 8) and np.all(np.diff(inside) > 0)
    cut = r["cells"][c[0]:]
    assert np.all(cnt[cut] < 8)


def test_disconnected_tail_is_dropped():
    """Only the seed's component is enumerated: a thin tail whose inside vertices do not
    connect to the disc's through face neighbours is not a candidate."""
    n = 24
    m = M.uniform_box((n, n, 1), (0, 0, -0.5), (1, 1, 0.5))
    om = oracle.OracleMesh(m.to_polymesh())
    h = 1.0 / n
    found = False
    for k in range(60):
        s = _solid(S.CIRCLE_TAIL, [2.2 * h, 4.0, 0.2 * h], center=(0.31 + 0.013 * k / n, 0.5 + 0.37 * h, 0.0),
                   q=S.quaternion_xyz((0, 0, 0.3 + 0.05 * k)))
        r = om.interact(np.array([s]), np.zeros((m.n_cells, 3)), capture=True)
        want, cnt = _brute_candidates(m, om, s)
        assert np.array_equal(np.sort(r["cells"]), want)
        if (cnt > 0).sum() > want.size:
            found = True
    assert found, "no configuration with a disconnected piece was exercised"


def test_seed_fallback_scans_in_index_order():
    n = 12
    m = M.uniform_box((n, n, n), (0, 0, 0), (1, 1, 1))
    om = oracle.OracleMesh(m.to_polymesh())
    h = 1.0 / n
    s = _solid(S.SPHERE, [1.3 * h, -3.5 * h, 0.0, 0.0], center=(0.5, 0.5, 0.5))
    r = om.interact(np.array([s]), np.zeros((m.n_cells, 3)), capture=True)
    near = om.nearest_cell(s["center"])
    assert near not in set(r["cells"].tolist())
    assert r["counts"][0, 3] == r["cells"].min()


def test_faithful_mode_equals_fast_mode():
    from sdfibm_amd import scenes
    sc = scenes.small_mixed_3d(16, seed=9)
    om = oracle.OracleMesh(sc.mesh.to_polymesh())
    a = om.interact(sc.solids, sc.U, rho_f=sc.rho_f, dt=sc.dt, mode=0, capture=True)
    b = om.interact(sc.solids, sc.U, rho_f=sc.rho_f, dt=sc.dt, mode=1, capture=True)
    for k in ("As", "Fs", "Ts", "Ct", "ft", "counts", "cells", "alpha"):
        assert np.array_equal(a[k], b[k]), k


def test_ct_last_solid_wins_and_clip():
    n = 16
    m = M.uniform_box((n, n, n), (0, 0, 0), (1, 1, 1))
    om = oracle.OracleMesh(m.to_polymesh())
    a = _solid(S.SPHERE, [0.2], center=(0.5, 0.5, 0.5))
    b = _solid(S.SPHERE, [0.15], center=(0.55, 0.5, 0.5))
    r = om.interact(np.array([a, b]), np.zeros((m.n_cells, 3)), capture=True)
    parts = split_candidates(r, 2)
    inside_b = parts[1][1][:parts[1][0][0]]
    assert np.all(r["Ct"][inside_b] == 1 + 4)
    assert np.all(r["As"] <= 1.0) and np.any(r["Ts"] > 1.0)
    cut_b = parts[1][1][parts[1][0][0]:]
    assert set(np.unique(r["Ct"][cut_b]).tolist()) <= {2.0, 3.0}


def test_rho_f_zero_when_off_fluid():
    from sdfibm_amd import scenes
    sc = scenes.small_mixed_3d(12, seed=2, n_each=1)
    om = oracle.OracleMesh(sc.mesh.to_polymesh())
    r = om.interact(sc.solids, sc.U, rho_f=0.0, dt=sc.dt)
    assert np.all(r["ft"] == 0.0)
    assert np.abs(r["Fs"]).max() > 0
