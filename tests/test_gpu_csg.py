"""CSG solids and the big-tile pass on the GPU against the CPU oracle (needs an MI355X).

CSG shapes (unions, differences and intersections of boxes, cylinders, ellipsoids and spheres)
small enough for the LDS tile pass, and ones whose tiles exceed it (more than 63 cells along x)
and take the big-tile pass; both precisions, both mesh kinds.  Tolerances as in
test_gpu_parity.py: candidate lists bit-exact; FP64 fractions 1e-12, Fs and force/torque 1e-9
relative; FP32 fractions 1e-6, Fs 1e-5, force/torque 1e-4 relative."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from _helpers import split_candidates
from sdfibm_amd import mesh as M
from sdfibm_amd import scenes
from sdfibm_amd import solids as S
from sdfibm_amd.cloud import SolidCloud
from test_gpu_parity import TOL

pytestmark = pytest.mark.gpu


def _scene(n, sizes_cells, seed, jitter=0.0):
    """CSG solids of the six families (sizes in cells) at random poses, overlapping allowed."""
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    mesh = M.jittered_box((n, n, n), (0, 0, 0), (1, 1, 1), jitter, seed) if jitter > 0 else \
        M.uniform_box((n, n, n), (0, 0, 0), (1, 1, 1))
    table = np.zeros(len(sizes_cells), dtype=S.CSG_DTYPE)
    for b, D in enumerate(sizes_cells):
        table[b] = scenes.csg_family(rng, b % 6, D * h)
    sol = S.make_solids(len(sizes_cells))
    sol["shape"] = S.CSG
    sol["prm"][:, 0] = np.arange(len(sizes_cells))
    sol["center"] = rng.uniform(0.3, 0.7, (len(sizes_cells), 3))
    sol["q"] = S.random_unit_quaternions(rng, len(sizes_cells))
    sol["vel"] = rng.normal(0, 0.1, (len(sizes_cells), 3))
    sol["omega"] = rng.normal(0, 1.0, (len(sizes_cells), 3))
    U = rng.normal(size=(mesh.n_cells, 3))
    return mesh, table, sol, U


def _check(mesh, table, sol, U, prec, structured=False):
    oracle.set_csg(table)
    om = oracle.OracleMesh(mesh.to_polymesh())
    ref = om.interact(sol, U, rho_f=1.2, dt=1e-3, capture=True, capacity=4 * mesh.n_cells)
    tol = TOL[prec]
    with SolidCloud(mesh, rho_f=1.2, precision=prec, force_structured=structured) as c:
        c.set_csg_shapes(table)
        c.set_solids(sol)
        c.set_velocity(U)
        ft = c.interact(0.0, 1e-3)
        got = c.fields()
        parts = split_candidates(ref, sol.size)
        for s in range(sol.size):
            counts, cells, alpha = c.candidates(s)
            rc, rcells, ralpha = parts[s]
            assert np.array_equal(counts, rc), (s, counts, rc)
            assert np.array_equal(cells, rcells), s
            assert np.max(np.abs(alpha - ralpha), initial=0.0) <= tol["a"], (s, np.max(np.abs(alpha - ralpha)))
        assert np.array_equal(got["Ct"], ref["Ct"])
        assert np.max(np.abs(got["As"] - ref["As"])) <= tol["a"]
        assert np.max(np.abs(got["Fs"] - ref["Fs"])) <= tol["f"] * np.abs(ref["Fs"]).max()
        scale = np.maximum(np.abs(ref["ft"]), tol["floor"] * np.abs(ref["ft"]).max())
        assert np.max(np.abs(ft - ref["ft"]) / scale) <= tol["ft"]
        return c.pair_counts()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("structured", [False, True])
def test_csg_small_tiles(prec, structured):
    """Twelve CSG solids of 8-20 cells (the LDS tile pass)."""
    mesh, table, sol, U = _scene(32, np.linspace(8, 20, 12), seed=3)
    _check(mesh, table, sol, U, prec, structured)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_csg_big_tiles(prec):
    """Three CSG solids of 66-74 cells on 96^3 (tiles beyond 63 cells along x: the big-tile
    pass) with two small ones; the cavity and the cross exercise disconnected rows."""
    mesh, table, sol, U = _scene(96, [70.0, 74.0, 66.0, 10.0, 12.0], seed=5)
    # families 0, 1, 2 big, 3 and 4 small; put the small ones away from the middle
    sol["center"][3] = (0.1, 0.1, 0.12)
    sol["center"][4] = (0.88, 0.9, 0.1)
    counts = _check(mesh, table, sol, U, prec)
    assert sum(counts) > 50000


def test_csg_big_tiles_jittered():
    """A big CSG solid (the cavity family) and a big ellipsoid on a jittered 80^3 mesh (the
    structured mesh kind), FP64."""
    mesh, table, sol, U = _scene(80, [70.0, 10.0, 10.0, 10.0, 72.0], seed=9, jitter=0.15)
    sol = sol[[0, 4]]
    sol["center"] = [(0.45, 0.5, 0.55), (0.52, 0.48, 0.5)]
    sol["shape"][1] = S.ELLIPSOID
    sol["prm"][1, :3] = (0.42, 0.2, 0.15)
    _check(mesh, table, sol, U[: mesh.n_cells], "fp64", structured=True)
