"""Run-time variants of the step, the autotuner and the captured interact (needs an MI355X).

Every variant computes the same fields bit for bit (only the order of the FP64 force/torque sums
changes); the autotuner keeps one of them and reuses its cached choice; interact replays from a
captured HIP graph equal eager interacts bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

from sdfibm_amd import mesh as M
from sdfibm_amd import scenes
from sdfibm_amd import solids as S
from sdfibm_amd.cloud import SolidCloud

pytestmark = pytest.mark.gpu


def _mixed_big_scene():
    """Two big CSG solids (tiles beyond the on-chip pass), sixty small spheres and a plane."""
    n = 96
    h = 1.0 / n
    m = M.uniform_box((n, n, n), (0, 0, 0), (1, 1, 1))
    rng = np.random.default_rng(23)
    table = np.zeros(2, dtype=S.CSG_DTYPE)
    table[0] = scenes.csg_family(rng, 0, 70 * h)
    table[1] = scenes.csg_family(rng, 4, 66 * h)
    sol = S.make_solids(63)
    sol["shape"][:2] = S.CSG
    sol["prm"][:2, 0] = (0, 1)
    sol["center"][:2] = [(0.3, 0.35, 0.5), (0.68, 0.62, 0.45)]
    sol["shape"][2:62] = S.SPHERE
    sol["prm"][2:62, 0] = rng.uniform(1.5, 3.0, 60) * h
    sol["center"][2:62] = rng.uniform(0.05, 0.95, (60, 3))
    sol["shape"][62] = S.PLANE
    sol["center"][62] = (0.5, 0.1, 0.5)
    sol["q"][:62] = S.random_unit_quaternions(rng, 62)
    sol["vel"][:62] = rng.normal(0, 0.1, (62, 3))
    sol["omega"][:62] = rng.normal(0, 1.0, (62, 3))
    U = rng.normal(size=(m.n_cells, 3))
    return m, table, sol, U


def _run(c, table, sol, U):
    c.set_csg_shapes(table)
    c.set_solids(sol)
    c.set_velocity(U)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_variants_agree(prec):
    m, table, sol, U = _mixed_big_scene()
    with SolidCloud(m, precision=prec) as c:
        _run(c, table, sol, U)
        c.set_phase_timing(True)
        ref_ft = c.interact(0.0, 1e-3)
        ref = c.fields()
        for w, ch, g in ((0, 1024, 4), (1, 256, 2), (0, 4096, 8), (1, 64, 1)):
            c.set_tuning(w, ch, g)
            ft = c.interact(0.0, 1e-3)
            f = c.fields()
            for k in ("As", "Fs", "Ts", "Ct"):
                np.testing.assert_array_equal(f[k], ref[k])
            np.testing.assert_allclose(ft, ref_ft, rtol=1e-12, atol=1e-14 * np.abs(ref_ft).max())
        assert c.timing()[6] > 0  # the big-tile pass ran (second stream)


def test_autotune_and_cache(tmp_path):
    m, table, sol, U = _mixed_big_scene()
    path = str(tmp_path / "tune.jsonl")
    with SolidCloud(m, precision="fp32") as c:
        _run(c, table, sol, U)
        ref_ft = c.interact(0.0, 1e-3)
        ref = c.fields()
        a = c.autotune(1e-3, path)
        assert a["source"] == "timed" and a["variants_timed"] == 9 and a["ms"] > 0
        b = c.autotune(1e-3, path)
        assert b["source"] == "cache" and b["variants_timed"] == 0
        assert (a["cells_wave"], a["big_chunk"], a["big_grid"]) == (b["cells_wave"], b["big_chunk"], b["big_grid"])
        ft = c.interact(0.0, 1e-3)
        for k in ("As", "Fs", "Ts", "Ct"):
            np.testing.assert_array_equal(c.fields()[k], ref[k])
        np.testing.assert_allclose(ft, ref_ft, rtol=1e-12, atol=1e-14 * np.abs(ref_ft).max())


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_interact_graph_replays_equal_eager(prec):
    m, table, sol, U = _mixed_big_scene()
    with SolidCloud(m, precision=prec) as c:
        _run(c, table, sol, U)
        c.set_tuning(0, 1024, 4)
        ft_e = c.interact(0.0, 1e-3)
        fe = c.fields()
        for n in (1, 5, 2):
            ft_g = c.interact_steps(0.0, 1e-3, n)
            np.testing.assert_array_equal(ft_g, ft_e)
            fg = c.fields()
            for k in ("As", "Fs", "Ts", "Ct"):
                np.testing.assert_array_equal(fg[k], fe[k])
            assert c.timing()[3] > 0
            c.interact(0.0, 1e-3)  # eager work between the replays
        pairs = c.pair_counts()
        assert sum(pairs) > 100000


def test_phase_timing_switch():
    """Phase timing is off by default (no events between the kernels: zeros), on request the
    phases are timed, and the results are the same either way."""
    sc = scenes.small_mixed_3d(32, seed=3, n_each=3)
    with SolidCloud(sc.mesh, rho_f=sc.rho_f, precision="fp32") as c:
        c.set_solids(sc.solids)
        c.set_velocity(sc.U)
        ft0 = c.interact(0.0, sc.dt)
        f0 = c.fields()
        assert np.all(c.timing() == 0)
        c.set_phase_timing(True)
        ft1 = c.interact(0.0, sc.dt)
        t = c.timing()
        assert t[3] > 0 and t[0] > 0 and t[2] > 0
        np.testing.assert_array_equal(ft1, ft0)
        for k in ("As", "Fs", "Ts", "Ct"):
            np.testing.assert_array_equal(c.fields()[k], f0[k])
        c.set_phase_timing(False)
        c.interact(0.0, sc.dt)
        assert np.all(c.timing() == 0)
