"""The coupled step loop as replays of one captured HIP graph (sdfibm_run_steps) against the
same steps issued one call at a time (interact, update_velocity, evolve, fix_internal: the
reference's loop, src/main.cpp:66-84, without the PISO solve).  Both run the same kernels on the
same inputs in the same order: solids, U and the fields agree bit for bit.  Replays are
interleaved with eager work and checked after every batch."""
from __future__ import annotations

import numpy as np
import pytest

from sdfibm_amd import scenes
from sdfibm_amd import solids as S
from sdfibm_amd._lib import SdfibmError
from sdfibm_amd.cloud import SolidCloud, motion_mask

pytestmark = pytest.mark.gpu


def _state(sol):
    return np.concatenate([sol["center"], sol["q"], sol["vel"], sol["omega"]], axis=1)


def _walls(inset):
    """Four fixed planes facing into the unit cube, `inset` inside its faces (x and y)."""
    w = S.make_solids(4)
    w["shape"] = S.PLANE
    for t, (c, deg) in enumerate((((0.5, inset, 0.5), 0.0), ((0.5, 1 - inset, 0.5), 180.0),
                                  ((inset, 0.5, 0.5), -90.0), ((1 - inset, 0.5, 0.5), 90.0))):
        w["center"][t] = c
        w["q"][t] = S.quaternion_xyz((0.0, 0.0, np.radians(deg)))
    return w


def _setup(c, sol, U, rho, masks, contact):
    c.set_solids(sol)
    c.set_velocity(U)
    c.set_dynamics(rho, masks)
    if contact:
        c.set_contact(*contact)


@pytest.mark.parametrize("jitter,prec,contact", [(0.0, "fp32", None), (0.2, "fp64", (40.0, 0.2)),
                                                 (0.0, "fp64", (40.0, 0.0))])
def test_graph_replays_match_eager_steps(jitter, prec, contact):
    sc = scenes.small_mixed_3d(24, seed=4, n_each=2, jitter=jitter)
    sol = np.concatenate([sc.solids[np.isin(sc.solids["shape"], [S.SPHERE, S.ELLIPSOID, S.BOX])], _walls(0.06)])
    ns = sol.size
    rho = np.full(ns, 1.8)
    masks = np.full(ns, -1, dtype=np.int32)
    masks[-4:] = motion_mask("b000000")
    sol["vel"][-4:] = 0.0
    sol["omega"][-4:] = 0.0
    g = np.array([0.0, -9.81, 0.0])
    dt = 1e-3
    with SolidCloud(sc.mesh, rho_f=sc.rho_f, precision=prec) as eager, \
            SolidCloud(sc.mesh, rho_f=sc.rho_f, precision=prec) as graph:
        _setup(eager, sol, sc.U, rho, masks, contact)
        _setup(graph, sol, sc.U, rho, masks, contact)
        t = 0.0
        for batch in (2, 3, 1):
            for _ in range(batch):
                ft_e = eager.interact(t, dt)
                eager.update_velocity(dt)
                eager.evolve(t, dt, g)
                eager.fix_internal()
                t += dt
            ft_g = graph.run_steps(t, dt, batch, g)
            np.testing.assert_array_equal(ft_g, ft_e)
            np.testing.assert_array_equal(_state(graph.solids()), _state(eager.solids()))
            np.testing.assert_array_equal(graph.velocity(), eager.velocity())
            fe, fg = eager.fields(), graph.fields()
            for key in ("As", "Fs", "Ts", "Ct"):
                np.testing.assert_array_equal(fg[key], fe[key])
            # eager work on the graph context between replays
            graph.interact(t, dt)
        moved = graph.solids()
        assert np.any(moved["center"][:-4] != sol["center"][:-4])
        np.testing.assert_array_equal(moved["center"][-4:], sol["center"][-4:])


def test_moving_plane_is_refused():
    sc = scenes.small_mixed_3d(16, seed=2, n_each=1)
    sol = np.concatenate([sc.solids[:3], _walls(0.1)])
    sol["vel"][-1] = (0.1, 0.0, 0.0)
    ns = sol.size
    with SolidCloud(sc.mesh, rho_f=1.0, precision="fp64") as c:
        c.set_solids(sol)
        c.set_velocity(sc.U)
        c.set_dynamics(2.0)  # no motion constraint on the planes
        with pytest.raises(SdfibmError, match="fixed plane"):
            c.run_steps(0.0, 1e-3, 2, (0, -9.81, 0))
        masks = np.full(ns, -1, dtype=np.int32)
        masks[-4:] = 0
        c.set_dynamics(2.0, masks)
        with pytest.raises(SdfibmError, match="fixed plane"):  # still has a velocity
            c.run_steps(0.0, 1e-3, 2, (0, -9.81, 0))


def test_pair_counts_total_the_candidate_sets():
    sc = scenes.small_mixed_3d(24, seed=6, n_each=2)
    sol = np.concatenate([sc.solids, _walls(0.08)])
    with SolidCloud(sc.mesh, rho_f=1.0, precision="fp64") as c:
        c.set_solids(sol)
        c.set_velocity(sc.U)
        c.interact(0.0, 1e-3)
        inside = cut = 0
        for s in range(sol.size):
            counts, _, _ = c.candidates(s)
            inside += int(counts[0])
            cut += int(counts[1] + counts[2])
        assert c.pair_counts() == (inside, cut)
        assert inside > 0 and cut > 0
