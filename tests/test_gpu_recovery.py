"""Recovery paths of the C ABI (needs an MI355X): a solid whose tile outgrows the on-chip tile
pass after set_solids, run_steps stopping at a failing step, and failed device allocations.

Each case compares against what a fresh context (or the same steps issued one call at a
time) gives: bit for bit, since both run the same kernels on the same inputs."""
from __future__ import annotations

import os

import numpy as np
import pytest

from sdfibm_amd import mesh as M
from sdfibm_amd import solids as S
from sdfibm_amd._lib import SdfibmError
from sdfibm_amd.cloud import SolidCloud

pytestmark = pytest.mark.gpu


def _long_ellipsoid_scene(n=(80, 80, 12)):
    """A 66-cell-long ellipsoid lying along y (its tile is 7 cells along x) among small
    spheres, on a uniform mesh of unit spacing scaled to the unit box along x."""
    h = 1.0 / n[0]
    m = M.uniform_box(n, (0, 0, 0), (1.0, n[1] * h, n[2] * h))
    rng = np.random.default_rng(5)
    sol = S.make_solids(6)
    sol["shape"] = S.SPHERE
    sol["prm"][:, 0] = 2.5 * h
    sol["center"] = [(0.15, 0.15, 6 * h), (0.85, 0.2, 6 * h), (0.2, 0.85, 6 * h), (0.8, 0.8, 6 * h),
                     (0.5, 0.5, 6 * h), (0.5, 0.5, 6 * h)]
    sol["shape"][5] = S.ELLIPSOID
    sol["prm"][5, :3] = (33.0 * h, 2.2 * h, 2.2 * h)
    sol["center"][5] = (0.5 + 0.3 * h, 0.5 + 0.1 * h, 6.2 * h)
    sol["q"] = (1.0, 0.0, 0.0, 0.0)
    sol["q"][5] = S.quaternion_xyz((0.0, 0.0, np.pi / 2))  # long axis along y
    sol["vel"] = rng.normal(0, 0.1, (6, 3))
    sol["omega"] = rng.normal(0, 1.0, (6, 3))
    sol["center"][4] = (0.3, 0.5, 6 * h)
    U = rng.normal(size=(m.n_cells, 3))
    return m, sol, U


def _same_results(a, b, fa, fb, ns):
    # force/torque: the cut-cell pass's workgroup size follows the mean tile size at set_solids,
    # which can differ between the two contexts -- only the order of the FP64 sums
    np.testing.assert_allclose(fa, fb, rtol=1e-12, atol=1e-14 * np.abs(fb).max())
    ga, gb = a.fields(), b.fields()
    for key in ("As", "Fs", "Ts", "Ct"):
        np.testing.assert_array_equal(ga[key], gb[key])
    for s in range(ns):
        ca, xa, aa = a.candidates(s)
        cb, xb, ab = b.candidates(s)
        assert np.array_equal(ca, cb) and np.array_equal(xa, xb) and np.array_equal(aa, ab), s


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("structured", [False, True])
def test_tile_outgrows_tile_pass_after_pose_update(prec, structured):
    """The ellipsoid turned to lie along x through the per-step pose update (same shapes: no
    re-classification at set_solids): its tile is now 67 cells along x, beyond the on-chip tile
    pass, and the step's classification sends it to the big-tile pass.  The result equals a fresh
    context's bit for bit."""
    m, sol, U = _long_ellipsoid_scene()
    turned = sol.copy()
    turned["q"][5] = (1.0, 0.0, 0.0, 0.0)
    with SolidCloud(m, precision=prec, force_structured=structured) as a, \
            SolidCloud(m, precision=prec, force_structured=structured) as b:
        a.set_solids(sol)
        a.set_velocity(U)
        a.interact(0.0, 1e-3)
        a.set_solids(turned)
        fa = a.interact(0.0, 1e-3)
        b.set_solids(turned)
        b.set_velocity(U)
        fb = b.interact(0.0, 1e-3)
        _same_results(a, b, fa, fb, sol.size)
        assert a.candidates(5)[0][:3].sum() > 500  # the long solid has its cells


def _state(c):
    s = c.solids()
    return np.concatenate([s["center"], s["q"], s["vel"], s["omega"]], axis=1), c.velocity()


def test_run_steps_keeps_last_valid_state():
    """The ellipsoid spins about z (no fluid force: rho_f = 0) so that its tile grows several
    times over during the graph replays, beyond the headroom run_steps sized the tile-code buffer
    with.  run_steps reports SDFIBM_ERR_CAPACITY, and the failing step and the steps after it
    leave the solids and U alone: they equal the state of a context that ran the same steps one
    run_steps call at a time (each call's eager pass grows the buffer) after some step j.  A second
    run_steps call grows the buffer and runs the remaining steps to the same final state."""
    m, sol, U = _long_ellipsoid_scene()
    sol["omega"] = 0.0
    sol["vel"] = 0.0
    sol["omega"][5] = (0.0, 0.0, 60.0)  # 0.06 rad per step: the tile quadruples within 15 steps
    dt, n = 1e-3, 40
    rho = np.full(sol.size, 2.0)
    with SolidCloud(m, rho_f=0.0, precision="fp64") as a, SolidCloud(m, rho_f=0.0, precision="fp64") as b:
        for c in (a, b):
            c.set_solids(sol)
            c.set_velocity(U)
            c.set_dynamics(rho)
        with pytest.raises(SdfibmError) as e:
            a.run_steps(0.0, dt, n)
        assert e.value.status == -8  # ERR_CAPACITY
        sa, ua = _state(a)
        hist = [_state(b)]
        for step in range(n):
            b.run_steps(step * dt, dt, 1)
            hist.append(_state(b))
        match = [j for j, (st, u) in enumerate(hist) if np.array_equal(st, sa) and np.array_equal(u, ua)]
        assert len(match) == 1 and 0 < match[0] < n, match
        j = match[0]
        a.run_steps(j * dt, dt, n - j)
        sa, ua = _state(a)
        np.testing.assert_array_equal(sa, hist[-1][0])
        np.testing.assert_array_equal(ua, hist[-1][1])


def test_failed_allocations_leave_a_clean_context():
    """Allocations refused by the failure-injection limit (SDFIBM_TEST_ALLOC_LIMIT) in set_solids,
    set_dynamics and set_contact: the calls return SDFIBM_ERR_NOMEM, the next interact / evolve
    return SDFIBM_ERR_STATE before launching anything, and the context works again (equal to a
    fresh one) once the allocations succeed."""
    n = 16
    h = 1.0 / n
    m = M.uniform_box((n, n, n), (0, 0, 0), (1, 1, 1))
    rng = np.random.default_rng(3)
    few = S.make_solids(4)
    few["shape"] = S.SPHERE
    few["prm"][:, 0] = 1.7 * h
    few["center"] = rng.uniform(0.2, 0.8, (4, 3))
    many = S.make_solids(40)
    many["shape"] = S.SPHERE
    many["prm"][:, 0] = 1.3 * h
    many["center"] = rng.uniform(0.15, 0.85, (40, 3))
    U = rng.normal(size=(m.n_cells, 3))
    key = "SDFIBM_TEST_ALLOC_LIMIT"
    try:
        with SolidCloud(m, precision="fp64") as c, SolidCloud(m, precision="fp64") as fresh:
            c.set_solids(few)
            c.set_velocity(U)
            c.interact(0.0, 1e-3)
            c.set_dynamics(np.full(4, 2.0))
            os.environ[key] = "0"
            with pytest.raises(SdfibmError) as e:
                c.set_solids(many)  # more solids than the buffers hold: reallocation refused
            assert e.value.status == -3
            for call in (lambda: c.interact(0.0, 1e-3), lambda: c.evolve(0.0, 1e-3)):
                with pytest.raises(SdfibmError) as e:
                    call()
                assert e.value.status == -6
            del os.environ[key]
            c.set_solids(many)
            fc = c.interact(0.0, 1e-3)
            fresh.set_solids(many)
            fresh.set_velocity(U)
            ff = fresh.interact(0.0, 1e-3)
            _same_results(c, fresh, fc, ff, many.size)
            # set_dynamics: the record buffer must grow from 4 to 40 solids
            os.environ[key] = "0"
            with pytest.raises(SdfibmError) as e:
                c.set_dynamics(np.full(40, 2.0))
            assert e.value.status == -3
            with pytest.raises(SdfibmError) as e:
                c.evolve(0.0, 1e-3)
            assert e.value.status == -6
            del os.environ[key]
            c.set_dynamics(np.full(40, 2.0))
            # contacts: the grid and sub-step buffers are refused
            os.environ[key] = "0"
            with pytest.raises(SdfibmError) as e:
                c.set_contact(10.0, 0.0)
            assert e.value.status == -3
            with pytest.raises(SdfibmError) as e:
                c.evolve(0.0, 1e-3)
            assert e.value.status == -6
            del os.environ[key]
            c.set_contact(10.0, 0.0)
            c.evolve(0.0, 1e-3)
    finally:
        os.environ.pop(key, None)
