"""The allocation-free velocity upload (sdfibm_set_velocity_cells_noalloc) against
sdfibm_set_velocity, through the C ABI (needs an MI355X).

Uploads U in pieces of at most the staging capacity while every library allocation is refused
(SDFIBM_TEST_ALLOC_LIMIT = 0), two calls in a row with the second smaller, and compares the
resident U and an interact on it with set_velocity's, bit for bit."""
from __future__ import annotations

import os

import numpy as np
import pytest

from sdfibm_amd import _lib
from sdfibm_amd import mesh as M
from sdfibm_amd import solids as S
from sdfibm_amd.cloud import SolidCloud

pytestmark = pytest.mark.gpu

MAXC = _lib.NOALLOC_MAX_CELLS


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_velocity_upload_without_allocation(prec):
    n = (112, 104, 96)  # 1,118,208 cells: more than the staging area holds
    m = M.uniform_box(n, (0, 0, 0), (1.0, 104 / 112, 96 / 112))
    N = m.n_cells
    assert N > MAXC
    rng = np.random.default_rng(17)
    U = rng.normal(size=(N, 3))
    sol = S.make_solids(3)
    sol["shape"] = S.SPHERE
    sol["prm"][:, 0] = (0.05, 0.03, 0.04)
    sol["center"] = [(0.3, 0.3, 0.4), (0.6, 0.5, 0.4), (0.5, 0.7, 0.3)]
    sol["vel"] = rng.normal(0, 0.1, (3, 3))
    key = "SDFIBM_TEST_ALLOC_LIMIT"
    with SolidCloud(m, precision=prec) as ref, SolidCloud(m, precision=prec) as c:
        ref.set_velocity(U)
        try:
            os.environ[key] = "0"  # any device allocation by the library now fails
            c.set_velocity_cells_noalloc(0, U[:MAXC])        # call 1: the largest piece
            c.set_velocity_cells_noalloc(MAXC, U[MAXC:])     # call 2: a smaller one, the rest
            with pytest.raises(_lib.SdfibmError) as e:
                c.set_velocity_cells_noalloc(0, U[:MAXC + 1])
            assert e.value.status == -1 and "staging area" in str(e.value)
            with pytest.raises(_lib.SdfibmError) as e:
                c.set_velocity_cells_noalloc(N - 10, U[:11])  # past the last cell
            assert e.value.status == -1
        finally:
            os.environ.pop(key, None)
        np.testing.assert_array_equal(c.velocity(), ref.velocity())
        # again, the second call smaller still: a few thousand cells in the middle change
        U2 = U.copy()
        U2[70000:76000] = rng.normal(size=(6000, 3))
        try:
            os.environ[key] = "0"
            c.set_velocity_cells_noalloc(69000, U2[69000:69000 + 20000])
            c.set_velocity_cells_noalloc(71000, U2[71000:74000])
        finally:
            os.environ.pop(key, None)
        ref.set_velocity(U2)
        np.testing.assert_array_equal(c.velocity(), ref.velocity())
        for x in (ref, c):
            x.set_solids(sol)
        fa, fb = c.interact(0.0, 1e-3), ref.interact(0.0, 1e-3)
        np.testing.assert_array_equal(fa, fb)
        ga, gb = c.fields(), ref.fields()
        for k in ("As", "Fs", "Ts", "Ct"):
            np.testing.assert_array_equal(ga[k], gb[k])
