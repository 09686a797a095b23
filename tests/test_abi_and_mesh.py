"""C-ABI library symbols (no GPU calls) and the mesh layer (CPU)."""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
from sdfibm_amd import _lib, foam_io
from sdfibm_amd import mesh as M
from sdfibm_amd import solids as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(sdfibm_\w+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    declared = _declared_functions()
    assert declared, "no declarations found"
    assert declared == set(_lib.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name


def test_status_strings_and_version():
    L = _lib.load()
    assert L.sdfibm_abi_version() == 1
    for code in range(0, -10, -1):
        assert L.sdfibm_status_string(code)
    assert L.sdfibm_status_string(-99) == b"unknown status"


def test_solid_record_layout():
    assert S.SOLID_DTYPE.itemsize == 160
    assert S.SOLID_DTYPE.fields["center"][1] == 56
    assert S.SOLID_DTYPE.fields["q"][1] == 80
    assert S.SOLID_DTYPE.fields["omega"][1] == 136


def test_uniform_box_geometry_closed_form():
    m = M.uniform_box((3, 4, 5), (0, 0, 0), (1.5, 2.0, 2.5))
    g = m.geometry()
    assert np.allclose(g["V"], 0.125, rtol=1e-14)
    ijk = np.array([(i, j, k) for k in range(5) for j in range(4) for i in range(3)], dtype=float)
    assert np.allclose(g["cc"], (ijk + 0.5) * 0.5, atol=1e-14)
    assert np.allclose(np.abs(g["fx_s"][:, 0]), 0.25)


def test_jittered_mesh_geometry_matches_oracle_and_fills_the_box():
    m = M.jittered_box((6, 5, 4), (0, 0, 0), (1, 1, 1), 0.2, seed=3)
    g = m.geometry()
    assert abs(g["V"].sum() - 1.0) < 1e-13  # boundary points stay on the box faces
    om = oracle.OracleMesh(m.to_polymesh())
    cc, V, fc, sf = om.geometry()
    assert np.allclose(cc, g["cc"], rtol=0, atol=1e-15)
    assert np.allclose(V, g["V"], rtol=1e-13, atol=0)


def test_jittered_mesh_geometry_invariants():
    """The face/cell geometry of non-planar faces (primitiveMesh's triangle fans): every cell's
    outward face area vectors sum to zero (closed surface) and the cell volumes fill the box.
    (The volume-weighted centres do not reproduce the box's first moment exactly: primitiveMesh
    treats each face as planar with its area vector and centre in the pyramid centroids; on this
    mesh the first moment is off by 3e-4 relative, which is the method, not an error.)"""
    n = (7, 6, 5)
    m = M.jittered_box(n, (0, 0, 0), (1, 2, 3), 0.3, seed=5)
    g = m.geometry()
    nx, ny, nz = n
    fx = g["fx_s"].reshape(nz, ny, nx + 1, 3)
    fy = g["fy_s"].reshape(nz, ny + 1, nx, 3)
    fz = g["fz_s"].reshape(nz + 1, ny, nx, 3)
    closure = (fx[:, :, 1:] - fx[:, :, :-1]) + (fy[:, 1:] - fy[:, :-1]) + (fz[1:] - fz[:-1])
    assert np.abs(closure).max() < 1e-15
    V, cc = g["V"], g["cc"]
    assert abs(V.sum() - 6.0) < 1e-12
    np.testing.assert_allclose((V[:, None] * cc).sum(axis=0), 6.0 * np.array([0.5, 1.0, 1.5]), rtol=1e-3)
    # the jitter is real: faces are not planar
    assert np.abs(g["fx_s"][:, 1:]).max() > 1e-3


def test_foam_scalar_field_roundtrip(tmp_path):
    v = np.random.default_rng(0).random(1000)
    for binary in (True, False):
        p = tmp_path / f"As_{binary}"
        foam_io.write_scalar_field(str(p), v, "As", binary=binary)
        w = foam_io.read_scalar_field(str(p))
        assert np.array_equal(v, w) if binary else np.allclose(v, w, rtol=0, atol=0)


def test_shape_registry():
    sid, prm = S.make_shape("Ellipsoid", radiusa=1, radiusb=2, radiusc=3)
    assert sid == S.ELLIPSOID and list(prm[:3]) == [1, 2, 3]
    with pytest.raises(S.UnknownShapeError):
        S.make_shape("Torus", radius=1)
    with pytest.raises(ValueError):
        S.make_shape("Sphere", radius=-1)
    q = S.quaternion_xyz((0.0, 0.0, np.pi / 3))
    assert np.allclose(q, [np.cos(np.pi / 6), 0, 0, np.sin(np.pi / 6)])
