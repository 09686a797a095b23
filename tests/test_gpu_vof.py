"""The VOF initialiser (tool_vof's writeVOF, tool_vof/solidcloud.cpp:141-173) on the GPU against
the field the reference wrote, tool_vof/example/0/alpha.water (tests/golden/toolvof.npz), from
the dictionary text restated in test_solid_dict.py."""
import os
import tempfile

import numpy as np
import pytest

from _helpers import toolvof_mesh
from sdfibm_amd import foam_io
from sdfibm_amd.solid_dict import load_solid_dict
from sdfibm_amd.vof import vof_field
from test_solid_dict import TOOLVOF_DICT

pytestmark = pytest.mark.gpu


def test_vof_tool_reproduces_alpha_water():
    mesh, g = toolvof_mesh()
    setup = load_solid_dict(TOOLVOF_DICT, is_text=True)
    alpha = vof_field(mesh, setup.solids, setup.two_d)
    assert np.max(np.abs(alpha - g["alpha"])) <= 1e-11
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "alpha.water")
        foam_io.write_scalar_field(path, alpha, "alpha.water")
        np.testing.assert_array_equal(foam_io.read_scalar_field(path), alpha)
