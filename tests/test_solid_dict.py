"""solidDict reading (SolidCloud::initFromDictionary, src/solidcloud.cpp:43-205; the VOF tool's
reader) and the VOF initialiser's mesh handling.  CPU only; the GPU run of the VOF tool on the
reference's alpha.water fixture is in tests/test_gpu_vof.py."""
import math

import numpy as np
import pytest

from _helpers import toolvof_solids
from sdfibm_amd import mesh as M
from sdfibm_amd import solids as S
from sdfibm_amd.foam_dict import FoamDictError, parse_foam_dict
from sdfibm_amd.solid_dict import load_solid_dict
from sdfibm_amd.vof import block_dims

# The VOF example's dictionary (tool_vof/example/solidDict), restated: four shapes, a plane shape,
# twelve solids and two planes.
TOOLVOF_DICT = """
FoamFile { version 2.0; format ascii; class dictionary; object solidDict; }
meta { on_twod 1; }
shapes
{
    shape1 { name circle1; type Circle; radius 0.3; }
    shape2 { name circle_tail1; type Circle_Tail; radius 0.3; ratio 1; thickness 0.1; }
    shape3 { name ellipse1; type Ellipse; radiusa 0.3; radiusb 0.2; }   // ellipse
    shape4 { name rectangle1; type Rectangle; radiusa 0.3; radiusb 0.2; }
    shape5 { name plane1; type Plane; }
}
/* solids: three of each shape in rotated copies */
solids
{
    solid1  { shp_name circle_tail1; pos (0.5 1.5 0); euler (0 0 -45); }
    solid2  { shp_name circle_tail1; pos (0.5 2.5 0); euler (0 0 0); }
    solid3  { shp_name circle_tail1; pos (0.5 3.5 0); euler (0 0 45); }
    solid4  { shp_name rectangle1; pos (1.5 1.5 0); euler (0 0 0); }
    solid5  { shp_name rectangle1; pos (1.5 2.5 0); euler (0 0 30); }
    solid6  { shp_name rectangle1; pos (1.5 3.5 0); euler (0 0 60); }
    solid7  { shp_name ellipse1; pos (2.5 1.5 0); euler (0 0 0); }
    solid8  { shp_name ellipse1; pos (2.5 2.5 0); euler (0 0 60); }
    solid9  { shp_name ellipse1; pos (2.5 3.5 0); euler (0 0 120); }
    solid10 { shp_name circle1; pos (3.5 1.0 0); euler (0 0 0); }
    solid11 { shp_name circle1; pos (3.5 2.5 0); euler (0 0 0); }
    solid12 { shp_name circle1; pos (3.5 3.5 0); euler (0 0 0); }
}
planes
{
    plane1 { shp_name plane1; pos (0 0 0); euler (0 0 15); }
    plane2 { shp_name plane1; pos (0 0 0); euler (0 0 -90); }
}
"""

SOLVER_DICT = """
meta { on_fluid 1; on_twod 1; gravity (0.0 -10.0 0.0); }
shapes { shape1 { name c1; type Circle; radius 0.05; } shape2 { name p1; type Plane; } }
materials { material1 { name mat1; type General; rho 3.0; } }
motions
{
    motion1 { name free2d; type Motion01Mask; mask b110001; }
    motion2 { name fixed; type Motion01Mask; mask b000000; }
}
solids
{
    solid1 { shp_name c1; mat_name mat1; mot_name free2d; pos (0.3 1.2 0); vel (0.1 0 0); omega (0 0 2); }
    solid2 { shp_name c1; mat_name mat1; mot_name free; pos (0.6 1.4 0); euler (0 0 90); }
    wall1  { shp_name p1; mat_name mat1; mot_name fixed; pos (0.5 0 0); }
}
"""


def test_parse_foam_dict_grammar():
    d = parse_foam_dict('a 1; b (1 2 (3 4)); c { d e; "f" 2.5; } // x\n/* y */ g h i;')
    assert d == {"a": 1, "b": [1, 2, [3, 4]], "c": {"d": "e", "f": 2.5}, "g": ["h", "i"]}
    for bad in ("a 1", "a { b 1;", "a (1 2;", "}", "a 1 }"):
        with pytest.raises(FoamDictError):
            parse_foam_dict(bad)


def test_toolvof_dictionary_gives_the_fixture_solids():
    setup = load_solid_dict(TOOLVOF_DICT, is_text=True)
    ref = toolvof_solids()
    assert setup.two_d and setup.solids.size == ref.size == 14
    for key in ("shape", "prm", "center", "q"):
        np.testing.assert_array_equal(setup.solids[key], ref[key])
    assert setup.names[-1] == "planes.plane2"


def test_solver_dictionary_materials_motions():
    s = load_solid_dict(SOLVER_DICT, is_text=True)
    assert s.on_fluid and s.two_d
    np.testing.assert_array_equal(s.gravity, [0.0, -10.0, 0.0])
    np.testing.assert_array_equal(s.rho, [3.0, 3.0, 3.0])
    assert list(s.masks) == [0b100011, -1, 0]
    np.testing.assert_array_equal(s.solids["vel"][0], [0.1, 0, 0])
    np.testing.assert_array_equal(s.solids["omega"][0], [0, 0, 2])
    np.testing.assert_allclose(s.solids["q"][1], [math.cos(math.pi / 4), 0, 0, math.sin(math.pi / 4)], atol=1e-15)
    assert s.solids["shape"][2] == S.PLANE


@pytest.mark.parametrize("text,err", [
    ("shapes { s { name a; type Torus; radius 1; } }", S.UnknownShapeError),
    ("meta { on_twod 1; } shapes { s { name a; type Circle; radius 1; } } solids { x { shp_name a; pos (0 0 1); } }",
     ValueError),
    ("motions { m { name a; type Motion01Mask; mask b12; } }", ValueError),
    ("shapes { s { name a; type Circle; radius 1; } } solids { x { shp_name b; pos (0 0 0); } }", KeyError),
])
def test_dictionary_errors(text, err):
    with pytest.raises(err):
        load_solid_dict(text, is_text=True)


@pytest.mark.parametrize("n", [(5, 4, 1), (3, 6, 2), (3, 1, 4), (4, 1, 1)])
def test_block_dims_of_blockmesh_numbering(n):
    pm = M.uniform_box(n, (0, 0, 0), (1, 1, 1)).to_polymesh()
    ok = []
    for c in block_dims(pm):
        try:
            M.structured_from_polymesh(pm, *c)
            ok.append(c)
        except ValueError:
            pass
    assert ok == [n]
