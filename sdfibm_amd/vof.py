"""VOF initialiser: the volume fraction of a solidDict's solids and planes on an OpenFOAM case's
mesh, written as a volScalarField (the reference's tool_vof, ``SolidCloud::writeVOF``,
tool_vof/solidcloud.cpp:141-173: clip(sum of fractions) per cell, no forcing).

    python -m sdfibm_amd.vof CASE_DIR [--dict solidDict] [--field alpha.water] [--ascii]

The mesh is read from CASE_DIR/constant/polyMesh: one blockMesh block (points and cells
i-fastest) runs on the structured path, any other mesh on the general polyMesh path; the
fractions run on the GPU through the coupling library in FP64.
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

from . import foam_io
from . import mesh as M
from .cloud import SolidCloud
from .solid_dict import load_solid_dict

__all__ = ["block_dims", "mesh_from_case", "vof_field", "main"]


def block_dims(pm) -> list:
    """Candidate (nx, ny, nz) of a single-block blockMesh polyMesh: cell 0 owns its +x, +y and +z
    internal faces, whose neighbours are cells 1, nx and nx*ny; with only two such faces the block
    is one cell thick in y or in z (both are candidates)."""
    nint = pm.neighbour.size
    own0 = np.nonzero(pm.owner[:nint] == 0)[0]
    nb = sorted(int(pm.neighbour[f]) for f in own0)
    n = pm.n_cells
    if len(nb) == 3:
        nx = nb[1]
        ny = nb[2] // nx
        return [(nx, ny, n // (nx * ny))]
    if len(nb) == 2:
        nx = nb[1]
        return [(nx, n // nx, 1), (nx, 1, n // nx)]
    return [(n, 1, 1)]


def mesh_from_case(case_dir: str):
    """The case's mesh: a StructuredHexMesh when it is one blockMesh block, else the PolyMesh."""
    pm = foam_io.read_polymesh(case_dir)
    for nx, ny, nz in block_dims(pm):
        if nx * ny * nz != pm.n_cells:
            continue
        try:
            return M.structured_from_polymesh(pm, nx, ny, nz)
        except ValueError:  # numbering of another block shape
            pass
    return pm


def vof_field(mesh, solids: np.ndarray, two_d: bool, precision: str = "fp64") -> np.ndarray:
    """clip(sum of the solids' cell fractions, 1) per cell."""
    with SolidCloud(mesh, two_d=two_d, rho_f=0.0, precision=precision) as c:
        c.set_solids(solids)
        c.set_velocity(np.zeros((mesh.n_cells, 3)))
        c.interact(0.0, 1.0)
        return c.fields()["As"]


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m sdfibm_amd.vof", description=__doc__.splitlines()[0])
    ap.add_argument("case")
    ap.add_argument("--dict", default="solidDict")
    ap.add_argument("--field", default="alpha.water")
    ap.add_argument("--ascii", action="store_true")
    a = ap.parse_args(argv)
    setup = load_solid_dict(os.path.join(a.case, a.dict))
    mesh = mesh_from_case(a.case)
    alpha = vof_field(mesh, setup.solids, setup.two_d)
    os.makedirs(os.path.join(a.case, "0"), exist_ok=True)
    out = os.path.join(a.case, "0", a.field)
    foam_io.write_scalar_field(out, alpha, a.field, binary=not a.ascii)
    print(f"wrote {out}: {mesh.n_cells} cells")
    return 0


if __name__ == "__main__":
    sys.exit(main())
