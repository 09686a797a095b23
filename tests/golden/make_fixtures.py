"""Write the golden fixtures under tests/golden/ from the data files the reference ships.

Run in a container where the reference checkout is mounted at /root/reference:

    python tests/golden/make_fixtures.py

Only data files are read (OpenFOAM binary lists), never code.  Outputs:

toolvof.npz      tool_vof/example: mesh points (80,802 x 3), alpha.water (40,000), and
                 SHA-256 digests of the mesh topology: internal faces/owners/neighbours in
                 file order, and per boundary patch the (owner, face points) rows sorted by
                 owner (the tests regenerate the topology and compare digests).  alpha.water is the
                 reference's smooth VOF field, clip(sum of As) over the solids and planes
                 of tool_vof/example/solidDict.
cylinder_block5.npz
                 examples/flow_past_cylinder/re200/0/As restricted to blockMeshDict block 5
                 (the uniform 120 x 120 block on [-6,6]^2 holding the circle r = 1): the
                 14,400 values and the count of non-zero values outside the block (0).
cylinder_re200.npz
                 the same case whole: all 78,400 values of 0/As, and the numbers of its
                 system/blockMeshDict (32 vertices; 9 hex blocks: vertex labels, cell counts,
                 simpleGrading ratios; the boundary patches' names, types and block faces),
                 from which the tests regenerate the nine-block mesh.
"""
from __future__ import annotations

import hashlib
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from sdfibm_amd import foam_io  # noqa: E402

REF = "/root/reference"


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    case = os.path.join(REF, "tool_vof", "example")
    pm = foam_io.read_polymesh(case)
    alpha = foam_io.read_scalar_field(os.path.join(case, "0", "alpha.water"))
    nint = pm.neighbour.size
    faces = pm.face_points.reshape(-1, 4).astype(np.int32)
    # boundary patches (startFace, nFaces) in file order; faces of a patch sorted by owner
    txt = open(os.path.join(case, "constant", "polyMesh", "boundary"), "rb").read().decode("latin1")
    starts = [int(x) for x in re.findall(r"startFace\s+(\d+);", txt)]
    sizes = [int(x) for x in re.findall(r"nFaces\s+(\d+);", txt)]
    bparts = []
    for st, nf in zip(starts, sizes):
        sl = slice(st, st + nf)
        order = np.argsort(pm.owner[sl], kind="stable")
        bparts.append(np.concatenate([pm.owner[sl][order][:, None].astype(np.int32), faces[sl][order]], axis=1))
    np.savez_compressed(
        os.path.join(HERE, "toolvof.npz"),
        points=pm.points, alpha=alpha,
        internal_faces_sha256=digest(faces[:nint]),
        internal_owner_sha256=digest(pm.owner[:nint].astype(np.int32)),
        neighbour_sha256=digest(pm.neighbour.astype(np.int32)),
        boundary_sha256=digest(np.concatenate(bparts)),
        patch_starts=np.array(starts), patch_sizes=np.array(sizes),
        dims=np.array([200, 200, 1]))
    As = foam_io.read_scalar_field(os.path.join(REF, "examples", "flow_past_cylinder", "re200", "0", "As"))
    blk = As[20800:20800 + 14400]
    outside = int(np.count_nonzero(As) - np.count_nonzero(blk))
    np.savez_compressed(os.path.join(HERE, "cylinder_block5.npz"), As=blk, outside_nonzero=outside,
                        n_total=As.size, offset=20800, dims=np.array([120, 120, 1]),
                        lo=np.array([-6.0, -6.0, -0.5]), hi=np.array([6.0, 6.0, 0.5]))
    from sdfibm_amd import blockmesh  # noqa: E402
    bm = blockmesh.read_block_mesh_dict(open(os.path.join(REF, "examples", "flow_past_cylinder", "re200", "system",
                                                          "blockMeshDict")).read())
    np.savez_compressed(os.path.join(HERE, "cylinder_re200.npz"), As=As, vertices=bm.vertices,
                        block_labels=np.array([b[0] for b in bm.blocks], dtype=np.int32),
                        block_cells=np.array([b[1] for b in bm.blocks], dtype=np.int32),
                        block_grading=np.array([b[2] for b in bm.blocks], dtype=np.float64),
                        patch_names=np.array([p[0] for p in bm.patches]),
                        patch_types=np.array([p[1] for p in bm.patches]),
                        patch_nfaces=np.array([len(p[2]) for p in bm.patches], dtype=np.int32),
                        patch_faces=np.array([f for p in bm.patches for f in p[2]], dtype=np.int32))
    print("wrote toolvof.npz, cylinder_block5.npz, cylinder_re200.npz")


if __name__ == "__main__":
    main()
