"""Multi-block hex meshes from a blockMeshDict, numbered as OpenFOAM's blockMesh numbers them.

The reference runs on whatever mesh OpenFOAM's ``blockMesh`` writes for its examples (e.g.
examples/flow_past_cylinder/re200/system/blockMeshDict: nine graded blocks).  This module
restates what the coupling step depends on, so that such a mesh can be regenerated without
OpenFOAM and handed to the polyMesh path (``SolidCloud(PolyMesh)``):

* block points: for each block, (i, j, k) i-fastest, at graded parametric coordinates
  (``simpleGrading``: per direction an expansion ratio R = last / first cell size, cell sizes
  in geometric progression) mapped into the hex by trilinear interpolation of its 8 vertices
  (exact for the straight-edged blocks of the examples); ``arc`` edges are mapped by
  transfinite interpolation between the block's curved edges;
* point merging: points of different blocks at the same position become one point, numbered
  in order of first appearance over the blocks' point lists (blockMesh's merge list);
* cells: block by block, i-fastest, each with the hex model's vertex order;
* faces: internal faces upper-triangular (for each cell in order, the faces it shares with
  higher-numbered cells, by increasing neighbour), listed as the owner's model face; then the
  boundary patches in dictionary order, each patch's block faces in the order listed, the
  cells of a block face in increasing order, each face as its cell's model face (outward);
  boundary faces no patch names go last, in a patch ``defaultFaces``.

Supported: ``convertToMeters``/``scale``, ``vertices``, ``blocks`` with ``hex`` and
``simpleGrading`` (three ratios), ``edges`` with ``arc`` (a point on the arc), ``boundary``.
"""
from __future__ import annotations

import numpy as np

from .foam_dict import _tokens
from .foam_io import PolyMesh
from .mesh import HEX_FACES, HEX_OFFSETS

__all__ = ["BlockMesh", "read_block_mesh_dict", "block_mesh"]

# local (i, j, k) axis of each hex-model face and its side
_FACE_AXIS = {"xmin": (0, 0), "xmax": (0, 1), "ymin": (1, 0), "ymax": (1, 1), "zmin": (2, 0), "zmax": (2, 1)}
_FACE_ORDER = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")


class BlockMesh:
    """The parsed dictionary: vertices [n, 3], blocks [(labels[8], (nx, ny, nz), (gx, gy, gz))],
    arcs {(a, b): point}, patches [(name, type, [face labels[4], ...])]."""

    def __init__(self, vertices, blocks, arcs, patches):
        self.vertices = np.asarray(vertices, dtype=np.float64)
        self.blocks = blocks
        self.arcs = arcs
        self.patches = patches


def _parse_list(tok, i):
    """A parenthesised list starting at tok[i] == '(' -> (nested python list, next index)."""
    assert tok[i] == "("
    out, i = [], i + 1
    while tok[i] != ")":
        if tok[i] == "(":
            sub, i = _parse_list(tok, i)
            out.append(sub)
        else:
            out.append(tok[i])
            i += 1
    return out, i + 1


def read_block_mesh_dict(text: str) -> BlockMesh:
    tok = list(_tokens(text))
    scale = 1.0
    vertices, blocks, arcs, patches = [], [], {}, []
    i = 0
    while i < len(tok):
        t = tok[i]
        if t in ("convertToMeters", "scale"):
            scale = float(tok[i + 1])
            i += 3
        elif t == "FoamFile":
            while tok[i] != "}":
                i += 1
            i += 1
        elif t == "vertices":
            lst, i = _parse_list(tok, i + 1)
            vertices = [[float(x) for x in v] for v in lst]
            i += 1 if i < len(tok) and tok[i] == ";" else 0
        elif t == "blocks":
            lst, i = _parse_list(tok, i + 1)
            q = 0
            while q < len(lst):
                if lst[q] != "hex":
                    raise ValueError(f"block type {lst[q]!r} not supported (hex only)")
                labels = [int(x) for x in lst[q + 1]]
                q += 2
                if isinstance(lst[q], str):  # cell zone name
                    q += 1
                n = tuple(int(x) for x in lst[q])
                q += 1
                if lst[q] != "simpleGrading":
                    raise ValueError(f"grading {lst[q]!r} not supported (simpleGrading only)")
                g = tuple(float(x) for x in lst[q + 1])
                if len(g) != 3:
                    raise ValueError("simpleGrading needs three expansion ratios")
                q += 2
                blocks.append((labels, n, g))
            i += 1 if i < len(tok) and tok[i] == ";" else 0
        elif t == "edges":
            lst, i = _parse_list(tok, i + 1)
            q = 0
            while q < len(lst):
                if lst[q] != "arc":
                    raise ValueError(f"edge type {lst[q]!r} not supported (arc only)")
                a, b = int(lst[q + 1]), int(lst[q + 2])
                arcs[(a, b)] = np.array([float(x) for x in lst[q + 3]])
                q += 4
            i += 1 if i < len(tok) and tok[i] == ";" else 0
        elif t in ("boundary", "patches"):
            if t == "patches":
                raise ValueError("the old 'patches' syntax is not supported")
            assert tok[i + 1] == "("
            i += 2
            while tok[i] != ")":
                name = tok[i]
                assert tok[i + 1] == "{"
                i += 2
                ptype, faces = "patch", []
                while tok[i] != "}":
                    if tok[i] == "type":
                        ptype = tok[i + 1]
                        i += 3
                    elif tok[i] == "faces":
                        lst, i = _parse_list(tok, i + 1)
                        faces = [[int(x) for x in f] for f in lst]
                        i += 1 if tok[i] == ";" else 0
                    else:  # other entries (neighbourPatch ...): skip to ';'
                        while tok[i] != ";":
                            i += 1
                        i += 1
                patches.append((name, ptype, faces))
                i += 1
            i += 1
            i += 1 if i < len(tok) and tok[i] == ";" else 0
        else:
            i += 1
    return BlockMesh(np.asarray(vertices) * scale, blocks, {k: v * scale for k, v in arcs.items()}, patches)


def _graded(n: int, ratio: float) -> np.ndarray:
    """Parametric coordinates 0..1 of the n + 1 points along an edge of n cells whose last cell
    is `ratio` times the first."""
    if n == 1 or abs(ratio - 1.0) < 1e-12:
        return np.arange(n + 1, dtype=np.float64) / n
    r = ratio ** (1.0 / (n - 1))
    return (1.0 - r ** np.arange(n + 1)) / (1.0 - r ** n)


def _arc_points(p0, p1, pm, lam):
    """Points at parameters lam (0..1, by angle) on the circular arc from p0 through pm to p1."""
    a, b = pm - p0, p1 - p0
    n = np.cross(a, b)
    nn = np.dot(n, n)
    if nn < 1e-300:
        return p0[None, :] + lam[:, None] * b[None, :]
    # circumcentre of (p0, pm, p1)
    c = p0 + (np.dot(a, a) * np.cross(b, n) + np.dot(b, b) * np.cross(n, a)) / (2.0 * nn)
    u0, u1, um = p0 - c, p1 - c, pm - c
    R = np.linalg.norm(u0)
    e1 = u0 / R
    e2 = np.cross(n / np.sqrt(nn), e1)
    ang = lambda u: np.arctan2(np.dot(u, e2), np.dot(u, e1)) % (2 * np.pi)
    t1, tm = ang(u1), ang(um)
    if tm > t1:  # the arc passes pm: go the other way round
        t1 -= 2 * np.pi
    t = lam * t1
    return c[None, :] + R * (np.cos(t)[:, None] * e1[None, :] + np.sin(t)[:, None] * e2[None, :])


def _block_points(bm: BlockMesh, labels, n, g):
    """Points of one block, (i, j, k) i-fastest: trilinear interpolation of the 8 vertices at
    graded parameters, corrected by the curved edges (transfinite interpolation)."""
    V = bm.vertices[labels]  # hex model order
    nx, ny, nz = n
    u, v, w = _graded(nx, g[0]), _graded(ny, g[1]), _graded(nz, g[2])
    uu = u[None, None, :]
    vv = v[None, :, None]
    ww = w[:, None, None]
    P = np.zeros((nz + 1, ny + 1, nx + 1, 3))
    for t, (oi, oj, ok) in enumerate(HEX_OFFSETS):
        wt = (uu if oi else 1 - uu) * (vv if oj else 1 - vv) * (ww if ok else 1 - ww)
        P += wt[..., None] * V[t][None, None, None, :]
    # curved edges: add each edge's deviation from its chord, blended across the block
    edges = [((0, 1), 0), ((3, 2), 0), ((4, 5), 0), ((7, 6), 0),
             ((0, 3), 1), ((1, 2), 1), ((4, 7), 1), ((5, 6), 1),
             ((0, 4), 2), ((1, 5), 2), ((2, 6), 2), ((3, 7), 2)]
    params = (u, v, w)
    for (a, b), axis in edges:
        la, lb = labels[a], labels[b]
        key, rev = ((la, lb), False) if (la, lb) in bm.arcs else (((lb, la), True) if (lb, la) in bm.arcs else (None, False))
        if key is None:
            continue
        lam = params[axis]
        pts = _arc_points(bm.vertices[key[0]], bm.vertices[key[1]], bm.arcs[key], 1.0 - lam if rev else lam)
        chord = bm.vertices[la][None, :] * (1 - lam[:, None]) + bm.vertices[lb][None, :] * lam[:, None]
        dev = pts - chord  # along the edge's own parameter
        # the edge's fixed corner coordinates in the other two axes
        oa = HEX_OFFSETS[a]
        others = [d for d in range(3) if d != axis]
        wts = []
        for d in others:
            p = params[d]
            wts.append(p if oa[d] else 1 - p)
        shape = [nz + 1, ny + 1, nx + 1]
        blend = np.ones(shape)
        for d, wd in zip(others, wts):
            s = [1, 1, 1]
            s[2 - d] = wd.size
            blend = blend * wd.reshape(s)
        s = [1, 1, 1, 3]
        s[2 - axis] = dev.shape[0]
        P += blend[..., None] * dev.reshape(s)
    return P.reshape(-1, 3)


def block_mesh(bm: BlockMesh) -> tuple[PolyMesh, dict]:
    """The polyMesh of a parsed blockMeshDict, and info: block cell offsets, patch face
    ranges (name -> (start, size))."""
    # points of every block, merged by position
    span = float(np.ptp(bm.vertices, axis=0).max()) or 1.0
    tol = 1e-9 * span
    labels_of_block, pts = [], []
    index = {}
    for labels, n, g in bm.blocks:
        P = _block_points(bm, labels, n, g)
        keys = np.round(P / tol).astype(np.int64)
        lab = np.empty(P.shape[0], dtype=np.int64)
        for q in range(P.shape[0]):
            k = (keys[q, 0], keys[q, 1], keys[q, 2])
            got = index.get(k)
            if got is None:
                got = len(pts)
                index[k] = got
                pts.append(P[q])
            lab[q] = got
        labels_of_block.append(lab)
    points = np.asarray(pts)
    # cells: 8 point labels in hex-model order, block by block
    cells, offsets = [], [0]
    for (labels, n, g), lab in zip(bm.blocks, labels_of_block):
        nx, ny, nz = n
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        i, j, k = i.ravel(), j.ravel(), k.ravel()
        cols = [lab[(i + o[0]) + (nx + 1) * ((j + o[1]) + (ny + 1) * (k + o[2]))] for o in HEX_OFFSETS]
        cells.append(np.stack(cols, axis=1))
        offsets.append(offsets[-1] + nx * ny * nz)
    cells = np.concatenate(cells)
    nc = cells.shape[0]
    # faces of every cell (model order), matched by their sorted point labels
    model = np.array([HEX_FACES[f] for f in _FACE_ORDER])  # [6, 4]
    cf = cells[:, model]  # [nc, 6, 4]
    key = np.sort(cf, axis=2).reshape(-1, 4)
    order = np.lexsort(key.T[::-1])
    ks = key[order]
    same = np.all(ks[1:] == ks[:-1], axis=1)
    partner = np.full(nc * 6, -1, dtype=np.int64)  # the other (cell*6 + face) sharing the face
    idx = np.nonzero(same)[0]
    partner[order[idx]] = order[idx + 1]
    partner[order[idx + 1]] = order[idx]
    flat = np.arange(nc * 6)
    cell_of = flat // 6
    nb_cell = np.where(partner >= 0, partner // 6, -1)
    # internal faces: owner = lower cell; per owner ascending, neighbours ascending
    own_sel = np.nonzero((nb_cell > cell_of))[0]
    sortkey = cell_of[own_sel] * (nc + 1) + nb_cell[own_sel]
    own_sel = own_sel[np.argsort(sortkey, kind="stable")]
    owner = [cell_of[own_sel]]
    neighbour = nb_cell[own_sel]
    fpts = [cf.reshape(-1, 4)[own_sel]]
    # boundary faces by patch
    bnd = np.nonzero(nb_cell < 0)[0]
    taken = np.zeros(nc * 6, dtype=bool)
    # block faces: (block, model face) -> sorted vertex labels of the block face
    bf_lookup = {}
    for b, (labels, n, g) in enumerate(bm.blocks):
        for f in _FACE_ORDER:
            verts = tuple(sorted(labels[v] for v in HEX_FACES[f]))
            bf_lookup.setdefault(verts, []).append((b, f))
    info = {"block_cell_offsets": offsets, "patches": {}}
    start = len(neighbour)
    for name, ptype, faces in bm.patches:
        sel_all = []
        for fv in faces:
            hits = bf_lookup.get(tuple(sorted(fv)))
            if not hits:
                raise ValueError(f"patch {name}: face {fv} is not a block face")
            for b, f in hits:
                labels, n, g = bm.blocks[b]
                nx, ny, nz = n
                axis, side = _FACE_AXIS[f]
                k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
                ijk = (i.ravel(), j.ravel(), k.ravel())
                on = ijk[axis] == (n[axis] - 1 if side else 0)
                cl = offsets[b] + np.nonzero(on)[0]  # increasing
                slots = cl * 6 + _FACE_ORDER.index(f)
                slots = slots[nb_cell[slots] < 0]
                sel_all.append(slots)
        sel = np.concatenate(sel_all) if sel_all else np.zeros(0, dtype=np.int64)
        taken[sel] = True
        owner.append(cell_of[sel])
        fpts.append(cf.reshape(-1, 4)[sel])
        info["patches"][name] = (start, sel.size)
        start += sel.size
    rest = bnd[~taken[bnd]]
    if rest.size:
        owner.append(cell_of[rest])
        fpts.append(cf.reshape(-1, 4)[rest])
        info["patches"]["defaultFaces"] = (start, rest.size)
    owner = np.concatenate(owner).astype(np.int32)
    fpts = np.concatenate(fpts).astype(np.int32)
    pm = PolyMesh(points=points, face_offsets=np.arange(0, 4 * owner.size + 1, 4, dtype=np.int64),
                  face_points=fpts.ravel(), owner=owner, neighbour=neighbour.astype(np.int32))
    return pm, info
