"""Structured hex meshes for the coupling step.

The reference reads its mesh through OpenFOAM (src/meshinfo.h:22-29: cellCells,
cellPoints, points, cell centres, volumes, face centres, face areas).  The
device path works on *logically structured* hex meshes — one blockMesh block,
points numbered i-fastest, then j, then k, and cells likewise — whose
connectivity is implicit in (i, j, k) and therefore costs no HBM traffic.
Two flavours:

* ``rectilinear``: points are the tensor product of three 1-D coordinate
  arrays; cell centres, volumes and face data are closed forms evaluated on
  the device.
* ``general``: arbitrary point positions (e.g. jittered interior vertices);
  cell centres, volumes, face centres and face area vectors are SoA arrays
  computed once here, the way OpenFOAM's primitiveMesh computes them.

For the oracle (and for OpenFOAM interoperability) a structured mesh is also
exported as a polyMesh: internal faces in blockMesh's upper-triangular order,
then the boundary patches xmin, xmax, ymin, ymax, zmin, zmax; face vertex
order from OpenFOAM's hex cell model (owner-outward normals).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .foam_io import PolyMesh

# OpenFOAM hex model vertex numbering: 0 (0,0,0) 1 (1,0,0) 2 (1,1,0) 3 (0,1,0)
# 4 (0,0,1) 5 (1,0,1) 6 (1,1,1) 7 (0,1,1); model faces with outward normals.
HEX_OFFSETS = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                        (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)], dtype=np.int64)
HEX_FACES = {
    "xmin": (0, 4, 7, 3), "xmax": (1, 2, 6, 5),
    "ymin": (0, 1, 5, 4), "ymax": (3, 7, 6, 2),
    "zmin": (0, 3, 2, 1), "zmax": (4, 5, 6, 7),
}
PATCH_ORDER = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")

__all__ = ["StructuredHexMesh", "uniform_box", "jittered_box", "HEX_FACES", "PATCH_ORDER",
           "polymesh_geometry", "structured_from_polymesh"]


@dataclass
class StructuredHexMesh:
    nx: int
    ny: int
    nz: int
    _points: np.ndarray | None         # (Np, 3) float64, i-fastest (built lazily when rectilinear)
    xs: np.ndarray | None = None       # 1-D coordinates when rectilinear
    ys: np.ndarray | None = None
    zs: np.ndarray | None = None
    nominal: tuple | None = None       # (origin(3), spacing(3), max |displacement| in cells) for jittered meshes
    _geom: dict = field(default_factory=dict, repr=False)

    @property
    def points(self) -> np.ndarray:
        if self._points is None:
            Z, Y, X = np.meshgrid(self.zs, self.ys, self.xs, indexing="ij")
            self._points = np.stack([X.reshape(-1), Y.reshape(-1), Z.reshape(-1)], axis=1)
        return self._points

    @property
    def n_cells(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def n_points(self) -> int:
        return (self.nx + 1) * (self.ny + 1) * (self.nz + 1)

    @property
    def rectilinear(self) -> bool:
        return self.xs is not None

    def point_index(self, i, j, k):
        return i + (self.nx + 1) * (j + (self.ny + 1) * k)

    def cell_index(self, i, j, k):
        return i + self.nx * (j + self.ny * k)

    # -- polyMesh export ---------------------------------------------------------------------
    def to_polymesh(self) -> PolyMesh:
        nx, ny, nz = self.nx, self.ny, self.nz
        nc = self.n_cells
        ci = np.arange(nc, dtype=np.int64)
        i = ci % nx
        j = (ci // nx) % ny
        k = ci // (nx * ny)

        def face_pts(ii, jj, kk, model):
            cols = []
            for v in model:
                o = HEX_OFFSETS[v]
                cols.append(self.point_index(ii + o[0], jj + o[1], kk + o[2]))
            return np.stack(cols, axis=1)

        # internal faces, upper-triangular: per owner cell, neighbours ascending (+x, +y, +z)
        own_l, nei_l, pts_l = [], [], []
        ordkey_l = []
        for d, (cond, step, model) in enumerate((
                (i < nx - 1, 1, HEX_FACES["xmax"]),
                (j < ny - 1, nx, HEX_FACES["ymax"]),
                (k < nz - 1, nx * ny, HEX_FACES["zmax"]))):
            sel = np.nonzero(cond)[0]
            own_l.append(sel)
            nei_l.append(sel + step)
            pts_l.append(face_pts(i[sel], j[sel], k[sel], model))
            ordkey_l.append(sel * 3 + d)
        own = np.concatenate(own_l)
        nei = np.concatenate(nei_l)
        fpts = np.concatenate(pts_l)
        order = np.argsort(np.concatenate(ordkey_l), kind="stable")
        own, nei, fpts = own[order], nei[order], fpts[order]
        # boundary patches
        b_own, b_pts = [], []
        for name in PATCH_ORDER:
            cond = {"xmin": i == 0, "xmax": i == nx - 1, "ymin": j == 0, "ymax": j == ny - 1,
                    "zmin": k == 0, "zmax": k == nz - 1}[name]
            sel = np.nonzero(cond)[0]
            b_own.append(sel)
            b_pts.append(face_pts(i[sel], j[sel], k[sel], HEX_FACES[name]))
        all_own = np.concatenate([own] + b_own).astype(np.int32)
        all_pts = np.concatenate([fpts] + b_pts).astype(np.int32)
        nf = all_own.size
        offsets = np.arange(0, 4 * nf + 1, 4, dtype=np.int64)
        return PolyMesh(self.points, offsets, all_pts.reshape(-1), all_own, nei.astype(np.int32))

    # -- geometry (OpenFOAM primitiveMesh restatement, vectorised) ---------------------------
    def geometry(self) -> dict:
        """Cell centres/volumes and the three face families' centres/area vectors.

        Face family ``x`` holds the faces normal to i (index (i, j, k) with i in 0..nx, area
        vector pointing to +i), likewise ``y``/``z``.  Values follow OpenFOAM's primitiveMesh
        face and cell geometry for the hex-model vertex orders above.
        """
        if self._geom:
            return self._geom
        nx, ny, nz = self.nx, self.ny, self.nz
        P = self.points.reshape(nz + 1, ny + 1, nx + 1, 3)

        def quad(p0, p1, p2, p3):
            return _quad_geometry(p0, p1, p2, p3)

        # x-faces at i = 0..nx: vertices (i,j,k),(i,j+1,k),(i,j+1,k+1),(i,j,k+1) -> model xmax order
        fx_c, fx_s = quad(P[:-1, :-1, :], P[:-1, 1:, :], P[1:, 1:, :], P[1:, :-1, :])
        # y-faces at j: (i,j,k),(i,j,k+1),(i+1,j,k+1),(i+1,j,k) -> model ymax order (3 7 6 2)
        fy_c, fy_s = quad(P[:-1, :, :-1], P[1:, :, :-1], P[1:, :, 1:], P[:-1, :, 1:])
        # z-faces at k: (i,j,k),(i+1,j,k),(i+1,j+1,k),(i,j+1,k) -> model zmax order (4 5 6 7)
        fz_c, fz_s = quad(P[:, :-1, :-1], P[:, :-1, 1:], P[:, 1:, 1:], P[:, 1:, :-1])
        # shapes: fx (nz, ny, nx+1, 3), fy (nz, ny+1, nx, 3), fz (nz+1, ny, nx, 3)
        cc, vol = _hex_cell_geometry(fx_c, fx_s, fy_c, fy_s, fz_c, fz_s, nx, ny, nz)
        self._geom = dict(cc=cc.reshape(-1, 3), V=vol.reshape(-1),
                          fx_c=fx_c.reshape(-1, 3), fx_s=fx_s.reshape(-1, 3),
                          fy_c=fy_c.reshape(-1, 3), fy_s=fy_s.reshape(-1, 3),
                          fz_c=fz_c.reshape(-1, 3), fz_s=fz_s.reshape(-1, 3))
        return self._geom

    def cell_centres(self) -> np.ndarray:
        return self.geometry()["cc"]

    def volumes(self) -> np.ndarray:
        return self.geometry()["V"]


def _cross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], axis=-1)


def _dot(a, b):
    return a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2]


def _quad_geometry(p0, p1, p2, p3):
    """Face centre / area vector of quads (p0..p3 in face order), primitiveMesh style."""
    pts = (p0, p1, p2, p3)
    fcen = (((p0 + p1) + p2) + p3) / 4.0
    sumN = np.zeros_like(p0)
    for a_i in range(4):
        a, b = pts[a_i], pts[(a_i + 1) % 4]
        sumN = sumN + _cross(b - a, fcen - a)
    nmag = np.sqrt(_dot(sumN, sumN))
    nhat = sumN / nmag[..., None]
    sumA = np.zeros(p0.shape[:-1])
    sumAc = np.zeros_like(p0)
    for a_i in range(4):
        a, b = pts[a_i], pts[(a_i + 1) % 4]
        c = (a + b) + fcen
        an = _dot(_cross(b - a, fcen - a), nhat)
        sumA = sumA + an
        sumAc = sumAc + an[..., None] * c
    fc = ((1.0 / 3.0) * sumAc) / sumA[..., None]
    sf = 0.5 * sumN
    return fc, sf


def _hex_cell_geometry(fx_c, fx_s, fy_c, fy_s, fz_c, fz_s, nx, ny, nz):
    """Cell centre and volume by face-pyramid decomposition (primitiveMesh cell geometry).

    The per-cell face summation follows the cell's OpenFOAM face order: owned faces
    (+x, +y, +z internal; then boundary patches xmin, xmax, ymin, ymax, zmin, zmax), then
    neighbour faces (-z, -y, -x internal).
    """
    # per cell faces: centre, area vector oriented OUT of the cell
    xm_c, xm_s = fx_c[:, :, :-1], -fx_s[:, :, :-1]
    xp_c, xp_s = fx_c[:, :, 1:], fx_s[:, :, 1:]
    ym_c, ym_s = fy_c[:, :-1, :], -fy_s[:, :-1, :]
    yp_c, yp_s = fy_c[:, 1:, :], fy_s[:, 1:, :]
    zm_c, zm_s = fz_c[:-1], -fz_s[:-1]
    zp_c, zp_s = fz_c[1:], fz_s[1:]
    shape = (nz, ny, nx)
    I = np.broadcast_to(np.arange(nx)[None, None, :], shape)
    J = np.broadcast_to(np.arange(ny)[None, :, None], shape)
    K = np.broadcast_to(np.arange(nz)[:, None, None], shape)
    faces = _cell_face_order(I, J, K, nx, ny, nz)
    fam = {"-x": (xm_c, xm_s), "+x": (xp_c, xp_s), "-y": (ym_c, ym_s), "+y": (yp_c, yp_s),
           "-z": (zm_c, zm_s), "+z": (zp_c, zp_s)}
    # cEst: mean of face centres, summed in the cell's face order
    faces = [[(name, mask) for name, mask in slot if mask.any()] for slot in faces]
    cest = np.zeros(shape + (3,))
    for slot in faces:
        for name, mask in slot:
            np.add(cest, fam[name][0], out=cest, where=mask[..., None])
    cest = cest / 6.0
    ctr = np.zeros(shape + (3,))
    vol = np.zeros(shape)
    for name in fam:  # each face of a cell once: pyramid volume x3 and weighted centre
        c, s = fam[name]
        pyr3 = _dot(s, c - cest)
        fam[name] = (pyr3, pyr3[..., None] * ((3.0 / 4.0) * c + (1.0 / 4.0) * cest))
    for slot in faces:
        for name, mask in slot:
            pyr3, wc = fam[name]
            np.add(ctr, wc, out=ctr, where=mask[..., None])
            np.add(vol, pyr3, out=vol, where=mask)
    ctr = ctr / vol[..., None]
    vol = vol * (1.0 / 3.0)
    return ctr, vol


def _cell_face_order(I, J, K, nx, ny, nz):
    """The 6 face slots of each cell in OpenFOAM order, as a list of slots; each slot lists
    (face name, mask) alternatives that are mutually exclusive per cell."""
    ix_int_p, iy_int_p, iz_int_p = I < nx - 1, J < ny - 1, K < nz - 1
    ix_int_m, iy_int_m, iz_int_m = I > 0, J > 0, K > 0
    order = []
    # owned internal faces, ascending neighbour
    owned = [("+x", ix_int_p), ("+y", iy_int_p), ("+z", iz_int_p)]
    boundary = [("-x", ~ix_int_m), ("+x", ~ix_int_p), ("-y", ~iy_int_m), ("+y", ~iy_int_p),
                ("-z", ~iz_int_m), ("+z", ~iz_int_p)]
    neigh = [("-z", iz_int_m), ("-y", iy_int_m), ("-x", ix_int_m)]
    # A cell's list is the concatenation of present entries.  Build 6 slots by position.
    seq = owned + boundary + neigh
    present = [m for _, m in seq]
    pos = np.zeros(I.shape, dtype=np.int64)
    slots = [[] for _ in range(6)]
    for (name, m) in seq:
        for s in range(6):
            slots[s].append((name, m & (pos == s)))
        pos = pos + m.astype(np.int64)
    del present
    return slots


def uniform_box(n, lo, hi) -> StructuredHexMesh:
    """Uniform rectilinear box mesh with ``n = (nx, ny, nz)`` cells spanning ``lo..hi``."""
    nx, ny, nz = (int(v) for v in n)
    xs = np.linspace(lo[0], hi[0], nx + 1)
    ys = np.linspace(lo[1], hi[1], ny + 1)
    zs = np.linspace(lo[2], hi[2], nz + 1)
    return rectilinear(xs, ys, zs)


def rectilinear(xs, ys, zs) -> StructuredHexMesh:
    xs, ys, zs = (np.asarray(a, dtype=np.float64) for a in (xs, ys, zs))
    return StructuredHexMesh(xs.size - 1, ys.size - 1, zs.size - 1, None, xs, ys, zs)


def jittered_box(n, lo, hi, amplitude: float, seed: int) -> StructuredHexMesh:
    """Uniform box whose interior vertices are displaced uniformly within ±amplitude·h per axis
    (boundary vertices stay on the boundary planes)."""
    base = uniform_box(n, lo, hi)
    nx, ny, nz = base.nx, base.ny, base.nz
    h = np.array([(hi[d] - lo[d]) / n[d] for d in range(3)])
    rng = np.random.Generator(np.random.PCG64(seed))
    P = base.points.reshape(nz + 1, ny + 1, nx + 1, 3).copy()
    disp = rng.uniform(-amplitude, amplitude, size=P.shape) * h
    interior = [np.zeros((nz + 1, ny + 1, nx + 1), bool) for _ in range(3)]
    kk, jj, ii = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    for d, (idx, nd) in enumerate(((ii, nx), (jj, ny), (kk, nz))):
        interior[d] = (idx > 0) & (idx < nd)
        if nd == 1:
            interior[d][:] = False
    # a coordinate moves only where the vertex is not on a boundary plane normal to that axis
    for d in range(3):
        P[..., d] += np.where(interior[d], disp[..., d], 0.0)
    m = StructuredHexMesh(nx, ny, nz, P.reshape(-1, 3))
    m.nominal = (tuple(lo), tuple(h), float(amplitude))
    return m


def polymesh_geometry(pm: PolyMesh):
    """Face centres/areas and cell centres/volumes of a general polyMesh (numpy, slow path)."""
    nf = pm.n_faces
    fc = np.zeros((nf, 3))
    sf = np.zeros((nf, 3))
    counts = np.diff(pm.face_offsets)
    for n in np.unique(counts):
        sel = np.nonzero(counts == n)[0]
        lab = np.stack([pm.face_points[pm.face_offsets[sel] + t] for t in range(n)], axis=1)
        P = pm.points[lab]
        if n == 4:
            c, s = _quad_geometry(P[:, 0], P[:, 1], P[:, 2], P[:, 3])
        else:
            raise NotImplementedError("only quad faces are handled by the vectorised path")
        fc[sel], sf[sel] = c, s
    return fc, sf


def structured_from_polymesh(pm: PolyMesh, nx: int, ny: int, nz: int) -> StructuredHexMesh:
    """Wrap a polyMesh that blockMesh wrote for a single hex block (points i-fastest) as a
    general structured mesh, after checking its point numbering and internal-face order."""
    npts = (nx + 1) * (ny + 1) * (nz + 1)
    if pm.points.shape[0] != npts or pm.n_cells != nx * ny * nz:
        raise ValueError("polyMesh sizes do not match the block dimensions")
    m = StructuredHexMesh(nx, ny, nz, np.ascontiguousarray(pm.points))
    ref = m.to_polymesh()
    if not (np.array_equal(ref.owner, pm.owner) and np.array_equal(ref.neighbour, pm.neighbour)):
        raise ValueError("owner/neighbour do not follow single-block blockMesh ordering")
    nint = pm.neighbour.size
    a = ref.face_points.reshape(-1, 4)[:nint]
    b = pm.face_points.reshape(-1, 4)[:nint]
    if not np.array_equal(a, b):
        raise ValueError("internal face vertex order differs from the hex-model convention")
    return m
