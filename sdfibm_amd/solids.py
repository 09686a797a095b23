"""Solid records and the shape registry for the coupling step.

The reference builds shapes through a name → creator registry
(src/genericfactory.h:18-40, src/libshape/shapefactory.h:11-39) from solidDict
``shapes`` entries, and each ``Solid`` points at one shape (src/solid.h:47-50).
Across the C ABI a solid is one flat record: the shape id, its parameters in
solidDict units, and the rigid-body state (centre, orientation quaternion,
velocity, angular velocity).  This module is the host-side registry: it maps
the reference's shape type names to ids, validates parameters, and turns
solidDict-style entries into records.
"""
from __future__ import annotations

import math

import numpy as np

# Shape ids (identical to SDFIBM_SHAPE_* in include/sdfibm_hip.h)
SPHERE, ELLIPSOID, CIRCLE, ELLIPSE, BOX, RECTANGLE, PLANE, CIRCLE_TAIL, CIRCLE_TWOTAIL, CSG = range(10)

SHAPE_IDS = {
    "Sphere": SPHERE, "Ellipsoid": ELLIPSOID, "Circle": CIRCLE, "Ellipse": ELLIPSE, "Box": BOX,
    "Rectangle": RECTANGLE, "Plane": PLANE, "Circle_Tail": CIRCLE_TAIL,
    "Circle_TwoTail": CIRCLE_TWOTAIL,
}
SHAPE_NAMES = {v: k for k, v in SHAPE_IDS.items()}
TWO_D_SHAPES = {CIRCLE, ELLIPSE, RECTANGLE, CIRCLE_TAIL, CIRCLE_TWOTAIL}

# Per-shape parameter slots: (solidDict key, slot) pairs; "com" fills 3 slots.
_PARAM_LAYOUT = {
    SPHERE: (("radius", 0), ("com", 1)),
    CIRCLE: (("radius", 0), ("com", 1)),
    ELLIPSOID: (("radiusa", 0), ("radiusb", 1), ("radiusc", 2)),
    ELLIPSE: (("radiusa", 0), ("radiusb", 1), ("com", 3)),
    RECTANGLE: (("radiusa", 0), ("radiusb", 1), ("com", 3)),
    BOX: (("radiusa", 0), ("radiusb", 1), ("radiusc", 2), ("com", 3)),
    PLANE: (),
    CIRCLE_TAIL: (("radius", 0), ("ratio", 1), ("thickness", 2), ("com", 3)),
    CIRCLE_TWOTAIL: (("radius", 0), ("ratio", 1), ("thickness", 2), ("com", 3)),
}

# CSG shapes (sdfibm_csg_shape in include/sdfibm_hip.h): primitives in the solid's local frame
# combined left to right with the reference's operators (src/libshape/sdf/sdf.h:133-142).
CSG_BOX, CSG_ELLIPSOID, CSG_CYLINDER, CSG_SPHERE = range(4)
CSG_UNION, CSG_DIFFERENCE, CSG_INTERSECTION = range(3)
CSG_MAX_PRIMS = 8
CSG_PRIM_DTYPE = np.dtype([("kind", "<i4"), ("op", "<i4"), ("axis", "<i4"), ("reserved", "<i4"),
                           ("center", "<f8", (3,)), ("size", "<f8", (3,))], align=True)
CSG_DTYPE = np.dtype([("n_prims", "<i4"), ("reserved", "<i4"), ("volume", "<f8"), ("moi", "<f8", (3,)),
                      ("radius_b", "<f8"), ("prim", CSG_PRIM_DTYPE, (CSG_MAX_PRIMS,))], align=True)
assert CSG_PRIM_DTYPE.itemsize == 64 and CSG_DTYPE.itemsize == 560

SOLID_DTYPE = np.dtype([
    ("shape", "<i4"), ("flags", "<i4"), ("prm", "<f8", (6,)), ("center", "<f8", (3,)),
    ("q", "<f8", (4,)), ("vel", "<f8", (3,)), ("omega", "<f8", (3,)),
], align=True)
assert SOLID_DTYPE.itemsize == 160

__all__ = ["SOLID_DTYPE", "SHAPE_IDS", "SHAPE_NAMES", "make_shape", "make_solids",
           "quaternion_xyz", "random_unit_quaternions", "shape_volume", "CSG_DTYPE", "csg_shape",
           "csg_inside_local"]


class UnknownShapeError(RuntimeError):
    """Raised for a shape type name the registry does not know (the reference's factory
    throws std::runtime_error for the same case, genericfactory.h:29-40)."""


def make_shape(type_name: str, **params) -> tuple:
    """Validate a solidDict ``shapes`` entry → (shape id, prm[6])."""
    if type_name not in SHAPE_IDS:
        raise UnknownShapeError(f"unknown shape type '{type_name}'")
    sid = SHAPE_IDS[type_name]
    prm = np.zeros(6)
    for key, slot in _PARAM_LAYOUT[sid]:
        if key == "com":
            com = params.get("com", (0.0, 0.0, 0.0))
            prm[slot:slot + 3] = np.asarray(com, dtype=np.float64)
        else:
            if key not in params:
                raise KeyError(f"shape {type_name} needs '{key}'")
            v = float(params[key])
            if key.startswith("radius") and not v > 0:
                raise ValueError(f"{type_name}.{key} must be > 0")
            prm[slot] = v
    return sid, prm


def quaternion_xyz(euler_rad) -> np.ndarray:
    """Orientation from Euler angles in the XYZ rotation sequence (what ``Solid::setOrientation``
    builds, src/solid.h:76-80): q = qx(a) * qy(b) * qz(c), each q_axis(t) = (cos t/2, sin t/2 · e)."""
    ax, ay, az = (float(a) for a in euler_rad)

    def axis_q(t, e):
        return np.array([math.cos(0.5 * t), *(math.sin(0.5 * t) * np.asarray(e, dtype=np.float64))])

    def mul(a, b):
        aw, av = a[0], a[1:]
        bw, bv = b[0], b[1:]
        w = aw * bw - (av[0] * bv[0] + av[1] * bv[1] + av[2] * bv[2])
        v = (aw * bv + bw * av) + np.array([av[1] * bv[2] - av[2] * bv[1],
                                            av[2] * bv[0] - av[0] * bv[2],
                                            av[0] * bv[1] - av[1] * bv[0]])
        return np.array([w, *v])

    q = axis_q(ax, (1, 0, 0))
    q = mul(q, axis_q(ay, (0, 1, 0)))
    q = mul(q, axis_q(az, (0, 0, 1)))
    return q


def random_unit_quaternions(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random rotations (Shoemake's subgroup algorithm); rows are (w, x, y, z)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    t2, t3 = 2 * np.pi * u2, 2 * np.pi * u3
    return np.stack([b * np.cos(t3), a * np.sin(t2), a * np.cos(t2), b * np.sin(t3)], axis=1)


def make_solids(n: int) -> np.ndarray:
    s = np.zeros(n, dtype=SOLID_DTYPE)
    s["q"][:, 0] = 1.0
    return s


def shape_volume(shape: int, prm) -> float:
    """Shape volume as the reference's shape constructors define it (m_volume); for 2-D
    shapes this is the area, and circle tails count only the circle (circle_tail.h:29)."""
    if shape == SPHERE:
        return 4.0 / 3.0 * math.pi * prm[0] ** 3
    if shape == ELLIPSOID:
        return 4.0 / 3.0 * math.pi * prm[0] * prm[1] * prm[2]
    if shape in (CIRCLE, CIRCLE_TAIL, CIRCLE_TWOTAIL):
        return math.pi * prm[0] ** 2
    if shape == ELLIPSE:
        return math.pi * prm[0] * prm[1]
    if shape == BOX:
        return 8.0 * prm[0] * prm[1] * prm[2]
    if shape == RECTANGLE:
        return 4.0 * prm[0] * prm[1]
    return 0.0


def csg_inside_local(shape: np.ndarray, p: np.ndarray) -> np.ndarray:
    """Strict inside test of a CSG shape at local points p [n, 3] (the library's predicate,
    vectorised): primitives combined left to right, union = or, difference = and not,
    intersection = and."""
    p = np.asarray(p, dtype=np.float64)
    res = None
    for t in range(int(shape["n_prims"])):
        g = shape["prim"][t]
        x, y, z = (p[:, d] - g["center"][d] for d in range(3))
        k = int(g["kind"])
        a = g["size"]
        if k == CSG_BOX:
            b = (np.abs(x) < a[0]) & (np.abs(y) < a[1]) & (np.abs(z) < a[2])
        elif k == CSG_ELLIPSOID:
            b = x * x * (1.0 / (a[0] * a[0])) + y * y * (1.0 / (a[1] * a[1])) + z * z * (1.0 / (a[2] * a[2])) < 1.0
        elif k == CSG_SPHERE:
            b = x * x + y * y + z * z < a[0] * a[0]
        elif k == CSG_CYLINDER:
            ax = int(g["axis"])
            u, v = [(y, z), (x, z), (x, y)][ax]
            w = (x, y, z)[ax]
            b = (u * u + v * v < a[0] * a[0]) & (np.abs(w) < a[1])
        else:
            raise ValueError(f"unknown CSG primitive kind {k}")
        if res is None:
            res = b
        elif int(g["op"]) == CSG_UNION:
            res = res | b
        elif int(g["op"]) == CSG_DIFFERENCE:
            res = res & ~b
        else:
            res = res & b
    return res


def _prim_box(g) -> tuple:
    a = np.asarray(g["size"], dtype=np.float64)
    k = int(g["kind"])
    if k == CSG_SPHERE:
        h = np.array([a[0]] * 3)
    elif k == CSG_CYLINDER:
        h = np.array([a[0]] * 3)
        h[int(g["axis"])] = a[1]
    else:
        h = a.copy()
    return np.asarray(g["center"]) - h, np.asarray(g["center"]) + h


def csg_shape(prims, *, volume=None, moi=None, radius_b=None, samples: int = 96) -> np.ndarray:
    """A CSG shape record from primitives given as dicts: kind ("box", "ellipsoid", "cylinder",
    "sphere"), op ("union", "difference", "intersection"; ignored for the first), center (local
    frame), size (box half sides; ellipsoid semi-axes; cylinder (radius, half length); sphere
    (radius,)), axis (cylinder: 0, 1 or 2).

    The mass properties the rigid-body update needs (IShape::m_volume, the diagonal of m_moi
    about the local origin per unit density, and m_radiusB for contacts) are taken as given, or
    integrated by the midpoint rule on a samples^3 grid over the shape's bounding box."""
    kinds = {"box": CSG_BOX, "ellipsoid": CSG_ELLIPSOID, "cylinder": CSG_CYLINDER, "sphere": CSG_SPHERE}
    ops = {"union": CSG_UNION, "difference": CSG_DIFFERENCE, "intersection": CSG_INTERSECTION}
    if not 1 <= len(prims) <= CSG_MAX_PRIMS:
        raise ValueError(f"a CSG shape has 1..{CSG_MAX_PRIMS} primitives")
    rec = np.zeros((), dtype=CSG_DTYPE)
    rec["n_prims"] = len(prims)
    for t, pr in enumerate(prims):
        g = rec["prim"][t]
        g["kind"] = kinds[pr["kind"]]
        g["op"] = ops[pr.get("op", "union")]
        g["axis"] = int(pr.get("axis", 2))
        g["center"] = pr.get("center", (0.0, 0.0, 0.0))
        size = list(pr["size"]) + [0.0] * (3 - len(pr["size"]))
        g["size"] = size
        if not all(v > 0 for v in size[: {CSG_BOX: 3, CSG_ELLIPSOID: 3, CSG_CYLINDER: 2, CSG_SPHERE: 1}[int(g["kind"])]]):
            raise ValueError("CSG primitive sizes must be > 0")
        if int(g["kind"]) == CSG_CYLINDER and g["axis"] not in (0, 1, 2):
            raise ValueError("cylinder axis must be 0, 1 or 2")
    lo, hi = _prim_box(rec["prim"][0])
    for t in range(1, len(prims)):
        a, b = _prim_box(rec["prim"][t])
        if int(rec["prim"][t]["op"]) == CSG_UNION:
            lo, hi = np.minimum(lo, a), np.maximum(hi, b)
        elif int(rec["prim"][t]["op"]) == CSG_INTERSECTION:
            lo, hi = np.maximum(lo, a), np.minimum(hi, b)
    if volume is None or moi is None or radius_b is None:
        n = int(samples)
        axes = [lo[d] + (np.arange(n) + 0.5) * (hi[d] - lo[d]) / n for d in range(3)]
        X, Y, Z = np.meshgrid(*axes, indexing="ij")
        P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
        ins = csg_inside_local(rec, P)
        dv = float(np.prod((hi - lo) / n))
        Q = P[ins]
        vol_n = ins.sum() * dv
        moi_n = np.array([((Q[:, 1] ** 2 + Q[:, 2] ** 2).sum()), ((Q[:, 0] ** 2 + Q[:, 2] ** 2).sum()),
                          ((Q[:, 0] ** 2 + Q[:, 1] ** 2).sum())]) * dv
        corners = np.array([[c0, c1, c2] for c0 in (lo[0], hi[0]) for c1 in (lo[1], hi[1]) for c2 in (lo[2], hi[2])])
        rb_n = float(np.sqrt((corners ** 2).sum(axis=1)).max())
    rec["volume"] = vol_n if volume is None else volume
    rec["moi"] = moi_n if moi is None else moi
    rec["radius_b"] = rb_n if radius_b is None else radius_b
    return rec
