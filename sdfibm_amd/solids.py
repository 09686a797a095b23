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
SPHERE, ELLIPSOID, CIRCLE, ELLIPSE, BOX, RECTANGLE, PLANE, CIRCLE_TAIL, CIRCLE_TWOTAIL = range(9)

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

SOLID_DTYPE = np.dtype([
    ("shape", "<i4"), ("flags", "<i4"), ("prm", "<f8", (6,)), ("center", "<f8", (3,)),
    ("q", "<f8", (4,)), ("vel", "<f8", (3,)), ("omega", "<f8", (3,)),
], align=True)
assert SOLID_DTYPE.itemsize == 160

__all__ = ["SOLID_DTYPE", "SHAPE_IDS", "SHAPE_NAMES", "make_shape", "make_solids",
           "quaternion_xyz", "random_unit_quaternions", "shape_volume"]


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
