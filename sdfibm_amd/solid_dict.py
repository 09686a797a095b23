"""solidDict -> solid records, densities and motion masks (the host side of
``SolidCloud::initFromDictionary``, src/solidcloud.cpp:43-205, and of the VOF tool's reader,
tool_vof/solidcloud.cpp).

Sections read: ``meta`` (on_fluid, on_twod, gravity), ``shapes`` (name, type, parameters),
``materials`` (type General, rho), ``motions`` (type Motion01Mask, mask), ``solids`` (shp_name,
mat_name, mot_name, pos, vel, euler in degrees, omega) and the VOF tool's ``planes`` (shp_name,
pos, euler), which follow the solids.  Errors follow the reference: an unknown shape type raises
(GenericFactory), an unknown material or motion type raises, a 2-D case needs z = 0 positions.
Forcers (for_name) are not part of this build and are ignored.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import solids as S
from .cloud import motion_mask
from .foam_dict import parse_foam_dict

__all__ = ["SolidSetup", "load_solid_dict"]


@dataclass
class SolidSetup:
    solids: np.ndarray                     # SOLID_DTYPE records: solids, then planes
    rho: np.ndarray                        # density per record (planes: 1)
    masks: np.ndarray                      # Motion01Mask bits per record (-1: free)
    two_d: bool = False
    on_fluid: bool = True
    gravity: np.ndarray = field(default_factory=lambda: np.zeros(3))
    names: list = field(default_factory=list)


def _vec(v, default=(0.0, 0.0, 0.0)):
    if v is None:
        return np.array(default, dtype=np.float64)
    a = np.asarray(v, dtype=np.float64).reshape(-1)
    if a.size != 3:
        raise ValueError(f"expected a 3-vector, got {v!r}")
    return a


def load_solid_dict(source: str, *, is_text: bool = False) -> SolidSetup:
    text = source if is_text else open(source).read()
    root = parse_foam_dict(text)
    meta = root.get("meta", {})
    two_d = bool(int(meta.get("on_twod", 0)))
    on_fluid = bool(int(meta.get("on_fluid", 1)))
    gravity = _vec(meta.get("gravity"))

    shapes = {}
    for key, d in root.get("shapes", {}).items():
        params = {k: v for k, v in d.items() if k not in ("name", "type")}
        shapes[str(d.get("name", key))] = S.make_shape(str(d["type"]), **params)

    mats = {}
    for key, d in root.get("materials", {}).items():
        if str(d.get("type")) != "General":
            raise ValueError(f"unrecognised material type {d.get('type')!r}")
        mats[str(d["name"])] = float(d["rho"])

    motions = {}
    for key, d in root.get("motions", {}).items():
        if str(d.get("type")) != "Motion01Mask":
            raise ValueError(f"motion type {d.get('type')!r} is not supported (Motion01Mask only)")
        motions[str(d["name"])] = motion_mask(str(d["mask"]))

    recs, rho, masks, names = [], [], [], []
    for section in ("solids", "planes"):
        for key, d in root.get(section, {}).items():
            shp = str(d["shp_name"])
            if shp not in shapes:
                raise KeyError(f"{section}.{key}: unknown shape name {shp!r}")
            sid, prm = shapes[shp]
            pos = _vec(d.get("pos"))
            if two_d and pos[2] != 0:
                raise ValueError(f"{section}.{key}: a solid must have z = 0 in a 2-D case")
            r = S.make_solids(1)[0]
            r["shape"] = sid
            r["prm"] = prm
            r["center"] = pos
            r["vel"] = _vec(d.get("vel"))
            r["omega"] = _vec(d.get("omega"))
            r["q"] = S.quaternion_xyz(_vec(d.get("euler")) * math.pi / 180.0)
            recs.append(r)
            mat = d.get("mat_name")
            rho.append(mats[str(mat)] if mat is not None else 1.0)
            mot = str(d.get("mot_name", "free"))
            masks.append(-1 if mot == "free" else motions[mot])
            names.append(f"{section}.{key}")
    sol = np.array(recs, dtype=S.SOLID_DTYPE) if recs else S.make_solids(0)
    return SolidSetup(sol, np.array(rho, dtype=np.float64), np.array(masks, dtype=np.int32), two_d, on_fluid,
                      gravity, names)
