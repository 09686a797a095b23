"""ctypes binding of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py's cpu_baseline leg.  The product package never imports
this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        i64 = C.c_int64
        dp = C.POINTER(C.c_double)
        L.orc_mesh_create.restype = P
        L.orc_mesh_create.argtypes = [i64, P, i64, P, P, P, i64, P]
        L.orc_mesh_create_structured.restype = P
        L.orc_mesh_create_structured.argtypes = [i64, i64, i64, P]
        L.orc_mesh_destroy.argtypes = [P]
        L.orc_mesh_ncells.restype = i64
        L.orc_mesh_ncells.argtypes = [P]
        L.orc_mesh_geometry.argtypes = [P, P, P, P, P]
        L.orc_mesh_cell_points.restype = i64
        L.orc_mesh_cell_points.argtypes = [P, P, P]
        L.orc_is_inside.restype = C.c_int
        L.orc_is_inside.argtypes = [P, P]
        L.orc_signed_distance.restype = C.c_double
        L.orc_signed_distance.argtypes = [P, P]
        L.orc_world_to_local.argtypes = [P, P, P]
        L.orc_cell_fraction.restype = C.c_double
        L.orc_cell_fraction.argtypes = [P, P, i64, C.c_int]
        L.orc_find_nearest_cell.restype = i64
        L.orc_find_nearest_cell.argtypes = [P, P]
        L.orc_interact.restype = C.c_int
        L.orc_interact.argtypes = [P, C.c_int, C.c_double, C.c_double, i64, P, P, P, P, P, P, P,
                                   C.c_int, P]
        L.orc_interact_range.restype = C.c_int
        L.orc_interact_range.argtypes = [P, C.c_int, C.c_double, C.c_double, i64, P, i64, i64, P,
                                         P, P, P, P, P, C.c_int]
        L.orc_update_velocity.argtypes = [i64, P, P, C.c_double]
        L.orc_fix_internal.argtypes = [P, P, P, i64, P]
        L.orc_mass_properties.argtypes = [P, P, P]
        L.orc_evolve.argtypes = [i64, P, P, P, C.c_double, C.c_double, P, C.c_double]
        L.orc_set_csg.argtypes = [P, i64]
        L.orc_bounding_radius.restype = C.c_double
        L.orc_bounding_radius.argtypes = [P]
        L.orc_contact_forces.argtypes = [i64, P, C.c_double, C.c_double, P]
        L.orc_evolve_contact.argtypes = [i64, P, P, P, C.c_double, C.c_double, P, C.c_double, C.c_double,
                                         C.c_double]
        del dp
        _lib = L
    return _lib


class _Capture(C.Structure):
    _fields_ = [("counts", C.c_void_p), ("cells", C.c_void_p), ("alpha", C.c_void_p),
                ("capacity", C.c_int64), ("used", C.c_int64)]


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class OracleMesh:
    def __init__(self, pm=None, structured=None):
        """pm: a PolyMesh; or structured=(nx, ny, nz, points) for a single hex block."""
        L = lib()
        if structured is not None:
            nx, ny, nz, pts = structured
            pts = np.ascontiguousarray(pts, dtype=np.float64)
            self.h = L.orc_mesh_create_structured(int(nx), int(ny), int(nz), _ptr(pts))
            if not self.h:
                raise ValueError("oracle rejected the mesh")
            self.n_cells = int(L.orc_mesh_ncells(self.h))
            self.n_points = pts.shape[0]
            nx, ny, nz = int(nx), int(ny), int(nz)
            self.n_faces = 3 * nx * ny * nz + nx * ny + ny * nz + nx * nz
            return
        self._keep = (np.ascontiguousarray(pm.points, dtype=np.float64),
                      np.ascontiguousarray(pm.face_offsets, dtype=np.int64),
                      np.ascontiguousarray(pm.face_points, dtype=np.int32),
                      np.ascontiguousarray(pm.owner, dtype=np.int32),
                      np.ascontiguousarray(pm.neighbour, dtype=np.int32))
        pts, off, fp, own, nei = self._keep
        self.h = L.orc_mesh_create(pts.shape[0], _ptr(pts), own.size, _ptr(off), _ptr(fp),
                                   _ptr(own), nei.size, _ptr(nei))
        if not self.h:
            raise ValueError("oracle rejected the mesh")
        self.n_cells = int(L.orc_mesh_ncells(self.h))
        self.n_faces = own.size
        self.n_points = pts.shape[0]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mesh_destroy(self.h)
            self.h = None

    def geometry(self):
        cc = np.zeros((self.n_cells, 3))
        V = np.zeros(self.n_cells)
        fc = np.zeros((self.n_faces, 3))
        sf = np.zeros((self.n_faces, 3))
        lib().orc_mesh_geometry(self.h, _ptr(cc), _ptr(V), _ptr(fc), _ptr(sf))
        return cc, V, fc, sf

    def cell_points(self):
        off = np.zeros(self.n_cells + 1, dtype=np.int64)
        n = lib().orc_mesh_cell_points(self.h, _ptr(off), None)
        lab = np.zeros(n, dtype=np.int32)
        lib().orc_mesh_cell_points(self.h, _ptr(off), _ptr(lab))
        return off, lab

    def nearest_cell(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        return int(lib().orc_find_nearest_cell(self.h, _ptr(p)))

    def cell_fraction(self, solid, cell, two_d=False):
        s = np.ascontiguousarray(solid)
        return float(lib().orc_cell_fraction(self.h, _ptr(s), int(cell), int(bool(two_d))))

    def interact(self, solids, U, *, two_d=False, rho_f=1.0, dt=1.0, mode=0, capture=False,
                 capacity=None):
        """One interact() call.  Returns dict(As, Fs, Ts, Ct, ft[, counts, cells, alpha])."""
        L = lib()
        solids = np.ascontiguousarray(solids)
        ns = solids.size
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.n_cells, 3)
        out = dict(As=np.zeros(self.n_cells), Fs=np.zeros((self.n_cells, 3)),
                   Ts=np.zeros(self.n_cells), Ct=np.zeros(self.n_cells), ft=np.zeros((ns, 6)))
        cap = None
        if capture:
            capacity = capacity or max(1024, 4 * self.n_cells)
            out["counts"] = np.zeros((ns, 4), dtype=np.int64)
            out["cells"] = np.zeros(capacity, dtype=np.int64)
            out["alpha"] = np.zeros(capacity)
            cap = _Capture(out["counts"].ctypes.data, out["cells"].ctypes.data,
                           out["alpha"].ctypes.data, capacity, 0)
        rc = L.orc_interact(self.h, int(bool(two_d)), float(rho_f), float(dt), ns, _ptr(solids),
                            _ptr(U), _ptr(out["As"]), _ptr(out["Fs"]), _ptr(out["Ts"]),
                            _ptr(out["Ct"]), _ptr(out["ft"]), int(mode),
                            C.byref(cap) if cap is not None else None)
        if rc != 0:
            raise RuntimeError(f"orc_interact failed: {rc}")
        if capture:
            out["cells"] = out["cells"][:cap.used]
            out["alpha"] = out["alpha"][:cap.used]
        return out

    def update_velocity(self, U, Fs, dt):
        U = np.ascontiguousarray(U, dtype=np.float64).copy()
        Fs = np.ascontiguousarray(Fs, dtype=np.float64)
        lib().orc_update_velocity(self.n_cells, _ptr(U), _ptr(Fs), float(dt))
        return U

    def fix_internal(self, U, Ct, solids):
        U = np.ascontiguousarray(U, dtype=np.float64).copy()
        Ct = np.ascontiguousarray(Ct, dtype=np.float64)
        solids = np.ascontiguousarray(solids)
        lib().orc_fix_internal(self.h, _ptr(U), _ptr(Ct), solids.size, _ptr(solids))
        return U

    def interact_range(self, solids, U, s0, s1, *, two_d=False, rho_f=1.0, dt=1.0, mode=0,
                       fields=None):
        L = lib()
        solids = np.ascontiguousarray(solids)
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.n_cells, 3)
        if fields is None:
            fields = dict(As=np.zeros(self.n_cells), Fs=np.zeros((self.n_cells, 3)),
                          Ts=np.zeros(self.n_cells), Ct=np.zeros(self.n_cells),
                          ft=np.zeros((solids.size, 6)))
        rc = L.orc_interact_range(self.h, int(bool(two_d)), float(rho_f), float(dt), solids.size,
                                  _ptr(solids), int(s0), int(s1), _ptr(U), _ptr(fields["As"]),
                                  _ptr(fields["Fs"]), _ptr(fields["Ts"]), _ptr(fields["Ct"]),
                                  _ptr(fields["ft"]), int(mode))
        if rc != 0:
            raise RuntimeError(f"orc_interact_range failed: {rc}")
        return fields


def set_csg(shapes):
    """The CSG shape table (records of the product's CSG_DTYPE layout) the oracle's CSG solids use."""
    global _csg_keep
    _csg_keep = np.ascontiguousarray(shapes)
    lib().orc_set_csg(_ptr(_csg_keep), int(_csg_keep.size))


_csg_keep = None


def is_inside(solid, p):
    s = np.ascontiguousarray(solid)
    p = np.ascontiguousarray(p, dtype=np.float64)
    return bool(lib().orc_is_inside(_ptr(s), _ptr(p)))


def signed_distance(solid, p):
    s = np.ascontiguousarray(solid)
    p = np.ascontiguousarray(p, dtype=np.float64)
    return float(lib().orc_signed_distance(_ptr(s), _ptr(p)))


def world_to_local(solid, p):
    s = np.ascontiguousarray(solid)
    p = np.ascontiguousarray(p, dtype=np.float64)
    out = np.zeros(3)
    lib().orc_world_to_local(_ptr(s), _ptr(p), _ptr(out))
    return out


# Rigid-body state per solid beyond the solid record (orc_dyn in sdfibm_oracle.h).
DYN_DTYPE = np.dtype([("rho", "f8"), ("mask", "i4"), ("first_fluid", "i4"), ("ff_old", "f8", 3),
                      ("ft_old", "f8", 3)])


def make_dyn(n, rho, mask=-1):
    d = np.zeros(n, dtype=DYN_DTYPE)
    d["rho"] = rho
    d["mask"] = mask
    d["first_fluid"] = 1
    return d


def mass_properties(solid):
    """(volume, principal moments per unit density) of one solid record."""
    s = np.ascontiguousarray(solid)
    v = np.zeros(1)
    moi = np.zeros(3)
    lib().orc_mass_properties(_ptr(s), _ptr(v), _ptr(moi))
    return float(v[0]), moi


def evolve(solids, dyn, fluid_ft, time, dt, gravity, rho_f):
    """SolidCloud::evolve on copies: returns (solids, dyn) after one step."""
    s = np.ascontiguousarray(solids).copy()
    d = np.ascontiguousarray(dyn, dtype=DYN_DTYPE).copy()
    ft = np.ascontiguousarray(fluid_ft, dtype=np.float64).reshape(-1, 6)
    g = np.ascontiguousarray(gravity, dtype=np.float64)
    lib().orc_evolve(s.size, _ptr(s), _ptr(d), _ptr(ft), float(time), float(dt), _ptr(g), float(rho_f))
    return s, d


def bounding_radius(solid) -> float:
    """IShape::m_radiusB of one solid record (plane: 0)."""
    a = np.ascontiguousarray(np.atleast_1d(solid)[:1]).copy()
    return float(lib().orc_bounding_radius(_ptr(a)))


def contact_forces(solids, k, damping=0.0):
    """Contact force [ns, 3] on every solid by brute force over all pairs (orc_contact_forces)."""
    s = np.ascontiguousarray(solids).copy()
    out = np.zeros((s.size, 3))
    lib().orc_contact_forces(s.size, _ptr(s), float(k), float(damping), _ptr(out))
    return out


def evolve_contact(solids, dyn, fluid_ft, time, dt, gravity, rho_f, k, damping=0.0):
    """SolidCloud::evolve with contacts on copies: returns (solids, dyn) after one step."""
    s = np.ascontiguousarray(solids).copy()
    d = np.ascontiguousarray(dyn, dtype=DYN_DTYPE).copy()
    ft = np.ascontiguousarray(fluid_ft, dtype=np.float64).reshape(-1, 6)
    g = np.ascontiguousarray(gravity, dtype=np.float64)
    lib().orc_evolve_contact(s.size, _ptr(s), _ptr(d), _ptr(ft), float(time), float(dt), _ptr(g), float(rho_f),
                             float(k), float(damping))
    return s, d
