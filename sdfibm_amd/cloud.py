"""Host-side mirror of the reference's ``SolidCloud`` coupling interface.

``SolidCloud.interact(time, dt)`` is ``SolidCloud::interact`` (src/solidcloud.cpp:435-464):
it zero-fills As/Fs/Ts/Ct, runs every solid's enumeration + fraction + forcing +
force/torque reduction, clips As, and stores each solid's fluid force and torque
(``Solid::setFluidForceAndTorque``, src/solid.h:136-140).  Here it calls the HIP
kernels through the C ABI; the fields stay on the device and are copied back only
on request.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .foam_io import PolyMesh
from .mesh import StructuredHexMesh
from .solids import CSG_DTYPE, SOLID_DTYPE

__all__ = ["SolidCloud", "DYNAMICS_DTYPE", "motion_mask"]

# sdfibm_dynamics (include/sdfibm_hip.h)
DYNAMICS_DTYPE = np.dtype([("rho", "f8"), ("motion_mask", "i4"), ("reserved", "i4")])


def motion_mask(mask: str) -> int:
    """Motion01Mask string "b<vx><vy><vz><wx><wy><wz>" -> bits (libmotion/motion01mask.h:27-52:
    must start with 'b', have 7 characters, each of the six 0 or 1)."""
    if len(mask) != 7 or mask[0] != "b" or any(ch not in "01" for ch in mask[1:]):
        raise ValueError(f"bad motion mask {mask!r}: expected b followed by six 0/1 characters")
    return sum(1 << k for k in range(6) if mask[k + 1] == "1")


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class SolidCloud:
    """Coupling-step context on the current GPU.

    mesh       StructuredHexMesh (rectilinear → closed-form geometry on the device; otherwise
               the host geometry arrays are uploaded once), or a PolyMesh (foam_io: a general
               polyMesh, e.g. several blockMesh blocks; the library derives its addressing and
               geometry)
    two_d      meta.on_twod
    rho_f      fluid density (0 reproduces on_fluid = 0)
    precision  "fp64" (parity mode) or "fp32" (resident fields and fractions in FP32; the
               inside tests and the force/torque reduction stay FP64)
    """

    def __init__(self, mesh, *, two_d: bool = False, rho_f: float = 1.0,
                 precision: str = "fp64", force_structured: bool = False):
        L = _lib.load()
        self.mesh = mesh
        self.n_cells = mesh.n_cells
        self._keep = []
        cfg = _lib.Config()
        cfg.two_d = 1 if two_d else 0
        cfg.precision = _lib.FP64 if precision == "fp64" else _lib.FP32
        cfg.rho_f = float(rho_f)
        self.precision = precision
        self.n_solids = 0
        if isinstance(mesh, PolyMesh):
            self._create_poly(L, mesh, cfg)
            return
        d = _lib.MeshDesc()
        d.nx, d.ny, d.nz = mesh.nx, mesh.ny, mesh.nz
        if mesh.rectilinear and not force_structured:
            d.kind = _lib.MESH_RECTILINEAR
            xs, ys, zs = (np.ascontiguousarray(a, dtype=np.float64) for a in (mesh.xs, mesh.ys, mesh.zs))
            self._keep += [xs, ys, zs]
            d.xs, d.ys, d.zs = xs.ctypes.data, ys.ctypes.data, zs.ctypes.data
        else:
            d.kind = _lib.MESH_STRUCTURED
            g = mesh.geometry()
            arrs = dict(points=np.ascontiguousarray(mesh.points, dtype=np.float64),
                        cc=np.ascontiguousarray(g["cc"]), V=np.ascontiguousarray(g["V"]),
                        fxc=np.ascontiguousarray(g["fx_c"]), fyc=np.ascontiguousarray(g["fy_c"]),
                        fzc=np.ascontiguousarray(g["fz_c"]), fxs=np.ascontiguousarray(g["fx_s"]),
                        fys=np.ascontiguousarray(g["fy_s"]), fzs=np.ascontiguousarray(g["fz_s"]))
            self._keep += list(arrs.values())
            d.points = arrs["points"].ctypes.data
            d.cell_centres = arrs["cc"].ctypes.data
            d.cell_volumes = arrs["V"].ctypes.data
            for f, key in enumerate("xyz"):
                d.face_centres[f] = arrs[f"f{key}c"].ctypes.data
                d.face_areas[f] = arrs[f"f{key}s"].ctypes.data
        h = C.c_void_p()
        st = L.sdfibm_create(C.byref(d), C.byref(cfg), C.byref(h))
        _lib.check(st, None)
        self._h = h
        self._keep = []  # the library copied the mesh

    def _create_poly(self, L, pm: PolyMesh, cfg):
        pts = np.ascontiguousarray(pm.points, dtype=np.float64)
        fo = np.ascontiguousarray(pm.face_offsets, dtype=np.int64)
        fp = np.ascontiguousarray(pm.face_points, dtype=np.int32)
        own = np.ascontiguousarray(pm.owner, dtype=np.int32)
        nei = np.ascontiguousarray(pm.neighbour, dtype=np.int32)
        d = _lib.PolyMeshDesc()
        d.n_points, d.n_faces, d.n_internal_faces, d.n_cells = pts.shape[0], own.size, nei.size, pm.n_cells
        d.points, d.face_offsets, d.face_points = pts.ctypes.data, fo.ctypes.data, fp.ctypes.data
        d.owner, d.neighbour = own.ctypes.data, nei.ctypes.data if nei.size else None
        h = C.c_void_p()
        st = L.sdfibm_create_polymesh(C.byref(d), C.byref(cfg), C.byref(h))
        _lib.check(st, None)
        self._h = h

    # -- lifecycle ---------------------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            _lib.load().sdfibm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state -------------------------------------------------------------------------------
    def set_solids(self, solids: np.ndarray):
        solids = np.ascontiguousarray(solids, dtype=SOLID_DTYPE)
        _lib.check(_lib.load().sdfibm_set_solids(self._h, _p(solids), solids.size), self._h)
        self.n_solids = solids.size

    def set_csg_shapes(self, shapes: np.ndarray):
        """The CSG shape table (records of solids.CSG_DTYPE, e.g. from solids.csg_shape); a CSG
        solid's prm[0] is its index here.  Call set_solids again afterwards."""
        shapes = np.ascontiguousarray(np.atleast_1d(shapes), dtype=CSG_DTYPE)
        _lib.check(_lib.load().sdfibm_set_csg_shapes(self._h, _p(shapes), shapes.size), self._h)

    def set_velocity(self, U: np.ndarray):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.n_cells, 3)
        _lib.check(_lib.load().sdfibm_set_velocity(self._h, _p(U)), self._h)

    def set_velocity_cells_noalloc(self, first: int, U_part: np.ndarray):
        """U of cells [first, first + len(U_part)) without any device allocation (at most
        _lib.NOALLOC_MAX_CELLS cells per call)."""
        U_part = np.ascontiguousarray(U_part, dtype=np.float64).reshape(-1, 3)
        _lib.check(_lib.load().sdfibm_set_velocity_cells_noalloc(self._h, int(first), U_part.shape[0], _p(U_part)),
                   self._h)

    # -- the coupling step -------------------------------------------------------------------
    def interact(self, time: float, dt: float) -> np.ndarray:
        ft = np.zeros((self.n_solids, 6))
        _lib.check(_lib.load().sdfibm_interact(self._h, float(time), float(dt), _p(ft)), self._h)
        return ft

    def interact_device(self, time: float, dt: float) -> None:
        """interact without copying force/torque back (it still waits for the step)."""
        _lib.check(_lib.load().sdfibm_interact(self._h, float(time), float(dt), None), self._h)

    def interact_steps(self, time: float, dt: float, n_steps: int) -> np.ndarray:
        """n_steps interacts of the current state as replays of one captured HIP graph; returns
        the last one's force/torque (timing()[3] is the replays' device time)."""
        ft = np.zeros((self.n_solids, 6))
        _lib.check(_lib.load().sdfibm_interact_steps(self._h, float(time), float(dt), int(n_steps), _p(ft)),
                   self._h)
        return ft

    def update_velocity(self, dt: float) -> None:
        """U -= Fs*dt on the device (the solver step after interact, src/main.cpp:70)."""
        _lib.check(_lib.load().sdfibm_update_velocity(self._h, float(dt)), self._h)

    def fix_internal(self) -> None:
        """SolidCloud::fixInternal: U = rigid velocity in cells wholly inside a solid."""
        _lib.check(_lib.load().sdfibm_fix_internal(self._h), self._h)

    # -- rigid-body update (SolidCloud::evolve) ---------------------------------------------
    def set_dynamics(self, rho, motion_mask=-1):
        """Solid densities and Motion01Mask bits (bit k = character k+1 of "b......"; -1 = free)."""
        d = np.zeros(self.n_solids, dtype=DYNAMICS_DTYPE)
        d["rho"] = rho
        d["motion_mask"] = motion_mask
        _lib.check(_lib.load().sdfibm_set_dynamics(self._h, _p(d), d.size), self._h)

    def set_contact(self, stiffness: float, damping: float = 0.0):
        """Contact spring-dashpot between solids during evolve (0, 0: no contacts, the
        reference's behaviour at its HEAD; 1e4, 0: the reference's linear law)."""
        _lib.check(_lib.load().sdfibm_set_contact(self._h, float(stiffness), float(damping)), self._h)

    def evolve(self, time: float, dt: float, gravity=(0.0, 0.0, 0.0)) -> None:
        g = np.ascontiguousarray(gravity, dtype=np.float64)
        _lib.check(_lib.load().sdfibm_evolve(self._h, float(time), float(dt), _p(g)), self._h)

    def run_steps(self, time: float, dt: float, n_steps: int, gravity=(0.0, 0.0, 0.0)) -> np.ndarray:
        """n_steps coupled steps (interact, U -= Fs dt, evolve, fixInternal) as replays of one
        captured HIP graph; returns the last step's force/torque [ns, 6]."""
        g = np.ascontiguousarray(gravity, dtype=np.float64)
        ft = np.zeros((self.n_solids, 6))
        _lib.check(_lib.load().sdfibm_run_steps(self._h, float(time), float(dt), int(n_steps), _p(g), _p(ft)),
                   self._h)
        return ft

    def solids(self) -> np.ndarray:
        out = np.zeros(self.n_solids, dtype=SOLID_DTYPE)
        _lib.check(_lib.load().sdfibm_get_solids(self._h, _p(out), out.size), self._h)
        return out

    # -- learned fraction surrogate ----------------------------------------------------------
    def set_surrogate(self, params, band=(0.02, 0.98), max_angle_deg=30.0):
        """Use the surrogate network (1473 FP32 parameters) for cut cells; params None: exact."""
        L = _lib.load()
        if params is None:
            _lib.check(L.sdfibm_set_surrogate(self._h, None, 0.0, 0.0, 0.0), self._h)
            return
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != 1473:
            raise ValueError("the surrogate has 1473 parameters")
        _lib.check(L.sdfibm_set_surrogate(self._h, _p(p), float(band[0]), float(band[1]), float(max_angle_deg)),
                   self._h)

    def surrogate_count(self) -> int:
        n = np.zeros(1, dtype=np.int64)
        _lib.check(_lib.load().sdfibm_surrogate_count(self._h, _p(n)), self._h)
        return int(n[0])

    def velocity(self) -> np.ndarray:
        U = np.zeros((self.n_cells, 3))
        _lib.check(_lib.load().sdfibm_get_velocity(self._h, _p(U)), self._h)
        return U

    def fields(self) -> dict:
        n = self.n_cells
        out = dict(As=np.zeros(n), Fs=np.zeros((n, 3)), Ts=np.zeros(n), Ct=np.zeros(n))
        _lib.check(_lib.load().sdfibm_get_fields(self._h, _p(out["As"]), _p(out["Fs"]),
                                                 _p(out["Ts"]), _p(out["Ct"])), self._h)
        return out

    def candidates(self, solid: int):
        """(counts[4], cells, alpha) of one solid in the reference's summation order."""
        L = _lib.load()
        counts = np.zeros(4, dtype=np.int64)
        _lib.check(L.sdfibm_candidates(self._h, int(solid), _p(counts), None, None, 0), self._h)
        n = int(counts[:3].sum())
        cells = np.zeros(n, dtype=np.int64)
        alpha = np.zeros(n)
        _lib.check(L.sdfibm_candidates(self._h, int(solid), _p(counts), _p(cells), _p(alpha), n),
                   self._h)
        return counts, cells, alpha

    def pair_counts(self) -> tuple:
        """(ALL_INSIDE pairs, cut pairs) of the last interact, summed over all solids."""
        out = np.zeros(2, dtype=np.int64)
        _lib.check(_lib.load().sdfibm_pair_counts(self._h, _p(out)), self._h)
        return int(out[0]), int(out[1])

    def set_tuning(self, cells_wave: int = -1, big_chunk: int = 1024, big_grid: int = 4):
        """Run-time variants of the step (see sdfibm_set_tuning)."""
        _lib.check(_lib.load().sdfibm_set_tuning(self._h, int(cells_wave), int(big_chunk), int(big_grid)), self._h)

    def autotune(self, dt: float, cache_path: str | None = None) -> dict:
        """Time the step's variants on the current solids and keep the fastest (see
        sdfibm_autotune); returns the choice."""
        import json

        buf = C.create_string_buffer(512)
        path = cache_path.encode() if cache_path else None
        _lib.check(_lib.load().sdfibm_autotune(self._h, float(dt), path, buf, 512), self._h)
        return json.loads(buf.value.decode())

    def buffer_capacity(self) -> tuple:
        """(cells the tile-code buffer holds, solids the per-solid buffers hold)."""
        out = np.zeros(2, dtype=np.int64)
        _lib.load().sdfibm_buffer_capacity(self._h, _p(out), 2)
        return int(out[0]), int(out[1])

    def set_phase_timing(self, on: bool = True) -> None:
        """HIP events between the phases of each interact (off by default: each record orders
        the kernels around it); timing() reads them."""
        _lib.check(_lib.load().sdfibm_set_phase_timing(self._h, 1 if on else 0), self._h)

    def timing(self) -> np.ndarray:
        """Phase times of the last interact in ms: setup+binning, tile pass, field pass, total,
        tile classification, cut cells + forcing (+ plane labelling), big-tile pass (second
        stream).  Structured meshes: zeros unless set_phase_timing(True); the replays'
        device time (index 3) after run_steps and interact_steps either way."""
        ms = np.zeros(7)
        _lib.load().sdfibm_last_timing(self._h, _p(ms), 7)
        return ms

    def device_pointers(self) -> dict:
        L = _lib.load()
        U3 = (C.c_void_p * 3)()
        F3 = (C.c_void_p * 3)()
        As, Ts = C.c_void_p(), C.c_void_p()
        Ct = C.c_void_p()
        _lib.check(L.sdfibm_field_pointers(self._h, U3, C.byref(As), F3, C.byref(Ts), C.byref(Ct)),
                   self._h)
        return dict(U=[U3[i] for i in range(3)], As=As.value, Fs=[F3[i] for i in range(3)],
                    Ts=Ts.value, Ct=Ct.value)
