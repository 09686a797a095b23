"""Seeded synthetic scenes for the coupling step: the benchmark configurations and small
parity cases.  Every generator is deterministic in its seed (numpy PCG64), so the HIP
path and the CPU oracle see bit-identical inputs.

sedimentation_2d   400x800x1 on [0,1]x[0,2]x[-0.5,0.5], 100 discs (d = 0.0625) on a jittered
                   10x10 lattice, four plane walls (normals into the fluid), rho_s/rho_f 1.5.
cloud_3d           uniform n^3 hex mesh on the unit cube, spheres and ellipsoids with lognormal
                   diameters (cells), random non-overlapping placement (largest first), random
                   unit quaternions, v ~ N(0, 0.1), omega ~ N(0, 1), and a divergence-free
                   Fourier velocity field (16 modes, rms 1).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import mesh as M
from . import solids as S

DEFAULT_SEED = 20261202


@dataclass
class Scene:
    name: str
    mesh: M.StructuredHexMesh
    solids: np.ndarray
    U: np.ndarray | None
    two_d: bool
    rho_f: float
    dt: float
    info: dict


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def fourier_velocity(points: np.ndarray, seed: int, n_modes: int = 16, kmax: int = 4) -> np.ndarray:
    """U(x) = sum_m (a_m x k_m) cos(2 pi n_m . x + phi_m), n_m in Z^3 \\ 0 with |n|_inf <= kmax.
    Each term is divergence-free ((a x k) . k = 0); the field is scaled to rms 1."""
    rng = _rng(seed)
    U = np.zeros_like(points)
    for _ in range(n_modes):
        while True:
            n = rng.integers(-kmax, kmax + 1, size=3)
            if np.any(n != 0):
                break
        a = rng.normal(size=3)
        phi = rng.uniform(0.0, 2 * np.pi)
        k = n / np.linalg.norm(n)
        amp = np.cross(a, k)
        ph = 2 * np.pi * (points @ n.astype(np.float64)) + phi
        U += np.cos(ph)[:, None] * amp[None, :]
    rms = math.sqrt(float(np.mean(np.sum(U * U, axis=1))))
    return U / rms


def cloud_3d(n: int = 256, n_solids: int = 10000, frac_sphere: float = 0.7, d_median: float = 8.0,
             d_sigma: float = 0.45, d_min: float = 4.0, d_max: float = 24.0, seed: int = DEFAULT_SEED,
             with_velocity: bool = True, gap_cells: float = 0.0, max_tries: int = 4000) -> Scene:
    """Spheres + ellipsoids in the unit cube (sizes in cells of h = 1/n)."""
    rng = _rng(seed)
    h = 1.0 / n
    mesh = M.uniform_box((n, n, n), (0, 0, 0), (1, 1, 1))
    d = np.exp(rng.normal(math.log(d_median), d_sigma, size=n_solids))
    while np.any((d < d_min) | (d > d_max)):  # truncate by resampling
        bad = (d < d_min) | (d > d_max)
        d[bad] = np.exp(rng.normal(math.log(d_median), d_sigma, size=int(bad.sum())))
    is_sphere = rng.random(n_solids) < frac_sphere
    ratios = rng.uniform(0.5, 1.0, size=(n_solids, 2))
    order = np.argsort(-d, kind="stable")
    radius_b = 0.5 * d * h  # bounding radius (max semi-axis)
    pos = np.full((n_solids, 3), np.nan)
    # uniform-grid hashing of placed centres
    cell = 2.0 * radius_b.max() + gap_cells * h
    ng = max(1, int(1.0 / cell))
    grid: dict = {}
    placed_r = np.zeros(n_solids)
    for idx in order:
        r = radius_b[idx]
        ok = False
        for _attempt in range(max_tries // 64):
            cand = rng.uniform(r, 1.0 - r, size=(64, 3))
            gi = np.minimum((cand * ng).astype(np.int64), ng - 1)
            good = np.ones(64, dtype=bool)
            for c in range(64):
                key0 = gi[c]
                for dz in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            lst = grid.get((key0[0] + dx, key0[1] + dy, key0[2] + dz))
                            if not lst:
                                continue
                            P = pos[lst]
                            rr = placed_r[lst] + r + gap_cells * h
                            if np.any(np.sum((P - cand[c]) ** 2, axis=1) < rr * rr):
                                good[c] = False
                                break
                        if not good[c]:
                            break
                    if not good[c]:
                        break
                if good[c]:
                    pos[idx] = cand[c]
                    placed_r[idx] = r
                    grid.setdefault(tuple(gi[c]), []).append(idx)
                    ok = True
                    break
            if ok:
                break
        if not ok:
            raise RuntimeError(f"could not place solid {idx} (d={d[idx]:.2f} cells) without overlap")
    sol = S.make_solids(n_solids)
    sol["center"] = pos
    sol["q"] = S.random_unit_quaternions(rng, n_solids)
    sol["vel"] = rng.normal(0.0, 0.1, size=(n_solids, 3))
    sol["omega"] = rng.normal(0.0, 1.0, size=(n_solids, 3))
    for s in range(n_solids):
        if is_sphere[s]:
            sol[s]["shape"] = S.SPHERE
            sol[s]["prm"][0] = 0.5 * d[s] * h
        else:
            a = 0.5 * d[s] * h
            sol[s]["shape"] = S.ELLIPSOID
            sol[s]["prm"][:3] = (a, a * ratios[s, 0], a * ratios[s, 1])
    U = None
    if with_velocity:
        U = fourier_velocity(mesh_cell_centres(mesh), seed + 1)
    vol = sum(S.shape_volume(int(x["shape"]), x["prm"]) for x in sol)
    info = dict(n=n, n_solids=n_solids, seed=seed, solid_volume_fraction=vol,
                n_spheres=int(is_sphere.sum()), d_cells=(float(d.min()), float(np.median(d)), float(d.max())))
    return Scene(f"cloud3d_{n}_{n_solids}", mesh, sol, U, False, 1.0, 1e-3, info)


def mesh_cell_centres(mesh: M.StructuredHexMesh) -> np.ndarray:
    """Cell centres; rectilinear meshes use midpoints (no full geometry pass)."""
    if mesh.rectilinear:
        cx = 0.5 * (mesh.xs[:-1] + mesh.xs[1:])
        cy = 0.5 * (mesh.ys[:-1] + mesh.ys[1:])
        cz = 0.5 * (mesh.zs[:-1] + mesh.zs[1:])
        Z, Y, X = np.meshgrid(cz, cy, cx, indexing="ij")
        return np.stack([X.reshape(-1), Y.reshape(-1), Z.reshape(-1)], axis=1)
    return mesh.cell_centres()


def walls_2d(lo=(0.0, 0.0), hi=(1.0, 2.0)) -> np.ndarray:
    """Four plane walls on the box faces, normals (local +y is outside) pointing into the fluid:
    bottom (euler 0), top (180), left (-90), right (+90) - the README convention."""
    w = S.make_solids(4)
    specs = [((0.5 * (lo[0] + hi[0]), lo[1], 0.0), 0.0),     # bottom: inside is y < lo
             ((0.5 * (lo[0] + hi[0]), hi[1], 0.0), 180.0),   # top: inside is y > hi
             ((lo[0], 0.5 * (lo[1] + hi[1]), 0.0), -90.0),   # left: inside is x < lo
             ((hi[0], 0.5 * (lo[1] + hi[1]), 0.0), 90.0)]    # right: inside is x > hi
    for t, (pos, ang) in enumerate(specs):
        w[t]["shape"] = S.PLANE
        w[t]["center"] = pos
        w[t]["q"] = S.quaternion_xyz((0.0, 0.0, math.radians(ang)))
    return w


def sedimentation_2d(nx: int = 400, ny: int = 800, n_discs: int = 100, diameter: float = 0.0625,
                     seed: int = DEFAULT_SEED, with_walls: bool = True) -> Scene:
    rng = _rng(seed)
    mesh = M.uniform_box((nx, ny, 1), (0.0, 0.0, -0.5), (1.0, 2.0, 0.5))
    h = 1.0 / nx
    r = 0.5 * diameter
    side = int(math.ceil(math.sqrt(n_discs)))
    # jittered lattice in the upper part of the box, non-overlap (gap >= 2h) by construction
    pitch_x = 1.0 / side
    pitch_y = 1.0 / side
    jit = max(0.0, 0.5 * (pitch_x - 2 * r - 2 * h))
    centres = []
    for t in range(n_discs):
        ix, iy = t % side, t // side
        cx = (ix + 0.5) * pitch_x + rng.uniform(-jit, jit)
        cy = 0.9 + (iy + 0.5) * pitch_y + rng.uniform(-jit, jit)
        centres.append((cx, cy, 0.0))
    disc = S.make_solids(n_discs)
    disc["shape"] = S.CIRCLE
    disc["prm"][:, 0] = r
    disc["center"] = np.array(centres)
    parts = [disc]
    if with_walls:
        parts.append(walls_2d((0.0, 0.0), (1.0, 2.0)))
    sol = np.concatenate(parts)
    U = np.zeros((mesh.n_cells, 3))
    info = dict(nx=nx, ny=ny, n_discs=n_discs, seed=seed, rho_ratio=1.5, gravity=(0.0, -10.0, 0.0))
    return Scene(f"sedimentation2d_{nx}x{ny}", mesh, sol, U, True, 1.0, 1e-4, info)


def small_mixed_3d(n: int = 24, seed: int = 7, n_each: int = 2, jitter: float = 0.0) -> Scene:
    """A small 3-D scene with every 3-D shape (sphere, ellipsoid, box, and z-extruded 2-D shapes)
    at random poses, overlapping allowed, for parity tests."""
    rng = _rng(seed)
    if jitter > 0:
        mesh = M.jittered_box((n, n, n), (0, 0, 0), (1, 1, 1), jitter, seed + 11)
    else:
        mesh = M.uniform_box((n, n, n), (0, 0, 0), (1, 1, 1))
    h = 1.0 / n
    kinds = []
    for _ in range(n_each):
        kinds += [("Sphere", dict(radius=rng.uniform(2, 5) * h, com=rng.uniform(-1, 1, 3) * h)),
                  ("Ellipsoid", dict(radiusa=rng.uniform(2, 6) * h, radiusb=rng.uniform(1.5, 4) * h,
                                     radiusc=rng.uniform(1.5, 4) * h)),
                  ("Box", dict(radiusa=rng.uniform(1.5, 4) * h, radiusb=rng.uniform(1.5, 4) * h,
                               radiusc=rng.uniform(1.5, 4) * h, com=rng.uniform(-1, 1, 3) * h)),
                  ("Circle", dict(radius=rng.uniform(1.5, 3.5) * h)),
                  ("Ellipse", dict(radiusa=rng.uniform(2, 4) * h, radiusb=rng.uniform(1.2, 2.5) * h)),
                  ("Rectangle", dict(radiusa=rng.uniform(2, 4) * h, radiusb=rng.uniform(1.2, 3) * h)),
                  ("Circle_Tail", dict(radius=rng.uniform(1.5, 3) * h, ratio=1.0, thickness=0.8 * h)),
                  ("Circle_TwoTail", dict(radius=rng.uniform(1.5, 3) * h, ratio=1.2, thickness=1.2 * h))]
    sol = S.make_solids(len(kinds))
    for t, (name, prm) in enumerate(kinds):
        sid, p = S.make_shape(name, **prm)
        sol[t]["shape"] = sid
        sol[t]["prm"] = p
        sol[t]["center"] = rng.uniform(0.25, 0.75, 3)
        if sid in S.TWO_D_SHAPES:
            sol[t]["q"] = S.quaternion_xyz((0.0, 0.0, rng.uniform(0, 2 * np.pi)))
        else:
            sol[t]["q"] = S.random_unit_quaternions(rng, 1)[0]
        sol[t]["vel"] = rng.normal(0, 0.1, 3)
        sol[t]["omega"] = rng.normal(0, 1.0, 3)
    U = rng.normal(0, 1.0, size=(mesh.n_cells, 3))
    return Scene(f"mixed3d_{n}", mesh, sol, U, False, 1.3, 1e-3, dict(seed=seed))
