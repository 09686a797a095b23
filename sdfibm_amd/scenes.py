"""Seeded synthetic scenes for the coupling step: the benchmark configurations and small
parity cases.  Every generator is deterministic in its seed (numpy PCG64), so the HIP
path and the CPU oracle see bit-identical inputs.

sedimentation_2d   400x800x1 on [0,1]x[0,2]x[-0.5,0.5], 100 discs (d = 0.0625) on a jittered
                   10x10 lattice, four plane walls (normals into the fluid), rho_s/rho_f 1.5.
cloud_3d           uniform n^3 hex mesh on the unit cube, spheres and ellipsoids with lognormal
                   diameters (cells), random non-overlapping placement (largest first), random
                   unit quaternions, v ~ N(0, 0.1), omega ~ N(0, 1), and a divergence-free
                   Fourier velocity field (16 modes, rms 1).
ellipsoids_jittered n^3 hex mesh of the unit cube with interior vertices jittered by U(+-0.2 h)
                   per axis, non-overlapping ellipsoids with semi-axes U[3, 10] cells,
                   omega ~ N(0, 5), six fixed plane walls (floor and ceiling 1.5 h inside the
                   faces, the side walls on them), Fourier velocity field (rms 0.05).
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
    csg: np.ndarray | None = None  # CSG shape table of the scene's CSG solids


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


def place_nonoverlapping(rng, radius_b, lo, hi, gap: float = 0.0, max_tries: int = 4000,
                         avoid=None) -> np.ndarray:
    """Random sequential placement of bounding spheres inside the box lo..hi, largest first,
    rejecting overlaps (plus `gap`) through a uniform-grid hash of the placed centres, and
    overlaps with the spheres `avoid` = (centres [m, 3], radii [m]) when given."""
    n_solids = radius_b.size
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    order = np.argsort(-radius_b, kind="stable")
    pos = np.full((n_solids, 3), np.nan)
    cell = 2.0 * radius_b.max() + gap
    ng = max(1, int(float((hi - lo).min()) / cell))
    grid: dict = {}
    placed_r = np.zeros(n_solids)
    for idx in order:
        r = radius_b[idx]
        ok = False
        for _attempt in range(max_tries // 64):
            cand = rng.uniform(lo + r, hi - r, size=(64, 3))
            gi = np.minimum(((cand - lo) / (hi - lo) * ng).astype(np.int64), ng - 1)
            good = np.ones(64, dtype=bool)
            if avoid is not None:
                ac, ar = avoid
                d2 = np.sum((cand[:, None, :] - ac[None, :, :]) ** 2, axis=2)
                good &= np.all(d2 >= (ar[None, :] + r + gap) ** 2, axis=1)
            for c in range(64):
                if not good[c]:
                    continue
                key0 = gi[c]
                for dz in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            lst = grid.get((key0[0] + dx, key0[1] + dy, key0[2] + dz))
                            if not lst:
                                continue
                            P = pos[lst]
                            rr = placed_r[lst] + r + gap
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
            raise RuntimeError(f"could not place solid {idx} (radius {r:.4g}) without overlap")
    return pos


def cloud_3d(n: int = 256, n_solids: int = 10000, frac_sphere: float = 0.7, d_median: float = 8.0,
             d_sigma: float = 0.45, d_min: float = 4.0, d_max: float = 24.0, seed: int = DEFAULT_SEED,
             with_velocity: bool = True, gap_cells: float = 0.0, max_tries: int = 4000,
             d_uniform: bool = False) -> Scene:
    """Spheres + ellipsoids in the unit cube (sizes in cells of h = 1/n).  Diameters lognormal
    (median, sigma) truncated to [d_min, d_max], or uniform on [d_min, d_max] (d_uniform)."""
    rng = _rng(seed)
    h = 1.0 / n
    mesh = M.uniform_box((n, n, n), (0, 0, 0), (1, 1, 1))
    if d_uniform:
        d = rng.uniform(d_min, d_max, size=n_solids)
    else:
        d = np.exp(rng.normal(math.log(d_median), d_sigma, size=n_solids))
    while np.any((d < d_min) | (d > d_max)):  # truncate by resampling
        bad = (d < d_min) | (d > d_max)
        d[bad] = np.exp(rng.normal(math.log(d_median), d_sigma, size=int(bad.sum())))
    is_sphere = rng.random(n_solids) < frac_sphere
    ratios = rng.uniform(0.5, 1.0, size=(n_solids, 2))
    radius_b = 0.5 * d * h  # bounding radius (max semi-axis)
    pos = place_nonoverlapping(rng, radius_b, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), gap_cells * h, max_tries)
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


def csg_family(rng, kind: int, D: float) -> np.ndarray:
    """One CSG shape of overall size D (world units): unions and differences of boxes,
    cylinders, ellipsoids and spheres.  kind 0: box with a cylindrical through-hole; 1: ellipsoid
    pierced by a rod (union); 2: cross of three bars (union); 3: cylinder with a slot cut out;
    4: ellipsoid with an inner spherical cavity (difference); 5: cube rounded by a sphere
    (intersection)."""
    u = rng.uniform(0.8, 1.0, 3)
    if kind == 0:
        prims = [dict(kind="box", size=(0.5 * D, 0.38 * D * u[0], 0.3 * D * u[1])),
                 dict(kind="cylinder", op="difference", size=(0.15 * D, 0.6 * D), axis=2, center=(0.1 * D, 0, 0))]
    elif kind == 1:
        prims = [dict(kind="ellipsoid", size=(0.32 * D, 0.3 * D * u[0], 0.28 * D * u[1])),
                 dict(kind="cylinder", op="union", size=(0.09 * D, 0.5 * D), axis=0)]
    elif kind == 2:
        w = 0.12 * D * u[0]
        prims = [dict(kind="box", size=(0.5 * D, w, w)), dict(kind="box", op="union", size=(w, 0.45 * D, w)),
                 dict(kind="box", op="union", size=(w, w, 0.4 * D))]
    elif kind == 3:
        prims = [dict(kind="cylinder", size=(0.35 * D, 0.5 * D * u[0]), axis=1),
                 dict(kind="box", op="difference", size=(0.16 * D, 0.6 * D, 0.08 * D), center=(0.3 * D, 0, 0))]
    elif kind == 4:
        prims = [dict(kind="ellipsoid", size=(0.5 * D, 0.4 * D * u[0], 0.36 * D * u[1])),
                 dict(kind="sphere", op="difference", size=(0.18 * D,), center=(0.12 * D, 0, 0))]
    else:
        prims = [dict(kind="box", size=(0.4 * D * u[0], 0.4 * D * u[1], 0.4 * D * u[2])),
                 dict(kind="sphere", op="intersection", size=(0.5 * D,))]
    return S.csg_shape(prims, samples=64)


def polydisperse_csg(n: int = 384, n_small: int = 100000, n_big: int = 50, seed: int = DEFAULT_SEED,
                     d_small=(2.0, 4.0), d_big=(40.0, 80.0), with_velocity: bool = True) -> Scene:
    """Small spheres (diameter uniform in d_small cells) and large CSG solids (overall size uniform
    in d_big cells, the six csg_family kinds in turn) in the unit cube of n^3 cells, placed
    without overlap of their bounding spheres (the large ones first, with a gap of 2 cells);
    random orientations, v ~ N(0, 0.1), omega ~ N(0, 1).  The solids are the CSG ones first
    (index order), then the spheres.  Returns the scene; its `csg` entry is the CSG table."""
    rng = _rng(seed)
    h = 1.0 / n
    mesh = M.uniform_box((n, n, n), (0, 0, 0), (1, 1, 1))
    Db = rng.uniform(d_big[0], d_big[1], n_big) * h
    table = np.zeros(n_big, dtype=S.CSG_DTYPE)
    for b in range(n_big):
        table[b] = csg_family(rng, b % 6, Db[b])
    rb_big = table["radius_b"].astype(np.float64)
    pos_big = place_nonoverlapping(rng, rb_big, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 2.0 * h)
    d = rng.uniform(d_small[0], d_small[1], n_small)
    r_small = 0.5 * d * h
    pos_small = place_nonoverlapping(rng, r_small, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.0,
                                     avoid=(pos_big, rb_big + 2.0 * h))
    ns = n_big + n_small
    sol = S.make_solids(ns)
    sol["shape"][:n_big] = S.CSG
    sol["prm"][:n_big, 0] = np.arange(n_big)
    sol["center"][:n_big] = pos_big
    sol["shape"][n_big:] = S.SPHERE
    sol["prm"][n_big:, 0] = r_small
    sol["center"][n_big:] = pos_small
    sol["q"] = S.random_unit_quaternions(rng, ns)
    sol["vel"] = rng.normal(0.0, 0.1, size=(ns, 3))
    sol["omega"] = rng.normal(0.0, 1.0, size=(ns, 3))
    U = fourier_velocity(mesh_cell_centres(mesh), seed + 1) if with_velocity else None
    vol = float(table["volume"].sum() + (4.0 / 3.0 * math.pi * r_small ** 3).sum())
    info = dict(n=n, n_small=n_small, n_big=n_big, seed=seed, solid_volume_fraction=vol,
                d_big_cells=(float(Db.min() / h), float(Db.max() / h)))
    return Scene(f"polydisperse_csg_{n}_{n_small}_{n_big}", mesh, sol, U, False, 1.0, 1e-3, info, table)


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


def walls_3d(inset_y: float, inset_side: float = 0.0) -> np.ndarray:
    """Six plane solids of the unit cube, normals (local +y) into the fluid, the solid side of each
    the half-space beyond it: floor and ceiling `inset_y` inside the y faces, the four side walls
    `inset_side` inside theirs.  (A cell inside two walls is forced by both every step, which
    the explicit loop without a fluid solve does not survive where three walls meet: keep the
    side walls on the faces, where they own no cell, when the floor and ceiling are inset.)"""
    w = S.make_solids(6)
    w["shape"] = S.PLANE
    a, b = inset_y, inset_side
    spec = (((0.5, a, 0.5), (0.0, 0.0, 0.0)), ((0.5, 1 - a, 0.5), (0.0, 0.0, np.pi)),
            ((b, 0.5, 0.5), (0.0, 0.0, -np.pi / 2)), ((1 - b, 0.5, 0.5), (0.0, 0.0, np.pi / 2)),
            ((0.5, 0.5, b), (np.pi / 2, 0.0, 0.0)), ((0.5, 0.5, 1 - b), (-np.pi / 2, 0.0, 0.0)))
    for t, (c, eul) in enumerate(spec):
        w["center"][t] = c
        w["q"][t] = S.quaternion_xyz(eul)
    return w


def _quat_R(q):
    """Rotation matrices [n, 3, 3] of unit quaternions (w, x, y, z)."""
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    return np.stack([np.stack([w * w + x * x - y * y - z * z, 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
                     np.stack([2 * (x * y + w * z), w * w - x * x + y * y - z * z, 2 * (y * z - w * x)], -1),
                     np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), w * w - x * x - y * y + z * z], -1)], -2)


def place_ellipsoids(rng, axes, R, lo, hi, gap: float, max_tries: int = 20000) -> np.ndarray:
    """Random sequential placement of oriented ellipsoids (semi-axes `axes`, rotations `R`) in the
    box lo..hi, largest first, each grown by gap/2 per axis for the test.  Overlap test: the
    Perram-Wertheim contact function F(l) = l (1 - l) r^T [(1 - l) A + l B]^-1 r, A = R diag(a^2)
    R^T; two ellipsoids are apart iff max over l in [0, 1] of F >= 1 (sampled at 33 values of
    l, which can only under-estimate the maximum: a sampled pass is never an overlap)."""
    n = axes.shape[0]
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    ax = axes + 0.5 * gap
    M = np.einsum("nij,nj,nkj->nik", R, ax ** 2, R)
    rb = ax.max(axis=1)
    lam = np.linspace(0.0, 1.0, 33)[1:-1]
    order = np.argsort(-rb, kind="stable")
    cell = 2.0 * rb.max()
    ng = max(1, int(float((hi - lo).min()) / cell))
    grid: dict = {}
    pos = np.full((n, 3), np.nan)
    for idx in order:
        r = axes[idx].max()
        ok = False
        for _attempt in range(max_tries // 64):
            cand = rng.uniform(lo + r, hi - r, size=(64, 3))
            gi = np.minimum(((cand - lo) / (hi - lo) * ng).astype(np.int64), ng - 1)
            for c in range(64):
                nb = []
                for dz in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            lst = grid.get((gi[c][0] + dx, gi[c][1] + dy, gi[c][2] + dz))
                            if lst:
                                nb.extend(lst)
                if nb:
                    nb = np.asarray(nb)
                    dvec = pos[nb] - cand[c]
                    near = np.linalg.norm(dvec, axis=1) < rb[nb] + rb[idx]
                    if np.any(near):
                        nb, dvec = nb[near], dvec[near]
                        S_ = (1 - lam)[None, :, None, None] * M[idx][None, None] + \
                            lam[None, :, None, None] * M[nb][:, None]  # [m, L, 3, 3]
                        x = np.linalg.solve(S_, np.broadcast_to(dvec[:, None, :, None], S_.shape[:2] + (3, 1)))
                        F = lam * (1 - lam) * np.einsum("mi,mli->ml", dvec, x[..., 0])
                        if np.any(F.max(axis=1) < 1.0):
                            continue
                pos[idx] = cand[c]
                grid.setdefault(tuple(gi[c]), []).append(idx)
                ok = True
                break
            if ok:
                break
        if not ok:
            raise RuntimeError(f"could not place ellipsoid {idx} without overlap")
    return pos


def ellipsoids_jittered(n: int = 200, n_solids: int = 2000, jitter: float = 0.2, a_min: float = 3.0,
                        a_max: float = 10.0, omega_sigma: float = 0.5, u_rms: float = 0.02,
                        seed: int = DEFAULT_SEED) -> Scene:
    """Non-overlapping ellipsoids (semi-axes U[a_min, a_max] cells, at least h/2 apart) among six
    fixed walls on a jittered hex mesh; the walls are the last six solids.  Overlapping solids
    force their shared cells twice per step (U -= (Fs_A + Fs_B) dt), which the explicit loop
    without a fluid solve does not survive; so the packing starts apart and the spin and flow
    are gentle (omega ~ N(0, 0.5), rms |U| 0.02): at omega ~ N(0, 5) the packed ellipsoids
    turn into each other within ~100 steps."""
    rng = _rng(seed)
    h = 1.0 / n
    mesh = M.jittered_box((n, n, n), (0, 0, 0), (1, 1, 1), jitter, seed + 11)
    axes = rng.uniform(a_min, a_max, size=(n_solids, 3)) * h
    inset = 1.5 * h
    q = S.random_unit_quaternions(rng, n_solids)
    pos = place_ellipsoids(rng, axes, _quat_R(q), (inset + 0.5 * h,) * 3, (1 - inset - 0.5 * h,) * 3, 0.5 * h,
                           max_tries=200000)
    sol = S.make_solids(n_solids)
    sol["shape"] = S.ELLIPSOID
    sol["prm"][:, :3] = axes
    sol["center"] = pos
    sol["q"] = q
    sol["vel"] = 0.0  # the solids start at rest and are carried by the flow
    sol["omega"] = rng.normal(0.0, omega_sigma, size=(n_solids, 3))
    sol = np.concatenate([sol, walls_3d(inset, 0.0)])
    U = u_rms * fourier_velocity(mesh_cell_centres(mesh), seed + 1)
    vol = float(np.sum(4.0 / 3.0 * np.pi * np.prod(axes, axis=1)))
    info = dict(n=n, n_solids=n_solids, n_walls=6, seed=seed, jitter=jitter, solid_volume_fraction=vol,
                semi_axes_cells=(a_min, a_max))
    return Scene(f"ellipsoids_jittered_{n}_{n_solids}", mesh, sol, U, False, 1.0, 1e-3, info)
