"""The benchmark configuration at its full size (256^3 cells, 10,000 spheres and ellipsoids, the
bench's own generator): properties that do not depend on the size.

* parity on a sample of solids: the solids do not overlap, so each solid's candidate set,
  fractions and force/torque are its own -- the oracle run on 16 sampled solids over the same
  full mesh and velocity gives them (tolerances of tests/test_gpu_parity.py);
* force balance: the solids' forces sum to rho_f * sum(Fs V) over the mesh;
* the fraction field lies in [0, 1], cells wholly inside a solid (Ct >= 4) have As = 1, and
  sum(As V) is the solids' volume to the discretisation error of the cut cells (the apex method
  fills the cut cells of these 4-24-cell convex solids 1.7 % short; bound: 3 %)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from _helpers import split_candidates
from sdfibm_amd import mesh as M
from sdfibm_amd import scenes
from sdfibm_amd import solids as S
from sdfibm_amd.cloud import SolidCloud

pytestmark = pytest.mark.gpu

TOL = {"fp64": dict(a=1e-12, ft=1e-9, bal=1e-11), "fp32": dict(a=1e-6, ft=1e-4, bal=1e-5)}


@pytest.fixture(scope="module")
def full():
    print("[fullsize] generating the 256^3 / 10,000-solid scene", flush=True)
    sc = scenes.cloud_3d(256, 10000)
    rng = np.random.default_rng(17)
    sample = np.sort(rng.choice(sc.solids.size, 16, replace=False))
    m = sc.mesh
    om = oracle.OracleMesh(structured=(m.nx, m.ny, m.nz, m.points))
    ref = om.interact(sc.solids[sample], sc.U, rho_f=sc.rho_f, dt=sc.dt, capture=True, capacity=4_000_000)
    return sc, sample, ref


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_full_size_properties(full, prec):
    sc, sample, ref = full
    tol = TOL[prec]
    n = sc.mesh.n_cells
    V = 1.0 / n
    print(f"[fullsize] {prec} interact", flush=True)
    with SolidCloud(sc.mesh, rho_f=sc.rho_f, precision=prec) as c:
        c.set_solids(sc.solids)
        c.set_velocity(sc.U)
        ft = c.interact(0.0, sc.dt)
        parts = split_candidates(ref, sample.size)
        for t, s in enumerate(sample):
            counts, cells, alpha = c.candidates(int(s))
            rc, rcells, ralpha = parts[t]
            assert np.array_equal(counts[:3], rc[:3]), (s, counts, rc)
            assert np.array_equal(cells, rcells), s
            assert np.max(np.abs(alpha - ralpha), initial=0.0) <= tol["a"], s
        scale = np.maximum(np.abs(ref["ft"]), 1e-3 * np.abs(ref["ft"]).max())
        assert np.max(np.abs(ft[sample] - ref["ft"]) / scale) <= tol["ft"]
        f = c.fields()
    # force balance between the solids and the field
    total = ft[:, :3].sum(axis=0)
    field = sc.rho_f * f["Fs"].sum(axis=0) * V
    assert np.max(np.abs(total - field)) <= tol["bal"] * np.abs(ft[:, :3]).sum()
    As, Ct = f["As"], f["Ct"]
    assert As.min() >= 0.0 and As.max() <= 1.0
    assert np.all(As[Ct >= 4] == 1.0)
    vol = sum(S.shape_volume(int(x["shape"]), x["prm"]) for x in sc.solids)
    assert abs(As.sum() * V - vol) <= 3e-2 * vol


def test_surrogate_rules_full_size(full):
    """The learned fraction on the benchmark configuration with the shipped parameters.  With
    every cell falling back to the exact fraction (the network still evaluated), the sampled
    solids meet max|dAs| <= 1e-3 and force/torque within 1e-4 of the oracle -- the only rule of
    this fallback family that reaches that bar: it serves no cell.  With the prescribed rule
    (prediction in [0.02, 0.98] and gradients more than 30 degrees apart -> exact) the network
    serves nearly all cut cells, the fractions stay in [0, 1] and the deviation on the sample is
    printed (the network's own error, above the bar)."""
    from sdfibm_amd import surrogate as SG

    sc, sample, ref = full
    parts = split_candidates(ref, sample.size)
    scale = np.maximum(np.abs(ref["ft"]), 1e-3 * np.abs(ref["ft"]).max())
    p = SG.load_params()

    def sample_errors(c, ft):
        da = 0.0
        for t, s in enumerate(sample):
            counts, cells, alpha = c.candidates(int(s))
            rc, rcells, ralpha = parts[t]
            assert np.array_equal(counts[:3], rc[:3]) and np.array_equal(cells, rcells), s
            da = max(da, float(np.max(np.abs(alpha - ralpha), initial=0.0)))
        return da, float(np.max(np.abs(ft[sample] - ref["ft"]) / scale))

    with SolidCloud(sc.mesh, rho_f=sc.rho_f, precision="fp32") as c:
        c.set_solids(sc.solids)
        c.set_velocity(sc.U)
        c.set_surrogate(p, band=(-1e30, 1e30), max_angle_deg=-1.0)
        da, dft = sample_errors(c, c.interact(0.0, sc.dt))
        assert c.surrogate_count() == 0
        assert da <= 1e-3 and dft <= 1e-4, (da, dft)
        c.set_surrogate(p, band=(0.02, 0.98), max_angle_deg=30.0)
        ft = c.interact(0.0, sc.dt)
        served = c.surrogate_count()
        As = c.fields()["As"]
        da2, dft2 = sample_errors(c, ft)
        print(f"[fullsize] surrogate, prescribed rule: {served} cut cells served, sample max|dalpha| "
              f"{da2:.3g}, force/torque {dft2:.3g}", flush=True)
        assert served > 1_000_000
        assert As.min() >= 0.0 and As.max() <= 1.0
        assert np.isfinite(da2) and da2 < 0.6


def test_polydisperse_csg_full_size():
    """Configuration C4 at its full size: 384^3 cells, 100,000 spheres of 2-4 cells and 50 CSG
    solids of 40-80 cells (the bench's generator), FP32.  The same size-independent properties:
    force balance between the solids and the field, As in [0, 1] and 1 in the cells wholly inside
    a solid, every CSG solid on the big-tile pass with its cells, the candidate sets of 3 CSG
    solids and 20 spheres against the oracle restricted to a box around each (the solids do not
    overlap), and sum(As V) against the solids' volumes: the CSG solids' within 1 % (the oracle
    on all 50: 0.09 % short); the apex method fills the cut cells of 2-4-cell spheres short, by
    21.7 % on 400 of them in the oracle, so the spheres' sum lies in [0.74, 0.82] of theirs."""
    print("[fullsize] generating the 384^3 polydisperse CSG scene", flush=True)
    sc = scenes.polydisperse_csg(384, 100000, 50, with_velocity=False)
    n = sc.mesh.n_cells
    V = 1.0 / n
    cc = scenes.mesh_cell_centres(sc.mesh)
    U = np.stack([np.sin(6.0 * cc[:, 1]), np.cos(5.0 * cc[:, 2]), 0.5 * cc[:, 0]], axis=1)
    del cc
    with SolidCloud(sc.mesh, rho_f=sc.rho_f, precision="fp32") as c:
        c.set_csg_shapes(sc.csg)
        c.set_solids(sc.solids)
        c.set_velocity(U)
        print("[fullsize] interact", flush=True)
        c.set_phase_timing(True)
        ft = c.interact(0.0, sc.dt)
        assert c.timing()[6] > 0
        big_pairs = [int(c.candidates(s)[0][:3].sum()) for s in range(50)]
        assert min(big_pairs) > 5000
        f = c.fields()
        # sampled solids against the oracle on the part of the mesh around each one
        oracle.set_csg(sc.csg)
        m = sc.mesh
        for s in [0, 17, 41] + list(range(50, 100050, 5000)):
            sol = sc.solids[s:s + 1]
            R = float(sc.csg["radius_b"][s]) if s < 50 else float(sol["prm"][0, 0])
            lo = [max(0, int(np.floor((sol["center"][0, d] - R) * 384)) - 2) for d in range(3)]
            hi = [min(384, int(np.ceil((sol["center"][0, d] + R) * 384)) + 2) for d in range(3)]
            sub = M.rectilinear(m.xs[lo[0]:hi[0] + 1], m.ys[lo[1]:hi[1] + 1], m.zs[lo[2]:hi[2] + 1])
            om = oracle.OracleMesh(structured=(sub.nx, sub.ny, sub.nz, sub.points))
            li = np.arange(sub.n_cells)
            gi = (lo[0] + li % sub.nx) + 384 * ((lo[1] + (li // sub.nx) % sub.ny) + 384 * (lo[2] + li // (sub.nx * sub.ny)))
            ref = om.interact(sol, U[gi], rho_f=sc.rho_f, dt=sc.dt, capture=True, capacity=4 * sub.n_cells)
            counts, cells, alpha = c.candidates(s)
            assert np.array_equal(counts[:3], ref["counts"][0, :3]), s
            assert np.array_equal(cells, gi[ref["cells"][:int(ref["counts"][0, :3].sum())]]), s
            assert np.max(np.abs(alpha - ref["alpha"][:alpha.size]), initial=0.0) <= 1e-6, s
    total = ft[:, :3].sum(axis=0)
    field = sc.rho_f * f["Fs"].sum(axis=0) * V
    assert np.max(np.abs(total - field)) <= 1e-5 * np.abs(ft[:, :3]).sum()
    As, Ct = f["As"], f["Ct"]
    assert As.min() >= 0.0 and As.max() <= 1.0
    assert np.all(As[Ct >= 4] == 1.0)
    # volumes, split between the CSG solids (cells whose centre lies within radius_b + h of a CSG
    # solid's centre: the spheres keep 2 cells away from that ball) and the spheres
    A3 = As.reshape(384, 384, 384)  # [z, y, x]
    h = 1.0 / 384
    v_csg = 0.0
    for s in range(50):
        c0 = sc.solids["center"][s]
        R = float(sc.csg["radius_b"][s]) + h
        lo = [max(0, int(np.floor((c0[d] - R) * 384))) for d in range(3)]
        hi = [min(384, int(np.ceil((c0[d] + R) * 384))) for d in range(3)]
        x = (np.arange(lo[0], hi[0]) + 0.5) * h - c0[0]
        y = (np.arange(lo[1], hi[1]) + 0.5) * h - c0[1]
        z = (np.arange(lo[2], hi[2]) + 0.5) * h - c0[2]
        ball = z[:, None, None] ** 2 + y[None, :, None] ** 2 + x[None, None, :] ** 2 <= R * R
        v_csg += float(A3[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]][ball].sum(dtype=np.float64)) * V
    v_sph = As.sum(dtype=np.float64) * V - v_csg
    exact_csg = float(sc.csg["volume"].sum())
    exact_sph = float((4.0 / 3.0 * np.pi * sc.solids["prm"][50:, 0] ** 3).sum())
    print(f"[fullsize] volume ratios: CSG {v_csg / exact_csg:.5f}, spheres {v_sph / exact_sph:.5f}", flush=True)
    assert abs(v_csg / exact_csg - 1.0) <= 1e-2
    assert 0.74 <= v_sph / exact_sph <= 0.82
