"""Benchmark of the coupling step (SolidCloud::interact) on one MI355X.

Workload (default): uniform 256^3 hex mesh on the unit cube (16.8M cells), 10,000 solids
(70 % spheres, 30 % ellipsoids, lognormal diameters 4-24 cells, non-overlapping, random
orientations, v ~ N(0,0.1), omega ~ N(0,1)), divergence-free Fourier velocity field,
FP32 resident fields and fractions, FP64 inside tests and force/torque reductions.
A step is one interact() over all solids with U and the solids already resident in HBM.

`--workload sedimentation2d` runs the reference's CPU-runnable case instead: 100 discs
(diameter 0.0625, density ratio 1.5, gravity -10 y) and four plane walls on a 400 x 800 x 1 mesh,
FP64; it times one interact and 200 full coupling steps (interact, U -= Fs dt, evolve,
fixInternal; the fluid solve is not part of the path) and the oracle's interact on the host.

Prints ONE JSON line (rank 0):
  value        IBM cell-updates/s = sum over solids of candidate cells / time per step
  roofline     the dominant kernel's algorithmic HBM bytes / its average duration (HIP events
               on the stream the kernels run on) against the 8 TB/s HBM3E peak
  cpu_baseline the CPU oracle in reference-cost mode (an N-sized visited array, std::set and a
               hash-map vertex cache per solid, FP64, one thread) on a bounded sample of the
               same solids, on this host
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
METRIC = "IBM cell-updates/sec + % HBM peak; surrogate vs plain speedup at max|ΔAs|≤1e-3"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def candidate_stats(cloud, ns):
    """(pairs, ALL_INSIDE pairs, cut pairs, per-solid pairs) of the last interact."""
    per = np.zeros(ns, dtype=np.int64)
    n_inside = n_cut = 0
    for s in range(ns):
        counts, _, _ = cloud.candidates(s)
        n_inside += int(counts[0])
        n_cut += int(counts[1] + counts[2])
        per[s] = int(counts[:3].sum())
    return n_inside + n_cut, n_inside, n_cut, per


def touched_stats(mesh, Ct):
    """Nc (cells touched by any solid) and Nv (their distinct vertices)."""
    nx, ny, nz = mesh.nx, mesh.ny, mesh.nz
    t = (Ct.reshape(nz, ny, nx) != 0)
    nc = int(t.sum())
    v = np.zeros((nz + 1, ny + 1, nx + 1), dtype=bool)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                v[dz:dz + nz, dy:dy + ny, dx:dx + nx] |= t
    return nc, int(v.sum())


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by the plain-copy calibration), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    for ln in open(path):
        f = ln.split()
        if len(f) >= 4 and f[0] == "traffic" and f[1] == workload and f[2] == kernel:
            return float(f[3])
    return None


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the CPU restatement: the checker and the timed CPU baseline, never shipped

    return oracle


def oracle_mesh(scene):
    """The oracle's restatement of the scene's mesh (and CSG table), built once per run."""
    oracle = _oracle()
    if getattr(scene, "csg", None) is not None:
        oracle.set_csg(scene.csg)
    m = scene.mesh
    t0 = time.time()
    om = oracle.OracleMesh(structured=(m.nx, m.ny, m.nz, m.points))
    log(f"[cpu] oracle mesh built in {time.time() - t0:.1f}s")
    return om


def cpu_baseline(scene, n_sample, budget_s, om=None):
    """Reference-cost CPU oracle on the first solids in index order until the sample or the
    time budget is used; returns cell-updates/s and the sample actually timed."""
    om = om or oracle_mesh(scene)
    U = scene.U
    fields = None
    done = 0
    pairs = 0
    elapsed = 0.0
    step = max(1, n_sample // 8)
    while done < n_sample and elapsed < budget_s:
        s1 = min(n_sample, done + step)
        fields = dict(As=np.zeros(om.n_cells), Fs=np.zeros((om.n_cells, 3)), Ts=np.zeros(om.n_cells),
                      Ct=np.zeros(om.n_cells), ft=np.zeros((scene.solids.size, 6)))
        t = time.perf_counter()
        om.interact_range(scene.solids, U, done, s1, rho_f=scene.rho_f, dt=scene.dt, mode=1, fields=fields)
        elapsed += time.perf_counter() - t
        done = s1
    # candidate pairs of the timed solids (fast bookkeeping; not timed)
    r = om.interact(scene.solids[:done], U, rho_f=scene.rho_f, dt=scene.dt, mode=0, capture=True,
                    capacity=8 * om.n_cells)
    pairs = int(r["counts"][:, :3].sum())
    return pairs / elapsed, done, pairs, elapsed


def copy_bandwidth(nbytes=2 << 30, reps=10):
    """Bytes moved per second (read + write) by a plain device-to-device copy on this GPU."""
    import torch

    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) * 1e-3 / reps
    del a, b
    return 2.0 * nbytes / sec


def cpu_baseline_fast(scene, per_solid_pairs, threads, budget_s, om=None):
    """The oracle with fast bookkeeping (stamped visited array, shared vertex cache), the solids
    split over `threads` host threads (the C++ runs outside the GIL), each with its own field
    arrays: an optimised CPU reference point beside the reference-cost one.  Times the first
    solids in index order, as many as fit the budget (estimated from a one-thread probe)."""
    import threading

    om = om or oracle_mesh(scene)
    U = scene.U
    nsol = scene.solids.size

    def fields():
        return dict(As=np.zeros(om.n_cells), Fs=np.zeros((om.n_cells, 3)), Ts=np.zeros(om.n_cells),
                    Ct=np.zeros(om.n_cells), ft=np.zeros((nsol, 6)))

    probe = min(nsol, 100)
    f0 = fields()
    t = time.perf_counter()
    om.interact_range(scene.solids, U, 0, probe, rho_f=scene.rho_f, dt=scene.dt, mode=0, fields=f0)
    per_solid = (time.perf_counter() - t) / probe
    n_take = int(min(nsol, max(probe, budget_s / per_solid * threads)))
    bounds = np.linspace(0, n_take, threads + 1).astype(int)
    fs = [fields() for _ in range(threads)]
    ths = [threading.Thread(target=om.interact_range,
                            args=(scene.solids, U, int(bounds[k]), int(bounds[k + 1])),
                            kwargs=dict(rho_f=scene.rho_f, dt=scene.dt, mode=0, fields=fs[k]))
           for k in range(threads)]
    t = time.perf_counter()
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    el = time.perf_counter() - t
    pairs = int(per_solid_pairs[:n_take].sum())  # the GPU's candidate counts (bit-exact with the oracle's)
    return pairs / el, n_take, pairs, el


def cpu_baseline_slabs(scene, per_solid_pairs, ranks, n_sample, budget_s):
    """The reference's own parallel mode, emulated on this host: the mesh decomposed into `ranks`
    z-slabs (decomposePar's `simple` split along z), one thread per slab, each running the
    reference-cost serial algorithm over every solid on its own sub-mesh -- a per-rank seed (the
    nearest slab cell), the per-solid slab-sized visited array and the seed-fallback scan of the
    whole slab for solids that do not reach it (src/cellenumerator.cpp:47-63) -- with the
    per-solid force/torque partials summed at the end (src/solidcloud.cpp:427-431).  Times the
    first solids in index order; the wall time is the slowest rank's."""
    import threading

    from sdfibm_amd import mesh as M

    oracle = _oracle()
    if getattr(scene, "csg", None) is not None:
        oracle.set_csg(scene.csg)
    m = scene.mesh
    nx, ny, nz = m.nx, m.ny, m.nz
    bounds = np.linspace(0, nz, ranks + 1).astype(int)
    P = m.points.reshape(nz + 1, ny + 1, nx + 1, 3)
    U3 = scene.U.reshape(nz, ny, nx, 3)
    meshes, Us = [], []
    for r in range(ranks):
        k0, k1 = int(bounds[r]), int(bounds[r + 1])
        if k1 <= k0:
            continue
        pts = np.ascontiguousarray(P[k0:k1 + 1].reshape(-1, 3))
        meshes.append(oracle.OracleMesh(structured=(nx, ny, k1 - k0, pts)))
        Us.append(np.ascontiguousarray(U3[k0:k1].reshape(-1, 3)))
    nsol = scene.solids.size
    # sample size from a probe of one rank (the most expensive: the slab with the most solids)
    probe = min(nsol, 4)
    fields = lambda om: dict(As=np.zeros(om.n_cells), Fs=np.zeros((om.n_cells, 3)), Ts=np.zeros(om.n_cells),
                             Ct=np.zeros(om.n_cells), ft=np.zeros((nsol, 6)))
    t = time.perf_counter()
    meshes[0].interact_range(scene.solids, Us[0], 0, probe, rho_f=scene.rho_f, dt=scene.dt, mode=1,
                             fields=fields(meshes[0]))
    per_solid = (time.perf_counter() - t) / probe
    n_take = int(min(n_sample, max(probe, budget_s / max(per_solid, 1e-9))))
    fs = [fields(om) for om in meshes]
    ths = [threading.Thread(target=om.interact_range, args=(scene.solids, Us[k], 0, n_take),
                            kwargs=dict(rho_f=scene.rho_f, dt=scene.dt, mode=1, fields=fs[k]))
           for k, om in enumerate(meshes)]
    t = time.perf_counter()
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    el = time.perf_counter() - t
    ft = sum(f["ft"] for f in fs)  # the per-solid reductions across ranks
    pairs = int(per_solid_pairs[:n_take].sum())
    return pairs / el, n_take, pairs, el, len(meshes), ft


def fp32_error_statement(scene, om, n_sample=16):
    """Errors of the FP32 step against the FP64 oracle on the same full mesh and velocity, for a
    sample of the solids (spread over the index range) given alone to both: candidate lists,
    fractions (absolute; relative where alpha >= 1e-3), As and Fs per element (Fs relative where
    |Fs| >= 1e-3 of its largest value) and force/torque (relative to max(|F|, 1e-3 max|F|))."""
    from sdfibm_amd.cloud import SolidCloud

    ns = scene.solids.size
    pick = np.unique(np.linspace(0, ns - 1, n_sample).astype(int))
    sol = scene.solids[pick]
    ref = om.interact(sol, scene.U, rho_f=scene.rho_f, dt=scene.dt, capture=True, capacity=4000000)
    with SolidCloud(scene.mesh, two_d=scene.two_d, rho_f=scene.rho_f, precision="fp32") as c:
        c.set_solids(sol)
        c.set_velocity(scene.U)
        ft = c.interact(0.0, scene.dt)
        g = c.fields()
        k = 0
        da_abs, da_rel = [], []
        for t in range(pick.size):
            counts, c_cells, alpha = c.candidates(t)
            nref = int(ref["counts"][t, :3].sum())
            r_cells, r_alpha = ref["cells"][k:k + nref], ref["alpha"][k:k + nref]
            k += nref
            if not (np.array_equal(counts[:3], ref["counts"][t, :3]) and np.array_equal(c_cells, r_cells)):
                return {"error": f"candidate cells of solid {pick[t]} differ"}
            d = np.abs(alpha - r_alpha)
            da_abs.append(d.max(initial=0.0))
            big = r_alpha >= 1e-3
            if big.any():
                da_rel.append((d[big] / r_alpha[big]).max())
    touched = ref["Ct"] != 0
    Fg, Fr = g["Fs"][touched], ref["Fs"][touched]
    fl = 1e-3 * np.abs(Fr).max()
    m = np.abs(Fr) >= fl
    rel = np.abs(Fg - Fr)[m] / np.abs(Fr)[m]
    scale = np.maximum(np.abs(ref["ft"]), 1e-3 * np.abs(ref["ft"]).max())
    dAs = np.abs(g["As"] - ref["As"])
    return {"solids": int(pick.size), "cells": int(touched.sum()), "candidate_lists_bit_exact": True,
            "alpha_max_abs": float(max(da_abs)), "alpha_max_rel_alpha_ge_1e-3": float(max(da_rel)),
            "As_max_abs": float(dAs.max()),
            "Fs_max_rel_above_1e-3_of_max": float(rel.max()), "Fs_p99_rel": float(np.quantile(rel, 0.99)),
            "Fs_max_abs_over_max": float(np.abs(Fg - Fr).max() / np.abs(Fr).max()),
            "force_torque_max_rel": float(np.max(np.abs(ft - ref["ft"]) / scale))}


def phase_pass(cloud, dt, steps):
    """Mean phase times (ms) of `steps` interacts with phase timing on.  The HIP events between
    the phases order the kernels around them (about 25 us per step on cloud3d), so the timed
    loops run without them and the phases come from a pass of their own."""
    cloud.set_phase_timing(True)
    ph = np.zeros(7)
    for _ in range(steps):
        cloud.interact_device(0.0, dt)
        ph += cloud.timing()
    cloud.set_phase_timing(False)
    return ph / steps


def surrogate_comparison(cloud, sc, ft0, As0, steps):
    """The same step with the learned fraction (MLP 11-32-32-1, BF16 MFMA) serving the cut cells
    of spheres and ellipsoids, at three fallback rules: the prescribed one (prediction in
    [0.02, 0.98] and a gradient angle above 30 degrees -> exact fraction), every in-band prediction
    falling back, and every cell falling back (the network still evaluated).  Reports each rule's
    step-time ratio and deviation from the plain method's As and force/torque, and the speedup of
    the fastest rule within max|dAs| <= 1e-3."""
    import torch

    from sdfibm_amd import surrogate as SG

    if not os.path.exists(SG.PARAMS_PATH):
        return {"skipped": "no trained parameter file"}
    p = SG.load_params()
    rules = {"plain": None, "prescribed": ((0.02, 0.98), 30.0), "in_band_exact": ((0.02, 0.98), -1.0),
             "all_exact": ((-1e30, 1e30), -1.0)}
    scale = np.maximum(np.abs(ft0), 1e-3 * np.abs(ft0).max())
    res = {}
    for name, rule in rules.items():
        if rule is None:
            cloud.set_surrogate(None)
        else:
            band, angle = rule
            cloud.set_surrogate(p, band=band, max_angle_deg=angle)
        for _ in range(3):
            cloud.interact_device(0.0, sc.dt)
        torch.cuda.synchronize()
        k = max(5, min(steps, 20))
        t0 = time.perf_counter()
        for _ in range(k):
            cloud.interact_device(0.0, sc.dt)
        torch.cuda.synchronize()
        sg_ms = 1e3 * (time.perf_counter() - t0) / k
        ph = phase_pass(cloud, sc.dt, k) * k
        ft1 = cloud.interact(0.0, sc.dt)
        As1 = cloud.fields()["As"]
        served = cloud.surrogate_count() if rule is not None else 0
        cloud.set_surrogate(None)
        dAs = float(np.abs(As1.astype(np.float64) - As0).max())
        if rule is None:
            plain_ms, plain_cells_ms = sg_ms, ph[5] / k
            continue
        res[name] = {"band": list(band) if band[0] > -1e29 else "all", "max_angle_deg": angle,
                     "served_cut_cells": int(served), "surrogate_ms_per_step": sg_ms, "speedup": plain_ms / sg_ms,
                     "k_cells_ms": ph[5] / k, "max_abs_dAs": dAs,
                     "force_torque_max_rel_err": float(np.max(np.abs(ft1 - ft0) / scale)),
                     "meets_max_dAs_1e-3": dAs <= 1e-3}
    ok = [r for r in res.values() if r["meets_max_dAs_1e-3"]]
    best = max(ok, key=lambda r: r["speedup"]) if ok else None
    out = {"plain_ms_per_step": plain_ms, "plain_k_cells_ms": plain_cells_ms, "rules": res,
           "speedup_at_max_dAs_1e-3": best["speedup"] if best else None,
           "served_cut_cells_at_max_dAs_1e-3": best["served_cut_cells"] if best else None,
           "amdahl_bound_no_k_cells": plain_ms / max(plain_ms - plain_cells_ms, 1e-9)}
    log(f"[bench] surrogate: {out}")
    return out


def bench_ellipsoids_loop(args):
    """Jittered-hex ellipsoid cloud with six walls: the coupled step loop (interact, U -= Fs dt,
    evolve, fixInternal) as replays of one captured HIP graph, against the same steps issued one
    call at a time."""
    import torch

    from sdfibm_amd import scenes
    from sdfibm_amd.cloud import SolidCloud, motion_mask

    n, nsol = args.n or 200, args.solids or 2000
    t0 = time.time()
    sc = None
    if args.scene_cache and os.path.exists(args.scene_cache):
        z = np.load(args.scene_cache, allow_pickle=False)
        if int(z["n"]) == n and int(z["ns"]) == nsol:
            from sdfibm_amd import mesh as M
            mesh = M.jittered_box((n,) * 3, (0, 0, 0), (1, 1, 1), 0.2, scenes.DEFAULT_SEED + 11)
            sc = scenes.Scene(f"ellipsoids_jittered_{n}_{nsol}", mesh, z["solids"], z["U"], False, 1.0, 1e-3, {})
    if sc is None:
        sc = scenes.ellipsoids_jittered(n, nsol)
        if args.scene_cache:
            np.savez(args.scene_cache, n=n, ns=nsol, solids=sc.solids, U=sc.U)
    ns = sc.solids.size
    log(f"[bench] scene {sc.name} ready in {time.time() - t0:.1f}s")
    dt, g = sc.dt, np.zeros(3)
    steps = args.loop_steps or 500
    rho = np.full(ns, 1.5)
    masks = np.full(ns, -1, dtype=np.int32)
    masks[nsol:] = motion_mask("b000000")
    cloud = SolidCloud(sc.mesh, rho_f=sc.rho_f, precision=args.precision)
    cloud.set_solids(sc.solids)
    cloud.set_velocity(sc.U)
    cloud.set_dynamics(rho, masks)
    if args.contact:
        cloud.set_contact(*args.contact)
    log(f"[bench] uploaded in {time.time() - t0:.1f}s")
    # per-kernel phases of one interact of the initial state
    cloud.interact(0.0, dt)
    ph = phase_pass(cloud, dt, 5)
    f = cloud.fields()
    nc, nv = touched_stats(sc.mesh, f["Ct"])
    pairs_init = sum(cloud.pair_counts())
    cloud.run_steps(0.0, dt, max(args.warmup, 1), g)
    batches = 5
    per = max(1, steps // batches)
    pairs = [sum(cloud.pair_counts())]
    dev_ms = 0.0
    t = time.perf_counter()
    for b in range(batches):
        cloud.run_steps(b * per * dt, dt, per, g)
        dev_ms += cloud.timing()[3]
        pairs.append(sum(cloud.pair_counts()))
    wall_ms = 1e3 * (time.perf_counter() - t)
    steps = batches * per
    ms_per_step = dev_ms / steps
    n_pairs = float(np.mean(pairs))
    # the same step issued call by call (each call synchronises)
    k_eager = min(50, steps)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for k in range(k_eager):
        cloud.interact_device(k * dt, dt)
        cloud.update_velocity(dt)
        cloud.evolve(k * dt, dt, g)
        cloud.fix_internal()
    torch.cuda.synchronize()
    eager_ms = 1e3 * (time.perf_counter() - t) / k_eager
    fin = cloud.solids()
    cloud.close()
    # the same initial interact through the general polyMesh path (the mesh handed over as
    # points/faces/owner/neighbour; addressing and geometry derived by the library)
    poly = None
    if not args.no_poly:
        t = time.perf_counter()
        pm = sc.mesh.to_polymesh()
        pc = SolidCloud(pm, rho_f=sc.rho_f, precision=args.precision)
        setup_s = time.perf_counter() - t
        pc.set_solids(sc.solids)
        pc.set_velocity(sc.U)
        pc.interact(0.0, dt)
        pairs_poly = sum(pc.pair_counts())
        fp = pc.fields()
        d_as = float(np.max(np.abs(fp["As"] - f["As"])))
        ct_diff = int(np.count_nonzero(fp["Ct"] != f["Ct"]))
        pt = []
        for _ in range(5):
            t = time.perf_counter()
            pc.interact_device(0.0, dt)
            pt.append(1e3 * (time.perf_counter() - t))
        tim = pc.timing()
        pc.close()
        poly = {"interact_ms": float(np.median(pt)), "device_ms": float(tim[3]),
                "phases_ms": {"pairs_labels_force": float(tim[0]), "cell_lists": float(tim[1]), "fields": float(tim[2])},
                "candidate_pairs": int(pairs_poly), "structured_candidate_pairs": int(pairs_init),
                "max_abs_dAs_vs_structured": d_as, "Ct_cells_differing_from_structured": ct_diff,
                "structured_interact_ms": float(ph[3]), "setup_s": setup_s,
                "note": "interact of the initial state, eager (three counts read back per step)"}
    rb = 4 if args.precision == "fp32" else 8
    N = sc.mesh.n_cells
    kernels = {"k_fields": {"ms": ph[2], "bytes": (5 * rb + 4) * N + 3 * rb * nc},
               "k_cells": {"ms": ph[5], "bytes": 3 * rb * nc + 3 * rb * nv}}
    dom = max(kernels, key=lambda k: kernels[k]["ms"])
    achieved = kernels[dom]["bytes"] / (kernels[dom]["ms"] * 1e-3)
    line = {
        "metric": METRIC, "value": n_pairs / (ms_per_step * 1e-3), "unit": "cell-updates/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": args.precision, "data": "synthetic (seeded generator)",
        "config": {"workload": f"ellipsoids3d_jittered {n}^3 hex (interior vertices jittered U(+-0.2h)), {nsol} "
                               f"ellipsoids (semi-axes U[3,10] cells, at rest, omega ~ N(0,0.5)) + 6 fixed plane walls, "
                               f"coupled step loop (interact, U -= Fs dt, evolve, fixInternal) as HIP graph "
                               f"replays, dt {dt}, {args.precision} fields",
                   "cells": N, "solids": int(ns), "candidate_pairs_mean": n_pairs,
                   "candidate_pairs_samples": pairs, "contact": list(args.contact) if args.contact else None,
                   "parallelism": "single GPU"},
        "ms_per_step_value_basis": "device time of the graph replays (HIP events), per coupled step",
        "wall_ms_per_step": wall_ms / steps,
        "eager_ms_per_step": eager_ms, "graph_speedup_vs_eager": eager_ms / ms_per_step,
        "phases_ms_initial_interact": {"setup_bin": ph[0], "k_classify": ph[4], "k_cells": ph[5], "k_fields": ph[2]},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": measured_traffic(dom, f"ellipsoids3d_jittered_{n}_{nsol}_{args.precision}"),
                     "algorithmic_bytes": kernels[dom]["bytes"]},
        "surrogate": "not applicable: the learned fraction serves cubic cells of rectilinear meshes only",
        "polymesh_path": poly,
        "mean_displacement": float(np.mean(np.linalg.norm(fin["center"][:nsol] - sc.solids["center"][:nsol], axis=1))),
    }
    if not args.no_cpu:
        rate, ns_t, pairs_t, el = cpu_baseline(sc, min(args.cpu_sample, 200), args.cpu_budget)
        line["cpu_baseline"] = {"value": rate, "unit": "cell-updates/s", "cores": 1, "kind": "port",
                                "sample": f"interact of the first {ns_t} of {ns} solids on the full mesh "
                                          f"({pairs_t} cell-updates, {el:.1f}s), reference-cost bookkeeping, "
                                          f"FP64, 1 thread",
                                "speedup": line["value"] / rate}
    return line


def host_info():
    """Cores and model of this host's CPU (the CPU baselines run here)."""
    model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return {"nproc": cores, "cpu_model": model}


def bench_polydisperse(args):
    """Configuration C4: small spheres (2-4 cells) and large CSG solids (40-80 cells) on a 384^3
    mesh: one interact per step, the size bins (the on-chip tile pass and the big-tile pass)
    timed apart, the autotuner's choice, and full-size checks of the result."""
    import torch

    from sdfibm_amd import scenes
    from sdfibm_amd.cloud import SolidCloud

    n = args.n if args.n != 256 else 384
    n_small = args.solids if args.solids != 10000 else 100000
    n_big = 50
    t0 = time.time()
    sc = None
    if args.scene_cache and os.path.exists(args.scene_cache):
        z = np.load(args.scene_cache, allow_pickle=False)
        if "csg" in z.files and int(z["n"]) == n and int(z["ns"]) == n_small:
            from sdfibm_amd import mesh as M
            sc = scenes.Scene(f"polydisperse_csg_{n}_{n_small}_{n_big}", M.uniform_box((n,) * 3, (0, 0, 0), (1, 1, 1)),
                              z["solids"], z["U"], False, 1.0, 1e-3, json.loads(str(z["info"])), z["csg"])
    if sc is None:
        sc = scenes.polydisperse_csg(n, n_small, n_big)
        if args.scene_cache:
            np.savez(args.scene_cache, n=n, ns=n_small, solids=sc.solids, U=sc.U, csg=sc.csg,
                     info=json.dumps(sc.info))
    log(f"[bench] scene {sc.name} ready in {time.time() - t0:.1f}s: {sc.info}")
    N = sc.mesh.n_cells
    ns = sc.solids.size
    cloud = SolidCloud(sc.mesh, rho_f=sc.rho_f, precision=args.precision)
    cloud.set_csg_shapes(sc.csg)
    cloud.set_solids(sc.solids)
    cloud.set_velocity(sc.U)
    tune = None if args.no_autotune else cloud.autotune(sc.dt, args.tune_cache or None)
    log(f"[bench] autotune: {tune}")
    for _ in range(args.warmup):
        cloud.interact_device(0.0, sc.dt)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        cloud.interact_device(0.0, sc.dt)
    torch.cuda.synchronize()
    ms_per_step = 1e3 * (time.perf_counter() - t) / args.steps
    ph = phase_pass(cloud, sc.dt, args.steps)
    cloud.interact_steps(0.0, sc.dt, args.steps)
    graph_ms = cloud.timing()[3] / args.steps
    ft = cloud.interact(0.0, sc.dt)
    f = cloud.fields()
    pairs_in, pairs_cut = cloud.pair_counts()
    big_pairs = [int(cloud.candidates(s)[0][:3].sum()) for s in range(n_big)]
    n_pairs = pairs_in + pairs_cut
    nc, nv = touched_stats(sc.mesh, f["Ct"])
    rb = 4 if args.precision == "fp32" else 8
    kernels = {"k_fields": {"ms": ph[2], "bytes": (5 * rb + 4) * N + 3 * rb * nc}}
    achieved = kernels["k_fields"]["bytes"] / (ph[2] * 1e-3)
    # full-size checks: force balance (the solids do not overlap), As range, volume
    V = (1.0 / n) ** 3
    F_cells = sc.rho_f * f["Fs"].sum(axis=0) * V
    F_solids = ft[:, :3].sum(axis=0)
    vol_solids = float(sc.csg["volume"].sum() + (4.0 / 3.0 * math.pi * sc.solids["prm"][n_big:, 0] ** 3).sum())
    small_ms = ph[4] + ph[5]
    line = {
        "metric": METRIC, "value": n_pairs / (ms_per_step * 1e-3), "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": args.precision, "data": "synthetic (seeded generator)",
        "config": {"workload": f"polydisperse_csg {n}^3 hex, {n_small} spheres (d ~ U[2,4] cells) + {n_big} CSG "
                               f"solids (unions/differences/intersections of boxes, cylinders, ellipsoids, spheres; "
                               f"40-80 cells), plain method, {args.precision} fields, fp64 reductions",
                   "cells": N, "solids": int(ns), "candidate_pairs": n_pairs, "inside_pairs": pairs_in,
                   "cut_pairs": pairs_cut, "touched_cells": nc,
                   "solid_volume_fraction": float(sc.info["solid_volume_fraction"]), "parallelism": "single GPU"},
        "graph_ms_per_step": graph_ms,
        "phases_ms": {"setup_bin": ph[0], "k_classify": ph[4], "k_cells": ph[5], "big_tile_pass": ph[6],
                      "k_fields": ph[2]},
        "size_bins": {
            "small": {"solids": int(ns - n_big), "candidate_pairs": int(n_pairs - sum(big_pairs)),
                      "ms": small_ms, "kernels": "k_classify (LDS bit rows) + k_cells, main stream"},
            "big": {"solids": n_big, "candidate_pairs": int(sum(big_pairs)), "ms": ph[6],
                    "kernels": "k_big_* (global-memory labelling, chunked cut cells), second stream",
                    "largest_solid_pairs": int(max(big_pairs))},
            "load_balance": {"big_over_small_ms": ph[6] / small_ms if small_ms > 0 else None,
                             "tile_phase_ms": ph[1], "critical_path": "big" if ph[6] > small_ms else "small",
                             "pairs_per_us": {"small": (n_pairs - sum(big_pairs)) / (small_ms * 1e3),
                                              "big": sum(big_pairs) / (ph[6] * 1e3) if ph[6] > 0 else None}},
        },
        "autotune": tune,
        "roofline": {"bound": "hbm", "kernel": "k_fields", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": measured_traffic("k_fields", f"polydisperse_csg_{n}_{n_small}_{n_big}_{args.precision}"),
                     "algorithmic_bytes": kernels["k_fields"]["bytes"]},
        "checks": {"force_balance_rel": float(np.abs(F_cells - F_solids).max() / max(np.abs(F_solids).max(), 1e-300)),
                   "As_min": float(f["As"].min()), "As_max": float(f["As"].max()),
                   "sum_As_V_over_solid_volume": float(f["As"].sum() * V / vol_solids)},
    }
    cloud.close()
    if not args.no_cpu:
        rate, ns_t, pairs_t, el = cpu_baseline(sc, args.cpu_sample, args.cpu_budget)
        line["cpu_baseline"] = {"value": rate, "unit": "cell-updates/s", "cores": 1, "kind": "port",
                                "sample": f"first {ns_t} of {ns} solids in index order (the {n_big} CSG solids "
                                          f"first) on the full {n}^3 mesh ({pairs_t} cell-updates, {el:.1f}s), "
                                          f"reference-cost bookkeeping, FP64, 1 thread",
                                "speedup": line["value"] / rate, "host": host_info()}
    return line


def bench_sedimentation(args):
    """The CPU-runnable case: one interact and a loop of full coupling steps, FP64."""
    import torch

    from sdfibm_amd import scenes
    from sdfibm_amd.cloud import SolidCloud, motion_mask

    sc = scenes.sedimentation_2d(400, 800, 100)
    ns = sc.solids.size
    dt = 5e-4  # examples/sedimentation/system/controlDict
    g = np.array(sc.info["gravity"])
    rho = np.full(ns, 1.5)
    masks = np.full(ns, motion_mask("b110001"), dtype=np.int32)
    masks[100:] = motion_mask("b000000")
    cloud = SolidCloud(sc.mesh, two_d=True, rho_f=1.0, precision="fp64")
    cloud.set_solids(sc.solids)
    cloud.set_velocity(sc.U)
    cloud.set_dynamics(rho, masks)
    for _ in range(args.warmup):
        cloud.interact_device(0.0, dt)
    torch.cuda.synchronize()
    t = time.perf_counter()
    reps = max(args.steps, 1)
    for _ in range(reps):
        cloud.interact_device(0.0, dt)
    torch.cuda.synchronize()
    ms_interact = 1e3 * (time.perf_counter() - t) / reps
    n_pairs = sum(int(cloud.candidates(s)[0][:3].sum()) for s in range(ns))
    t = time.perf_counter()
    for k in range(args.loop_steps):
        cloud.interact_device(k * dt, dt)
        cloud.update_velocity(dt)
        cloud.evolve(k * dt, dt, g)
        cloud.fix_internal()
    torch.cuda.synchronize()
    ms_step = 1e3 * (time.perf_counter() - t) / args.loop_steps
    fin = cloud.solids()
    # the same loop continued as replays of one captured HIP graph
    cloud.run_steps(args.loop_steps * dt, dt, 2, g)  # capture + warm-up
    t = time.perf_counter()
    cloud.run_steps((args.loop_steps + 2) * dt, dt, args.loop_steps, g)
    ms_graph = 1e3 * (time.perf_counter() - t) / args.loop_steps
    dev_graph = cloud.timing()[3] / args.loop_steps
    # and with the contact stage on (discs and walls; the reference's linear law, k = 1e4, plus
    # damping): its cost per step inside the graph
    cloud.set_contact(1e4, 2.0)
    cloud.run_steps((2 * args.loop_steps + 2) * dt, dt, 2, g)
    cloud.run_steps((2 * args.loop_steps + 4) * dt, dt, args.loop_steps, g)
    dev_graph_contact = cloud.timing()[3] / args.loop_steps
    cloud.close()
    line = {
        "metric": METRIC, "value": n_pairs / (ms_interact * 1e-3), "unit": "cell-updates/s", "n_gpus": 1,
        "steps": reps, "warmup": args.warmup, "ms_per_step": ms_interact, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp64", "data": "synthetic (seeded generator)",
        "config": {"workload": "sedimentation2d 400x800x1, 100 discs d=0.0625 + 4 plane walls, rho_s/rho_f 1.5, "
                               "g = -10 y, dt 5e-4, fp64", "cells": sc.mesh.n_cells, "solids": int(ns),
                   "candidate_pairs": n_pairs, "parallelism": "single GPU"},
        "loop": {"steps": args.loop_steps, "ms_per_coupling_step": ms_step,
                 "mean_disc_drop": float(np.mean(sc.solids["center"][:100, 1] - fin["center"][:100, 1])),
                 "graph_ms_per_coupling_step": ms_graph, "graph_device_ms_per_step": dev_graph,
                 "graph_speedup": ms_step / ms_graph,
                 "graph_device_ms_per_step_with_contacts": dev_graph_contact},
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle  # the CPU restatement: timed here as the baseline, never shipped

        om = oracle.OracleMesh(structured=(sc.mesh.nx, sc.mesh.ny, sc.mesh.nz, sc.mesh.points))
        t = time.perf_counter()
        om.interact(sc.solids, sc.U, two_d=True, rho_f=1.0, dt=dt, mode=1)
        el = time.perf_counter() - t
        line["cpu_baseline"] = {"value": n_pairs / el, "unit": "cell-updates/s", "cores": 1, "kind": "port",
                                "sample": f"one full interact ({n_pairs} cell-updates, {el:.2f}s), reference-cost "
                                          "bookkeeping, FP64, 1 thread",
                                "speedup": line["value"] * el / n_pairs}
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--solids", type=int, default=10000)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--cpu-sample", type=int, default=400)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the fast and z-slab CPU baselines (0: this host's cores, at most 16)")
    ap.add_argument("--no-fp64", action="store_true", help="cloud3d fp32: skip the FP64 point")
    ap.add_argument("--scene-cache", default="", help="npz file to reuse the generated scene")
    ap.add_argument("--workload", default="cloud3d",
                    choices=["cloud3d", "sedimentation2d", "ellipsoids3d_jittered", "polydisperse_csg"])
    ap.add_argument("--tune-cache", default="", help="autotuner cache file (JSON lines); default: time every run")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default variants (profiling runs)")
    ap.add_argument("--loop-steps", type=int, default=None,
                    help="coupled steps timed (sedimentation2d: 200, ellipsoids3d_jittered: 500)")
    ap.add_argument("--contact", type=float, nargs=2, default=None, metavar=("K", "C"),
                    help="ellipsoids3d_jittered: contact stiffness and damping (default: no contacts)")
    ap.add_argument("--small-spheres", action="store_true",
                    help="cloud3d: spheres of uniform diameter 2-4 cells instead of the lognormal mix")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="cloud3d fp32: skip the learned-fraction comparison run")
    ap.add_argument("--no-poly", action="store_true",
                    help="ellipsoids3d_jittered: skip the same interact through the general polyMesh path")
    ap.add_argument("--copy-test", type=float, default=0.0,
                    help="only run a plain device copy of this many GB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the coupling step is benchmarked on a single MI355X "
              "(a multi-GPU run would shard the mesh into z-slabs; not part of this build).",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from sdfibm_amd import scenes
    from sdfibm_amd.cloud import SolidCloud

    if args.copy_test > 0:
        # counter calibration: a device-to-device copy and an elementwise negation, each
        # reading and writing 4 bytes per element exactly once
        n = int(args.copy_test * 1e9) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        for _ in range(5):
            torch.neg(a, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 5}), flush=True)
        return

    if args.workload == "sedimentation2d":
        args.loop_steps = args.loop_steps or 200
        print(json.dumps(bench_sedimentation(args)), flush=True)
        return
    if args.workload == "polydisperse_csg":
        print(json.dumps(bench_polydisperse(args)), flush=True)
        return
    if args.workload == "ellipsoids3d_jittered":
        if args.n == 256 and args.solids == 10000:  # the cloud3d defaults
            args.n, args.solids = 200, 2000
        print(json.dumps(bench_ellipsoids_loop(args)), flush=True)
        return

    t0 = time.time()
    sc = None
    if args.scene_cache and os.path.exists(args.scene_cache):
        z = np.load(args.scene_cache, allow_pickle=False)
        small = int(z["small"]) if "small" in z.files else 0
        if int(z["n"]) == args.n and int(z["ns"]) == args.solids and small == int(args.small_spheres):
            from sdfibm_amd import mesh as M
            sc = scenes.Scene(f"cloud3d_{args.n}_{args.solids}",
                              M.uniform_box((args.n,) * 3, (0, 0, 0), (1, 1, 1)), z["solids"], z["U"],
                              False, 1.0, 1e-3, json.loads(str(z["info"])))
    if sc is None:
        if args.small_spheres:  # 2-4 cell spheres, uniform diameters
            sc = scenes.cloud_3d(args.n, args.solids, frac_sphere=1.0, d_min=2.0, d_max=4.0, d_uniform=True)
        else:
            sc = scenes.cloud_3d(args.n, args.solids)
        if args.scene_cache:
            np.savez(args.scene_cache, n=args.n, ns=args.solids, small=int(args.small_spheres), solids=sc.solids, U=sc.U,
                     info=json.dumps({k: (float(v) if isinstance(v, (np.floating,)) else v)
                                      for k, v in sc.info.items()}))
    log(f"[bench] scene {sc.name} ready in {time.time() - t0:.1f}s: {sc.info}")
    N = sc.mesh.n_cells
    cloud = SolidCloud(sc.mesh, two_d=sc.two_d, rho_f=sc.rho_f, precision=args.precision)
    cloud.set_solids(sc.solids)
    cloud.set_velocity(sc.U)
    tune = None if args.no_autotune else cloud.autotune(sc.dt, args.tune_cache or None)
    log(f"[bench] autotune: {tune}")
    for _ in range(args.warmup):
        cloud.interact_device(0.0, sc.dt)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        cloud.interact_device(0.0, sc.dt)
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t_start
    ms_per_step = 1e3 * t_total / args.steps
    phases = phase_pass(cloud, sc.dt, args.steps)
    log(f"[bench] {ms_per_step:.3f} ms/step; phases (setup+bin, tiles, fields, sum, classify, cells) ms = {phases}")

    # the same interact as replays of one captured HIP graph (no host synchronisation between steps)
    cloud.interact_steps(0.0, sc.dt, args.steps)
    graph_ms = cloud.timing()[3] / args.steps
    # a host-coupled caller's step: the solids' new poses (set_solids, same shapes) and the
    # interact with the force/torque copy back; and with U handed over from the host as well
    k_h = 10
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(k_h):
        cloud.set_solids(sc.solids)
        cloud.interact(0.0, sc.dt)
    host_ms = 1e3 * (time.perf_counter() - t) / k_h
    t = time.perf_counter()
    for _ in range(k_h):
        cloud.set_solids(sc.solids)
        cloud.set_velocity(sc.U)
        cloud.interact(0.0, sc.dt)
    host_u_ms = 1e3 * (time.perf_counter() - t) / k_h
    ft = cloud.interact(0.0, sc.dt)
    f = cloud.fields()
    n_pairs, n_inside, n_cut, per_solid = candidate_stats(cloud, sc.solids.size)
    nc, nv = touched_stats(sc.mesh, f["Ct"])
    value = n_pairs / (ms_per_step * 1e-3)
    rb = 4 if args.precision == "fp32" else 8
    # Algorithmic HBM bytes (each array element counted once): the field pass writes
    # As, Fs(3), Ts (rb each) and Ct (int32) for every cell and reads U(3) of the touched
    # cells; the cut-cell/forcing pass reads U(3) of the touched cells and the vertices (3)
    # of the touched cells in the reference's mesh model; the whole step adds the cell
    # centre (3) and volume of the touched cells.
    B_total = (5 * rb + 4) * N + (3 * rb + 3 * rb + rb) * nc + 3 * rb * nv
    kernels = {
        "k_fields": {"ms": phases[2], "bytes": (5 * rb + 4) * N + 3 * rb * nc},
        "k_cells": {"ms": phases[5], "bytes": 3 * rb * nc + 3 * rb * nv},
        "k_classify": {"ms": phases[4], "bytes": None},
        "setup_bin": {"ms": phases[0], "bytes": None},
    }
    dom = max(kernels, key=lambda k: kernels[k]["ms"] if kernels[k]["bytes"] else -1.0)
    achieved = kernels[dom]["bytes"] / (kernels[dom]["ms"] * 1e-3)
    traffic = measured_traffic(dom, f"cloud3d_{args.n}_{args.solids}_{args.precision}")
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic (seeded generator)",
        "config": {"workload": f"cloud3d {args.n}^3 hex, {args.solids} solids "
                               + ("(spheres, d ~ U[2,4] cells), " if args.small_spheres else
                                  "(70% sphere / 30% ellipsoid), ")
                               + f"plain method, {args.precision} fields, fp64 reductions",
                   "cells": N, "solids": int(sc.solids.size), "candidate_pairs": n_pairs,
                   "inside_pairs": n_inside, "cut_pairs": n_cut, "touched_cells": nc,
                   "touched_vertices": nv, "solid_volume_fraction": float(sc.info["solid_volume_fraction"]),
                   "parallelism": "single GPU"},
        "phases_basis": "HIP events between the phases, in a pass of its own after the timed loop "
                        "(the events order the kernels around them; the timed loop runs without them)",
        "phases_ms": {"setup_bin": phases[0], "k_classify": phases[4], "k_cells": phases[5],
                      "k_fields": phases[2]},
        "graph_ms_per_step": graph_ms,
        "host_coupled_ms_per_step": {"set_solids+interact+ft_copy": host_ms,
                                     "set_solids+set_velocity+interact+ft_copy": host_u_ms},
        "autotune": tune,
        "mesh_cells_per_s": N / (ms_per_step * 1e-3),
        "hbm_frac_whole_step": B_total / (ms_per_step * 1e-3) / HBM_PEAK,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "algorithmic_bytes": kernels[dom]["bytes"]},
    }
    copy_bw = copy_bandwidth()
    line["roofline"]["copy_peak_measured"] = copy_bw / 1e9
    line["roofline"]["frac_of_copy_peak"] = achieved / copy_bw
    # the two kernels with algorithmic bytes run within a few per cent of each other: both lines
    per_kernel = {}
    for k in ("k_fields", "k_cells"):
        a = kernels[k]["bytes"] / (kernels[k]["ms"] * 1e-3)
        per_kernel[k] = {"ms": kernels[k]["ms"], "achieved": a / 1e9, "frac": a / HBM_PEAK,
                         "algorithmic_bytes": kernels[k]["bytes"],
                         "traffic": measured_traffic(k, f"cloud3d_{args.n}_{args.solids}_{args.precision}")}
    per_kernel["k_cells"]["limit"] = "VALU issue (about two thirds of the SIMD issue rate) and the scattered U gathers of the candidate cells (measured traffic about 3x the algorithmic bytes)"
    line["roofline"]["per_kernel"] = per_kernel
    if args.precision == "fp32" and not args.no_surrogate:
        line["surrogate"] = surrogate_comparison(cloud, sc, ft, f["As"], args.steps)
    if args.precision == "fp32" and not args.no_fp64:
        # the FP64 point of the same scene (FP64 fractions and fields)
        with SolidCloud(sc.mesh, two_d=sc.two_d, rho_f=sc.rho_f, precision="fp64") as c64:
            c64.set_solids(sc.solids)
            c64.set_velocity(sc.U)
            c64.autotune(sc.dt)
            for _ in range(2):
                c64.interact_device(0.0, sc.dt)
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(args.steps):
                c64.interact_device(0.0, sc.dt)
            torch.cuda.synchronize()
            ms64 = 1e3 * (time.perf_counter() - t) / args.steps
        line["fp64_point"] = {"ms_per_step": ms64, "value": n_pairs / (ms64 * 1e-3), "unit": "cell-updates/s"}
    if not args.no_cpu:
        host = host_info()
        om = oracle_mesh(sc)
        if args.precision == "fp32":
            line["fp32_vs_fp64_oracle"] = fp32_error_statement(sc, om)
            log(f"[bench] fp32 errors: {line['fp32_vs_fp64_oracle']}")
        rate, ns_t, pairs_t, el = cpu_baseline(sc, args.cpu_sample, args.cpu_budget, om=om)
        line["cpu_baseline"] = {"value": rate, "unit": "cell-updates/s", "cores": 1, "kind": "port",
                                "sample": f"first {ns_t} of {sc.solids.size} solids in index order on the full "
                                          f"{args.n}^3 mesh ({pairs_t} cell-updates, {el:.1f}s), reference-cost "
                                          f"bookkeeping, FP64, 1 thread",
                                "speedup": value / rate, "host": host}
        threads = int(args.cpu_threads) if args.cpu_threads > 0 else min(host["nproc"], 16)
        rate_f, ns_f, pairs_f, el_f = cpu_baseline_fast(sc, per_solid, threads, args.cpu_budget / 2, om=om)
        line["cpu_baseline_fast"] = {"value": rate_f, "unit": "cell-updates/s", "cores": threads, "kind": "port",
                                     "sample": f"first {ns_f} of {sc.solids.size} solids ({pairs_f} cell-updates, "
                                               f"{el_f:.1f}s), fast bookkeeping, FP64, solids split over "
                                               f"{threads} threads",
                                     "speedup": value / rate_f, "host": host}
        del om
        rate_s, ns_s, pairs_s, el_s, nr, _ = cpu_baseline_slabs(sc, per_solid, threads, args.cpu_sample,
                                                                  args.cpu_budget)
        line["cpu_baseline_slabs"] = {"value": rate_s, "unit": "cell-updates/s", "cores": nr, "kind": "port",
                                      "sample": f"first {ns_s} of {sc.solids.size} solids on {nr} z-slab sub-meshes "
                                                f"run concurrently ({pairs_s} cell-updates, {el_s:.1f}s): the "
                                                f"reference's MPI decomposition emulated, reference-cost "
                                                f"bookkeeping per rank, FP64",
                                      "speedup": value / rate_s, "host": host}
        best = max(rate, rate_s)
        line["speedup_vs_reference_cpu"] = {"value": value / best, "basis": "best of the one-thread and the "
                                            "z-slab reference-cost runs", "gate": 30.0, "meets": value / best >= 30.0}
    cloud.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
