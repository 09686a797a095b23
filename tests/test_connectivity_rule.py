"""The tile pass skips its flood fill when a cheap test on the rows of active cells says the
active set is connected (every nonzero row one run of cells; per slice k, nonzero rows minus
overlapping j-neighbour pairs equal to one; nonempty slices consecutive, each overlapping the one
before it in some row).  The test must never say "connected" for a disconnected set: checked here
against scipy's face-connectivity labelling on random sets, sets cut from random ellipsoids, and
sets built to be disconnected.  (The GPU tests compare the resulting candidate lists with the
oracle; this checks the rule itself on many more sets.)"""
from __future__ import annotations

import numpy as np
from scipy import ndimage


def rule_says_connected(act: np.ndarray) -> bool:
    """act[k, j, i] bool.  Restates the device test on 64-bit row words."""
    nz, ny, nx = act.shape
    rows = np.zeros((nz, ny), dtype=object)
    for k in range(nz):
        for j in range(ny):
            w = 0
            for i in range(nx):
                if act[k, j, i]:
                    w |= 1 << i
            rows[k, j] = w
    n_nonzero = n_links = 0
    slice_nz = [False] * nz
    slice_link = [False] * nz
    for k in range(nz):
        for j in range(ny):
            a = rows[k, j]
            if a == 0:
                continue
            t = a >> ((a & -a).bit_length() - 1)
            if t & (t + 1):
                return False  # more than one run
            n_nonzero += 1
            slice_nz[k] = True
            if j > 0 and a & rows[k, j - 1]:
                n_links += 1
            if k > 0 and a & rows[k - 1, j]:
                slice_link[k] = True
    ks = [k for k in range(nz) if slice_nz[k]]
    if not ks:
        return False
    if ks[-1] - ks[0] + 1 != len(ks):
        return False
    if any(not slice_link[k] for k in ks[1:]):
        return False
    return n_nonzero - n_links == len(ks)


def n_components(act):
    return ndimage.label(act)[1]


def _active_from_vertices(ins):
    """cells with >= 1 inside vertex from a vertex mask ins[k, j, i] of shape (nz+1, ny+1, nx+1)"""
    a = np.zeros(tuple(s - 1 for s in ins.shape), dtype=bool)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                a |= ins[dz:dz + a.shape[0], dy:dy + a.shape[1], dx:dx + a.shape[2]]
    return a


def test_random_sets_never_false_positive():
    rng = np.random.default_rng(0)
    said = 0
    for trial in range(3000):
        shape = tuple(rng.integers(1, 7, 3))
        act = rng.random(shape) < rng.uniform(0.2, 0.9)
        if rule_says_connected(act):
            said += 1
            assert n_components(act) == 1, act
    assert said > 100  # the rule does accept a fair share of connected random sets


def test_ellipsoid_sets():
    """Active sets of random thin and thick ellipsoids: never a false "connected"; the thick
    ones (all semi-axes >= 1.5 cells) are accepted."""
    rng = np.random.default_rng(1)
    accepted_thick = 0
    n_thick = 0
    for trial in range(400):
        thick = trial % 2 == 0
        axes = rng.uniform(1.5, 6.0, 3) if thick else np.r_[rng.uniform(2.0, 7.0), rng.uniform(0.05, 0.6, 2)]
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        n = int(np.ceil(2 * axes.max())) + 3
        c = n / 2 + rng.uniform(-0.5, 0.5, 3)
        g = np.stack(np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing="ij"), -1)
        p = (g - c) @ R  # local coordinates (rows are R^T (g - c))
        ins = ((p / axes) ** 2).sum(-1) < 1.0
        act = _active_from_vertices(ins)
        if not act.any():
            continue
        ok = rule_says_connected(act)
        if ok:
            assert n_components(act) == 1
        if thick:
            n_thick += 1
            accepted_thick += ok
    assert accepted_thick >= 0.9 * n_thick, (accepted_thick, n_thick)


def test_constructed_disconnected_sets_rejected():
    a = np.zeros((3, 4, 8), dtype=bool)
    a[0, 0, 0:3] = True
    a[0, 0, 5:7] = True  # two runs in one row
    assert not rule_says_connected(a)
    b = np.zeros((3, 4, 8), dtype=bool)
    b[0, 0, 0:3] = True
    b[0, 1, 4:6] = True  # j-neighbours that do not overlap
    assert not rule_says_connected(b) and n_components(b) == 2
    c = np.zeros((3, 4, 8), dtype=bool)
    c[0, 1, 2:4] = True
    c[2, 1, 2:4] = True  # a gap slice
    assert not rule_says_connected(c) and n_components(c) == 2
    d = np.zeros((3, 4, 8), dtype=bool)
    d[0, 1, 2:4] = True
    d[1, 2, 2:4] = True  # consecutive slices whose rows do not overlap
    assert not rule_says_connected(d) and n_components(d) == 2
    e = np.zeros((2, 4, 8), dtype=bool)
    e[0, 0, 0:2] = True
    e[0, 2, 0:2] = True
    e[1, 0:3, 0:2] = True  # connected through the next slice only: the rule declines (fill runs)
    assert n_components(e) == 1 and not rule_says_connected(e)
