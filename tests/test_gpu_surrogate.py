"""Learned fraction surrogate on the GPU: the MFMA inference and the training kernels' gradient
against the numpy reference (oracle/surrogate_ref.py), training, and its use in interact.

Tolerances: inference accumulates in FP32 in another order than the reference, which can move a
BF16 rounding of a hidden activation by one unit (2^-8 relative): 99 % of outputs within 1e-5 and
all within 2e-2 of the output scale.  Gradients: FP32 sums of up to 4096 terms, 1e-3 relative to
the gradient's norm per component."""
from __future__ import annotations

import numpy as np
import pytest

import surrogate_ref as R
from sdfibm_amd import scenes
from sdfibm_amd import surrogate as SG
from sdfibm_amd.cloud import SolidCloud

pytestmark = pytest.mark.gpu


def _params(seed):
    p = SG.init_params(seed)
    rng = np.random.default_rng(seed)
    p += rng.normal(0, 0.05, p.size).astype(np.float32)  # non-zero biases
    return p


def test_samples_distribution():
    X, y = SG.samples(7, 50000)
    assert np.all(np.isfinite(X)) and np.all(np.isfinite(y))
    g = X[:, 8:11]
    np.testing.assert_allclose(np.linalg.norm(g, axis=1), 1.0, atol=1e-5)
    assert np.all(X[:, 11:] == 0)
    # cells contain a surface point: mostly cut (labels strictly inside (0, 1))
    assert np.mean((y > 0) & (y < 1)) > 0.9 and y.min() >= 0 and y.max() < 1.5
    # a cell holding a surface point has corners on both sides in most samples (the ellipsoid's
    # first-order distance is not 1-Lipschitz, so no diagonal bound holds in general)
    assert np.mean((X[:, :8].min(1) < 0) & (X[:, :8].max(1) > 0)) > 0.9


def test_mfma_inference_matches_reference():
    p = _params(11)
    X, _ = SG.samples(3, 4096)
    got = SG.evaluate(p, X)
    ref, _ = R.forward(p, X)
    scale = max(np.abs(ref).max(), 1.0)
    d = np.abs(got - ref) / scale
    assert np.quantile(d, 0.99) <= 1e-5 and d.max() <= 2e-2, (np.quantile(d, 0.99), d.max())


def test_training_gradient_matches_reference():
    p = _params(5)
    X, y = SG.samples(9, 4096)
    g, ss = SG.grad(p, X, y)
    gr, ssr = R.grad(p, X, y)
    assert ss == pytest.approx(ssr, rel=1e-4)
    assert np.max(np.abs(g - gr)) <= 1e-3 * np.linalg.norm(gr), np.max(np.abs(g - gr)) / np.linalg.norm(gr)


def test_training_reduces_error():
    p0 = SG.init_params(2)
    Xv, yv = SG.samples(123, 20000)
    e0 = np.mean((SG.evaluate(p0, Xv) - yv) ** 2)
    p, rep = SG.train(p0, seed=4, n_train=262144, n_val=20000, batch=16384, epochs=4, lr=3e-3)
    e1 = np.mean((SG.evaluate(p, Xv) - yv) ** 2)
    assert rep["val_mse"] < 0.1 * e0 and e1 < 0.1 * e0, (e0, e1, rep)


def test_interact_surrogate_switches():
    """Never falling back (angle 180) serves eligible cut cells by the network; always falling
    back (infinite band, negative angle) reproduces the exact fractions bit for bit."""
    sc = scenes.small_mixed_3d(32, seed=2, n_each=3)
    sol = sc.solids[np.isin(sc.solids["shape"], [0, 1])]  # spheres and ellipsoids
    p = _params(8)
    with SolidCloud(sc.mesh, rho_f=sc.rho_f, precision="fp32") as c:
        c.set_solids(sol)
        c.set_velocity(sc.U)
        ft0 = c.interact(0.0, sc.dt)
        f0 = c.fields()
        n_cut = 0
        for s in range(sol.size):
            counts, _, _ = c.candidates(s)
            n_cut += int(counts[1] + counts[2])
        c.set_surrogate(p, band=(-1e30, 1e30), max_angle_deg=-1.0)
        ft1 = c.interact(0.0, sc.dt)
        f1 = c.fields()
        assert c.surrogate_count() == 0
        np.testing.assert_array_equal(f1["As"], f0["As"])
        np.testing.assert_array_equal(ft1, ft0)
        c.set_surrogate(p, band=(0.0, 1.0), max_angle_deg=180.0)
        c.interact(0.0, sc.dt)
        served = c.surrogate_count()
        assert 0 < served <= n_cut
        c.set_surrogate(None)
        c.interact(0.0, sc.dt)
        np.testing.assert_array_equal(c.fields()["As"], f0["As"])


def test_shipped_parameters_accuracy():
    """The parameter file shipped with the package (trained on the GPU on 10M cells, 10 epochs):
    error on fresh cells of the training distribution, and it beats the untrained network."""
    p = SG.load_params()
    X, y = SG.samples(2024, 100000)
    e = np.abs(SG.evaluate(p, X) - y)
    e0 = np.abs(SG.evaluate(SG.init_params(0), X) - y)
    assert e.mean() < 0.015 and np.quantile(e, 0.99) < 0.06 and e.max() < 0.3, (e.mean(), e.max())
    assert e.mean() < 0.1 * e0.mean()
