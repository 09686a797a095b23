"""Scene generators: the ellipsoid placement keeps ellipsoids apart (checked by dense point
sampling, independent of the Perram-Wertheim test the placement uses), and the contact function
reduces to the sphere distance for spheres.  CPU only."""
import numpy as np

from sdfibm_amd import scenes
from sdfibm_amd import solids as S


def test_placed_ellipsoids_do_not_overlap():
    rng = np.random.default_rng(4)
    n = 60
    axes = rng.uniform(0.03, 0.1, (n, 3))
    q = S.random_unit_quaternions(rng, n)
    R = scenes._quat_R(q)
    pos = scenes.place_ellipsoids(rng, axes, R, (0, 0, 0), (1, 1, 1), 0.0)
    assert np.all(np.isfinite(pos))
    # dense sample of the box: no point inside two ellipsoids
    P = rng.uniform(0, 1, (400000, 3))
    count = np.zeros(P.shape[0], dtype=np.int64)
    for s in range(n):
        loc = (P - pos[s]) @ R[s]  # local coordinates (R^T (p - c))
        count += np.sum((loc / axes[s]) ** 2, axis=1) < 1.0
    assert count.max() <= 1
    # the packing is not trivially sparse
    assert np.mean(count) > 0.05


def test_contact_function_of_spheres():
    """Two spheres of radii a, b placed at distance just above / below a + b."""
    a, b = 0.1, 0.2
    axes = np.array([[a] * 3, [b] * 3])
    R = np.stack([np.eye(3)] * 2)
    M = np.einsum("nij,nj,nkj->nik", R, axes ** 2, R)
    lam = np.linspace(0, 1, 1001)[1:-1]
    for dist, apart in ((0.3001, True), (0.2999, False)):
        r = np.array([dist, 0, 0])
        F = [l * (1 - l) * r @ np.linalg.solve((1 - l) * M[0] + l * M[1], r) for l in lam]
        assert (max(F) >= 1.0) == apart
