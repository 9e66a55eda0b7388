"""Scene generators on the CPU: the stress scene's triangle edges point in
directions uniform on the unit sphere.  For a uniform direction every
component has E[c^2] = 1/3 and E[c^4] = 1/5 (z is uniform on [-1, 1]);
directions normalised from cube samples lean toward the diagonals and give
E[c^4] = 0.180, which the fourth-moment bound rejects."""
import numpy as np

from aurora_amd import scenes


def test_stress_edge_directions_are_uniform_on_the_sphere():
    arr = scenes.stress_scene(30000, 30000, seed=1)
    q = arr.quads.astype(np.float64)
    for cols in (slice(3, 6), slice(6, 9)):
        e = q[:, cols]
        d = e / np.linalg.norm(e, axis=1, keepdims=True)
        # standard errors at n = 30000: 0.0033 (mean), 0.0017 (c^2), 0.0015 (c^4)
        assert np.all(np.abs(d.mean(0)) < 0.015)
        assert np.all(np.abs((d ** 2).mean(0) - 1.0 / 3.0) < 0.008)
        assert np.all(np.abs((d ** 4).mean(0) - 0.2) < 0.007)


def test_stress_edge_lengths_and_radii_follow_their_ranges():
    arr = scenes.stress_scene(30000, 30000, seed=2)
    q = arr.quads.astype(np.float64)
    for cols in (slice(3, 6), slice(6, 9)):
        ln = np.linalg.norm(q[:, cols], axis=1)
        assert ln.min() > 0.001 * (1 - 1e-6) and ln.max() < 0.05 * (1 + 1e-6)
    r = arr.spheres[:, 3].astype(np.float64)
    assert r.min() >= 1e-3 * (1 - 1e-6) and r.max() <= 1e-1 * (1 + 1e-6)
    # log-uniform radii: log10 r is uniform on [-3, -1], mean -2
    assert abs(np.log10(r).mean() + 2.0) < 0.02
    assert np.all(arr.spheres[:, :3] >= 0) and np.all(arr.spheres[:, :3] <= 1)
