"""Fan-patch meshes and their oracle fixtures (CPU)."""
import os
import sys

import numpy as np

from doubleadapparabolicfem_amd import meshes

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_fan_fixtures  # noqa: E402


def test_fan_mesh_is_a_conforming_ccw_triangulation():
    m = meshes.fan_mesh(4, [0.0, 1.0], seed=3)
    v = m.vertices[m.cells]
    area = 0.5 * ((v[:, 1, 0] - v[:, 0, 0]) * (v[:, 2, 1] - v[:, 0, 1])
                  - (v[:, 2, 0] - v[:, 0, 0]) * (v[:, 1, 1] - v[:, 0, 1]))
    assert area.min() > 0 and abs(area.sum() - 1.0) < 1e-13
    # conforming: every edge has one or two cells, two exactly off the boundary of the square
    edges = {}
    for c in m.cells:
        for i, j in ((1, 2), (2, 0), (0, 1)):
            key = (min(c[i], c[j]), max(c[i], c[j]))
            edges[key] = edges.get(key, 0) + 1
    for (a, b), cnt in edges.items():
        on_bd = any(np.all(np.isclose(m.vertices[[a, b], k], t)) for k in (0, 1) for t in (0.0, 1.0))
        assert cnt == (1 if on_bd else 2)
    # centre valences are the sums of four side splits in 1..16
    val = np.bincount(m.cells.ravel())
    centres = val[-16:]
    assert centres.min() >= 4 and centres.max() <= 64 and centres.max() > 40


def test_graded_times():
    t = meshes.graded_times(2, 10)
    assert t[0] == 0.0 and t[1] == 2.0 ** -10 and t[-1] == 1.0
    assert np.all(np.diff(t) > 0) and t.size == 13


def test_fixture_matches_oracle():
    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "fan_p1.npz"))
    out = make_fan_fixtures.compute("fan_p1")
    assert np.array_equal(fx["params"], out["params"])
    assert np.allclose(out["x"], fx["x"], rtol=1e-12, atol=1e-14 * np.abs(fx["x"]).max())
    assert np.allclose(out["flux_by_cell"], fx["flux_by_cell"], rtol=1e-10, atol=1e-16)
