"""Oracle G-estimator pinned by the reference's published initial-mesh values
(algorithm_data_plot/{1d,2d}/smooth_initial.csv, adaptive runs, 6 significant
digits), by the degree-p invariant and by exact interpolation (no GPU)."""
import json
import os

import numpy as np
import pytest

from oracle import cases, g_estimator as G
from oracle import mesh as M
from oracle.assembly import assemble, spatial_geometry
from oracle.basis import spacetime_lagrange
from oracle.estimator import main_solve

G_ANCHORS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "g_anchors.json")))


@pytest.mark.parametrize("rec", G_ANCHORS, ids=lambda a: f"{a['dim']}d-p{a['p']}")
def test_published_g_anchor(rec):
    stm = cases.uniform_mesh(rec["dim"], rec["outer"])
    g2, lG, _, _ = G.g_estimator(stm, rec["p"], cases.smooth_initial(stm))
    assert len(lG) == rec["g_true_dofs"]
    assert abs(g2 - rec["g_lambda_difference_squared"]) <= 6e-6 * rec["g_lambda_difference_squared"]


@pytest.mark.parametrize("dim,p,outer", [(1, 2, 2), (2, 1, 1), (2, 2, 1)])
def test_same_degree_g_vanishes(dim, p, outer):
    """With G = Y the solve reproduces the first block row of the saddle system,
    which the discrete solution satisfies: the difference norm is round-off."""
    stm = cases.uniform_mesh(dim, outer)
    prob = cases.smooth_initial(stm)
    d = M.spacetime_dofs(stm, p)
    x = main_solve(assemble(stm, d, prob))
    g2, lam2 = G.same_degree_g(stm, p, prob, x)
    assert lam2 > 1e-4
    assert abs(g2) <= 1e-24


@pytest.mark.parametrize("which", ["x", "y"])
@pytest.mark.parametrize("dim,p,outer", [(1, 1, 1), (1, 3, 1), (2, 1, 1), (2, 2, 1)])
def test_prolongation_is_the_same_function(dim, p, outer, which):
    """Nodal interpolation into degree p+1 reproduces a degree-p function: both
    coefficient vectors evaluate to the same values at random points of every
    space-time cell."""
    stm = cases.uniform_mesh(dim, outer)
    dp, dq = M.spacetime_dofs(stm, p), M.spacetime_dofs(stm, p + 1)
    rng = np.random.default_rng(11)
    v = rng.standard_normal(dp.nX if which == "x" else dp.nY)
    w = G.prolong_nodal(stm, dp, dq, v, which)
    pts = rng.random((7, dim + 1))
    if dim == 2:
        pts[:, :2] *= 0.5  # inside the reference triangle
    vp, _ = spacetime_lagrange(dim, p, pts)
    vq, _ = spacetime_lagrange(dim, p + 1, pts)
    mp = dp.x_map if which == "x" else dp.y_map
    mq = dq.x_map if which == "x" else dq.y_map
    fp = np.where(mp >= 0, v[np.maximum(mp, 0)], 0.0) @ vp.T
    fq = np.where(mq >= 0, w[np.maximum(mq, 0)], 0.0) @ vq.T
    assert np.max(np.abs(fp - fq)) <= 1e-12 * max(1.0, np.max(np.abs(fp)))


def test_planted_sign_mistake_fails_anchor(monkeypatch):
    """Plant a plausible mistake in the oracle -- add B^T I u instead of
    subtracting it -- and check that the published anchor rejects it."""
    rec = [a for a in G_ANCHORS if (a["dim"], a["p"]) == (2, 2)][0]
    stm = cases.uniform_mesh(2, 0)
    prob = cases.smooth_initial(stm)
    orig = G.prolong_nodal

    def flipped(stm_, dp, dq, vec, which):
        out = orig(stm_, dp, dq, vec, which)
        return -out if which == "x" else out

    monkeypatch.setattr(G, "prolong_nodal", flipped)
    g2, _, _, _ = G.g_estimator(stm, 2, prob)
    assert abs(g2 - rec["g_lambda_difference_squared"]) > 1e-2 * rec["g_lambda_difference_squared"]
