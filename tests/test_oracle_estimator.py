"""Oracle estimator pinned by the reference's published uniform-run totals
(algorithm_data_plot/{1d,2d}/smooth_initial.csv, 6 significant digits) and by
the equilibration invariant div sigma = Pi(ell - d_t u) (no GPU)."""
import json
import os

import numpy as np
import pytest

from oracle import cases, estimator, mesh as M

ANCHORS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "anchors.json")))
FAST = {(1, 1, o) for o in range(4)} | {(1, 2, o) for o in range(3)} | {(1, 3, o) for o in range(2)} \
    | {(1, 4, 0)} | {(2, 1, 0), (2, 1, 1), (2, 2, 0), (2, 3, 0)}
SLOW = {(2, 1, 2), (2, 2, 1)}


def _cases(keys):
    return [a for a in ANCHORS if (a["dim"], a["p"], a["outer"]) in keys]


def _check(rec):
    stm = cases.uniform_mesh(rec["dim"], rec["outer"])
    d = M.spacetime_dofs(stm, rec["p"])
    r = estimator.estimate(stm, d, cases.smooth_initial(stm))
    rel = lambda a, b: abs(a - b) / abs(b)
    # published with 6 significant digits
    assert rel(r.lambda_y2, rec["lambda_y_squared"]) < 6e-6
    assert rel(r.eta2, rec["eta_squared"]) < 6e-6
    assert rel(r.flux2, rec["y_flux_squared"]) < 6e-6
    assert rel(r.y_est2, rec["y_estimator_squared"]) < 6e-6
    # tensor meshes: slab reconstruction reproduces lambda (published ~1e-31)
    assert r.recon2 < 1e-20
    # equilibration: divergence residual at round-off (published 1e-16 .. 1e-13)
    assert r.div_resid_l2 < 1e-11
    assert np.all(r.flux_by_cell >= 0)


@pytest.mark.parametrize("rec", _cases(FAST), ids=lambda a: f"{a['dim']}d-p{a['p']}-o{a['outer']}")
def test_published_anchor(rec):
    _check(rec)


@pytest.mark.slow
@pytest.mark.parametrize("rec", _cases(SLOW), ids=lambda a: f"{a['dim']}d-p{a['p']}-o{a['outer']}")
def test_published_anchor_slow(rec):
    _check(rec)


def test_planted_orientation_mistake_fails_anchor(monkeypatch):
    """Plant a plausible mistake in the oracle -- ignore the edge orientation
    signs of the RT edge moments -- and check the published anchor rejects it."""
    rec = _cases({(2, 2, 0)})[0]
    stm = cases.uniform_mesh(2, 0)
    d = M.spacetime_dofs(stm, 2)
    orig = estimator.build_patches_spatial

    def unsigned(mesh, p):
        pts = orig(mesh, p)
        for pt in pts:
            pt.flux_map = [[(li, pi, 1.0) for (li, pi, _) in ent] for ent in pt.flux_map]
        return pts

    monkeypatch.setattr(estimator, "build_patches_spatial", unsigned)
    try:
        r = estimator.estimate(stm, d, cases.smooth_initial(stm))
        bad = abs(r.flux2 - rec["y_flux_squared"]) / rec["y_flux_squared"]
    except RuntimeError:
        bad = np.inf
    assert bad > 1e-3
