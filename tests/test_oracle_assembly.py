"""Oracle element matrices pinned by Kronecker closed forms; DOF counts pinned by
the published X/Y true-DOF columns (no GPU)."""
import json
import os

import numpy as np
import pytest

from oracle import assembly, basis, cases, mesh as M, quadrature

ANCHORS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "anchors.json")))


def _seg_mats(p, h):
    x, w = quadrature.gauss_legendre_01(p + 3)
    v, d = basis.segment_lagrange(p, x)
    Mt = h * np.einsum("q,qi,qj->ij", w, v, v)
    Dt = np.einsum("q,qi,qj->ij", w, d, v)          # (d/dt psi_i, phi_j)
    Kt = np.einsum("q,qi,qj->ij", w, d, d) / h
    return Mt, Dt, Kt


def _tri_mats(p, V):
    qp, qw = quadrature.duffy_triangle(2 * p + 4)   # higher than the oracle's rule
    v, g = basis.triangle_lagrange(p, qp)
    J = np.stack([V[1] - V[0], V[2] - V[0]], axis=1)
    det = abs(np.linalg.det(J))
    gx = np.einsum("ab,qib->qia", np.linalg.inv(J).T, g)
    Ms = det * np.einsum("q,qi,qj->ij", qw, v, v)
    Ks = det * np.einsum("q,qia,qja->ij", qw, gx, gx)
    return Ms, Ks


@pytest.mark.parametrize("p", [1, 2, 3])
def test_prism_local_kronecker(p):
    V = np.array([[0.1, 0.2], [0.7, 0.25], [0.3, 0.9]])
    sm = M.SpatialMesh(2, V, np.array([[0, 1, 2]]))
    stm = M.SpaceTimeMesh(sm, np.array([0.3, 0.55]))
    prob = assembly.Problem(assembly.identity_M(sm), 0.0, lambda x: x[:, 0] * 0)
    Ax, B, f, Bdt = assembly.local_blocks(stm, p, prob)
    Ms, Ks = _tri_mats(p, V)
    Mt, Dt, _ = _seg_mats(p, 0.25)
    assert np.allclose(Ax[0], np.kron(Ks, Mt), rtol=0, atol=1e-13 * np.abs(Ax).max())
    assert np.allclose(B[0], np.kron(Ms, Dt) + np.kron(Ks, Mt), rtol=0, atol=1e-13 * np.abs(B).max())
    assert np.allclose(Bdt[0], np.kron(Ms, Dt), rtol=0, atol=1e-13 * np.abs(B).max())


@pytest.mark.parametrize("p", [1, 2, 3])
def test_rect_local_kronecker(p):
    sm = M.interval_mesh(1, 0.2, 0.45)
    stm = M.SpaceTimeMesh(sm, np.array([0.1, 0.4]))
    prob = assembly.Problem(np.full((1, 1, 1), 2.5), 1.5, lambda x: x[:, 0] * 0)
    Ax, B, f, _ = assembly.local_blocks(stm, p, prob)
    Mx, _, Kx = _seg_mats(p, 0.25)
    Mt, Dt, _ = _seg_mats(p, 0.3)
    assert np.allclose(Ax[0], 2.5 * np.kron(Kx, Mt), atol=1e-13)
    assert np.allclose(B[0], np.kron(Mx, Dt) + 2.5 * np.kron(Kx, Mt), atol=1e-13)
    xq, wq = quadrature.gauss_legendre_01(p + 2)
    vx, _ = basis.segment_lagrange(p, xq)
    assert np.allclose(f[0], 1.5 * np.kron(0.25 * wq @ vx, 0.3 * wq @ vx), atol=1e-14)


@pytest.mark.parametrize("rec", [a for a in ANCHORS if a["dim"] == 1 or a["outer"] <= 4],
                         ids=lambda a: f"{a['dim']}d-p{a['p']}-o{a['outer']}")
def test_dof_counts_match_published(rec):
    stm = cases.uniform_mesh(rec["dim"], rec["outer"])
    sd = M.spatial_dofs(stm.space, rec["p"])
    m = stm.n_intervals
    p = rec["p"]
    assert (m * p + 1) * sd.n_free == rec["x_true_dofs"]
    assert m * (p + 1) * sd.n_free == rec["y_true_dofs_after"]


def test_global_blocks_symmetry_and_sign():
    stm = cases.uniform_mesh(2, 1)
    d = M.spacetime_dofs(stm, 2)
    S = assembly.assemble(stm, d, cases.smooth_initial(stm))
    A = S.A.toarray()
    assert np.allclose(A, A.T, atol=1e-15)
    assert np.linalg.eigvalsh(A).min() > 0
    C = S.C.toarray()
    assert np.allclose(C, C.T) and np.linalg.eigvalsh(C).max() <= 1e-14
    # B applied to a function constant in time is the spatial stiffness part only:
    # sum over rows of Bdt in each X column block vanishes for time-constant Y fields
    ones_time = np.ones(d.nY)
    assert np.allclose((S.Bdt @ ones_time)[d.sd.n_free:-d.sd.n_free], 0.0, atol=1e-13)
