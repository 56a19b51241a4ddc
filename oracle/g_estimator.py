"""G-estimator on tensor meshes (oracle; test infrastructure only).

Restates compute_g_estimator_for_accepted_inner_iteration_
(algorithm/adaptive_driver.hpp:2680-2905):

1. G is the Y space one degree higher in space and time, discontinuous in time
   (FiniteElementTraits<p+1, p+1> with SpaceOnlyPolicy, adaptive_driver.hpp:2697-2710;
   the same policy as Y, algorithm/runners/detail/runner_case_execution.hpp:47-50);
2. A_G lambda_G = f_G - B(X, G)^T u_delta with A_G the Y-block bilinear form on G
   and B = B_dt + B_A the mixed form between X (degree p) and G
   (adaptive_driver.hpp:2732-2800), solved directly here;
3. G^2 = (lambda_G - I lambda_delta)^T A_G (lambda_G - I lambda_delta), where I is
   nodal interpolation into G (prolong_true_coefficients_nodal, :2832-2836,
   quadratic form at :2860-2866), clamped at 0 after the negative check (:2869-2883).

B(X_p, G)^T u_delta equals B(X_{p+1}, G)^T (I_X u_delta): nodal interpolation
reproduces the degree-p function exactly, and the degree-(p+1) cell rule
integrates the polynomial integrand exactly (M piecewise constant).  So the
right-hand side is formed from the degree-(p+1) saddle blocks.

Pinned by the published initial-mesh G values of the reference's adaptive runs
(algorithm_data_plot/{1d,2d}/smooth_initial.csv), by the degree-p invariant
(G at the same degree vanishes: the first block row of the saddle system), and
by exact interpolation of polynomials (tests/test_oracle_g_estimator.py).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse.linalg as spl

from .assembly import assemble, spatial_geometry
from .basis import segment_nodes, spacetime_lagrange, triangle_nodes
from .estimator import main_solve
from .mesh import SpaceTimeDofs, SpaceTimeMesh, spacetime_dofs


def _ref_nodes(dim, q):
    """Reference space x time nodes of the degree-q element, local order
    i_space * (q+1) + i_time (the order of spacetime_lagrange)."""
    sn = np.atleast_2d(triangle_nodes(q)) if dim == 2 else np.asarray(segment_nodes(q), dtype=float).reshape(-1, 1)
    tn = np.asarray(segment_nodes(q), dtype=float)
    Ns, Nt = len(sn), len(tn)
    pts = np.zeros((Ns * Nt, dim + 1))
    pts[:, :dim] = np.repeat(sn, Nt, axis=0)
    pts[:, dim] = np.tile(tn, Ns)
    return pts


def prolong_nodal(stm: SpaceTimeMesh, dp: SpaceTimeDofs, dq: SpaceTimeDofs, vec, which: str):
    """Nodal interpolation of a degree-p X ("x") or Y ("y") coefficient vector
    into the degree-q space: the degree-p function evaluated at the degree-q
    nodes, cell by cell."""
    mesh = stm.space
    d = mesh.dim
    pts = _ref_nodes(d, dq.p)
    vals, _ = spacetime_lagrange(d, dp.p, pts)        # [nodes_q, basis_p]
    mp = dp.x_map if which == "x" else dp.y_map
    mq = dq.x_map if which == "x" else dq.y_map
    out = np.zeros(dq.nX if which == "x" else dq.nY)
    loc = np.where(mp >= 0, vec[np.maximum(mp, 0)], 0.0)  # [cells, basis_p]
    nodal = loc @ vals.T                               # [cells, nodes_q]
    ok = mq >= 0
    out[mq[ok]] = nodal[ok]
    return out


def g_estimator(stm: SpaceTimeMesh, p: int, prob, x=None):
    """Returns (G^2 raw, lambda_G, I lambda_delta, rhs) for the degree-p solution x
    (solved here when None)."""
    dp = spacetime_dofs(stm, p)
    if x is None:
        x = main_solve(assemble(stm, dp, prob))
    lam, u = x[:dp.nY], x[dp.nY:]
    dq = spacetime_dofs(stm, p + 1)
    Sq = assemble(stm, dq, prob)
    Iu = prolong_nodal(stm, dp, dq, u, "x")
    Il = prolong_nodal(stm, dp, dq, lam, "y")
    rhs = Sq.f - Sq.B.T @ Iu
    lG = spl.spsolve(Sq.A.tocsc(), rhs)
    diff = lG - Il
    g2 = float(diff @ (Sq.A @ diff))
    return g2, lG, Il, rhs


def same_degree_g(stm: SpaceTimeMesh, p: int, prob, x):
    """The G construction with G = Y (no degree raise): lambda_G then solves the
    first block row A lambda = f - B^T u of the saddle system, which the exact
    discrete solution satisfies, so G^2 vanishes up to round-off."""
    dp = spacetime_dofs(stm, p)
    S = assemble(stm, dp, prob)
    lam, u = x[:dp.nY], x[dp.nY:]
    lG = spl.spsolve(S.A.tocsc(), S.f - S.B.T @ u)
    diff = lG - lam
    return float(diff @ (S.A @ diff)), float(lam @ (S.A @ lam))
