"""The adaptive 1+1D X loop on the GPU (AdaptiveX1D: the library's reference numbering,
GPU element-quadrature assembly, caller's-CSR MINRES, GPU X indicators, Doerfler marking)
against the published adaptive runs and against the oracle on the same spaces:

* every step's X / Y true-DoF counts equal the published ones, and eta^2, lambda_Y^2 match
  them to their 6 published digits (algorithm_data_plot/1d/smooth_initial.csv:336-352 and
  the p = 2, 3 rows of tests/golden/adaptive1d_anchors.json);
* every step's K and b equal the oracle's element assembly on the same spaces to 1e-12
  (relative to max |K|), and the GPU indicators of the GPU solution equal the oracle's
  indicators of that solution to 1e-10 per cell.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import doubleadapparabolicfem_amd as apf
from doubleadapparabolicfem_amd import adaptive1d as A
from oracle import adaptive1d as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "adaptive1d_anchors.json")


def published(p, mode="adaptive_standard"):
    """published records of the 1+1D smooth_initial run of degree p in the given mode"""
    return [a for a in json.load(open(GOLD)) if a["p"] == p and a["mode"] == mode]


@pytest.fixture(scope="module")
def ctx():
    return apf.Context(0)


def u0(x):
    return np.sin(np.pi * x)


@pytest.mark.parametrize("p,steps", [(1, 17), (2, 12), (3, 10)])
def test_gpu_adaptive_run_reproduces_published_history(ctx, p, steps):
    gold = published(p)[:steps]
    run = A.AdaptiveX1D(ctx, p=p, theta_x=0.4, tol=1e-10, max_it=40000)
    T = O.Tree()
    OX = O.Space(T, p, True, T.leaves())
    OX.refine(list(OX.active))
    for a in gold:
        print(f"p={p} outer={a['outer']} X={run.X.n_dofs}", flush=True)
        OY = O.Space(T, p, False, OX.active)
        assert (run.X.n_dofs, run.Y.n_dofs) == (a["x_true_dofs"], a["y_true_dofs"]), a["source"]
        assert list(run.X.active) == OX.active and list(run.Y.active) == OY.active
        # assembly parity on these spaces
        pair = A.SpaceTimePair(run.X, run.Y)
        gs = apf.GeneralSystem(ctx, pair)
        gs.assemble(u0)
        rowptr, col, vals, b = gs.download()
        Ko, bo = O.assemble(OX, OY, u0)
        Kg = sp.csr_matrix((vals, col, rowptr), shape=Ko.shape)
        scale = abs(Ko).max()
        assert abs(Kg - Ko).max() <= 1e-12 * scale
        assert np.abs(b - bo).max() <= 1e-12 * max(1.0, np.abs(bo).max())
        del gs
        rec = run.step()
        assert rec["minres_rel_residual"] <= 1e-10
        assert rec["eta_squared"] == pytest.approx(a["eta_squared"], rel=6e-6), a["source"]
        assert rec["lambda_y_squared"] == pytest.approx(a["lambda_y_squared"], rel=6e-6), a["source"]
        # indicators of the GPU solution, oracle vs GPU
        l2o, t2o = O.x_indicators(OX, OY, rec["solution"], u0)
        lxo = np.zeros(len(OX.active))
        np.add.at(lxo, OY.ancestors_in(OX), l2o)
        np.testing.assert_allclose(rec["eta2_by_x_cell"], lxo + t2o, rtol=1e-10, atol=1e-14)
        assert run.X.n_dofs == a["x_true_dofs_after"], a["source"]
        OX.refine([int(c) for c in rec["marked_x_cells"]])


@pytest.mark.parametrize("p,steps", [(1, 8), (2, 6), (3, 5)])
def test_gpu_graph_norm_matches_oracle_and_preconditions_minres(ctx, p, steps):
    """Hhat = -C + B_dt diag(A_y)^-1 B_dt^T + A_x on the adaptive meshes of the published run,
    GPU against the oracle (1e-12 relative to max |Hhat|); MINRES with Hhat as the X block
    reaches the 1e-10 true residual and the direct solve's solution"""
    run = A.AdaptiveX1D(ctx, p=p, theta_x=0.4, tol=1e-10, max_it=40000)
    T = O.Tree()
    OX = O.Space(T, p, True, T.leaves())
    OX.refine(list(OX.active))
    for k in range(steps):
        OY = O.Space(T, p, False, OX.active)
        pair = A.SpaceTimePair(run.X, run.Y)
        gs = apf.GeneralSystem(ctx, pair)
        gs.assemble(u0)
        H = gs.graph_norm()
        rowptr, col, vals = H.download()
        Hg = sp.csr_matrix((vals, col, rowptr), shape=(run.X.n_dofs, run.X.n_dofs)).toarray()
        Ko, bo = O.assemble(OX, OY, u0)
        Ho = O.graph_norm(OX, OY, Ko)
        scale = np.abs(Ho).max()
        print(f"p={p} step={k} nX={run.X.n_dofs} |Hg-Ho|/max={np.abs(Hg - Ho).max() / scale:.2e}", flush=True)
        assert np.abs(Hg - Ho).max() <= 1e-12 * scale
        assert np.abs(Hg - Hg.T).max() <= 1e-13 * scale
        x, st = gs.matrix().minres(gs.rhs(), gs.nY, hhat=H, tol=1e-10, max_it=40000, precond=apf.APF_FP64)
        assert st["true_relative_residual"] <= 1e-10
        xo = O.solve(Ko, bo)
        assert np.linalg.norm(x - xo) <= 1e-7 * np.linalg.norm(xo)
        del H, gs
        rec = run.step()
        OX.refine([int(c) for c in rec["marked_x_cells"]])


@pytest.mark.parametrize("p,steps,max_direct", [(1, 10, 64), (2, 7, 64), (3, 6, 64), (2, 6, 4)])
def test_gpu_y_estimator_on_adaptive_meshes(ctx, p, steps, max_direct):
    """The Y estimator on the reference's adaptive meshes (Y cells sliced into the time
    slabs they span, apf_general_estimator_create_sliced): the published y_estimator^2 of
    every step to its 6 digits, and the oracle's slab-group estimator on the same spaces and
    the same GPU solution to 1e-10 (totals, and per source cell to 1e-10 of the largest cell value).  max_direct = 4 gives the
    larger slab meshes coarse meshes, so the slab solve runs its multigrid."""
    gold = published(p)[:steps]
    run = A.AdaptiveX1D(ctx, p=p, theta_x=0.4, tol=1e-10, max_it=40000)
    T = O.Tree()
    OX = O.Space(T, p, True, T.leaves())
    OX.refine(list(OX.active))
    for a in gold:
        OY = O.Space(T, p, False, OX.active)
        pair = A.SpaceTimePair(run.X, run.Y)
        gs = apf.GeneralSystem(ctx, pair)
        gs.assemble(u0)
        x, st = gs.matrix().minres(gs.rhs(), gs.nY, tol=1e-10, max_it=40000, precond=apf.APF_FP64)
        ge = apf.GeneralEstimator(gs, A.slab_groups(run.Y, max_direct=max_direct))
        ge.assemble(A.smooth_initial_u0)
        est = ge.estimate(x, apf.APF_FP64, 0, per_cell=True)
        print(f"p={p} outer={a['outer']} X={run.X.n_dofs} y_est2={est['y_est2']:.6g} "
              f"(published {a['y_estimator_squared']:.6g}), slab its {est['slab_iterations']}", flush=True)
        assert est["y_est2"] == pytest.approx(a["y_estimator_squared"], rel=6e-6), a["source"]
        eo = O.estimate_y(OX, OY, x, u0)
        assert est["y_est2"] == pytest.approx(eo["y_est2"], rel=1e-10)
        for key in ("flux_by_cell", "recon_by_cell"):
            # per cell to 1e-10 of the largest cell value; reconstruction terms at round-off
            # (the slabs of a uniform Y mesh) to 1e-12 of the total
            diff = np.abs(est[key] - eo[key]).max()
            assert diff <= 1e-10 * np.abs(eo[key]).max() + 1e-12 * eo["y_est2"], (key, diff)
        del ge, gs
        rec = run.step()
        OX.refine([int(c) for c in rec["marked_x_cells"]])


def test_gpu_adaptive_loop_reports_y_estimator(ctx):
    """AdaptiveX1D(with_y_estimator=True): each step's record carries the Y estimator of the
    step's solution (the published y_estimator^2 to its 6 digits) beside the X indicators."""
    gold = published(2)[:5]
    run = A.AdaptiveX1D(ctx, p=2, theta_x=0.4, tol=1e-10, max_it=40000, with_y_estimator=True)
    for a in gold:
        rec = run.step()
        est = rec["y_estimator"]
        print(f"outer={a['outer']} y_est2={rec['y_estimator_squared']:.6g} slab its {est['slab_iterations']}",
              flush=True)
        assert rec["y_estimator_squared"] == pytest.approx(a["y_estimator_squared"], rel=6e-6), a["source"]
        assert rec["eta_squared"] == pytest.approx(a["eta_squared"], rel=6e-6), a["source"]
        assert est["flux_by_cell"].sum() + est["recon_by_cell"].sum() == pytest.approx(est["y_est2"], rel=1e-12)


FORCE = "adaptive_force_effective_rho_no_inner_refinement"


@pytest.mark.parametrize("p,steps", [(1, 28), (2, 22), (3, 20)])
def test_gpu_adaptive_run_reproduces_published_no_inner_history(ctx, p, steps):
    """The published 1+1D runs without inner Y refinement (Y = X's mesh every outer
    iteration, algorithm_data_plot/1d/smooth_initial.csv), as far as a step solves in a few
    seconds: X up to 44,562 (p = 1), 31,379 (p = 2) and 15,356 (p = 3) true DoFs.  Every
    step's X / Y true-DoF counts equal the published ones, and eta^2, lambda_Y^2 match their 6
    published digits; MINRES (Hhat-preconditioned) reaches the reference's 1e-10 acceptance."""
    gold = published(p, FORCE)[:steps]
    assert len(gold) == steps
    run = A.AdaptiveX1D(ctx, p=p, theta_x=0.4, tol=1e-10, max_it=100000)
    for a in gold:
        rec = run.step()
        print(f"p={p} outer={a['outer']} X={rec['x_true_dofs']} its={rec['minres_iterations']} "
              f"solve={rec['minres_seconds']:.2f}s", flush=True)
        assert (rec["x_true_dofs"], rec["y_true_dofs"]) == (a["x_true_dofs"], a["y_true_dofs"]), a["source"]
        assert rec["minres_rel_residual"] <= 1e-10
        assert rec["eta_squared"] == pytest.approx(a["eta_squared"], rel=6e-6), a["source"]
        assert rec["lambda_y_squared"] == pytest.approx(a["lambda_y_squared"], rel=6e-6), a["source"]
        assert rec["x_true_dofs_after"] == a["x_true_dofs_after"], a["source"]


INNER_GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "adaptive1d_inner_anchors.json")


@pytest.mark.parametrize("p,outer_steps", [(1, 26), (2, 26), (3, 27)])
def test_gpu_adaptive_driver_with_inner_loop_reproduces_published_history(ctx, p, outer_steps):
    """The whole adaptive driver on the GPU, inner Y loop included (Y estimator against its
    threshold, Doerfler marking of the Y cells with theta_y = 0.4, Y refinement with its
    grading closure), against every inner-loop record of the published adaptive_standard
    runs through outer iteration 25 / 25 / 26 at p = 1 / 2 / 3 (X up to 29,431 / 66,983 /
    67,739 true DoFs), the inner iterations that refine Y included: outer and inner indices, X
    and Y true-DoF counts before and after each Y refinement exactly, eta^2, lambda_Y^2,
    y_estimator^2 and its threshold to their 6 published digits.  MINRES starts from the
    previous solve's solution prolonged onto the refined spaces."""
    gold = [a for a in json.load(open(INNER_GOLD))
            if a["p"] == p and a["mode"] == "adaptive_standard" and a["outer"] < outer_steps]
    run = A.AdaptiveX1D(ctx, p=p, theta_x=0.4, rho=1.0, tol=1e-10, max_it=100000, inner_y_loop=True,
                        theta_y=0.4)
    k = 0
    for outer in range(outer_steps):
        rec = run.step()
        for r in rec["inner"]:
            a = gold[k]
            k += 1
            print(f"p={p} outer={outer} inner={r['inner_iteration']} X={r['x_true_dofs']} "
                  f"Y={r['y_true_dofs_before']}->{r.get('y_true_dofs_after', r['y_true_dofs'])} "
                  f"its={r['minres_iterations']}", flush=True)
            assert (a["outer"], a["inner"]) == (outer, r["inner_iteration"]), a["source"]
            assert (r["x_true_dofs"], r["y_true_dofs_before"]) == (a["x_true_dofs"], a["y_true_dofs_before"]), \
                a["source"]
            assert r.get("y_true_dofs_after", r["y_true_dofs"]) == a["y_true_dofs_after"], a["source"]
            for key in ("eta_squared", "lambda_y_squared", "y_estimator_squared", "y_estimator_threshold_squared"):
                assert r[key] == pytest.approx(a[key], rel=6e-6), (key, a["source"])
    assert k == len(gold)


@pytest.mark.parametrize("p,outer_steps", [(1, 22), (2, 21)])
def test_gpu_g_estimator_on_adaptive_meshes(ctx, p, outer_steps):
    """G^2 of every inner iteration on the adaptive meshes (degree-(p+1) element assembly on
    (X_{p+1}, G), G solve by conjugate gradients on the device, energy of lambda_G - I lambda)
    against the published adaptive_g_effective_rho runs, whose meshes follow the standard
    runs' history: G true DoFs exactly and G^2 to its 6 digits through outer iteration 21 / 20
    at p = 1 / 2, including the inner iterations that refine Y; and against the oracle's G
    (sparse direct solve) on the same GPU solution to 1e-8 on the first steps."""
    gold = [a for a in json.load(open(INNER_GOLD))
            if a["p"] == p and a["mode"] == "adaptive_g_effective_rho" and a["outer"] < outer_steps]
    run = A.AdaptiveX1D(ctx, p=p, tol=1e-10, max_it=100000, inner_y_loop=True, with_g_estimator=True)
    T = O.Tree()
    OX = O.Space(T, p, True, T.leaves())
    OX.refine(list(OX.active))
    k = 0
    for outer in range(outer_steps):
        if outer < 6:
            OY = O.Space(T, p, False, OX.active)
        rec = run.step()
        for r in rec["inner"]:
            a = gold[k]
            k += 1
            print(f"p={p} outer={outer} inner={r['inner_iteration']} G={r['g_true_dofs']} "
                  f"g2={r['g_lambda_difference_squared']:.6g} (published {a['g_lambda_difference_squared']:.6g}) "
                  f"its={r['g_solver_iterations']}", flush=True)
            assert (a["outer"], a["inner"], a["x_true_dofs"]) == (outer, r["inner_iteration"], r["x_true_dofs"])
            assert r["g_true_dofs"] == a["g_true_dofs"], a["source"]
            assert r["g_lambda_difference_squared"] == pytest.approx(a["g_lambda_difference_squared"], rel=6e-6), \
                a["source"]
            if outer < 6:
                go = O.g_estimate(OX, OY, r["solution"], u0)
                assert go["g_true_dofs"] == r["g_true_dofs"]
                assert r["g_lambda_difference_squared"] == pytest.approx(go["g_lambda_difference_squared"], rel=1e-8)
        if outer < 6:
            OX.refine([int(c) for c in rec["marked_x_cells"]])
    assert k == len(gold)
