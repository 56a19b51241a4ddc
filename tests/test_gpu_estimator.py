"""GPU parity of the equilibrated-flux estimator against the oracle (through
the C ABI), and the full GPU pipeline against the published anchors."""
import json
import os

import numpy as np
import pytest

import doubleadapparabolicfem_amd as apf
from doubleadapparabolicfem_amd import meshes
from oracle import assembly as OA, estimator as OE

from gpu_helpers import oracle_setup

pytestmark = pytest.mark.gpu
ANCHORS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "anchors.json")))


@pytest.fixture(scope="module")
def ctx():
    return apf.Context(0)


# (2, 3, 1) with the checkerboard kappa is the C4 patch population: p=3 union-jack
# interior patches of order 209 per time mode, factored by block elimination
CASES = [(1, 1, 3), (1, 2, 2), (1, 3, 1), (2, 1, 0), (2, 1, 1), (2, 2, 0), (2, 2, 1), (2, 3, 0), (2, 1, 2),
         (2, 3, 1)]


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("prec", [apf.APF_FP64, apf.APF_FP32_REFINE], ids=["fp64", "fp32ref"])
@pytest.mark.parametrize("dim,p,outer", CASES)
def test_estimator_matches_oracle(ctx, dim, p, outer, prec):
    """Same input solution (the oracle's direct solve) on both sides: every
    estimator value agrees to 1e-10 relative (the mixed path after 2 FP64
    refinement steps), per cell and in total."""
    m = meshes.uniform_mesh(dim, outer)
    coeff = meshes.smooth_initial(dim)
    if dim == 2 and outer >= 1:
        coeff = apf.Coefficients(meshes.checkerboard_kappa(m, 4, seed=3), 0.0, coeff.u0)
    print(f"estimator dim={dim} p={p} outer={outer} prec={prec}", flush=True)
    stm, od, prob = oracle_setup(m, p, coeff)
    S = OA.assemble(stm, od, prob)
    x = OE.main_solve(S)
    ref = OE.estimate(stm, od, prob, S, x)
    sp = apf.Space(m, p)
    sysm = apf.System(ctx, sp)
    sysm.assemble(coeff)
    sysm.set_solution(x)
    est = sysm.estimate(prec, 2)
    assert est["max_patch_residual"] <= 1e-10
    for key, rv in (("flux2", ref.flux2), ("lambda_y2", ref.lambda_y2), ("trace2", ref.trace2),
                    ("eta2", ref.eta2), ("y_est2", ref.y_est2)):
        assert _rel(est[key], rv) <= 1e-10, (key, est[key], rv)
    fc = est["flux_by_cell"]
    assert np.all(np.abs(fc - ref.flux_by_cell) <= 1e-10 * np.abs(ref.flux_by_cell) + 1e-14 * ref.flux2)
    assert est["div_residual_l2"] <= 1e-11
    assert est["recon2"] <= 1e-25


@pytest.mark.parametrize("dim,p,outer", [(1, 2, 2), (2, 1, 1), (2, 2, 1), (2, 3, 1)])
def test_full_pivot_retry_matches_oracle(ctx, dim, p, outer, monkeypatch):
    """The retry of patches that miss the acceptance (FP64 Gauss-Jordan with complete
    pivoting, rank threshold 1e-12, as the reference's FullPivLU retry), forced on every
    patch: the same values as the oracle to 1e-10 and every patch accepted.  Without the
    switch no patch takes the retry."""
    m = meshes.uniform_mesh(dim, outer)
    coeff = meshes.smooth_initial(dim)
    if dim == 2:
        coeff = apf.Coefficients(meshes.checkerboard_kappa(m, 4, seed=3), 0.0, coeff.u0)
    print(f"full-pivot retry dim={dim} p={p} outer={outer}", flush=True)
    stm, od, prob = oracle_setup(m, p, coeff)
    S = OA.assemble(stm, od, prob)
    x = OE.main_solve(S)
    ref = OE.estimate(stm, od, prob, S, x)
    sp = apf.Space(m, p)
    sysm = apf.System(ctx, sp)
    sysm.assemble(coeff)
    sysm.set_solution(x)
    plain = sysm.estimate(apf.APF_FP32_REFINE, 2)
    assert plain["fullpiv_patches"] == 0
    monkeypatch.setenv("APF_PATCH_FULLPIV", "1")
    est = sysm.estimate(apf.APF_FP32_REFINE, 2)
    assert est["fullpiv_patches"] == len(sp.patch_sizes()[0])
    assert est["max_patch_residual"] <= 1e-10
    for key, rv in (("flux2", ref.flux2), ("lambda_y2", ref.lambda_y2), ("trace2", ref.trace2),
                    ("eta2", ref.eta2), ("y_est2", ref.y_est2)):
        assert _rel(est[key], rv) <= 1e-10, (key, est[key], rv)
    fc = est["flux_by_cell"]
    assert np.all(np.abs(fc - ref.flux_by_cell) <= 1e-10 * np.abs(ref.flux_by_cell) + 1e-14 * ref.flux2)
    assert est["div_residual_l2"] <= 1e-11


@pytest.mark.parametrize("rec", [a for a in ANCHORS if (a["dim"], a["p"], a["outer"]) in
                                 {(1, 1, 3), (1, 2, 3), (1, 3, 2), (2, 1, 2), (2, 2, 2), (2, 3, 1)}],
                         ids=lambda a: f"{a['dim']}d-p{a['p']}-o{a['outer']}")
def test_gpu_pipeline_reproduces_published_anchor(ctx, rec):
    """Assembly -> MINRES (1e-10) -> estimator, all on the device."""
    m = meshes.uniform_mesh(rec["dim"], rec["outer"])
    sp = apf.Space(m, rec["p"])
    assert (sp.nX, sp.nY) == (rec["x_true_dofs"], rec["y_true_dofs_after"])
    sysm = apf.System(ctx, sp)
    sysm.assemble(meshes.smooth_initial(rec["dim"]))
    sysm.minres(tol=1e-10, max_it=50000, check_interval=100)
    est = sysm.estimate(apf.APF_FP32_REFINE, 2)
    for key, col in (("lambda_y2", "lambda_y_squared"), ("eta2", "eta_squared"), ("flux2", "y_flux_squared"),
                     ("y_est2", "y_estimator_squared")):
        assert _rel(est[key], rec[col]) <= 6e-6, (key, est[key], rec[col])


def test_gpu_pipeline_reproduces_published_anchor_at_full_c2_size(ctx):
    """The C2 workload itself: 2+1D p=2, union-jack 64x64 x 32 intervals
    (2,596,769 unknowns), assembly -> multigrid MINRES (1e-10) -> mixed estimator,
    with the slab reconstruction taken from the solution and solved, against the
    published record algorithm_data_plot/2d/smooth_initial.csv:740."""
    rec = [a for a in ANCHORS if (a["dim"], a["p"], a["outer"]) == (2, 2, 5)][0]
    print("full C2 pipeline", flush=True)
    m = meshes.uniform_mesh(2, 5)
    sp = apf.Space(m, 2, coarse=meshes.uniform_hierarchy(2, 5))
    assert (sp.nX, sp.nY) == (rec["x_true_dofs"], rec["y_true_dofs_after"])
    sysm = apf.System(ctx, sp)
    sysm.assemble(meshes.smooth_initial(2))
    _, st = sysm.minres(tol=1e-10, max_it=2000, check_interval=10, download=False)
    assert st["true_relative_residual"] <= 1e-10
    for slab in (apf.APF_SLAB_FROM_SOLUTION, apf.APF_SLAB_SOLVE):
        est = sysm.estimate(apf.APF_FP32_REFINE, 2, per_cell=False, slab=slab)
        assert est["max_patch_residual"] <= 1e-10
        for key, col in (("lambda_y2", "lambda_y_squared"), ("eta2", "eta_squared"), ("flux2", "y_flux_squared"),
                         ("y_est2", "y_estimator_squared")):
            assert _rel(est[key], rec[col]) <= 6e-6, (slab, key, est[key], rec[col])
    # the slab solve starts from lambda, which already meets the first block row to the
    # MINRES tolerance: its reconstruction term stays at the level of that residual
    assert est["slab_stop"] in (apf.APF_SLAB_STOP_TOL, apf.APF_SLAB_STOP_FLOOR)
    assert est["recon2"] <= 1e-12 * est["lambda_y2"]


def test_non_finite_patch_input_fails_the_acceptance(ctx):
    """A NaN in one solution entry reaches the right-hand sides of the patches
    around its cell: those patches fail the 1e-10 acceptance (the reference's
    finiteness checks, dense_local_error_blocks_2d.hpp:931-951) and the estimate
    returns APF_RESIDUAL_TOO_LARGE instead of NaN indicators."""
    print("planted NaN in the estimator input", flush=True)
    m = meshes.uniform_mesh(2, 1)
    coeff = meshes.smooth_initial(2)
    stm, od, prob = oracle_setup(m, 2, coeff)
    x = OE.main_solve(OA.assemble(stm, od, prob))
    sp = apf.Space(m, 2)
    sysm = apf.System(ctx, sp)
    sysm.assemble(coeff)
    for prec in (apf.APF_FP64, apf.APF_FP32_REFINE):
        sysm.set_solution(x)
        sysm.estimate(prec, 2)  # finite input: accepted
        bad = x.copy()
        bad[od.nY // 2] = np.nan
        sysm.set_solution(bad)
        with pytest.raises(apf.APFError) as e:
            sysm.estimate(prec, 2)
        assert e.value.code == apf.APF_RESIDUAL_TOO_LARGE, e.value


@pytest.mark.parametrize("dim,p,outer", [(1, 1, 3), (1, 2, 2), (2, 2, 1), (2, 1, 2)])
def test_slab_reconstruction_solve_matches_oracle(ctx, dim, p, outer):
    """lambda perturbed by 10%: the slab solve recovers the oracle's per-slab
    direct solves (lambda_tilde depends on u only), so the reconstruction term
    ||lambda - lambda_tilde||_Y^2 and the flux terms agree per cell."""
    m = meshes.uniform_mesh(dim, outer)
    coeff = meshes.smooth_initial(dim)
    print(f"slab solve dim={dim} p={p} outer={outer}", flush=True)
    stm, od, prob = oracle_setup(m, p, coeff)
    S = OA.assemble(stm, od, prob)
    x = OE.main_solve(S)
    rng = np.random.default_rng(7)
    x[:od.nY] *= 1.0 + 0.1 * rng.standard_normal(od.nY)
    ref = OE.estimate(stm, od, prob, S, x)
    sp = apf.Space(m, p, coarse=meshes.uniform_hierarchy(dim, outer))
    sysm = apf.System(ctx, sp)
    sysm.assemble(coeff)
    sysm.set_solution(x)
    est = sysm.estimate(apf.APF_FP64, 0, slab=apf.APF_SLAB_SOLVE)
    assert est["slab_relative_residual"] <= 1e-12, est["slab_relative_residual"]
    assert est["slab_stop"] in (apf.APF_SLAB_STOP_TOL, apf.APF_SLAB_STOP_FLOOR)
    assert ref.recon2 > 1e-6 * ref.lambda_y2
    for key, rv in (("recon2", ref.recon2), ("flux2", ref.flux2), ("y_est2", ref.y_est2)):
        assert _rel(est[key], rv) <= 1e-10, (key, est[key], rv)
    rc = est["recon_by_cell"]
    assert np.all(np.abs(rc - ref.recon_by_cell) <= 1e-9 * np.abs(ref.recon_by_cell) + 1e-13 * ref.recon2)


@pytest.mark.parametrize("prec", [apf.APF_FP64, apf.APF_FP32_REFINE], ids=["fp64", "fp32ref"])
@pytest.mark.parametrize("dim,p,outer", [(1, 2, 3), (2, 2, 2), (2, 3, 1)])
def test_repeated_estimates_are_identical(ctx, dim, p, outer, prec):
    """A system's first estimate in a precision also times its tile candidates
    (column blockings of the solve, slab groupings of the right-hand sides);
    later estimates run the chosen tiles only.  All calls give the same per-cell
    values bit for bit: the tiling changes no sum order."""
    print(f"estimate overlap dim={dim} p={p} outer={outer} prec={prec}", flush=True)
    m = meshes.uniform_mesh(dim, outer)
    sp = apf.Space(m, p, coarse=meshes.uniform_hierarchy(dim, outer))
    sysm = apf.System(ctx, sp)
    sysm.assemble(meshes.smooth_initial(dim))
    sysm.minres(tol=1e-10, max_it=600, check_interval=10)
    runs = [sysm.estimate(prec, 2) for _ in range(3)]
    for r in runs[1:]:
        for key in ("flux_by_cell", "resid_by_cell", "recon_by_cell"):
            assert np.array_equal(r[key], runs[0][key]), key
        for key in ("eta2", "flux2", "lambda_y2"):
            assert r[key] == runs[0][key], key


def test_estimate_snapshot_in_the_reference_format(ctx, tmp_path):
    """The per-cell estimator components of a GPU estimate, written in the
    reference's cellwise-components format and read back, sum to the
    estimate's totals."""
    from doubleadapparabolicfem_amd import snapshot

    print("estimate snapshot", flush=True)
    m = meshes.uniform_mesh(2, 1)
    sp = apf.Space(m, 2, coarse=meshes.uniform_hierarchy(2, 1))
    sysm = apf.System(ctx, sp)
    sysm.assemble(meshes.smooth_initial(2))
    sysm.minres(tol=1e-10, max_it=600, check_interval=10)
    est = sysm.estimate(apf.APF_FP32_REFINE, 2)
    path = tmp_path / "cellwise_error_components.bin"
    snapshot.write_cellwise_error_components_binary(est, path)
    with open(path, "rb") as f:
        h = np.fromfile(f, dtype=np.int32, count=3)
        n = int(h[1])
        ids = np.fromfile(f, dtype=np.int32, count=n)
        flux, recon, resid, total = (np.fromfile(f, dtype=np.float64, count=n) for _ in range(4))
    assert h[0] == 1 and h[2] == 4 and n == sp.n_space_cells * sp.n_intervals
    assert sorted(ids.tolist()) == list(range(n))
    assert _rel(flux.sum(), est["flux2"]) < 1e-12
    assert _rel(recon.sum(), est["recon2"]) < 1e-12
    assert _rel(total.sum(), est["y_est2"]) < 1e-12
    assert np.all(np.diff(total) <= 0)


@pytest.mark.parametrize("prec", [apf.APF_FP64, apf.APF_FP32_REFINE], ids=["fp64", "fp32ref"])
def test_overlapped_factor_and_slab_solve_match_sequential(ctx, prec, monkeypatch):
    """The factor phase on a second stream beside the slab reconstruction (forced on for
    both precisions) gives bit-identical estimates to the phases in turn."""
    m = meshes.uniform_mesh(2, 3)
    sp = apf.Space(m, 2, coarse=meshes.uniform_hierarchy(2, 3))
    sysm = apf.System(ctx, sp)
    sysm.assemble(meshes.smooth_initial(2))
    sysm.minres(tol=1e-10, max_it=2000, check_interval=10)
    sysm.estimate(prec, 2, slab=apf.APF_SLAB_SOLVE)  # first call: tiles and precisions chosen
    monkeypatch.setenv("APF_EST_OVERLAP", "0")
    seq = sysm.estimate(prec, 2, slab=apf.APF_SLAB_SOLVE)
    monkeypatch.setenv("APF_EST_OVERLAP", "1")
    ovl = sysm.estimate(prec, 2, slab=apf.APF_SLAB_SOLVE)
    for key in ("flux2", "recon2", "eta2", "y_est2", "max_patch_residual"):
        assert ovl[key] == seq[key], key
    for key in ("flux_by_cell", "resid_by_cell", "recon_by_cell"):
        assert np.array_equal(ovl[key], seq[key]), key
    assert ovl["seconds_total"] > 0
