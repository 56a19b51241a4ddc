"""GPU parity of the main-system path (assembly of K, f, g and MINRES)
against the oracle, through the C ABI."""
import numpy as np
import pytest

import doubleadapparabolicfem_amd as apf
from doubleadapparabolicfem_amd import meshes
from oracle import assembly as OA, estimator as OE

from gpu_helpers import oracle_setup

pytestmark = pytest.mark.gpu

CASES = [(1, 1, 3), (1, 2, 2), (1, 3, 1), (2, 1, 1), (2, 2, 1), (2, 3, 0), (2, 2, 2)]


@pytest.fixture(scope="module")
def ctx():
    return apf.Context(0)


def _coeff(dim, kind):
    if kind == "smooth":
        return meshes.smooth_initial(dim)
    return None


@pytest.mark.parametrize("mode", [apf.APF_ASSEMBLE_KRON, apf.APF_ASSEMBLE_SCATTER], ids=["kron", "scatter"])
@pytest.mark.parametrize("dim,p,outer", CASES)
def test_assembly_matches_oracle(ctx, dim, p, outer, mode):
    m = meshes.uniform_mesh(dim, outer)
    coeff = meshes.smooth_initial(dim)
    if dim == 2 and outer >= 1:
        coeff = apf.Coefficients(meshes.checkerboard_kappa(m, 4, seed=7), 0.3, coeff.u0)
    print(f"assembly dim={dim} p={p} outer={outer} mode={mode}", flush=True)
    sp = apf.Space(m, p)
    sysm = apf.System(ctx, sp)
    sysm.assemble(coeff, mode)
    vals, b = sysm.download()
    stm, od, prob = oracle_setup(m, p, coeff)
    S = OA.assemble(stm, od, prob)
    K = S.K()
    rowptr, col = sp.pattern()
    assert np.array_equal(K.indices, col)
    scale = np.abs(K.data).max()
    err = np.abs(vals - K.data)
    # 1e-12 relative per entry; structural zeros / cancellation measured against the matrix scale
    assert np.all(err <= 1e-12 * np.abs(K.data) + 1e-14 * scale), err.max()
    rhs = S.rhs()
    assert np.all(np.abs(b - rhs) <= 1e-12 * np.abs(rhs) + 1e-15 * np.abs(rhs).max())


@pytest.mark.parametrize("prec", [apf.APF_FP64, apf.APF_FP32_REFINE], ids=["pc64", "pc32"])
@pytest.mark.parametrize("dim,p,outer", [(1, 1, 3), (1, 2, 3), (2, 1, 2), (2, 2, 1), (2, 3, 1)])
def test_minres_matches_direct_solve(ctx, dim, p, outer, prec):
    m = meshes.uniform_mesh(dim, outer)
    coeff = meshes.smooth_initial(dim)
    print(f"minres dim={dim} p={p} outer={outer} prec={prec}", flush=True)
    sp = apf.Space(m, p)
    sysm = apf.System(ctx, sp)
    sysm.assemble(coeff)
    x, st = sysm.minres(tol=1e-10, max_it=20000, check_interval=50, precond=prec)
    assert st["converged"] and st["true_relative_residual"] <= 1e-10
    stm, od, prob = oracle_setup(m, p, coeff)
    S = OA.assemble(stm, od, prob)
    xd = OE.main_solve(S)
    K = S.K()
    b = S.rhs()
    assert np.linalg.norm(K @ x - b) <= 1.01e-10 * np.linalg.norm(b)
    assert np.linalg.norm(x - xd) <= 1e-7 * np.linalg.norm(xd)


def test_minres_failure_is_reported(ctx):
    m = meshes.uniform_mesh(2, 2)
    sp = apf.Space(m, 2)
    sysm = apf.System(ctx, sp)
    sysm.assemble(meshes.smooth_initial(2))
    with pytest.raises(apf.NotConverged) as ei:
        sysm.minres(tol=1e-12, max_it=3, check_interval=3)
    assert ei.value.stats["iterations"] == 3


def test_minres_graph_replays_match_eager(ctx):
    """The hipGraph of three MINRES iterations is replayed many times with eager
    true-residual checks in between; every replay must reproduce the eager
    launch sequence, so the final iterate is bitwise identical."""
    m = meshes.uniform_mesh(2, 3)
    sp = apf.Space(m, 2, coarse=meshes.uniform_hierarchy(2, 3))
    sysm = apf.System(ctx, sp)
    sysm.assemble(meshes.smooth_initial(2))
    sysm.use_graph(True)
    xg, sg = sysm.minres(tol=1e-10, max_it=300, check_interval=9)
    sysm.use_graph(False)
    xe, se = sysm.minres(tol=1e-10, max_it=300, check_interval=9)
    assert sg["iterations"] == se["iterations"]
    assert np.array_equal(xg, xe)


@pytest.mark.parametrize("prec", [apf.APF_FP64, apf.APF_FP32_REFINE], ids=["mg64", "mg32"])
@pytest.mark.parametrize("dim,p,outer", [(1, 1, 5), (1, 2, 4), (2, 1, 3), (2, 2, 3), (2, 3, 2)])
def test_multigrid_minres_converges_and_matches(ctx, dim, p, outer, prec):
    m = meshes.uniform_mesh(dim, outer)
    coeff = meshes.smooth_initial(dim)
    if dim == 2:
        coeff = apf.Coefficients(meshes.checkerboard_kappa(m, 4, seed=11), 0.0, coeff.u0)
    print(f"mg minres dim={dim} p={p} outer={outer} prec={prec}", flush=True)
    sp = apf.Space(m, p, coarse=meshes.uniform_hierarchy(dim, outer))
    sysm = apf.System(ctx, sp)
    sysm.assemble(coeff)
    x, st = sysm.minres(tol=1e-10, max_it=600, check_interval=10, precond=prec)
    assert st["converged"] and st["true_relative_residual"] <= 1e-10
    assert st["iterations"] <= 150
    stm, od, prob = oracle_setup(m, p, coeff)
    S = OA.assemble(stm, od, prob)
    xd = OE.main_solve(S)
    assert np.linalg.norm(x - xd) <= 1e-7 * np.linalg.norm(xd)


def _pattern_case(name):
    if name == "uj-p1":
        return meshes.uniform_mesh(2, 2), 1
    if name == "uj-p3":
        return meshes.uniform_mesh(2, 1), 3
    if name == "uj-p4":
        return meshes.union_jack_mesh(4, np.linspace(0.0, 1.0, 3)), 4
    if name == "interval-p2":
        return meshes.uniform_mesh(1, 3), 2
    if name == "fan-graded-p2":
        return meshes.fan_mesh(2, meshes.graded_times(1, 4), seed=5), 2
    return meshes.segment_batch_mesh([8, 16, 24], np.linspace(0.0, 1.0, 4)), 1


@pytest.mark.parametrize("name", ["uj-p1", "uj-p3", "uj-p4", "interval-p2", "fan-graded-p2", "batch-p1"])
def test_device_pattern_equals_host_pattern(ctx, name):
    """The column indices generated on the device are bit-identical to the host pattern."""
    m, p = _pattern_case(name)
    print(f"pattern {name}", flush=True)
    sp = apf.Space(m, p)
    sysm = apf.System(ctx, sp)
    rp_d, col_d = sysm.download_pattern()
    rp_h, col_h = sp.pattern()
    assert np.array_equal(rp_d, rp_h)
    assert np.array_equal(col_d, col_h)


@pytest.mark.parametrize("name", ["uj-p1", "uj-p3", "uj-p4", "interval-p2", "fan-graded-p2", "batch-p1"])
def test_matrix_free_product_equals_csr(ctx, name):
    """y = K x from K's Kronecker factors (the product of the MINRES iterations) against
    the assembled CSR K on random x: every entry within 1e-14 of (|K| |x|)_i."""
    import scipy.sparse as sps
    m, p = _pattern_case(name)
    print(f"matrix-free product {name}", flush=True)
    sp = apf.Space(m, p)
    sysm = apf.System(ctx, sp)
    coeff = meshes.smooth_initial(m.dim)
    if m.dim == 2:
        coeff = apf.Coefficients(meshes.checkerboard_kappa(m, 4, seed=5), 0.0, coeff.u0)
    sysm.assemble(coeff)
    vals, _ = sysm.download()
    rowptr, col = sp.pattern()
    Kabs = sps.csr_matrix((np.abs(vals), col, rowptr), shape=(sp.n, sp.n))
    x = np.random.default_rng(3).standard_normal(sp.n)
    y_csr = sysm.matvec(x, matrix_free=False)
    y_mf = sysm.matvec(x, matrix_free=True)
    bound = Kabs @ np.abs(x)
    assert np.all(np.abs(y_mf - y_csr) <= 1e-14 * bound + 1e-300), np.max(np.abs(y_mf - y_csr) / (bound + 1e-300))


def test_minres_matrix_free_and_csr_products_agree(ctx, monkeypatch):
    """MINRES with the matrix-free Lanczos products and with the CSR products: both
    accepted on the assembled K's true residual, the same iterations (to the check
    interval) and solutions within 1e-8."""
    m = meshes.uniform_mesh(2, 3)
    sp = apf.Space(m, 2, coarse=meshes.uniform_hierarchy(2, 3))
    sysm = apf.System(ctx, sp)
    sysm.assemble(meshes.smooth_initial(2))
    x_mf, s_mf = sysm.minres(tol=1e-10, max_it=600, check_interval=10)
    monkeypatch.setenv("APF_MINRES_MATFREE", "0")
    x_csr, s_csr = sysm.minres(tol=1e-10, max_it=600, check_interval=10)
    assert s_mf["converged"] and s_csr["converged"]
    assert abs(s_mf["iterations"] - s_csr["iterations"]) <= 10
    assert np.linalg.norm(x_mf - x_csr) <= 1e-8 * np.linalg.norm(x_csr)
