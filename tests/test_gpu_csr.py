"""MINRES on a K handed over by the caller (apf_csr_minres_solve), through the C ABI:
the oracle assembles K, f, g and the graph-norm matrix Hhat the way the reference's
assembler does, and the GPU solve is held to the oracle's direct solve."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg  # noqa: F401  (sp.linalg)

import doubleadapparabolicfem_amd as apf
from doubleadapparabolicfem_amd import meshes
from oracle import assembly as OA, estimator as OE

from gpu_helpers import oracle_setup

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return apf.Context(0)


def _system(dim, p, outer, kappa=False):
    m = meshes.uniform_mesh(dim, outer)
    coeff = meshes.smooth_initial(dim)
    if kappa:
        coeff = apf.Coefficients(meshes.checkerboard_kappa(m, 4, seed=11), 0.0, coeff.u0)
    stm, od, prob = oracle_setup(m, p, coeff)
    S = OA.assemble(stm, od, prob)
    return S


def graph_norm_hhat(S):
    """Hhat = -C_signed + B_dt diag(A_y)^-1 B_dt^T + A_x, the reference's graph-norm
    approximation (src/linear_algebra/preconditioners/schur_approximations.hpp:233-287),
    formed by the caller."""
    dA = 1.0 / S.A.diagonal()
    H = -S.C + S.Bdt @ sp.diags(dA) @ S.Bdt.T + S.Ax
    return sp.csr_matrix(H)


@pytest.mark.parametrize("prec", [apf.APF_FP64, apf.APF_FP32_REFINE], ids=["pc64", "pc32"])
@pytest.mark.parametrize("with_hhat", [True, False], ids=["hhat", "schur-diag"])
@pytest.mark.parametrize("dim,p,outer,kappa", [(1, 1, 3, False), (1, 2, 2, False), (1, 3, 1, False),
                                               (2, 1, 1, True), (2, 2, 1, False), (2, 1, 2, True)])
def test_csr_minres_matches_direct_solve(ctx, dim, p, outer, kappa, with_hhat, prec):
    print(f"csr minres dim={dim} p={p} outer={outer} kappa={kappa} hhat={with_hhat} prec={prec}", flush=True)
    S = _system(dim, p, outer, kappa)
    K = S.K()
    b = S.rhs()
    xd = OE.main_solve(S)
    nY = S.A.shape[0]
    Km = apf.CsrMatrix(ctx, K)
    assert (Km.n, Km.nnz) == (K.shape[0], K.nnz)
    H = apf.CsrMatrix(ctx, graph_norm_hhat(S)) if with_hhat else None
    x, st = Km.minres(b, nY, hhat=H, tol=1e-10, max_it=20000, check_interval=20, degree=6, precond=prec)
    assert st["converged"] == 1
    assert st["true_relative_residual"] <= 1e-10
    assert np.linalg.norm(K @ x - b) <= 1e-10 * np.linalg.norm(b) * (1 + 1e-6)
    assert np.linalg.norm(x - xd) <= 1e-7 * np.linalg.norm(xd)


def test_csr_triangle_storage_is_the_symmetric_completion(ctx):
    """Only the lower triangle of K in column-major storage (the part Eigen::MINRES with
    Eigen::Lower reads) gives the same matrix and the same solution."""
    print("csr triangle storage", flush=True)
    S = _system(2, 2, 1)
    K = S.K()
    b = S.rhs()
    nY = S.A.shape[0]
    L = sp.tril(K).tocsc()
    L.sort_indices()
    full = apf.CsrMatrix(ctx, K)
    tri = apf.CsrMatrix(ctx, outer=L.indptr, inner=L.indices, values=L.data, storage=apf.APF_CSR_TRIANGLE)
    assert (tri.n, tri.nnz) == (full.n, full.nnz)
    xf, sf = full.minres(b, nY, tol=1e-10, check_interval=20, degree=4)
    xt, stt = tri.minres(b, nY, tol=1e-10, check_interval=20, degree=4)
    assert stt["converged"] == 1
    assert np.linalg.norm(xt - xf) <= 1e-8 * np.linalg.norm(xf)


def test_csr_matches_the_system_solution_and_repeats_bitwise(ctx):
    """K, b downloaded from a GPU-assembled system and handed back as a caller's matrix:
    the CSR solve agrees with the system's multigrid MINRES, and a second solve of the
    same input is bit-identical."""
    print("csr vs system minres", flush=True)
    m = meshes.uniform_mesh(2, 2)
    sp_ = apf.Space(m, 2, coarse=meshes.uniform_hierarchy(2, 2))
    sysm = apf.System(ctx, sp_)
    sysm.assemble(meshes.smooth_initial(2))
    vals, b = sysm.download()
    rowptr, col = sysm.download_pattern()
    xs, _ = sysm.minres(tol=1e-11, max_it=2000, check_interval=10)
    K = sp.csr_matrix((vals, col, rowptr), shape=(sp_.n, sp_.n))
    Km = apf.CsrMatrix(ctx, K)
    x1, s1 = Km.minres(b, sp_.nY, tol=1e-11, check_interval=25, degree=8)
    x2, s2 = Km.minres(b, sp_.nY, tol=1e-11, check_interval=25, degree=8)
    assert np.array_equal(x1, x2) and s1["iterations"] == s2["iterations"]
    assert np.linalg.norm(x1 - xs) <= 1e-8 * np.linalg.norm(xs)


def test_csr_rejects_bad_input(ctx):
    print("csr argument checks", flush=True)
    K = sp.identity(4, format="csr")
    Km = apf.CsrMatrix(ctx, K)
    with pytest.raises(apf.APFError, match="n_lambda"):
        Km.minres(np.ones(4), 0)
    with pytest.raises(apf.APFError, match="Chebyshev degree"):
        Km.minres(np.ones(4), 2, degree=0)
    with pytest.raises(apf.APFError, match="inner index"):
        apf.CsrMatrix(ctx, outer=[0, 1, 2], inner=[0, 5], values=[1.0, 1.0])
    # an indefinite lambda block is refused before any iteration
    bad = sp.csr_matrix(np.diag([-1.0, 1.0, 1.0, 1.0]))
    with pytest.raises(apf.APFError, match="not positive definite"):
        apf.CsrMatrix(ctx, bad).minres(np.ones(4), 2)


@pytest.mark.parametrize("with_hhat", [True, False], ids=["hhat", "schur-diag"])
def test_csr_graph_and_block_streams_match_eager_bitwise(ctx, monkeypatch, with_hhat):
    """The iterations replayed from a hipGraph, with the X block's Chebyshev chain on a second
    stream beside the lambda block's, give the same iterate bit for bit as eager launches on
    one stream (APF_CSR_GRAPH=0, APF_CSR_FORK=0), at the same iteration count, with a check
    interval that is no multiple of 3 (graph replays plus eager iterations per batch); so does
    each switch alone."""
    print(f"csr graph vs eager hhat={with_hhat}", flush=True)
    S = _system(2, 2, 1)
    K = S.K()
    b = S.rhs()
    nY = S.A.shape[0]
    Km = apf.CsrMatrix(ctx, K)
    H = apf.CsrMatrix(ctx, graph_norm_hhat(S)) if with_hhat else None
    out = {}
    for g, f in ((1, 1), (0, 0), (1, 0), (0, 1)):
        monkeypatch.setenv("APF_CSR_GRAPH", str(g))
        monkeypatch.setenv("APF_CSR_FORK", str(f))
        out[(g, f)] = Km.minres(b, nY, hhat=H, tol=1e-10, max_it=20000, check_interval=20, degree=8)
    x0, s0 = out[(0, 0)]
    assert s0["converged"] == 1
    for key, (x, st) in out.items():
        assert st["iterations"] == s0["iterations"], key
        assert np.array_equal(x, x0), key
    # the lambda block's steps on K's lambda rows instead of the compact block: the same sums with
    # other lane partials, so the same solution to round-off
    monkeypatch.setenv("APF_CSR_COMPACT", "0")
    xk, sk = Km.minres(b, nY, hhat=H, tol=1e-10, max_it=20000, check_interval=20, degree=8)
    assert sk["converged"] == 1 and abs(sk["iterations"] - s0["iterations"]) <= 20
    assert np.linalg.norm(xk - x0) <= 1e-8 * np.linalg.norm(x0)


def test_block_pcg_solves_the_leading_block_and_graph_matches_eager(ctx, monkeypatch):
    """apf_csr_block_pcg_solve on K's SPD leading block A_y: the true residual meets the
    tolerance, the solution equals scipy's direct solve to 1e-9, apf_csr_block_energy equals
    d^T A_y d, apf_csr_spmv equals K x; batches replayed from a hipGraph give the eager
    iterate bit for bit (APF_CSR_GRAPH=0)."""
    print("csr block pcg", flush=True)
    S = _system(2, 2, 1)
    K = S.K().tocsr()
    nY = S.A.shape[0]
    A = K[:nY, :nY].tocsc()
    rng = np.random.default_rng(7)
    b = rng.standard_normal(nY)
    Km = apf.CsrMatrix(ctx, K)
    monkeypatch.setenv("APF_CSR_GRAPH", "1")
    x1, s1 = Km.block_pcg(b, nY, tol=1e-12, check_interval=20, degree=8)
    monkeypatch.setenv("APF_CSR_GRAPH", "0")
    x0, s0 = Km.block_pcg(b, nY, tol=1e-12, check_interval=20, degree=8)
    assert s1["converged"] == 1 and s1["true_relative_residual"] <= 1e-12
    assert s1["iterations"] == s0["iterations"] and np.array_equal(x1, x0)
    xd = sp.linalg.spsolve(A, b)
    assert np.linalg.norm(x1 - xd) <= 1e-9 * np.linalg.norm(xd)
    d = rng.standard_normal(nY)
    assert Km.block_energy(d, nY) == pytest.approx(float(d @ (A @ d)), rel=1e-12)
    v = rng.standard_normal(K.shape[0])
    np.testing.assert_allclose(Km.spmv(v), K @ v, rtol=1e-12, atol=1e-12 * np.abs(K @ v).max())
