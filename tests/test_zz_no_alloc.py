"""The allocation-free CSR MINRES (apf_csr_minres_solve_noalloc), through the C ABI,
against the allocating entry point apf_csr_minres_solve on the same matrix: two calls in
a row, the second with a smaller system, each bit-identical to the allocating solve
(same kernels in the same order) and so within the 1e-10 bar of the oracle's direct
solve that the allocating entry is held to."""
import numpy as np
import pytest
import scipy.sparse as sp

import doubleadapparabolicfem_amd as apf
from doubleadapparabolicfem_amd import meshes
from oracle import assembly as OA, estimator as OE

from gpu_helpers import oracle_setup

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return apf.Context(0)


def _saddle(dim, p, outer):
    m = meshes.uniform_mesh(dim, outer)
    stm, od, prob = oracle_setup(m, p, meshes.smooth_initial(dim))
    return OA.assemble(stm, od, prob)


@pytest.mark.parametrize("prec", [apf.APF_FP64, apf.APF_FP32_REFINE], ids=["pc64", "pc32"])
def test_noalloc_equals_allocating_solve_on_two_calls(ctx, prec, monkeypatch):
    # first call: 2+1D p=2 (n = 4,725); second: 1+1D p=1, smaller (n = 3,007)
    for dim, p, outer in ((2, 2, 2), (1, 1, 4)):
        print(f"noalloc csr minres dim={dim} p={p} outer={outer} prec={prec}", flush=True)
        S = _saddle(dim, p, outer)
        K = S.K()
        b = S.rhs()
        nY = S.A.shape[0]
        n = K.shape[0]
        assert n <= apf.APF_NOALLOC_MAX_N
        Km = apf.CsrMatrix(ctx, K)
        # the allocating solve reading K's lambda rows, as the allocation-free one must (it cannot
        # build the compact lambda block); graph replays and the two streams give its eager iterate
        monkeypatch.setenv("APF_CSR_COMPACT", "0")
        x_ref, st_ref = Km.minres(b, nY, tol=1e-10, max_it=20000, check_interval=20, degree=6, precond=prec)
        monkeypatch.delenv("APF_CSR_COMPACT")
        d_b, d_x = apf.DeviceArray(ctx, n), apf.DeviceArray(ctx, n)
        d_b.upload(b)
        d_x.upload(np.full(n, np.nan))  # the solve must not read d_x without use_x0
        st = apf.csr_minres_noalloc(Km, d_b, d_x, nY, tol=1e-10, max_it=20000, check_interval=20, degree=6,
                                    precond=prec)
        x = d_x.download()
        assert st["converged"] == 1 and st["iterations"] == st_ref["iterations"]
        assert np.array_equal(x, x_ref)
        assert st["true_relative_residual"] <= 1e-10
        xd = OE.main_solve(S)
        assert np.linalg.norm(x - xd) <= 1e-7 * np.linalg.norm(xd)


def test_noalloc_initial_guess_and_size_limit(ctx):
    print("noalloc csr minres: initial guess and size limit", flush=True)
    S = _saddle(1, 2, 3)
    K = S.K()
    b = S.rhs()
    nY = S.A.shape[0]
    n = K.shape[0]
    Km = apf.CsrMatrix(ctx, K)
    xd = OE.main_solve(S)
    d_b, d_x = apf.DeviceArray(ctx, n), apf.DeviceArray(ctx, n)
    d_b.upload(b)
    d_x.upload(xd)  # start at the direct solution: accepted before any iteration
    st = apf.csr_minres_noalloc(Km, d_b, d_x, nY, tol=1e-10, check_interval=20, use_x0=True)
    assert st["converged"] == 1 and st["iterations"] == 0
    assert np.array_equal(d_x.download(), xd)
    big = apf.CsrMatrix(ctx, sp.identity(apf.APF_NOALLOC_MAX_N + 1, format="csr"))
    e_b, e_x = apf.DeviceArray(ctx, big.n), apf.DeviceArray(ctx, big.n)
    with pytest.raises(apf.APFError, match="scratch limit"):
        apf.csr_minres_noalloc(big, e_b, e_x, 1)
