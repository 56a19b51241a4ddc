"""The CPU baseline's C++/OpenMP restatement (oracle/cpu_baseline.cpp) computes what it
is timed for: its element assembly equals the oracle's K outside the t = 0 block (which
it does not assemble), its patch solves pass the 1e-10 acceptance, and its MINRES (the
Chebyshev-Jacobi preconditioned algorithm of apf_csr_minres_solve) reaches 1e-10."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import assembly as OA, cases, cpu_baseline as CB, mesh as OM


@pytest.mark.parametrize("dim,p,outer", [(1, 2, 2), (2, 1, 1), (2, 2, 2)])
def test_cpu_assembly_matches_oracle(dim, p, outer):
    stm = cases.uniform_mesh(dim, outer)
    tc = CB.TensorCase(stm, p)
    tc.assemble(2, 1)
    od = OM.spacetime_dofs(stm, p)
    K = OA.assemble(stm, od, cases.smooth_initial(stm)).K()
    assert np.array_equal(K.indptr, tc.rowptr) and np.array_equal(K.indices, tc.col)
    D = (sp.csr_matrix((tc.val, tc.col, tc.rowptr), shape=K.shape) - K).tocoo()
    keep = ~((D.row >= od.nY) & (D.col >= od.nY))
    assert np.abs(D.data[keep]).max() <= 1e-13 * abs(K).max()


def test_cpu_patch_solves_and_minres():
    stm = cases.uniform_mesh(2, 1)
    tc = CB.TensorCase(stm, 2)
    orders = tc.patch_orders()
    assert set(orders.tolist()) <= {99, 183, 189, 363}
    _, stat = CB.patch_solves(orders, 2)
    assert stat <= 1e-10
    od = OM.spacetime_dofs(stm, 2)
    S = OA.assemble(stm, od, cases.smooth_initial(stm))
    K = S.K()
    _, it, rel = CB.minres(K.indptr.astype(np.int64), K.indices.astype(np.int32), K.data, S.rhs(), od.nY, 2, 2000,
                          check_interval=20)
    assert rel <= 1e-10 and it < 2000
