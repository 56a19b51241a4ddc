"""Non-tensor (layered) space-time meshes on the host, no GPU: the cell restriction rows
describe continuous X functions across the hanging interfaces, the Y cells lie in their
X ancestors, the oracle's subspace assembly reduces to the tensor oracle when nothing
hangs, and the product's host-side CSR pattern (apf_general_pattern) is bit-identical
to the oracle's symbolic pattern."""
import numpy as np
import pytest

import doubleadapparabolicfem_amd as apf
from doubleadapparabolicfem_amd import layered as L
from doubleadapparabolicfem_amd import meshes
from oracle import assembly as OA, layered as OL, mesh as OM
from oracle.basis import segment_lagrange, spatial_lagrange

CASES = [
    (1, 1, [(0.0, 0.25, 16), (0.25, 0.5, 8), (0.5, 1.0, 4)], [(0, 1), (1, 2), (0, 1)]),
    (1, 3, [(0.0, 0.5, 8), (0.5, 1.0, 2)], None),
    (2, 1, L.graded_layers(2, 2), None),
    (2, 2, [(0.0, 0.25, 8), (0.25, 0.5, 4), (0.5, 1.0, 4)], [(0, 1), (1, 2), (0, 1)]),
    (2, 3, [(0.0, 0.5, 4), (0.5, 1.0, 2)], [(0, 2), (1, 1)]),
]


def _ids(c):
    return f"{c[0]}d-p{c[1]}-{len(c[2])}layers" + ("-yref" if c[3] else "")


def _eval_x(lm, u, k, pts, tau):
    """X function with true coefficients u on layer k's cells at points pts, time tau in [0, 1]."""
    X = lm.X
    t0, t1, n = lm.x_layers[k]
    base = int(np.argmax(X.layer == k))
    c, xi = L.locate(lm.dim, n, pts)
    sv, _ = spatial_lagrange(lm.dim, lm.p, xi)
    tv, _ = segment_lagrange(lm.p, np.full(len(pts), tau))
    Nt = lm.p + 1
    out = np.zeros(len(pts))
    for a in range(len(pts)):
        for s in range(sv.shape[1]):
            for it in range(Nt):
                d, w = X.rows(base + c[a], s * Nt + it)
                out[a] += sv[a, s] * tv[a, it] * np.dot(w, u[d])
    return out


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_x_is_continuous_across_hanging_interfaces(case):
    """At every interface between layers the X function evaluated on the cells below
    (at tau = 1) and above (tau = 0) agrees: the hanging-node rows of the finer side
    interpolate the coarser side's function."""
    dim, p, xl, yr = case
    lm = L.layered_mesh(dim, p, xl, yr)
    rng = np.random.default_rng(1)
    u = rng.standard_normal(lm.X.n_dofs)
    pts = rng.uniform(0.0, 1.0, size=(40, dim))
    n_hanging = 0
    for k in range(len(xl) - 1):
        below = _eval_x(lm, u, k, pts, 1.0)
        above = _eval_x(lm, u, k + 1, pts, 0.0)
        assert np.allclose(below, above, rtol=0, atol=1e-12 * np.abs(u).max()), k
        n_hanging += xl[k][2] != xl[k + 1][2]
    assert n_hanging >= 1
    # hanging rows carry more than one master somewhere
    assert np.diff(lm.X.restr_ptr).max() > 1


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_y_cells_lie_in_their_x_ancestors(case):
    dim, p, xl, yr = case
    lm = L.layered_mesh(dim, p, xl, yr)
    X, Y = lm.X, lm.Y
    for c in range(Y.n_cells):
        a = lm.y_ancestor[c]
        assert X.times[a, 0] <= Y.times[c, 0] and Y.times[c, 1] <= X.times[a, 1]
        V = Y.vertices[Y.cells[c]]
        W = X.vertices[X.cells[a]]
        if dim == 1:
            assert W.min() - 1e-14 <= V.min() and V.max() <= W.max() + 1e-14
        else:
            J = np.stack([W[1] - W[0], W[2] - W[0]], axis=1)
            xi = np.linalg.solve(J, (V - W[0]).T).T
            assert np.all(xi >= -1e-12) and np.all(xi.sum(axis=1) <= 1 + 1e-12)


@pytest.mark.parametrize("dim,p", [(1, 2), (2, 1), (2, 2)])
def test_oracle_reduces_to_tensor_oracle_without_hanging_nodes(dim, p):
    """Layers of one resolution and Y on the X cells: the layered numbering is the tensor
    numbering and P is the identity, so the subspace assembly equals the tensor oracle
    (pinned by closed forms and the published anchors) bit for bit."""
    n = 4
    lm = L.layered_mesh(dim, p, [(0.0, 0.5, n), (0.5, 1.0, n)])
    u0 = meshes.smooth_initial(dim).u0
    S = OL.assemble(lm, u0=u0)
    sm = OM.interval_mesh(n) if dim == 1 else OM.union_jack_mesh(n)
    stm = OM.SpaceTimeMesh(sm, np.array([0.0, 0.5, 1.0]))
    St = OA.assemble(stm, OM.spacetime_dofs(stm, p), OA.Problem(OA.identity_M(sm), 0.0, u0))
    assert (S.K() != St.K()).nnz == 0
    assert np.array_equal(S.rhs(), St.rhs())


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_host_pattern_equals_oracle_pattern(case):
    dim, p, xl, yr = case
    lm = L.layered_mesh(dim, p, xl, yr)
    rowptr, col = apf.general_pattern(lm)
    r2, c2 = OL.symbolic_pattern(lm)
    assert np.array_equal(rowptr, r2)
    assert np.array_equal(col, c2)
    # and it holds every entry of the oracle's matrix (outside it only the round-off of
    # basis values that vanish at the fine nodes in exact arithmetic)
    K = OL.assemble(lm, u0=meshes.smooth_initial(dim).u0).K()
    import scipy.sparse as sp
    Pm = sp.csr_matrix((np.ones(len(col)), col, rowptr), shape=K.shape)
    assert abs(K - K.multiply(Pm)).max() <= 1e-13 * abs(K).max()


def test_graded_layers_follow_the_split_policy():
    """two refinement rounds per level (spatial bisection, then space-time split) of the
    cells with t0 = 0: layer resolutions double and layer lengths halve toward t = 0,
    the bottom two layers share the finest resolution"""
    layers = L.graded_layers(2, 3)
    assert layers == [(0.0, 0.125, 16), (0.125, 0.25, 16), (0.25, 0.5, 8), (0.5, 1.0, 4)]


def test_host_pattern_rejects_bad_rows():
    lm = L.layered_mesh(2, 1, [(0.0, 0.5, 4), (0.5, 1.0, 2)])
    lm.X.restr_dof = lm.X.restr_dof.copy()
    lm.X.restr_dof[0] = lm.X.n_dofs + 5
    with pytest.raises(apf.APFError, match="true dof out of range"):
        apf.general_pattern(lm)
