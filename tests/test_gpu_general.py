"""GPU element-local quadrature assembly on non-tensor space-time meshes (hanging X
nodes at layer interfaces, Y cells finer than their X ancestor), through the C ABI,
against the oracle's subspace assembly P^T K_fine P; MINRES on the assembled K."""
import numpy as np
import pytest
import scipy.sparse as sp

import doubleadapparabolicfem_amd as apf
from doubleadapparabolicfem_amd import layered as L
from doubleadapparabolicfem_amd import meshes
from oracle import estimator as OE, layered as OL

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return apf.Context(0)


def kappa_grid(dim, k, seed):
    """piecewise-constant M = kappa I on a k^dim grid, kappa = 10^U(-1,1)"""
    rng = np.random.default_rng(seed)
    kap = 10.0 ** rng.uniform(-1.0, 1.0, size=k ** dim)
    return kap[:, None, None] * np.eye(dim)[None, :, :]


CASES = [
    (1, 1, [(0.0, 0.25, 16), (0.25, 0.5, 8), (0.5, 1.0, 4)], [(0, 1), (1, 2), (0, 1)], False),
    (1, 2, [(0.0, 0.125, 16), (0.125, 0.25, 8), (0.25, 1.0, 4)], [(1, 1), (0, 2), (0, 1)], True),
    (1, 3, [(0.0, 0.5, 8), (0.5, 1.0, 4)], None, True),
    (2, 1, L.graded_layers(4, 2), None, True),
    (2, 1, [(0.0, 0.25, 8), (0.25, 1.0, 4)], [(1, 2), (0, 1)], False),
    (2, 2, [(0.0, 0.25, 8), (0.25, 0.5, 4), (0.5, 1.0, 4)], [(0, 1), (1, 2), (0, 1)], True),
    (2, 3, [(0.0, 0.5, 8), (0.5, 1.0, 4)], [(0, 1), (1, 1)], True),
]


def _ids(c):
    return f"{c[0]}d-p{c[1]}-{len(c[2])}layers" + ("-yref" if c[3] else "") + ("-kappa" if c[4] else "")


def poly_u0(x):
    """u0 = prod 4 x_i (1 - x_i): the bottom rules of the coarse X cells and of the fine
    oracle mesh both integrate u0 psi exactly, so g is compared at round-off (with
    sin(pi x) sin(pi y) the two quadratures differ by ~1e-9 where the bottom X cells are
    coarser than the finest Y cells)"""
    return np.prod(4.0 * x * (1.0 - x), axis=1)


def _assemble(ctx, case, ell=0.0):
    dim, p, xl, yr, kap = case
    lm = L.layered_mesh(dim, p, xl, yr)
    u0 = poly_u0
    k = 4
    Mg = kappa_grid(dim, k, seed=17) if kap else None
    gs = apf.GeneralSystem(ctx, lm)
    gs.assemble(u0, Mg, k, ell)
    S = OL.assemble(lm, Mg, k, ell, u0)
    return lm, gs, S


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_general_assembly_matches_oracle(ctx, case):
    """every entry of K and b within 1e-12 relative (plus 1e-14 of the matrix scale for
    entries that cancel) of P^T K_fine P, on the bit-identical pattern"""
    print(f"general assembly {_ids(case)}", flush=True)
    lm, gs, S = _assemble(ctx, case, ell=0.7)
    rowptr, col, vals, b = gs.download()
    r2, c2 = OL.symbolic_pattern(lm)
    assert np.array_equal(rowptr, r2) and np.array_equal(col, c2)
    K = S.K()
    Kd = sp.csr_matrix((vals, col, rowptr), shape=K.shape)
    D = (Kd - K).tocoo()
    scale = abs(K).max()
    Kabs = abs(K).tocsr()
    mag = np.asarray(Kabs[D.row, D.col]).ravel()
    assert np.all(np.abs(D.data) <= 1e-12 * mag + 1e-14 * scale), np.abs(D.data).max()
    rhs = S.rhs()
    assert np.all(np.abs(b - rhs) <= 1e-12 * np.abs(rhs) + 1e-15 * np.abs(rhs).max())


def test_general_path_equals_tensor_path_on_a_tensor_mesh(ctx):
    """Layers of one resolution, Y on the X cells: the same numbering as the tensor path,
    and the element-quadrature kernel agrees with the tensor path's assembly entrywise."""
    print("general vs tensor path", flush=True)
    dim, p = 2, 2
    lm = L.layered_mesh(dim, p, [(0.0, 0.5, 8), (0.5, 1.0, 8)])
    u0 = meshes.smooth_initial(dim).u0
    gs = apf.GeneralSystem(ctx, lm)
    gs.assemble(u0)
    rowptr, col, vals, b = gs.download()
    m = meshes.union_jack_mesh(8, [0.0, 0.5, 1.0])
    sp_ = apf.Space(m, p)
    sysm = apf.System(ctx, sp_)
    sysm.assemble(meshes.smooth_initial(dim))
    v2, b2 = sysm.download()
    r2, c2 = sp_.pattern()
    assert np.array_equal(rowptr, r2) and np.array_equal(col, c2)
    scale = np.abs(v2).max()
    assert np.all(np.abs(vals - v2) <= 1e-12 * np.abs(v2) + 1e-14 * scale)
    assert np.all(np.abs(b - b2) <= 1e-12 * np.abs(b2) + 1e-15 * np.abs(b2).max())


@pytest.mark.parametrize("case", [CASES[1], CASES[3], CASES[5]], ids=_ids)
def test_minres_on_the_general_system(ctx, case):
    """K assembled on the device, solved on the device (caller's-CSR MINRES on a view of
    K): within 1e-7 of the oracle's direct solve of its own system, true residual <= 1e-10"""
    print(f"general minres {_ids(case)}", flush=True)
    lm, gs, S = _assemble(ctx, case)
    _, _, _, b = gs.download()
    K = gs.matrix()
    x, st = K.minres(b, gs.nY, tol=1e-10, max_it=50000, check_interval=20, degree=8)
    assert st["converged"] == 1
    xd = OE.main_solve(S)
    assert np.linalg.norm(x - xd) <= 1e-7 * np.linalg.norm(xd)


def _groups(lm):
    """slab groups of a layered mesh as tensor meshes with their multigrid hierarchies"""
    out = []
    for g in L.slab_groups(lm):
        n, times = g["n"], g["times"]
        mk = meshes.interval_mesh if lm.dim == 1 else meshes.union_jack_mesh
        coarse, k = [], n // 2
        while k >= 2:
            coarse.append(mk(k, [0.0, 1.0]))
            k //= 2
        out.append((mk(n, times), coarse, g["first_y_cell"]))
    return out


EST_CASES = [CASES[0], CASES[1], CASES[3], CASES[5], CASES[6]]


@pytest.mark.parametrize("prec", [apf.APF_FP64, apf.APF_FP32_REFINE], ids=["fp64", "fp32ref"])
@pytest.mark.parametrize("case", EST_CASES, ids=_ids)
def test_general_estimator_matches_oracle(ctx, case, prec):
    """The estimator on a non-tensor mesh (patches per (slab, vertex) on each slab's own
    mesh, u_delta through the X ancestors, slab reconstruction from that u) against the
    oracle's slab-group decomposition with direct slab solves, on the oracle's direct
    solution with lambda perturbed by 10% (lambda_tilde depends on u only, so the
    reconstruction term is far from zero): every total and per-cell value within 1e-10
    relative."""
    print(f"general estimator {_ids(case)} prec={prec}", flush=True)
    dim, p, xl, yr, kap = case
    lm, gs, S = _assemble(ctx, case, ell=0.3)
    x = OE.main_solve(S)
    nY = lm.Y.n_dofs
    x[:nY] *= 1.0 + 0.1 * np.random.default_rng(9).standard_normal(nY)
    k = 4
    Mg = kappa_grid(dim, k, seed=17) if kap else None
    ge = apf.GeneralEstimator(gs, _groups(lm))
    ge.assemble(poly_u0, Mg, k, 0.3)
    est = ge.estimate(x, prec, 2)
    ref = OL.estimate(lm, x, Mg, k, 0.3, poly_u0)
    assert est["max_patch_residual"] <= 1e-10
    assert ref["recon2"] > 1e-8 * ref["y_est2"]  # the slabs' reconstruction is not lambda itself
    for key in ("flux2", "recon2", "lambda_y2", "trace2", "eta2", "y_est2"):
        assert abs(est[key] - ref[key]) <= 1e-10 * abs(ref[key]), (key, est[key], ref[key])
    for key in ("flux_by_cell", "recon_by_cell"):
        a, b = est[key], ref[key]
        assert np.all(np.abs(a - b) <= 1e-9 * np.abs(b) + 1e-13 * np.abs(b).sum()), key
