"""Batches of independent 1+1D problems in one system (segments that share no
vertex): every problem's solution and estimator values equal the oracle's
solve of that problem alone."""
import numpy as np
import pytest

import doubleadapparabolicfem_amd as apf
from doubleadapparabolicfem_amd import meshes
from oracle import assembly as OA, estimator as OE

from gpu_helpers import oracle_setup

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return apf.Context(0)


def _problem_coefficients(mesh, kappa, coef):
    def u0(x):
        k = np.arange(1, coef.size + 1)
        return np.sum(coef[None, :] * np.sin(np.pi * x[:, :1] * k[None, :]), axis=1)
    return apf.Coefficients(np.full((mesh.cells.shape[0], 1, 1), kappa), 0.0, u0)


@pytest.mark.parametrize("p", [1, 2])
def test_batch_equals_independent_solves(ctx, p):
    nx = [8, 16, 24]
    times = np.linspace(0.0, 1.0, 5)
    rng = np.random.default_rng(17 + p)
    kappa = 10.0 ** rng.uniform(-1.0, 1.0, len(nx))
    coef = rng.standard_normal((len(nx), 5))
    bm = meshes.segment_batch_mesh(nx, times)
    sp = apf.Space(bm, p, coarse=meshes.segment_batch_hierarchy(nx, 2))
    print(f"batch p={p} nx={nx} n={sp.n}", flush=True)
    nf_tot, ns_tot, m, Nt = sp.info["nf"], sum(nx), len(times) - 1, p + 1
    # oracle: every problem alone on its own interval
    refs, xs, nfs = [], [], []
    for i, n in enumerate(nx):
        mi = meshes.interval_mesh(n, times)
        stm, od, prob = oracle_setup(mi, p, _problem_coefficients(mi, kappa[i], coef[i]))
        S = OA.assemble(stm, od, prob)
        x = OE.main_solve(S)
        refs.append(OE.estimate(stm, od, prob, S, x))
        xs.append((x, od))
        nfs.append(od.sd.n_free)
    assert sum(nfs) == nf_tot
    # the oracle solutions in the batch numbering (free nodes of segment i follow those of 0..i-1)
    xb = np.zeros(sp.n)
    off = 0
    for (x, od), nfi in zip(xs, nfs):
        for k in range(m):
            for l in range(Nt):
                src = (k * Nt + l) * nfi
                dst = (k * Nt + l) * nf_tot + off
                xb[dst:dst + nfi] = x[src:src + nfi]
        for tn in range(m * p + 1):
            src = od.nY + tn * nfi
            dst = sp.nY + tn * nf_tot + off
            xb[dst:dst + nfi] = x[src:src + nfi]
        off += nfi
    sysm = apf.System(ctx, sp)
    sysm.assemble(meshes.segment_batch_coefficients(bm, nx, kappa, coef))
    # MINRES on the whole batch reaches the independent direct solves
    xg, st = sysm.minres(tol=1e-12, max_it=20000, check_interval=10)
    assert st["true_relative_residual"] <= 1e-12
    assert np.linalg.norm(xg - xb) <= 1e-9 * np.linalg.norm(xb)
    # estimator on the oracle's solution: per problem, per cell, to 1e-10
    sysm.set_solution(xb)
    for prec in (apf.APF_FP64, apf.APF_FP32_REFINE):
        est = sysm.estimate(prec, 2)
        cell_off = 0
        for i, (n, ref) in enumerate(zip(nx, refs)):
            got = est["flux_by_cell"].reshape(m, ns_tot)[:, cell_off:cell_off + n].ravel()
            arr = ref.flux_by_cell
            assert np.all(np.abs(got - arr) <= 1e-10 * np.abs(arr) + 1e-14 * arr.max()), (prec, i)
            # exact solution: the slab reconstruction reproduces lambda, both sides ~ round-off
            got = est["recon_by_cell"].reshape(m, ns_tot)[:, cell_off:cell_off + n].ravel()
            assert got.max() <= 1e-25 and ref.recon_by_cell.max() <= 1e-25
            cell_off += n
        tot = sum(r.flux2 for r in refs)
        assert abs(est["flux2"] - tot) <= 1e-10 * tot
        assert abs(est["lambda_y2"] - sum(r.lambda_y2 for r in refs)) <= 1e-10 * sum(r.lambda_y2 for r in refs)
