"""Fan-patch meshes (vertex patches of up to 64 cells, orders up to 1665 per
time mode) on a time grid graded toward t = 0: the estimator's large-patch path
(Gauss-Jordan in global memory, solve tiles in a global workspace) and MINRES
against the oracle's results committed under tests/golden (written by
tests/golden/make_fan_fixtures.py)."""
import os

import numpy as np
import pytest

import doubleadapparabolicfem_amd as apf
from doubleadapparabolicfem_amd import meshes

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def ctx():
    return apf.Context(0)


@pytest.mark.parametrize("name", ["fan_p1", "fan_p2", "fan_p3"])
def test_fan_patches_match_oracle(ctx, name):
    fx = np.load(os.path.join(GOLDEN, name + ".npz"))
    p, q, seed, smin, smax, lmin, lmax = (int(v) for v in fx["params"])
    m = meshes.fan_mesh(q, meshes.graded_times(lmin, lmax), seed, smin, smax)
    sp = apf.Space(m, p, coarse=meshes.fan_hierarchy(q))
    nf, ns_, nc, bd = sp.patch_sizes()
    order = nf + ns_ + (bd == 0)
    print(f"{name}: n={sp.n} patches={order.size} max cells={nc.max()} max order={order.max()}", flush=True)
    assert nc.max() >= 24
    coeff = meshes.indicator_initial()
    sysm = apf.System(ctx, sp)
    sysm.assemble(coeff)
    xr = fx["x"]
    x, st = sysm.minres(tol=1e-10, max_it=20000, check_interval=20)
    assert st["true_relative_residual"] <= 1e-10
    assert np.linalg.norm(x - xr) <= 1e-7 * np.linalg.norm(xr)
    sysm.set_solution(xr)
    ref_f, ref_r = fx["flux_by_cell"], fx["recon_by_cell"]
    for prec in (apf.APF_FP64, apf.APF_FP32_REFINE):
        est = sysm.estimate(prec, 2)
        print(f"  prec={prec} max patch residual {est['max_patch_residual']:.2e} "
              f"extra refinements {est['extra_refinements']}", flush=True)
        assert est["max_patch_residual"] <= 1e-10
        assert abs(est["flux2"] - float(fx["flux2"])) <= 1e-10 * float(fx["flux2"])
        assert abs(est["lambda_y2"] - float(fx["lambda_y2"])) <= 1e-10 * float(fx["lambda_y2"])
        fc = est["flux_by_cell"]
        assert np.all(np.abs(fc - ref_f) <= 1e-10 * np.abs(ref_f) + 1e-14 * ref_f.sum()), prec
        assert est["recon_by_cell"].max() <= 1e-25 and ref_r.max() <= 1e-25
