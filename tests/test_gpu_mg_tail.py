"""The fused coarse tail of the multigrid V-cycle (k_mg_tail: the levels from the
first one whose vectors fit one workgroup's LDS down to the dense coarsest solve, one
workgroup per column) against the per-level launches (APF_MG_TAIL=0) of the same
V-cycle: MINRES takes the same iterations to the same solution, and the slab
reconstruction solve (the Y columns' V-cycle alone) the same lambda_tilde."""
import os

import numpy as np
import pytest

import doubleadapparabolicfem_amd as apf
from doubleadapparabolicfem_amd import meshes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return apf.Context(0)


class _tail_env:
    """APF_MG_TAIL for the multigrid builds inside the block (read when a hierarchy is built)"""

    def __init__(self, on):
        self.on = on

    def __enter__(self):
        self.old = os.environ.get("APF_MG_TAIL")
        os.environ["APF_MG_TAIL"] = "1" if self.on else "0"

    def __exit__(self, *exc):
        if self.old is None:
            del os.environ["APF_MG_TAIL"]
        else:
            os.environ["APF_MG_TAIL"] = self.old


def _run(ctx, dim, p, outer, tail, prec):
    """a fresh system: MINRES, then (FP64 multigrid) an estimate with the slab solve, whose
    preconditioner is a second (FP32) hierarchy built in the same block"""
    m = meshes.uniform_mesh(dim, outer)
    coeff = meshes.smooth_initial(dim)
    if dim == 2:
        coeff = apf.Coefficients(meshes.checkerboard_kappa(m, 4, seed=11), 0.0, coeff.u0)
    sp = apf.Space(m, p, coarse=meshes.uniform_hierarchy(dim, outer))
    sysm = apf.System(ctx, sp)
    sysm.assemble(coeff)
    with _tail_env(tail):
        x, st = sysm.minres(tol=1e-10, max_it=600, check_interval=10, precond=prec)
        est = sysm.estimate(apf.APF_FP64, 2, per_cell=False, slab=apf.APF_SLAB_SOLVE)
    return x, st, est


# (2, 2, 4): P2 on 32x32 squares, the tail starts below the fine level; (2, 2, 3) and
# (1, 1, 5): the whole V-cycle fits, the tail starts at the fine level
@pytest.mark.parametrize("prec", [apf.APF_FP64, apf.APF_FP32_REFINE], ids=["mg64", "mg32"])
@pytest.mark.parametrize("dim,p,outer", [(2, 2, 4), (2, 2, 3), (1, 1, 5)])
def test_fused_tail_matches_per_level_launches(ctx, dim, p, outer, prec):
    print(f"mg tail dim={dim} p={p} outer={outer} prec={prec}", flush=True)
    x_on, st_on, e_on = _run(ctx, dim, p, outer, True, prec)
    x_off, st_off, e_off = _run(ctx, dim, p, outer, False, prec)
    assert st_on["converged"] and st_off["converged"]
    assert st_on["iterations"] == st_off["iterations"]
    rel = np.linalg.norm(x_on - x_off) / np.linalg.norm(x_off)
    print(f"  iterations {st_on['iterations']}, rel diff {rel:.3e}, bitwise {np.array_equal(x_on, x_off)}",
          flush=True)
    assert rel <= 1e-12
    # slab reconstruction: PCG with the Y columns' V-cycle only
    assert e_on["slab_relative_residual"] <= 1e-11 and e_off["slab_relative_residual"] <= 1e-11
    assert e_on["slab_iterations"] == e_off["slab_iterations"]
    for key in ("flux2", "eta2", "y_est2"):
        assert abs(e_on[key] - e_off[key]) <= 1e-10 * abs(e_off[key]), key
