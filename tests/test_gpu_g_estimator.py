"""GPU parity of the G-estimator against the oracle (through the C ABI).

The published G values are on the initial meshes, which have no coarser mesh
for the multigrid of the slab solve; the oracle carries those anchors
(test_oracle_g_estimator.py) and the GPU path is held to the oracle on refined
meshes, with the same degree-p solution on both sides."""
import numpy as np
import pytest

import doubleadapparabolicfem_amd as apf
from doubleadapparabolicfem_amd import meshes
from oracle import assembly as OA, estimator as OE, g_estimator as OG

from gpu_helpers import oracle_setup

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return apf.Context(0)


def _systems(ctx, m, dim, outer, p, coeff):
    sp = apf.Space(m, p)
    sysp = apf.System(ctx, sp)
    sysp.assemble(coeff)
    sq = apf.Space(m, p + 1, coarse=meshes.uniform_hierarchy(dim, outer))
    sysq = apf.System(ctx, sq)
    sysq.assemble(coeff)
    return sp, sysp, sq, sysq


@pytest.mark.parametrize("kappa", [False, True], ids=["M=I", "checkerboard"])
@pytest.mark.parametrize("dim,p,outer", [(1, 1, 2), (1, 2, 2), (1, 3, 1), (2, 1, 1), (2, 2, 1), (2, 1, 2)])
def test_g_estimator_matches_oracle(ctx, dim, p, outer, kappa):
    """G^2 on the oracle's direct solution: the slab solve run to its FP64
    round-off floor (tol 1e-15) against the oracle's direct solve agrees to 1e-10
    relative."""
    if kappa and dim == 1:
        pytest.skip("the checkerboard coefficient is two-dimensional")
    m = meshes.uniform_mesh(dim, outer)
    coeff = meshes.smooth_initial(dim)
    if kappa:
        coeff = apf.Coefficients(meshes.checkerboard_kappa(m, 4, seed=3), 0.0, coeff.u0)
    print(f"g-estimator dim={dim} p={p} outer={outer} kappa={kappa}", flush=True)
    stm, od, prob = oracle_setup(m, p, coeff)
    x = OE.main_solve(OA.assemble(stm, od, prob))
    g2_ref, lG, _, _ = OG.g_estimator(stm, p, prob, x)
    _, sysp, _, sysq = _systems(ctx, m, dim, outer, p, coeff)
    sysp.set_solution(x)
    g = sysp.g_estimate(sysq, tol=1e-15, max_it=300)
    assert g["g_true_dofs"] == len(lG)
    assert g2_ref > 0
    assert abs(g["lambda_difference_squared"] - g2_ref) <= 1e-10 * g2_ref, (g, g2_ref)
    assert g["solver_relative_residual"] <= 1e-12, g
    assert g["solver_stop"] in (apf.APF_SLAB_STOP_TOL, apf.APF_SLAB_STOP_FLOOR), g
    assert g["lambda_difference"] == pytest.approx(np.sqrt(g["lambda_difference_squared"]), rel=1e-15)


def test_g_estimator_after_minres_and_repeat(ctx):
    """Full GPU pipeline (assembly, MINRES to 1e-10, G): close to the oracle on
    its direct solution, and a second call on the same inputs is bit-identical."""
    dim, p, outer = 2, 2, 2
    m = meshes.uniform_mesh(dim, outer)
    coeff = meshes.smooth_initial(dim)
    print(f"g-estimator after minres dim={dim} p={p} outer={outer}", flush=True)
    stm, od, prob = oracle_setup(m, p, coeff)
    x = OE.main_solve(OA.assemble(stm, od, prob))
    g2_ref = OG.g_estimator(stm, p, prob, x)[0]
    sp = apf.Space(m, p, coarse=meshes.uniform_hierarchy(dim, outer))
    sysp = apf.System(ctx, sp)
    sysp.assemble(coeff)
    xs, _ = sysp.minres(tol=1e-10, max_it=600, check_interval=10)
    sq = apf.Space(m, p + 1, coarse=meshes.uniform_hierarchy(dim, outer))
    sysq = apf.System(ctx, sq)
    sysq.assemble(coeff)
    a = sysp.g_estimate(sysq)
    b = sysp.g_estimate(sysq)
    assert a["lambda_difference_squared"] == b["lambda_difference_squared"]
    assert abs(a["lambda_difference_squared"] - g2_ref) <= 1e-6 * g2_ref, (a, g2_ref)
    # another degree-p system on the same mesh with the same solution: the G system's
    # interpolation, built on its first call, gives the same value bit for bit
    sysp2 = apf.System(ctx, apf.Space(m, p))
    sysp2.assemble(coeff)
    sysp2.set_solution(xs)
    c = sysp2.g_estimate(sysq)
    assert c["lambda_difference_squared"] == a["lambda_difference_squared"]


def test_g_estimator_rejects_mismatched_systems(ctx):
    m = meshes.uniform_mesh(2, 1)
    coeff = meshes.smooth_initial(2)
    print("g-estimator argument checks", flush=True)
    sp = apf.Space(m, 1)
    sysp = apf.System(ctx, sp)
    sysp.assemble(coeff)
    same = apf.System(ctx, apf.Space(m, 1, coarse=meshes.uniform_hierarchy(2, 1)))
    same.assemble(coeff)
    with pytest.raises(apf.APFError, match="degree p\\+1"):
        sysp.g_estimate(same)
    other = apf.System(ctx, apf.Space(meshes.uniform_mesh(2, 2), 2, coarse=meshes.uniform_hierarchy(2, 2)))
    other.assemble(coeff)
    with pytest.raises(apf.APFError, match="different meshes"):
        sysp.g_estimate(other)
    unassembled = apf.System(ctx, apf.Space(m, 2, coarse=meshes.uniform_hierarchy(2, 1)))
    with pytest.raises(apf.APFError, match="assemble"):
        sysp.g_estimate(unassembled)


def test_g_estimator_with_more_multigrid_tiles_than_default_partials(ctx):
    """1+1D p=1 on 2^19 cells x 2 intervals: the degree-2 G system has 1,048,575
    free nodes, i.e. 16,384 multigrid transpose tiles of 64 rows, each writing one
    dot partial, four times the 4 x 1024 partials a system starts with.  The
    partial buffer grows with the multigrid at setup: the slab solve converges,
    G is finite and positive, and a repeated call is bit-identical."""
    n = 2 ** 19
    print(f"g-estimator with {n} cells (large multigrid tile count)", flush=True)
    m = meshes.interval_mesh(n, [0.0, 0.5, 1.0])
    coarse = [meshes.interval_mesh(2 ** j, [0.0, 1.0]) for j in range(18, 0, -1)]
    coeff = meshes.smooth_initial(1)
    sp_ = apf.Space(m, 1, coarse=coarse)
    sysp = apf.System(ctx, sp_)
    sysp.assemble(coeff)
    # any degree-p coefficient vector serves (the cells are 2e-6 x 0.5, too thin for the
    # multigrid MINRES to reach 1e-10 in a test's time)
    sysp.set_solution(np.random.default_rng(5).standard_normal(sp_.n))
    sq = apf.Space(m, 2, coarse=coarse)
    assert -(-sq.nf // 64) > 4 * 1024
    sysq = apf.System(ctx, sq)
    sysq.assemble(coeff)
    a = sysp.g_estimate(sysq, tol=1e-12, max_it=300)
    b = sysp.g_estimate(sysq, tol=1e-12, max_it=300)
    assert np.isfinite(a["lambda_difference_squared"]) and a["lambda_difference_squared"] > 0
    assert a["solver_stop"] in (apf.APF_SLAB_STOP_TOL, apf.APF_SLAB_STOP_FLOOR), a
    assert a["lambda_difference_squared"] == b["lambda_difference_squared"]
