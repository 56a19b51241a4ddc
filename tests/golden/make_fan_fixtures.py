"""Oracle results on fan-patch meshes (vertex patches of up to 64 cells), for
the GPU parity test of the large-patch estimator path.

Each case is a fan mesh (q x q squares, side splits uniform in smin..smax,
seeded), a time grid graded toward t = 0, u0 = indicator of [0.25,0.75]^2,
M = I, ell = 0.  The oracle assembles the space-time system, solves it
directly and runs the estimator; the fixture holds the case parameters, the
solution and the per-cell estimator values.  The oracle is this repository's
own CPU restatement; no reference code is run.

Usage:  python tests/golden/make_fan_fixtures.py [case ...]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

CASES = {
    # name: (p, q, seed, smin, smax, lmin, lmax)
    "fan_p1": (1, 2, 3, 1, 16, 1, 3),
    "fan_p2": (2, 1, 0, 16, 16, 1, 3),
    "fan_p3": (3, 1, 0, 16, 16, 1, 3),
}


def compute(name):
    from doubleadapparabolicfem_amd import meshes
    from oracle import assembly as OA, estimator as OE, mesh as OM
    p, q, seed, smin, smax, lmin, lmax = CASES[name]
    m = meshes.fan_mesh(q, meshes.graded_times(lmin, lmax), seed, smin, smax)
    sm = OM.SpatialMesh(2, m.vertices.astype(float), m.cells.astype(np.int64))
    stm = OM.SpaceTimeMesh(sm, m.times)
    od = OM.spacetime_dofs(stm, p)
    prob = OA.Problem(OA.identity_M(sm), 0.0, meshes.indicator_initial().u0)
    S = OA.assemble(stm, od, prob)
    x = OE.main_solve(S)
    ref = OE.estimate(stm, od, prob, S, x)
    return dict(params=np.array(CASES[name], dtype=np.int64), x=x, flux_by_cell=ref.flux_by_cell,
                recon_by_cell=ref.recon_by_cell, flux2=ref.flux2, lambda_y2=ref.lambda_y2)


def main(names):
    for name in names:
        out = compute(name)
        dst = os.path.join(HERE, name + ".npz")
        np.savez_compressed(dst, **out)
        print(f"wrote {dst}: n = {out['x'].size}, flux2 = {out['flux2']:.6e}")


if __name__ == "__main__":
    main(sys.argv[1:] or list(CASES))
