"""Uniform space-time meshes of the published runs (oracle; test infrastructure only).

Uniform outer iteration k of the reference's `uniform_standard` runs
(algorithm/examples/*/smooth_initial*; one uniform level from the roots,
algorithm/examples/space_time_2d/initial_data_2d_examples.hpp:70-75):
* 1+1D: 2^(k+1) x 2^(k+1) rectangles on (0,1)^2;
* 2+1D: k = 0 -> the unit square cut by both diagonals (4 triangles), 1 interval;
        k >= 1 -> union-jack 2^(k+1) x 2^(k+1) squares, 2^k intervals.
Pinned by the published X/Y true-DOF counts.
"""
import numpy as np

from . import mesh as M
from .assembly import Problem, identity_M


def uniform_mesh(dim: int, outer: int) -> M.SpaceTimeMesh:
    if dim == 1:
        n = 2 ** (outer + 1)
        return M.SpaceTimeMesh(M.interval_mesh(n), np.linspace(0.0, 1.0, n + 1))
    if outer == 0:
        return M.SpaceTimeMesh(M.criss_cross_mesh(1), np.linspace(0.0, 1.0, 2))
    n = 2 ** (outer + 1)
    m = 2 ** outer
    return M.SpaceTimeMesh(M.union_jack_mesh(n), np.linspace(0.0, 1.0, m + 1))


def smooth_initial(stm: M.SpaceTimeMesh) -> Problem:
    """M = I, ell = 0, u0 = prod sin(pi x_i)
    (algorithm/examples/space_time_1d/initial_data_1d_examples.hpp:9-63,
     algorithm/examples/space_time_2d/initial_data_2d_examples.hpp:15-80)."""
    if stm.space.dim == 1:
        u0 = lambda x: np.sin(np.pi * x[:, 0])
    else:
        u0 = lambda x: np.sin(np.pi * x[:, 0]) * np.sin(np.pi * x[:, 1])
    return Problem(identity_M(stm.space), 0.0, u0)
