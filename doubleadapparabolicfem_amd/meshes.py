"""Structured tensor space-time meshes and problem data of the hot-path workloads.

These are the active meshes the reference reaches by uniform refinement
(algorithm/support/space_time_2d.hpp:91-114 roots, NVB bisection with the
union-jack outcome; algorithm/support/space_time_1d.hpp:14-24) and the
problem data of algorithm/examples/space_time_{1d,2d}/initial_data_*_examples.hpp.
"""
from __future__ import annotations

import numpy as np

from . import Coefficients, Mesh


def interval_mesh(n: int, times) -> Mesh:
    x = np.arange(n + 1) / n
    cells = np.stack([np.arange(n), np.arange(1, n + 1)], axis=1)
    return Mesh(1, x[:, None], cells.astype(np.int32), np.asarray(times, dtype=np.float64))


def union_jack_mesh(n: int, times) -> Mesh:
    """n x n squares of (0,1)^2; square (i,j) is cut along the diagonal through
    its corner with even i+j (counter-clockwise triangles)."""
    j, i = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    verts = np.stack([i.ravel() / n, j.ravel() / n], axis=1)
    vid = lambda a, b: b * (n + 1) + a
    cells = np.zeros((2 * n * n, 3), dtype=np.int32)
    t = 0
    for jj in range(n):
        for ii in range(n):
            if (ii + jj) % 2 == 0:
                cells[t] = (vid(ii, jj), vid(ii + 1, jj), vid(ii + 1, jj + 1))
                cells[t + 1] = (vid(ii, jj), vid(ii + 1, jj + 1), vid(ii, jj + 1))
            else:
                cells[t] = (vid(ii, jj), vid(ii + 1, jj), vid(ii, jj + 1))
                cells[t + 1] = (vid(ii + 1, jj), vid(ii + 1, jj + 1), vid(ii, jj + 1))
            t += 2
    return Mesh(2, verts, cells, np.asarray(times, dtype=np.float64))


def criss_cross_mesh(n: int, times) -> Mesh:
    """n x n squares each cut by both diagonals (4 triangles around a centre vertex)."""
    vid = lambda a, b: b * (n + 1) + a
    verts = [(a / n, b / n) for b in range(n + 1) for a in range(n + 1)]
    nv0 = len(verts)
    cells = []
    for jj in range(n):
        for ii in range(n):
            c = nv0 + jj * n + ii
            verts.append(((ii + 0.5) / n, (jj + 0.5) / n))
            a, b, cc, d = vid(ii, jj), vid(ii + 1, jj), vid(ii + 1, jj + 1), vid(ii, jj + 1)
            cells += [(a, b, c), (b, cc, c), (cc, d, c), (d, a, c)]
    return Mesh(2, np.array(verts), np.array(cells, dtype=np.int32), np.asarray(times, dtype=np.float64))


def uniform_mesh(dim: int, outer: int) -> Mesh:
    """Active mesh of uniform outer iteration `outer` of the reference's
    uniform_standard runs (one uniform level from the roots):
    1+1D: 2^(outer+1) rectangles per direction;
    2+1D: outer 0 -> unit square cut by both diagonals, one interval;
          outer >= 1 -> union-jack 2^(outer+1) squares per side, 2^outer intervals."""
    if dim == 1:
        n = 2 ** (outer + 1)
        return interval_mesh(n, np.linspace(0.0, 1.0, n + 1))
    if outer == 0:
        return criss_cross_mesh(1, np.array([0.0, 1.0]))
    n = 2 ** (outer + 1)
    return union_jack_mesh(n, np.linspace(0.0, 1.0, 2 ** outer + 1))


def uniform_hierarchy(dim: int, outer: int):
    """Nested coarser spatial meshes of uniform_mesh(dim, outer), fine to coarse
    (for the multigrid preconditioner): halve the resolution down to 2 cells
    per direction."""
    out = []
    if dim == 1:
        n = 2 ** (outer + 1) // 2
        while n >= 2:
            out.append(interval_mesh(n, [0.0, 1.0]))
            n //= 2
        return out
    if outer == 0:
        return out
    n = 2 ** (outer + 1) // 2
    while n >= 2:
        out.append(union_jack_mesh(n, [0.0, 1.0]))
        n //= 2
    return out


def smooth_initial(dim: int) -> Coefficients:
    """M = I, ell = 0, u0 = prod_i sin(pi x_i) (smooth_initial examples)."""
    if dim == 1:
        return Coefficients(None, 0.0, lambda x: np.sin(np.pi * x[:, 0]))
    return Coefficients(None, 0.0, lambda x: np.sin(np.pi * x[:, 0]) * np.sin(np.pi * x[:, 1]))


def checkerboard_kappa(mesh: Mesh, k: int = 8, seed: int = 4) -> np.ndarray:
    """M = kappa(x) I, kappa piecewise constant on a k x k spatial checkerboard,
    kappa = 10^U(-1,1) per square (std::mt19937_64-independent: numpy PCG64 with `seed`)."""
    rng = np.random.default_rng(seed)
    vals = 10.0 ** rng.uniform(-1.0, 1.0, size=(k, k))
    cen = mesh.vertices[mesh.cells].mean(axis=1)
    ix = np.minimum((cen[:, 0] * k).astype(int), k - 1)
    iy = np.minimum((cen[:, 1] * k).astype(int), k - 1)
    kap = vals[iy, ix]
    d = mesh.dim
    return kap[:, None, None] * np.eye(d)[None, :, :]


def segment_batch_mesh(nx, times) -> Mesh:
    """Independent 1+1D problems in one mesh: problem i lives on the segment
    [2i, 2i+1] cut into nx[i] equal intervals; segments share no vertex, so the
    problems decouple and each segment's ends are its Dirichlet boundary."""
    verts, cells, off = [], [], 0
    for i, n in enumerate(nx):
        verts.append(2.0 * i + np.arange(n + 1) / n)
        cells.append(off + np.stack([np.arange(n), np.arange(1, n + 1)], axis=1))
        off += n + 1
    return Mesh(1, np.concatenate(verts)[:, None], np.concatenate(cells).astype(np.int32),
                np.asarray(times, dtype=np.float64))


def segment_batch_hierarchy(nx, levels):
    """Nested coarser batch meshes (every segment halved `levels` times)."""
    out = []
    for lev in range(1, levels + 1):
        out.append(segment_batch_mesh([n >> lev for n in nx], [0.0, 1.0]))
    return out


def segment_batch_problems(n_problems=256, n_t=32, seed0=5000, dof_range=(1e3, 1e5), multiple=8):
    """Synthetic batch of independent 1+1D heat problems: per problem i (numpy
    PCG64 seeded with seed0 + i) a target X-DOF count log-uniform in dof_range,
    kappa = 10^U(-1,1), u0 = sum_{k=1..5} c_k sin(k pi x) with c_k ~ N(0,1),
    ell = 0; all share the uniform time grid of n_t intervals.  Returns
    (nx per problem, kappa per problem, coefficients [n_problems, 5])."""
    nx, kap, coef = [], [], []
    for i in range(n_problems):
        rng = np.random.default_rng(seed0 + i)
        target = np.exp(rng.uniform(np.log(dof_range[0]), np.log(dof_range[1])))
        n = max(multiple, int(round((target / (n_t + 1) + 1) / multiple)) * multiple)
        nx.append(n)
        kap.append(10.0 ** rng.uniform(-1.0, 1.0))
        coef.append(rng.standard_normal(5))
    return nx, np.array(kap), np.array(coef)


def segment_batch_coefficients(mesh: Mesh, nx, kappa, coef) -> Coefficients:
    """M = kappa_i per cell of segment i; u0 = sum_k c_ik sin(k pi (x - 2i))."""
    seg_of_cell = np.repeat(np.arange(len(nx)), nx)
    M = kappa[seg_of_cell][:, None, None] * np.ones((1, 1, 1))

    def u0(x):
        i = np.floor(x[:, 0] / 2.0).astype(int)
        xl = x[:, 0] - 2.0 * i
        k = np.arange(1, coef.shape[1] + 1)
        return np.sum(coef[i] * np.sin(np.pi * xl[:, None] * k[None, :]), axis=1)

    return Coefficients(M, 0.0, u0)



def fan_mesh(q: int, times, seed: int = 3, smin: int = 1, smax: int = 16) -> Mesh:
    """Fan-patch mesh of (0,1)^2: q x q squares, every square side split into s
    equal segments (s uniform in smin..smax per side, numpy PCG64 with `seed`),
    and every square a fan of triangles around its centre vertex.  The centre's
    vertex patch has K = (sum of its four side counts) cells, 4..64 at the
    default range; the side and corner patches have 4..8.  The mesh is
    conforming: both squares of a shared side see the same split."""
    rng = np.random.default_rng(seed)
    sh = rng.integers(smin, smax + 1, size=(q + 1, q))  # side y = j/q, x in [i/q, (i+1)/q]
    sv = rng.integers(smin, smax + 1, size=(q + 1, q))  # side x = i/q, y in [j/q, (j+1)/q]
    verts = [(i / q, j / q) for j in range(q + 1) for i in range(q + 1)]
    corner = lambda i, j: j * (q + 1) + i
    hp, vp = {}, {}
    for j in range(q + 1):
        for i in range(q):
            s = sh[j, i]
            hp[j, i] = list(range(len(verts), len(verts) + s - 1))
            verts += [((i + k / s) / q, j / q) for k in range(1, s)]
    for i in range(q + 1):
        for j in range(q):
            s = sv[i, j]
            vp[i, j] = list(range(len(verts), len(verts) + s - 1))
            verts += [(i / q, (j + k / s) / q) for k in range(1, s)]
    cells = []
    for j in range(q):
        for i in range(q):
            c = len(verts)
            verts.append(((i + 0.5) / q, (j + 0.5) / q))
            ring = ([corner(i, j)] + hp[j, i] + [corner(i + 1, j)] + vp[i + 1, j] + [corner(i + 1, j + 1)]
                    + hp[j + 1, i][::-1] + [corner(i, j + 1)] + vp[i, j][::-1])
            cells += [(ring[k], ring[(k + 1) % len(ring)], c) for k in range(len(ring))]
    return Mesh(2, np.array(verts), np.array(cells, dtype=np.int32), np.asarray(times, dtype=np.float64))


def graded_times(lmin: int = 2, lmax: int = 10) -> np.ndarray:
    """Time grid of [0,1] graded toward t = 0: uniform intervals of length
    2^-lmin, the first one bisected toward 0 down to length 2^-lmax."""
    head = [0.0] + [2.0 ** -l for l in range(lmax, lmin - 1, -1)]
    tail = [k / 2 ** lmin for k in range(2, 2 ** lmin + 1)]
    return np.array(head + tail)


def indicator_initial() -> Coefficients:
    """M = I, ell = 0, u0 = indicator of [0.25, 0.75]^2."""
    def u0(x):
        inside = (x[:, 0] >= 0.25) & (x[:, 0] <= 0.75) & (x[:, 1] >= 0.25) & (x[:, 1] <= 0.75)
        return inside.astype(np.float64)
    return Coefficients(None, 0.0, u0)


def fan_hierarchy(q: int):
    """Nested coarser meshes of fan_mesh(q, ...) for the multigrid
    preconditioner, fine to coarse: every fan triangle lies in the triangle
    (two corners of its square side, centre) of criss_cross_mesh(q), which lies
    in a half of union_jack_mesh(q); then union-jack halvings down to 2."""
    out = [criss_cross_mesh(q, [0.0, 1.0])]
    n = q
    while n >= 2:
        out.append(union_jack_mesh(n, [0.0, 1.0]))
        n //= 2
    return out
