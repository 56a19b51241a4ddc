"""Space-time meshes that are not tensor products: layers graded toward t = 0.

The reference refines the cells that touch t = 0 by its generation split policy
(src/mesh/refinement/split_policy.hpp:8-26: a spatial bisection, then a
space-time split, alternately).  Starting from a uniform union-jack mesh, the
cells with t0 = 0 refined twice are union-jack cells of half the size over the
lower half of their time interval; repeated, this gives time layers
[t_k, t_k+1] whose spatial meshes are uniform union-jack (1+1D: uniform
interval) meshes that get finer toward t = 0.  Where a fine layer meets a
coarser one, X (continuous in space and time) has hanging nodes: the fine
side's interface nodes that are not nodes of the coarse mesh take the value of
the coarse side's function, as the reference's constrained DoFs do
(src/finite_element/assembly/constrained_dofs.hpp:11-226).  Y (continuous in
space, discontinuous in time) can be refined further than X, in space and in
time; each Y cell then lies in one X cell, its X ancestor
(assembly/detail/active_cell_locator.hpp).

The output is what the reference's FESpace hands to its assembler for each cell:
the cell geometry and the cell restriction rows u_cell = C_c u_true
(constrained_dofs.hpp:36-56): a free local DoF has the row [(true dof, 1.0)],
a Dirichlet one an empty row, a hanging one its masters with weights, ordered
by true dof.  Local DoF index i_space * (p+1) + i_time
(basis/functions/triangular_prism_lagrange.hpp:32-35), spatial nodes in the
equidistant order (nodes/triangle_equidistant.hpp:39-53).  The true-DoF
numbering is this module's own (time level by time level, nodes by first
appearance), not the reference's.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

TOL_W = 1e-14  # hanging-node weights below this (the coarse basis vanishes there) are dropped


def _tuples(dim, p):
    if dim == 1:
        return [(p - b, b) for b in range(p + 1)]
    return [(p - b - c, b, c) for c in range(p + 1) for b in range(p + 1 - c)]


def _factor(p, a, s):
    val = np.ones_like(s)
    for m in range(a):
        val = val * (p * s - m) / (m + 1.0)
    return val


def spatial_basis(dim, p, xi):
    """P_p Lagrange basis on the reference cell at points xi [n, dim] -> [n, Ns]
    (closed form: products of barycentric factors prod_{m<a} (p lam - m)/(m+1))."""
    xi = np.atleast_2d(xi)
    if dim == 1:
        lam = [1.0 - xi[:, 0], xi[:, 0]]
    else:
        lam = [1.0 - xi[:, 0] - xi[:, 1], xi[:, 0], xi[:, 1]]
    out = []
    for t in _tuples(dim, p):
        v = np.ones(len(xi))
        for a, l in zip(t, lam):
            v = v * _factor(p, a, l)
        out.append(v)
    return np.stack(out, axis=1)


def segment_basis(p, s):
    """Equidistant Lagrange basis on [0, 1] at s [n] -> [n, p+1]."""
    s = np.atleast_1d(s)
    nodes = np.arange(p + 1) / p
    out = np.ones((len(s), p + 1))
    for i in range(p + 1):
        for j in range(p + 1):
            if j != i:
                out[:, i] *= (s - nodes[j]) / (nodes[i] - nodes[j])
    return out


def uniform_spatial(dim, n):
    """Vertices and cells of the uniform mesh with n cells per direction: intervals,
    or union-jack triangles (square (i, j) cut through its corner with even i + j,
    counter-clockwise), in the order of meshes.union_jack_mesh."""
    if dim == 1:
        return (np.arange(n + 1) / n)[:, None], np.stack([np.arange(n), np.arange(1, n + 1)], axis=1)
    j, i = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    verts = np.stack([i.ravel() / n, j.ravel() / n], axis=1)
    jj, ii = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    v = lambda a, b: b * (n + 1) + a
    even = (ii + jj) % 2 == 0
    t0 = np.where(even[:, None], np.stack([v(ii, jj), v(ii + 1, jj), v(ii + 1, jj + 1)], 1),
                  np.stack([v(ii, jj), v(ii + 1, jj), v(ii, jj + 1)], 1))
    t1 = np.where(even[:, None], np.stack([v(ii, jj), v(ii + 1, jj + 1), v(ii, jj + 1)], 1),
                  np.stack([v(ii + 1, jj), v(ii + 1, jj + 1), v(ii, jj + 1)], 1))
    cells = np.stack([t0, t1], axis=1).reshape(-1, 3)
    return verts, cells


def locate(dim, n, x):
    """Cell of the uniform mesh (n per direction) that contains each point x [k, dim],
    and the point's reference coordinates in it."""
    verts, cells = uniform_spatial(dim, n)
    x = np.atleast_2d(x)
    if dim == 1:
        c = np.minimum((x[:, 0] * n).astype(np.int64), n - 1)
        xi = (x[:, 0] - verts[cells[c, 0], 0]) * n
        return c, xi[:, None]
    i = np.minimum((x[:, 0] * n).astype(np.int64), n - 1)
    j = np.minimum((x[:, 1] * n).astype(np.int64), n - 1)
    best_c = np.zeros(len(x), dtype=np.int64)
    best_xi = np.zeros((len(x), 2))
    best_m = np.full(len(x), -np.inf)
    for h in (0, 1):
        c = 2 * (j * n + i) + h
        V = verts[cells[c]]
        J = np.stack([V[:, 1] - V[:, 0], V[:, 2] - V[:, 0]], axis=2)  # columns
        xi = np.linalg.solve(J, (x - V[:, 0])[:, :, None])[:, :, 0]
        m = np.minimum(np.minimum(xi[:, 0], xi[:, 1]), 1.0 - xi[:, 0] - xi[:, 1])
        better = m > best_m
        best_c = np.where(better, c, best_c)
        best_xi = np.where(better[:, None], xi, best_xi)
        best_m = np.where(better, m, best_m)
    return best_c, best_xi


@dataclass
class CellSpace:
    """Cells and restriction rows of one space (the arrays of apf_cell_space_desc)."""
    dim: int
    p: int
    vertices: np.ndarray    # [nv, dim]
    cells: np.ndarray       # [nc, dim+1] int32
    times: np.ndarray       # [nc, 2]
    restr_ptr: np.ndarray   # [nc * n_loc + 1] int64
    restr_dof: np.ndarray   # int32
    restr_w: np.ndarray     # float64
    n_dofs: int
    layer: np.ndarray       # [nc] layer (X) or sub-interval (Y) of each cell
    res: np.ndarray         # [nc] cells per direction of the cell's spatial mesh
    local_cell: np.ndarray  # [nc] index of the cell in its spatial mesh

    @property
    def n_cells(self):
        return len(self.cells)

    def rows(self, c, i):
        """true dofs and weights of local dof i of cell c"""
        a, b = self.restr_ptr[c * self.n_loc + i], self.restr_ptr[c * self.n_loc + i + 1]
        return self.restr_dof[a:b], self.restr_w[a:b]

    @property
    def n_loc(self):
        Ns = self.p + 1 if self.dim == 1 else (self.p + 1) * (self.p + 2) // 2
        return Ns * (self.p + 1)


@dataclass
class LayeredMesh:
    dim: int
    p: int
    x_layers: list          # [(t0, t1, n)] bottom to top
    y_layers: list          # [(t0, t1, n, x_layer)] bottom to top
    X: CellSpace
    Y: CellSpace
    y_ancestor: np.ndarray  # [nY cells] int32
    x_bottom: np.ndarray    # X cells with t0 = 0

    @property
    def n(self):
        return self.X.n_dofs + self.Y.n_dofs


class _MeshNodes:
    """P_p nodes of a uniform mesh: per (cell, local node) an integer key of the node's
    position on the grid of spacing 1 / (N p), and the free-node numbering by first
    appearance (boundary nodes excluded)."""

    def __init__(self, dim, p, n, N):
        self.dim, self.p, self.n = dim, p, n
        verts, cells = uniform_spatial(dim, n)
        self.verts, self.cells = verts, cells
        T = np.array(_tuples(dim, p), dtype=float) / p                 # [Ns, dim+1]
        X = np.einsum("sv,cvd->csd", T, verts[cells])                   # [nc, Ns, dim]
        self.coords = X
        G = N * p
        ik = np.rint(X * G).astype(np.int64)
        self.key = ik[..., 0] if dim == 1 else ik[..., 0] * (G + 1) + ik[..., 1]
        self.G = G
        bd = np.any((ik == 0) | (ik == G), axis=-1)
        flat = self.key.ravel()
        uk, first = np.unique(flat, return_index=True)
        order = np.argsort(first, kind="stable")
        self.boundary_of = dict(zip(flat[first], bd.ravel()[first]))
        free_keys = [k for k in flat[np.sort(first)] if not self.boundary_of[k]]
        self.free_keys = np.array(free_keys, dtype=np.int64)
        self.sorted_keys = np.sort(self.free_keys)
        self.sorted_ids = np.argsort(self.free_keys, kind="stable")
        self.all_keys = uk  # sorted, boundary included
        del order

    @property
    def n_free(self):
        return len(self.free_keys)

    def free_id(self, keys):
        """free id of each key, -1 for a boundary node, -2 for a key not in this mesh"""
        keys = np.asarray(keys)
        pos = np.searchsorted(self.sorted_keys, keys)
        pos_c = np.minimum(pos, len(self.sorted_keys) - 1)
        hit = (len(self.sorted_keys) > 0) & (self.sorted_keys[pos_c] == keys)
        out = np.where(hit, self.sorted_ids[pos_c], -2)
        pos2 = np.minimum(np.searchsorted(self.all_keys, keys), len(self.all_keys) - 1)
        in_mesh = self.all_keys[pos2] == keys
        return np.where(hit, out, np.where(in_mesh, -1, -2))


def layered_mesh(dim, p, x_layers, y_refine=None):
    """x_layers: [(t0, t1, n)] bottom to top, contiguous, n cells per direction (powers
    of two times a common base, so the meshes of neighbouring layers are nested).
    y_refine: per X layer (r, s): the Y mesh of the layer has n 2^r cells per
    direction and s sub-intervals (default (0, 1): Y on the X cells)."""
    m = len(x_layers)
    if y_refine is None:
        y_refine = [(0, 1)] * m
    for k in range(m - 1):
        assert abs(x_layers[k][1] - x_layers[k + 1][0]) < 1e-15
    ns_all = [L[2] for L in x_layers] + [L[2] * 2 ** r for L, (r, _) in zip(x_layers, y_refine)]
    N = max(ns_all)
    nodes = {}

    def mesh_nodes(n):
        if n not in nodes:
            nodes[n] = _MeshNodes(dim, p, n, N)
        return nodes[n]

    Ns = p + 1 if dim == 1 else (p + 1) * (p + 2) // 2
    Nt = p + 1
    n_loc = Ns * Nt

    # ---- X: time levels, the mesh that owns each level, and the level's true dofs
    levels = []        # (t, owner resolution)
    for k, (t0, t1, n) in enumerate(x_layers):
        for i in range(p):
            t = t0 + (t1 - t0) * i / p
            if i == 0 and k > 0:
                levels.append((t, min(n, x_layers[k - 1][2])))
            else:
                levels.append((t, n))
    levels.append((x_layers[-1][1], x_layers[-1][2]))
    level_base = []
    nX = 0
    for t, n in levels:
        level_base.append(nX)
        nX += mesh_nodes(n).n_free

    xv, xc, xt, xl, xres, xloc = [], [], [], [], [], []
    rows = []   # per (cell, local): list of (dof, w)
    voff = 0
    for k, (t0, t1, n) in enumerate(x_layers):
        mn = mesh_nodes(n)
        nc = len(mn.cells)
        xv.append(mn.verts)
        xc.append(mn.cells + voff)
        voff += len(mn.verts)
        xt.append(np.tile([t0, t1], (nc, 1)))
        xl.append(np.full(nc, k))
        xres.append(np.full(nc, n))
        xloc.append(np.arange(nc))
        r = [[None] * n_loc for _ in range(nc)]
        for it in range(Nt):
            lev = k * p + it
            own = mesh_nodes(levels[lev][1])
            fid = own.free_id(mn.key)          # [nc, Ns]
            base = level_base[lev]
            hang = np.argwhere(fid == -2)
            if len(hang):
                pts = mn.coords[hang[:, 0], hang[:, 1]]
                cc, xi = locate(dim, own.n, pts)
                w = spatial_basis(dim, p, xi)  # [h, Ns]
                mkeys = own.key[cc]            # [h, Ns]
                mfid = own.free_id(mkeys)
            for c in range(nc):
                for s in range(Ns):
                    f = fid[c, s]
                    if f >= 0:
                        r[c][s * Nt + it] = [(base + int(f), 1.0)]
                    elif f == -1:
                        r[c][s * Nt + it] = []
            for h, (c, s) in enumerate(hang):
                acc = {}
                for j in range(Ns):
                    if mfid[h, j] >= 0 and abs(w[h, j]) > TOL_W:
                        d = base + int(mfid[h, j])
                        acc[d] = acc.get(d, 0.0) + w[h, j]
                r[c][s * Nt + it] = sorted(acc.items())
        rows.extend(row for cell in r for row in cell)
    X = _cell_space(dim, p, xv, xc, xt, xl, xres, xloc, rows, nX)

    # ---- Y: sub-intervals of each layer, time-discontinuous, no hanging nodes
    yv, yc, yt, yl, yres, yloc, anc = [], [], [], [], [], [], []
    rows = []
    y_layers = []
    nY = 0
    voff = 0
    x_cell_base = np.concatenate([[0], np.cumsum([len(mesh_nodes(L[2]).cells) for L in x_layers])])
    for k, ((t0, t1, n), (rr, sdiv)) in enumerate(zip(x_layers, y_refine)):
        ny = n * 2 ** rr
        mn = mesh_nodes(ny)
        nc = len(mn.cells)
        # X ancestor of each Y cell: the layer's X cell containing its centroid
        cen = mn.coords.mean(axis=1)
        ac, _ = locate(dim, n, cen)
        for j in range(sdiv):
            a = t0 + (t1 - t0) * j / sdiv
            b = t0 + (t1 - t0) * (j + 1) / sdiv
            y_layers.append((a, b, ny, k))
            yv.append(mn.verts)
            yc.append(mn.cells + voff)
            voff += len(mn.verts)
            yt.append(np.tile([a, b], (nc, 1)))
            yl.append(np.full(nc, len(y_layers) - 1))
            yres.append(np.full(nc, ny))
            yloc.append(np.arange(nc))
            anc.append(x_cell_base[k] + ac)
            fid = mn.free_id(mn.key)
            nf = mn.n_free
            r = [[None] * n_loc for _ in range(nc)]
            for c in range(nc):
                for s in range(Ns):
                    for it in range(Nt):
                        f = fid[c, s]
                        r[c][s * Nt + it] = [(nY + it * nf + int(f), 1.0)] if f >= 0 else []
            nY += Nt * nf
            rows.extend(row for cell in r for row in cell)
    Y = _cell_space(dim, p, yv, yc, yt, yl, yres, yloc, rows, nY)
    x_bottom = np.nonzero(X.times[:, 0] == 0.0)[0].astype(np.int32)
    return LayeredMesh(dim, p, list(x_layers), y_layers, X, Y, np.concatenate(anc).astype(np.int32), x_bottom)


def _cell_space(dim, p, v, c, t, l, res, loc, rows, n_dofs):
    ptr = np.zeros(len(rows) + 1, dtype=np.int64)
    ptr[1:] = np.cumsum([len(r) for r in rows])
    dof = np.array([d for r in rows for d, _ in r], dtype=np.int32)
    w = np.array([x for r in rows for _, x in r], dtype=np.float64)
    return CellSpace(dim, p, np.concatenate(v), np.concatenate(c).astype(np.int32), np.concatenate(t),
                     ptr, dof, w, int(n_dofs), np.concatenate(l), np.concatenate(res), np.concatenate(loc))


def graded_layers(n_top, levels):
    """X layers of the mesh that `levels` double rounds of the reference's split policy
    give when every round refines the cells with t0 = 0 (a spatial bisection, then a
    space-time split; two bisections of a union-jack mesh give the union-jack mesh of
    half the size), starting from n_top cells per direction on [0, 1]:
    [1/2, 1] at 2 n_top, [1/4, 1/2] at 4 n_top, ..., [2^-L, 2^-(L-1)] and [0, 2^-L] at
    2^L n_top (L = levels), bottom to top."""
    layers = [(0.0, 2.0 ** -levels, n_top * 2 ** levels)]
    for j in range(levels, 0, -1):
        layers.append((2.0 ** -j, 2.0 ** -(j - 1), n_top * 2 ** j))
    return layers


def slab_groups(lm):
    """Runs of consecutive Y sub-intervals with one spatial mesh: per group the number
    of cells per direction, the time nodes and the first Y cell (the Y cells of a group
    are its mesh's cells, sub-interval by sub-interval)."""
    groups = []
    first = 0
    ncell = lambda n: n if lm.dim == 1 else 2 * n * n
    for (a, b, n, _) in lm.y_layers:
        if groups and groups[-1]["n"] == n:
            groups[-1]["times"].append(b)
        else:
            groups.append({"n": n, "times": [a, b], "first_y_cell": first})
        first += ncell(n)
    return groups
