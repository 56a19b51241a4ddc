"""Tensor space-time meshes and Lagrange DOF numbering (oracle; test infrastructure only).

The reference stores an adaptive prism/rectangle tree (src/mesh/mesh.hpp) and
numbers DOFs on the active antichain (src/finite_element/fespace/dofs/
dof_distribution_2d.hpp).  For the hot-path parity cases the active meshes are
tensor products  T_h x {t_0 < ... < t_m}  (uniform union-jack / criss-cross
triangulations in 2+1D, uniform intervals in 1+1D), so this module builds them
directly.  The DOF numbering below is this build's own; it is a permutation of
the reference's, so matrices are compared entrywise after mapping and all
estimator quantities are numbering-independent.  Counts are pinned by the
published X/Y true-DOF columns of algorithm_data_plot/{1d,2d}/smooth_initial.csv.

Spaces (src/finite_element/fespace/policies.hpp:9-28):
* X: continuous in space and time, P_p(space) x P_p(time), zero Dirichlet data
  on the spatial boundary;
* Y: continuous in space, discontinuous in time, same element.

Numbering (shared bit-for-bit with the product's host code):
* spatial nodes: vertices, then (p-1) nodes per edge (edges sorted by
  (min vertex, max vertex), nodes ordered from the min vertex), then interior
  nodes per triangle in reference tuple order; free = not on the boundary,
  numbered by first appearance when scanning cells in order and their local
  nodes in reference order (keeps neighbours close in memory);
* X dof = time_node * n_free + free_node, time_node = k*p + l;
* Y dof = (k*(p+1) + l) * n_free + free_node.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .basis import triangle_tuples


@dataclass
class SpatialMesh:
    dim: int
    vertices: np.ndarray          # [nv, dim]
    cells: np.ndarray             # [nc, dim+1] vertex ids (CCW in 2D)

    @property
    def n_cells(self):
        return len(self.cells)


def interval_mesh(n: int, a: float = 0.0, b: float = 1.0) -> SpatialMesh:
    x = a + (b - a) * np.arange(n + 1) / n
    cells = np.array([[i, i + 1] for i in range(n)], dtype=np.int64)
    return SpatialMesh(1, x[:, None], cells)


def union_jack_mesh(n: int) -> SpatialMesh:
    """n x n squares on (0,1)^2, square (i,j) cut along the diagonal through its
    corner with even i+j (the NVB union-jack pattern: even-parity vertices have
    valence 8, odd ones valence 4)."""
    vid = lambda i, j: j * (n + 1) + i
    verts = np.array([(i / n, j / n) for j in range(n + 1) for i in range(n + 1)])
    cells = []
    for j in range(n):
        for i in range(n):
            if (i + j) % 2 == 0:
                cells.append((vid(i, j), vid(i + 1, j), vid(i + 1, j + 1)))
                cells.append((vid(i, j), vid(i + 1, j + 1), vid(i, j + 1)))
            else:
                cells.append((vid(i, j), vid(i + 1, j), vid(i, j + 1)))
                cells.append((vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)))
    return SpatialMesh(2, verts, np.array(cells, dtype=np.int64))


def criss_cross_mesh(n: int) -> SpatialMesh:
    """n x n squares each cut by both diagonals (4 triangles around a centre vertex)."""
    vid = lambda i, j: j * (n + 1) + i
    verts = [(i / n, j / n) for j in range(n + 1) for i in range(n + 1)]
    nv0 = len(verts)
    cells = []
    for j in range(n):
        for i in range(n):
            c = nv0 + j * n + i
            verts.append(((i + 0.5) / n, (j + 0.5) / n))
            a, b, cc, d = vid(i, j), vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)
            cells += [(a, b, c), (b, cc, c), (cc, d, c), (d, a, c)]
    return SpatialMesh(2, np.array(verts), np.array(cells, dtype=np.int64))


@dataclass
class SpatialDofs:
    """P_p Lagrange node numbering on a conforming simplicial mesh."""
    p: int
    n_nodes: int
    cell_nodes: np.ndarray        # [nc, Ns] node id per local basis function
    boundary: np.ndarray          # [n_nodes] bool
    free_index: np.ndarray        # [n_nodes] -> free id or -1
    n_free: int
    node_coords: np.ndarray       # [n_nodes, dim]
    edges: np.ndarray = field(default=None)           # [ne, 2] (2D only)
    boundary_vertex: np.ndarray = field(default=None)  # [nv] bool
    boundary_edge: np.ndarray = field(default=None)    # [ne] bool


def spatial_dofs(mesh: SpatialMesh, p: int) -> SpatialDofs:
    nv = len(mesh.vertices)
    if mesh.dim == 1:
        nc = mesh.n_cells
        n_nodes = nv + nc * (p - 1)
        cell_nodes = np.zeros((nc, p + 1), dtype=np.int64)
        coords = np.zeros((n_nodes, 1))
        coords[:nv] = mesh.vertices
        for c, (v0, v1) in enumerate(mesh.cells):
            cell_nodes[c, 0] = v0
            cell_nodes[c, p] = v1
            for j in range(1, p):
                nid = nv + c * (p - 1) + (j - 1)
                cell_nodes[c, j] = nid
                coords[nid] = mesh.vertices[v0] + (mesh.vertices[v1] - mesh.vertices[v0]) * j / p
        # boundary = vertices of exactly one cell (the interval ends; every
        # segment's ends when the mesh holds several disjoint segments)
        bvert = np.bincount(mesh.cells.ravel(), minlength=nv) == 1
        boundary = np.zeros(n_nodes, dtype=bool)
        boundary[:nv] = bvert
        free_index = _first_appearance(cell_nodes, boundary, n_nodes)
        return SpatialDofs(p, n_nodes, cell_nodes, boundary, free_index, int((free_index >= 0).sum()), coords,
                           None, bvert, None)

    cells = mesh.cells
    nc = len(cells)
    edge_set = {}
    count = {}
    for c in range(nc):
        for (a, b) in ((cells[c, 0], cells[c, 1]), (cells[c, 1], cells[c, 2]), (cells[c, 2], cells[c, 0])):
            key = (min(a, b), max(a, b))
            count[key] = count.get(key, 0) + 1
    edges = np.array(sorted(count.keys()), dtype=np.int64)
    for e, key in enumerate(map(tuple, edges)):
        edge_set[key] = e
    ne = len(edges)
    bedge = np.array([count[tuple(e)] == 1 for e in edges], dtype=bool)
    bvert = np.zeros(nv, dtype=bool)
    bvert[edges[bedge].ravel()] = True
    tuples = triangle_tuples(p)
    interior_tuples = [t for t in tuples if min(t) > 0]
    ni = len(interior_tuples)
    n_nodes = nv + ne * (p - 1) + nc * ni
    cell_nodes = np.zeros((nc, len(tuples)), dtype=np.int64)
    coords = np.zeros((n_nodes, 2))
    coords[:nv] = mesh.vertices
    boundary = np.zeros(n_nodes, dtype=bool)
    boundary[:nv] = bvert
    for c in range(nc):
        V = cells[c]
        for k, t in enumerate(tuples):
            nz = [i for i in range(3) if t[i] > 0]
            if len(nz) == 1:
                nid = V[nz[0]]
            elif len(nz) == 2:
                i, j = nz
                vi, vj = V[i], V[j]
                e = edge_set[(min(vi, vj), max(vi, vj))]
                off = t[j] if vi < vj else t[i]
                nid = nv + e * (p - 1) + (off - 1)
                boundary[nid] = bedge[e]
            else:
                nid = nv + ne * (p - 1) + c * ni + interior_tuples.index(t)
            cell_nodes[c, k] = nid
            coords[nid] = (t[0] * mesh.vertices[V[0]] + t[1] * mesh.vertices[V[1]]
                           + t[2] * mesh.vertices[V[2]]) / p
    free_index = _first_appearance(cell_nodes, boundary, n_nodes)
    return SpatialDofs(p, n_nodes, cell_nodes, boundary, free_index, int((free_index >= 0).sum()), coords,
                       edges, bvert, bedge)


def _first_appearance(cell_nodes, boundary, n_nodes):
    """Free ids in order of first appearance over (cell, local node)."""
    free_index = -np.ones(n_nodes, dtype=np.int64)
    nf = 0
    for nid in cell_nodes.ravel():
        if not boundary[nid] and free_index[nid] < 0:
            free_index[nid] = nf
            nf += 1
    return free_index


@dataclass
class SpaceTimeMesh:
    space: SpatialMesh
    times: np.ndarray             # [m+1]

    @property
    def n_intervals(self):
        return len(self.times) - 1

    @property
    def n_cells(self):
        return self.space.n_cells * self.n_intervals


@dataclass
class SpaceTimeDofs:
    """X and Y DOF maps per space-time cell (cell = k*n_space_cells + s)."""
    p: int
    sd: SpatialDofs
    nX: int
    nY: int
    x_map: np.ndarray             # [n_cells, n_loc] global X dof or -1
    y_map: np.ndarray             # [n_cells, n_loc] global Y dof or -1


def spacetime_dofs(stm: SpaceTimeMesh, p: int) -> SpaceTimeDofs:
    sd = spatial_dofs(stm.space, p)
    m = stm.n_intervals
    ns = stm.space.n_cells
    Ns = sd.cell_nodes.shape[1]
    Nt = p + 1
    nf = sd.n_free
    nX = (m * p + 1) * nf
    nY = m * (p + 1) * nf
    fi = sd.free_index[sd.cell_nodes]           # [ns, Ns]
    x_map = np.zeros((m * ns, Ns * Nt), dtype=np.int64)
    y_map = np.zeros((m * ns, Ns * Nt), dtype=np.int64)
    for k in range(m):
        for l in range(Nt):
            xm = np.where(fi >= 0, (k * p + l) * nf + fi, -1)
            ym = np.where(fi >= 0, (k * Nt + l) * nf + fi, -1)
            x_map[k * ns:(k + 1) * ns, l::Nt] = xm
            y_map[k * ns:(k + 1) * ns, l::Nt] = ym
    return SpaceTimeDofs(p, sd, nX, nY, x_map, y_map)
