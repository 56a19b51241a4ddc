"""1+1D adaptive space-time meshes, their X and Y spaces, and the adaptive X loop.

Mesh1D and FESpace1D wrap the library's restatement of the reference's 1+1D cell tree,
refinement closure and DoF numbering (apf_mesh1d_*, apf_fespace1d_*): cell ids, true-DoF
numbers and restriction rows are the reference's.  An FESpace1D carries the arrays of
apf_cell_space_desc, so a pair (X, Y) feeds GeneralSystem (the GPU element-quadrature
assembly) directly.

AdaptiveX1D runs the reference's adaptive driver (algorithm/adaptive_driver.hpp:130-1230)
on that pair: X = FESpace(SpaceTimePolicy) on the root cell refined `x_uniform_levels`
times, Y = FESpace(SpaceOnlyPolicy) initialised on X's active cells
(algorithm/examples/space_time_1d/space_time_1d_example.hpp:113-132 and
adaptive_driver.hpp:1129).  Per outer iteration the saddle system is assembled and solved on
the GPU and the X indicator eta^2 = lambda_Y^2 + initial-trace^2 per X cell
(indicator_utilities.hpp:435-473) is computed on the GPU.  With the inner Y loop the Y
estimator decides whether Y is accepted or Doerfler-marked and refined, and the solve
repeats; without it Y is X's mesh.  X is then Doerfler-marked with theta_x
(adaptive_driver.hpp:974-1012) and refined (x_space_.refine, :1117), and Y is
re-initialised on X's active cells.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import (APF_BAD_ARG, APF_FP64, APFError, Context, GeneralEstimator, GeneralSystem, _check, _ptr,
               combine_estimator, doerfler_mark, lib, y_estimator_threshold)


class Mesh1D:
    """Root cell [x0, x1] x [t0, t1] and its refinement tree (apf_mesh1d_*)."""

    def __init__(self, x0=0.0, x1=1.0, t0=0.0, t1=1.0, _handle=None):
        h = _handle
        if h is None:
            h = ctypes.c_void_p()
            _check(lib().apf_mesh1d_create(x0, x1, t0, t1, ctypes.byref(h)), None, "apf_mesh1d_create")
        self.handle = h
        self._destroy = lib().apf_mesh1d_destroy

    @classmethod
    def slab(cls, source: "Mesh1D", source_cells, t_begin, t_end):
        """the mesh of one time slab: the source cells sliced to [t_begin, t_end]
        (apf_mesh1d_create_slab)"""
        c = np.ascontiguousarray(source_cells, dtype=np.int32)
        h = ctypes.c_void_p()
        _check(lib().apf_mesh1d_create_slab(source.handle, _ptr(c), len(c), float(t_begin), float(t_end),
                                            ctypes.byref(h)), None, "apf_mesh1d_create_slab")
        return cls(_handle=h)

    def __del__(self):
        if getattr(self, "handle", None):
            self._destroy(self.handle)
            self.handle = None

    def info(self):
        n, nx, nt = (ctypes.c_int64() for _ in range(3))
        _check(lib().apf_mesh1d_info(self.handle, ctypes.byref(n), ctypes.byref(nx), ctypes.byref(nt)))
        return n.value, nx.value, nt.value

    def vertices(self):
        _, nx, nt = self.info()
        x, t = np.zeros(nx), np.zeros(nt)
        _check(lib().apf_mesh1d_vertices(self.handle, _ptr(x), _ptr(t)))
        return x, t

    def cells(self):
        """dict of per-cell arrays: parent, generation, is_leaf, vertex_ids [n, 4], box [n, 4]"""
        n, _, _ = self.info()
        out = dict(parent=np.zeros(n, np.int32), generation=np.zeros(n, np.int32), is_leaf=np.zeros(n, np.int32),
                   vertex_ids=np.zeros((n, 4), np.int32), box=np.zeros((n, 4)))
        _check(lib().apf_mesh1d_cells(self.handle, _ptr(out["parent"]), _ptr(out["generation"]), _ptr(out["is_leaf"]),
                                      _ptr(out["vertex_ids"]), _ptr(out["box"])))
        return out

    def leaf_cells(self):
        return np.flatnonzero(self.cells()["is_leaf"]).astype(np.int32)


class FESpace1D:
    """X (continuous_in_time) or Y space of degree p on a Mesh1D, with the arrays of
    apf_cell_space_desc: vertices [nv, 1], cells [n_active, 2] spatial vertex ids,
    times [n_active, 2], restr_ptr / restr_dof / restr_w, n_dofs (true DoFs)."""

    dim = 1

    def __init__(self, mesh: Mesh1D, p: int, continuous_in_time: bool, active_cells):
        self.mesh, self.p, self.continuous_in_time = mesh, p, bool(continuous_in_time)
        act = np.ascontiguousarray(active_cells, dtype=np.int32)
        h = ctypes.c_void_p()
        _check(lib().apf_fespace1d_create(mesh.handle, p, int(self.continuous_in_time), _ptr(act), len(act),
                                          ctypes.byref(h)), None, "apf_fespace1d_create")
        self.handle = h
        self._destroy = lib().apf_fespace1d_destroy
        self._refresh()

    def __del__(self):
        if getattr(self, "handle", None):
            self._destroy(self.handle)
            self.handle = None

    @property
    def n_loc(self):
        return (self.p + 1) ** 2

    @property
    def n_cells(self):
        return len(self.active)

    def _refresh(self):
        na, nt, nc, ne = (ctypes.c_int64() for _ in range(4))
        _check(lib().apf_fespace1d_info(self.handle, ctypes.byref(na), ctypes.byref(nt), ctypes.byref(nc),
                                        ctypes.byref(ne)), None, "apf_fespace1d_info")
        self.n_dofs, self.n_constrained = nt.value, nc.value
        self.active = np.zeros(na.value, np.int32)
        _check(lib().apf_fespace1d_active_cells(self.handle, _ptr(self.active)))
        self.cell_dofs = np.zeros((na.value, self.n_loc), np.int32)
        _check(lib().apf_fespace1d_cell_dofs(self.handle, _ptr(self.cell_dofs)))
        self.restr_ptr = np.zeros(na.value * self.n_loc + 1, np.int64)
        self.restr_dof = np.zeros(ne.value, np.int32)
        self.restr_w = np.zeros(ne.value)
        _check(lib().apf_fespace1d_restriction(self.handle, _ptr(self.restr_ptr), _ptr(self.restr_dof),
                                               _ptr(self.restr_w)), None, "apf_fespace1d_restriction")
        self.con_ptr = np.zeros(nc.value + 1, np.int64)
        nce = 0
        if nc.value:
            _check(lib().apf_fespace1d_constraints(self.handle, _ptr(self.con_ptr), None, None))
            nce = int(self.con_ptr[-1])
        self.con_masters = np.zeros(nce, np.int32)
        self.con_weights = np.zeros(nce)
        if nc.value:
            _check(lib().apf_fespace1d_constraints(self.handle, _ptr(self.con_ptr), _ptr(self.con_masters),
                                                   _ptr(self.con_weights)))
        x, _ = self.mesh.vertices()
        self.vertices = x[:, None].copy()
        cl = self.mesh.cells()
        self.cells = np.ascontiguousarray(cl["vertex_ids"][self.active, :2])
        self.times = np.ascontiguousarray(cl["box"][self.active, 2:])
        self.box = np.ascontiguousarray(cl["box"][self.active])
        self.generation = cl["generation"][self.active]

    def refine(self, marked):
        """FESpace::refine of the marked active cells (in marking order)"""
        m = np.ascontiguousarray(marked, dtype=np.int32)
        _check(lib().apf_fespace1d_refine(self.handle, _ptr(m), len(m)), None, "apf_fespace1d_refine")
        self._refresh()

    def neighbours(self, temporal: bool):
        """[n_active, 2, max] neighbour cell ids across the spatial or temporal faces, -1 padded
        (the mesh file's layout)"""
        mx = ctypes.c_int()
        _check(lib().apf_fespace1d_neighbours(self.handle, int(temporal), ctypes.byref(mx), None))
        out = np.zeros((self.n_cells, 2, mx.value), np.int32)
        _check(lib().apf_fespace1d_neighbours(self.handle, int(temporal), ctypes.byref(mx), _ptr(out)))
        return out

    def local_coefficients(self, coef):
        """[n_active, (p+1)^2] local values of the function with true-DoF vector coef"""
        coef = np.asarray(coef, dtype=np.float64)
        vals = np.zeros(len(self.restr_ptr) - 1)
        seg = np.repeat(np.arange(len(vals)), np.diff(self.restr_ptr))
        np.add.at(vals, seg, self.restr_w * coef[self.restr_dof])
        return vals.reshape(self.n_cells, self.n_loc)

    def ancestors_of(self, fine: "FESpace1D"):
        """position in this space's active list of each active cell's first active ancestor in `fine`"""
        out = np.zeros(fine.n_cells, np.int32)
        _check(lib().apf_fespace1d_ancestors(self.handle, fine.handle, _ptr(out)), None, "apf_fespace1d_ancestors")
        return out


class SpaceTimePair:
    """X and Y spaces on one mesh, in the form GeneralSystem takes."""

    dim = 1

    def __init__(self, X: FESpace1D, Y: FESpace1D):
        if X.p != Y.p:
            raise ValueError("X and Y need the same degree")
        self.X, self.Y, self.p = X, Y, X.p
        self.y_ancestor = X.ancestors_of(Y)


def slab_groups(Y: FESpace1D, max_direct=64):
    """The estimator's time slabs (between consecutive time endpoints of Y's cells, each
    cell sliced into the slabs it spans: time_slabs/time_slab_builder.hpp:53-120), runs of
    slabs with one spatial partition grouped as tensor meshes, for GeneralEstimator:
    [(Mesh of the group, coarse meshes, source Y cell of each group cell)].  A partition with
    more than `max_direct` intervals gets coarse meshes of every other vertex (nested: each
    coarse interval is a union of fine ones) for the slab solve's multigrid; smaller ones are
    solved on one level."""
    from . import Mesh

    box = Y.box
    times = np.unique(np.concatenate([box[:, 2], box[:, 3]]))
    runs = []
    for k in range(len(times) - 1):
        t0, t1 = times[k], times[k + 1]
        inside = np.flatnonzero((box[:, 2] <= t0) & (box[:, 3] >= t1))
        inside = inside[np.argsort(box[inside, 0], kind="stable")]
        part = tuple(map(tuple, box[inside, :2]))
        if runs and runs[-1][0] == part:
            runs[-1][1].append(t1)
            runs[-1][2].append(inside)
        else:
            runs.append((part, [t0, t1], [inside]))
    out = []
    for part, ts, srcs in runs:
        verts = np.array([part[0][0]] + [b for (_, b) in part])
        ns = len(part)
        mesh = Mesh(1, verts[:, None], np.stack([np.arange(ns), np.arange(1, ns + 1)], axis=1).astype(np.int32),
                    np.array(ts))
        coarse = []
        v = verts
        while len(v) - 1 > max_direct:
            keep = v[::2] if (len(v) - 1) % 2 == 0 else np.concatenate([v[::2], v[-1:]])
            v = keep
            nc = len(v) - 1
            coarse.append(Mesh(1, v[:, None], np.stack([np.arange(nc), np.arange(1, nc + 1)], axis=1).astype(np.int32),
                               np.array([0.0, 1.0])))  # finest coarse mesh first
        out.append((mesh, coarse, np.concatenate(srcs).astype(np.int32)))
    return out


class TimeSlabSpace:
    """The reference's time-slab space of a Y space (TimeSlabBuilder::initialize,
    time_slabs/time_slab_builder.hpp:53-200): slab times = the distinct time endpoints of the
    active cells, ascending; slab k holds, in the source's active order, every active cell
    spanning [t_k, t_{k+1}], sliced to it, on a mesh of its own with a Y space (continuous in
    space, discontinuous in time) of the source's degree.  Attributes: times, slabs (list of
    dict(mesh, space, source_cells))."""

    def __init__(self, Y: FESpace1D):
        box = Y.box
        self.times = np.unique(np.concatenate([box[:, 2], box[:, 3]]))
        self.slabs = []
        for k in range(len(self.times) - 1):
            t0, t1 = self.times[k], self.times[k + 1]
            inside = np.flatnonzero((box[:, 2] <= t0) & (box[:, 3] >= t1))
            src = Y.active[inside].astype(np.int32)
            mesh = Mesh1D.slab(Y.mesh, src, t0, t1)
            space = FESpace1D(mesh, Y.p, False, np.arange(len(src), dtype=np.int32))
            self.slabs.append(dict(mesh=mesh, space=space, source_cells=src, t_begin=t0, t_end=t1))

    @property
    def n_slabs(self):
        return len(self.slabs)

    def edge_patches(self):
        """The 1+1D patches of every slab (TimeSlabEdgePatchBuilder,
        time_slabs/time_slab_edge_patch_builder.hpp:349-527): the slab's cells sorted by
        (x_begin, x_end, slab cell id); a boundary patch on the first cell's left vertex, one
        interior patch per shared vertex of consecutive cells, a boundary patch on the last
        cell's right vertex; patch ids consecutive over the slabs.  Per patch: dict(patch_id,
        slab_id, spatial_vertex_id (slab mesh), kind (0 interior, 1 boundary), x_vertex,
        t_begin, t_end, cells = [dict(slab_cell_id, source_cell_id, left_vertex_id,
        right_vertex_id, side (0 left of the vertex, 1 right of it), x_begin, x_end)])."""
        out = []
        for k, sl in enumerate(self.slabs):
            sp = sl["space"]
            vids = sp.cells
            cells = [dict(slab_cell_id=i, source_cell_id=int(sl["source_cells"][i]), left_vertex_id=int(vids[i, 0]),
                          right_vertex_id=int(vids[i, 1]), x_begin=float(sp.box[i, 0]), x_end=float(sp.box[i, 1]))
                     for i in range(sp.n_cells)]
            cells.sort(key=lambda c: (c["x_begin"], c["x_end"], c["slab_cell_id"]))
            if not cells:
                continue

            def patch(kind, vid, xv, members):
                return dict(patch_id=len(out), slab_id=k, spatial_vertex_id=vid, kind=kind, x_vertex=xv,
                            t_begin=float(sl["t_begin"]), t_end=float(sl["t_end"]), cells=members)

            out.append(patch(1, cells[0]["left_vertex_id"], cells[0]["x_begin"], [dict(cells[0], side=1)]))
            for a, b in zip(cells[:-1], cells[1:]):
                if a["right_vertex_id"] != b["left_vertex_id"] or abs(a["x_end"] - b["x_begin"]) > 1e-12:
                    raise ValueError("edge_patches: adjacent slab cells do not share a spatial vertex")
                out.append(patch(0, a["right_vertex_id"], a["x_end"], [dict(a, side=0), dict(b, side=1)]))
            out.append(patch(1, cells[-1]["right_vertex_id"], cells[-1]["x_end"], [dict(cells[-1], side=0)]))
        return out

    def restrict(self, Y: FESpace1D, coef):
        """per slab, the true coefficients of the Y function `coef` restricted to the slab
        (nodal values at the slab space's DoFs, each taken on its source cell: exact, the
        slab cells being slices of the source cells at the same degree)"""
        pos = {int(c): i for i, c in enumerate(Y.active)}
        return [prolong_nodal(Y, coef, sl["space"], cell_map=np.array([pos[int(c)] for c in sl["source_cells"]]))
                for sl in self.slabs]


def _lagrange_values(p, s):
    """[len(s), p+1] values of the degree-p Lagrange basis on the equidistant nodes of [0, 1]"""
    s = np.asarray(s, dtype=np.float64)
    nodes = np.arange(p + 1) / p
    V = np.ones((s.size, p + 1))
    for a in range(p + 1):
        for k in range(p + 1):
            if k != a:
                V[:, a] *= (s - nodes[k]) / (nodes[a] - nodes[k])
    return V


def prolong_nodal(src: FESpace1D, coef, dst: FESpace1D, cell_map=None):
    """True coefficients in `dst` of the `src` function with true coefficients `coef`: its
    value at the node of each free dst DoF, taken on the first active dst cell (in active
    order, local order) where that DoF is free, evaluated on the src cell containing it
    (fespace/prolongation.hpp:15-97).  dst's active cells lie inside src's, on the same mesh,
    or cell_map gives the position in src's active list of each dst cell's container."""
    anc = src.ancestors_of(dst) if cell_map is None else np.asarray(cell_map, dtype=np.int64)
    loc = src.local_coefficients(coef)
    p, q = src.p, dst.p
    ii, jj = np.divmod(np.arange(dst.n_loc), q + 1)     # local = i_space * (q+1) + i_time
    bd, bs = dst.box, src.box[anc]
    x = bd[:, 0:1] + (bd[:, 1:2] - bd[:, 0:1]) * (ii / q)[None, :]
    t = bd[:, 2:3] + (bd[:, 3:4] - bd[:, 2:3]) * (jj / q)[None, :]
    sx = ((x - bs[:, 0:1]) / (bs[:, 1:2] - bs[:, 0:1])).ravel()
    st = ((t - bs[:, 2:3]) / (bs[:, 3:4] - bs[:, 2:3])).ravel()
    Lx, Lt = _lagrange_values(p, sx), _lagrange_values(p, st)
    c = loc[np.repeat(anc, dst.n_loc)].reshape(-1, p + 1, p + 1)
    vals = np.einsum("na,nab,nb->n", Lx, c, Lt)
    # the free dofs: restriction rows [(d, 1.0)]
    cnt = np.diff(dst.restr_ptr)
    one = np.flatnonzero(cnt == 1)
    free = one[dst.restr_w[dst.restr_ptr[one]] == 1.0]
    dof = dst.restr_dof[dst.restr_ptr[free]]
    first = np.full(dst.n_dofs, -1, np.int64)
    order = np.argsort(dof, kind="stable")
    ds = dof[order]
    head = np.concatenate([[True], ds[1:] != ds[:-1]])
    first[ds[head]] = free[order][head]
    if (first < 0).any():
        raise ValueError("prolong_nodal: a target true DoF is not free on any cell")
    return vals[first]


def g_estimate(ctx: Context, X: FESpace1D, Y: FESpace1D, x, u0=None, tol=1e-12, max_it=100000, degree=8):
    """The G estimator of an accepted (or any) inner iteration on the 1+1D adaptive pair
    (compute_g_estimator_for_accepted_inner_iteration_, adaptive_driver.hpp:2680-2905): G =
    the Y space of degree p+1 in space and time on Y's cells; A_G lambda_G = f_G - B(X,G)^T
    u_delta; G^2 = (lambda_G - I lambda_delta)^T A_G (lambda_G - I lambda_delta), with I the
    nodal prolongation.  A_G, f_G and B(X_{p+1},G) come from the degree-(p+1) GeneralSystem
    on (X_{p+1}, G); B(X,G)^T u = B(X_{p+1},G)^T (I u) since the prolongation of u is the
    same function and both rules integrate the polynomial integrand exactly.  The G system
    is solved by conjugate gradients on the device (apf_csr_block_pcg_solve) from I lambda."""
    p = X.p
    if p + 1 > 3:
        raise ValueError("g_estimate: the degree-(p+1) system needs p <= 2 on 1+1D adaptive meshes")
    u0 = smooth_initial_u0 if u0 is None else u0
    t0 = time.perf_counter()
    x = np.asarray(x, dtype=np.float64)
    lam, u = x[:Y.n_dofs], x[Y.n_dofs:]
    Xq = FESpace1D(X.mesh, p + 1, True, X.active)
    G = FESpace1D(X.mesh, p + 1, False, Y.active)
    Iu = prolong_nodal(X, u, Xq)
    Il = prolong_nodal(Y, lam, G)
    gs = GeneralSystem(ctx, SpaceTimePair(Xq, G))
    gs.assemble(u0)
    K = gs.matrix()
    nG = gs.nY
    rhs = gs.rhs()[:nG] - K.spmv(np.concatenate([np.zeros(nG), Iu]))[:nG]
    lg, st = K.block_pcg(rhs, nG, tol=tol, max_it=max_it, degree=degree, x0=Il)
    d = lg - Il
    g2 = K.block_energy(d, nG)
    coeff2 = float(d @ d)
    if g2 < -1000.0 * np.finfo(float).eps * max(1.0, coeff2):
        raise APFError(APF_BAD_ARG, f"g_estimate: negative G^2 ({g2})")
    del K, gs
    return dict(g_true_dofs=G.n_dofs, g_lambda_difference_squared=max(0.0, g2),
                g_solver_iterations=st["iterations"], g_solver_relative_residual=st["true_relative_residual"],
                g_seconds=time.perf_counter() - t0)


def smooth_initial_u0(x):
    """u0 = sin(pi x) (algorithm/examples/space_time_1d/initial_data_1d_examples.hpp:9-63: M = 1, ell = 0)"""
    return np.sin(np.pi * np.asarray(x)[..., 0])


class AdaptiveX1D:
    """The reference's adaptive driver on 1+1D meshes, on the GPU (adaptive_driver.hpp:130-1230).

    Each outer iteration solves on (X, Y), computes the X indicators and, with
    inner_y_loop=True, runs the inner Y loop: the Y estimator by source cell (flux^2 +
    reconstruction^2) against its threshold rho^2 (lambda_Y^2 + (1 + 1/(rho + 1/2)^2) trace^2);
    above it the Y cells are Doerfler-marked with theta_y and Y is refined with its grading
    closure (time_slab_adaptive_refinement.hpp:240-338, no local time-slab closure), and the
    iteration repeats; an empty marking stops the loop.  Without the inner loop Y is X's mesh
    (the published runs without inner refinement).  Then the X cells are marked with theta_x on
    the last iteration's indicators, X is refined and Y re-initialised on X's active cells."""

    def __init__(self, ctx: Context, p=1, theta_x=0.4, rho=1.0, x_uniform_levels=1, u0=smooth_initial_u0,
                 tol=1e-10, max_it=40000, with_y_estimator=False, estimator_precision=APF_FP64, hhat=True,
                 cheb_degree=24, inner_y_loop=False, theta_y=0.4, max_inner=50, with_g_estimator=False,
                 initial_guess=True, check_interval=20):
        """hhat: precondition the X block with the graph norm Hhat formed on the device (the
        reference's choice, parabolic_graph_norm_minres.hpp:46-181), else with the Schur
        complement's diagonal; cheb_degree: Chebyshev steps of each block; with_g_estimator: G^2
        of every inner iteration (g_estimate, the published adaptive_g_effective_rho runs);
        initial_guess: start MINRES from the previous solve's solution, prolonged nodally onto the
        refined spaces (the reference's use_adaptive_initial_guess, which its production runs leave
        off because their solves are direct; for MINRES it saves 40-65% of the iterations on the
        larger meshes); check_interval: MINRES iterations between true-residual checks."""
        self.ctx, self.p, self.theta_x, self.rho, self.u0 = ctx, p, theta_x, rho, u0
        self.hhat, self.cheb_degree = hhat, cheb_degree
        self.with_y_estimator, self.estimator_precision = with_y_estimator or inner_y_loop, estimator_precision
        self.inner_y_loop, self.theta_y, self.max_inner = inner_y_loop, theta_y, max_inner
        self.with_g_estimator = with_g_estimator
        self.initial_guess, self._prev = initial_guess, None
        self.check_interval = check_interval
        self.tol, self.max_it = tol, max_it
        self.mesh = Mesh1D(0.0, 1.0, 0.0, 1.0)
        self.X = FESpace1D(self.mesh, p, True, self.mesh.leaf_cells())
        for _ in range(x_uniform_levels):
            self.X.refine(self.X.active)
        self.Y = FESpace1D(self.mesh, p, False, self.X.active)
        self.history = []

    def _solve(self):
        """assemble, solve, X indicators (and the Y estimator) on the current pair"""
        pair = SpaceTimePair(self.X, self.Y)
        gs = GeneralSystem(self.ctx, pair)
        ms = gs.assemble(self.u0)
        K = gs.matrix()
        t0 = time.perf_counter()
        x0 = self._guess() if self.initial_guess else None
        H = gs.graph_norm() if self.hhat else None
        x, stats = K.minres(gs.rhs(), gs.nY, hhat=H, tol=self.tol, max_it=self.max_it, degree=self.cheb_degree,
                            precond=APF_FP64, x0=x0, check_interval=self.check_interval)
        if self.initial_guess:
            self._prev = (FESpace1D(self.mesh, self.p, True, self.X.active),
                          FESpace1D(self.mesh, self.p, False, self.Y.active), x)
        t_solve = time.perf_counter() - t0
        del H
        lam2_y, tr2_x = gs.x_indicators(x, self.u0)
        lam2_x = np.zeros(self.X.n_cells)
        np.add.at(lam2_x, pair.y_ancestor, lam2_y)
        eta2_x = lam2_x + tr2_x
        lam2, tr2 = float(lam2_y.sum()), float(tr2_x.sum())
        rec = dict(x_true_dofs=self.X.n_dofs, y_true_dofs=self.Y.n_dofs, x_cells=self.X.n_cells,
                   y_cells=self.Y.n_cells, lambda_y_squared=lam2, initial_trace_squared=tr2,
                   eta_squared=float(eta2_x.sum()),
                   y_estimator_threshold_squared=y_estimator_threshold(self.rho, lam2, tr2),
                   minres_iterations=stats["iterations"], minres_rel_residual=stats["true_relative_residual"],
                   minres_seconds=t_solve, assembly_ms=ms, solution=x, eta2_by_x_cell=eta2_x)
        if self.with_y_estimator:
            t0 = time.perf_counter()
            ge = GeneralEstimator(gs, slab_groups(self.Y))
            ge.assemble(self.u0)
            est = ge.estimate(x, self.estimator_precision, 2, per_cell=True)
            rec["y_estimator"] = est
            rec["y_estimator_squared"] = est["y_est2"]
            rec["y_estimator_by_cell"] = combine_estimator(est["flux_by_cell"], est["recon_by_cell"])
            rec["estimator_seconds"] = time.perf_counter() - t0
            del ge
        if self.with_g_estimator:
            rec.update(g_estimate(self.ctx, self.X, self.Y, x, self.u0))
        del K
        del gs
        return rec

    def _guess(self):
        """the previous solution prolonged onto the current spaces (u always: X only grows; lambda
        where every Y cell lies inside a previous Y cell, else zero)"""
        if self._prev is None:
            return None
        Xp, Yp, xp = self._prev
        u = prolong_nodal(Xp, xp[Yp.n_dofs:], self.X)
        try:
            lam = prolong_nodal(Yp, xp[:Yp.n_dofs], self.Y)
        except APFError:
            lam = np.zeros(self.Y.n_dofs)
        return np.concatenate([lam, u])

    def step(self, refine=True):
        """one outer iteration (with its inner Y loop when enabled); returns the record of the
        accepted (last) inner iteration, with the list of all of them under "inner"."""
        inner = []
        for it in range(self.max_inner if self.inner_y_loop else 1):
            rec = self._solve()
            rec["inner_iteration"] = it
            rec["y_true_dofs_before"] = self.Y.n_dofs
            inner.append(rec)
            if not self.inner_y_loop:
                break
            rec["accepted"] = rec["y_estimator_squared"] <= rec["y_estimator_threshold_squared"]
            if rec["accepted"]:
                break
            marked_y, _ = doerfler_mark(rec["y_estimator_by_cell"], self.theta_y,
                                        ids=self.Y.active.astype(np.int64))
            rec["marked_y_cells"] = np.asarray(marked_y)
            if len(marked_y) == 0 or it + 1 >= self.max_inner:
                break
            self.Y.refine(marked_y)
            rec["y_true_dofs_after"] = self.Y.n_dofs
        rec["inner"] = inner
        rec.setdefault("y_true_dofs_after", self.Y.n_dofs)
        if refine:
            marked, _ = doerfler_mark(rec["eta2_by_x_cell"], self.theta_x, ids=self.X.active.astype(np.int64))
            rec["marked_x_cells"] = np.asarray(marked)
            self.X.refine(marked)
            self.Y = FESpace1D(self.mesh, self.p, False, self.X.active)
            rec["x_true_dofs_after"] = self.X.n_dofs
        self.history.append(rec)
        return rec
