"""Main saddle-point assembly (oracle; test infrastructure only).

Restates, per space-time cell K (X cell == Y cell on tensor meshes, so the X
basis is the pretabulated one, src/finite_element/assembly/main_system/
dense_local_main_blocks_2d.hpp:208-216):

* A_y(i,j) = sum_q w_q |J| grad_x phi_i . M grad_x phi_j
  (src/finite_element/assembly/detail/local_bilinear_forms.hpp:24-81)
* B(i,j)   = sum_q w_q |J| (d_t psi_i phi_j + grad_x psi_i . M grad_x phi_j),
  rows = X basis, cols = Y basis
  (src/finite_element/assembly/detail/local_mixed_bilinear_forms.hpp:39-153)
* f(i)     = sum_q w_q |J| ell phi_i
  (src/finite_element/assembly/detail/local_linear_forms.hpp:18-46)
* C(i,j)   = - sum_q w_q |T| psi_i(x_q,0) psi_j(x_q,0) on cells touching t=0
  (src/finite_element/assembly/detail/local_trace_bilinear_forms.hpp:20-49)
* g(i)     = g_scale * sum_q w_q |T| u0(x_q) psi_i(x_q,0), g_scale = -1
  (src/finite_element/assembly/detail/local_trace_linear_forms.hpp:22-52;
   src/finite_element/system/solve_main_system.hpp:1833)

Scatter: Dirichlet rows/cols are dropped and duplicate (row, col) entries are
summed (src/finite_element/assembly/scatter.hpp:66-115, constrained_dofs.hpp:94-102).
The pattern is the full element-coupling pattern (the two-pass path,
src/finite_element/assembly/main_system/two_pass_full_saddle_2d.hpp:421-459),
not the |a_ij| > 1e-15 pruned triplet pattern.

The spatial geometry follows CellGeometry<Space,2>::make
(src/finite_element/geometry/cell_geometry.hpp:161-207): J = [v1-v0, v2-v0],
grad_x = J^{-T} grad_ref, d_t = d_tau / h_t.

Pinned by the Kronecker closed forms A_y = K_s(M) (x) M_t, B = M_s (x) D_t +
K_s (x) M_t on each cell (tests/test_oracle_assembly.py).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

from .basis import spacetime_lagrange
from .mesh import SpaceTimeDofs, SpaceTimeMesh
from .quadrature import cell_rule, space_rule


def spatial_geometry(mesh):
    """Per spatial cell: J [nc,d,d], detJ [nc], invJT [nc,d,d], origin [nc,d]."""
    V = mesh.vertices[mesh.cells]                 # [nc, d+1, d]
    J = np.stack([V[:, i + 1] - V[:, 0] for i in range(mesh.dim)], axis=-1)  # [nc,d,d]
    if mesh.dim == 1:
        det = J[:, 0, 0]
        invJT = 1.0 / J
    else:
        det = J[:, 0, 0] * J[:, 1, 1] - J[:, 0, 1] * J[:, 1, 0]
        invJT = np.zeros_like(J)
        invJT[:, 0, 0] = J[:, 1, 1] / det
        invJT[:, 0, 1] = -J[:, 1, 0] / det
        invJT[:, 1, 0] = -J[:, 0, 1] / det
        invJT[:, 1, 1] = J[:, 0, 0] / det
    return J, det, invJT, V[:, 0]


@dataclass
class Problem:
    """Problem data: M as one SPD tensor per spatial cell ([ns,d,d]), constant ell,
    u0 a callable on spatial points [n,d] -> [n]."""
    M: np.ndarray
    ell: float
    u0: object


def identity_M(mesh):
    d = mesh.dim
    return np.broadcast_to(np.eye(d), (mesh.n_cells, d, d)).copy()


@dataclass
class SaddleBlocks:
    A: sp.csr_matrix   # nY x nY
    B: sp.csr_matrix   # nX x nY
    C: sp.csr_matrix   # nX x nX
    f: np.ndarray      # nY
    g: np.ndarray      # nX
    Bdt: sp.csr_matrix  # nX x nY, time-derivative part of B
    Ax: sp.csr_matrix   # nX x nX, int M grad_x psi_i . grad_x psi_j

    def K(self):
        """Compose [A_y B^T; B C] keeping explicit zeros (full pattern)."""
        A = self.A.tocoo()
        B = self.B.tocoo()
        C = self.C.tocoo()
        nY = A.shape[0]
        r = np.concatenate([A.row, B.col, B.row + nY, C.row + nY])
        c = np.concatenate([A.col, B.row + nY, B.col, C.col + nY])
        v = np.concatenate([A.data, B.data, B.data, C.data])
        n = nY + C.shape[0]
        K = sp.coo_matrix((v, (r, c)), shape=(n, n)).tocsr()
        K.sort_indices()
        return K

    def rhs(self):
        return np.concatenate([self.f, self.g])


def local_blocks(stm: SpaceTimeMesh, p: int, prob: Problem, cells=None):
    """Local A_y, B, f (and B_dt, A_x) for the given space-time cells (default: all)."""
    mesh = stm.space
    d = mesh.dim
    ns = mesh.n_cells
    if cells is None:
        cells = np.arange(stm.n_cells)
    s = cells % ns
    k = cells // ns
    ht = (stm.times[1:] - stm.times[:-1])[k]
    _, det, invJT, _ = spatial_geometry(mesh)
    pts, w = cell_rule(d, p)
    vals, grads = spacetime_lagrange(d, p, pts)                # [nq,n], [nq,n,d+1]
    gx = np.einsum("cab,qib->cqia", invJT[s], grads[:, :, :d])  # [nc,nq,n,d]
    dt = grads[None, :, :, d] / ht[:, None, None]                # [nc,nq,n]
    dmu = np.abs(det[s])[:, None] * ht[:, None] * w[None, :]     # [nc,nq]
    Mgx = np.einsum("cab,cqib->cqia", prob.M[s], gx)
    Ax = np.einsum("cq,cqia,cqja->cij", dmu, gx, Mgx)
    Bdt = np.einsum("cq,cqi,qj->cij", dmu, dt, vals)
    B = Bdt + Ax
    f = prob.ell * np.einsum("cq,qi->ci", dmu, vals)
    return Ax, B, f, Bdt


def local_trace(stm: SpaceTimeMesh, p: int, prob: Problem, g_scale: float = -1.0):
    """C and g on the cells of the first interval (the t=0 layer)."""
    mesh = stm.space
    d = mesh.dim
    _, det, _, _ = spatial_geometry(mesh)
    xs, ws = space_rule(d, p)
    pts = np.concatenate([xs, np.zeros((len(ws), 1))], axis=1)
    vals, _ = spacetime_lagrange(d, p, pts)                    # [nb, n]
    meas = np.abs(det)[:, None] * ws[None, :]                  # [ns, nb]
    C = -np.einsum("cq,qi,qj->cij", meas, vals, vals)
    # physical bottom points: v0 + J xi
    J, _, _, v0 = spatial_geometry(mesh)
    xq = v0[:, None, :] + np.einsum("cab,qb->cqa", J, xs)
    u0q = prob.u0(xq.reshape(-1, d)).reshape(mesh.n_cells, len(ws))
    g = g_scale * np.einsum("cq,cq,qi->ci", meas, u0q, vals)
    return C, g, xq, meas, vals


def _triplets(loc, rmap, cmap):
    nc, nr, ncol = loc.shape
    R = np.repeat(rmap[:, :, None], ncol, axis=2)
    Cc = np.repeat(cmap[:, None, :], nr, axis=1)
    mask = (R >= 0) & (Cc >= 0)
    return R[mask], Cc[mask], loc[mask]


def _build(trips, shape):
    """Sum duplicate triplets; explicit zeros are kept, so the pattern is the
    full element-coupling pattern (sparse '+' would drop exact zeros)."""
    if not trips:
        return sp.csr_matrix(shape)
    r = np.concatenate([t[0] for t in trips])
    c = np.concatenate([t[1] for t in trips])
    v = np.concatenate([t[2] for t in trips])
    M = sp.coo_matrix((v, (r, c)), shape=shape).tocsr()
    M.sum_duplicates()
    M.sort_indices()
    return M


def _scatter_vec(loc, rmap, n):
    mask = rmap >= 0
    out = np.zeros(n)
    np.add.at(out, rmap[mask], loc[mask])
    return out


def assemble(stm: SpaceTimeMesh, dofs: SpaceTimeDofs, prob: Problem, chunk: int = 8192) -> SaddleBlocks:
    p = dofs.p
    nX, nY = dofs.nX, dofs.nY
    tA, tB, tBdt, tAx = [], [], [], []
    f = np.zeros(nY)
    for c0 in range(0, stm.n_cells, chunk):
        cells = np.arange(c0, min(c0 + chunk, stm.n_cells))
        Al, Bl, fl, Bdtl = local_blocks(stm, p, prob, cells)
        ym = dofs.y_map[cells]
        xm = dofs.x_map[cells]
        tA.append(_triplets(Al, ym, ym))
        tB.append(_triplets(Bl, xm, ym))
        tBdt.append(_triplets(Bdtl, xm, ym))
        tAx.append(_triplets(Al, xm, xm))
        f += _scatter_vec(fl, ym, nY)
    Cl, gl, *_ = local_trace(stm, p, prob)
    ns = stm.space.n_cells
    xm0 = dofs.x_map[:ns]
    C = _build([_triplets(Cl, xm0, xm0)], (nX, nX))
    g = _scatter_vec(gl, xm0, nX)
    return SaddleBlocks(_build(tA, (nY, nY)), _build(tB, (nX, nY)), C, f, g,
                        _build(tBdt, (nX, nY)), _build(tAx, (nX, nX)))
