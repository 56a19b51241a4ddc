"""Equilibrated-flux Y-estimator on tensor meshes (oracle; test infrastructure only).

Restates compute_adaptive_test_space_iteration
(src/finite_element/time_slabs/time_slab_adaptive_loop.hpp:727-2029) for
tensor meshes:

1. main solve K [lambda; u] = [f; g] (here: sparse direct; the converged
   MINRES/PARDISO solution is the same up to the solver tolerance);
2. slab reconstruction  A_slab lt = f_slab - (B^T u)_slab per time slab
   (src/finite_element/time_slabs/time_slab_reconstruction.hpp:584-735);
   on tensor meshes the slab space is Y restricted to the slab;
3. one mixed patch problem per (slab, spatial vertex a)
   (src/finite_element/time_slabs/time_slab_equilibrated_flux_reconstruction.hpp:
   1490-1538 (1+1D), 3300-3380 (2+1D)):
      [A  B^T 0] [s]   [f]     A = int M^-1 s_i.s_j,  B = int q_i div s_j
      [B  0   c] [r] = [g]     f_i = -int psi_a grad(theta).s_i
      [0  c^T 0] [m]   [0]     g_i = int [psi_a (ell - d_t u) - grad psi_a . M grad(theta)] q_i
   theta = lt + u; c = mean-zero rows (interior patches only), solved by dense
   LU with partial pivoting (src/linear_algebra/eigen_backend/dense_solver.hpp:61-76);
   flux space: RT_p x P_p(time) with sigma.n = 0 on the outer patch boundary,
   free on internal edges and (boundary patches) on dOmega
   (src/finite_element/error_fespace/patch_rt_flux_space_2d.hpp:467-487);
4. per-cell terms (src/finite_element/time_slabs/time_slab_equilibrated_flux_indicators.hpp:
   34-190, 203-525): flux = (s + M grad theta).M^-1 (s + M grad theta),
   residual = ell - d_t u - div s, and the Y-norm ||lambda - lt||^2_Y
   (src/finite_element/time_slabs/time_slab_error_norms.hpp:371-537);
5. eta^2 = lambda_Y^2 + trace^2, trace^2 = int_{t=0} (u0 - u)^2
   (algorithm/detail/indicator_utilities.hpp:339-392, 444-472).

Pinned against the published uniform-run totals in
algorithm_data_plot/{1d,2d}/smooth_initial.csv (tests/test_oracle_estimator.py)
and by the divergence-residual invariant div s = Pi(ell - d_t u).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import scipy.sparse.linalg as spl

from .assembly import Problem, SaddleBlocks, assemble, local_trace, spatial_geometry
from .basis import legendre01, spacetime_lagrange
from .mesh import SpaceTimeDofs, SpaceTimeMesh
from .quadrature import cell_rule
from .rt import REF_EDGES, RTTriangle, flux_1d_ref, orthonormal_triangle_scalar, triangle_monomials


@dataclass
class Patch:
    slab: int
    vertex: int
    boundary: bool
    cells: np.ndarray                 # spatial cell ids
    n_flux: int                       # spatial flux dofs
    n_scalar: int                     # spatial scalar dofs
    flux_map: list = field(default_factory=list)   # per cell: [(local, patch dof, sign)]
    scalar_off: list = field(default_factory=list)  # per cell: first scalar dof


def build_patches_spatial(mesh, p):
    """Spatial vertex patches (shared by every slab of a tensor mesh).

    Topology follows TimeSlabVertexPatchSet (src/finite_element/time_slabs/
    time_slab_vertex_patch_builder.hpp:24-140, 517-648) and the edge freedom
    rule of PatchRTFluxSpace2D::edge_should_be_free_.
    """
    d = mesh.dim
    nv = len(mesh.vertices)
    cells_of = [[] for _ in range(nv)]
    for c, V in enumerate(mesh.cells):
        for v in V:
            cells_of[v].append(c)
    if d == 1:
        xs = mesh.vertices[:, 0]
        bvert = (xs == xs.min()) | (xs == xs.max())
        facet_count = {}
        for V in mesh.cells:
            for v in V:
                facet_count[v] = facet_count.get(v, 0) + 1
    else:
        ecount = {}
        for V in mesh.cells:
            for (i, j) in REF_EDGES:
                key = (min(V[i], V[j]), max(V[i], V[j]))
                ecount[key] = ecount.get(key, 0) + 1
        bvert = np.zeros(nv, dtype=bool)
        for (a, b), cnt in ecount.items():
            if cnt == 1:
                bvert[a] = bvert[b] = True
    patches = []
    for a in range(nv):
        cells = np.array(sorted(cells_of[a]), dtype=np.int64)
        is_b = bool(bvert[a])
        pt = Patch(-1, a, is_b, cells, 0, 0)
        if d == 1:
            # facets are vertices; count within the patch
            inpatch = {}
            for c in cells:
                for v in mesh.cells[c]:
                    inpatch[v] = inpatch.get(v, 0) + 1
            free_facets = sorted(v for v, cnt in inpatch.items()
                                 if cnt == 2 or (is_b and bvert[v]))
            slot = {v: i for i, v in enumerate(free_facets)}
            nfree = len(free_facets)
            off = nfree
            for c in cells:
                entries = []
                for e_loc, v in enumerate(mesh.cells[c]):
                    if v in slot:
                        entries.append((e_loc, slot[v], 1.0))
                for k in range(p):
                    entries.append((2 + k, off, 1.0))
                    off += 1
                pt.flux_map.append(entries)
            pt.n_flux = off
            ns_loc = p + 1
        else:
            inpatch = {}
            for c in cells:
                V = mesh.cells[c]
                for (i, j) in REF_EDGES:
                    key = (min(V[i], V[j]), max(V[i], V[j]))
                    inpatch[key] = inpatch.get(key, 0) + 1
            free_edges = sorted(k for k, cnt in inpatch.items()
                                if cnt == 2 or (is_b and ecount[k] == 1))
            slot = {k: i for i, k in enumerate(free_edges)}
            off = len(free_edges) * (p + 1)
            nint = p * (p + 1)
            for c in cells:
                V = mesh.cells[c]
                entries = []
                for e_loc, (i, j) in enumerate(REF_EDGES):
                    A, B = V[i], V[j]
                    key = (min(A, B), max(A, B))
                    if key not in slot:
                        continue
                    for k in range(p + 1):
                        sign = 1.0 if A < B else -((-1.0) ** k)
                        entries.append((e_loc * (p + 1) + k, slot[key] * (p + 1) + k, sign))
                for k in range(nint):
                    entries.append((3 * (p + 1) + k, off, 1.0))
                    off += 1
                pt.flux_map.append(entries)
            pt.n_flux = off
            ns_loc = (p + 1) * (p + 2) // 2
        pt.scalar_off = [i * ns_loc for i in range(len(cells))]
        pt.n_scalar = ns_loc * len(cells)
        patches.append(pt)
    return patches


class LocalBases:
    """Reference-cell tables of the flux/scalar bases at the cell rule points."""

    def __init__(self, dim, p):
        self.dim, self.p = dim, p
        pts, w = cell_rule(dim, p)
        self.pts, self.w = pts, w
        xs = pts[:, :dim]
        self.tau = pts[:, dim]
        self.Lt, _ = legendre01(p, self.tau)                      # [nq, p+1]
        if dim == 1:
            V, D = flux_1d_ref(p, xs[:, 0])
            self.flux_val = V[:, :, None]                         # [nq, nf, 1]
            self.flux_div = D                                     # d/dxi
            S, _ = legendre01(p, xs[:, 0])
            self.scal = S
            self.bary = np.stack([1.0 - xs[:, 0], xs[:, 0]], axis=1)
            self.dbary = np.array([[-1.0], [1.0]])                # d bary / d xi
        else:
            rt = RTTriangle(p)
            V, D = rt.eval_ref(xs)
            self.flux_val, self.flux_div = V, D
            R = orthonormal_triangle_scalar(p)
            self.scal = triangle_monomials(p, xs) @ R
            self.bary = np.stack([1.0 - xs[:, 0] - xs[:, 1], xs[:, 0], xs[:, 1]], axis=1)
            self.dbary = np.array([[-1.0, -1.0], [1.0, 0.0], [0.0, 1.0]])
        self.lag_val, self.lag_grad = spacetime_lagrange(dim, p, pts)


@dataclass
class EstimatorResult:
    lambda_y2: float
    trace2: float
    eta2: float
    flux2: float
    recon2: float
    y_est2: float
    div_resid_l2: float
    flux_by_cell: np.ndarray
    resid_by_cell: np.ndarray
    recon_by_cell: np.ndarray
    lam: np.ndarray = None
    u: np.ndarray = None
    lt: np.ndarray = None
    patch_sigma: dict = None


def main_solve(S: SaddleBlocks):
    x = spl.spsolve(S.K().tocsc(), S.rhs())
    return x


def slab_reconstruction(stm, dofs, S: SaddleBlocks, u):
    """Per-slab solve A_slab lt = f_slab - (B^T u)_slab (tensor meshes)."""
    nf = dofs.sd.n_free
    Nt = dofs.p + 1
    rhs = S.f - S.B.T @ u
    lt = np.zeros(dofs.nY)
    A = S.A.tocsr()
    for k in range(stm.n_intervals):
        idx = np.arange(k * Nt * nf, (k + 1) * Nt * nf)
        Ak = A[idx][:, idx].tocsc()
        lt[idx] = spl.spsolve(Ak, rhs[idx])
    return lt


def _local_coeffs(vec, m):
    return np.where(m >= 0, vec[np.maximum(m, 0)], 0.0)


def estimate(stm: SpaceTimeMesh, dofs: SpaceTimeDofs, prob: Problem, S: SaddleBlocks = None,
             x=None, keep_sigma=False) -> EstimatorResult:
    mesh = stm.space
    d, p = mesh.dim, dofs.p
    ns = mesh.n_cells
    m = stm.n_intervals
    if S is None:
        S = assemble(stm, dofs, prob)
    if x is None:
        x = main_solve(S)
    lam, u = x[:dofs.nY], x[dofs.nY:]
    lt = slab_reconstruction(stm, dofs, S, u)
    lambda_y2 = float(lam @ (S.A @ lam))

    # trace term on the t=0 layer
    _, _, xq, meas, tvals = local_trace(stm, p, prob)
    uq = np.einsum("qi,ci->cq", tvals, _local_coeffs(u, dofs.x_map[:ns]))
    u0q = prob.u0(xq.reshape(-1, d)).reshape(ns, -1)
    trace2 = float(np.sum(meas * (u0q - uq) ** 2))

    J, det, invJT, _ = spatial_geometry(mesh)
    Minv = np.linalg.inv(prob.M)
    LB = LocalBases(d, p)
    nq = len(LB.w)
    ht_all = stm.times[1:] - stm.times[:-1]

    # per space-time cell quadrature data: grad theta, d_t u
    def cell_fields(c):
        s, k = c % ns, c // ns
        ht = ht_all[k]
        cl = _local_coeffs(lt, dofs.y_map[c])
        cu = _local_coeffs(u, dofs.x_map[c])
        gref_l = np.einsum("qia,i->qa", LB.lag_grad, cl)
        gref_u = np.einsum("qia,i->qa", LB.lag_grad, cu)
        grad_theta = (gref_l[:, :d] + gref_u[:, :d]) @ invJT[s].T       # [nq, d]
        dtu = gref_u[:, d] / ht
        return grad_theta, dtu

    # physical flux basis per spatial cell (Piola)
    def flux_phys(s):
        if d == 1:
            return LB.flux_val, LB.flux_div / J[s, 0, 0]
        val = np.einsum("ab,qib->qia", J[s], LB.flux_val) / det[s]
        return val, LB.flux_div / det[s]

    patches = build_patches_spatial(mesh, p)
    Nt = p + 1
    # sigma coefficients per (slab, vertex): [n_flux, Nt]
    sigma = {}
    for k in range(m):
        ht = ht_all[k]
        for pt in patches:
            nF, nS = pt.n_flux * Nt, pt.n_scalar * Nt
            A = np.zeros((nF, nF))
            Bm = np.zeros((nS, nF))
            f = np.zeros(nF)
            g = np.zeros(nS)
            cvec = np.zeros((Nt, nS))
            for ci, s in enumerate(pt.cells):
                c = k * ns + s
                dmu = np.abs(det[s]) * ht * LB.w                              # [nq]
                fv, fd = flux_phys(s)
                loc = mesh.cells[s]
                a_loc = int(np.nonzero(loc == pt.vertex)[0][0])
                psi = LB.bary[:, a_loc]
                gpsi = invJT[s] @ LB.dbary[a_loc]                             # [d]
                grad_theta, dtu = cell_fields(c)
                # patch-level basis on this cell: flux (space-time), scalar
                Fm = np.zeros((nq, pt.n_flux, d))
                Fd = np.zeros((nq, pt.n_flux))
                for (li, pi, sg) in pt.flux_map[ci]:
                    Fm[:, pi] += sg * fv[:, li]
                    Fd[:, pi] += sg * fd[:, li]
                nsl = LB.scal.shape[1]
                Sm = np.zeros((nq, pt.n_scalar))
                Sm[:, pt.scalar_off[ci]:pt.scalar_off[ci] + nsl] = LB.scal
                Fst = (Fm[:, :, None, :] * LB.Lt[:, None, :, None]).reshape(nq, nF, d)
                Fdst = (Fd[:, :, None] * LB.Lt[:, None, :]).reshape(nq, nF)
                Sst = (Sm[:, :, None] * LB.Lt[:, None, :]).reshape(nq, nS)
                # the cell's own columns only (space-time index = patch dof * Nt + time mode)
                fp = np.array(sorted({pi for (_, pi, _) in pt.flux_map[ci]}), dtype=np.int64)
                fc = (fp[:, None] * Nt + np.arange(Nt)[None, :]).ravel()
                sc = ((pt.scalar_off[ci] + np.arange(nsl))[:, None] * Nt + np.arange(Nt)[None, :]).ravel()
                Fc = Fst[:, fc]
                MiF = np.einsum("ab,qib->qia", Minv[s], Fc)
                A[np.ix_(fc, fc)] += np.einsum("q,qia,qja->ij", dmu, Fc, MiF, optimize=True)
                Bm[np.ix_(sc, fc)] += np.einsum("q,qi,qj->ij", dmu, Sst[:, sc], Fdst[:, fc], optimize=True)
                f += np.einsum("q,qia,qa->i", dmu, Fst, -psi[:, None] * grad_theta)
                Mgt = grad_theta @ prob.M[s].T
                gval = psi * (prob.ell - dtu) - Mgt @ gpsi
                g += np.einsum("q,qi->i", dmu * gval, Sst)
                for kk in range(Nt):
                    cvec[kk] += _mean_row(Sm, LB, det[s], kk, Nt)
            if pt.boundary:
                Kmat = np.block([[A, Bm.T], [Bm, np.zeros((nS, nS))]])
                rhs = np.concatenate([f, g])
            else:
                Kmat = np.block([[A, Bm.T, np.zeros((nF, Nt))],
                                 [Bm, np.zeros((nS, nS)), cvec.T],
                                 [np.zeros((Nt, nF)), cvec, np.zeros((Nt, Nt))]])
                rhs = np.concatenate([f, g, np.zeros(Nt)])
            sol = np.linalg.solve(Kmat, rhs)
            res = np.abs(Kmat @ sol - rhs).max()
            if res > 1e-10 * max(1.0, np.abs(rhs).max()):
                raise RuntimeError(f"patch residual {res} too large (slab {k}, vertex {pt.vertex})")
            sigma[(k, pt.vertex)] = sol[:nF].reshape(pt.n_flux, Nt)

    # per-cell diagnostics
    flux_by = np.zeros(stm.n_cells)
    resid_by = np.zeros(stm.n_cells)
    recon_by = np.zeros(stm.n_cells)
    pidx = {pt.vertex: pt for pt in patches}
    for c in range(stm.n_cells):
        s, k = c % ns, c // ns
        ht = ht_all[k]
        dmu = np.abs(det[s]) * ht * LB.w
        fv, fd = flux_phys(s)
        grad_theta, dtu = cell_fields(c)
        sig = np.zeros((nq, d))
        dsig = np.zeros(nq)
        for v in mesh.cells[s]:
            pt = pidx[v]
            ci = int(np.nonzero(pt.cells == s)[0][0])
            coef = sigma[(k, v)]
            for (li, pi, sg) in pt.flux_map[ci]:
                tv = LB.Lt @ coef[pi]                      # [nq]
                sig += sg * fv[:, li] * tv[:, None]
                dsig += sg * fd[:, li] * tv
        mis = sig + grad_theta @ prob.M[s].T
        flux_by[c] = np.sum(dmu * np.einsum("qa,ab,qb->q", mis, Minv[s], mis))
        resid_by[c] = np.sum(dmu * (prob.ell - dtu - dsig) ** 2)
        e = _local_coeffs(lam - lt, dofs.y_map[c])
        ge = np.einsum("qia,i->qa", LB.lag_grad, e)[:, :d] @ invJT[s].T
        recon_by[c] = np.sum(dmu * np.einsum("qa,ab,qb->q", ge, prob.M[s], ge))
    flux2 = float(flux_by.sum())
    recon2 = float(recon_by.sum())
    return EstimatorResult(lambda_y2, trace2, lambda_y2 + trace2, flux2, recon2, flux2 + recon2,
                           float(np.sqrt(resid_by.sum())), flux_by, resid_by, recon_by,
                           lam, u, lt, sigma if keep_sigma else None)


def _mean_row(Sm, LB, detv, kk, Nt):
    """Mean-zero row for time mode kk: int_omega q_j (spatial), placed on mode kk."""
    nq, nsp = Sm.shape
    # spatial integral via the space-time rule (time weights sum to 1)
    row_sp = np.einsum("q,qi->i", np.abs(detv) * LB.w, Sm)
    out = np.zeros((nsp, Nt))
    out[:, kk] = row_sp
    return out.reshape(-1)
