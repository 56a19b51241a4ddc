"""Local flux / scalar bases for the patch problems (oracle; test infrastructure only).

The reference uses RT_p on triangles with a contravariant Piola map
(src/finite_element/basis/functions/raviart_thomas_triangle.hpp:34-86;
N = (p+1)(p+3), p+1 moments per edge, p(p+1) interior), and in 1+1D the
continuous flux space span{hat at the patch vertex} + cell bubbles
xi(1-xi) L_k(xi), k < p (src/finite_element/error_fespace/patch_flux_space_1d.hpp:62-93),
i.e. continuous P_{p+1}.  The patch solution sigma_a does not depend on the
basis chosen for these spaces, so this module builds its own moment-dual
bases:

* 2D: DOFs  N_{e,k}(s) = int_e s.n_e L_k ds  (edge e traversed counter-clockwise,
  L_k shifted Legendre in the edge parameter) and interior moments against
  (P_{p-1})^2 monomials; the nodal basis is the inverse Vandermonde in the
  monomial basis (P_p)^2 + x P~_p.
* 1D: endpoint values and the bubbles above.

Scalar spaces: P_p per cell (discontinuous), orthonormalised monomials on the
reference cell (2D: Gram-Schmidt in L2(T^); 1D: shifted Legendre).
"""
from __future__ import annotations

import numpy as np

from .basis import legendre01
from .quadrature import duffy_triangle, gauss_legendre_01

# reference triangle vertices and counter-clockwise edges, edge e opposite vertex e
REF_V = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
REF_EDGES = [(1, 2), (2, 0), (0, 1)]


def _rt_monomials(p: int):
    """Spanning set of RT_p: list of (kind, a, b) with
    kind 0 -> (x^a y^b, 0), kind 1 -> (0, x^a y^b), kind 2 -> x^a y^b (x, y) with a+b = p."""
    out = []
    for tot in range(p + 1):
        for a in range(tot, -1, -1):
            b = tot - a
            out.append((0, a, b))
            out.append((1, a, b))
    for a in range(p, -1, -1):
        out.append((2, a, p - a))
    return out


def _eval_monomials(p: int, xy):
    """values [n, N, 2], divergence [n, N] of the RT_p spanning set."""
    xy = np.atleast_2d(xy)
    # centred at the barycentre: RT_p is translation invariant and the
    # Vandermonde is far better conditioned than with raw monomials
    x, y = xy[:, 0] - 1.0 / 3.0, xy[:, 1] - 1.0 / 3.0
    mons = _rt_monomials(p)
    n = len(xy)
    V = np.zeros((n, len(mons), 2))
    D = np.zeros((n, len(mons)))

    def pw(t, e):
        return t ** e if e > 0 else np.ones_like(t)

    for j, (kind, a, b) in enumerate(mons):
        m = pw(x, a) * pw(y, b)
        dmx = a * pw(x, a - 1) * pw(y, b) if a > 0 else np.zeros_like(x)
        dmy = b * pw(x, a) * pw(y, b - 1) if b > 0 else np.zeros_like(x)
        if kind == 0:
            V[:, j, 0] = m
            D[:, j] = dmx
        elif kind == 1:
            V[:, j, 1] = m
            D[:, j] = dmy
        else:
            V[:, j, 0] = x * m
            V[:, j, 1] = y * m
            D[:, j] = 2.0 * m + x * dmx + y * dmy
    return V, D


class RTTriangle:
    """Moment-dual RT_p basis on the reference triangle."""

    def __init__(self, p: int):
        self.p = p
        self.N = (p + 1) * (p + 3)
        self.n_edge = 3 * (p + 1)
        mons = _rt_monomials(p)
        assert len(mons) == self.N
        Vdm = np.zeros((self.N, self.N))
        s, ws = gauss_legendre_01(p + 2)
        Lk, _ = legendre01(p, s)
        row = 0
        for (ia, ib) in REF_EDGES:
            A, B = REF_V[ia], REF_V[ib]
            t = B - A
            nrm = np.array([t[1], -t[0]])
            pts = A[None, :] + s[:, None] * t[None, :]
            val, _ = _eval_monomials(p, pts)
            flux = val @ nrm                                  # [ns, N]
            for k in range(p + 1):
                Vdm[row] = (ws * Lk[:, k]) @ flux
                row += 1
        qp, qw = duffy_triangle(2 * p + 2)
        val, _ = _eval_monomials(p, qp)
        for tot in range(p):
            for a in range(tot, -1, -1):
                b = tot - a
                m = qp[:, 0] ** a * qp[:, 1] ** b
                Vdm[row] = (qw * m) @ val[:, :, 0]
                row += 1
                Vdm[row] = (qw * m) @ val[:, :, 1]
                row += 1
        assert row == self.N
        self.coef = np.linalg.inv(Vdm)                        # mono_i -> basis_j

    def eval_ref(self, xy):
        V, D = _eval_monomials(self.p, xy)
        return np.einsum("nia,ij->nja", V, self.coef), D @ self.coef


def orthonormal_triangle_scalar(p: int):
    """Coefficient matrix R so that phi = m @ R is L2(T^)-orthonormal, m = monomials deg <= p."""
    qp, qw = duffy_triangle(2 * p + 1)
    m = triangle_monomials(p, qp)
    G = (m * qw[:, None]).T @ m
    L = np.linalg.cholesky(G)
    return np.linalg.inv(L).T


def triangle_monomials(p: int, xy):
    xy = np.atleast_2d(xy)
    cols = []
    for tot in range(p + 1):
        for a in range(tot, -1, -1):
            cols.append(xy[:, 0] ** a * xy[:, 1] ** (tot - a))
    return np.stack(cols, axis=1)


def flux_1d_ref(p: int, xi):
    """1D flux basis on [0,1]: [left end, right end, bubbles k<p]: values, d/dxi."""
    xi = np.atleast_1d(xi)
    L, dL = legendre01(p, xi)
    n = len(xi)
    V = np.zeros((n, p + 2))
    D = np.zeros((n, p + 2))
    V[:, 0] = 1.0 - xi
    D[:, 0] = -1.0
    V[:, 1] = xi
    D[:, 1] = 1.0
    bub = xi * (1.0 - xi)
    dbub = 1.0 - 2.0 * xi
    for k in range(p):
        V[:, 2 + k] = bub * L[:, k]
        D[:, 2 + k] = dbub * L[:, k] + bub * dL[:, k]
    return V, D
