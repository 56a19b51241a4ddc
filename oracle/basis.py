"""Lagrange bases on the reference segment, triangle, quadrilateral and prism (oracle).

Restates
* equidistant segment nodes i/p (src/finite_element/nodes/segment_equidistant.hpp:34-45);
* equidistant triangle nodes, ``c`` outer loop, ``b`` inner, point (b/p, c/p),
  barycentric tuple (a,b,c) with a = p-b-c
  (src/finite_element/nodes/triangle_equidistant.hpp:39-53);
* prism index ``i_space*(Q+1) + i_time`` (src/finite_element/basis/functions/
  triangular_prism_lagrange.hpp:32-35,44-85);
* quadrilateral (1+1D) index ``i_space*(Q+1) + i_time``
  (src/finite_element/basis/functions/quadrilateral_lagrange.hpp:29-60).

The reference builds the nodal basis through a Dubiner-Vandermonde QR inverse
(src/finite_element/basis/polynomials/triangular_lagrange.hpp:31-63); the
nodal basis is unique, so here it is written in closed form as the product of
barycentric factors prod_{m<a} (p*lam - m)/(m+1).  Pinned by the Kronecker
property at the nodes and the partition of unity (tests/test_oracle_tables.py).
"""
from __future__ import annotations

import numpy as np


def _factor(p: int, a: int, s):
    """prod_{m<a} (p s - m)/(m+1) and its derivative in s."""
    s = np.asarray(s, dtype=float)
    val = np.ones_like(s)
    der = np.zeros_like(s)
    for m in range(a):
        f = (p * s - m) / (m + 1.0)
        der = der * f + val * (p / (m + 1.0))
        val = val * f
    return val, der


def segment_nodes(p: int):
    return np.arange(p + 1) / p


def segment_lagrange(p: int, s):
    """values [n, p+1], derivatives [n, p+1] of the equidistant Lagrange basis."""
    s = np.atleast_1d(np.asarray(s, dtype=float))
    nodes = segment_nodes(p)
    vals = np.ones((len(s), p + 1))
    ders = np.zeros((len(s), p + 1))
    for i in range(p + 1):
        v = np.ones_like(s)
        d = np.zeros_like(s)
        for j in range(p + 1):
            if j == i:
                continue
            f = (s - nodes[j]) / (nodes[i] - nodes[j])
            d = d * f + v / (nodes[i] - nodes[j])
            v = v * f
        vals[:, i] = v
        ders[:, i] = d
    return vals, ders


def triangle_tuples(p: int):
    """Barycentric tuples (a,b,c) in the reference node order."""
    out = []
    for c in range(p + 1):
        for b in range(p + 1 - c):
            out.append((p - b - c, b, c))
    return out


def triangle_nodes(p: int):
    return np.array([(b / p, c / p) for (_, b, c) in triangle_tuples(p)])


def triangle_lagrange(p: int, xy):
    """values [n, N], gradients [n, N, 2] of the P_p Lagrange basis on the reference triangle."""
    xy = np.atleast_2d(np.asarray(xy, dtype=float))
    l1 = xy[:, 0]
    l2 = xy[:, 1]
    l0 = 1.0 - l1 - l2
    tuples = triangle_tuples(p)
    N = len(tuples)
    vals = np.zeros((len(xy), N))
    grads = np.zeros((len(xy), N, 2))
    for k, (a, b, c) in enumerate(tuples):
        fa, da = _factor(p, a, l0)
        fb, db = _factor(p, b, l1)
        fc, dc = _factor(p, c, l2)
        vals[:, k] = fa * fb * fc
        # d/dx: l0 -> -1, l1 -> +1 ; d/dy: l0 -> -1, l2 -> +1
        grads[:, k, 0] = -da * fb * fc + fa * db * fc
        grads[:, k, 1] = -da * fb * fc + fa * fb * dc
    return vals, grads


def spatial_lagrange(dim: int, p: int, x):
    """Spatial P_p basis: values [n, Ns], gradients [n, Ns, dim]."""
    if dim == 1:
        v, d = segment_lagrange(p, np.asarray(x).reshape(-1))
        return v, d[:, :, None]
    return triangle_lagrange(p, x)


def spacetime_lagrange(dim: int, p: int, pts):
    """Tensor space x time basis at reference points [n, dim+1].

    Returns values [n, Ns*(p+1)] and gradients [n, Ns*(p+1), dim+1] (spatial
    components then d/dtau), index = i_space*(p+1) + i_time.
    """
    pts = np.atleast_2d(np.asarray(pts, dtype=float))
    sv, sg = spatial_lagrange(dim, p, pts[:, :dim])
    tv, td = segment_lagrange(p, pts[:, dim])
    n = len(pts)
    Ns = sv.shape[1]
    Nt = p + 1
    vals = (sv[:, :, None] * tv[:, None, :]).reshape(n, Ns * Nt)
    grads = np.zeros((n, Ns, Nt, dim + 1))
    grads[..., :dim] = sg[:, :, None, :] * tv[:, None, :, None]
    grads[..., dim] = sv[:, :, None] * td[:, None, :]
    return vals, grads.reshape(n, Ns * Nt, dim + 1)


def legendre01(p: int, s):
    """Shifted Legendre L_k(2s-1), k=0..p: values [n, p+1], derivatives [n, p+1].

    Same family as detail::shifted_legendre_family
    (src/finite_element/error_fespace/patch_dof_map.hpp:60-80).
    """
    s = np.atleast_1d(np.asarray(s, dtype=float))
    x = 2.0 * s - 1.0
    P = np.zeros((len(s), p + 1))
    D = np.zeros((len(s), p + 1))
    P[:, 0] = 1.0
    if p >= 1:
        P[:, 1] = x
        D[:, 1] = 1.0
    for k in range(1, p):
        P[:, k + 1] = ((2 * k + 1) * x * P[:, k] - k * P[:, k - 1]) / (k + 1)
        D[:, k + 1] = D[:, k - 1] + (2 * k + 1) * P[:, k]
    return P, 2.0 * D
