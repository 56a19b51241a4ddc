"""Quadrature rules (oracle; test infrastructure only).

Restates
* ``gauss_legendre_1d<Order>``  (src/quadrature/gauss_legendre_1d.hpp:157-172):
  Order-point Gauss-Legendre on [0,1] (points 0.5(xi+1), weights 0.5 w).
* ``duffy_triangle<Degree>`` (src/quadrature/reference_triangle_duffy.hpp:24-61):
  order = (Degree+3)//2, points (u, (1-u) v), weights w_u w_v (1-u), u outer.
* ``make_space_time_tensor_product_rule`` (src/quadrature/tensor_product_quadrature.hpp:10-33):
  space index outer, time index inner (k = is*n_time + it).
* degrees used by the runner (algorithm/runners/detail/runner_case_execution.hpp:29-37):
  1+1D: QSpace = 2p+1 Gauss points, QTime = 2p+1 Gauss points;
  2+1D: QSpace = 4p+1 (Duffy exactness degree), QTime = 2p+1 Gauss points.

The reference tabulates the Gauss nodes as 16-digit decimals; here they come
from ``numpy.polynomial.legendre.leggauss`` (Golub-Welsch), so points agree to
~1e-16, far inside the 1e-12 assembly tolerance.  Pinned by exactness tests.
"""
from __future__ import annotations

import numpy as np


def gauss_legendre_01(order: int):
    """``order``-point Gauss-Legendre rule on [0,1]: (points[order], weights[order])."""
    x, w = np.polynomial.legendre.leggauss(order)
    return 0.5 * (x + 1.0), 0.5 * w


def duffy_triangle(degree: int):
    """Collapsed (Duffy) rule on conv{(0,0),(1,0),(0,1)}, exact to total ``degree``."""
    order = (degree + 3) // 2
    u, wu = gauss_legendre_01(order)
    pts = []
    wts = []
    for i in range(order):
        for j in range(order):
            pts.append((u[i], (1.0 - u[i]) * u[j]))
            wts.append(wu[i] * wu[j] * (1.0 - u[i]))
    return np.array(pts), np.array(wts)


def space_rule(dim: int, p: int):
    """Spatial rule used for cell and bottom (t=0) integrals at degree p."""
    if dim == 1:
        x, w = gauss_legendre_01(2 * p + 1)
        return x[:, None], w
    return duffy_triangle(4 * p + 1)


def time_rule(p: int):
    return gauss_legendre_01(2 * p + 1)


def cell_rule(dim: int, p: int):
    """Space-time tensor rule: points [nq, dim+1] (space coords then t), weights [nq]."""
    xs, ws = space_rule(dim, p)
    ts, wt = time_rule(p)
    ns, nt = len(ws), len(wt)
    pts = np.zeros((ns * nt, dim + 1))
    wts = np.zeros(ns * nt)
    for i in range(ns):
        for j in range(nt):
            k = i * nt + j
            pts[k, :dim] = xs[i]
            pts[k, dim] = ts[j]
            wts[k] = ws[i] * wt[j]
    return pts, wts
