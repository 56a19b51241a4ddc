"""Oracle tables pinned by closed forms (no GPU).

Gauss-Legendre / Duffy exactness (src/quadrature/reference_triangle_duffy.hpp:16-20),
Lagrange Kronecker property and partition of unity
(src/finite_element/nodes/triangle_equidistant.hpp:39-53), RT moment duality.
"""
import math

import numpy as np
import pytest

from oracle import basis, quadrature, rt


@pytest.mark.parametrize("order", range(1, 13))
def test_gauss_legendre_exact(order):
    x, w = quadrature.gauss_legendre_01(order)
    for k in range(2 * order):
        assert abs(np.sum(w * x ** k) - 1.0 / (k + 1)) < 1e-14


@pytest.mark.parametrize("deg", [5, 9, 13, 17])
def test_duffy_exact(deg):
    pts, w = quadrature.duffy_triangle(deg)
    assert len(w) == ((deg + 3) // 2) ** 2
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            assert abs(np.sum(w * pts[:, 0] ** a * pts[:, 1] ** b) - exact) < 1e-14


def test_duffy_not_exact_beyond_degree():
    # a plausible wrong order (too few points) must fail: degree 2*order polynomial in u
    pts, w = quadrature.duffy_triangle(5)   # order 4: exact to 5 (total), not 9
    a = 9
    exact = math.factorial(a) / math.factorial(a + 2)
    assert abs(np.sum(w * pts[:, 0] ** a) - exact) > 1e-8


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_triangle_lagrange_kronecker(p):
    nodes = basis.triangle_nodes(p)
    v, g = basis.triangle_lagrange(p, nodes)
    assert np.allclose(v, np.eye(len(nodes)), atol=1e-13)
    rnd = np.random.default_rng(p).random((20, 2)) * 0.5
    v, g = basis.triangle_lagrange(p, rnd)
    assert np.allclose(v.sum(1), 1.0, atol=1e-13)
    assert np.allclose(g.sum(1), 0.0, atol=1e-12)
    # gradient against central differences
    h = 1e-6
    for d in range(2):
        e = np.zeros(2)
        e[d] = h
        fd = (basis.triangle_lagrange(p, rnd + e)[0] - basis.triangle_lagrange(p, rnd - e)[0]) / (2 * h)
        assert np.allclose(fd, g[:, :, d], atol=1e-7)


def test_triangle_node_order_matches_reference():
    # c outer, b inner, point (b/p, c/p): p = 2 -> (0,0),(.5,0),(1,0),(0,.5),(.5,.5),(0,1)
    nodes = basis.triangle_nodes(2)
    assert np.allclose(nodes, [[0, 0], [.5, 0], [1, 0], [0, .5], [.5, .5], [0, 1]])


@pytest.mark.parametrize("p", [1, 2, 3])
def test_prism_index_and_tensor(p):
    pts = np.array([[0.2, 0.3, 0.7], [0.1, 0.1, 0.1]])
    v, g = basis.spacetime_lagrange(2, p, pts)
    sv, sg = basis.triangle_lagrange(p, pts[:, :2])
    tv, td = basis.segment_lagrange(p, pts[:, 2])
    for i in range(sv.shape[1]):
        for j in range(p + 1):
            k = i * (p + 1) + j
            assert np.allclose(v[:, k], sv[:, i] * tv[:, j])
            assert np.allclose(g[:, k, 2], sv[:, i] * td[:, j])


@pytest.mark.parametrize("p", [1, 2, 3])
def test_rt_moment_duality(p):
    T = rt.RTTriangle(p)
    assert T.N == (p + 1) * (p + 3)
    s, ws = quadrature.gauss_legendre_01(p + 2)
    L, _ = basis.legendre01(p, s)
    for e, (ia, ib) in enumerate(rt.REF_EDGES):
        A, B = rt.REF_V[ia], rt.REF_V[ib]
        t = B - A
        nrm = np.array([t[1], -t[0]])
        val, _ = T.eval_ref(A[None, :] + s[:, None] * t[None, :])
        mom = np.einsum("q,qk,qja,a->kj", ws, L, val, nrm)
        expect = np.zeros((p + 1, T.N))
        for k in range(p + 1):
            expect[k, e * (p + 1) + k] = 1.0
        assert np.allclose(mom, expect, atol=1e-12)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_rt_divergence_in_Pp_and_gauss(p):
    # int_T div s = sum of outward edge fluxes (Gauss) for every basis function
    T = rt.RTTriangle(p)
    qp, qw = quadrature.duffy_triangle(2 * p + 2)
    _, D = T.eval_ref(qp)
    tot = qw @ D
    expect = np.zeros(T.N)
    for e in range(3):
        expect[e * (p + 1)] = 1.0     # only the L_0 moment carries net flux
    assert np.allclose(tot, expect, atol=1e-12)
