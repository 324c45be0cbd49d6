"""Pins of the oracle's reference-element data against the reference's own
tables and tests (quadrature tables, Legendre/Lagrange known answers)."""
import json
import os

import numpy as np
import pytest

import oracle as O


def test_quadrature_order_rules():
    # src/numerics/quadrature/tools.cpp:7-24, quadrilateral.cpp:100-105
    for p in range(0, 8):
        assert O.lib().orc_nq1d_from_qorder(O.lib().orc_hex_qorder(p)) == p + 2
        assert O.lib().orc_nq1d_from_qorder(O.lib().orc_quad_face_qorder(p)) == p + 2
        # mass matrix / IC / L2 error rule: quadrature order 2p
        assert O.lib().orc_nq1d_from_qorder(O.lib().orc_hex_qorder(2 * p)) == 2 * p + 2


def test_gl_rule_matches_reference_tables(golden_dir):
    tab = json.load(open(os.path.join(golden_dir, "gl_tables.json")))
    checked = 0
    for key, ent in tab.items():
        if key == "defects":
            continue
        n = int(key)
        x, w = O.gl_rule(n)
        xr = np.array([float(v) for v in ent["pts"]])
        wr = np.array([float(v) for v in ent["wts"]])
        # the 32-digit literals round to the nearest double: agree to 1 ulp
        assert np.all(np.abs(x - xr) <= np.spacing(np.abs(xr)) + 1e-300), n
        assert np.all(np.abs(w - wr) <= np.spacing(np.abs(wr))), n
        checked += 1
    assert checked >= 23
    # the reference's 19-point node table is defective (pts(15) never set)
    assert tab["defects"]["19"]["unassigned"] == [15]


def test_gl_rule_independent_check():
    # numpy's Golub-Welsch rule as an independent route; exactness on
    # monomials up to degree 2n-1
    for n in range(1, 26):
        x, w = O.gl_rule(n)
        xn, wn = np.polynomial.legendre.leggauss(n)
        assert np.allclose(x, xn, rtol=0, atol=1e-14)
        assert np.allclose(w, wn, rtol=0, atol=1e-14)
        for k in range(0, 2 * n):
            exact = 0.0 if k % 2 else 2.0 / (k + 1)
            assert abs(np.dot(w, x ** k) - exact) < 1e-13


@pytest.mark.parametrize("dim,area", [(1, 2.0), (2, 4.0), (3, 8.0)])
def test_weights_sum_to_area(dim, area):
    # test/numerics/quadrature/test_quadrature.cpp:19-122
    for order in range(1, 20):
        n = O.lib().orc_nq1d_from_qorder(2 * order + 1)
        x, w = O.gl_rule(n)
        s = np.sum(w) ** dim if dim > 1 else np.sum(w)
        assert abs(s - area) < 1e-13


def test_equidistant_nodes():
    # test/numerics/basis/test_basis.cpp:25-38; test/numerics/test_nodes.cpp:11-25
    x = np.zeros(2)
    O.lib().orc_equidistant_nodes(0.0, 10.0, 2, x)
    assert list(x) == [0.0, 10.0]
    x = np.zeros(3)
    O.lib().orc_equidistant_nodes(-1.0, 1.0, 3, x)
    assert list(x) == [-1.0, 0.0, 1.0]


def _legendre_fixture(golden_dir):
    return json.load(open(os.path.join(golden_dir, "legendre_p4.json")))


def test_legendre_3d_p4_known_answers(golden_dir):
    g = _legendre_fixture(golden_dir)
    nb = 125
    val, grad = O.basis_3d("LegendreHex", 4, np.array(g["points_3d"]))
    ev = np.array(g["values_3d"]).reshape(2, nb)
    eg = np.array(g["grads_3d"]).reshape(3, 2, nb)
    assert np.max(np.abs(val - ev)) < g["tolerance"]
    # all three gradient components (the reference checks only x and y)
    for d in range(3):
        assert np.max(np.abs(grad[:, :, d] - eg[d])) < g["tolerance"], d


def test_legendre_2d_p4_known_answers(golden_dir):
    # 2D vectors: the hex basis restricted to z = 0 slices is the quad basis
    # times P_k(0); take k = 0 (P_0 = 1) columns.
    g = _legendre_fixture(golden_dir)
    pts = np.array([[x, y, 0.0] for x, y in g["points_2d"]])
    val, grad = O.basis_3d("LegendreHex", 4, pts)
    ev = np.array(g["values_2d"]).reshape(2, 25)
    eg = np.array(g["grads_2d"]).reshape(2, 2, 25)
    assert np.max(np.abs(val[:, :25] - ev)) < g["tolerance"]
    for d in range(2):
        assert np.max(np.abs(grad[:, :25, d] - eg[d])) < g["tolerance"]


def test_legendre_closed_forms():
    # test_basis.cpp:308-357 (p=1 exact) and orthogonality on a GL rule
    v, gr = O.basis_1d("LegendreHex", 1, -1.0)
    assert list(v) == [1.0, -1.0] and list(gr) == [0.0, 1.0]
    x, w = O.gl_rule(10)
    for p in range(1, 8):
        V = np.array([O.basis_1d("LegendreHex", p, xi)[0] for xi in x])
        M = V.T @ np.diag(w) @ V
        assert np.allclose(M, np.diag(2.0 / (2 * np.arange(p + 1) + 1)), atol=1e-14)
        # derivative check by central differences of the polynomial
        h = 1e-6
        for xi in (-0.7, 0.1, 0.9):
            d = (O.basis_1d("LegendreHex", p, xi + h)[0] - O.basis_1d("LegendreHex", p, xi - h)[0]) / (2 * h)
            assert np.allclose(O.basis_1d("LegendreHex", p, xi)[1], d, atol=1e-7)


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_lagrange_hex_nodal_and_gradients(p):
    # test_basis.cpp:217-303: nodal property; p=1 gradients
    n = p + 1
    xn = np.linspace(-1, 1, n)
    pts = np.array([[xn[i], xn[j], xn[k]] for k in range(n) for j in range(n) for i in range(n)])
    val, grad = O.basis_3d("LagrangeHex", p, pts)
    assert np.allclose(val, np.eye(n ** 3), atol=1e-13)
    # gradient vs finite differences at a generic point
    x0 = np.array([[0.13, -0.41, 0.77]])
    _, g0 = O.basis_3d("LagrangeHex", p, x0)
    h = 1e-6
    for d in range(3):
        xp = x0.copy(); xp[0, d] += h
        xm = x0.copy(); xm[0, d] -= h
        fd = (O.basis_3d("LagrangeHex", p, xp)[0] - O.basis_3d("LagrangeHex", p, xm)[0]) / (2 * h)
        assert np.allclose(g0[0, :, d], fd[0], atol=1e-7)
    if p == 1:
        v, gr = O.basis_1d("LagrangeHex", 1, -1.0)
        assert np.allclose(gr, [-0.5, 0.5])
