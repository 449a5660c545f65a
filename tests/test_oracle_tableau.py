"""Pins the Radau IIA(5) constants (oracle radau_oracle.c and the HIP header
radau_core.hpp, both restating radau2a.c:101-231) against closed forms:

* c = ((4 - sqrt 6)/10, (4 + sqrt 6)/10, 1) -- the Radau IIA collocation nodes;
* A_ij = integral_0^{c_i} l_j(s) ds for the Lagrange basis l_j on c;
* gamma, alpha +- i beta = eigenvalues of A^{-1};
* T Tinv = I and Tinv A^{-1} T = [[gamma, 0, 0], [0, alpha, -beta], [0, beta, alpha]];
* TinvAinv = T^{-1} A^{-1};
* e_1..e_3 = e_0 * (-(13 + 7 sqrt 6)/3, (-13 + 7 sqrt 6)/3, -1/3) (RADAU5's
  DD1..DD3) and the estimator e_0 phi(0) + sum_k w_k phi(c_k), w = A^T e_{1..3},
  annihilates polynomials of degree <= 2 (local error O(h^4), rkELO = 4).
A transposed or mistyped constant fails one of these.
"""
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def closed_form():
    s6 = math.sqrt(6.0)
    c = np.array([(4 - s6) / 10, (4 + s6) / 10, 1.0])
    A = np.zeros((3, 3))
    for j in range(3):
        others = [c[m] for m in range(3) if m != j]
        # l_j(s) = prod (s - c_m)/(c_j - c_m)
        poly = np.poly1d([1.0])
        for cm in others:
            poly = poly * np.poly1d([1.0, -cm]) / (c[j] - cm)
        ip = poly.integ()
        for i in range(3):
            A[i, j] = ip(c[i]) - ip(0.0)
    return c, A


def check_tableau(tab, tol=2e-15):
    c, A = closed_form()
    assert np.allclose(tab["c"], c, rtol=0, atol=tol)
    assert np.allclose(tab["A"], A, rtol=0, atol=4 * tol)
    # stiff accuracy: b = last row of A; row sums = c
    assert np.allclose(A.sum(1), c, atol=tol)
    Ainv = np.linalg.inv(tab["A"])
    ev = np.linalg.eigvals(Ainv)
    real = ev[np.abs(ev.imag) < 1e-9].real
    cplx = ev[ev.imag > 1e-9]
    assert real.size == 1 and cplx.size == 1
    assert abs(real[0] - tab["gamma"]) < 1e-13
    assert abs(cplx[0].real - tab["alpha"]) < 1e-13
    assert abs(cplx[0].imag - tab["beta"]) < 1e-13
    T = tab["T"]
    Tinv = np.linalg.inv(T)
    D = Tinv @ Ainv @ T
    expect = np.array([[tab["gamma"], 0, 0], [0, tab["alpha"], -tab["beta"]], [0, tab["beta"], tab["alpha"]]])
    assert np.abs(D - expect).max() < 1e-12
    assert np.abs(tab["TinvAinv"] - Tinv @ Ainv).max() < 1e-12
    s6 = math.sqrt(6.0)
    e = tab["e"]
    dd = np.array([-(13 + 7 * s6) / 3, (-13 + 7 * s6) / 3, -1.0 / 3.0])
    assert np.allclose(e[1:] / e[0], dd, rtol=1e-15, atol=0)
    # estimator annihilates polynomials of degree <= 2
    w = A.T @ e[1:]
    for p in range(3):
        val = e[0] * (1.0 if p == 0 else 0.0) + (w * c ** p).sum()
        assert abs(val) < 1e-14, (p, val)
    # degree 3 is not annihilated: the estimate is genuinely O(h^4)
    assert abs((w * c ** 3).sum()) > 1e-3
    assert tab["ELO"] == 4


def test_oracle_tableau_closed_forms(oracle_lib):
    check_tableau(oracle_lib.tableau())


def _header_constants():
    src = open(os.path.join(ROOT, "accelerint_amd", "csrc", "radau_core.hpp")).read()
    vals = {}
    for name, expr in re.findall(r"constexpr double (rk\w+) = ([^;]+);", src):
        if re.fullmatch(r"[-+0-9.eE* ]+", expr.strip()):
            vals[name] = eval(expr)  # numeric literal arithmetic only
    return vals


def test_device_tableau_closed_forms():
    v = _header_constants()
    tab = {
        "A": np.array([[v["rkA%d%d" % (i, j)] for j in range(3)] for i in range(3)]),
        "c": np.array([v["rkC0"], v["rkC1"], v["rkC2"]]),
        "e": np.array([v["rkE0"], v["rkE1"], v["rkE2"], v["rkE3"]]),
        "gamma": v["rkGamma"], "alpha": v["rkAlpha"], "beta": v["rkBeta"],
        "T": np.array([[v["rkT%d%d" % (i, j)] for j in range(3)] for i in range(3)]),
        "TinvAinv": np.array([[v["rkTiAi%d%d" % (i, j)] for j in range(3)] for i in range(3)]),
        "ELO": v["rkELO"],
    }
    check_tableau(tab)


def test_device_and_oracle_constants_identical(oracle_lib):
    v = _header_constants()
    tab = oracle_lib.tableau()
    assert v["rkGamma"] == tab["gamma"] and v["rkAlpha"] == tab["alpha"] and v["rkBeta"] == tab["beta"]
    assert np.array_equal(np.array([[v["rkA%d%d" % (i, j)] for j in range(3)] for i in range(3)]), tab["A"])
