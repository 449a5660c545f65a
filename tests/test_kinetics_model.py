"""Pins the synthetic kinetics callbacks (oracle/mech_oracle.c; the device
versions are compared against these in tests/test_gpu_parity.py) by what the
physics fixes, independently of the formulas being checked:

* every reaction conserves elements, hence mass: sum_k dY_k/dt = 0;
* adiabatic constant pressure conserves the mixture enthalpy:
  d/dt sum_k Y_k h_k(T)/W_k = cp dT/dt + sum_k h_k/W_k dY_k/dt = 0, with h_k
  computed here from the NASA-like polynomial by quadrature of cp;
* second law: each net rate of progress has the sign of the reaction's
  affinity -sum_k nu_k mu_k, mu_k = g_k + RT ln(C_k / C0);
* the analytic Jacobian equals central finite differences of dydt, and the
  reference's own finite-difference Jacobian (generic/fd_jacob.c:24-127).
"""
import math

import numpy as np
import pytest
from scipy.integrate import quad

from accelerint_amd import mechanism as M

R = M.R_GAS


@pytest.fixture(scope="module", params=[14, 54])
def net(request):
    return M.network(request.param)


def test_element_balance(net):
    F = net.formulas
    for j in range(net.nr):
        sl = net.slots[j]
        lhs = sum(F[s] for s in sl[:3] if s >= 0)
        rhs = sum(F[s] for s in sl[3:] if s >= 0)
        assert np.array_equal(lhs, rhs)
        assert (sl[:3] >= 0).sum() >= 2


def test_mass_and_enthalpy_conservation(oracle_lib, net):
    om = oracle_lib.Mechanism.kinetics(net)
    y, P = M.initial_conditions(net, 8, seed=3)
    a0, a1, a2, bh, bs = net.thermo.T
    for c in range(8):
        f = oracle_lib.dydt(om, 0.0, P[c], y[:, c])
        T, Y = y[0, c], y[1:, c]
        dY = f[1:]
        scale = np.abs(dY).max()
        assert abs(dY.sum()) <= 1e-12 * scale
        # enthalpy per species by quadrature of cp from 298.15 K
        h = np.array([R * (quad(lambda x, k=k: a0[k] + a1[k] * x + a2[k] * x * x, 298.15, T)[0]
                           + (a0[k] * 298.15 + a1[k] * 298.15 ** 2 / 2 + a2[k] * 298.15 ** 3 / 3 + bh[k]))
                      for k in range(net.ns)])
        cp = R * (Y * (a0 + a1 * T + a2 * T * T) / net.W).sum()
        dh = cp * f[0] + (h / net.W * dY).sum()
        assert abs(dh) <= 1e-9 * (np.abs(h / net.W * dY).sum())


def test_rates_follow_affinity(oracle_lib, net):
    om = oracle_lib.Mechanism.kinetics(net)
    y, P = M.initial_conditions(net, 4, seed=9)
    for c in range(4):
        T, Y = y[0, c], y[1:, c]
        q = oracle_lib.kinetics_rates(om, T, P[c], Y)
        S = (Y / net.W).sum()
        C = 1e-6 * P[c] / (R * T) * (Y / net.W) / S
        C0 = 1e-6 * M.P_ATM / (R * T)
        mu_RT = net.g_RT(T) + np.log(C / C0)
        for j in range(net.nr):
            sl = net.slots[j]
            aff = sum(mu_RT[s] for s in sl[:3] if s >= 0) - sum(mu_RT[s] for s in sl[3:] if s >= 0)
            if abs(aff) > 1e-6:
                assert np.sign(q[j]) == np.sign(aff)


def test_jacobian_vs_central_differences(oracle_lib, net):
    om = oracle_lib.Mechanism.kinetics(net)
    y, P = M.initial_conditions(net, 3, seed=4)
    for c in range(3):
        J = oracle_lib.jacobian(om, 0.0, P[c], y[:, c])
        Jfd = np.zeros_like(J)
        for j in range(net.n):
            h = 1e-6 * max(abs(y[j, c]), 1e-4)
            yp, ym = y[:, c].copy(), y[:, c].copy()
            yp[j] += h
            ym[j] -= h
            Jfd[:, j] = (oracle_lib.dydt(om, 0, P[c], yp) - oracle_lib.dydt(om, 0, P[c], ym)) / (2 * h)
        scale = np.abs(J).max(1, keepdims=True)
        assert (np.abs(J - Jfd) <= 1e-4 * scale).all()


def test_reference_fd_jacobian_on_van_der_pol(oracle_lib):
    """generic/fd_jacob.c:24-127 (first order) on a well-scaled problem.  On the
    kinetics networks its perturbation r0/ewt (ewt = ATOL + RTOL|y|, the
    reference multiplies where CVODE divides) is far too large for species at
    1e-7 mass fraction, so it is only checked here."""
    om = oracle_lib.Mechanism.van_der_pol()
    y = np.array([1.3, -0.4])
    J = oracle_lib.jacobian(om, 0.0, 50.0, y)
    Jfd = oracle_lib.fd_jacobian(om, 0.0, 50.0, y)
    assert np.allclose(Jfd, J, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("order", [2, 4, 6])
def test_fd_stencil_moments(oracle_lib, order):
    """Central stencils of FD_ORD 2, 4, 6 (generic/fd_jacob.c:33-66): the
    weights reproduce the derivative exactly on polynomials of degree <= order
    (moment 1 equals 1, moments 0 and 2..order vanish) and not beyond."""
    x, w = oracle_lib.fd_stencil(order)
    assert len(x) == order and np.array_equal(x, -x[::-1]) and np.array_equal(w, -w[::-1])
    for m in range(order + 1):
        assert abs(np.sum(w * x ** m) - (1.0 if m == 1 else 0.0)) <= 1e-14, m
    assert abs(np.sum(w * x ** (order + 1))) > 1e-3


@pytest.mark.parametrize("order", [1, 2, 4, 6])
def test_reference_fd_orders_on_van_der_pol(oracle_lib, order):
    """Every FD_ORD of generic/fd_jacob.c against the analytic Jacobian."""
    om = oracle_lib.Mechanism.van_der_pol()
    y = np.array([1.3, -0.4])
    J = oracle_lib.jacobian(om, 0.0, 50.0, y)
    Jfd = oracle_lib.fd_jacobian(om, 0.0, 50.0, y, order=order)
    assert np.allclose(Jfd, J, rtol=1e-5, atol=1e-6)


def test_van_der_pol_jacobian(oracle_lib):
    om = oracle_lib.Mechanism.van_der_pol()
    y = np.array([1.3, -0.4])
    mu = 250.0
    J = oracle_lib.jacobian(om, 0.0, mu, y)
    assert np.allclose(J, [[0.0, 1.0], [-2 * mu * y[0] * y[1] - 1, mu * (1 - y[0] ** 2)]])


def test_generator_is_seeded():
    a, b = M.network(14), M.network(14)
    assert np.array_equal(a.slots, b.slots) and np.array_equal(a.lnA, b.lnA)
    ya, Pa = M.initial_conditions(a, 16, seed=1)
    assert np.allclose(ya[1:].sum(0), 1.0)
    assert (ya[0] >= 1000).all() and (ya[0] <= 1600).all()
    assert (Pa >= M.P_ATM).all() and (Pa <= 25 * M.P_ATM).all()
