"""Pins the oracle's Integrator weights (libcntr/cntr/integration.hpp:128-514).

Checks: the reference's own decimal tables (tests/golden/integration_tables.json,
extracted from libcntr/cntr/integration.cpp:152-1184), the known answers of
libcntr/test/integration.cpp:14-148, and exactness properties that a wrong index or
sign would break.
"""
import json
import math
import os

import numpy as np
import pytest

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "integration_tables.json")))


@pytest.mark.parametrize("k", range(0, 6))
def test_gregory_tables_match_reference(orc, k):
    t = orc.integrator_tables(k)
    k1, k2 = k + 1, 2 * (k + 1)
    g = t["greg"].reshape(k2, k2)
    ref = GOLD["gregory"][str(k)]
    start = np.array(ref["start"]).reshape(k1, k1)
    sigma = np.array(ref["sigma"]).reshape(k1, k1)
    omega = np.array(ref["omega"])
    np.testing.assert_allclose(t["omega"], omega, rtol=0, atol=2e-16 * 4)
    np.testing.assert_allclose(g[:k1, :k1], start, rtol=0, atol=1e-15)
    for n in range(k1):
        np.testing.assert_allclose(g[k1 + n, :k1], sigma[n], rtol=0, atol=1e-15)
        for i in range(k1, k1 + n + 1):
            assert abs(g[k1 + n, i] - omega[k1 + n - i]) < 1e-15


@pytest.mark.parametrize("k", range(0, 6))
def test_bd_weights_match_reference(orc, k):
    t = orc.integrator_tables(k)
    np.testing.assert_allclose(t["bd"], GOLD["bd"][str(k + 1)], rtol=0, atol=1e-14)


@pytest.mark.parametrize("k", range(2, 6))
def test_rcorr_matches_reference(orc, k):
    t = orc.integrator_tables(k)
    np.testing.assert_allclose(t["rcorr"], GOLD["rcorr"][str(k)], rtol=0, atol=1e-15)


def test_gregory_weight_accessor_rule(orc):
    # integration.hpp:382-390: simple form for n >= 2k+2
    k = 5
    om = orc.integrator_tables(k)["omega"]
    n = 20
    for j in range(n + 1):
        w = orc.gregory_weight(k, n, j)
        exp = om[j] if j <= k else (om[n - j] if j >= n - k else 1.0)
        assert w == exp
    # exactness: Gregory rule integrates polynomials of degree <= k exactly
    for n in range(1, 30):
        for p in range(k + 1):
            s = sum(orc.gregory_weight(k, n, j) * j ** p for j in range(max(n, k) + 1))
            assert abs(s - n ** (p + 1) / (p + 1)) < 1e-9 * max(1, n ** (p + 1))


def _poly_eval(P, k, x, f):
    return sum(f[l] * sum(P[a * (k + 1) + l] * x ** a for a in range(k + 1)) for l in range(k + 1))


def test_known_answers_reference_integration_cpp(orc):
    # libcntr/test/integration.cpp:14-148, k=5, h=0.1, eps=1e-4
    k, h, eps = 5, 0.1, 1e-4
    t = orc.integrator_tables(k)
    P, D, W, bd = t["P"], t["D"].reshape(6, 6), t["W"].reshape(6, 6, 6), t["bd"]
    ft = [math.cos(h * i) for i in range(k + 1)]
    err = sum(abs(math.cos((i + 0.5) * h) - _poly_eval(P, k, i + 0.5, ft)) for i in range(k))
    assert err < eps
    err = sum(abs(-math.sin(h * i) - sum(D[i, l] * ft[l] for l in range(6)) / h) for i in range(6))
    assert err < eps
    err = 0.0
    for i in range(1, k + 1):
        for j in range(i):
            ia = sum(h * W[i, j, l] * ft[l] for l in range(6))
            err += abs(math.sin(h * j) - math.sin(h * i) - ia)
    assert err < eps
    t1 = 0.5
    df = sum(bd[l] * math.cos(t1 - l * h) / h for l in range(k + 2))
    assert abs(df + math.sin(t1)) < eps
    # rcorr known answers (test/integration.cpp:102-121)
    ntau, beta, e1, e2 = 100, 5.0, -0.1, 0.2
    dtau = beta / ntau

    def fermi_exp(b, tau, w):
        f = lambda x: 0.0 if x * b > 100 else (1.0 if x * b < -100 else 1.0 / (1.0 + math.exp(x * b)))
        return math.exp(w * tau) * f(w) if w < 0 else math.exp((tau - b) * w) * f(-w)

    A = [-fermi_exp(beta, m * dtau, -e1) for m in range(ntau + 1)]
    B = [-fermi_exp(beta, m * dtau, -e2) for m in range(ntau + 1)]
    R = {1: 0.0137659, 2: 0.0274638, 3: 0.0410948, 4: 0.0546598}
    rc = t["rcorr"].reshape(k - 1, 6, 6)
    err = 0.0
    for m in range(1, k):
        r = sum(dtau * rc[m - 1, j, l] * A[j] * B[l] for j in range(6) for l in range(6))
        err += abs(R[m] - r)
    assert err < eps
    # integrate(): sum_i |sin(hi) - h*I.integrate(cos, i)|
    nt, hh = 100, 0.1
    f = [math.cos(hh * i) for i in range(nt + 1)]
    err = sum(abs(math.sin(hh * i) - hh * sum(orc.gregory_weight(k, i, j) * f[j] for j in range(i + 1)))
              for i in range(k + 1, nt + 1))
    assert err < eps
