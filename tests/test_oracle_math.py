"""CPU pins for the oracle's special functions and tables.

Each check compares the oracle with something it does not compute itself:
scipy's prolate spheroidal functions, closed forms of Gauss-Legendre and
Chebyshev interpolation, and the reference's fitted residual tables
(cpp/src/dmk/aot_kernels.cpp:229-317) against the PSWF fit target
1 - int_0^{(x+1)/2} psi0 / int_0^1 psi0 (cpp/src/dmk/direct.cpp:172-174).
"""
import numpy as np
import pytest
import scipy.special as sp

BETAS = {3: 7.2462000846862793, 6: 13.739999771118164, 9: 20.736000061035156, 12: 27.870000839233398}


@pytest.mark.parametrize("digits", [3, 6, 9, 12])
def test_pswf_matches_scipy_shape(oracle_mod, digits):
    beta = BETAS[digits]
    x = np.linspace(0.0, 0.95, 20)
    s = np.array([sp.pro_ang1(0, 0, beta, v)[0] for v in x])
    o = np.array([oracle_mod.pswf(beta, v) for v in x])
    # every use in DMK is a ratio, so compare psi(x)/psi(0)
    np.testing.assert_allclose(o / o[0], s / s[0], rtol=0, atol=5e-9)


@pytest.mark.parametrize("digits", [6, 9])
def test_pswf_unit_norm_and_integral(oracle_mod, digits):
    beta = BETAS[digits]
    x, w = oracle_mod.gauss_legendre(200)
    f = np.array([oracle_mod.pswf(beta, v) for v in x])
    assert abs(np.sum(w * f * f) - 1.0) < 1e-12  # unit L2 norm on [-1, 1]
    # int_0^r psi by quadrature on [0, r] equals the Legendre antiderivative
    r = 0.37
    f2 = np.array([oracle_mod.pswf(beta, 0.5 * r * (v + 1)) for v in x])
    assert abs(0.5 * r * np.sum(w * f2) - oracle_mod.pswf_integral(beta, r)) < 1e-14
    c0, c1, c2, c4 = oracle_mod.pswf_intvals(beta)
    assert abs(c0 - oracle_mod.pswf_integral(beta, 1.0)) < 1e-14


def test_gauss_legendre_exactness(oracle_mod):
    x, w = oracle_mod.gauss_legendre(40)
    for k in range(0, 79):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.sum(w * x ** k) - exact) < 1e-14


@pytest.mark.parametrize("digits,nc", [(3, 6), (6, 12), (9, 18), (12, 25)])
def test_residual_tables_fit_pswf_target(oracle_mod, digits, nc):
    beta = BETAS[digits]
    c = oracle_mod.residual_coeffs(digits)
    assert len(c) == nc
    xs = np.linspace(-1.0, 1.0, 801)
    i1 = oracle_mod.pswf_integral(beta, 1.0)
    target = np.array([1.0 - oracle_mod.pswf_integral(beta, 0.5 * (v + 1)) / i1 for v in xs])
    p = np.polyval(c[::-1], xs)
    assert np.max(np.abs(p - target)) < 10.0 ** (-digits)
    # endpoint behaviour: pure 1/r at r -> 0, vanishing residual at r = bs
    assert abs(p[0] - 1.0) < 10.0 ** (-digits)
    assert abs(p[-1]) < 10.0 ** (-digits)


def test_beta_table(oracle_mod):
    assert oracle_mod.beta(1e-3) == 7.2462000846862793
    assert oracle_mod.beta(1e-6) == 13.739999771118164
    assert oracle_mod.beta(1e-9) == 20.736000061035156
    assert oracle_mod.beta(1e-12) == 27.870000839233398


@pytest.mark.parametrize("p", [9, 18, 28])
def test_chebyshev_tables_closed_forms(oracle_mod, p):
    c2p, p2c, vinv = oracle_mod.cheb_tables(p)
    # inverse Vandermonde at Chebyshev nodes: (2/p) T_k(x_i), halved for k = 0
    xs = np.cos(np.pi * ((p - 1 - np.arange(p)) + 0.5) / p)
    T = np.array([np.cos(k * np.arccos(xs)) for k in range(p)])  # T[k, i]
    closed = (2.0 / p) * T
    closed[0] *= 0.5
    np.testing.assert_allclose(vinv, closed, atol=1e-13)
    # p2c re-expands the parent's T_j on the child's half exactly (degree < p)
    t = np.linspace(-1, 1, 7)
    for child in (0, 7):
        for d in range(3):
            M = p2c[child, d].T  # stored column-major: M[k, j]
            side = (child >> d) & 1
            for j in (0, 3, p - 1):
                coef = M[:, j]
                val = np.polynomial.chebyshev.chebval(t, coef)
                xp = 0.5 * t + (0.5 if side else -0.5)
                assert np.max(np.abs(val - np.cos(j * np.arccos(xp)))) < 1e-11
            # child -> parent is the transpose
            np.testing.assert_array_equal(c2p[child, d], p2c[child, d].T)
