"""Pins the oracle's 1D rules and Lagrange basis (and the product's own
host-side tables) by mathematics and by the reference's LagrangeBasis."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O


@pytest.mark.parametrize("n", range(1, 12))
def test_gauss_exact_to_degree_2n_minus_1(n):
    x, w = O.gauss_rule(n)
    for k in range(2 * n):
        assert abs(np.sum(w * x ** k) - 1.0 / (k + 1)) < 1e-14
    # error on x^(2n) is the closed form (n!)^4 / ((2n+1) ((2n)!)^2) on [0,1]
    # (a rule with more points, or wrong nodes, gives a different number)
    from math import factorial
    err = 1.0 / (2 * n + 1) - np.sum(w * x ** (2 * n))
    exact = factorial(n) ** 4 / ((2 * n + 1) * factorial(2 * n) ** 2)
    if exact > 1e-13:
        assert abs(err - exact) <= 1e-6 * exact + 1e-15
    assert np.all(np.diff(x) > 0) and np.allclose(x + x[::-1], 1.0, atol=1e-15)


@pytest.mark.parametrize("n", range(2, 12))
def test_gll_exact_to_degree_2n_minus_3(n):
    x, w = O.gll_rule(n)
    assert x[0] == 0.0 and x[-1] == 1.0
    for k in range(2 * n - 2):
        assert abs(np.sum(w * x ** k) - 1.0 / (k + 1)) < 1e-14
    if n <= 6:  # GLL is not exact one degree higher
        assert abs(np.sum(w * x ** (2 * n - 2)) - 1.0 / (2 * n - 1)) > 1e-10


def test_gll_known_tables():
    # p = 2: nodes 0, 1/2, 1, weights 1/6, 2/3, 1/6 on [0,1] (4/3, 1/3 on [-1,1])
    x, w = O.gll_rule(3)
    assert np.allclose(x, [0, 0.5, 1], atol=1e-16) and np.allclose(w, [1 / 6, 2 / 3, 1 / 6])
    # p = 3: interior nodes (1 -+ 1/sqrt(5))/2, weights 1/12, 5/12
    x, w = O.gll_rule(4)
    assert np.allclose(x[1:3], [(1 - 5 ** -0.5) / 2, (1 + 5 ** -0.5) / 2], atol=1e-16)
    assert np.allclose(w, [1 / 12, 5 / 12, 5 / 12, 1 / 12], atol=1e-16)


def test_lagrange_bitwise_against_reference_golden(golden):
    xs = golden["lag_x"]
    for name in ["x2", "y3", "z4"] + [f"gll{n}" for n in range(2, 10)]:
        nodes = golden[f"lag_{name}_nodes"]
        val, der = golden[f"lag_{name}_value"], golden[f"lag_{name}_deriv"]
        n = len(nodes)
        for j in range(n):
            v = np.array([O.lib().orc_lagrange_value(nodes, n, j, x) for x in xs])
            d = np.array([O.lib().orc_lagrange_derivative(nodes, n, j, x) for x in xs])
            assert np.array_equal(v, val[j]) and np.array_equal(d, der[j]), name


def test_partition_of_unity_reference_kat():
    """sum_factorization/tests/test_lagrange.cc:19-46: |sum(sum_j l_j(x) - 1)| < 1e-6
    for 2/3/4-node sets on x in [-2, 0.5] step 0.05 (float)."""
    sets = [[-2.0, 0.5], [-2.0, 0.0, 1.0], [-3.0, 0.5, 1.0, 1.5]]
    x = np.float32(-2.0)
    sums = [np.float32(0)] * 3
    while x <= np.float32(0.5):
        for k, nodes in enumerate(sets):
            nd = np.array(nodes, dtype=np.float64)
            s = sum(O.lib().orc_lagrange_value(nd, len(nd), j, float(x)) for j in range(len(nd)))
            sums[k] = np.float32(sums[k] + np.float32(s) - np.float32(1.0))
        x = np.float32(x + np.float32(0.05))
    assert all(abs(s) < 1e-6 for s in sums)


@pytest.mark.parametrize("p,nq,quad", [(p, p + 1, O.GLL) for p in range(1, 9)] +
                         [(p, p + 2, O.GAUSS) for p in range(1, 9)] +
                         [(p, p + 1, O.GAUSS) for p in range(1, 9)])
def test_basis_tables_invariants(p, nq, quad):
    B, D, xq, wq = O.basis_tables(p, nq, quad)
    Bm, Dm = B.reshape(nq, p + 1), D.reshape(nq, nq)
    assert np.allclose(Bm.sum(axis=1), 1.0, atol=1e-13)           # partition of unity
    assert np.allclose(Dm @ np.ones(nq), 0.0, atol=1e-11)         # d/dx 1 = 0
    assert np.allclose(Dm @ xq, 1.0, atol=1e-11)                  # d/dx x = 1
    assert np.allclose(Dm @ xq ** (nq - 1), (nq - 1) * xq ** (nq - 2), atol=1e-9)
    sup = O.gll_rule(p + 1)[0]
    assert np.allclose(Bm @ sup ** p, xq ** p, atol=1e-12)        # interpolates x^p exactly
    if quad == O.GLL and nq == p + 1:
        assert np.array_equal(Bm, np.eye(nq))                    # collocation: B = I exactly


@pytest.mark.parametrize("p,nq,quad", [(p, p + 1, O.GLL) for p in range(1, 9)] +
                         [(p, p + 2, O.GAUSS) for p in range(1, 9)])
def test_product_tables_match_oracle(p, nq, quad):
    """libmfx's host-side tables (its own implementation) agree with the
    oracle's to rounding (host-only entry point, no GPU needed)."""
    from benchmarks_amd._lib import lib
    B = np.zeros(nq * (p + 1)); D = np.zeros(nq * nq); xq = np.zeros(nq); wq = np.zeros(nq)
    st = lib().mfx_basis_tables(p, nq, quad, B.ctypes.data, D.ctypes.data, xq.ctypes.data,
                                wq.ctypes.data)
    assert st == 0
    Bo, Do, xo, wo = O.basis_tables(p, nq, quad)
    assert np.allclose(B, Bo, rtol=0, atol=1e-14) and np.allclose(xq, xo, rtol=0, atol=1e-15)
    assert np.allclose(wq, wo, rtol=0, atol=1e-15) and np.allclose(D, Do, rtol=1e-12, atol=1e-12)
