"""Pins the multigrid oracle (oracle/mg_oracle.py) by what the mathematics
fixes: transfers reproduce coarse-space functions exactly, Chebyshev steps
equal the closed-form Chebyshev error polynomial, the Lanczos estimate meets
the exact spectrum, and the V-cycle contracts independently of h."""
import numpy as np
import pytest

from oracle import mg_oracle as M
from oracle import oracle as O
from test_oracle_mesh import assembled_matrix


def _line_coords(n, p):
    """Parametric coordinate t in [0,1] of the nodes of a Q_p node line on n cells."""
    sup, _ = O.gll_rule(p + 1)
    t = [(e + sup[l]) / n for e in range(n) for l in range(p)] + [1.0]
    return np.array(t)


@pytest.mark.parametrize("nc,pc,nf,pf", [(3, 2, 3, 4), (4, 1, 4, 2), (2, 4, 2, 8), (5, 3, 5, 6),
                                          (3, 1, 6, 1), (2, 2, 4, 2), (3, 4, 6, 4),
                                          (2, 1, 4, 4), (2, 2, 8, 3), (3, 1, 3, 1 + 3)])
def test_prolongation_reproduces_coarse_polynomials(nc, pc, nf, pf):
    """A global polynomial of degree pc in the parametric coordinate lies in
    both spaces: P must map its coarse nodal values to its fine ones."""
    P = M.prolongation_1d(nc, pc, nf, pf)
    tc, tf = _line_coords(nc, pc), _line_coords(nf, pf)
    for k in range(pc + 1):
        assert np.allclose(P @ tc ** k, tf ** k, rtol=0, atol=1e-13), k
    assert np.allclose(P.sum(axis=1), 1.0, atol=1e-14)


def test_prolongation_is_nodal_injection_where_nodes_coincide():
    P = M.prolongation_1d(4, 1, 8, 1)
    for m in range(5):
        row = P[2 * m]
        assert row[m] == 1.0 and np.count_nonzero(row) == 1
    assert np.allclose(P[1], [0.5, 0.5, 0, 0, 0])


def test_prolongation_rejects_unrelated_lines():
    with pytest.raises(AssertionError):
        M.prolongation_1d(3, 1, 4, 2)   # 4 cells are no refinement of 3
    with pytest.raises(AssertionError):
        M.prolongation_1d(3, 2, 3, 2)   # same line


def test_prolongation_3d_is_exact_for_trilinear_functions():
    fine = O.make_problem(2, 3, O.GLL, (3, 2, 4), deform=0.05)
    coarse = O.make_problem(1, 2, O.GLL, (3, 2, 4), deform=0.05)
    P = M.prolongation_3d(fine, coarse)

    def nodes(prob):
        lines = [_line_coords(prob.box.n[a], prob.p) for a in range(3)]
        z, y, x = np.meshgrid(lines[2], lines[1], lines[0], indexing="ij")
        return x.ravel(), y.ravel(), z.ravel()
    xf, yf, zf = nodes(fine)
    xc, yc, zc = nodes(coarse)
    f = lambda x, y, z: 1 + 2 * x - y + 3 * z + x * y * z - 0.5 * y * z
    assert np.allclose(P @ f(xc, yc, zc), f(xf, yf, zf), atol=1e-13)


def _small_spd():
    P = O.make_problem(2, 3, O.GAUSS, (3, 3, 2), deform=0.05)
    A = assembled_matrix(P)
    return P, A


def test_chebyshev_equals_error_polynomial():
    """From x0 = 0: x_k = (I - p_k(D^-1 A)) A^-1 b with p_k the scaled
    Chebyshev polynomial; checked through the eigendecomposition of the
    symmetric D^-1/2 A D^-1/2."""
    prob, A = _small_spd()
    dinv = 1.0 / np.diag(A)
    b = O.splitmix64_uniform(11, prob.N)
    b[prob.constrained.astype(bool)] = 0.0
    lmax = 1.3 * np.linalg.eigvalsh(np.sqrt(dinv)[:, None] * A * np.sqrt(dinv)[None, :]).max()
    for degree in (1, 2, 5, 8):
        x = M.chebyshev(lambda v: A @ v, dinv, b, None, degree, lmax / 15, lmax)
        S = np.sqrt(dinv)
        lam, V = np.linalg.eigh(S[:, None] * A * S[None, :])
        pk = M.chebyshev_error_polynomial(lam, degree, lmax / 15, lmax)
        e0 = np.linalg.solve(A, b)  # error of x0 = 0
        # D^-1 A = S (S A S) S^-1, so p(D^-1 A) e0 = S V diag(p) V^T S^-1 e0
        ek = S * (V @ (pk * (V.T @ (e0 / S))))
        assert np.allclose(e0 - x, ek, rtol=1e-10, atol=1e-12 * np.abs(e0).max()), degree


def test_chebyshev_polynomial_bound():
    lmin, lmax = 0.1, 1.5
    lam = np.linspace(lmin, lmax, 2001)
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    for k in (1, 3, 5):
        # |T_k| <= 1 on [-1, 1], so |p_k| <= 1 / T_k(theta / delta) on the interval
        T = np.polynomial.chebyshev.Chebyshev.basis(k)
        assert np.abs(M.chebyshev_error_polynomial(lam, k, lmin, lmax)).max() <= (
            1.0 / abs(T(theta / delta))) * (1 + 1e-12)
        assert M.chebyshev_error_polynomial(0.0, k, lmin, lmax) == pytest.approx(1.0)


def test_lanczos_estimate_meets_spectrum():
    prob, A = _small_spd()
    dinv = 1.0 / np.diag(A)
    S = np.sqrt(dinv)
    exact = np.linalg.eigvalsh(S[:, None] * A * S[None, :]).max()
    r0 = O.splitmix64_uniform(0x5EED, prob.N)
    est12 = M.lanczos_lambda_max(lambda v: A @ v, dinv, r0, 12)
    est_full = M.lanczos_lambda_max(lambda v: A @ v, dinv, r0, prob.N)
    assert 0.85 * exact <= est12 <= exact * (1 + 1e-10)
    assert est_full == pytest.approx(exact, rel=1e-8)


def test_level_descriptors():
    d = M.level_descriptors(4, 6, O.GAUSS, 16, 2000)
    assert [(l[0], l[3][0]) for l in d] == [(4, 16), (2, 16), (1, 16), (1, 8)]
    assert [l[1] for l in d] == [6, 3, 2, 2]
    d = M.level_descriptors(1, 3, O.GAUSS, 16, 100)
    assert [(l[0], l[1], l[3][0]) for l in d] == [(1, 3, 16), (1, 2, 8), (1, 2, 4)]
    d = M.level_descriptors(3, 4, O.GLL, (6, 4, 2), 10)
    assert [(l[0], l[3]) for l in d] == [(3, (6, 4, 2)), (1, (6, 4, 2)), (1, (3, 2, 1))]


def _vcycle_rate(mg, its=25):
    """Spectral radius of I - M A on the free DoFs by power iteration."""
    prob = mg.levels[0]["problem"]
    cons = prob.constrained.astype(bool)
    e = O.splitmix64_uniform(3, prob.N)
    e[cons] = 0
    rate = 0.0
    for _ in range(its):
        e_new = e - mg.vmult(prob.vmult(e))
        e_new[cons] = 0
        rate = np.linalg.norm(e_new) / np.linalg.norm(e)
        e = e_new / np.linalg.norm(e_new)
    return rate


def test_vcycle_contracts_h_independently():
    rates = []
    for n in (4, 8):
        mg = M.Multigrid(2, 3, O.GAUSS, n, deform=0.05, coarse_max_dofs=8)
        assert mg.levels[-1]["problem"].box.n[0] == 2
        rates.append(_vcycle_rate(mg))
    assert max(rates) < 0.25, rates
    assert abs(rates[0] - rates[1]) < 0.1, rates


def test_fine_smoother_rule_is_exact_on_affine_cells():
    """Gauss q = p+1 integrates the stiffness of affine cells exactly, so a
    fine level that smooths with q = p+1 under a q = p+2 operator gives the
    same V-cycle to rounding on an undeformed mesh; on deformed cells it does
    not, and PCG still converges to the q = p+2 operator's solution."""
    r = O.splitmix64_uniform(5, M.Multigrid(4, 6, O.GAUSS, 4, coarse_max_dofs=30).levels[0]["problem"].N)
    a = M.Multigrid(4, 6, O.GAUSS, 4, coarse_max_dofs=30)
    b = M.Multigrid(4, 6, O.GAUSS, 4, coarse_max_dofs=30, fine_smoother_nq=5)
    assert b.levels[0]["smooth"].nq == 5 and b.levels[0]["problem"].nq == 6
    za, zb = a.vmult(r), b.vmult(r)
    assert np.linalg.norm(za - zb) <= 1e-11 * np.linalg.norm(za)
    a = M.Multigrid(4, 6, O.GAUSS, 4, deform=0.05, coarse_max_dofs=30)
    b = M.Multigrid(4, 6, O.GAUSS, 4, deform=0.05, coarse_max_dofs=30, fine_smoother_nq=5)
    za, zb = a.vmult(r), b.vmult(r)
    assert np.linalg.norm(za - zb) >= 1e-4 * np.linalg.norm(za)
    prob = b.levels[0]["problem"]
    rhs = prob.rhs(1)
    rhs[prob.constrained.astype(bool)] = 0
    x, steps, hist = b.pcg(rhs, rel_tol=1e-10)
    assert hist[-1] <= 1e-10 * hist[0] and steps <= 15
    xj, _, _ = prob.cg(rhs, inv_diag=prob.inverse_diagonal(), rel_tol=1e-13, abs_tol=1e-300)
    assert np.linalg.norm(x - xj) <= 1e-8 * np.linalg.norm(xj)


def test_mg_pcg_solves():
    mg = M.Multigrid(4, 6, O.GAUSS, 4, deform=0.05, coarse_max_dofs=30)
    prob = mg.levels[0]["problem"]
    b = prob.rhs(1)
    b[prob.constrained.astype(bool)] = 0
    x, steps, hist = mg.pcg(b, rel_tol=1e-10)
    assert hist[-1] <= 1e-10 * hist[0] and steps <= 15
    xj, r, _ = prob.cg(b, inv_diag=prob.inverse_diagonal(), rel_tol=1e-13, abs_tol=1e-300)
    assert np.linalg.norm(x - xj) <= 1e-8 * np.linalg.norm(xj)
    assert r.steps > 3 * steps
