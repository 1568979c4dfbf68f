"""Pins the oracle's mesh, DoF map, geometry, operator, diagonal and
right-hand side by counts, closed forms, finite differences and an assembled
matrix built independently in numpy."""
import numpy as np
import pytest

from oracle import oracle as O
from test_oracle_kernels import dense_element_matrix

# DoF counts of the reference's mesh cycles (KAT-2):
# bakeoff_problems_dealii/tests/scaling_bp35_kokkos_kernel/results/4_GPU.out:802 and
# CEED_bp/results/1xGH200_P4.txt:635-648
KAT2 = {8: 18513, 9: 35937, 10: 70785, 11: 139425, 12: 274625, 13: 545025, 14: 1081665,
        15: 2146689, 16: 4276737, 17: 8520321, 18: 16974593, 19: 33883137, 20: 67634433,
        21: 135005697}


def cycle_mesh(c):
    nref, rem = c // 3, c % 3
    n = tuple((2 if d < rem else 1) * (1 << nref) for d in range(3))
    hi = tuple(2.8 if d < rem else 0.9 for d in range(3))
    return n, (-1.0, -1.0, -1.0), hi


@pytest.mark.parametrize("cycle", sorted(KAT2))
def test_dof_counts_kat2(cycle):
    n, lo, hi = cycle_mesh(cycle)
    b = O.make_box(n, lo, hi)
    assert O.lib().orc_box_n_dofs(O.C.byref(b), 4) == KAT2[cycle]


@pytest.mark.parametrize("p", [1, 2, 4])
def test_map_nodes_and_constraints(p):
    n = (3, 2, 4)
    P = O.make_problem(p, p + 1, O.GLL, n, deform=0.05)
    nm = p + 1
    # every unconstrained DoF is reached, constrained ones never
    valid = P.map[P.map != O.INVALID]
    reached = np.zeros(P.N, bool)
    reached[valid] = True
    assert np.array_equal(reached, P.constrained == 0)
    # boundary count N - prod(n_d p - 1)
    assert int(P.constrained.sum()) == P.N - np.prod([k * p - 1 for k in n])
    # the physical node of map(e, l) is cell e's l-th lexicographic node
    coords = P.dof_coordinates()
    sup = O.gll_rule(nm)[0]
    cx = P.cellX.reshape(P.E, 8, 3)
    for e in [0, 5, P.E - 1]:
        for l in range(nm ** 3):
            g = P.map[e * nm ** 3 + l]
            if g == O.INVALID:
                continue
            xi = (sup[l % nm], sup[(l // nm) % nm], sup[l // nm ** 2])
            w = [(1 - xi[0]) if not (v & 1) else xi[0] for v in range(8)]
            phi = np.array([((xi[0] if v & 1 else 1 - xi[0]) * (xi[1] if v & 2 else 1 - xi[1]) *
                             (xi[2] if v & 4 else 1 - xi[2])) for v in range(8)])
            assert np.allclose(phi @ cx[e], coords[g], atol=1e-14)
            del w


def test_geometry_affine_closed_form():
    n = (2, 3, 4)
    lo, hi = (0.0, -1.0, 0.5), (2.0, 0.0, 3.5)
    p, nq = 3, 5
    P = O.make_problem(p, nq, O.GAUSS, n, lo=lo, hi=hi)
    h = [(hi[d] - lo[d]) / n[d] for d in range(3)]
    _, wq = O.gauss_rule(nq)
    w3 = np.einsum("z,y,x->zyx", wq, wq, wq).ravel()
    G = P.G.reshape(P.E, 6, nq ** 3)
    det = h[0] * h[1] * h[2]
    assert np.allclose(G[:, 0], w3 * det / h[0] ** 2, rtol=1e-14)
    assert np.allclose(G[:, 3], w3 * det / h[1] ** 2, rtol=1e-14)
    assert np.allclose(G[:, 5], w3 * det / h[2] ** 2, rtol=1e-14)
    assert np.abs(G[:, [1, 2, 4]]).max() < 1e-15


def test_geometry_sheared_cell_and_orientation():
    """A general affine (sheared, rotated) cell: G = w det(J) J^-1 J^-T with
    J = M diag(h); catches A^T A in place of A A^T and swapped axes."""
    nq = 3
    _, _, xq, wq = O.basis_tables(2, nq, O.GAUSS)
    M = np.array([[1.0, 0.4, -0.2], [0.1, 0.9, 0.3], [-0.3, 0.2, 1.2]])
    ref = np.array([[(v >> d) & 1 for d in range(3)] for v in range(8)], dtype=float)
    X = ref @ M.T + np.array([0.3, -0.1, 0.7])
    G = np.zeros(6 * nq ** 3)
    O.lib().orc_geometry(1, X.ravel().copy(), nq, xq, wq, G)
    Ainv = np.linalg.inv(M)          # J = M for the unit reference cube
    GG = np.linalg.det(M) * Ainv @ Ainv.T
    w3 = np.einsum("z,y,x->zyx", wq, wq, wq).ravel()
    comps = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    Gr = G.reshape(6, nq ** 3)
    for c, (a, b) in enumerate(comps):
        assert np.allclose(Gr[c], w3 * GG[a, b], rtol=1e-13, atol=1e-15)
    # the transposed product differs for this cell
    wrong = np.linalg.det(M) * Ainv.T @ Ainv
    assert not np.allclose(wrong, GG)


def test_geometry_trilinear_finite_difference():
    P = O.make_problem(2, 3, O.GLL, (2, 2, 2), deform=0.08)
    nq = 3
    cx = P.cellX.reshape(P.E, 8, 3)

    def xmap(e, xi):
        phi = np.array([((xi[0] if v & 1 else 1 - xi[0]) * (xi[1] if v & 2 else 1 - xi[1]) *
                         (xi[2] if v & 4 else 1 - xi[2])) for v in range(8)])
        return phi @ cx[e]
    G = P.G.reshape(P.E, 6, nq ** 3)
    e, eps = 5, 1e-6
    for q in [0, 13, 26]:
        xi = np.array([P.xq[q % 3], P.xq[(q // 3) % 3], P.xq[q // 9]])
        J = np.zeros((3, 3))
        for j in range(3):
            d = np.zeros(3); d[j] = eps
            J[:, j] = (xmap(e, xi + d) - xmap(e, xi - d)) / (2 * eps)
        A = np.linalg.inv(J)
        w = P.wq[q % 3] * P.wq[(q // 3) % 3] * P.wq[q // 9]
        GG = w * np.linalg.det(J) * A @ A.T
        comps = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
        for c, (a, b) in enumerate(comps):
            assert abs(G[e, c, q] - GG[a, b]) < 1e-8


def assembled_matrix(P):
    """Dense global matrix from element matrices with the reference's
    constraint treatment: masked rows/columns, identity on constrained DoFs."""
    nm3 = (P.p + 1) ** 3
    A = np.zeros((P.N, P.N))
    B = P.B if not P.collocation else np.eye(P.nq).ravel()
    for e in range(P.E):
        K = dense_element_matrix(P.p, P.nq, B, P.D, P.G[e * 6 * P.nq ** 3:(e + 1) * 6 * P.nq ** 3])
        m = P.map[e * nm3:(e + 1) * nm3]
        ok = m != O.INVALID
        idx = m[ok].astype(np.int64)
        A[np.ix_(idx, idx)] += K[np.ix_(ok, ok)]
    c = P.constrained.astype(bool)
    A[c, :] = 0
    A[:, c] = 0
    A[c, c] = 1.0
    return A


@pytest.mark.parametrize("p,nq,quad", [(1, 2, O.GLL), (2, 3, O.GLL), (3, 4, O.GLL),
                                       (2, 4, O.GAUSS), (3, 5, O.GAUSS)])
def test_vmult_and_diagonal_match_assembled_matrix(p, nq, quad):
    P = O.make_problem(p, nq, quad, (2, 3, 2), deform=0.05)
    A = assembled_matrix(P)
    x = O.splitmix64_uniform(3, P.N)
    y = P.vmult(x, threads=2)
    assert np.allclose(y, A @ x, rtol=1e-12, atol=1e-12 * np.abs(A @ x).max())
    assert np.allclose(P.diagonal(), np.diag(A), rtol=1e-12)
    c = P.constrained.astype(bool)
    assert np.array_equal(y[c], x[c])                       # identity rows
    assert np.allclose(A, A.T)                              # symmetric
    assert np.linalg.eigvalsh(A).min() > 0                  # positive definite


def test_rhs_f1_integrates_volume():
    """With no constraint zeroing, sum_i b_i = int_Omega 1 = volume; here the
    constrained entries are zeroed, so compare against the free part computed
    from the volume minus an independent boundary-node sum."""
    for quad, nq, p in [(O.GAUSS, 4, 2), (O.GLL, 3, 2)]:
        P = O.make_problem(p, nq, quad, (2, 2, 3), lo=(0, 0, 0), hi=(2.0, 1.0, 1.5))
        b = P.rhs(0)
        # full (unconstrained) load vector via an all-valid map
        full_map = P.map.copy()
        nm = p + 1
        Nx, Ny = 2 * p + 1, 2 * p + 1
        for e in range(P.E):
            ex, ey, ez = e % 2, (e // 2) % 2, e // 4
            for l in range(nm ** 3):
                a, bb, c = l % nm, (l // nm) % nm, l // nm ** 2
                full_map[e * nm ** 3 + l] = (ex * p + a) + Nx * ((ey * p + bb) + Ny * (ez * p + c))
        bf = np.zeros(P.N)
        O.lib().orc_rhs(nm, P.nq, P.B, P.xq, P.wq, P.cellX, full_map, P.E, P.N, 0, bf)
        assert abs(bf.sum() - 3.0) < 1e-13
        free = P.constrained == 0
        assert np.allclose(b[free], bf[free], rtol=1e-14)
        assert np.all(b[~free] == 0)
