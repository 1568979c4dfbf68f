"""Pins the network oracle (oracle/nn_oracle.py): each transposed step is the
exact adjoint of its forward step, a centred delta kernel is the identity,
the bias-free ReLU network is positively homogeneous, and the backward pass
matches central finite differences of the forward pass."""
import numpy as np
import pytest

from oracle import nn_oracle as N


def _rand(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)


@pytest.mark.parametrize("dims", [(5, 4, 7), (6, 6, 6), (3, 2, 5)])
@pytest.mark.parametrize("stride", [1, 2])
def test_conv_adjoint(dims, stride):
    out_dims = dims if stride == 1 else N.coarse_dims(dims)
    X = _rand(dims + (3,), 1)
    W = _rand((4, 27, 3), 2)
    G = _rand(out_dims + (4,), 3)
    lhs = np.sum(N.conv(X, W, stride, out_dims) * G)
    rhs = np.sum(X * N.conv_adjoint(G, W, stride, dims))
    assert lhs == pytest.approx(rhs, rel=1e-12)
    dW = N.conv_weight_grad(X, G, stride, 4, 3)
    E = _rand(W.shape, 4)
    assert np.sum(dW * E) == pytest.approx(np.sum(N.conv(X, E, stride, out_dims) * G), rel=1e-12)


@pytest.mark.parametrize("fine", [(5, 4, 7), (6, 6, 6)])
def test_up_adjoint(fine):
    coarse = N.coarse_dims(fine)
    X = _rand(coarse + (2,), 5)
    G = _rand(fine + (2,), 6)
    assert np.sum(N.up(X, fine) * G) == pytest.approx(np.sum(X * N.up_adjoint(G, coarse)), rel=1e-12)


def test_delta_kernel_is_identity_and_stride_subsamples():
    X = _rand((5, 6, 7, 3), 7)
    W = np.zeros((3, 27, 3))
    W[:, 13, :] = np.eye(3)  # centre tap (1,1,1)
    assert np.array_equal(N.conv(X, W, 1, X.shape[:3]), X)
    Y = N.conv(X, W, 2, N.coarse_dims(X.shape[:3]))
    assert np.array_equal(Y, X[::2, ::2, ::2])
    W2 = np.zeros((3, 27, 3))
    W2[:, 14, :] = np.eye(3)  # tap (2,1,1): reads x + 1
    Y2 = N.conv(X, W2, 1, X.shape[:3])
    assert np.array_equal(Y2[:, :, :-1], X[:, :, 1:]) and not Y2[:, :, -1].any()


def test_network_is_positively_homogeneous():
    P = N.split_params(N.init_params(3))
    x = _rand((9, 9, 9), 8)
    y1, _ = N.forward(P, x)
    y2, _ = N.forward(P, 2.5 * x)
    assert np.allclose(y2, 2.5 * y1, rtol=1e-12, atol=1e-12)


def test_backward_matches_finite_differences():
    flat = N.init_params(1)
    x = _rand((7, 6, 5), 9)
    w = _rand((7, 6, 5), 10)

    def loss(f):
        y, _ = N.forward(N.split_params(f), x)
        return np.sum(w * y)
    y, cache = N.forward(N.split_params(flat), x)
    g = N.flatten(N.backward(N.split_params(flat), cache, w))
    rng = np.random.default_rng(11)
    starts = np.cumsum([0] + [27 * ci * co for _, ci, co in N.LAYERS])
    idx = np.concatenate([rng.integers(starts[i], starts[i + 1], 6) for i in range(len(N.LAYERS))])
    eps = 1e-6
    for k in idx:
        fp, fm = flat.copy(), flat.copy()
        fp[k] += eps
        fm[k] -= eps
        fd = (loss(fp) - loss(fm)) / (2 * eps)
        assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-8 * max(1.0, abs(fd))), k


def test_init_statistics():
    flat = N.init_params(3)
    P = N.split_params(flat)
    assert flat.size == N.n_params() == 360288
    for name, ci, co in N.LAYERS:
        assert np.std(P[name]) == pytest.approx(np.sqrt(2.0 / (27 * ci)), rel=0.15)


def test_bf16_rounding_is_nearest_even():
    """Against torch's own bfloat16 cast, and on exact ties."""
    import torch
    v = _rand(4096, 12) * np.exp(_rand(4096, 13) * 4)
    want = torch.tensor(v, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(N.bf16(v), want)
    ties = np.array([1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, -(1 + 2.0 ** -8), 2.0 ** -8 * (1 + 2.0 ** -8)])
    assert np.array_equal(N.bf16(ties), [1.0, 1 + 2.0 ** -6, -1.0, 2.0 ** -8])


# --- the two-level cycle and flexible CG (nn_oracle.two_grid_apply, fcg) ---

def _two_grid_problem():
    """Q2 on 4^3 deformed cells with the Q1 grid of 2^3 cells as coarse grid
    (one free coarse DoF); small enough for dense checks."""
    from oracle import mg_oracle as M
    from oracle import oracle as O
    fine = O.make_problem(2, 3, O.GLL, 4, deform=0.05)
    coarse = O.make_problem(1, 2, O.GLL, 2, deform=0.05)
    return fine, coarse, M.prolongation_3d(fine, coarse)


def _dense(prob):
    free = np.flatnonzero(prob.constrained == 0)
    A = np.zeros((free.size, free.size))
    for j, dof in enumerate(free):
        e = np.zeros(prob.N)
        e[dof] = 1.0
        A[:, j] = prob.vmult(e)[free]
    return free, 0.5 * (A + A.T)


def test_two_grid_cycle_without_network_is_the_squared_chebyshev_polynomial():
    """With all weights zero the network adds nothing (omega = 0) and the
    cycle is k Chebyshev steps from zero followed by k more from that
    iterate: the error A^-1 r - z is T(M)^2 A^-1 r, M = D^-1 A and T the
    scaled Chebyshev polynomial, evaluated through the eigenvectors of
    D^-1/2 A D^-1/2 (a dropped smoothing pass, a wrong interval or a
    misplaced Jacobi scaling fails this)."""
    from oracle import mg_oracle as M
    fine, coarse, P = _two_grid_problem()
    lmin, lmax = N.smoother_interval(fine)
    r = np.random.default_rng(3).standard_normal(fine.N)
    k = 3
    z, omega = N.two_grid_apply(fine, coarse, P, np.zeros(N.n_params()), 1.0, k, lmin, lmax, r)
    assert omega == 0.0
    free, A = _dense(fine)
    d = np.diag(A)
    S = A / np.sqrt(np.outer(d, d))
    lam, V = np.linalg.eigh(S)
    T = M.chebyshev_error_polynomial(lam, k, lmin, lmax)
    e0 = np.linalg.solve(A, r[free])
    e2 = (V * T ** 2) @ (V.T @ (np.sqrt(d) * e0)) / np.sqrt(d)
    assert np.linalg.norm((e0 - z[free]) - e2) <= 1e-11 * np.linalg.norm(e0)
    cons = fine.constrained.astype(bool)
    assert np.array_equal(z[cons], r[cons])


def test_two_grid_correction_is_energy_optimal():
    """With a nonzero network the correction omega c lowers the energy norm
    of the error, and omega is its minimiser along c: moving it by 10% either
    way raises the energy."""
    fine, coarse, P = _two_grid_problem()
    lmin, lmax = N.smoother_interval(fine)
    r = np.random.default_rng(4).standard_normal(fine.N)
    r[fine.constrained.astype(bool)] = 0.0
    params = N.init_params(5)
    z, omega, z_pre, z_corr = N.two_grid_apply(fine, coarse, P, params, 0.05, 2, lmin, lmax, r,
                                               parts=True)
    free, A = _dense(fine)
    e0 = np.zeros(fine.N)
    e0[free] = np.linalg.solve(A, r[free])

    def energy(x):
        e = (e0 - x)[free]
        return e @ A @ e
    assert omega != 0.0
    assert energy(z_corr) < energy(z_pre)
    c = (z_corr - z_pre) / omega
    for f in (0.9, 1.1):
        assert energy(z_pre + f * omega * c) > energy(z_corr)


def test_two_grid_cycle_is_positively_homogeneous():
    """The network sees r / ||r|| and its output is scaled by ||r||; the rest
    of the cycle is linear: cycle(2 r) = 2 cycle(r)."""
    fine, coarse, P = _two_grid_problem()
    lmin, lmax = N.smoother_interval(fine)
    r = np.random.default_rng(6).standard_normal(fine.N)
    params = N.init_params(7)
    z1, _ = N.two_grid_apply(fine, coarse, P, params, 0.05, 2, lmin, lmax, r)
    z2, _ = N.two_grid_apply(fine, coarse, P, params, 0.05, 2, lmin, lmax, 2.0 * r)
    assert np.linalg.norm(z2 - 2.0 * z1) <= 1e-13 * np.linalg.norm(z1)


@pytest.mark.parametrize("m", [2, 4])
def test_fcg_with_a_linear_preconditioner_is_pcg(m):
    """With a fixed SPD preconditioner (Jacobi) flexible CG is PCG in exact
    arithmetic once it keeps the previous direction (m >= 2: the slot being
    replaced is not projected out): its iterates match the oracle's
    deal.II-semantics CG (orc_cg) step by step."""
    from oracle import oracle as O
    prob = O.make_problem(3, 4, O.GLL, 3, deform=0.05)
    b = O.splitmix64_uniform(8, prob.N)
    b[prob.constrained.astype(bool)] = 0.0
    dinv = prob.inverse_diagonal()
    for steps in (1, 2, 6):
        x, hist = N.fcg(prob.vmult, lambda r: dinv * r, b, steps, m=m)
        xo, res, ho = prob.cg(b, inv_diag=dinv, max_steps=steps, abs_tol=1e-300, rel_tol=1e-14)
        assert res.steps == steps
        assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo), (m, steps)
        assert np.allclose(hist, ho, rtol=1e-9)
