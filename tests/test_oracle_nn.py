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
