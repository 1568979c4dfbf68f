"""CPU restatement (numpy, FP64) of the learned preconditioner: its network,
the two-level cycle it serves as coarse solve in, and the flexible CG
around it (the last two at the end of this file).

Test infrastructure only: tests/ import it, the product does not.

No reference counterpart: the reference has no learned component.  The
network is a 4-level 3D U-Net on the nodal grid of a structured Q_p mesh
(configuration C5: 16 -> 32 -> 64 channels, 3^3 kernels, no biases), with
additive skips:

    e0 = relu(conv_a(x))                 1 -> 16, level 0
    e1 = relu(conv_b(e0, stride 2))     16 -> 32, level 1
    e2 = relu(conv_c(e1, stride 2))     32 -> 64, level 2
    e3 = relu(conv_d(e2, stride 2))     64 -> 64, level 3
    d2 = relu(conv_e(up(e3)) + e2)      64 -> 64, level 2
    d1 = relu(conv_f(up(d2)) + e1)      64 -> 32, level 1
    d0 = relu(conv_g(up(d1)) + e0)      32 -> 16, level 0
    y  = conv_h(d0)                     16 -> 1,  level 0

Tensors are channels-last arrays [z][y][x][c]; a level has ceil(n/2) nodes
per direction of the level above.  conv (a correlation with zero padding 1):
    out[p][o] = sum_{t, c} W[o][t][c] in[s p + off_t][c],
with tap t = tx + 3 ty + 9 tz, off_t = (tx-1, ty-1, tz-1) and stride s.
up(v)[i] = v[i >> 1] per direction.  Backward is the exact adjoint of each
step (checked against finite differences in tests/test_oracle_nn.py).
"""
from __future__ import annotations

import numpy as np

LAYERS = [("a", 1, 16), ("b", 16, 32), ("c", 32, 64), ("d", 64, 64),
          ("e", 64, 64), ("f", 64, 32), ("g", 32, 16), ("h", 16, 1)]
TAPS = [(tx - 1, ty - 1, tz - 1) for tz in range(3) for ty in range(3) for tx in range(3)]


def n_params() -> int:
    return sum(27 * ci * co for _, ci, co in LAYERS)


def split_params(flat: np.ndarray) -> dict:
    """Flat parameter vector (layer order, each [cout][tap][cin]) -> dict."""
    out, k = {}, 0
    for name, ci, co in LAYERS:
        n = 27 * ci * co
        out[name] = flat[k:k + n].reshape(co, 27, ci)
        k += n
    return out


def coarse_dims(d):
    return tuple((n + 1) // 2 for n in d)


def _shift_view(X, off, stride, out_dims):
    """in[s p + off] for every output p (zero outside): array out_dims + (c,)."""
    oz, oy, ox = out_dims
    nz, ny, nx, c = X.shape
    out = np.zeros((oz, oy, ox, c), dtype=X.dtype)
    dx, dy, dz = off
    z = stride * np.arange(oz) + dz
    y = stride * np.arange(oy) + dy
    x = stride * np.arange(ox) + dx
    mz, my, mx = (z >= 0) & (z < nz), (y >= 0) & (y < ny), (x >= 0) & (x < nx)
    out[np.ix_(mz, my, mx)] = X[np.ix_(z[mz], y[my], x[mx])]
    return out


def up(X, fine_dims):
    """Nearest up-sampling: out[i] = X[i >> 1] per direction."""
    fz, fy, fx = fine_dims
    return X[np.arange(fz) >> 1][:, np.arange(fy) >> 1][:, :, np.arange(fx) >> 1]


def up_adjoint(G, coarse):
    """Adjoint of up: sum of the (1 or 2 per direction) fine children."""
    out = np.zeros(tuple(coarse) + (G.shape[3],), dtype=G.dtype)
    fz, fy, fx = G.shape[:3]
    np.add.at(out, (np.arange(fz)[:, None, None] >> 1, np.arange(fy)[None, :, None] >> 1,
                    np.arange(fx)[None, None, :] >> 1), G)
    return out


def conv(X, W, stride, out_dims):
    co = W.shape[0]
    out = np.zeros(tuple(out_dims) + (co,), dtype=np.float64)
    for t, off in enumerate(TAPS):
        out += _shift_view(X, off, stride, out_dims) @ W[:, t, :].T
    return out


def conv_adjoint(G, W, stride, in_dims):
    """d(in) for out = conv(in, W, stride): scatter of W^T G to s p + off."""
    ci = W.shape[2]
    dX = np.zeros(tuple(in_dims) + (ci,), dtype=np.float64)
    oz, oy, ox = G.shape[:3]
    nz, ny, nx = in_dims
    for t, (dx, dy, dz) in enumerate(TAPS):
        contrib = G @ W[:, t, :]  # (oz, oy, ox, ci)
        z = stride * np.arange(oz) + dz
        y = stride * np.arange(oy) + dy
        x = stride * np.arange(ox) + dx
        mz, my, mx = (z >= 0) & (z < nz), (y >= 0) & (y < ny), (x >= 0) & (x < nx)
        dX[np.ix_(z[mz], y[my], x[mx])] += contrib[np.ix_(mz, my, mx)]
    return dX


def conv_weight_grad(X, G, stride, co, ci):
    out_dims = G.shape[:3]
    dW = np.zeros((co, 27, ci))
    for t, off in enumerate(TAPS):
        S = _shift_view(X, off, stride, out_dims).reshape(-1, ci)
        dW[:, t, :] = G.reshape(-1, co).T @ S
    return dW


def bf16(v: np.ndarray) -> np.ndarray:
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    u = np.ascontiguousarray(v, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(P: dict, x: np.ndarray, bf16_storage: bool = False):
    """x: [z][y][x] field.  Returns (y field, cache of activations).  With
    bf16_storage the weights and the stored activations e0..d0 are rounded
    to bfloat16 as the GPU forward pass stores them (sums stay exact)."""
    d0 = x.shape
    d1, d2, d3 = coarse_dims(d0), coarse_dims(coarse_dims(d0)), coarse_dims(coarse_dims(coarse_dims(d0)))
    if bf16_storage:
        P = {k: bf16(v) for k, v in P.items()}
        relu = lambda v: bf16(np.maximum(v, 0.0))
    else:
        relu = lambda v: np.maximum(v, 0.0)
    c = {"x": x[..., None].astype(np.float64)}
    c["e0"] = relu(conv(c["x"], P["a"], 1, d0))
    c["e1"] = relu(conv(c["e0"], P["b"], 2, d1))
    c["e2"] = relu(conv(c["e1"], P["c"], 2, d2))
    c["e3"] = relu(conv(c["e2"], P["d"], 2, d3))
    c["d2"] = relu(conv(up(c["e3"], d2), P["e"], 1, d2) + c["e2"])
    c["d1"] = relu(conv(up(c["d2"], d1), P["f"], 1, d1) + c["e1"])
    c["d0"] = relu(conv(up(c["d1"], d0), P["g"], 1, d0) + c["e0"])
    y = conv(c["d0"], P["h"], 1, d0)[..., 0]
    return y, c


def backward(P: dict, c: dict, dy: np.ndarray) -> dict:
    """Gradients of sum(dy * y) with respect to every layer's weights."""
    g = {}
    d0 = dy.shape
    d1, d2, d3 = c["e1"].shape[:3], c["e2"].shape[:3], c["e3"].shape[:3]
    G = dy[..., None]
    g["h"] = conv_weight_grad(c["d0"], G, 1, 1, 16)
    gd0 = conv_adjoint(G, P["h"], 1, d0) * (c["d0"] > 0)          # pre-activation of d0
    ge0 = gd0.copy()                                              # skip into e0
    u1 = up(c["d1"], d0)
    g["g"] = conv_weight_grad(u1, gd0, 1, 16, 32)
    gd1 = up_adjoint(conv_adjoint(gd0, P["g"], 1, d0), d1) * (c["d1"] > 0)
    ge1 = gd1.copy()
    u2 = up(c["d2"], d1)
    g["f"] = conv_weight_grad(u2, gd1, 1, 32, 64)
    gd2 = up_adjoint(conv_adjoint(gd1, P["f"], 1, d1), d2) * (c["d2"] > 0)
    ge2 = gd2.copy()
    u3 = up(c["e3"], d2)
    g["e"] = conv_weight_grad(u3, gd2, 1, 64, 64)
    ge3 = up_adjoint(conv_adjoint(gd2, P["e"], 1, d2), d3) * (c["e3"] > 0)
    g["d"] = conv_weight_grad(c["e2"], ge3, 2, 64, 64)
    ge2 = (ge2 + conv_adjoint(ge3, P["d"], 2, d2)) * (c["e2"] > 0)
    g["c"] = conv_weight_grad(c["e1"], ge2, 2, 64, 32)
    ge1 = (ge1 + conv_adjoint(ge2, P["c"], 2, d1)) * (c["e1"] > 0)
    g["b"] = conv_weight_grad(c["e0"], ge1, 2, 32, 16)
    ge0 = (ge0 + conv_adjoint(ge1, P["b"], 2, d0)) * (c["e0"] > 0)
    g["a"] = conv_weight_grad(c["x"], ge0, 1, 16, 1)
    return g


def flatten(g: dict) -> np.ndarray:
    return np.concatenate([g[name].ravel() for name, _, _ in LAYERS])


def init_params(seed: int) -> np.ndarray:
    """He-normal weights (std sqrt(2 / (27 cin))) from a Box-Muller transform
    of the SplitMix64 stream of `seed` -- the product's initializer."""
    from .oracle import splitmix64_uniform
    n = n_params()
    u = (splitmix64_uniform(seed, 2 * n) + 1.0) * 0.5  # [0, 1)
    u1 = np.maximum(u[0::2], 2.0 ** -53)
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u[1::2])
    out, k = np.zeros(n), 0
    for _, ci, co in LAYERS:
        m = 27 * ci * co
        out[k:k + m] = z[k:k + m] * np.sqrt(2.0 / (27 * ci))
        k += m
    return out


# ---------------------------------------------------------------------------
# The learned preconditioner as a solver component, composed from this
# module's network and the multigrid oracle's Chebyshev smoother, Lanczos
# estimate and transfers (oracle/mg_oracle.py).  No reference counterpart.
# ---------------------------------------------------------------------------

def smoother_interval(fine, iters: int = 12, safety: float = 1.2, smoothing_range: float = 15.0,
                      seed: int = 0x5EED):
    """[lambda_max / range, lambda_max] with lambda_max = safety x the
    `iters`-step Lanczos estimate of the largest eigenvalue of D^-1 A, from
    the SplitMix64 vector of `seed` with constrained rows zeroed: the
    interval of the cycle's Chebyshev smoother."""
    from . import mg_oracle as M
    from . import oracle as O
    r0 = O.splitmix64_uniform(seed, fine.N)
    r0[fine.constrained.astype(bool)] = 0.0
    lmax = safety * M.lanczos_lambda_max(fine.vmult, fine.inverse_diagonal(), r0, iters)
    return lmax / smoothing_range, lmax


def two_grid_apply(fine, coarse, P, params, out_scale, degree, lmin, lmax, r,
                   bf16_storage=True, parts=False):
    """FP64 restatement of the two-level cycle with the network as the coarse
    solve (benchmarks_amd/csrc/nn.cpp, precondition_two_level with a coarse
    operator): from rb = r with constrained rows 0,
      z   = `degree` Chebyshev(Jacobi) steps on A z = rb from zero,
      rho = rb - A z,   rho_c = P^T rho (coarse constrained rows 0),
      c   = P (||rho_c|| s net(rho_c / ||rho_c||)) (constrained rows 0 on
            both grids),
      z  += omega c,  omega = c.rho / c.Ac (0 if c.Ac <= 0),
      z   = `degree` Chebyshev steps from z; z_c = r_c on constrained rows.
    fine, coarse: oracle Problems of one box; P: the prolongation between
    them (mg_oracle.prolongation_3d); params: flat network weights.  Returns
    (z, omega), or with parts=True also the iterates before and after the
    correction."""
    from . import mg_oracle as M
    cons = fine.constrained.astype(bool)
    cons_c = coarse.constrained.astype(bool)
    dinv = fine.inverse_diagonal()
    rb = np.array(r, dtype=np.float64)
    rb[cons] = 0.0
    z = M.chebyshev(fine.vmult, dinv, rb, None, degree, lmin, lmax)
    z_pre = z.copy()
    rho = rb - fine.vmult(z)
    rc = P.T @ rho
    rc[cons_c] = 0.0
    rn = float(np.linalg.norm(rc))
    omega = 0.0
    if rn > 0.0:
        cd = tuple(n * coarse.p + 1 for n in coarse.box.n)  # x, y, z node counts
        y, _ = forward(split_params(params), (rc / rn).reshape(cd[::-1]), bf16_storage)
        ec = rn * out_scale * y.ravel()
        ec[cons_c] = 0.0
        c = P @ ec
        c[cons] = 0.0
        cac = float(c @ fine.vmult(c))
        omega = float(c @ rho) / cac if cac > 0.0 else 0.0
        z = z + omega * c
    z_corr = z.copy()
    z = M.chebyshev(fine.vmult, dinv, rb, z, degree, lmin, lmax)
    z[cons] = r[cons]
    if parts:
        return z, omega, z_pre, z_corr
    return z, omega


def fcg(vmult, precondition, b, max_steps, m: int = 4, x0=None):
    """Flexible CG with up to m stored directions (Notay's FCG(m)) as
    mfx_nn_fcg_solve runs it: z = precondition(r); p = z minus its
    A-projections on the stored directions other than the slot it replaces,
    p = z - sum_j (z.q_j / p_j.q_j) p_j with q_j = A p_j; alpha = p.r / p.q;
    x += alpha p; r -= alpha q.  Returns (x, residual norms)."""
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64)
    r = b - vmult(x)
    Ps, Qs, pq = [None] * m, [None] * m, [0.0] * m
    hist = [float(np.linalg.norm(r))]
    for steps in range(max_steps):
        z = precondition(r)
        slot, stored = steps % m, min(steps, m)
        p = z.copy()
        for j in range(stored):
            if j != slot:
                p -= (z @ Qs[j]) / pq[j] * Ps[j]
        q = vmult(p)
        Ps[slot], Qs[slot], pq[slot] = p, q, float(p @ q)
        alpha = (p @ r) / pq[slot]
        x += alpha * p
        r -= alpha * q
        hist.append(float(np.linalg.norm(r)))
    return x, hist
