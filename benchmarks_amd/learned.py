"""Training data and training loop of the learned preconditioner
(benchmarks_amd.laplace.LearnedPreconditioner).

Data: residual -> error pairs (r, e = A^-1 r) on the node grid of a box
operator.  r is the nodal interpolant of a sum of `n_modes` random Fourier
modes sin(pi k1 x) sin(pi k2 y) sin(pi k3 z), integer wave numbers in
1..k_max (default 8) and amplitudes uniform in [-1, 1] drawn from the
SplitMix64 stream of `seed` (evaluated by the library's mfx_vec_sine_modes),
zero on the Dirichlet boundary; e solves A e = r in FP64 to 1e-10, by the
FP64 multigrid-preconditioned CG when a PreconditionMG is given (the
measured configuration), else by Jacobi-CG.  Pairs are stored as the rows of
two [count, N] FP64 tensors.

Training: the network sees x = r_c / ||r_c|| on crops r_c of the fields and
learns t = e_c / (s ||r_c||), s = mean ||e|| / ||r|| over the fields (the
output scale); mean squared error, Adam, all through the library's kernels;
then the energy-norm loss on whole fields.

train_coarse does both on the node grid of a coarse box operator (Q1 on half
the fine cells in the measured configuration) and attaches the network as
the coarse solve of the two-grid preconditioner: the wave numbers of the
training fields then reach half the coarse grid's Nyquist number
(k_max = cells / 2), which covers what a restricted smoothed residual holds.
Given the fine operator, half of the pairs are instead what the cycle hands
the network (restricted_fields): a smooth fine field, the cycle's own FP32
Chebyshev pre-smoothing and residual, restricted to the coarse grid by the
library (mfx_nn_restricted_residual), and the exact coarse solution.  The
mix took 31-32 flexible-CG steps on the measured problem with 32 pairs (30
with 4096) against 36-37 for coarse fields alone and 39-51 for restricted
fields alone.
"""
from __future__ import annotations

import time

import torch

from .laplace import (DiagonalMatrix, PreconditionMG, ReductionControl, SolverCG, sine_modes)


def _stream(seed: int):
    """u in [0, 1) from SplitMix64: z = seed + k 0x9E3779B97F4A7C15, standard mix."""
    M = (1 << 64) - 1
    k = 0
    while True:
        k += 1
        z = (seed + k * 0x9E3779B97F4A7C15) & M
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        yield (z >> 11) * 2.0 ** -53


def _modes(u, n_modes: int, k_max: int):
    """Wave numbers (three per mode, each in 1..k_max) and amplitudes in
    [-1, 1] of `n_modes` modes, drawn from the stream in the order k1 k2 k3 a."""
    ks, amps = [], []
    for _ in range(n_modes):
        ks += [1 + min(int(next(u) * k_max), k_max - 1) for _ in range(3)]
        amps.append(2.0 * next(u) - 1.0)
    return ks, amps


def smooth_rhs(op, xyz, u, n_modes, k_max: int = 8, out=None):
    """A smooth field of `n_modes` random sine modes at the operator's DoF
    coordinates xyz, constrained rows zero (into `out` when given)."""
    f = op.initialize_dof_vector() if out is None else out
    ks, amps = _modes(u, n_modes, k_max)
    sine_modes(xyz, ks, amps, f)
    op.zero_constrained(f)
    return f


def _solver(op_or_mg):
    if isinstance(op_or_mg, PreconditionMG):
        return op_or_mg.operator(), op_or_mg
    return op_or_mg, DiagonalMatrix(op_or_mg.compute_inverse_diagonal())


def smooth_fields(op_or_mg, count: int, seed: int = 3, n_modes: int = 32, rel_tol: float = 1e-10,
                  k_max: int = 8, out=None):
    """`count` pairs (r, e) of the operator: the rows of two [count, N] FP64
    tensors (written into `out` = (R, E) when given)."""
    op, precond = _solver(op_or_mg)
    n = op.m()
    R, E = out if out is not None else (
        torch.empty(count, n, dtype=torch.float64, device="cuda"),
        torch.empty(count, n, dtype=torch.float64, device="cuda"))
    xyz = op.dof_coordinates()
    u = _stream(seed)
    for i in range(count):
        smooth_rhs(op, xyz, u, n_modes, k_max, out=R[i])
        E[i].zero_()
        SolverCG(ReductionControl(100000, 1e-300, rel_tol)).solve(op, E[i], R[i], precond)
    return R, E


def restricted_fields(net, fine, coarse, count: int, degree: int = 4, seed: int = 13,
                      n_modes: int = 32, k_max: int = 8, rel_tol: float = 1e-10, out=None):
    """`count` pairs (rho_c, e_c) as the two-grid cycle presents them to the
    network: a smooth field r on the fine operator's nodes (random Fourier
    modes up to k_max), the cycle's own `degree` FP32 Chebyshev(Jacobi) steps
    from zero and residual rho = r - A z, restricted to the coarse box
    (R = P^T; mfx_nn_restricted_residual), and e_c solving A_c e_c = rho_c in
    FP64 (multigrid CG when `coarse` is a PreconditionMG, else Jacobi-CG) to
    rel_tol.  Attaches the coarse operator and the cycle mode to `net`."""
    op_c, precond = _solver(coarse)
    net.set_coarse(op_c)
    net.set_mode(degree, True)
    n = op_c.m()
    R, E = out if out is not None else (
        torch.empty(count, n, dtype=torch.float64, device="cuda"),
        torch.empty(count, n, dtype=torch.float64, device="cuda"))
    xyz = fine.dof_coordinates()
    u = _stream(seed)
    r = fine.initialize_dof_vector()
    for i in range(count):
        smooth_rhs(fine, xyz, u, n_modes, k_max, out=r)
        net.restricted_residual(fine, r, R[i])
        E[i].zero_()
        SolverCG(ReductionControl(100000, 1e-300, rel_tol)).solve(op_c, E[i], R[i], precond)
    return R, E


def _as_rows(x):
    """Pairs given as a list of vectors or as a [count, N] tensor -> tensor."""
    return x if isinstance(x, torch.Tensor) else torch.stack(list(x))


def output_scale(rs, es) -> float:
    """Mean of ||e|| / ||r|| over the pairs."""
    R, E = _as_rows(rs), _as_rows(es)
    return (torch.linalg.norm(E, dim=1) / torch.linalg.norm(R, dim=1)).mean().item()


def train_on_fields(net, dims, rs, es, steps: int, batch: int, crop, lr: float, seed: int = 1):
    """Adam steps on random crops (edge `crop` nodes, or whole fields when
    crop is None) of the (r, e) fields; dims = (nx, ny, nz) of the fields.
    Sets the network's output scale.  Returns loss statistics and timing."""
    nx, ny, nz = dims
    R, E = _as_rows(rs), _as_rows(es)
    s = output_scale(R, E)
    net.output_scale = s
    R = R.view(-1, nz, ny, nx)
    E = E.view(-1, nz, ny, nx)
    count = R.shape[0]
    c = (nx, ny, nz) if crop is None else (min(crop, nx), min(crop, ny), min(crop, nz))
    u = _stream(seed + 1000)
    losses = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        xs, ts = [], []
        for _ in range(batch):
            f = min(int(next(u) * count), count - 1)
            ox = min(int(next(u) * (nx - c[0] + 1)), nx - c[0])
            oy = min(int(next(u) * (ny - c[1] + 1)), ny - c[1])
            oz = min(int(next(u) * (nz - c[2] + 1)), nz - c[2])
            rc = R[f, oz:oz + c[2], oy:oy + c[1], ox:ox + c[0]]
            ec = E[f, oz:oz + c[2], oy:oy + c[1], ox:ox + c[0]]
            nr = torch.linalg.norm(rc).clamp_min(1e-300)
            xs.append((rc / nr).float())
            ts.append((ec / (s * nr)).float())
        x = torch.stack(xs).contiguous()
        t = torch.stack(ts).contiguous()
        losses.append(net.train_step(c, x, t, lr))
    torch.cuda.synchronize()
    k = max(1, min(10, len(losses) // 4))
    return {"steps": steps, "batch": batch, "crop": c, "seconds": time.perf_counter() - t0,
            "loss_first": sum(losses[:k]) / k, "loss_last": sum(losses[-k:]) / k,
            "output_scale": s}


def train_energy(net, op, rs, es, steps: int, batch: int, lr: float, seed: int = 1):
    """Adam steps on the energy-norm loss (pred - e)^T A (pred - e) / e^T A e
    of whole fields of `op` (the loss the preconditioner is judged by);
    keeps the output scale already set (or sets it from the fields)."""
    R, E = _as_rows(rs), _as_rows(es)
    if net.output_scale == 1.0:
        net.output_scale = output_scale(R, E)
    count = R.shape[0]
    u = _stream(seed + 2000)
    losses = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        idx = torch.tensor([min(int(next(u) * count), count - 1) for _ in range(batch)],
                           device=R.device)
        losses.append(net.train_energy_step(op, R[idx].contiguous(), E[idx].contiguous(), lr))
    torch.cuda.synchronize()
    k = max(1, min(10, len(losses) // 4))
    return {"steps": steps, "batch": batch, "seconds": time.perf_counter() - t0,
            "loss_first": sum(losses[:k]) / k, "loss_last": sum(losses[-k:]) / k,
            "output_scale": net.output_scale}


def train_coarse(net, coarse, n_fields: int = 32, k_max=None, mse_steps: int = 200,
                 energy_steps: int = 600, mse_batch: int = 8, energy_batch: int = 2,
                 seed: int = 3, fine=None, smoother_degree: int = 4, fine_k_max=None):
    """Train `net` on (r, e) pairs of a coarse box operator: `coarse` is a
    PreconditionMG on that box (the pairs come from its FP64 multigrid CG;
    the operator is coarse.operator()) or the box operator itself (Jacobi-CG).
    `n_fields` pairs: smooth right-hand sides with wave numbers up to k_max
    (default half the smallest cell count), or, given the fine operator, half
    of them restricted pre-smoothed fine residuals (restricted_fields, fine
    wave numbers up to fine_k_max, default half the fine cell count, and
    `smoother_degree` Chebyshev steps as in the cycle).  Mean squared error
    on whole fields, then the energy-norm loss; the network is attached as
    the coarse solve (net.set_coarse).  Returns data and training
    statistics."""
    op = coarse.operator() if isinstance(coarse, PreconditionMG) else coarse
    cells = op.box_cells
    if cells is None:
        raise ValueError("the coarse operator must be a box operator")
    dims = tuple(c * op.degree + 1 for c in cells)
    if k_max is None:
        k_max = max(1, min(cells) // 2)
    n_fine = n_fields // 2 if fine is not None else 0
    if fine is not None and fine_k_max is None:
        if fine.box_cells is None:
            raise ValueError("the fine operator must be a box operator")
        fine_k_max = max(1, min(fine.box_cells) // 2)
    n = op.m()
    R = torch.empty(n_fields, n, dtype=torch.float64, device="cuda")
    E = torch.empty(n_fields, n, dtype=torch.float64, device="cuda")
    n_coarse = n_fields - n_fine
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    smooth_fields(coarse, n_coarse, seed=seed, n_modes=32, k_max=k_max,
                  out=(R[:n_coarse], E[:n_coarse]))
    if n_fine:
        restricted_fields(net, fine, coarse, n_fine, degree=smoother_degree, seed=seed + 10,
                          n_modes=32, k_max=fine_k_max, out=(R[n_coarse:], E[n_coarse:]))
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    st_mse = train_on_fields(net, dims, R, E, steps=mse_steps, batch=mse_batch, crop=None,
                             lr=2e-3, seed=seed)
    st_en = train_energy(net, op, R, E, steps=energy_steps, batch=energy_batch, lr=1e-3,
                         seed=seed)
    net.set_coarse(op)
    return {"coarse_dims": dims, "k_max": k_max, "n_fields": n_fields,
            "n_restricted_fields": n_fine, "fine_k_max": fine_k_max, "data_s": t_data,
            "data_solver": "multigrid CG" if op is not coarse else "Jacobi-CG",
            "mse": st_mse, "energy": st_en}
