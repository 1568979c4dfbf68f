"""Training data and training loop of the learned preconditioner
(benchmarks_amd.laplace.LearnedPreconditioner).

Data: residual -> error pairs (r, e = A^-1 r) on the node grid of a box
operator.  r is the nodal interpolant of a sum of `n_modes` random Fourier
modes sin(pi k1 x) sin(pi k2 y) sin(pi k3 z), integer wave numbers in
1..k_max (default 8) and amplitudes uniform in [-1, 1] drawn from the SplitMix64 stream of
`seed`, zero on the Dirichlet boundary; e solves A e = r in FP64 to 1e-10,
by the FP64 multigrid-preconditioned CG when a PreconditionMG is given (the
measured configuration), else by Jacobi-CG.

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
the network (restricted_fields): a smooth fine field after the cycle's own
Chebyshev pre-smoothing, its residual restricted to the coarse grid, and the
exact coarse solution.  The mix took 31-32 flexible-CG steps on the measured
problem against 36-37 for coarse fields alone and 39-51 for restricted
fields alone.
"""
from __future__ import annotations

import math
import time

import torch

from .laplace import (DiagonalMatrix, PreconditionMG, ReductionControl, SolverCG, chebyshev_f32,
                      transfer_restrict)


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


def smooth_rhs(op, xyz, u, n_modes, k_max: int = 8):
    f = torch.zeros(op.m(), dtype=torch.float64, device="cuda")
    for _ in range(n_modes):
        k = [1 + min(int(next(u) * k_max), k_max - 1) for _ in range(3)]
        a = 2.0 * next(u) - 1.0
        f += a * (torch.sin(math.pi * k[0] * xyz[:, 0]) * torch.sin(math.pi * k[1] * xyz[:, 1])
                  * torch.sin(math.pi * k[2] * xyz[:, 2]))
    op.zero_constrained(f)
    return f


def smooth_fields(op_or_mg, count: int, seed: int = 3, n_modes: int = 32, rel_tol: float = 1e-10,
                  k_max: int = 8):
    """`count` pairs (r, e) as FP64 vectors of the operator's DoFs."""
    if isinstance(op_or_mg, PreconditionMG):
        op, precond = op_or_mg.operator(), op_or_mg
    else:
        op, precond = op_or_mg, DiagonalMatrix(op_or_mg.compute_inverse_diagonal())
    xyz = op.dof_coordinates()
    u = _stream(seed)
    rs, es = [], []
    for _ in range(count):
        r = smooth_rhs(op, xyz, u, n_modes, k_max)
        e = op.initialize_dof_vector()
        SolverCG(ReductionControl(100000, 1e-300, rel_tol)).solve(op, e, r, precond)
        rs.append(r)
        es.append(e)
    return rs, es


def restricted_fields(net, fine, coarse, count: int, degree: int = 4, seed: int = 13,
                      n_modes: int = 32, k_max: int = 8, rel_tol: float = 1e-10):
    """`count` pairs (rho_c, e_c) as the two-grid cycle presents them to the
    network: a smooth field r on the fine operator's nodes (random Fourier
    modes up to k_max), `degree` FP32 Chebyshev(Jacobi) steps z from zero on
    the network's own smoother interval, rho = r - A z restricted to the
    coarse box (R = P^T), and e_c solving A_c e_c = rho_c in FP64 (multigrid
    CG when `coarse` is a PreconditionMG, else Jacobi-CG) to rel_tol."""
    if isinstance(coarse, PreconditionMG):
        op_c, precond = coarse.operator(), coarse
    else:
        op_c, precond = coarse, DiagonalMatrix(coarse.compute_inverse_diagonal())
    info = net.smoother_info(fine)
    inv32 = fine.compute_inverse_diagonal().float()
    xyz = fine.dof_coordinates()
    u = _stream(seed)
    rs, es = [], []
    for _ in range(count):
        r = smooth_rhs(fine, xyz, u, n_modes, k_max)
        z = torch.zeros(fine.m(), dtype=torch.float32, device="cuda")
        chebyshev_f32(fine, inv32, z, r.float(), degree, info["lambda_min"], info["lambda_max"],
                      True)
        az = fine.initialize_dof_vector()
        fine.vmult(az, z.double())
        rho = r - az
        fine.zero_constrained(rho)
        rc = transfer_restrict(fine, op_c, rho)
        op_c.zero_constrained(rc)
        e = op_c.initialize_dof_vector()
        SolverCG(ReductionControl(100000, 1e-300, rel_tol)).solve(op_c, e, rc, precond)
        rs.append(rc)
        es.append(e)
    return rs, es


def train_on_fields(net, dims, rs, es, steps: int, batch: int, crop, lr: float, seed: int = 1):
    """Adam steps on random crops (edge `crop` nodes, or whole fields when
    crop is None) of the (r, e) fields; dims = (nx, ny, nz) of the fields.
    Sets the network's output scale.  Returns loss statistics and timing."""
    nx, ny, nz = dims
    ratios = [(torch.linalg.norm(e) / torch.linalg.norm(r)).item() for r, e in zip(rs, es)]
    s = sum(ratios) / len(ratios)
    net.output_scale = s
    R = [r.view(nz, ny, nx) for r in rs]
    E = [e.view(nz, ny, nx) for e in es]
    c = (nx, ny, nz) if crop is None else (min(crop, nx), min(crop, ny), min(crop, nz))
    u = _stream(seed + 1000)
    losses = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        xs, ts = [], []
        for _ in range(batch):
            f = min(int(next(u) * len(R)), len(R) - 1)
            ox = min(int(next(u) * (nx - c[0] + 1)), nx - c[0])
            oy = min(int(next(u) * (ny - c[1] + 1)), ny - c[1])
            oz = min(int(next(u) * (nz - c[2] + 1)), nz - c[2])
            rc = R[f][oz:oz + c[2], oy:oy + c[1], ox:ox + c[0]]
            ec = E[f][oz:oz + c[2], oy:oy + c[1], ox:ox + c[0]]
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
    if net.output_scale == 1.0:
        ratios = [(torch.linalg.norm(e) / torch.linalg.norm(r)).item() for r, e in zip(rs, es)]
        net.output_scale = sum(ratios) / len(ratios)
    u = _stream(seed + 2000)
    losses = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        idx = [min(int(next(u) * len(rs)), len(rs) - 1) for _ in range(batch)]
        r = torch.stack([rs[i] for i in idx]).contiguous()
        e = torch.stack([es[i] for i in idx]).contiguous()
        losses.append(net.train_energy_step(op, r, e, lr))
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
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rs, es = smooth_fields(coarse, n_fields - n_fine, seed=seed, n_modes=32, k_max=k_max)
    if n_fine:
        rf, ef = restricted_fields(net, fine, coarse, n_fine, degree=smoother_degree,
                                   seed=seed + 10, n_modes=32, k_max=fine_k_max)
        rs, es = rs + rf, es + ef
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    st_mse = train_on_fields(net, dims, rs, es, steps=mse_steps, batch=mse_batch, crop=None,
                             lr=2e-3, seed=seed)
    st_en = train_energy(net, op, rs, es, steps=energy_steps, batch=energy_batch, lr=1e-3,
                         seed=seed)
    net.set_coarse(op)
    return {"coarse_dims": dims, "k_max": k_max, "n_fields": n_fields,
            "n_restricted_fields": n_fine, "fine_k_max": fine_k_max, "data_s": t_data,
            "data_solver": "multigrid CG" if op is not coarse else "Jacobi-CG",
            "mse": st_mse, "energy": st_en}
