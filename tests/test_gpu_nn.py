"""The learned preconditioner's network on the GPU (mfx_nn_*) against the
FP64 network oracle (oracle/nn_oracle.py).

Tolerances: the forward pass stores activations and weights in bfloat16
(8-bit mantissa); against the oracle with the same bfloat16 storage it
agrees to 1e-2 relative (2-norm) -- rare one-ulp rounding differences of an
activation propagate -- and to 4e-2 against plain FP64.  Weight gradients
(FP32 backward on the stored activations) agree per layer to 2e-2 with the
bfloat16-storage oracle.  Training and flexible CG are checked by what they
must achieve: a falling loss, and fewer FCG steps with the trained network
than without it."""
import numpy as np
import pytest

from oracle import nn_oracle as N
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DIMS = (13, 11, 9)  # x, y, z nodes


def _t(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device="cuda")


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _net(seed=3):
    from benchmarks_amd.laplace import LearnedPreconditioner
    net = LearnedPreconditioner(seed)
    flat = N.init_params(seed)
    net.set_params(flat)
    return net, flat.astype(np.float32).astype(np.float64)


def test_initialization_matches_oracle(gpu):
    from benchmarks_amd.laplace import LearnedPreconditioner
    net = LearnedPreconditioner(5)
    assert net.n_params() == N.n_params()
    assert np.allclose(net.get_params(), N.init_params(5), rtol=1e-6, atol=1e-7)


def test_forward_matches_oracle(gpu):
    net, flat = _net()
    x = np.random.default_rng(1).standard_normal(DIMS[::-1])
    y = net.forward(DIMS, _t(x)).cpu().numpy().astype(np.float64)
    yb, _ = N.forward(N.split_params(flat), x, bf16_storage=True)
    y64, _ = N.forward(N.split_params(flat), x)
    assert _rel(y, yb) <= 1e-2
    assert _rel(y, y64) <= 4e-2


def test_forward_on_odd_and_even_grids(gpu):
    net, flat = _net(4)
    for dims in [(8, 8, 8), (17, 5, 6), (2, 3, 4)]:
        x = np.random.default_rng(2).standard_normal(dims[::-1])
        y = net.forward(dims, _t(x)).cpu().numpy().astype(np.float64)
        yb, _ = N.forward(N.split_params(flat), x, bf16_storage=True)
        assert _rel(y, yb) <= 1e-2, dims


def test_gradient_matches_oracle(gpu):
    net, flat = _net()
    rng = np.random.default_rng(3)
    x = rng.standard_normal(DIMS[::-1])
    w = rng.standard_normal(DIMS[::-1])
    g = net.gradient(DIMS, _t(x), _t(w)).cpu().numpy().astype(np.float64)
    P = {k: N.bf16(v) for k, v in N.split_params(flat).items()}
    _, cache = N.forward(P, x, bf16_storage=True)
    go = N.flatten(N.backward(P, cache, w))
    starts = np.cumsum([0] + [27 * ci * co for _, ci, co in N.LAYERS])
    for i, (name, _, _) in enumerate(N.LAYERS):
        a, b = starts[i], starts[i + 1]
        assert _rel(g[a:b], go[a:b]) <= 2e-2, name


def test_network_is_positively_homogeneous_on_gpu(gpu):
    net, _ = _net()
    x = _t(np.random.default_rng(4).standard_normal(DIMS[::-1]))
    y1 = net.forward(DIMS, x)
    y2 = net.forward(DIMS, 4.0 * x)  # a power of two: the bf16 roundings scale exactly
    assert (y2 - 4.0 * y1).abs().max().item() <= 1e-6 * y1.abs().max().item()


def test_training_reduces_loss(gpu):
    import torch
    net, _ = _net(7)
    rng = np.random.default_rng(5)
    dims = (12, 12, 12)
    x = _t(rng.standard_normal((4,) + dims))
    # a smooth linear target map: a 3-point average along x of the input
    xa = x.cpu().numpy()
    t = _t((np.roll(xa, 1, axis=3) + xa + np.roll(xa, -1, axis=3)) / 3.0)
    losses = [net.train_step(dims, x, t, 2e-3) for _ in range(150)]
    assert losses[-1] < 0.5 * losses[0], (losses[0], losses[-1])
    assert torch.isfinite(torch.tensor(losses)).all()


def _problem(p=2, n=10):
    from benchmarks_amd.laplace import LaplaceOperator
    return LaplaceOperator(p, p + 1, O.GLL, n, deform=0.05)


def test_preconditioner_passes_constrained_rows(gpu):
    from benchmarks_amd.laplace import LearnedPreconditioner, random_vector
    A = _problem()
    net = LearnedPreconditioner(2)
    r = random_vector(A.m(), 3)
    z = A.initialize_dof_vector()
    net.vmult(A, z, r)
    cons = A.device_array("constrained", __import__("torch").int32, A.info.n_constrained).long()
    assert (z[cons] == r[cons]).all()
    z2 = A.initialize_dof_vector()
    net.vmult(A, z2, 3.0 * r)  # z = ||r|| s net(r / ||r||): exactly linear in ||r||
    assert ((z2 - 3.0 * z).abs().max() / z.abs().max()).item() <= 1e-12


def test_trained_network_reduces_fcg_steps(gpu):
    """Train on (residual, error) pairs of the operator itself with the
    energy-norm loss, then compare flexible-CG steps of the two-level
    preconditioner (one Chebyshev step on each side of the network) with the
    same cycle without the network and with Jacobi: what the trained network
    adds.  Measured: 22 steps against 66 without the network, 118 Jacobi."""
    from benchmarks_amd.laplace import (DiagonalMatrix, LaplaceOperator, LearnedPreconditioner,
                                        NoConvergence, ReductionControl, SolverCG, random_vector)
    from benchmarks_amd.learned import smooth_fields, train_energy
    A = LaplaceOperator(2, 3, O.GLL, 12, deform=0.05, with_fp32=True)
    inv = A.compute_inverse_diagonal()
    rs, es = smooth_fields(A, 24, seed=3, n_modes=32)
    trained = LearnedPreconditioner(11)
    stats = train_energy(trained, A, rs, es, steps=400, batch=4, lr=2e-3, seed=1)
    assert stats["loss_last"] < 0.1 * stats["loss_first"], stats
    without = LearnedPreconditioner(11)
    trained.set_mode(1, True)
    without.set_mode(1, False)
    b = random_vector(A.m(), 21)
    A.zero_constrained(b)
    steps = {}
    for name, P in [("trained", trained), ("without", without), ("jacobi", DiagonalMatrix(inv))]:
        x = A.initialize_dof_vector()
        c = ReductionControl(3000, 1e-300, 1e-8)
        SolverCG(c).solve(A, x, b, P)
        assert c.last_value() <= 1e-8 * c.initial_value()
        steps[name] = c.last_step()
    print("FCG steps", steps)
    assert steps["trained"] < 0.6 * steps["without"], steps
    assert steps["trained"] < 0.5 * steps["jacobi"], steps


def test_coarse_correction_matches_composition(gpu):
    """The two-grid application with the network on a coarse box operator,
    recomposed from the library's separately tested pieces: Chebyshev steps
    (mfx_chebyshev_f32), restriction and prolongation (checked against the
    multigrid oracle), the network (checked against the network oracle) and
    the energy-optimal weight omega = c.rho / c.Ac.  The pieces run in FP64
    here and in FP32 inside the preconditioner: agreement to 1e-4 relative,
    and omega to 1e-4 (an omitted or misplaced transfer, weight or
    constraint is an O(1) difference)."""
    import torch
    from benchmarks_amd.laplace import (LaplaceOperator, LearnedPreconditioner, chebyshev_f32,
                                        random_vector, transfer_prolongate, transfer_restrict)
    A = LaplaceOperator(2, 3, O.GLL, (8, 6, 4), deform=0.05, with_fp32=True)
    Cop = LaplaceOperator(1, 2, O.GLL, (4, 3, 2), deform=0.05, with_fp32=True)
    cdims = (5, 4, 3)
    net = LearnedPreconditioner(5)
    net.set_coarse(Cop)
    deg = 2
    net.set_mode(deg, True)
    r = random_vector(A.m(), 7)
    A.zero_constrained(r)
    z = A.initialize_dof_vector()
    net.vmult(A, z, r)
    info = net.smoother_info(A)
    assert info["omega"] != 0.0
    cons = A.device_array("constrained", torch.int32, A.info.n_constrained).long()
    dinv = A.compute_inverse_diagonal().float()
    r32 = r.float()
    r32[cons] = 0.0
    zc = torch.zeros_like(r32)
    chebyshev_f32(A, dinv, zc, r32, deg, info["lambda_min"], info["lambda_max"], True)
    az = torch.empty_like(zc)
    A.vmult(az, zc)
    rho = (r32 - az).double()
    rc = transfer_restrict(A, Cop, rho)
    Cop.zero_constrained(rc)
    rn = torch.linalg.norm(rc)
    y = net.forward(cdims, (rc / rn).float().contiguous())
    ec = (rn * net.output_scale) * y.double()
    Cop.zero_constrained(ec)
    c = transfer_prolongate(A, Cop, ec)
    A.zero_constrained(c)
    ac = A.initialize_dof_vector()
    A.vmult(ac, c)
    omega = (c @ rho / (c @ ac)).item()
    assert abs(omega - info["omega"]) <= 1e-4 * abs(omega), (omega, info["omega"])
    zc += (omega * c).float()
    chebyshev_f32(A, dinv, zc, r32, deg, info["lambda_min"], info["lambda_max"], False)
    want = zc.double()
    want[cons] = r[cons]
    err = (torch.linalg.norm(want - z) / torch.linalg.norm(want)).item()
    assert err <= 1e-4, err
    # without the network the coarse operator plays no part (the FP32
    # operator's atomic scatter leaves the last bits order-dependent)
    plain = LearnedPreconditioner(5)
    plain.set_mode(deg, False)
    z0 = A.initialize_dof_vector()
    plain.vmult(A, z0, r)
    net.set_mode(deg, False)
    z1 = A.initialize_dof_vector()
    net.vmult(A, z1, r)
    assert (torch.linalg.norm(z0 - z1) / torch.linalg.norm(z0)).item() <= 1e-6


def test_trained_coarse_network_beats_jacobi_in_work(gpu):
    """The measured configuration at small size: Q4 on 16^3 cells, the
    network as the coarse solve on the Q1 grid of 8^3 cells, trained on
    (residual, error) pairs of that Q1 operator from its FP64 multigrid CG,
    half of them restricted pre-smoothed fine residuals (train_coarse).
    What the
    trained network adds: flexible-CG steps against the same cycle without
    it, and fine-operator applications (2 degree + 2 per FCG step, 1 per
    Jacobi-CG step) against Jacobi-CG."""
    from benchmarks_amd.laplace import (DiagonalMatrix, LaplaceOperator, LearnedPreconditioner,
                                        PreconditionMG, ReductionControl, SolverCG)
    from benchmarks_amd.learned import _stream, smooth_rhs, train_coarse
    A = LaplaceOperator(4, 5, O.GLL, 16, deform=0.05, with_fp32=True, geometry=1)
    Cmg = PreconditionMG(1, 2, O.GLL, 8, deform=0.05)
    net = LearnedPreconditioner(3)
    st = train_coarse(net, Cmg, n_fields=32, mse_steps=100, energy_steps=300, seed=3, fine=A,
                      smoother_degree=4)
    assert st["data_solver"] == "multigrid CG" and st["n_restricted_fields"] == 16
    assert st["energy"]["loss_last"] < 0.5 * st["energy"]["loss_first"], st["energy"]
    deg = 4
    net.set_mode(deg, True)
    without = LearnedPreconditioner(3)
    without.set_mode(deg, False)
    b = smooth_rhs(A, A.dof_coordinates(), _stream(4), 32)
    steps = {}
    for name, P in [("trained", net), ("without", without),
                    ("jacobi", DiagonalMatrix(A.compute_inverse_diagonal()))]:
        x = A.initialize_dof_vector()
        c = ReductionControl(5000, 1e-300, 1e-8)
        SolverCG(c).solve(A, x, b, P)
        assert c.last_value() <= 1e-8 * c.initial_value()
        steps[name] = c.last_step()
    print("FCG steps", steps)
    assert steps["trained"] < 0.5 * steps["without"], steps
    assert (2 * deg + 2) * steps["trained"] < steps["jacobi"], steps


def test_flexible_cg_equals_pcg_for_linear_preconditioner(gpu):
    """With the network switched off the preconditioner is a fixed Chebyshev
    polynomial, so flexible CG must take PCG's steps: compare with the
    oracle's Jacobi-CG bound (it cannot be slower than 1 Chebyshev step per
    Jacobi step) and with a rerun (deterministic)."""
    from benchmarks_amd.laplace import (DiagonalMatrix, LaplaceOperator, LearnedPreconditioner,
                                        ReductionControl, SolverCG, random_vector)
    A = LaplaceOperator(3, 4, O.GLL, 6, deform=0.05, with_fp32=True)
    b = random_vector(A.m(), 2)
    A.zero_constrained(b)
    P = LearnedPreconditioner(1)
    P.set_mode(2, False)
    runs = []
    for _ in range(2):
        x = A.initialize_dof_vector()
        c = ReductionControl(3000, 1e-300, 1e-9)
        SolverCG(c).solve(A, x, b, P)
        runs.append(c.last_step())
    xj = A.initialize_dof_vector()
    cj = ReductionControl(3000, 1e-300, 1e-9)
    SolverCG(cj).solve(A, xj, b, DiagonalMatrix(A.compute_inverse_diagonal()))
    assert runs[0] == runs[1]
    assert runs[0] < cj.last_step()


def test_coarse_network_on_two_fine_operators_in_turn(gpu):
    """One network and coarse operator used with two fine operators of
    different sizes one after the other, the first destroyed before the
    second is made (so the second may reuse its address): the transfer and
    its scratch are rebuilt for the second operator, and the application
    equals a fresh network's on it (FP32 atomic sums: 1e-5)."""
    import gc

    import torch
    from benchmarks_amd.laplace import LaplaceOperator, LearnedPreconditioner, random_vector
    Cop = LaplaceOperator(1, 2, O.GLL, 4, deform=0.05, with_fp32=True)
    net = LearnedPreconditioner(5)
    net.set_coarse(Cop)
    net.set_mode(2, True)
    A1 = LaplaceOperator(2, 3, O.GLL, 4, deform=0.05, with_fp32=True)
    r1 = random_vector(A1.m(), 3)
    A1.zero_constrained(r1)
    z1 = A1.initialize_dof_vector()
    net.vmult(A1, z1, r1)
    del A1, r1, z1
    gc.collect()
    A2 = LaplaceOperator(2, 3, O.GLL, 8, deform=0.05, with_fp32=True)
    r2 = random_vector(A2.m(), 4)
    A2.zero_constrained(r2)
    z2 = A2.initialize_dof_vector()
    net.vmult(A2, z2, r2)
    fresh = LearnedPreconditioner(5)
    fresh.set_coarse(Cop)
    fresh.set_mode(2, True)
    z3 = A2.initialize_dof_vector()
    fresh.vmult(A2, z3, r2)
    assert (torch.linalg.norm(z2 - z3) / torch.linalg.norm(z3)).item() <= 1e-5


def _fallback_problem(n=8, p=2):
    from benchmarks_amd.laplace import PreconditionMG, random_vector
    mg = PreconditionMG(p, p + 1, O.GLL, n, deform=0.05, geometry=1)
    A = mg.operator()
    b = random_vector(A.m(), 12)
    A.zero_constrained(b)
    return mg, A, b


@pytest.mark.parametrize("degree", [0, 3])
def test_fallback_takes_over_from_a_dead_network(gpu, degree):
    """A zeroed network (host-side switch: all weights 0) contributes
    nothing.  Alone (smoother degree 0) it gives a zero direction, which
    breaks flexible CG down at step 0; inside the two-grid cycle its
    correction is rejected (omega = 0) at every application, and after
    `window` of them the solve switches.  With the multigrid fallback set,
    the solve converges in a few multigrid steps and reports the switch;
    without it, degree 0 raises NoConvergence."""
    from benchmarks_amd.laplace import (LearnedPreconditioner, NoConvergence, ReductionControl,
                                        SolverCG)
    mg, A, b = _fallback_problem()
    net = LearnedPreconditioner(4)
    net.set_params(np.zeros(net.n_params(), dtype=np.float32))
    net.set_mode(degree, True)
    if degree == 0:
        with pytest.raises(NoConvergence):
            SolverCG(ReductionControl(200, 1e-300, 1e-8)).solve(A, A.initialize_dof_vector(), b,
                                                                net)
    window = 4
    net.set_fallback(mg, window=window)
    x = A.initialize_dof_vector()
    c = ReductionControl(200, 1e-300, 1e-8)
    SolverCG(c).solve(A, x, b, net)
    assert c.last_value() <= 1e-8 * c.initial_value()
    xm = A.initialize_dof_vector()
    cm = ReductionControl(200, 1e-300, 1e-8)
    SolverCG(cm).solve(A, xm, b, mg)
    expect = 0 if degree == 0 else window
    assert net.fallback_step == expect, (net.fallback_step, expect)
    assert c.last_step() <= expect + cm.last_step() + 2, (c.last_step(), cm.last_step())
    # the FP64 residual of the returned solution meets the tolerance
    r = A.initialize_dof_vector()
    A.vmult(r, x)
    assert (np.linalg.norm((b - r).cpu().numpy()) <= 1.01e-8 * np.linalg.norm(b.cpu().numpy()))


def test_fallback_stays_off_for_a_working_preconditioner(gpu):
    """The Chebyshev-only cycle (network off) reduces the residual steadily:
    the stagnation test (less than 2x over the window) never fires, and the
    rejected-correction test does not apply without the network."""
    from benchmarks_amd.laplace import LearnedPreconditioner, ReductionControl, SolverCG
    mg, A, b = _fallback_problem()
    net = LearnedPreconditioner(4)
    net.set_mode(3, False)
    net.set_fallback(mg)
    c = ReductionControl(2000, 1e-300, 1e-8)
    SolverCG(c).solve(A, A.initialize_dof_vector(), b, net)
    assert net.fallback_step == -1 and c.last_step() > 8


def test_fallback_rejects_a_foreign_multigrid(gpu):
    from benchmarks_amd._lib import MfxError
    from benchmarks_amd.laplace import LearnedPreconditioner, ReductionControl, SolverCG
    mg, A, b = _fallback_problem()
    other, _, _ = _fallback_problem(n=6)
    net = LearnedPreconditioner(4)
    net.set_mode(2, False)
    net.set_fallback(other)
    with pytest.raises(MfxError):
        SolverCG(ReductionControl(50, 1e-300, 1e-8)).solve(A, A.initialize_dof_vector(), b, net)


def _cycle_setup(seed=5):
    from benchmarks_amd.laplace import LaplaceOperator, LearnedPreconditioner
    from oracle import mg_oracle as M
    n, nc = (8, 6, 4), (4, 3, 2)
    A = LaplaceOperator(2, 3, O.GLL, n, deform=0.05, with_fp32=True)
    Cop = LaplaceOperator(1, 2, O.GLL, nc, deform=0.05, with_fp32=True)
    fine = O.make_problem(2, 3, O.GLL, n, deform=0.05)
    coarse = O.make_problem(1, 2, O.GLL, nc, deform=0.05)
    net = LearnedPreconditioner(seed)
    net.set_coarse(Cop)
    return A, Cop, fine, coarse, M.prolongation_3d(fine, coarse), net


def test_two_grid_cycle_matches_oracle(gpu):
    """The library's two-level cycle (FP32 smoothing and transfers, network
    on BF16 MFMA) against its FP64 restatement composed from the network and
    multigrid oracles (nn_oracle.two_grid_apply, bfloat16 storage in the
    network), same weights and right-hand side.  The Chebyshev interval
    agrees to 1e-6 (both Lanczos estimates run in FP64); omega and z to 1e-2:
    FP32 residuals move a few bfloat16 roundings of the network's
    activations.  A dropped smoothing pass, transfer, weight or constraint
    is an O(1) difference."""
    A, Cop, fine, coarse, P, net = _cycle_setup()
    deg = 2
    net.set_mode(deg, True)
    params = net.get_params().astype(np.float64)
    r = O.splitmix64_uniform(17, A.m())
    import torch
    z = A.initialize_dof_vector()
    net.vmult(A, z, torch.as_tensor(r, device="cuda"))
    info = net.smoother_info(A)
    lmin, lmax = N.smoother_interval(fine)
    assert abs(info["lambda_max"] - lmax) <= 1e-6 * lmax
    zo, omega = N.two_grid_apply(fine, coarse, P, params, net.output_scale, deg, lmin, lmax, r)
    assert abs(info["omega"] - omega) <= 1e-2 * abs(omega), (info["omega"], omega)
    assert _rel(z.cpu().numpy(), zo) <= 1e-2, _rel(z.cpu().numpy(), zo)


@pytest.mark.parametrize("steps", [1, 2])
def test_fcg_steps_match_oracle(gpu, steps):
    """The first flexible-CG steps with the two-level cycle (mfx_nn_fcg_solve
    stopped by the step limit) against the oracle's FCG(4) around the
    oracle's cycle: the iterate to 1e-2 (the cycle's own agreement)."""
    import torch
    from benchmarks_amd.laplace import NoConvergence, ReductionControl, SolverCG
    A, Cop, fine, coarse, P, net = _cycle_setup(6)
    deg = 2
    net.set_mode(deg, True)
    params = net.get_params().astype(np.float64)
    b = O.splitmix64_uniform(19, A.m())
    b[fine.constrained.astype(bool)] = 0.0
    x = A.initialize_dof_vector()
    with pytest.raises(NoConvergence):
        SolverCG(ReductionControl(steps, 1e-300, 1e-14)).solve(
            A, x, torch.as_tensor(b, device="cuda"), net)
    lmin, lmax = N.smoother_interval(fine)
    s = net.output_scale
    xo, hist = N.fcg(fine.vmult,
                     lambda r: N.two_grid_apply(fine, coarse, P, params, s, deg, lmin, lmax, r)[0],
                     b, steps)
    assert hist[-1] < hist[0]
    assert _rel(x.cpu().numpy(), xo) <= 1e-2, _rel(x.cpu().numpy(), xo)


def test_learned_fcg_on_the_configured_problem(gpu):
    """The learned-preconditioner configuration at its size: Q4 BP5 on 64^3
    deformed cells (16,974,593 DoFs), the network as the coarse solve on the
    Q1 grid of 32^3 cells (33^3 nodes), trained on 32 pairs from the FP64
    multigrid CG of that grid (half restricted pre-smoothed fine residuals;
    the bench trains on 4096), smoother degree 4, a smooth right-hand side.
    Flexible CG to 1e-8 with the multigrid fallback armed: the residual
    recomputed in FP64 from the returned solution, by the library and by the
    oracle's CPU operator, meets 1e-8, the fallback
    never engages, the trained network takes fewer steps than the same cycle
    without it, and fewer fine-operator applications (2k + 2 per FCG step:
    2k Chebyshev, one for omega, FCG's own) than Jacobi-CG (one per step)."""
    import torch
    from benchmarks_amd.laplace import (DiagonalMatrix, LearnedPreconditioner, PreconditionMG,
                                        ReductionControl, SolverCG)
    from benchmarks_amd.learned import _stream, smooth_rhs, train_coarse
    n, deg = 64, 4
    mg = PreconditionMG(4, 5, O.GLL, n, deform=0.05, geometry=1)
    A = mg.operator()
    assert A.m() == 16974593
    cmg = PreconditionMG(1, 2, O.GLL, n // 2, deform=0.05)
    net = LearnedPreconditioner(3)
    st = train_coarse(net, cmg, n_fields=32, mse_steps=200, energy_steps=600, seed=3, fine=A,
                      smoother_degree=deg)
    assert st["coarse_dims"] == (33, 33, 33) and st["data_solver"] == "multigrid CG"
    net.set_mode(deg, True)
    net.set_fallback(mg)
    b = smooth_rhs(A, A.dof_coordinates(), _stream(4), 32)
    steps = {}
    without = LearnedPreconditioner(3)
    without.set_mode(deg, False)
    for name, P in [("learned", net), ("without", without),
                    ("jacobi", DiagonalMatrix(A.compute_inverse_diagonal()))]:
        x = A.initialize_dof_vector()
        c = ReductionControl(20000, 1e-300, 1e-8)
        SolverCG(c).solve(A, x, b, P)
        steps[name] = c.last_step()
        if name == "learned":
            r = A.initialize_dof_vector()
            A.vmult(r, x)
            res = (torch.linalg.norm(b - r) / torch.linalg.norm(b)).item()
            assert res <= 1e-8, res
            assert net.fallback_step == -1
            # the same certificate from the oracle's own operator on the CPU
            # (its own mesh, map and stored metric terms)
            Po = O.make_problem(4, 5, O.GLL, n, deform=0.05)
            xh, bh = x.cpu().numpy(), b.cpu().numpy()
            res_o = np.linalg.norm(bh - Po.vmult(xh, threads=16)) / np.linalg.norm(bh)
            assert res_o <= 1.01e-8, res_o
            del Po
    print("C5 steps", steps)
    assert steps["learned"] < steps["without"], steps
    assert (2 * deg + 2) * steps["learned"] < steps["jacobi"], steps
