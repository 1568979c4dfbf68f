"""Pins the oracle's CG: iteration counts of the reference logs (KAT-1),
agreement with a direct solve, and h^(p+1) convergence of the manufactured
solution."""
import numpy as np
import pytest

from oracle import oracle as O
from test_oracle_mesh import assembled_matrix, cycle_mesh

# CEED_bp/results/1xGH200_P4.txt:635-638 (q = 6) and
# bakeoff_problems_dealii/tests/scaling_bp35_kokkos_kernel/results/1_GPU.out:639-642 (q = 5)
KAT1 = {8: 92, 9: 99, 10: 153, 11: 170}


@pytest.mark.parametrize("cycle", sorted(KAT1))
@pytest.mark.parametrize("nq", [6, 5])
def test_cg_iterations_reference_logs(cycle, nq):
    n, lo, hi = cycle_mesh(cycle)
    P = O.make_problem(4, nq, O.GAUSS, n, lo=lo, hi=hi)
    x, r, hist = P.cg(P.rhs(0), rel_tol=1e-9, abs_tol=1e-16)
    assert r.converged and r.steps == KAT1[cycle]
    assert hist[-1] <= 1e-9 * hist[0] < hist[-2]


def test_cg_matches_direct_solve():
    P = O.make_problem(2, 3, O.GLL, (3, 2, 2), deform=0.05)
    A = assembled_matrix(P)
    b = O.splitmix64_uniform(4, P.N)
    b[P.constrained.astype(bool)] = 0
    x, r, _ = P.cg(b, rel_tol=1e-13, abs_tol=1e-300)
    assert r.converged
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-11)
    xj, rj, _ = P.cg(b, inv_diag=P.inverse_diagonal(), rel_tol=1e-13, abs_tol=1e-300)
    assert np.allclose(xj, x, rtol=1e-9, atol=1e-10)


def test_cg_failure_reports_max_steps():
    P = O.make_problem(2, 3, O.GLL, 3)
    b = P.rhs(1)
    x, r, _ = P.cg(b, rel_tol=1e-14, max_steps=3)
    assert not r.converged and r.steps == 3


@pytest.mark.parametrize("p", [2, 4])
def test_manufactured_solution_convergence(p):
    """Q_p, Gauss p+2: L2 error of u = sin(pi x) sin(pi y) sin(pi z) falls
    like h^(p+1) (the configuration C1 problem on coarser meshes)."""
    errs = []
    for n in (2, 4):
        P = O.make_problem(p, p + 2, O.GAUSS, n)
        x, r, _ = P.cg(P.rhs(1), inv_diag=P.inverse_diagonal(), rel_tol=1e-12)
        errs.append(P.l2_error_sin(x))
    rate = np.log2(errs[0] / errs[1])
    assert rate > p + 0.5, (errs, rate)


def test_cg_max_steps_zero_fails_at_step_zero():
    """deal.II's SolverControl::check fails at step 0 when max_steps is 0 and
    the start is not converged: no step is taken and x keeps its value."""
    P = O.make_problem(2, 3, O.GLL, 3)
    b = P.rhs(1)
    x0 = O.splitmix64_uniform(5, P.N)
    x, r, hist = P.cg(b, rel_tol=1e-10, max_steps=0, x0=x0)
    assert not r.converged and r.steps == 0
    assert np.array_equal(x, x0)
    assert len(hist) == 1
