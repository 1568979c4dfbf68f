"""mfx_cg_solve_noalloc (the CG solve whose work buffers live in the
library's fixed device array) against mfx_cg_solve, both through the C-ABI:
two no-allocation solves in a row, the second on a smaller operator, each
taking mfx_cg_solve's steps and solution (same kernels; only the order of the
atomic sums can differ, so the iterates agree to 1e-12); the calls take no
device memory; an operator over the limit is rejected before any work."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _call(fn, op, b, inv, rel_tol=1e-10, x0=None, x=None):
    """Solve into x (a new zero vector, or a copy of x0, when not given)."""
    from benchmarks_amd import _lib
    if x is None:
        x = torch.zeros_like(b) if x0 is None else x0.clone()
    res = _lib.CgResult()
    st = getattr(_lib.lib(), fn)(op._h, C.c_void_p(x.data_ptr()), C.c_void_p(b.data_ptr()),
                                 None if inv is None else C.c_void_p(inv.data_ptr()), 100000,
                                 1e-300, rel_tol, C.byref(res),
                                 C.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return st, x, res


def _problem(p, n, geometry, seed):
    from benchmarks_amd.laplace import LaplaceOperator, random_vector
    op = LaplaceOperator(p, p + 1, O.GLL, n, deform=0.05, geometry=geometry)
    b = random_vector(op.m(), seed)
    op.zero_constrained(b)
    return op, b, op.compute_inverse_diagonal()


def test_noalloc_cg_equals_cg_on_two_calls_in_a_row(gpu):
    from benchmarks_amd import _lib
    L = _lib.lib()
    assert L.mfx_cg_noalloc_max_dofs() == 1 << 21
    # a few dozen steps each: the atomic summation order cannot move a step
    # count this short across the threshold
    big = _problem(4, (4, 3, 3), 1, 3)  # 17*13*13 = 2873 DoFs, patch kernel
    small = _problem(2, (5, 4, 3), 0, 4)  # 11*9*7 = 693 DoFs, per-cell kernel
    # the first call on a device looks up the array's address (and may load
    # the code object that holds it); every later call takes no memory
    st, _, _ = _call("mfx_cg_solve_noalloc", small[0], small[1], small[2])
    assert st == 0
    xs = [torch.zeros_like(big[1]), torch.zeros_like(small[1])]
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    out = []
    for (op, b, inv), x in zip((big, small), xs):
        st, x, res = _call("mfx_cg_solve_noalloc", op, b, inv, x=x)
        assert st == 0, L.mfx_last_error()
        out.append((x, res))
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0, (free0, free1)
    for (op, b, inv), (x, res) in zip((big, small), out):
        st, xr, rr = _call("mfx_cg_solve", op, b, inv)
        assert st == 0
        assert res.converged == 1 and res.steps == rr.steps, (res.steps, rr.steps)
        assert res.initial_value == pytest.approx(rr.initial_value, rel=1e-14)
        assert res.last_value <= 1e-10 * res.initial_value
        assert (torch.linalg.norm(x - xr) / torch.linalg.norm(xr)).item() <= 1e-12


def test_noalloc_cg_identity_preconditioner_and_start_vector(gpu):
    """PreconditionIdentity (no inverse diagonal) from a nonzero start, and a
    step-limited solve, as mfx_cg_solve does them."""
    op, b, _ = _problem(3, (4, 5, 3), 0, 6)
    x0 = torch.as_tensor(O.splitmix64_uniform(7, op.m()), device="cuda")
    st, x, res = _call("mfx_cg_solve_noalloc", op, b, None, x0=x0)
    st_r, xr, rr = _call("mfx_cg_solve", op, b, None, x0=x0)
    assert st == st_r == 0 and res.steps == rr.steps
    assert (torch.linalg.norm(x - xr) / torch.linalg.norm(xr)).item() <= 1e-12
    from benchmarks_amd import _lib
    res = _lib.CgResult()
    x = torch.zeros_like(b)
    st = _lib.lib().mfx_cg_solve_noalloc(op._h, C.c_void_p(x.data_ptr()), C.c_void_p(b.data_ptr()),
                                         None, 3, 1e-300, 1e-14, C.byref(res), None)
    assert st == 4 and res.steps == 3  # MFX_ERR_NOT_CONVERGED at the step limit


def test_noalloc_cg_rejects_operators_over_the_limit(gpu):
    from benchmarks_amd import _lib
    op, b, inv = _problem(2, 64, 1, 5)  # 129^3 = 2,146,689 DoFs > 2^21
    assert op.m() > _lib.lib().mfx_cg_noalloc_max_dofs()
    st, x, _ = _call("mfx_cg_solve_noalloc", op, b, inv)
    assert st == 1  # MFX_ERR_INVALID_ARG
    assert b"limit" in _lib.lib().mfx_last_error()
    assert float(x.abs().max()) == 0.0
