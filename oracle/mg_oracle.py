"""CPU restatement (numpy, FP64) of the mixed-precision multigrid solver.

Test infrastructure only: tests/ (and nothing in the product) import it.

The reference has no multigrid on its hot path: its Chebyshev/multigrid
include is unused (CEED_bp/src/bp3.cc:27) and its solves are identity/Jacobi
CG (CEED_bp/src/bp3.cc:264-288, bp5_kokkos/benchmark.cc:336-368).  This
module therefore restates published methods, pinned in tests by what the
mathematics fixes:

* nodal-interpolation transfers between Q_pf on nf cells and Q_pc on nc
  cells of one structured box (nf == nc: p-transfer; nf == 2 nc: h-transfer),
  a tensor product P = P_z (x) P_y (x) P_x of 1D matrices, restriction P^T;
* Chebyshev iteration with a Jacobi preconditioner (Saad, Iterative Methods
  for Sparse Linear Systems, 2nd ed., Algorithm 12.1): after k steps from x0
  the error is T_k((theta - D^-1 A)/delta) / T_k(theta/delta) e0, theta and
  delta the centre and half-width of [lambda_min, lambda_max];
* the Lanczos tridiagonal matrix from the coefficients of preconditioned CG
  (Saad, section 6.7.3) for the largest eigenvalue of D^-1 A;
* the V-cycle with pre- and post-smoothing and a direct coarse solve, and
  CG preconditioned by it with deal.II's ReductionControl stopping rule.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import oracle as O


def lagrange_values(p: int, x: float) -> np.ndarray:
    """Degree-p nodal basis on the Gauss-Lobatto support points of [0,1] at x."""
    sup, _ = O.gll_rule(p + 1)
    return np.array([O.lib().orc_lagrange_value(sup, p + 1, m, x) for m in range(p + 1)])


def prolongation_1d(nc: int, pc: int, nf: int, pf: int) -> np.ndarray:
    """Dense 1D prolongation (nf pf + 1) x (nc pc + 1), nf = r nc: coarse
    cell ef // r, reference coordinate (ef mod r + xi_f) / r."""
    assert nf % nc == 0 and (nf > nc or pf > pc)
    r = nf // nc
    supf, _ = O.gll_rule(pf + 1)
    P = np.zeros((nf * pf + 1, nc * pc + 1))
    for j in range(nf * pf + 1):
        ef = min(j // pf, nf - 1)
        l = j - ef * pf
        ec = ef // r
        xi = ((ef % r) + supf[l]) / r
        P[j, ec * pc: ec * pc + pc + 1] += lagrange_values(pc, xi)
    return P


def prolongation_3d(fine: O.Problem, coarse: O.Problem) -> sp.csr_matrix:
    """P = P_z (x) P_y (x) P_x for lexicographic (x fastest) node numbering."""
    Ps = [sp.csr_matrix(prolongation_1d(coarse.box.n[a], coarse.p, fine.box.n[a], fine.p))
          for a in range(3)]
    return sp.kron(Ps[2], sp.kron(Ps[1], Ps[0])).tocsr()


def chebyshev(vmult, dinv, b, x0, degree, lmin, lmax):
    """`degree` Chebyshev(Jacobi) updates of x towards A x = b (FP64)."""
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64)
    r = b - vmult(x) if x0 is not None else b.copy()
    d = dinv * r / theta
    x = x + d
    rho_old = delta / theta
    for _ in range(1, degree):
        r = b - vmult(x)
        rho = 1.0 / (2.0 * theta / delta - rho_old)
        d = rho * rho_old * d + (2.0 * rho / delta) * dinv * r
        x = x + d
        rho_old = rho
    return x


def chebyshev_error_polynomial(lam, degree, lmin, lmax):
    """T_k((theta - lam)/delta) / T_k(theta/delta)."""
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    T = np.polynomial.chebyshev.Chebyshev.basis(degree)
    return T((theta - np.asarray(lam)) / delta) / T(theta / delta)


def lanczos_lambda_max(vmult, dinv, r0, iters):
    """Largest eigenvalue of D^-1 A from `iters` steps of Jacobi-preconditioned
    CG on A x = r0 through the Lanczos matrix T (T_ii = 1/a_i + b_{i-1}/a_{i-1},
    T_{i,i+1} = sqrt(b_i)/a_i)."""
    r = r0.copy()
    z = dinv * r
    p = z.copy()
    gamma = r @ z
    g0 = gamma
    alpha, beta = [], []
    for _ in range(iters):
        if gamma <= 1e-28 * g0:
            break
        ap = vmult(p)
        pap = p @ ap
        if not pap > 0:
            break
        a = gamma / pap
        r = r - a * ap
        z = dinv * r
        gn = r @ z
        bb = gn / gamma
        p = z + bb * p
        gamma = gn
        alpha.append(a)
        beta.append(bb)
    k = len(alpha)
    T = np.zeros((k, k))
    for i in range(k):
        T[i, i] = 1.0 / alpha[i] + (beta[i - 1] / alpha[i - 1] if i > 0 else 0.0)
        if i + 1 < k:
            T[i, i + 1] = T[i + 1, i] = np.sqrt(beta[i]) / alpha[i]
    return float(np.linalg.eigvalsh(T).max())


def level_descriptors(p, nq, quad, n, coarse_max_dofs):
    """(degree, nq, quad, cells) of each level: degrees p, p/2, ..., 1 on the
    fine cells, then Q1 on halved meshes while the free DoFs exceed the
    direct-solve limit and every cell count is even.  Below the fine level a
    Gauss hierarchy uses q = p+1 points (exact for the stiffness of affine
    cells); a GLL one keeps q - p of the fine level."""
    n = tuple(n) if not isinstance(n, int) else (n, n, n)
    extra = 1 if quad == O.GAUSS else nq - p
    out = [(p, nq, quad, n)]
    while p > 1:
        p = max(1, p // 2)
        out.append((p, p + extra, quad, n))

    def free(deg, cells):
        return int(np.prod([c * deg - 1 for c in cells]))
    while free(p, n) > coarse_max_dofs and all(c % 2 == 0 for c in n):
        n = tuple(c // 2 for c in n)
        out.append((p, p + extra, quad, n))
    return out


class Multigrid:
    """FP64 restatement of the product's V-cycle preconditioner."""

    def __init__(self, p, nq, quad, n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), deform=0.0,
                 smoother_degree=5, smoothing_range=15.0, eig_iterations=12, eig_safety=1.2,
                 coarse_max_dofs=2000, eig_seed=0x5EED, fine_smoother_nq=None):
        """fine_smoother_nq: Gauss points of the fine level's smoothing
        operator (Chebyshev, and the residual handed to the next level) when
        it differs from the fine operator's; the outer PCG keeps the fine
        operator."""
        self.degree = smoother_degree
        self.levels = []
        for (dp, dnq, dq, dn) in level_descriptors(p, nq, quad, n, coarse_max_dofs):
            prob = O.make_problem(dp, dnq, dq, dn, lo, hi, deform)
            smooth = prob
            if not self.levels and fine_smoother_nq and fine_smoother_nq != nq:
                smooth = O.make_problem(dp, fine_smoother_nq, dq, dn, lo, hi, deform)
            self.levels.append({"problem": prob, "smooth": smooth,
                                "dinv": smooth.inverse_diagonal()})
        for l in range(len(self.levels) - 1):
            self.levels[l]["P"] = prolongation_3d(self.levels[l]["problem"],
                                                  self.levels[l + 1]["problem"])
        for L in self.levels[:-1]:
            prob = L["smooth"]
            r0 = O.splitmix64_uniform(eig_seed, prob.N)
            r0[prob.constrained.astype(bool)] = 0.0
            L["lmax"] = eig_safety * lanczos_lambda_max(prob.vmult, L["dinv"], r0, eig_iterations)
            L["lmin"] = L["lmax"] / smoothing_range
        C = self.levels[-1]["problem"]
        free = np.flatnonzero(C.constrained == 0)
        A = np.zeros((free.size, free.size))
        for j, dof in enumerate(free):
            e = np.zeros(C.N)
            e[dof] = 1.0
            A[:, j] = C.vmult(e)[free]
        A = 0.5 * (A + A.T)
        self.free = free
        self.Ainv = np.linalg.inv(A) if free.size else A

    def vcycle(self, b, l=0):
        L = self.levels[l]
        prob = L["smooth"]
        if l == len(self.levels) - 1:
            x = np.zeros(prob.N)
            x[self.free] = self.Ainv @ b[self.free]
            return x
        x = chebyshev(prob.vmult, L["dinv"], b, None, self.degree, L["lmin"], L["lmax"])
        r = b - prob.vmult(x)
        bc = L["P"].T @ r
        bc[self.levels[l + 1]["problem"].constrained.astype(bool)] = 0.0
        x = x + L["P"] @ self.vcycle(bc, l + 1)
        return chebyshev(prob.vmult, L["dinv"], b, x, self.degree, L["lmin"], L["lmax"])

    def vmult(self, r):
        """z = V-cycle(r) with constrained rows passed through."""
        prob = self.levels[0]["problem"]
        cons = prob.constrained.astype(bool)
        rb = r.copy()
        rb[cons] = 0.0
        z = self.vcycle(rb)
        z[cons] = r[cons]
        return z

    def pcg(self, b, x0=None, max_steps=1000, abs_tol=1e-300, rel_tol=1e-10):
        """CG on the fine operator preconditioned by one V-cycle per step;
        returns (x, steps, residual history)."""
        A = self.levels[0]["problem"]
        x = np.zeros(A.N) if x0 is None else np.array(x0, dtype=np.float64)
        r = b - A.vmult(x)
        res = np.linalg.norm(r)
        tol = max(abs_tol, rel_tol * res)
        hist = [res]
        if res <= tol:
            return x, 0, hist
        z = self.vmult(r)
        p = z.copy()
        gamma = r @ z
        steps = 0
        while steps < max_steps:
            ap = A.vmult(p)
            alpha = gamma / (p @ ap)
            x += alpha * p
            r -= alpha * ap
            steps += 1
            res = np.linalg.norm(r)
            hist.append(res)
            if res <= tol:
                break
            z = self.vmult(r)
            gn = r @ z
            p = z + (gn / gamma) * p
            gamma = gn
        return x, steps, hist
