// bp_driver.cc -- the reference's BP3 timing driver on the mfx C++ API.
//
// Mirrors LaplaceProblem::run/solve of CEED_bp/src/bp3.cc:246-329, 423-534:
// for each mesh cycle c (n_refine = c/3, rem = c%3, box [-1,-1,-1]-p2 with
// p2_d = 2.8 and 2 subdivisions for d < rem, else 0.9 and 1; :452-474) whose
// DoF count lies in [minsize, maxsize]: right-hand side f = 1, 10 CG solves
// with ReductionControl(1e9, 1e-16, 1e-9) and PreconditionIdentity (best
// time kept), then 5 x n_mv matvecs (n_mv = 200 below 1e7 DoFs, else 50),
// and the same convergence table.
//
//   bp_driver <degree> [minsize] [maxsize] [nq]      (nq defaults to degree+2)
//   bp_driver --bp5 <degree> <n> [reps]              BP5: GLL, deformed unit cube,
//                                                    Jacobi-CG to 1e-8 from a random rhs
//   bp_driver --mg <degree> <n> [nq]                 mixed-precision BP3: Gauss nq
//                                                    (degree+2), deformed unit cube,
//                                                    f = 3 pi^2 prod sin(pi x_d), FP64 CG
//                                                    to 1e-10 preconditioned by the FP32
//                                                    multigrid V-cycle, next to
//                                                    Jacobi-CG to 1e-8
//   bp_driver --batch <count> [seed]                 batched BP5 solves: count instances,
//                                                    cells per direction log-uniform in
//                                                    [4, 32], degree uniform in 2..6
//                                                    (SplitMix64 seed, default 2),
//                                                    per-instance Jacobi-CG to 1e-8
//   bp_driver --fcg <degree> <n> <smoother_degree>   BP5 on n^3 deformed cells, flexible
//                                                    CG to 1e-8 with the learned
//                                                    preconditioner's Chebyshev cycle
//                                                    (network off), next to Jacobi-CG
//   bp_driver --verify <bk3|bk5> <nq> <nelmt>        the verification fixture of
//       CEED_BK/src/BK3/serial_verification.cc:25-79 and
//       sum_factorization/tests/BK5/serial_verification.cc:16-47 (G = 2, in = 3,
//       cos() shape tables) through the GPU E-vector kernels; prints the norm
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/mfx.hpp"

using Clock = std::chrono::steady_clock;
using VectorD = mfx::Vector<double>;

static void check_hip_sync() { mfx::check_hip(hipDeviceSynchronize()); }

static double seconds_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

static int run_bp3(unsigned degree, std::size_t minsize, std::size_t maxsize, unsigned nq) {
  std::printf("Testing FE_Q<3>(%u), QGauss(%u)\nNo. of GPUs: 1\n", degree, nq);
  std::printf(" cells     dofs    matvec   CG_tot_time CG_time/iters cg_its cg_reduction\n");
  for (unsigned cycle = 0; cycle < 38; ++cycle) {
    const unsigned n_refine = cycle / 3, remainder = cycle % 3;
    mfx::BoxMesh mesh;
    std::size_t projected = 1;
    for (unsigned d = 0; d < 3; ++d) {
      mesh.p1[d] = -1.0;
      mesh.p2[d] = d < remainder ? 2.8 : 0.9;
      mesh.n_cells[d] = (d < remainder ? 2u : 1u) << n_refine;
      projected *= std::size_t(mesh.n_cells[d]) * degree + 1;
    }
    if (projected < minsize) continue;
    if (projected > maxsize) {
      std::printf("Projected size %zu higher than max size, terminating.\n", projected);
      break;
    }
    mfx::LaplaceOperator<double> op(mesh, degree, nq, mfx::Quadrature::gauss);
    VectorD x, b;
    op.initialize_dof_vector(x);
    op.compute_rhs(b, 0);
    double time_cg = 1e10;
    unsigned its = 0;
    double reduction = 0;
    for (int i = 0; i < 10; ++i) {
      mfx::ReductionControl control(1000000000u, 1e-16, 1e-9);
      mfx::SolverCG<VectorD> cg(control);
      x = 0.;
      check_hip_sync();
      const auto t0 = Clock::now();
      cg.solve(op, x, b, mfx::PreconditionIdentity());
      time_cg = std::min(time_cg, seconds_since(t0));
      its = control.last_step();
      reduction = std::pow(control.last_value() / control.initial_value(), 1.0 / its);
    }
    double best_mv = 1e10;
    for (int i = 0; i < 5; ++i) {
      const unsigned n_mv = op.m() < 10000000 ? 200 : 50;
      check_hip_sync();
      const auto t0 = Clock::now();
      for (unsigned k = 0; k < n_mv; ++k) op.vmult(x, b);
      check_hip_sync();
      best_mv = std::min(best_mv, seconds_since(t0) / n_mv);
    }
    std::printf("%7u %9llu %.3e %.3e %.3e %6u %.3e\n", op.n_cells(),
                (unsigned long long)op.m(), best_mv, time_cg, time_cg / its, its, reduction);
    std::printf("RESULT cycle=%u cells=%u dofs=%llu matvec_s=%.6e dofs_per_s=%.6e cg_its=%u "
                "cg_s=%.6e\n",
                cycle, op.n_cells(), (unsigned long long)op.m(), best_mv, op.m() / best_mv, its,
                time_cg);
    std::fflush(stdout);
  }
  return 0;
}

static int run_bp5(unsigned degree, unsigned n, int reps) {
  mfx::BoxMesh mesh;
  for (int d = 0; d < 3; ++d) mesh.n_cells[d] = n;
  mesh.deform = 0.05;
  mfx::LaplaceOperator<double> op(mesh, degree, degree + 1, mfx::Quadrature::gauss_lobatto);
  VectorD b, x;
  op.initialize_dof_vector(b);
  op.initialize_dof_vector(x);
  mfx::check(mfx_vec_random(b.get_values(), b.size(), 1, nullptr));
  mfx::check(mfx_vec_zero_constrained(op.handle(), b.get_values(), nullptr));
  mfx::DiagonalMatrix<VectorD> jacobi;
  op.compute_inverse_diagonal(jacobi.get_vector());
  double best = 1e10;
  unsigned its = 0;
  for (int r = 0; r < reps; ++r) {
    mfx::ReductionControl control(100000, 1e-300, 1e-8);
    mfx::SolverCG<VectorD> cg(control);
    x = 0.;
    check_hip_sync();
    const auto t0 = Clock::now();
    cg.solve(op, x, b, jacobi);
    best = std::min(best, seconds_since(t0));
    its = control.last_step();
  }
  double best_mv = 1e10;
  for (int i = 0; i < 4; ++i) {
    check_hip_sync();
    const auto t0 = Clock::now();
    for (int k = 0; k < 50; ++k) op.vmult(x, b);
    check_hip_sync();
    best_mv = std::min(best_mv, seconds_since(t0) / 50);
  }
  std::printf(" p |  q | n_element |     n_dofs |     time/it |   dofs/s/it | itCG | time/matvec\n");
  std::printf("%2u | %2u |%10u |%11llu | %11.4e | %11.4e | %4u | %11.4e\n", degree, degree + 1,
              op.n_cells(), (unsigned long long)op.m(), best / its, op.m() / best * its, its,
              best_mv);
  std::printf("RESULT bp5 p=%u cells=%u dofs=%llu matvec_s=%.6e cg_its=%u cg_s=%.6e\n", degree,
              op.n_cells(), (unsigned long long)op.m(), best_mv, its, best);
  return 0;
}

// Mixed-precision multigrid CG against Jacobi-CG on one BP3 problem.
static int run_mg(unsigned degree, unsigned n, unsigned nq) {
  mfx::BoxMesh mesh;
  for (int d = 0; d < 3; ++d) mesh.n_cells[d] = n;
  mesh.deform = 0.05;
  check_hip_sync();
  auto t0 = Clock::now();
  mfx::PreconditionMG mg(mesh, degree, nq, mfx::Quadrature::gauss);
  check_hip_sync();
  const double t_setup = seconds_since(t0);
  const auto &op = mg.get_operator();
  VectorD b, x, xj;
  op.compute_rhs(b, 1);
  mfx::check(mfx_vec_zero_constrained(op.handle(), b.get_values(), nullptr));
  op.initialize_dof_vector(x);
  op.initialize_dof_vector(xj);
  std::printf("levels:");
  for (unsigned l = 0; l < mg.n_levels(); ++l) {
    const auto info = mg.level(l);
    std::printf(" Q%d/%u^3", info.degree, info.n_cells[0]);
  }
  std::printf("\n");
  double t_mg = 0.0;
  unsigned its_mg = 0;
  for (int r = 0; r < 2; ++r) {  // the first solve loads the kernels
    mfx::ReductionControl control(200, 1e-300, 1e-10);
    mfx::SolverCG<VectorD> cg(control);
    x = 0.;
    check_hip_sync();
    t0 = Clock::now();
    cg.solve(op, x, b, mg);
    t_mg = seconds_since(t0);
    its_mg = control.last_step();
  }
  mfx::DiagonalMatrix<VectorD> jacobi;
  op.compute_inverse_diagonal(jacobi.get_vector());
  mfx::ReductionControl cj(100000, 1e-300, 1e-8);
  mfx::SolverCG<VectorD> cg(cj);
  check_hip_sync();
  t0 = Clock::now();
  cg.solve(op, xj, b, jacobi);
  const double t_j = seconds_since(t0);
  VectorD diff = x;
  std::vector<double> hx = x.to_host(), hj = xj.to_host();
  for (std::size_t i = 0; i < hx.size(); ++i) hx[i] -= hj[i];
  diff.import_elements(hx);
  std::printf("RESULT mg p=%u cells=%u dofs=%llu setup_s=%.6e mg_its=%u mg_s=%.6e jacobi_its=%u "
              "jacobi_s=%.6e rel_diff=%.6e\n",
              degree, op.n_cells(), (unsigned long long)op.m(), t_setup, its_mg, t_mg,
              cj.last_step(), t_j, diff.l2_norm() / xj.l2_norm());
  return 0;
}

// u in [0, 1) from the SplitMix64 stream of `seed` (k-th draw: z = seed +
// k 0x9E3779B97F4A7C15, standard mix, (z >> 11) 2^-53)
struct SplitMix {
  std::uint64_t seed, k = 0;
  double operator()() {
    std::uint64_t z = seed + (++k) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return double(z >> 11) * 0x1.0p-53;
  }
};

static int run_batch(unsigned count, std::uint64_t seed) {
  SplitMix u{seed};
  std::vector<mfx::BatchInstance> inst(count);
  for (auto &in : inst) {
    const double t = u();
    const unsigned n = (unsigned)std::lround(std::exp(std::log(4.0) + t * (std::log(32.0) - std::log(4.0))));
    in.degree = 2 + std::min(unsigned(u() * 5), 4u);
    for (int k = 0; k < 3; ++k) in.mesh.n_cells[k] = n;
    in.mesh.deform = 0.05;
  }
  auto t0 = Clock::now();
  mfx::BatchedLaplace batch(inst, true);
  check_hip_sync();
  const double t_setup = seconds_since(t0);
  VectorD b, x, inv;
  batch.initialize_dof_vector(b);
  batch.initialize_dof_vector(x);
  mfx::check(mfx_vec_random(b.get_values(), b.size(), seed, nullptr));
  const auto off = batch.offsets();
  for (unsigned i = 0; i < count; ++i)
    mfx::check(mfx_vec_zero_constrained(mfx_batch_instance_operator(batch.handle(), (int)i),
                                        b.get_values() + off[i], nullptr));
  batch.compute_inverse_diagonal(inv);
  batch.solve(x, b, &inv, 100000, 1e-300, 1e-2);  // loads the kernels
  x = 0.;
  check_hip_sync();
  t0 = Clock::now();
  const auto res = batch.solve(x, b, &inv, 100000, 1e-300, 1e-8);
  const double t_solve = seconds_since(t0);
  int steps_max = 0, steps_total = 0, converged = 0;
  for (const auto &r : res) {
    steps_max = std::max(steps_max, r.steps);
    steps_total += r.steps;
    converged += r.converged;
  }
  std::printf("variants:");
  for (unsigned p = 2; p <= 6; ++p) {
    mfx_batch_tuning t{};
    if (mfx_batch_get_tuning(batch.handle(), (int)p, &t) != MFX_OK) continue;
    std::printf(" p%u:%s", p, t.geometry[t.variant] == MFX_GEOM_ON_THE_FLY ? "on_the_fly" : "stored");
  }
  std::printf("\nRESULT batch instances=%u dofs=%llu setup_s=%.6e solve_s=%.6e steps_max=%d "
              "steps_total=%d converged=%d\n",
              count, (unsigned long long)batch.m(), t_setup, t_solve, steps_max, steps_total,
              converged);
  return converged == (int)count ? 0 : 1;
}

static int run_fcg(unsigned degree, unsigned n, int smoother_degree) {
  mfx::BoxMesh mesh;
  for (int k = 0; k < 3; ++k) mesh.n_cells[k] = n;
  mesh.deform = 0.05;
  mfx::LaplaceOperator<double> op(mesh, degree, degree + 1, mfx::Quadrature::gauss_lobatto, true,
                                  mfx::Geometry::on_the_fly);
  VectorD b, x;
  op.initialize_dof_vector(b);
  op.initialize_dof_vector(x);
  mfx::check(mfx_vec_random(b.get_values(), b.size(), 1, nullptr));
  mfx::check(mfx_vec_zero_constrained(op.handle(), b.get_values(), nullptr));
  mfx::LearnedPreconditioner P(1);
  P.set_mode(smoother_degree, false);
  mfx::ReductionControl cf(5000, 1e-300, 1e-8);
  mfx::SolverCG<VectorD> fcg(cf);
  fcg.solve(op, x, b, P);
  mfx::DiagonalMatrix<VectorD> jacobi;
  op.compute_inverse_diagonal(jacobi.get_vector());
  mfx::ReductionControl cj(100000, 1e-300, 1e-8);
  mfx::SolverCG<VectorD> cg(cj);
  x = 0.;
  cg.solve(op, x, b, jacobi);
  std::printf("RESULT fcg p=%u cells=%u dofs=%llu smoother_degree=%d fcg_its=%u jacobi_its=%u\n",
              degree, op.n_cells(), (unsigned long long)op.m(), smoother_degree, cf.last_step(),
              cj.last_step());
  return 0;
}

// E-vector kernels on the reference's verification fixture.
static int run_verify(const std::string &kind, unsigned nq, unsigned nelmt) {
  const bool bk5 = kind == "bk5";
  const unsigned nm = bk5 ? nq : nq - 1;
  const std::size_t nm3 = std::size_t(nm) * nm * nm, nq3 = std::size_t(nq) * nq * nq;
  std::vector<double> basis(std::size_t(nq) * nm), dbasis(std::size_t(nq) * nq);
  for (unsigned p = 0; p < nq; ++p)
    for (unsigned i = 0; i < nm; ++i) basis[p * nm + i] = std::cos(double(p * nm + i));
  for (unsigned i = 0; i < nq; ++i)
    for (unsigned p = 0; p < nq; ++p) dbasis[i * nq + p] = std::cos(double(i * nq + p));
  VectorD B(basis.size()), D(dbasis.size()), G(nelmt * 6 * nq3), in(nelmt * nm3), out(nelmt * nm3);
  B.import_elements(basis);
  D.import_elements(dbasis);
  G.import_elements(std::vector<double>(G.size(), 2.0));
  in.import_elements(std::vector<double>(in.size(), 3.0));
  if (bk5)
    mfx::check(mfx_bk5_evec_f64(nq - 1, D.get_values(), G.get_values(), in.get_values(),
                                out.get_values(), nelmt, nullptr));
  else
    mfx::check(mfx_bk3_evec_f64(nm - 1, nq, B.get_values(), D.get_values(), G.get_values(),
                                in.get_values(), out.get_values(), nelmt, nullptr));
  check_hip_sync();
  std::printf("RESULT verify %s nq=%u nelmt=%u norm=%.17g\n", kind.c_str(), nq, nelmt, out.l2_norm());
  return 0;
}

int main(int argc, char **argv) {
  try {
    if (argc > 1 && std::string(argv[1]) == "--verify") {
      if (argc < 5) {
        std::printf("Usage: bp_driver --verify bk3|bk5 nq nelmt\n");
        return 1;
      }
      return run_verify(argv[2], std::atoi(argv[3]), std::atoi(argv[4]));
    }
    if (argc > 1 && std::string(argv[1]) == "--mg") {
      if (argc < 4) {
        std::printf("Usage: bp_driver --mg degree n [nq]\n");
        return 1;
      }
      const unsigned degree = std::atoi(argv[2]);
      return run_mg(degree, std::atoi(argv[3]), argc > 4 ? std::atoi(argv[4]) : degree + 2);
    }
    if (argc > 1 && std::string(argv[1]) == "--batch") {
      if (argc < 3) {
        std::printf("Usage: bp_driver --batch count [seed]\n");
        return 1;
      }
      return run_batch(std::atoi(argv[2]), argc > 3 ? std::atoll(argv[3]) : 2);
    }
    if (argc > 1 && std::string(argv[1]) == "--fcg") {
      if (argc < 5) {
        std::printf("Usage: bp_driver --fcg degree n smoother_degree\n");
        return 1;
      }
      return run_fcg(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]));
    }
    if (argc > 1 && std::string(argv[1]) == "--bp5") {
      if (argc < 4) {
        std::printf("Usage: bp_driver --bp5 degree n [reps]\n");
        return 1;
      }
      return run_bp5(std::atoi(argv[2]), std::atoi(argv[3]), argc > 4 ? std::atoi(argv[4]) : 3);
    }
    if (argc == 1) {
      std::printf("Expected at least one argument.\nUsage:\n./bp_driver degree minsize maxsize [nq]\n");
      return 1;
    }
    const unsigned degree = std::atoi(argv[1]);
    const std::size_t minsize = argc > 2 ? std::atoll(argv[2]) : 1;
    const std::size_t maxsize = argc > 3 ? std::atoll(argv[3]) : std::size_t(-1);
    const unsigned nq = argc > 4 ? std::atoi(argv[4]) : degree + 2;
    return run_bp3(degree, minsize, maxsize, nq);
  } catch (std::exception &exc) {
    std::fprintf(stderr, "\n----------------------------------------------------\n"
                         "Exception on processing: \n%s\nAborting!\n"
                         "----------------------------------------------------\n",
                 exc.what());
    return 1;
  }
}
