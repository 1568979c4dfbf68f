// cg_state.hpp -- device-resident state of one conjugate-gradient solve.
//
// Semantics of deal.II SolverCG + ReductionControl (as called at
// CEED_bp/src/bp3.cc:264-288 and bp5_kokkos/benchmark.cc:355-368):
//   r = b - A x, tol = max(abs_tol, rel_tol ||r_0||_2); stop when
//   ||r_k||_2 <= tol (done = 1) or k >= max_steps / NaN / p.Ap == 0 (done = 2).
//   Step k: v = P r; gamma = r.v; beta = gamma / gamma_old (k > 1);
//   p = v + beta p; v = A p; alpha = gamma / p.v; x += alpha p; r -= alpha v.
// The r.v of step k+1 is computed in the update pass of step k, so each
// step needs two global reductions (p.Ap, and r.r with r.Pr).  The x update
// of step k runs in the direction pass of step k+1, which reads p anyway
// (one pass over N fewer per step), and in cg_finish after the last step.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mfx {

struct CgState {
  double gamma;    // r.(P r) of the current residual
  double beta;     // for the next direction update
  double alpha;    // step length of the current step
  double tol;      // absolute stopping tolerance
  double initial;  // ||r_0||
  double last;     // ||r_k||
  double *history; // optional device array of ||r_k||, k = 0..history_len-1
  int history_len;
  int steps;       // last_step()
  int max_steps;
  int done;        // 0 running, 1 converged, 2 failed
  int first;       // 1 before the first direction update
  int breakdown;   // p.Ap == 0 or NaN
};

uint32_t cg_vec_blocks(uint64_t n);
uint32_t cg_cons_blocks(uint64_t nc);
void cg_init(CgState *st, const double *b, double *r, const double *invd, uint64_t n,
             double *part, double abs_tol, double rel_tol, int max_steps, hipStream_t s);
// x += alpha p of the previous step (not on the first), p = P r + beta p, Ap = 0
void cg_direction(const CgState *st, const double *r, const double *invd, double *p, double *ap,
                  double *x, uint64_t n, hipStream_t s);
void cg_constrained(const CgState *st, const uint32_t *list, uint64_t nc, const double *p,
                    double *ap, double *part, hipStream_t s);
// alpha = gamma / (sum of the n partials of p.Ap); with many partials the
// cg_alpha_extra(n) slots after them hold chunk sums
void cg_alpha(CgState *st, double *part, uint32_t n, hipStream_t s);
uint32_t cg_alpha_extra(uint32_t n);
// r -= alpha Ap with the partials of r.r and r.Pr
void cg_update(const CgState *st, double *r, const double *ap, const double *invd, uint64_t n,
               double *part, hipStream_t s);
void cg_check(CgState *st, const double *part, uint32_t nb, hipStream_t s);
// the last step's x += alpha p (nothing if no step was taken)
void cg_finish(const CgState *st, double *x, const double *p, uint64_t n, hipStream_t s);

}  // namespace mfx
