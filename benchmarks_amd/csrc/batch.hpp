// batch.hpp -- kernels of the batched CG (batch_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mfx {

// A contiguous DoF range of one instance; vector kernels run one workgroup
// per segment, so every partial sum belongs to exactly one instance.
struct Seg {
  uint64_t begin, end;
  uint32_t inst, pad;
};

// An instance's p.Ap partials from the batched patch kernel: `n`
// consecutive doubles from `part` (workgroups x waves of its degree's launch).
struct PapRange {
  double *part;  // zeroed again by the alpha kernel after reading
  uint64_t n;
};

// Per-instance CG state (deal.II SolverCG + ReductionControl semantics).
struct BatchCg {
  double gamma, beta, alpha, tol, initial, last;
  int steps, max_steps, done, converged, breakdown, first;
};

// r = b - t; partial sums of r.r and r.(dinv r) per segment
void launch_b_residual(const Seg *segs, uint32_t n_segs, double *r, const double *b,
                       const double *t, const double *dinv, double *part_rr, double *part_rz,
                       hipStream_t s);
// per instance: ||r0||, tolerance, gamma = r.Pr, done when already converged;
// flags[i] = done, *n_active = instances still iterating
void launch_b_init(BatchCg *st, const uint32_t *inst_seg, uint32_t n_inst, const double *part_rr,
                   const double *part_rz, double abs_tol, double rel_tol, int max_steps,
                   int *flags, int *n_active, hipStream_t s);
// x += alpha p of the previous step (not on the first), p = dinv r + beta p
// (beta = 0 on the first step); ap = 0 on the free rows and ap = p on the
// constrained ones (cmask[i] != 0: the operator's identity rows), whose
// sum of p^2 goes to part_c per segment
void launch_b_direction(const Seg *segs, uint32_t n_segs, const BatchCg *st, const double *r,
                        const double *dinv, const uint8_t *cmask, double *p, double *ap, double *x,
                        double *part_c, hipStream_t s);
// after the iteration: the last step's x += alpha p of every instance that took a step
void launch_b_finish(const Seg *segs, uint32_t n_segs, const BatchCg *st, double *x,
                     const double *p, hipStream_t s);
// partial sums of p.ap per segment
void launch_b_pap(const Seg *segs, uint32_t n_segs, const BatchCg *st, const double *p,
                  const double *ap, double *part, hipStream_t s);
// alpha = gamma / p.ap per instance
void launch_b_alpha(BatchCg *st, const uint32_t *inst_seg, uint32_t n_inst, const double *part,
                    hipStream_t s);
// alpha = gamma / p.ap per instance, p.ap = the instance's patch-kernel
// partials (in index order, then set to zero for the next step) + its
// segments' constrained-row partials of launch_b_direction (in order)
void launch_b_alpha_fused(BatchCg *st, const PapRange *ranges, const uint32_t *inst_seg,
                          uint32_t n_inst, const double *part_c, hipStream_t s);
// mask[list[k]] = 1 for k < n
void launch_mark_u8(const uint32_t *list, uint64_t n, uint8_t *mask, hipStream_t s);
// r -= alpha ap; partial sums of r.r and r.(dinv r)
void launch_b_update(const Seg *segs, uint32_t n_segs, const BatchCg *st, double *r,
                     const double *ap, const double *dinv, double *part_rr, double *part_rz,
                     hipStream_t s);
// convergence test, beta, gamma; flags and *n_active as in launch_b_init
void launch_b_check(BatchCg *st, const uint32_t *inst_seg, uint32_t n_inst, const double *part_rr,
                    const double *part_rz, int *flags, int *n_active, hipStream_t s);
// dst[i] = src[i] + offset for valid entries (invalid stays invalid)
void launch_offset_map(const uint32_t *src, uint32_t *dst, uint64_t n, uint32_t offset,
                       hipStream_t s);
void launch_fill_u16(uint16_t *dst, uint16_t v, uint64_t n, hipStream_t s);

}  // namespace mfx
