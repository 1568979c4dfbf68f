// batch_kernels.hip -- vector and scalar steps of the batched CG: many
// independent problems concatenated into one DoF vector, each with its own
// CG scalars and convergence state.
//
// Vector kernels run one workgroup per segment (a contiguous range inside
// one instance), skip segments of converged instances, and write one partial
// per segment.  Their segment loops are unrolled by 4 so that each thread has
// several independent loads in flight.  The scalar kernels run one workgroup per instance, which sums
// that instance's segment partials in a fixed order, so every reduction is
// deterministic.
#include <algorithm>

#include "batch.hpp"
#include "mfx_internal.hpp"

namespace mfx {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ void block_sum2(double &a, double &b) {
  __shared__ double ra[kBlock / 64], rb[kBlock / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    b += __shfl_xor(b, o, 64);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    ra[w] = a;
    rb[w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sa = 0.0, sb = 0.0;
#pragma unroll
    for (int i = 0; i < kBlock / 64; ++i) {
      sa += ra[i];
      sb += rb[i];
    }
    a = sa;
    b = sb;
  }
}

__global__ void __launch_bounds__(kBlock)
k_b_residual(const Seg *__restrict__ segs, double *__restrict__ r, const double *__restrict__ b,
             const double *__restrict__ t, const double *__restrict__ dinv,
             double *__restrict__ part_rr, double *__restrict__ part_rz) {
  const Seg sg = segs[blockIdx.x];
  double rr = 0.0, rz = 0.0;
#pragma unroll 4
  for (uint64_t i = sg.begin + threadIdx.x; i < sg.end; i += kBlock) {
    const double v = b[i] - t[i];
    r[i] = v;
    rr += v * v;
    rz += v * dinv[i] * v;
  }
  block_sum2(rr, rz);
  if (threadIdx.x == 0) {
    part_rr[blockIdx.x] = rr;
    part_rz[blockIdx.x] = rz;
  }
}

// Sums of an instance's segment partials: one workgroup per instance,
// threads stride over the segments, fixed-order tree within the workgroup.
__device__ __forceinline__ void instance_sums(const uint32_t *__restrict__ inst_seg, uint32_t i,
                                              const double *__restrict__ pa,
                                              const double *__restrict__ pb, double &a,
                                              double &b) {
  a = 0.0;
  b = 0.0;
  for (uint32_t k = inst_seg[i] + threadIdx.x; k < inst_seg[i + 1]; k += kBlock) {
    a += pa[k];
    if (pb) b += pb[k];
  }
  block_sum2(a, b);  // valid in thread 0
}

__global__ void __launch_bounds__(kBlock)
k_b_init(BatchCg *__restrict__ st, const uint32_t *__restrict__ inst_seg,
         const double *__restrict__ part_rr, const double *__restrict__ part_rz, double abs_tol,
         double rel_tol, int max_steps, int *__restrict__ flags) {
  const uint32_t i = blockIdx.x;
  double rr, rz;
  instance_sums(inst_seg, i, part_rr, part_rz, rr, rz);
  if (threadIdx.x == 0) {
    BatchCg c{};
    c.initial = sqrt(rr);
    c.last = c.initial;
    c.tol = fmax(abs_tol, rel_tol * c.initial);
    c.gamma = rz;
    c.max_steps = max_steps;
    c.first = 1;
    if (c.initial <= c.tol) {
      c.done = 1;
      c.converged = 1;
    } else if (max_steps <= 0) {
      c.done = 1;
    }
    st[i] = c;
    flags[i] = c.done;
  }
}

// number of instances still iterating
__global__ void __launch_bounds__(kBlock)
k_b_count(const int *__restrict__ flags, uint32_t n_inst, int *__restrict__ n_active) {
  double a = 0.0, unused = 0.0;
  for (uint32_t i = threadIdx.x; i < n_inst; i += kBlock) a += flags[i] ? 0.0 : 1.0;
  block_sum2(a, unused);
  if (threadIdx.x == 0) *n_active = (int)a;
}

// x += alpha p of the previous step (it needs the p read here anyway), then
// p = dinv r + beta p; ap = 0 on the free rows (the apply accumulates
// there) and ap = p on the constrained rows (the identity rows of the
// operator), whose share of p.ap, the sum of p^2 over them, goes to one
// partial per segment
__global__ void __launch_bounds__(kBlock)
k_b_direction(const Seg *__restrict__ segs, const BatchCg *__restrict__ st,
              const double *__restrict__ r, const double *__restrict__ dinv,
              const uint8_t *__restrict__ cmask, double *__restrict__ p, double *__restrict__ ap,
              double *__restrict__ x, double *__restrict__ part_c) {
  const Seg sg = segs[blockIdx.x];
  const BatchCg &c = st[sg.inst];
  if (c.done) return;
  const bool first = c.first != 0;
  const double beta = first ? 0.0 : c.beta, alpha = c.alpha;
  double pc = 0.0, unused = 0.0;
#pragma unroll 4
  for (uint64_t i = sg.begin + threadIdx.x; i < sg.end; i += kBlock) {
    const double pi = p[i];
    if (!first) x[i] += alpha * pi;
    const double pn = dinv[i] * r[i] + (beta != 0.0 ? beta * pi : 0.0);
    p[i] = pn;
    const bool cons = cmask[i] != 0;
    ap[i] = cons ? pn : 0.0;
    pc += cons ? pn * pn : 0.0;
  }
  block_sum2(pc, unused);
  if (threadIdx.x == 0) part_c[blockIdx.x] = pc;
}

// the last step's x += alpha p of every instance that took a step
__global__ void __launch_bounds__(kBlock)
k_b_finish(const Seg *__restrict__ segs, const BatchCg *__restrict__ st, double *__restrict__ x,
           const double *__restrict__ p) {
  const Seg sg = segs[blockIdx.x];
  const BatchCg &c = st[sg.inst];
  if (c.steps == 0) return;
  const double alpha = c.alpha;
#pragma unroll 4
  for (uint64_t i = sg.begin + threadIdx.x; i < sg.end; i += kBlock) x[i] += alpha * p[i];
}

__global__ void __launch_bounds__(kBlock)
k_b_pap(const Seg *__restrict__ segs, const BatchCg *__restrict__ st,
        const double *__restrict__ p, const double *__restrict__ ap, double *__restrict__ part) {
  const Seg sg = segs[blockIdx.x];
  if (st[sg.inst].done) return;
  double s = 0.0, unused = 0.0;
#pragma unroll 4
  for (uint64_t i = sg.begin + threadIdx.x; i < sg.end; i += kBlock) s += p[i] * ap[i];
  block_sum2(s, unused);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kBlock)
k_b_alpha(BatchCg *__restrict__ st, const uint32_t *__restrict__ inst_seg,
          const double *__restrict__ part) {
  const uint32_t i = blockIdx.x;
  if (st[i].done) return;
  double pap, unused;
  instance_sums(inst_seg, i, part, nullptr, pap, unused);
  if (threadIdx.x == 0) {
    BatchCg &c = st[i];
    if (!(pap > 0.0)) {
      c.breakdown = 1;
      c.alpha = 0.0;
    } else {
      c.alpha = c.gamma / pap;
    }
  }
}

__global__ void __launch_bounds__(kBlock)
k_b_alpha_fused(BatchCg *__restrict__ st, const PapRange *__restrict__ ranges,
                const uint32_t *__restrict__ inst_seg, const double *__restrict__ part_c) {
  const uint32_t i = blockIdx.x;
  if (st[i].done) return;
  const PapRange rg = ranges[i];
  double a = 0.0, b = 0.0;
  for (uint64_t k = threadIdx.x; k < rg.n; k += kBlock) {
    a += rg.part[k];
    rg.part[k] = 0.0;  // the patch kernel writes only the slots it uses
  }
  for (uint32_t k = inst_seg[i] + threadIdx.x; k < inst_seg[i + 1]; k += kBlock) b += part_c[k];
  block_sum2(a, b);
  if (threadIdx.x == 0) {
    BatchCg &c = st[i];
    const double pap = a + b;
    if (!(pap > 0.0)) {
      c.breakdown = 1;
      c.alpha = 0.0;
    } else {
      c.alpha = c.gamma / pap;
    }
  }
}

// r -= alpha ap (x += alpha p follows in the next direction pass or k_b_finish)
__global__ void __launch_bounds__(kBlock)
k_b_update(const Seg *__restrict__ segs, const BatchCg *__restrict__ st, double *__restrict__ r,
           const double *__restrict__ ap, const double *__restrict__ dinv,
           double *__restrict__ part_rr, double *__restrict__ part_rz) {
  const Seg sg = segs[blockIdx.x];
  const BatchCg &c = st[sg.inst];
  if (c.done) return;
  const double alpha = c.alpha;
  double rr = 0.0, rz = 0.0;
#pragma unroll 4
  for (uint64_t i = sg.begin + threadIdx.x; i < sg.end; i += kBlock) {
    const double v = r[i] - alpha * ap[i];
    r[i] = v;
    rr += v * v;
    rz += v * dinv[i] * v;
  }
  block_sum2(rr, rz);
  if (threadIdx.x == 0) {
    part_rr[blockIdx.x] = rr;
    part_rz[blockIdx.x] = rz;
  }
}

__global__ void __launch_bounds__(kBlock)
k_b_check(BatchCg *__restrict__ st, const uint32_t *__restrict__ inst_seg,
          const double *__restrict__ part_rr, const double *__restrict__ part_rz,
          int *__restrict__ flags) {
  const uint32_t i = blockIdx.x;
  if (st[i].done) return;
  double rr, rz;
  instance_sums(inst_seg, i, part_rr, part_rz, rr, rz);
  if (threadIdx.x == 0) {
    BatchCg &c = st[i];
    c.steps += 1;
    c.last = sqrt(rr);
    if (c.breakdown) {
      c.done = 1;
    } else if (c.last <= c.tol) {
      c.done = 1;
      c.converged = 1;
    } else if (c.steps >= c.max_steps) {
      c.done = 1;
    } else {
      c.beta = rz / c.gamma;
      c.gamma = rz;
      c.first = 0;
    }
    flags[i] = c.done;
  }
}

__global__ void k_offset_map(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst,
                             uint64_t n, uint32_t offset) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = src[i];
    dst[i] = v == kInvalid ? kInvalid : v + offset;
  }
}

__global__ void k_fill_u16(uint16_t *__restrict__ dst, uint16_t v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = v;
}

__global__ void k_mark_u8(const uint32_t *__restrict__ list, uint64_t n, uint8_t *__restrict__ mask) {
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (uint64_t)gridDim.x * blockDim.x)
    mask[list[k]] = 1;
}

inline unsigned grid_for(uint64_t n) {
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + kBlock - 1) / kBlock, 65536));
}

}  // namespace

void launch_b_residual(const Seg *segs, uint32_t n_segs, double *r, const double *b,
                       const double *t, const double *dinv, double *part_rr, double *part_rz,
                       hipStream_t s) {
  if (!n_segs) return;
  hipLaunchKernelGGL(k_b_residual, dim3(n_segs), dim3(kBlock), 0, s, segs, r, b, t, dinv,
                     part_rr, part_rz);
  MFX_HIP(hipGetLastError());
}

void launch_b_init(BatchCg *st, const uint32_t *inst_seg, uint32_t n_inst, const double *part_rr,
                   const double *part_rz, double abs_tol, double rel_tol, int max_steps,
                   int *flags, int *n_active, hipStream_t s) {
  hipLaunchKernelGGL(k_b_init, dim3(n_inst), dim3(kBlock), 0, s, st, inst_seg, part_rr, part_rz,
                     abs_tol, rel_tol, max_steps, flags);
  hipLaunchKernelGGL(k_b_count, dim3(1), dim3(kBlock), 0, s, flags, n_inst, n_active);
  MFX_HIP(hipGetLastError());
}

void launch_b_direction(const Seg *segs, uint32_t n_segs, const BatchCg *st, const double *r,
                        const double *dinv, const uint8_t *cmask, double *p, double *ap, double *x,
                        double *part_c, hipStream_t s) {
  if (!n_segs) return;
  hipLaunchKernelGGL(k_b_direction, dim3(n_segs), dim3(kBlock), 0, s, segs, st, r, dinv, cmask, p,
                     ap, x, part_c);
  MFX_HIP(hipGetLastError());
}

void launch_b_finish(const Seg *segs, uint32_t n_segs, const BatchCg *st, double *x,
                     const double *p, hipStream_t s) {
  if (!n_segs) return;
  hipLaunchKernelGGL(k_b_finish, dim3(n_segs), dim3(kBlock), 0, s, segs, st, x, p);
  MFX_HIP(hipGetLastError());
}

void launch_b_pap(const Seg *segs, uint32_t n_segs, const BatchCg *st, const double *p,
                  const double *ap, double *part, hipStream_t s) {
  if (!n_segs) return;
  hipLaunchKernelGGL(k_b_pap, dim3(n_segs), dim3(kBlock), 0, s, segs, st, p, ap, part);
  MFX_HIP(hipGetLastError());
}

void launch_b_alpha(BatchCg *st, const uint32_t *inst_seg, uint32_t n_inst, const double *part,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_b_alpha, dim3(n_inst), dim3(kBlock), 0, s, st, inst_seg, part);
  MFX_HIP(hipGetLastError());
}

void launch_b_alpha_fused(BatchCg *st, const PapRange *ranges, const uint32_t *inst_seg,
                          uint32_t n_inst, const double *part_c, hipStream_t s) {
  hipLaunchKernelGGL(k_b_alpha_fused, dim3(n_inst), dim3(kBlock), 0, s, st, ranges, inst_seg,
                     part_c);
  MFX_HIP(hipGetLastError());
}

void launch_mark_u8(const uint32_t *list, uint64_t n, uint8_t *mask, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_mark_u8, dim3(grid_for(n)), dim3(kBlock), 0, s, list, n, mask);
  MFX_HIP(hipGetLastError());
}

void launch_b_update(const Seg *segs, uint32_t n_segs, const BatchCg *st, double *r,
                     const double *ap, const double *dinv, double *part_rr, double *part_rz,
                     hipStream_t s) {
  if (!n_segs) return;
  hipLaunchKernelGGL(k_b_update, dim3(n_segs), dim3(kBlock), 0, s, segs, st, r, ap, dinv, part_rr,
                     part_rz);
  MFX_HIP(hipGetLastError());
}

void launch_b_check(BatchCg *st, const uint32_t *inst_seg, uint32_t n_inst, const double *part_rr,
                    const double *part_rz, int *flags, int *n_active, hipStream_t s) {
  hipLaunchKernelGGL(k_b_check, dim3(n_inst), dim3(kBlock), 0, s, st, inst_seg, part_rr, part_rz,
                     flags);
  hipLaunchKernelGGL(k_b_count, dim3(1), dim3(kBlock), 0, s, flags, n_inst, n_active);
  MFX_HIP(hipGetLastError());
}

void launch_offset_map(const uint32_t *src, uint32_t *dst, uint64_t n, uint32_t offset,
                       hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_offset_map, dim3(grid_for(n)), dim3(kBlock), 0, s, src, dst, n, offset);
  MFX_HIP(hipGetLastError());
}

void launch_fill_u16(uint16_t *dst, uint16_t v, uint64_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill_u16, dim3(grid_for(n)), dim3(kBlock), 0, s, dst, v, n);
  MFX_HIP(hipGetLastError());
}

}  // namespace mfx
