// vector.hip -- vector kernels and the device-resident conjugate-gradient
// steps.  All reductions are two-stage and deterministic: every workgroup
// writes one partial (fixed association inside the group), and a single
// workgroup sums the partials in index order.  CG scalars (alpha, beta,
// gamma, the convergence test) never leave the GPU; a `done` flag makes
// every CG kernel return at once after convergence, so the host can enqueue
// iterations in batches and check the flag once per batch.
#include "mfx_internal.hpp"
#include "cg_state.hpp"

namespace mfx {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Workgroup sum of up to two values; results valid in thread 0.
__device__ __forceinline__ void block_sum2(double &a, double &b) {
  __shared__ double ra[kBlock / 64], rb[kBlock / 64];
  a = wave_sum(a);
  b = wave_sum(b);
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

// Sum partials[0..n) in a fixed order with one workgroup.
__device__ __forceinline__ double sum_partials(const double *__restrict__ part, uint32_t n) {
  double s = 0.0, unused = 0.0;
  for (uint32_t i = threadIdx.x; i < n; i += kBlock) s += part[i];
  block_sum2(s, unused);
  return s;  // valid in thread 0
}

__global__ void k_fill(double *x, double v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    x[i] = v;
}
__global__ void k_fill_f32(float *x, float v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    x[i] = v;
}

template <typename T>
__global__ void k_copy_listed(T *__restrict__ dst, const T *__restrict__ src,
                              const uint32_t *__restrict__ list, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t j = list[i];
    dst[j] = src[j];
  }
}

__global__ void k_set_listed(double *__restrict__ dst, double v,
                             const uint32_t *__restrict__ list, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[list[i]] = v;
}

__global__ void k_invert_diagonal(double *d, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    d[i] = d[i] > 0.0 ? 1.0 / d[i] : 1.0;
}

// SplitMix64 stream, element i uses state seed + (i+1) * golden gamma.
__global__ void k_random(double *x, uint64_t n, uint64_t seed) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    x[i] = (double)(z >> 11) * 0x1.0p-53 * 2.0 - 1.0;
  }
}

// Sum of sine modes at given points (x, y, z interleaved):
// out_i = sum_m a_m sin(pi k0_m x_i) sin(pi k1_m y_i) sin(pi k2_m z_i),
// the modes summed in index order.
__global__ void k_sine_modes(const double *__restrict__ xyz, uint64_t n, SineModes m,
                             double *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    double f = 0.0;
    for (int k = 0; k < m.n; ++k)
      f += m.a[k] * (sinpi(m.k[3 * k] * x) * sinpi(m.k[3 * k + 1] * y) * sinpi(m.k[3 * k + 2] * z));
    out[i] = f;
  }
}

__global__ void k_gather(const uint32_t *__restrict__ map, uint64_t n,
                         const double *__restrict__ src, double *__restrict__ evec) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t j = map[i];
    evec[i] = (j == kInvalid) ? 0.0 : src[j];
  }
}

__global__ void k_scatter_add(const uint32_t *__restrict__ map, uint64_t n,
                              const double *__restrict__ evec, double *__restrict__ dst) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t j = map[i];
    if (j != kInvalid) unsafeAtomicAdd(dst + j, evec[i]);
  }
}

__global__ void k_dot_partials(const double *__restrict__ x, const double *__restrict__ y,
                               uint64_t n, double *__restrict__ part) {
  double s = 0.0, unused = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    s += x[i] * y[i];
  block_sum2(s, unused);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_dot_final(const double *__restrict__ part, uint32_t n, double *result) {
  const double s = sum_partials(part, n);
  if (threadIdx.x == 0) *result = s;
}

// ---------------------------------------------------------------------------
// CG (deal.II SolverCG semantics, see cg_state.hpp)
// ---------------------------------------------------------------------------

// r = b - r (r holds A x0 on entry); partials of r.r and r.(P r).
__global__ void k_cg_init(CgState *st, const double *__restrict__ b, double *__restrict__ r,
                          const double *__restrict__ invd, uint64_t n, double *__restrict__ part) {
  double rr = 0.0, rz = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const double ri = b[i] - r[i];
    r[i] = ri;
    rr += ri * ri;
    rz += ri * (invd ? invd[i] * ri : ri);
  }
  block_sum2(rr, rz);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = rr;
    part[gridDim.x + blockIdx.x] = rz;
  }
}

// ||r_0||, tolerance, step-0 check.
__global__ void k_cg_start(CgState *st, const double *__restrict__ part, uint32_t nb,
                           double abs_tol, double rel_tol, int max_steps) {
  const double rr = sum_partials(part, nb);
  __syncthreads();
  const double rz = sum_partials(part + nb, nb);
  if (threadIdx.x == 0) {
    const double rn = sqrt(rr);
    st->initial = rn;
    st->last = rn;
    st->tol = fmax(abs_tol, rel_tol * rn);
    st->gamma = rz;
    st->beta = 0.0;
    st->first = 1;
    st->steps = 0;
    st->max_steps = max_steps;
    st->breakdown = 0;
    if (st->history && st->history_len > 0) st->history[0] = rn;
    // deal.II's SolverControl::check fails at step 0 when max_steps <= 0
    st->done = (rn <= st->tol) ? 1 : ((isnan(rn) || max_steps <= 0) ? 2 : 0);
  }
}

// x += alpha p (the previous step's update, which needs the old p this
// kernel reads anyway), then p = P r + beta p (p = P r on the first step);
// Ap = 0.
__global__ void k_cg_direction(const CgState *st, const double *__restrict__ r,
                               const double *__restrict__ invd, double *__restrict__ p,
                               double *__restrict__ ap, double *__restrict__ x, uint64_t n) {
  if (st->done) return;
  const double beta = st->beta, alpha = st->alpha;
  const bool first = st->first != 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const double v = invd ? invd[i] * r[i] : r[i];
    if (first) {
      p[i] = v;
    } else {
      const double pi = p[i];
      x[i] += alpha * pi;
      p[i] = v + beta * pi;
    }
    ap[i] = 0.0;
  }
}

// The last step's x += alpha p, after the iteration has stopped.
__global__ void k_cg_finish(const CgState *st, double *__restrict__ x,
                            const double *__restrict__ p, uint64_t n) {
  if (st->steps == 0) return;
  const double alpha = st->alpha;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    x[i] += alpha * p[i];
}

// Identity rows: Ap_c = p_c, partial sum of p_c^2 (the constrained part of p.Ap).
__global__ void k_cg_constrained(const CgState *st, const uint32_t *__restrict__ list,
                                 uint64_t nc, const double *__restrict__ p,
                                 double *__restrict__ ap, double *__restrict__ part) {
  if (st->done) return;
  double s = 0.0, unused = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nc;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t j = list[i];
    const double v = p[j];
    ap[j] = v;
    s += v * v;
  }
  block_sum2(s, unused);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// alpha = gamma / p.Ap
// sums of chunks of the partials: chunk b = part[b chunk, (b + 1) chunk)
__global__ void k_cg_chunks(const CgState *st, const double *__restrict__ part, uint32_t n,
                            uint32_t chunk, double *__restrict__ out) {
  if (st->done) return;
  const uint32_t b0 = blockIdx.x * chunk, b1 = min(n, b0 + chunk);
  double s = 0.0, unused = 0.0;
  for (uint32_t i = b0 + threadIdx.x; i < b1; i += kBlock) s += part[i];
  block_sum2(s, unused);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

__global__ void k_cg_alpha(CgState *st, const double *__restrict__ part, uint32_t n) {
  if (st->done) return;
  const double pap = sum_partials(part, n);
  if (threadIdx.x == 0) {
    st->alpha = st->gamma / pap;
    if (!(pap > 0.0) && !(pap < 0.0)) st->breakdown = 1;
  }
}

// r -= alpha Ap; partials of r.r and r.(P r).  (x += alpha p is applied by
// the next direction update, or by k_cg_finish after the last step.)
__global__ void k_cg_update(const CgState *st, double *__restrict__ r,
                            const double *__restrict__ ap, const double *__restrict__ invd,
                            uint64_t n, double *__restrict__ part) {
  if (st->done) return;
  const double alpha = st->alpha;
  double rr = 0.0, rz = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const double ri = r[i] - alpha * ap[i];
    r[i] = ri;
    rr += ri * ri;
    rz += ri * (invd ? invd[i] * ri : ri);
  }
  block_sum2(rr, rz);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = rr;
    part[gridDim.x + blockIdx.x] = rz;
  }
}

// Convergence test of step k and beta for step k+1.
__global__ void k_cg_check(CgState *st, const double *__restrict__ part, uint32_t nb) {
  if (st->done) return;
  const double rr = sum_partials(part, nb);
  __syncthreads();
  const double rz = sum_partials(part + nb, nb);
  if (threadIdx.x == 0) {
    const double rn = sqrt(rr);
    const int it = st->steps + 1;
    st->steps = it;
    st->last = rn;
    if (st->history && it < st->history_len) st->history[it] = rn;
    if (rn <= st->tol) {
      st->done = 1;
    } else if (it >= st->max_steps || isnan(rn) || st->breakdown) {
      st->done = 2;
    } else {
      st->beta = rz / st->gamma;
      st->gamma = rz;
      st->first = 0;
    }
  }
}

inline unsigned vec_grid(uint64_t n) {
  const uint64_t g = (n + kBlock * 4 - 1) / (kBlock * 4);
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(g, 4096));
}

}  // namespace

void launch_fill(double *x, double v, uint64_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill, dim3(vec_grid(n)), dim3(kBlock), 0, s, x, v, n);
  MFX_HIP(hipGetLastError());
}
void launch_fill_f32(float *x, float v, uint64_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill_f32, dim3(vec_grid(n)), dim3(kBlock), 0, s, x, v, n);
  MFX_HIP(hipGetLastError());
}
template <typename T>
void launch_copy_listed(T *dst, const T *src, const uint32_t *list, uint64_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_copy_listed<T>, dim3(vec_grid(n)), dim3(kBlock), 0, s, dst, src, list, n);
  MFX_HIP(hipGetLastError());
}
template void launch_copy_listed<double>(double *, const double *, const uint32_t *, uint64_t,
                                         hipStream_t);
template void launch_copy_listed<float>(float *, const float *, const uint32_t *, uint64_t,
                                        hipStream_t);
void launch_set_listed(double *dst, double v, const uint32_t *list, uint64_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_set_listed, dim3(vec_grid(n)), dim3(kBlock), 0, s, dst, v, list, n);
  MFX_HIP(hipGetLastError());
}
void launch_invert_diagonal(double *d, uint64_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_invert_diagonal, dim3(vec_grid(n)), dim3(kBlock), 0, s, d, n);
  MFX_HIP(hipGetLastError());
}
void launch_random(double *x, uint64_t n, uint64_t seed, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_random, dim3(vec_grid(n)), dim3(kBlock), 0, s, x, n, seed);
  MFX_HIP(hipGetLastError());
}
void launch_gather(const uint32_t *map, uint64_t n, const double *src, double *evec,
                   hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_gather, dim3(vec_grid(n)), dim3(kBlock), 0, s, map, n, src, evec);
  MFX_HIP(hipGetLastError());
}
void launch_scatter_add(const uint32_t *map, uint64_t n, const double *evec, double *dst,
                        hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_scatter_add, dim3(vec_grid(n)), dim3(kBlock), 0, s, map, n, evec, dst);
  MFX_HIP(hipGetLastError());
}
void launch_sine_modes(const double *xyz, uint64_t n, const SineModes &m, double *out,
                       hipStream_t s) {
  MFX_REQUIRE(m.n >= 0 && m.n <= SineModes::kMax, "sine modes: count out of range");
  if (!n) return;
  hipLaunchKernelGGL(k_sine_modes, dim3(vec_grid(n)), dim3(kBlock), 0, s, xyz, n, m, out);
  MFX_HIP(hipGetLastError());
}
uint32_t dot_n_blocks(uint64_t n) { return vec_grid(n); }
void launch_dot(const double *x, const double *y, uint64_t n, double *partials, double *result,
                hipStream_t s) {
  const unsigned g = vec_grid(n);
  hipLaunchKernelGGL(k_dot_partials, dim3(g), dim3(kBlock), 0, s, x, y, n, partials);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kBlock), 0, s, partials, g, result);
  MFX_HIP(hipGetLastError());
}

// ---- CG launchers ----
uint32_t cg_vec_blocks(uint64_t n) { return vec_grid(n); }
uint32_t cg_cons_blocks(uint64_t nc) { return nc ? std::min<uint32_t>(vec_grid(nc), 256) : 0; }

void cg_init(CgState *st, const double *b, double *r, const double *invd, uint64_t n,
             double *part, double abs_tol, double rel_tol, int max_steps, hipStream_t s) {
  const unsigned g = vec_grid(n);
  hipLaunchKernelGGL(k_cg_init, dim3(g), dim3(kBlock), 0, s, st, b, r, invd, n, part);
  hipLaunchKernelGGL(k_cg_start, dim3(1), dim3(kBlock), 0, s, st, part, g, abs_tol, rel_tol,
                     max_steps);
  MFX_HIP(hipGetLastError());
}
void cg_direction(const CgState *st, const double *r, const double *invd, double *p, double *ap,
                  double *x, uint64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_direction, dim3(vec_grid(n)), dim3(kBlock), 0, s, st, r, invd, p, ap, x,
                     n);
}
void cg_finish(const CgState *st, double *x, const double *p, uint64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_finish, dim3(vec_grid(n)), dim3(kBlock), 0, s, st, x, p, n);
  MFX_HIP(hipGetLastError());
}
void cg_constrained(const CgState *st, const uint32_t *list, uint64_t nc, const double *p,
                    double *ap, double *part, hipStream_t s) {
  if (!nc) return;
  hipLaunchKernelGGL(k_cg_constrained, dim3(cg_cons_blocks(nc)), dim3(kBlock), 0, s, st, list,
                     nc, p, ap, part);
}
// More partials than one workgroup sums quickly (the cell-wise Gauss apply
// writes one per workgroup: ~1.6e5 at 67M DoFs, 0.23 ms in k_cg_alpha
// alone): chunks of kAlphaChunk are summed in parallel first, each in a
// fixed order, into the slots after the partials, and alpha sums those.
constexpr uint32_t kAlphaDirect = 4096, kAlphaChunk = 2048;
uint32_t cg_alpha_extra(uint32_t n) {
  return n > kAlphaDirect ? (n + kAlphaChunk - 1) / kAlphaChunk : 0;
}
void cg_alpha(CgState *st, double *part, uint32_t n, hipStream_t s) {
  const uint32_t r = cg_alpha_extra(n);
  if (r) {
    hipLaunchKernelGGL(k_cg_chunks, dim3(r), dim3(kBlock), 0, s, st, part, n, kAlphaChunk,
                       part + n);
    hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(kBlock), 0, s, st, part + n, r);
  } else {
    hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(kBlock), 0, s, st, part, n);
  }
}
void cg_update(const CgState *st, double *r, const double *ap, const double *invd, uint64_t n,
               double *part, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_update, dim3(vec_grid(n)), dim3(kBlock), 0, s, st, r, ap, invd, n, part);
}
void cg_check(CgState *st, const double *part, uint32_t nb, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_check, dim3(1), dim3(kBlock), 0, s, st, part, nb);
}

}  // namespace mfx
