// mg_kernels.hip -- kernels of the multigrid preconditioner: separable grid
// transfers, fused Chebyshev steps, precision conversion, the dense FP64
// coarse solve and the vector updates of the outer preconditioned CG.
//
// Transfers.  On a structured box with lexicographic node numbering, both
// the p-transfer (same cells, degree p_f -> p_c) and the h-transfer (Q_p on
// n cells -> n/2 cells) are tensor products of 1D interpolation matrices,
// P = P_z (x) P_y (x) P_x, so prolongation is three passes of a 1D matrix
// along one axis each, and restriction R = P^T is the same three passes with
// the transposed 1D matrices.  Each pass reads its input once per stencil
// entry from a neighbourhood of consecutive threads (coalesced along x for
// every axis) and writes its output once.
#include <algorithm>

#include "mfx_internal.hpp"
#include "mg.hpp"

namespace mfx {
namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(uint64_t n) {
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + kBlock - 1) / kBlock, 65536));
}

template <typename T, bool ADD>
__global__ void k_axis_pass(const T *__restrict__ in, T *__restrict__ out, uint32_t n0, uint32_t n1,
                            uint32_t n2, int axis, uint32_t in_len, const int *__restrict__ rowptr,
                            const int *__restrict__ col, const T *__restrict__ val) {
  const uint64_t total = (uint64_t)n0 * n1 * n2;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t i0 = (uint32_t)(t % n0);
    const uint64_t q = t / n0;
    const uint32_t i1 = (uint32_t)(q % n1), i2 = (uint32_t)(q / n1);
    uint32_t row;
    uint64_t base, stride;
    if (axis == 0) {
      row = i0;
      base = (uint64_t)in_len * q;
      stride = 1;
    } else if (axis == 1) {
      row = i1;
      base = i0 + (uint64_t)n0 * in_len * i2;
      stride = n0;
    } else {
      row = i2;
      base = i0 + (uint64_t)n0 * i1;
      stride = (uint64_t)n0 * n1;
    }
    T acc = 0;
    for (int k = rowptr[row]; k < rowptr[row + 1]; ++k) acc += val[k] * in[base + stride * col[k]];
    if (ADD)
      out[t] += acc;
    else
      out[t] = acc;
  }
}

__global__ void k_cheb_step(float *__restrict__ x, float *__restrict__ d,
                            const float *__restrict__ b, const float *__restrict__ t,
                            const float *__restrict__ dinv, float c1, float c2, int x_zero,
                            uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const float r = t ? b[i] - t[i] : b[i];
    const float dn = (c1 != 0.0f ? c1 * d[i] : 0.0f) + c2 * dinv[i] * r;
    d[i] = dn;
    x[i] = x_zero ? dn : x[i] + dn;
  }
}

__global__ void k_residual_f32(float *__restrict__ r, const float *__restrict__ b,
                               const float *__restrict__ t, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    r[i] = b[i] - t[i];
}

__global__ void k_to_double(const float *__restrict__ in, double *__restrict__ out, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (double)in[i];
}

__global__ void k_set_listed_f32(float *__restrict__ dst, float v, const uint32_t *__restrict__ list,
                                 uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[list[i]] = v;
}

// one workgroup per row of the coarse inverse
__global__ void __launch_bounds__(kBlock)
k_dense_solve(const double *__restrict__ Ainv, const uint32_t *__restrict__ free_idx,
              uint32_t n_free, const float *__restrict__ b, float *__restrict__ x) {
  __shared__ double red[kBlock / 64];
  const uint32_t row = blockIdx.x;
  double acc = 0.0;
  for (uint32_t j = threadIdx.x; j < n_free; j += kBlock)
    acc += Ainv[(size_t)row * n_free + j] * (double)b[free_idx[j]];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) s += red[w];
    x[free_idx[row]] = (float)s;
  }
}

__global__ void k_pcg_update(double *__restrict__ x, double *__restrict__ r,
                             const double *__restrict__ p, const double *__restrict__ ap,
                             double alpha, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    r[i] -= alpha * ap[i];
  }
}

__global__ void k_pcg_direction(double *__restrict__ p, const double *__restrict__ z, double beta,
                                uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = z[i] + beta * p[i];
}

__global__ void k_axpy_f64(double *__restrict__ y, double a, const double *__restrict__ x, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    y[i] += a * x[i];
}

__global__ void k_scale(double *__restrict__ z, const double *__restrict__ r,
                        const double *__restrict__ dinv, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    z[i] = dinv[i] * r[i];
}

__global__ void k_residual_f64(double *__restrict__ r, const double *__restrict__ b,
                               const double *__restrict__ t, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    r[i] = b[i] - t[i];
}

}  // namespace

template <typename T>
void launch_axis_pass(const AxisMatrix<T> &M, int axis, const uint32_t out_dims[3], const T *in,
                      T *out, bool add, hipStream_t s) {
  MFX_REQUIRE(axis >= 0 && axis < 3, "axis must be 0, 1 or 2");
  MFX_REQUIRE(M.rowptr && M.col && M.val && in && out, "transfer pass: null pointer");
  MFX_REQUIRE((uint32_t)M.rows == out_dims[axis], "transfer pass: matrix rows != output length");
  const uint64_t n = (uint64_t)out_dims[0] * out_dims[1] * out_dims[2];
  if (n == 0) return;
  if (add)
    hipLaunchKernelGGL((k_axis_pass<T, true>), dim3(grid_for(n)), dim3(kBlock), 0, s, in, out,
                       out_dims[0], out_dims[1], out_dims[2], axis, (uint32_t)M.cols, M.rowptr,
                       M.col, M.val);
  else
    hipLaunchKernelGGL((k_axis_pass<T, false>), dim3(grid_for(n)), dim3(kBlock), 0, s, in, out,
                       out_dims[0], out_dims[1], out_dims[2], axis, (uint32_t)M.cols, M.rowptr,
                       M.col, M.val);
  MFX_HIP(hipGetLastError());
}
template void launch_axis_pass<float>(const AxisMatrix<float> &, int, const uint32_t[3],
                                      const float *, float *, bool, hipStream_t);
template void launch_axis_pass<double>(const AxisMatrix<double> &, int, const uint32_t[3],
                                       const double *, double *, bool, hipStream_t);

void launch_cheb_step(float *x, float *d, const float *b, const float *t, const float *dinv,
                      float c1, float c2, bool x_zero, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_cheb_step, dim3(grid_for(n)), dim3(kBlock), 0, s, x, d, b, t, dinv, c1, c2,
                     (int)x_zero, n);
  MFX_HIP(hipGetLastError());
}

void launch_residual_f32(float *r, const float *b, const float *t, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_residual_f32, dim3(grid_for(n)), dim3(kBlock), 0, s, r, b, t, n);
  MFX_HIP(hipGetLastError());
}

void launch_to_double(const float *in, double *out, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_to_double, dim3(grid_for(n)), dim3(kBlock), 0, s, in, out, n);
  MFX_HIP(hipGetLastError());
}

void launch_set_listed_f32(float *dst, float v, const uint32_t *list, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_set_listed_f32, dim3(grid_for(n)), dim3(kBlock), 0, s, dst, v, list, n);
  MFX_HIP(hipGetLastError());
}

void launch_dense_solve(const double *Ainv, const uint32_t *free_idx, uint32_t n_free,
                        const float *b, float *x, hipStream_t s) {
  if (n_free == 0) return;
  hipLaunchKernelGGL(k_dense_solve, dim3(n_free), dim3(kBlock), 0, s, Ainv, free_idx, n_free, b,
                     x);
  MFX_HIP(hipGetLastError());
}

void launch_pcg_update(double *x, double *r, const double *p, const double *ap, double alpha,
                       uint64_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_pcg_update, dim3(grid_for(n)), dim3(kBlock), 0, s, x, r, p, ap, alpha, n);
  MFX_HIP(hipGetLastError());
}

void launch_axpy_f64(double *y, double a, const double *x, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_axpy_f64, dim3(grid_for(n)), dim3(kBlock), 0, s, y, a, x, n);
  MFX_HIP(hipGetLastError());
}

void launch_scale(double *z, const double *r, const double *dinv, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(kBlock), 0, s, z, r, dinv, n);
  MFX_HIP(hipGetLastError());
}

void launch_residual_f64(double *r, const double *b, const double *t, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_residual_f64, dim3(grid_for(n)), dim3(kBlock), 0, s, r, b, t, n);
  MFX_HIP(hipGetLastError());
}

void launch_pcg_direction(double *p, const double *z, double beta, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_pcg_direction, dim3(grid_for(n)), dim3(kBlock), 0, s, p, z, beta, n);
  MFX_HIP(hipGetLastError());
}

}  // namespace mfx
