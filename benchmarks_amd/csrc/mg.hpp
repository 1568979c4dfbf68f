// mg.hpp -- launchers of the multigrid kernels (mg_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/mfx.h"

namespace mfx {

// operator applications with vmult semantics (zero, apply, identity rows)
void vmult64(mfx_op &op, double *dst, const double *src, hipStream_t s);
void vmult32(mfx_op &op, float *dst, const float *src, hipStream_t s);
// `degree` Chebyshev(Jacobi) steps towards A x = b in FP32 on [lmin, lmax]
// (t, d: scratch vectors of the operator's size)
void chebyshev(mfx_op &op, float *x, const float *b, const float *dinv, float *t, float *d,
               int degree, double lmin, double lmax, bool zero_initial, hipStream_t s);
// largest eigenvalue of D^-1 A from `iters` CG (Lanczos) steps, FP64
double estimate_lambda_max(mfx_op &op, const double *inv_diag, int iters, hipStream_t s);

// A 1D transfer matrix in CSR form (device arrays), applied along one axis of
// a lexicographic (x fastest) node grid.
template <typename T>
struct AxisMatrix {
  int rows = 0, cols = 0;
  const int *rowptr = nullptr;
  const int *col = nullptr;
  const T *val = nullptr;
};

// out[i0,i1,i2] (+)= sum_k M[i_axis][k] in[..k..]: `out_dims` are the output
// grid's node counts, the input has M.cols nodes along `axis`.
template <typename T>
void launch_axis_pass(const AxisMatrix<T> &M, int axis, const uint32_t out_dims[3], const T *in,
                      T *out, bool add, hipStream_t s);

// One Chebyshev step on FP32 vectors:
//   r = b - t (r = b when t is null), d = c1 d + c2 dinv r, x = x + d
// (x = d when x_zero: the iterate was zero).  c1 = 0 starts the recurrence.
void launch_cheb_step(float *x, float *d, const float *b, const float *t, const float *dinv,
                      float c1, float c2, bool x_zero, uint64_t n, hipStream_t s);
// r = b - t
void launch_residual_f32(float *r, const float *b, const float *t, uint64_t n, hipStream_t s);
void launch_to_double(const float *in, double *out, uint64_t n, hipStream_t s);
void launch_set_listed_f32(float *dst, float v, const uint32_t *list, uint64_t n, hipStream_t s);
// x[free[i]] = sum_j Ainv[i][j] b[free[j]] in FP64 (the other entries of x are
// left alone)
void launch_dense_solve(const double *Ainv, const uint32_t *free_idx, uint32_t n_free,
                        const float *b, float *x, hipStream_t s);
// x += alpha p, r -= alpha ap
void launch_pcg_update(double *x, double *r, const double *p, const double *ap, double alpha,
                       uint64_t n, hipStream_t s);
// z = dinv * r (elementwise)
void launch_scale(double *z, const double *r, const double *dinv, uint64_t n, hipStream_t s);
// r = b - t
void launch_residual_f64(double *r, const double *b, const double *t, uint64_t n, hipStream_t s);
// y += a x
void launch_axpy_f64(double *y, double a, const double *x, uint64_t n, hipStream_t s);
// p = z + beta p
void launch_pcg_direction(double *p, const double *z, double beta, uint64_t n, hipStream_t s);

}  // namespace mfx
