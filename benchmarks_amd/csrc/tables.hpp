// tables.hpp -- compile-time 1D tables of the line-form patch kernel
// (patch.hip) and the even-odd matrix application it uses.
//
// The GLL nodes and collocation derivative matrix D on [0, 1] are evaluated
// by constexpr code, so that the kernel takes the entries as scalar
// immediates (no loads, nothing to spill).  gll_d_table_error() compares
// them with the run-time basis tables (basis.cpp), which every operator
// creation checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace mfx {

// 1/x from the hardware reciprocal estimate refined by Newton steps
// (e = 1 - x r; r += r e), two steps in FP64, one in FP32.
__device__ __forceinline__ double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}
__device__ __forceinline__ float recip(float x) {
  float r = __builtin_amdgcn_rcpf(x);
  const float e = __builtin_fmaf(-x, r, 1.0f);
  return __builtin_fmaf(r, e, r);
}

constexpr double cx_cos(double t) {  // Taylor series, |t| <= pi
  double s = 1.0, term = 1.0;
  for (int k = 1; k < 40; ++k) {
    term *= -t * t / ((2.0 * k - 1.0) * (2.0 * k));
    s += term;
  }
  return s;
}

// Gauss-Lobatto-Legendre nodes on [0, 1]: the Legendre-recurrence Newton
// iteration x <- x - (x P_m(x) - P_{m-1}(x)) / (n P_m(x)), m = n - 1, from
// the Chebyshev-Gauss-Lobatto points; its fixed points are the end points
// and the roots of P_m'.  Mapped to [0, 1] and made exactly symmetric.
// D[i][j] = l_j'(x_i) by barycentric differentiation.
template <int N>
struct GllD {
  double x[N] = {};
  double D[N][N] = {};
  constexpr GllD() {
    const double pi = 3.14159265358979323846;
    double z[N] = {};
    for (int i = 0; i < N; ++i) z[i] = cx_cos(pi * i / (N - 1));
    for (int it = 0; it < 100; ++it)
      for (int i = 0; i < N; ++i) {
        double p0 = 1.0, p1 = z[i];
        for (int k = 2; k < N; ++k) {
          const double p2 = ((2.0 * k - 1.0) * z[i] * p1 - (k - 1.0) * p0) / k;
          p0 = p1;
          p1 = p2;
        }
        z[i] -= (z[i] * p1 - p0) / (N * p1);
      }
    for (int i = 0; i < N; ++i) x[i] = 0.5 * (1.0 - z[i]);
    for (int i = 0; i < N / 2; ++i) {
      const double h = 0.5 * (x[i] + (1.0 - x[N - 1 - i]));
      x[i] = h;
      x[N - 1 - i] = 1.0 - h;
    }
    x[0] = 0.0;
    x[N - 1] = 1.0;
    if (N & 1) x[N / 2] = 0.5;
    derivative_matrix(x, D);
  }
  // D[i][j] = (w_j / w_i) / (x_i - x_j), D[i][i] = -sum_{j != i} D[i][j],
  // w_j = 1 / prod_{k != j} (x_j - x_k)
  static constexpr void derivative_matrix(const double (&x)[N], double (&D)[N][N]) {
    double w[N] = {};
    for (int j = 0; j < N; ++j) {
      double pr = 1.0;
      for (int k = 0; k < N; ++k)
        if (k != j) pr *= x[j] - x[k];
      w[j] = 1.0 / pr;
    }
    for (int i = 0; i < N; ++i) {
      double sum = 0.0;
      for (int j = 0; j < N; ++j)
        if (j != i) {
          D[i][j] = (w[j] / w[i]) / (x[i] - x[j]);
          sum += D[i][j];
        }
      D[i][i] = -sum;
    }
  }
};

// out[i] = sum_a M(i, a) in[a] for an O x I matrix with centro-symmetry
// M(O-1-i, I-1-a) = S M(i, a), S = +1 (interpolation B) or -1 (derivative
// D), by the even-odd split: with e_a = in[a] + in[m-a], o_a = in[a] -
// in[m-a] (a < I/2, m = I-1), the half sums and differences
//   P_i = (out[i] + out[O-1-i]) / 2,  Q_i = (out[i] - out[O-1-i]) / 2
// are products of I/2-wide matrices with e and o (the middle input of odd I
// joins the even part), i < O/2; the middle output of odd O is one such
// product.  About half the multiply-adds of the plain product; M(i, a) must
// be a compile-time constant after unrolling.
template <int O, int I, int S, typename T, typename M>
__device__ __forceinline__ void eo_mat(M mat, const T *in, T *out) {
  static_assert(S == 1 || S == -1, "centro-symmetry sign");
  constexpr int HI = I / 2, HO = O / 2, mi = I - 1, mo = O - 1;
  T e[HI > 0 ? HI : 1], o[HI > 0 ? HI : 1];
#pragma unroll
  for (int a = 0; a < HI; ++a) {
    e[a] = in[a] + in[mi - a];
    o[a] = in[a] - in[mi - a];
  }
#pragma unroll
  for (int i = 0; i < HO; ++i) {
    T p = 0, q = 0;
#pragma unroll
    for (int a = 0; a < HI; ++a) {
      const T cf = T(0.5 * (mat(i, a) + S * mat(i, mi - a)));
      const T cg = T(0.5 * (mat(i, a) - S * mat(i, mi - a)));
      if constexpr (S == 1) {
        p += cf * e[a];
        q += cg * o[a];
      } else {
        p += cf * o[a];
        q += cg * e[a];
      }
    }
    if constexpr (I & 1) {
      if constexpr (S == 1)
        p += T(mat(i, HI)) * in[HI];
      else
        q += T(mat(i, HI)) * in[HI];
    }
    out[i] = p + q;
    out[mo - i] = p - q;
  }
  if constexpr (O & 1) {
    T v = 0;
#pragma unroll
    for (int a = 0; a < HI; ++a) v += T(mat(HO, a)) * (S == 1 ? e[a] : o[a]);
    if constexpr ((I & 1) && S == 1) v += T(mat(HO, HI)) * in[HI];
    out[HO] = v;
  }
}

// Even-odd tables of a centro-symmetric (S = 1) or centro-antisymmetric
// (S = -1) O x I matrix M, as eo_mat forms them, stored in this order:
//   cf[i][a] = (M[i][a] + S M[i][I-1-a]) / 2, i < O/2, a < I/2
//   cg[i][a] = (M[i][a] - S M[i][I-1-a]) / 2
//   M[i][I/2] (I odd), i < O/2;  M[O/2][a] (O odd), a < I/2;
//   M[O/2][I/2] (both odd).
constexpr int eo_table_size(int O, int I) {
  return 2 * (O / 2) * (I / 2) + ((I & 1) ? O / 2 : 0) + ((O & 1) ? I / 2 : 0) +
         (((I & 1) && (O & 1)) ? 1 : 0);
}

// out = M in with M given by its even-odd tables (eo_table_size(O, I)
// values at tab, read with compile-time offsets, so through the scalar cache)
template <int O, int I, int S, typename T>
__device__ __forceinline__ void eo_tab(const T *__restrict__ tab, const T *in, T *out) {
  static_assert(S == 1 || S == -1, "centro-symmetry sign");
  constexpr int HI = I / 2, HO = O / 2, mi = I - 1, mo = O - 1;
  constexpr int CG = HO * HI, MC = 2 * HO * HI, MR = MC + ((I & 1) ? HO : 0), CT = MR + HI;
  T e[HI > 0 ? HI : 1], o[HI > 0 ? HI : 1];
#pragma unroll
  for (int a = 0; a < HI; ++a) {
    e[a] = in[a] + in[mi - a];
    o[a] = in[a] - in[mi - a];
  }
#pragma unroll
  for (int i = 0; i < HO; ++i) {
    T p = 0, q = 0;
#pragma unroll
    for (int a = 0; a < HI; ++a) {
      if constexpr (S == 1) {
        p += tab[i * HI + a] * e[a];
        q += tab[CG + i * HI + a] * o[a];
      } else {
        p += tab[i * HI + a] * o[a];
        q += tab[CG + i * HI + a] * e[a];
      }
    }
    if constexpr (I & 1) {
      if constexpr (S == 1)
        p += tab[MC + i] * in[HI];
      else
        q += tab[MC + i] * in[HI];
    }
    out[i] = p + q;
    out[mo - i] = p - q;
  }
  if constexpr (O & 1) {
    T v = 0;
#pragma unroll
    for (int a = 0; a < HI; ++a) v += tab[MR + a] * (S == 1 ? e[a] : o[a]);
    if constexpr ((I & 1) && S == 1) v += tab[CT] * in[HI];
    out[HO] = v;
  }
}

// max |table - run-time| of the compile-time GLL derivative matrix of n
// points (D row-major n x n)
double gll_d_table_error(int n, const double *D);
// Even-odd tables (eo_tab's layout) of the row-major O x I matrix M with
// centro-symmetry sign S, appended to out; returns the largest deviation of
// M from that symmetry relative to its largest entry.
double eo_tables(const double *M, int O, int I, int S, std::vector<double> &out);

}  // namespace mfx
