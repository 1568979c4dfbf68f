// ref_harness.cc -- C entry points around the reference's own header-only
// serial kernels and Lagrange basis, compiled where they lie in the reference
// tree (include root passed by oracle/Makefile).  The output library lives in
// oracle/_ref/ and is test infrastructure: it produces golden vectors and the
// timed CPU baseline; the product never loads it.
//
// Wrapped reference code:
//   BK3::Serial::SumFactorization  CEED_BK/include/kernels/BK3/serial_kernels.hpp:9-197
//   BK5::Serial::SumFactorization  sum_factorization/include/kernels/BK5/serial_kernels.hpp:9-87
//   LagrangeBasis                  sum_factorization/include/math/lagrange_basis.hpp:8-70
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "CEED_BK/include/kernels/BK3/serial_kernels.hpp"
#include "sum_factorization/include/kernels/BK5/serial_kernels.hpp"
#include "sum_factorization/include/math/lagrange_basis.hpp"

extern "C" {

// Element layout of the reference: in/out[e][i][j][k], nm = nq - 1.
double ref_bk3_serial(unsigned nq, unsigned nelmt, const double *basis,
                      const double *dbasis, const double *G, double *in,
                      double *out) {
  return BK3::Serial::SumFactorization<double>(nq, nq, nq, nelmt, basis, basis,
                                               basis, dbasis, dbasis, dbasis,
                                               G, in, out);
}

double ref_bk5_serial(unsigned nq, unsigned nelmt, const double *dbasis,
                      const double *G, const double *in, double *out) {
  return BK5::Serial::SumFactorization<double>(nq, nq, nq, nelmt, dbasis,
                                               dbasis, dbasis, G, in, out);
}

double ref_lagrange_value(const double *nodes, int n, int j, double x) {
  LagrangeBasis<double> basis(std::vector<double>(nodes, nodes + n));
  return basis(j, x);
}

double ref_lagrange_derivative(const double *nodes, int n, int j, double x) {
  LagrangeBasis<double> basis(std::vector<double>(nodes, nodes + n));
  return basis.derivative(j, x);
}

float ref_lagrange_value_f32(const float *nodes, int n, int j, float x) {
  LagrangeBasis<float> basis(std::vector<float>(nodes, nodes + n));
  return basis(j, x);
}

}  // extern "C"
