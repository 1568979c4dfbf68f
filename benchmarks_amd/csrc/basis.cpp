// basis.cpp -- 1D quadrature rules and Lagrange shape data on [0,1], host side.
//
// The shape data deal.II's MatrixFree hands the reference kernel
// (shape_values, co_shape_gradients; CEED_bp/include/portable_laplace_operator.h:147,
// consumed at CEED_bp/include/bk3_kokkos_kernel.h:156-161, 243-245):
//   B[q][m] = l_m(x_q): FE_Q nodal basis on the p+1 Gauss-Lobatto support
//             points, evaluated at the quadrature points;
//   D[q][k] = derivative at x_q of the Lagrange polynomial through the
//             quadrature points that is 1 at x_k (collocation derivative).
// Lagrange values follow the product formula of LagrangeBasis
// (sum_factorization/include/math/lagrange_basis.hpp:29-65).
#include <algorithm>
#include <cmath>
#include <vector>

#include "mfx_internal.hpp"

namespace mfx {
namespace {

using ld = long double;
constexpr ld kPiL = 3.141592653589793238462643383279502884L;

// Legendre P_n and its derivative at x (|x| < 1) from the recurrence.
void legendre(int n, ld x, ld &P, ld &dP) {
  ld pm1 = 1.0L, p = x;
  if (n == 0) {
    P = 1.0L;
    dP = 0.0L;
    return;
  }
  for (int k = 2; k <= n; ++k) {
    const ld pn = ((2 * k - 1) * x * p - (k - 1) * pm1) / k;
    pm1 = p;
    p = pn;
  }
  P = p;
  dP = n * (x * p - pm1) / (x * x - 1.0L);
}

// Write the symmetric rule (nodes t on [-1,1], weights v) as [0,1] arrays.
void store(int n, const std::vector<ld> &t, const std::vector<ld> &v, double *x, double *w) {
  for (int i = 0; i < n; ++i) {
    x[i] = static_cast<double>(0.5L * (1.0L + t[i]));
    w[i] = static_cast<double>(0.5L * v[i]);
  }
}

}  // namespace

// Gauss-Legendre: roots of P_n (Newton), weights 2 / ((1 - x^2) P_n'^2).
void gauss_rule(int n, double *x, double *w) {
  MFX_REQUIRE(n >= 1 && n <= 32, "gauss_rule: 1 <= n <= 32");
  std::vector<ld> t(n), v(n);
  for (int i = 0; i < (n + 1) / 2; ++i) {
    ld z = std::cos(kPiL * (i + 0.75L) / (n + 0.5L)), P = 0, dP = 1;
    for (int it = 0; it < 64; ++it) {
      legendre(n, z, P, dP);
      const ld dz = P / dP;
      z -= dz;
      if (std::fabs(dz) <= 1e-30L) break;
    }
    legendre(n, z, P, dP);
    const ld wt = 2.0L / ((1.0L - z * z) * dP * dP);
    t[i] = -z;
    t[n - 1 - i] = z;
    v[i] = v[n - 1 - i] = wt;
  }
  if (n & 1) t[n / 2] = 0.0L;
  store(n, t, v, x, w);
}

// Gauss-Lobatto-Legendre: end points and roots of P_{n-1}' (Newton with
// P'' from Legendre's equation), weights 2 / (m (m+1) P_m^2), m = n-1.
void gll_rule(int n, double *x, double *w) {
  MFX_REQUIRE(n >= 2 && n <= 32, "gll_rule: 2 <= n <= 32");
  const int m = n - 1;
  std::vector<ld> t(n), v(n);
  t[0] = -1.0L;
  t[n - 1] = 1.0L;
  for (int i = 1; i <= (n - 1) / 2; ++i) {
    ld z = std::cos(kPiL * i / m);
    for (int it = 0; it < 64; ++it) {
      ld P, dP;
      legendre(m, z, P, dP);
      const ld d2P = (2.0L * z * dP - m * (m + 1.0L) * P) / (1.0L - z * z);
      const ld dz = dP / d2P;
      z -= dz;
      if (std::fabs(dz) <= 1e-30L) break;
    }
    t[i] = -z;
    t[n - 1 - i] = z;
  }
  if (n & 1) t[n / 2] = 0.0L;
  for (int i = 0; i < n; ++i) {
    ld P = 1.0L;
    if (i == 0 || i == n - 1) {
      P = (i == 0 && (m & 1)) ? -1.0L : 1.0L;
    } else {
      ld dP;
      legendre(m, t[i], P, dP);
    }
    v[i] = 2.0L / (m * (m + 1.0L) * P * P);
  }
  for (int i = 0; i < n / 2; ++i) v[i] = v[n - 1 - i] = 0.5L * (v[i] + v[n - 1 - i]);
  store(n, t, v, x, w);
  x[0] = 0.0;
  x[n - 1] = 1.0;
}

namespace {
double lagrange_value(const double *c, int n, int j, double x) {
  double den = 1.0, num = 1.0;
  for (int m = 0; m < n; ++m)
    if (m != j) den *= c[j] - c[m];
  for (int m = 0; m < n; ++m)
    if (m != j) num *= (x - c[m]);
  return num / den;
}
double lagrange_derivative(const double *c, int n, int j, double x) {
  double sum = 0.0;
  for (int i = 0; i < n; ++i) {
    if (i == j) continue;
    double prod = 1.0;
    for (int m = 0; m < n; ++m)
      if (m != i && m != j) prod *= (x - c[m]) / (c[j] - c[m]);
    sum += (1.0 / (c[j] - c[i])) * prod;
  }
  return sum;
}
}  // namespace

void lagrange_values(int p, double x, double *values) {
  MFX_REQUIRE(p >= 1 && p <= 8, "degree must be in 1..8");
  std::vector<double> sup(p + 1), sw(p + 1);
  gll_rule(p + 1, sup.data(), sw.data());
  for (int m = 0; m <= p; ++m) values[m] = lagrange_value(sup.data(), p + 1, m, x);
}

void basis_tables(int p, int nq, int quad, double *B, double *D, double *xq, double *wq,
                  double *support) {
  MFX_REQUIRE(p >= 1 && p <= 8, "degree must be in 1..8");
  MFX_REQUIRE(nq >= p + 1 && nq <= p + 2, "n_q_points_1d must be p+1 or p+2");
  const int nm = p + 1;
  std::vector<double> sw(nm);
  gll_rule(nm, support, sw.data());
  if (quad == MFX_QUAD_GAUSS)
    gauss_rule(nq, xq, wq);
  else
    gll_rule(nq, xq, wq);
  for (int q = 0; q < nq; ++q)
    for (int m = 0; m < nm; ++m) B[q * nm + m] = lagrange_value(support, nm, m, xq[q]);
  for (int q = 0; q < nq; ++q)
    for (int k = 0; k < nq; ++k) D[q * nq + k] = lagrange_derivative(xq, nq, k, xq[q]);
}

}  // namespace mfx
