// tables.cpp -- host-side comparison of the compile-time 1D tables
// (tables.hpp) with the run-time basis tables.
#include "tables.hpp"

#include <algorithm>
#include <cmath>

#include "mfx_internal.hpp"

namespace mfx {

namespace {

template <int N>
double gll_err(const double *D) {
  const GllD<N> t{};
  double err = 0.0;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) err = std::max(err, std::fabs(t.D[i][j] - D[i * N + j]));
  return err;
}

}  // namespace

double gll_d_table_error(int n, const double *D) {
  switch (n) {
    case 2: return gll_err<2>(D);
    case 3: return gll_err<3>(D);
    case 4: return gll_err<4>(D);
    case 5: return gll_err<5>(D);
    case 6: return gll_err<6>(D);
    case 7: return gll_err<7>(D);
    case 8: return gll_err<8>(D);
    case 9: return gll_err<9>(D);
    default: throw Error{MFX_ERR_UNSUPPORTED, "gll_d_table_error: 2 <= n <= 9"};
  }
}

double eo_tables(const double *M, int O, int I, int S, std::vector<double> &out) {
  const int HO = O / 2, HI = I / 2, mi = I - 1;
  double mx = 0.0, dev = 0.0;
  for (int i = 0; i < O; ++i)
    for (int a = 0; a < I; ++a) {
      mx = std::max(mx, std::fabs(M[i * I + a]));
      dev = std::max(dev, std::fabs(M[i * I + a] - S * M[(O - 1 - i) * I + (I - 1 - a)]));
    }
  for (int i = 0; i < HO; ++i)
    for (int a = 0; a < HI; ++a) out.push_back(0.5 * (M[i * I + a] + S * M[i * I + mi - a]));
  for (int i = 0; i < HO; ++i)
    for (int a = 0; a < HI; ++a) out.push_back(0.5 * (M[i * I + a] - S * M[i * I + mi - a]));
  if (I & 1)
    for (int i = 0; i < HO; ++i) out.push_back(M[i * I + HI]);
  if (O & 1)
    for (int a = 0; a < HI; ++a) out.push_back(M[HO * I + a]);
  if ((I & 1) && (O & 1)) out.push_back(M[HO * I + HI]);
  return mx > 0.0 ? dev / mx : 0.0;
}

}  // namespace mfx
